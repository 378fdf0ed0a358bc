"""The entry budget of the tile-segment binning (the reference's scan,
shaders.js:321-370), restated in oracle.tile_budget, against hand-worked
cases: padding of long segments, whole-tile drops, a later tile that still
fits after a drop, chunk boundaries, and empty tiles counted once the
running total is past the budget. The GPU side is in test_gpu_parity.py."""
import numpy as np

import oracle as O


def test_padded_size():
    assert [O.padded_size(c) for c in (0, 1, 2047, 2048, 2049, 4096, 4097)] == [0, 1, 2047, 2048, 4096, 4096, 8192]


def test_everything_fits():
    cnt = np.array([5, 0, 3000, 7])
    kept, dropped = O.tile_budget(cnt, 5 + 4096 + 7)
    assert kept.tolist() == [5, 0, 3000, 7] and dropped == 0


def test_padded_segment_drops_whole_tiles_and_later_small_tiles_fit():
    # 3000 keys pad to 4096: 10 + 4096 > 4100 drops tile 1; tile 2 (3) fits
    # again (10 + 3 <= 4100); tile 3 (90) does not (13 + 90 <= 4100: fits);
    # tile 4 (4090) does not (103 + 4090 > 4100)
    cnt = np.array([10, 3000, 3, 90, 4090])
    kept, dropped = O.tile_budget(cnt, 4100)
    assert kept.tolist() == [10, 0, 3, 90, 0] and dropped == 2


def test_chunks_start_at_the_unbudgeted_prefix():
    # 130 tiles = chunks of 64, 64, 2. Chunk 0 holds one 100-key tile that
    # the budget of 60 drops; chunk 1 still starts at 100 (the prefix before
    # drops), so all its 64 tiles (empty ones included) count as dropped;
    # so do chunk 2's.
    cnt = np.zeros(130, np.int64)
    cnt[0] = 100
    kept, dropped = O.tile_budget(cnt, 60)
    assert not kept.any()
    assert dropped == 1 + 64 + 2
    # with a budget of 100 the big tile fits; nothing after it is dropped
    kept, dropped = O.tile_budget(cnt, 100)
    assert kept[0] == 100 and dropped == 0


def test_apply_tile_budget():
    ts = np.array([0, 2, 2, 5, 6], np.uint32)         # tiles of 2, 0, 3, 1 keys
    ids = np.arange(6, dtype=np.uint32)
    keys = np.arange(6, dtype=np.uint64) + 100
    i2, k2, ts2, dropped = O.apply_tile_budget(ids, keys, ts, 3)   # 2 fits, 2 + 3 > 3 drops, 2 + 1 fits
    assert i2.tolist() == [0, 1, 5] and k2.tolist() == [100, 101, 105]
    assert ts2.tolist() == [0, 2, 2, 2, 3] and dropped == 1
