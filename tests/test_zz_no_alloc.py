"""gs_export_ply_noalloc: the PLY export without any device allocation.

Its rows come from the same kernel as gs_export_ply, staged chunk by chunk in
the library's fixed device array, so the file must equal gs_export_ply's byte
for byte. Two calls in a row: a trainer whose rows span several chunks (SH
degree 3, 150k Gaussians: 62 floats a row, 67650 rows a chunk), then a smaller
trainer (SH degree 1, 3000 Gaussians), so the second call runs over what the
first left in the array. The camera limit of the array is an error, not a
truncated bake."""
import numpy as np
import pytest

import gsabi
import scenes


def _trainer(n, deg, seed, ncams=4, W=96, H=64):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 16), np.float32)
    p[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    p[:, 2] += 4.0
    p[:, 3:6] = rng.normal(-4, 0.5, (n, 3))
    p[:, 6:10] = rng.normal(0, 1, (n, 4))
    p[:, 10:14] = rng.normal(0, 1, (n, 4))
    K = (deg + 1) ** 2 - 1
    sh = rng.normal(0, 0.05, (n, 3 * K)).astype(np.float32) if K else None
    cams = scenes.ring_cameras(ncams, 4.0, W, H, 80.0)
    images = [{"tw": W, "th": H, "rgb": np.full(W * H * 3, 0.5, np.float32)} for _ in cams]
    t = gsabi.GSTrainer(shDeg=deg, maxSplats=n)
    g = {"data": p, "n": n}
    if sh is not None:
        g["sh"] = sh
    t.setup(g, cams, images, W, H, 4.0)
    return t


@pytest.mark.gpu
def test_noalloc_export_equals_the_allocating_one_twice():
    for n, deg, seed, steps in ((150000, 3, 1, 3), (3000, 1, 2, 2)):
        print(f"noalloc export: n={n} shDeg={deg}", flush=True)
        t = _trainer(n, deg, seed)
        for ci in range(steps):
            t.stepOnce(ci)
        for bake in (True, False):
            want = t.exportPly(bake=bake)
            got = t.exportPly(bake=bake, noalloc=True)
            assert len(got) == len(want), (n, bake)
            assert got == want, (n, deg, bake)
        t.destroy()


@pytest.mark.gpu
def test_noalloc_export_rejects_more_cameras_than_its_array_holds():
    print("noalloc export: 4097 cameras", flush=True)
    n, W, H = 64, 16, 16
    t = _trainer(n, 0, 3, ncams=4097, W=W, H=H)
    with pytest.raises(gsabi.GsError, match="4096 camera centres"):
        t.exportPly(bake=True, noalloc=True)
    assert t.exportPly(bake=True) == t.exportPly(bake=True)   # the allocating form has no such limit
    t.destroy()
