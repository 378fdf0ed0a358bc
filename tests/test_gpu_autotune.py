"""Autotuner (gs_autotune): times the 9 raster shapes (tiles of 8, 16 and 32
pixels x LDS batches of 64, 128 and 256 keys) on given views and keeps the
fastest eligible one.

* The sweep leaves the trainer as it was: a trainer that runs it between two
  steps ends bit-identical to a twin that does not (params, SH, Adam moments
  through the next step's update, iteration count, loss window).
* The entries are complete and the choice follows the rule: the fastest
  measured shape that drops no more keys to the entry budget than 16-pixel
  tiles do; a shape dropping more is flagged, not chosen.
* Config 4 (1M Gaussians, 8 views from 800x450 to 1920x1080, bf16 storage,
  one batched pass): every shape is measured and printed, the 8-pixel ones
  too (their views above 16,384 tiles bin by the radix path inside the
  batched step).
"""
import numpy as np
import pytest

import gsabi
import scenes

pytestmark = pytest.mark.gpu

SIZES = [(160, 96), (200, 120), (128, 128), (96, 64)]


def _small(n=8000, seed=0):
    p, sh = scenes.gaussians(n, seed, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))
    rng = np.random.default_rng(seed + 3)
    cams, images = [], []
    for k, (W, H) in enumerate(SIZES):
        a = 2 * np.pi * k / len(SIZES)
        R, t = scenes.look_at((4.0 * np.sin(a), 0.2, -4.0 * np.cos(a)))
        cams.append({"imgIdx": k, "R": R, "t": t, "f": 1.1 * W, "cx": W / 2, "cy": H / 2, "w": W, "h": H})
        images.append({"tw": W, "th": H, "rgb": rng.uniform(0, 1, W * H * 3).astype(np.float32)})
    return p, sh, cams, images


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _check_rule(entries):
    assert [(e["tile"], e["batch"]) for e in entries] == [(t, b) for t in (8, 16, 32) for b in (64, 128, 256)]
    d16 = entries[5]["dropped"]
    assert entries[5]["status"] == gsabi.GS_TUNE_MEASURED
    for e in entries:
        assert e["keys"] > 0
        want = gsabi.GS_TUNE_MORE_DROPS if e["dropped"] > d16 else gsabi.GS_TUNE_MEASURED
        assert e["status"] == want, e
        if want == gsabi.GS_TUNE_MEASURED:
            assert e["ms"] > 0
    ok = [e for e in entries if e["status"] == gsabi.GS_TUNE_MEASURED]
    return min(ok, key=lambda e: e["ms"])


@pytest.mark.parametrize("views", [[1], [0, 2, 3]])
def test_sweep_leaves_state_and_applies_the_rule(views):
    p, sh, cams, images = _small()
    tr = []
    for _ in range(2):
        t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], viewsPerStep=3, seed=5)
        t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 128, 2.0)
        t.iter = 2500
        tr.append(t)
    a, b = tr
    for t in tr:
        t.stepViews([0, 1, 2])
        t.stepOnce(3)
    print(f"autotune over views {views}", flush=True)
    res = b.autotune(views, reps=2)
    best = _check_rule(res)
    info = b.info()
    assert (info.tile, info.batch) == (best["tile"], best["batch"])
    assert info.iter == a.info().iter
    for e in res:
        print(f"  tile {e['tile']:2d} batch {e['batch']:3d} status {e['status']} {e['ms']:.3f} ms "
              f"keys {e['keys']} dropped {e['dropped']}", flush=True)
    b.setRasterShape(16, 256)
    for t in tr:
        t.stepViews([2, 0, 1])
        t.stepOnce(1)
    ga, gb = a.readGaussians(), b.readGaussians()
    np.testing.assert_array_equal(_bits(ga["data"]), _bits(gb["data"]))
    np.testing.assert_array_equal(_bits(ga["sh"]), _bits(gb["sh"]))
    assert b.readLoss() == pytest.approx(a.readLoss(), rel=1e-12)   # double atomics
    for t in tr:
        t.destroy()


def test_shapes_dropping_more_keys_are_not_chosen():
    p, sh, cams, images = _small(n=20000, seed=2)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0])
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 128, 2.0)
    keys16 = t.autotune([1], reps=1)[5]["keys"]
    t.destroy()
    # a budget of half the 16-pixel keys: every shape drops, the 8-pixel ones most
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], entriesCap=keys16 // 2)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 128, 2.0)
    res = t.autotune([1], reps=1)
    for e in res:
        print(f"  tile {e['tile']:2d} batch {e['batch']:3d} status {e['status']} keys {e['keys']} "
              f"dropped {e['dropped']}", flush=True)
    best = _check_rule(res)
    assert res[5]["dropped"] > 0
    assert all(e["status"] == gsabi.GS_TUNE_MORE_DROPS for e in res[:3])
    assert (t.info().tile, t.info().batch) == (best["tile"], best["batch"])
    t.destroy()


def test_config4_batched_sweep():
    p, sh, cams = scenes.config4(n=1_000_000, seed=0)
    images = [{"tw": c["w"], "th": c["h"], "rgb": np.full(c["w"] * c["h"] * 3, 0.4, np.float32)} for c in cams]
    maxW, maxH = max(c["w"] for c in cams), max(c["h"] for c in cams)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], precision=gsabi.GS_PREC_BF16, viewsPerStep=8)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, maxW, maxH, 8.0)
    t.iter = 2500
    print("config 4 batched sweep", flush=True)
    res = t.autotune(list(range(8)), reps=3)
    for e in res:
        print(f"  tile {e['tile']:2d} batch {e['batch']:3d} status {e['status']} {e['ms']:.3f} ms "
              f"keys {e['keys']} dropped {e['dropped']}", flush=True)
    best = _check_rule(res)
    big = any(((c["w"] + 7) // 8) * ((c["h"] + 7) // 8) > 16384 for c in cams)
    assert big and all(e["status"] != gsabi.GS_TUNE_TOO_MANY_TILES for e in res)
    assert (t.info().tile, t.info().batch) == (best["tile"], best["batch"])
    t.stepViews(list(range(8)))
    assert np.isfinite(t.readGaussians()["data"]).all()
    t.destroy()


def test_shape_and_autotune_arguments_are_checked():
    p, sh, cams, images = _small(n=2000)
    with pytest.raises(gsabi.GsError):
        gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], tile=12)
    with pytest.raises(gsabi.GsError):
        gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], batch=96)
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], viewsPerStep=2)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 128, 2.0)
    for tile, batch in ((16, 32), (4, 256), (64, 64)):
        with pytest.raises(gsabi.GsError):
            t.setRasterShape(tile, batch)
    assert (t.info().tile, t.info().batch) == (16, 256)
    with pytest.raises(gsabi.GsError):
        t.autotune([0, 1, 2], reps=1)          # more views than viewsPerStep
    with pytest.raises(gsabi.GsError):
        t.autotune([-1], reps=1)               # the sweep takes explicit cameras
    with pytest.raises(gsabi.GsError):
        t.autotune([len(cams)], reps=1)
    t.setRasterShape(32, 64)
    assert (t.info().tile, t.info().batch) == (32, 64)
    t.stepViews([0, 1])
    assert np.isfinite(t.readGaussians()["data"]).all()
    t.destroy()
