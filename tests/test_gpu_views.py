"""Batched views (gs_step_views): several training views per step, rasterized
forward and backward in one persistent launch each, gradients added across
views, one Adam step; replayed as a hipGraph when the step's shape repeats.

* Bit-exact against the restatement's batched step (helpers.OracleTrainer.
  step_views) on uneven view sizes: every view's image, T, end, dL/dC and
  fixed-point sums, the summed gradients, and the params and SH after Adam,
  over several steps, in f32 and 16-bit storage, at the default raster shape
  and at 8- and 32-pixel tiles with smaller LDS batches; with 8-pixel tiles
  on a view too large for the tile-segment binning, which then bins by the
  radix path inside the batched step.
* Graph replay equals eager execution, step after step, with single-view
  steps, renders and readbacks in between.
* At full size (config 4: 1M Gaussians, 8 views from 800x450 to 1920x1080,
  bf16 storage): every view's sums equal a single-view step's on the same
  parameters, and the summed gradient equals those steps' gradients added
  in view order.
"""
import numpy as np
import pytest

import gsabi
import scenes
from helpers import OracleTrainer

pytestmark = pytest.mark.gpu

SIZES = [(160, 96), (96, 64), (200, 120), (130, 70)]


def _scene(n=6000, seed=0, sizes=SIZES):
    p, sh = scenes.gaussians(n, seed, (-1.5, -1.5, -1.5), (1.5, 1.5, 1.5))
    cams = []
    rng = np.random.default_rng(seed + 5)
    images = []
    for k, (W, H) in enumerate(sizes):
        a = 2 * np.pi * k / len(sizes)
        R, t = scenes.look_at((4.0 * np.sin(a), 0.3, -4.0 * np.cos(a)))
        cams.append({"imgIdx": k, "R": R, "t": t, "f": 1.1 * W, "cx": W / 2, "cy": H / 2, "w": W, "h": H})
        rgb = rng.uniform(0, 1, W * H * 3).astype(np.float32)
        rgb[rng.random(W * H * 3) < 0.02] = -1.0     # some invalid pixels
        images.append({"tw": W, "th": H, "rgb": rgb})
    return p, sh, cams, images


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.mark.parametrize("precision,tile,batch", [("f32", 16, 256), ("bf16", 16, 256), ("f32", 8, 128),
                                                  ("bf16", 32, 64)])
def test_views_bit_exact_against_restatement(precision, tile, batch):
    print(f"batched views: {precision}, tile {tile}, batch {batch}", flush=True)
    p, sh, cams, images = _scene()
    prec = {"f32": gsabi.GS_PREC_F32, "bf16": gsabi.GS_PREC_BF16}[precision]
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], viewsPerStep=4, precision=prec, tile=tile, batch=batch)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 120, 2.0)
    ref = OracleTrainer(p, sh, cams, images, 2.0, shDeg=2, precision=precision, tile=tile)
    t.iter = ref.iter = 1996      # the second step starts at 2000: all SH bands from there
    for s, order in enumerate(([0, 1, 2, 3], [2, 0, 3], [3, 1, 0, 2])):
        print(f"batched step {s}: views {order}", flush=True)
        outs = t.step_views_debug(order)
        per, gF, sG = ref.step_views(order)
        for k, (o, r) in enumerate(zip(outs, per)):
            lab = f"step {s} view {k}"
            np.testing.assert_array_equal(_bits(o["img"]), _bits(r["img"]), err_msg=lab)
            np.testing.assert_array_equal(_bits(o["T"]), _bits(r["T"]), err_msg=lab)
            np.testing.assert_array_equal(o["end"], r["end"], err_msg=lab)
            np.testing.assert_array_equal(_bits(o["gC"][:, :3]), _bits(r["gC"]), err_msg=lab)
            np.testing.assert_array_equal(o["accum"], r["accum"], err_msg=lab)
            tr = o["tileRanges"].astype(np.int64)
            np.testing.assert_array_equal(tr[:, 1] - tr[:, 0], np.diff(r["ts"].astype(np.int64)), err_msg=lab)
        np.testing.assert_array_equal(_bits(outs[0]["gradF"][:, :14]), _bits(gF[:, :14]), err_msg=f"step {s}")
        np.testing.assert_array_equal(_bits(outs[0]["shGrad"]), _bits(sG), err_msg=f"step {s}")
        g = t.readGaussians()
        np.testing.assert_array_equal(_bits(g["data"]), _bits(ref.P), err_msg=f"params step {s}")
        np.testing.assert_array_equal(_bits(g["sh"]), _bits(ref.SH), err_msg=f"sh step {s}")
        assert t.info().iter == ref.iter
    t.destroy()


def test_view_above_segment_tiles_bit_exact():
    """8-pixel tiles on a 1040x1040 view: 16,900 tiles, more than the
    tile-segment binning's 16,384, so that view bins by the depth-then-tile
    radix path into its own slot while the others take the batched segment
    launches. Bit-exact against the restatement (eager debug steps), and a
    trainer replaying the step from a hipGraph ends in the same state."""
    sizes = [(1040, 1040), (160, 96), (200, 120)]
    p, sh, cams, images = _scene(n=6000, seed=4, sizes=sizes)
    kw = dict(shDeg=2, maxSplats=p.shape[0], viewsPerStep=3, tile=8, batch=128)
    t = gsabi.GSTrainer(graph=0, **kw)
    g = gsabi.GSTrainer(graph=1, **kw)
    for x in (t, g):
        x.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 1040, 1040, 2.0)
    ref = OracleTrainer(p, sh, cams, images, 2.0, shDeg=2, precision="f32", tile=8)
    t.iter = g.iter = ref.iter = 1996
    for s, order in enumerate(([0, 1, 2], [1, 0, 2], [2, 1, 0])):
        print(f"batched step {s}: views {order}", flush=True)
        outs = t.step_views_debug(order)
        g.stepViews(order)
        per, gF, sG = ref.step_views(order)
        for k, (o, r) in enumerate(zip(outs, per)):
            lab = f"step {s} view {order[k]}"
            np.testing.assert_array_equal(_bits(o["img"]), _bits(r["img"]), err_msg=lab)
            np.testing.assert_array_equal(_bits(o["T"]), _bits(r["T"]), err_msg=lab)
            np.testing.assert_array_equal(o["end"], r["end"], err_msg=lab)
            np.testing.assert_array_equal(o["accum"], r["accum"], err_msg=lab)
            tr = o["tileRanges"].astype(np.int64)
            np.testing.assert_array_equal(tr[:, 1] - tr[:, 0], np.diff(r["ts"].astype(np.int64)), err_msg=lab)
        np.testing.assert_array_equal(_bits(outs[0]["gradF"][:, :14]), _bits(gF[:, :14]), err_msg=f"step {s}")
        a, b = t.readGaussians(), g.readGaussians()
        np.testing.assert_array_equal(_bits(a["data"]), _bits(ref.P), err_msg=f"params step {s}")
        np.testing.assert_array_equal(_bits(a["sh"]), _bits(ref.SH), err_msg=f"sh step {s}")
        np.testing.assert_array_equal(_bits(b["data"]), _bits(ref.P), err_msg=f"graph params step {s}")
        np.testing.assert_array_equal(_bits(b["sh"]), _bits(ref.SH), err_msg=f"graph sh step {s}")
    t.destroy()
    g.destroy()


def test_graph_replay_and_launch_forms_equal_eager():
    """Trainers from the same start and seed: graph replay (first step eager,
    second captured, later ones replayed) with the overlapped launches (each
    view's raster on a second stream beside the next views' binning), graph
    replay with the persistent launches, and eager persistent launches.
    Between batched steps, single-view steps, a render and readbacks run on
    the same buffers. State bit-identical after every step."""
    p, sh, cams, images = _scene(n=5000, seed=3, sizes=[(128, 96)] * 4)
    tr = []
    for graph, launch in ((1, gsabi.GS_VIEWS_OVERLAP), (1, gsabi.GS_VIEWS_PERSISTENT),
                          (0, gsabi.GS_VIEWS_PERSISTENT), (0, gsabi.GS_VIEWS_OVERLAP)):
        t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], viewsPerStep=3, graph=graph, viewsLaunch=launch,
                            seed=11)
        t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 128, 96, 2.0)
        t.iter = 2500
        tr.append(t)
    for s in range(8):
        print(f"graph replay step {s}", flush=True)
        for t in tr:
            t.stepViews([-1, -1, -1])
            if s % 3 == 2:
                t.stepOnce()
                t.renderView(cams[1])
        ref = tr[2].readGaussians()
        loss = tr[2].readLoss()
        for k in (0, 1, 3):
            a = tr[k].readGaussians()
            np.testing.assert_array_equal(_bits(a["data"]), _bits(ref["data"]), err_msg=f"trainer {k} step {s}")
            np.testing.assert_array_equal(_bits(a["sh"]), _bits(ref["sh"]), err_msg=f"trainer {k} step {s}")
            assert tr[k].readLoss() == pytest.approx(loss, rel=1e-12)   # double atomics
    for t in tr:
        t.destroy()


def test_config4_batched_sums_equal_single_view_steps():
    p, sh, cams = scenes.config4(n=1_000_000, seed=0)
    images = [{"tw": c["w"], "th": c["h"], "rgb": np.full(c["w"] * c["h"] * 3, 0.4, np.float32)} for c in cams]
    maxW, maxH = max(c["w"] for c in cams), max(c["h"] for c in cams)
    kw = dict(shDeg=2, maxSplats=p.shape[0], precision=gsabi.GS_PREC_BF16)
    t = gsabi.GSTrainer(viewsPerStep=8, **kw)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, maxW, maxH, 8.0)
    t.iter = 2500
    outs = t.step_views_debug(list(range(8)), want=("accum",))
    t.destroy()
    total = None
    seen = None
    for k in range(8):
        print(f"config 4 single view {k} ({cams[k]['w']}x{cams[k]['h']})", flush=True)
        s = gsabi.GSTrainer(**kw)
        s.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, maxW, maxH, 8.0)
        s.iter = 2500
        g = s.step_debug(k, want=("accum", "gradF", "cnt"))
        s.destroy()
        np.testing.assert_array_equal(outs[k]["accum"], g["accum"], err_msg=f"view {k}")
        assert outs[k]["numKeys"] == g["numKeys"] and g["numKeys"] > 100_000
        gv = g["gradF"][:, :14]
        mask = g["cnt"] > 0
        if total is None:
            total = np.zeros_like(gv)
            seen = np.zeros(mask.shape, bool)
        first, later = mask & ~seen, mask & seen
        total[first] = gv[first]
        total[later] = total[later] + gv[later]
        seen |= mask
    np.testing.assert_array_equal(_bits(outs[0]["gradF"][:, :14]), _bits(total))


def test_views_with_an_empty_view_and_repeated_views():
    """A batch holding a view that sees no Gaussian (camera facing away: no
    keys, every tile range empty) and the same view twice: bit-exact against
    the restatement's batched step, through the one-launch-per-stage binning."""
    p, sh, cams, images = _scene(n=4000, seed=7)
    away = dict(cams[1])
    R, tvec = scenes.look_at((0.0, 0.0, -4.0), center=(0.0, 0.0, -9.0))   # looks away from the scene
    away.update(R=R, t=tvec, imgIdx=1)
    cams = [cams[0], away, cams[2], cams[3]]
    t = gsabi.GSTrainer(shDeg=2, maxSplats=p.shape[0], viewsPerStep=4)
    t.setup({"data": p, "n": p.shape[0], "sh": sh}, cams, images, 200, 120, 2.0)
    ref = OracleTrainer(p, sh, cams, images, 2.0, shDeg=2)
    t.iter = ref.iter = 2400
    for s, order in enumerate(([1, 0, 2], [2, 2, 1, 3])):
        print(f"batched step {s}: views {order}", flush=True)
        outs = t.step_views_debug(order)
        per, gF, sG = ref.step_views(order)
        for k, (o, r) in enumerate(zip(outs, per)):
            lab = f"step {s} view {k}"
            assert o["numKeys"] == len(r["ids"]), lab
            np.testing.assert_array_equal(_bits(o["img"]), _bits(r["img"]), err_msg=lab)
            np.testing.assert_array_equal(o["accum"], r["accum"], err_msg=lab)
        assert outs[0 if order[0] == 1 else 2]["numKeys"] == 0
        np.testing.assert_array_equal(_bits(outs[0]["gradF"][:, :14]), _bits(gF[:, :14]), err_msg=f"step {s}")
        g = t.readGaussians()
        np.testing.assert_array_equal(_bits(g["data"]), _bits(ref.P), err_msg=f"params step {s}")
        np.testing.assert_array_equal(_bits(g["sh"]), _bits(ref.SH), err_msg=f"sh step {s}")
    t.destroy()
