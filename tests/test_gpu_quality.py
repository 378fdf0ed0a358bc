"""Training quality of mixed precision against f32 on the GPU: the metric's
"mixed vs FP32, <= 0.05 dB PSNR" requirement, checked on a reduced
benchmark scene (200k Gaussians, 16 views at 800x450) after 1500 steps from
the same start, camera sequence and seed."""
import numpy as np
import pytest

import gsabi
import scenes

N, VIEWS, W, H, F = 200_000, 16, 800, 450, 550.0
STEPS = 1500


def _scene():
    p, sh, cams = scenes.config2(n=N, seed=0, views=VIEWS, W=W, H=H, f=F)
    gp, gsh, _ = scenes.config2(n=N, seed=1, views=VIEWS, W=W, H=H, f=F)
    gt = gsabi.GSTrainer(shDeg=2, maxSplats=N)
    blank = [{"tw": W, "th": H, "rgb": np.zeros(W * H * 3, np.float32)}]
    gt.setup({"data": gp, "n": N, "sh": gsh}, [dict(cams[0], imgIdx=0)], blank, W, H, 8.0)
    images = []
    for c in cams:
        img = gt.renderView(c)
        images.append({"tw": W, "th": H, "rgb": np.clip(img[..., :3], 0, 1).reshape(-1).copy()})
    gt.destroy()
    return p, sh, cams, images


def _train(p, sh, cams, images, precision):
    t = gsabi.GSTrainer(shDeg=2, maxSplats=N, posLr=1.6e-4, seed=7, precision=precision)
    t.setup({"data": p, "n": N, "sh": sh}, cams, images, W, H, 8.0)
    t.iter = 2000
    before = float(np.mean([t.evalCamPsnr(ci) for ci in range(VIEWS)]))
    for _ in range(STEPS):
        t.stepOnce()
    after = float(np.mean([t.evalCamPsnr(ci) for ci in range(VIEWS)]))
    t.destroy()
    return before, after


@pytest.mark.gpu
def test_mixed_precision_psnr_within_005_db_of_f32():
    p, sh, cams, images = _scene()
    b32, a32 = _train(p, sh, cams, images, gsabi.GS_PREC_F32)
    bmx, amx = _train(p, sh, cams, images, gsabi.GS_PREC_MIXED)
    print(f"PSNR f32 {b32:.3f} -> {a32:.3f} dB, mixed {bmx:.3f} -> {amx:.3f} dB, delta {amx - a32:+.4f} dB")
    assert a32 > b32 + 0.5                 # the run trains
    assert abs(bmx - b32) <= 0.01          # same start up to the bf16 rounding of the initial SH
    assert abs(amx - a32) <= 0.05
