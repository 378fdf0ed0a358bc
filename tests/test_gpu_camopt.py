"""Camera optimisation on the GPU (camOpt / expComp), through the C ABI,
against the restatement (oracle/ and oracle/camopt.py).

- The per-camera gradient rows (x64 fixed-point i32: pose and exposure of
  the trained camera, the shared log-focal) equal the restatement's after
  every step, bit for bit, in f32 and mixed precision; so do the params.
- The camera Adam step at iteration 1525 (trainer.js:468-470, :476-537)
  gives the same double-precision poses, focal and exposures; camera 0 and
  the holdout stay put; later steps train through the updated cameras and
  stay bit-exact.
- evalCamPsnrRefined (trainer.js:583-612): the refined pose's PSNR equals
  the restatement's, and neither the Gaussians nor the stored camera change.
"""
import math

import numpy as np
import pytest

import camopt as CO
import gsabi
import scenes
from helpers import OracleTrainer, f32
from test_gpu_configs import _targets

N = 8000
RADIUS = 8.0


def _scene():
    sizes = [(320, 240)] * 4
    p, sh, cams = scenes.config4(n=N, seed=0, sizes=sizes)
    gp, gsh, _ = scenes.config4(n=N, seed=1, sizes=sizes)
    images, W, H = _targets(gp, gsh, cams, 2, N)
    cams = [dict(c) for c in cams]
    cams[2]["g"], cams[2]["b"] = 0.05, -0.02         # a camera with its own exposure
    return p, sh, cams, images, W, H


def _pair(precision, cam_opt=True, exp_comp=True):
    p, sh, cams, images, W, H = _scene()
    prec = gsabi.GS_PREC_MIXED if precision == "mixed" else gsabi.GS_PREC_F32
    t = gsabi.GSTrainer(shDeg=2, maxSplats=N, precision=prec, camOpt=int(cam_opt), expComp=int(exp_comp))
    t.setup({"data": p, "n": N, "sh": sh}, cams, images, W, H, RADIUS)
    ref = OracleTrainer(p, sh, cams, images, RADIUS, opts={"camOpt": cam_opt, "expComp": exp_comp}, shDeg=2,
                        precision=precision)
    return t, ref, cams


def _same_params(t, ref, msg):
    out = t.readGaussians()
    np.testing.assert_array_equal(out["data"].view(np.uint32), ref.P.view(np.uint32), err_msg=msg)
    np.testing.assert_array_equal(out["sh"].view(np.uint32), ref.SH.view(np.uint32), err_msg=msg)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["f32", "mixed"])
def test_gpu_camera_gradients_bit_exact(precision):
    t, ref, cams = _pair(precision)
    t.iter = ref.iter = 1400
    for k, ci in enumerate((0, 1, 2, 1, 3)):
        print(f"{precision} step {k} camera {ci}", flush=True)
        t.stepOnce(ci)
        r = ref.step(ci)
        g = t.readCamGrads()
        np.testing.assert_array_equal(g, r["camGrad"], err_msg=f"step {k} camera {ci}")
        assert np.all(g[ci] != 0) and g[len(cams), 0] != 0
        _same_params(t, ref, f"step {k}")
    t.destroy()


@pytest.mark.gpu
def test_gpu_camera_adam_step_bit_exact():
    t, ref, cams = _pair("f32")
    t.iter = ref.iter = 1520
    t.holdout = ref.holdout = 3
    st0 = [t.camState(ci) for ci in range(len(cams))]
    for k, ci in enumerate((1, 2, 0, 3, 1, 2, 1)):       # iterations 1521..1527; the Adam step at 1525
        print(f"step {k} camera {ci} iter {1521 + k}", flush=True)
        t.stepOnce(ci)
        r = ref.step(ci)
        np.testing.assert_array_equal(t.readCamGrads(), ref.cam_grads(), err_msg=f"step {k}")
        _same_params(t, ref, f"step {k}")
        if 1521 + k == 1525:
            assert not ref.cam_grads().any()
    for ci in range(len(cams)):
        got = t.camState(ci)
        want = ref.camopt.cams[ci]
        assert got["R"] == want["R"] and got["t"] == want["t"], ci
        assert got["f"] == want["f"] and got["g"] == want["g"] and got["b"] == want["b"], ci
    # camera 0 (the gauge anchor) and the holdout keep their pose; the
    # others moved; the shared focal scaled every camera
    for ci in (0, 3):
        assert t.camState(ci)["R"] == st0[ci]["R"] and t.camState(ci)["t"] == st0[ci]["t"]
    for ci in (1, 2):
        assert t.camState(ci)["R"] != st0[ci]["R"] and t.camState(ci)["t"] != st0[ci]["t"]
        assert t.camState(ci)["g"] != st0[ci]["g"]
    assert all(t.camState(ci)["f"] != st0[ci]["f"] for ci in range(len(cams)))
    t.destroy()


@pytest.mark.gpu
def test_gpu_eval_cam_psnr_refined():
    t, ref, cams = _pair("f32", cam_opt=False, exp_comp=True)
    t.iter = ref.iter = 2000
    for ci in (0, 1):
        t.stepOnce(ci)
        ref.step(ci)
    before = t.readGaussians()
    st0 = t.camState(2)
    ci, steps = 2, 4
    base = CO.CamState(cams).cams[ci]        # the trainer's start: f32 inputs as doubles

    def grad(R, tt, g, b):
        return ref.eval_pass(ci, {"R": R, "t": tt, "g": g, "b": b})[1]

    R, tt, g, b = CO.refine_pose(base["R"], base["t"], grad, steps, f32(RADIUS), exp_comp=True)
    want = ref.eval_pass(ci, {"R": R, "t": tt, "g": g, "b": b})[0]
    got = t.evalCamPsnrRefined(ci, steps)
    print(f"refined PSNR {got:.6f} dB (restatement {want:.6f} dB), plain {t.evalCamPsnr(ci):.6f} dB")
    # the pose steps come from integer gradient sums and double host math on
    # both sides; the PSNR's squared-error sum is a double atomic on the GPU
    assert abs(got - want) <= 1e-9 * abs(want)
    after = t.readGaussians()
    np.testing.assert_array_equal(after["data"].view(np.uint32), before["data"].view(np.uint32))
    assert t.camState(2) == st0
    assert not t.readCamGrads().any()
    t.destroy()
