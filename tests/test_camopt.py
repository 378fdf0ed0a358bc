"""Camera pose, focal and exposure optimisation (camOpt / expComp):
gradients from the loss (shaders.js:614-616) and the chain (:834-858), the
per-camera Adam of trainer.js:476-537 and the test-time refinement of
evalCamPsnrRefined (:583-612).

CPU: the restatement's analytic camera gradients against central
differences of the double-precision loss (the reference's gradCheckPose
protocol, gradcheck.js:15-90: rotation R <- exp(w^) R, translation, shared
log-focal, exposure log-gain and bias), on the reference rig and on a rotated
and translated copy of it with a non-unit exposure; x64 fixed-point sums against plain sums; the
camera Adam (oracle/camopt.py) against closed forms. The GPU tests are in
test_gpu_camopt.py."""
import math

import numpy as np
import pytest

import camopt as CO
import oracle as O
import scenes


def _rig(rotated):
    params, cams, images, W, H, radius, opts = scenes.gradcheck_rig()
    cam = dict(cams[0])
    if rotated:
        # the same camera-space scene seen from a rotated, translated camera:
        # p_world = R0^T (p_cam - t0)
        R0 = CO.rodrigues([0.3, -0.2, 0.5])
        t0 = [0.1, -0.2, 0.3]
        R0f = np.array(R0, np.float32).astype(np.float64).reshape(3, 3)
        t0f = np.array(t0, np.float32).astype(np.float64)
        params = params.copy()
        params[:, :3] = ((params[:, :3].astype(np.float64) - t0f) @ R0f).astype(np.float32)
        cam["R"], cam["t"] = R0, t0
        cam["g"], cam["b"] = 0.08, -0.03      # a non-unit exposure gain and a bias
    o = O.make_opts(eCut=opts["eCut"], aMin=opts["aMin"], radClamp=opts["radClamp"],
                    anisoReg=opts["anisoReg"], shDeg=0)
    tgt = O.pack_targets(images[0]["rgb"])
    return params, cam, tgt, o


def _analytic(st, cam_d, params, tgt, fixed=False):
    cam = O.make_cam(cam_d)
    proj, rect, cnt = st.project(cam, params, None)
    ids, ts, _ = st.bin_sort(cam, proj, rect, cnt)
    img, T, end = st.render_fwd(cam, proj, ids, ts)
    gC, valid, stats, efix, esum = st.loss_charb(cam, img, tgt, cam_grads=True)
    g = st.render_bwd(cam, proj, ids, ts, T, end, gC, valid, fixed=fixed)
    kw = {"gfix": g} if fixed else {"gsum": g}
    _, _, cfix, csum = st.chain(cam, params, None, proj, cam_grads=True, **kw)
    return cfix, csum, efix, esum, int(valid.sum()), int((cnt > 0).sum())


def _loss(st, cam_d, params, tgt):
    return O.loss_of(st, O.make_cam(cam_d), params, None, tgt)


def _perturbed(cam, k, h):
    """gradcheck.js:64-76: pose(sign) for quantity k at step h."""
    c = dict(cam)
    if k < 3:
        w = [0.0, 0.0, 0.0]
        w[k] = h
        c["R"] = CO.m3mul(CO.rodrigues(w), [float(x) for x in cam["R"]])
    elif k < 6:
        t = [float(x) for x in cam["t"]]
        t[k - 3] += h
        c["t"] = t
    elif k == 6:
        c["f"] = cam["f"] * math.exp(h)
    elif k == 7:
        c["g"] = cam.get("g", 0.0) + h
    else:
        c["b"] = cam.get("b", 0.0) + h
    return c


NAMES = ["rot.x", "rot.y", "rot.z", "t.x", "t.y", "t.z", "logf", "expGain", "expBias"]
STEPS = [1e-3, 1e-3, 1e-3, 2e-3, 2e-3, 2e-3, 1e-3, 5e-3, 5e-3]


@pytest.mark.parametrize("rotated,smooth", [(False, False), (False, True), (True, True)])
def test_camera_gradients_match_central_differences(oracle_lib, rotated, smooth):
    """Analytic camera gradients (double-precision restatement) against
    central differences of the double-precision loss.

    A pose change moves every splat at once, so on the reference rig (eCut 9,
    aMin 1e-4) the step crosses many alpha cutoffs; each crossing is a jump
    in the loss, and their count grows with h, which leaves the differences
    biased by ~1e-3 at every h: there the reference's own check applies
    (its rig and camera, relative error < 0.05, gradcheck.js:87). With the
    cutoffs pushed out of
    the way (eCut 40, aMin 1e-14) the loss is smooth and the differences
    converge as h^2 until the f32 rounding of the perturbed camera fields
    dominates: there the bound is 1e-3, on the reference camera and on the
    rotated, translated one with a non-unit exposure (which a transposed
    rotation, a dropped Jacobian term or a dropped gain factor fails)."""
    params, cam, tgt, o = _rig(rotated)
    if smooth:
        o = O.make_opts(eCut=40.0, aMin=1e-14, radClamp=10.0, anisoReg=0.0, shDeg=0)
    st = O.Stage(o, "f64")
    _, csum, _, esum, _, _ = _analytic(st, cam, params, tgt)
    analytic = list(csum[:6]) + [csum[6], esum[0], esum[1]]
    # errors relative to the component, or to a tenth of its block's norm
    # (rotation, translation) for a component much smaller than the rest
    block = [np.linalg.norm(analytic[0:3])] * 3 + [np.linalg.norm(analytic[3:6])] * 3 + [0.0] * 3
    report = []
    for k in range(9):
        best = None
        for h in ((STEPS[k], STEPS[k] / 4) if smooth else (STEPS[k], STEPS[k] / 4, STEPS[k] / 16)):
            lp = _loss(st, _perturbed(cam, k, h), params, tgt)
            lm = _loss(st, _perturbed(cam, k, -h), params, tgt)
            fd = (lp - lm) / (2 * h)
            den = max(abs(analytic[k]), abs(fd), 0.1 * block[k])
            e = abs(analytic[k] - fd) / den
            best = e if best is None else min(best, e)
        report.append((NAMES[k], analytic[k], fd, best))
    for name, a, fd, e in report:
        assert abs(a) > 10.0, (name, a)          # every quantity has a real gradient on this rig
        assert e < (1e-3 if smooth else 5e-2), report


def test_camera_fixed_point_matches_sums(oracle_lib):
    """camAdd's i32(clamp(v * 64)) truncates each contribution by less than
    one unit: the fixed-point totals stay within (contributions / 64) of the
    plain sums of the same f32 terms."""
    params, cam, tgt, o = _rig(True)
    st = O.Stage(o, "f32")
    cfix, csum, efix, esum, npx, nvis = _analytic(st, cam, params, tgt, fixed=True)
    assert np.all(np.abs(cfix / 64.0 - csum) <= nvis / 64.0 + 1e-9)
    assert np.all(np.abs(efix / 64.0 - esum) <= npx / 64.0 + 1e-9)
    assert np.all(np.abs(csum) > 1.0) and np.all(np.abs(esum) > 1.0)
    # the f32 twin is close to the double-precision restatement (relative to
    # the largest term: the small ones are sums of large cancelling parts)
    _, c64, _, e64, _, _ = _analytic(O.Stage(o, "f64"), cam, params, tgt, fixed=False)
    assert np.max(np.abs(csum - c64)) < 5e-3 * np.max(np.abs(c64))
    assert np.max(np.abs(esum - e64)) < 5e-3 * np.max(np.abs(e64))


def test_exposure_gradient_closed_form(oracle_lib):
    """One valid pixel, no splats: C = 0 (black), err = bias - target, so
    d/d(bias) = sum_c err_c / sqrt(err_c^2 + 0.03^2) and d/d(log gain) = 0."""
    cam = O.make_cam({"R": [1, 0, 0, 0, 1, 0, 0, 0, 1], "t": [0, 0, 0], "f": 50.0, "cx": 1.0, "cy": 1.0,
                      "w": 1, "h": 1, "g": 0.0, "b": 0.02})
    img = np.zeros((1, 4), np.float64)
    tgt = O.pack_targets(np.array([[0.2, 0.4, 0.8]], np.float32))
    st = O.Stage(O.make_opts(shDeg=0), "f64")
    _, _, _, efix, esum = st.loss_charb(cam, img, tgt, cam_grads=True)
    tc = np.array([round(0.2 * 255), round(0.4 * 255), round(0.8 * 255)]) / 255.0
    err = float(np.float32(0.02)) - tc
    want = np.sum(err / np.sqrt(err * err + 0.0009))
    assert esum[0] == 0.0 and abs(esum[1] - want) < 1e-12
    assert efix[1] == int(want * 64)


# ------------------------------------------------------------ camera Adam

def test_rodrigues_closed_forms():
    rng = np.random.default_rng(3)
    assert CO.rodrigues([0.0, 0.0, 0.0]) == [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0]
    for _ in range(20):
        w = rng.normal(size=3) * rng.uniform(0.01, 2.5)
        R = np.array(CO.rodrigues(list(w))).reshape(3, 3)
        th = np.linalg.norm(w)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14
        np.testing.assert_allclose(R @ w, w, atol=1e-14)                       # the axis is fixed
        assert abs((np.trace(R) - 1) / 2 - math.cos(th)) < 1e-14              # by angle |w|
        # right-handed: a vector orthogonal to the axis turns towards w x v
        v = np.cross(w, rng.normal(size=3))
        v /= np.linalg.norm(v)
        np.testing.assert_allclose(R @ v, math.cos(th) * v + math.sin(th) * np.cross(w / th, v), atol=1e-14)
    A = list(rng.normal(size=9))
    B = list(rng.normal(size=9))
    np.testing.assert_allclose(np.array(CO.m3mul(A, B)).reshape(3, 3),
                               np.array(A).reshape(3, 3) @ np.array(B).reshape(3, 3), atol=1e-14)


def _cams(n=4):
    return [{"R": CO.rodrigues([0.1 * i, 0.2, -0.1]), "t": [0.5 * i, 0.1, 2.0], "f": 300.0 + i,
             "cx": 160, "cy": 120, "w": 320, "h": 240, "g": 0.0, "b": 0.0} for i in range(n)]


def test_camera_adam_first_step_closed_form():
    """At camStep 1 Adam's update is lr * g / (|g| + eps): each pose slot moves
    by its learning rate against the gradient's sign. Camera 0 (the gauge
    anchor) and the holdout do not move; the shared focal moves all cameras;
    exposures move only with expComp (trainer.js:489-535)."""
    st = CO.CamState(_cams())
    before = [dict(m) for m in st.cams]
    g = np.zeros((5, 8), np.int32)
    g[:4] = [64, -64, 0, 128, 0, -64, 640, -64]
    g[4, 0] = 6400
    it, horizon, radius = 3000, 60000, 2.5
    st.apply(g, it, horizon, radius, holdout=2, exp_comp=False)
    decay = 0.02 ** min(1, it / (0.75 * horizon))
    rotLr, trnLr, focLr = 2e-4 * decay, 2e-4 * radius * decay, 1e-4 * decay
    for ci in (0, 2):
        assert st.cams[ci]["R"] == before[ci]["R"] and st.cams[ci]["t"] == before[ci]["t"]
    for ci in (1, 3):
        m, b = st.cams[ci], before[ci]
        Rw = np.array(CO.rodrigues([-rotLr, rotLr, 0.0])).reshape(3, 3)
        np.testing.assert_allclose(np.array(m["R"]).reshape(3, 3), Rw @ np.array(b["R"]).reshape(3, 3),
                                   rtol=0, atol=1e-15)
        np.testing.assert_allclose(m["t"], [b["t"][0] - trnLr, b["t"][1], b["t"][2] + trnLr], rtol=0, atol=1e-15)
        assert m["g"] == 0.0 and m["b"] == 0.0
    assert abs(st.logfScale + focLr) < 1e-18
    for m, b in zip(st.cams, before):
        assert abs(m["f"] - b["f0"] * math.exp(-focLr)) < 1e-12

    # expComp: gain and bias step too, and clamp at +-0.1 / +-0.05
    st = CO.CamState(_cams())
    for k in range(400):
        st.apply(g, it, horizon, radius, exp_comp=True)
    for ci in (1, 2, 3):
        assert st.cams[ci]["g"] == -0.1 and st.cams[ci]["b"] == 0.05
    assert st.cams[0]["g"] == 0.0
    # a constant gradient makes every Adam step exactly lr in size
    assert abs(st.logfScale + 400 * focLr) < 1e-12
    # the shared log-focal clamps at +-0.3 (lr 1e-4 at decay 1)
    st = CO.CamState(_cams())
    for k in range(3100):
        st.apply(g, 0, horizon, radius)
    assert st.logfScale == -0.3
    assert st.cams[0]["f"] == st.cams[0]["f0"] * math.exp(-0.3)


def test_camera_adam_zero_gradient_is_identity():
    st = CO.CamState(_cams())
    before = [dict(m) for m in st.cams]
    st.apply(np.zeros((5, 8), np.int32), 2000, 60000, 1.0, exp_comp=True)
    for m, b in zip(st.cams, before):
        assert m["t"] == b["t"] and m["f"] == b["f0"]
        np.testing.assert_allclose(m["R"], b["R"], rtol=0, atol=0)


def test_refine_pose_descends_on_a_quadratic():
    """evalCamPsnrRefined's Adam loop on a synthetic quadratic in the
    translation: the gradient 64 * (t - t*) steps t towards t*."""
    target = [0.3, -0.2, 0.1]

    def grad(R, t, g, b):
        return [0, 0, 0] + [int(64 * 64 * (t[k] - target[k])) for k in range(3)] + [0, 0]

    # translation lr 3e-4 * sceneRadius = 3e-3: 400 steps reach the target
    R, t, g, b = CO.refine_pose(CO.rodrigues([0.0, 0.0, 0.0]), [0.0, 0.0, 0.0], grad, 400, 10.0)
    np.testing.assert_allclose(t, target, atol=5e-3)
    assert R == [1.0, 0, 0, 0, 1.0, 0, 0, 0, 1.0] and g == 0.0 and b == 0.0


def test_wrap_i32():
    assert CO.wrap_i32([2 ** 31, -2 ** 31 - 1, 5]).tolist() == [-2 ** 31, 2 ** 31 - 1, 5]
