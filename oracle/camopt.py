"""CPU restatement of the reference's camera optimiser: the per-camera Adam
of GSTrainer._applyCamGrads (src/gs/trainer.js:476-537), the test-time pose
refinement of evalCamPsnrRefined (:583-612) and the rotation helpers it uses
(rodrigues, m3mul: src/sfm/geometry.js:108-114, :150-161), in JS-number
(double) arithmetic.

Test infrastructure only: tests/ use it for the expected results of the
trainer's camera optimiser. Math.sqrt / Math.cos / Math.sin / Math.exp / Math.pow
are taken from Python's math module, which calls the same libm as the
trainer's host code; V8's own implementations may differ from libm in the
last ulp (parity unpinned there).
"""
import math

import numpy as np

B1, B2, EPS = 0.9, 0.99, 1e-15


def m3mul(A, B):
    """geometry.js:108-114, row-major 3x3."""
    return [A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j] + A[i * 3 + 2] * B[6 + j]
            for i in range(3) for j in range(3)]


def rodrigues(w):
    """geometry.js:150-161: axis-angle -> rotation matrix."""
    th = math.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
    if th < 1e-12:
        return [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    k = [w[0] / th, w[1] / th, w[2] / th]
    c, s = math.cos(th), math.sin(th)
    v = 1 - c
    return [
        c + k[0] * k[0] * v, k[0] * k[1] * v - k[2] * s, k[0] * k[2] * v + k[1] * s,
        k[1] * k[0] * v + k[2] * s, c + k[1] * k[1] * v, k[1] * k[2] * v - k[0] * s,
        k[2] * k[0] * v - k[1] * s, k[2] * k[1] * v + k[0] * s, c + k[2] * k[2] * v,
    ]


def _clamp(x, lo, hi):
    return max(lo, min(hi, x))


def wrap_i32(x):
    """Two's-complement wrap of integer sums (the reference's atomic<i32>)."""
    x = np.asarray(x, np.int64)
    return (((x + 2 ** 31) % 2 ** 32) - 2 ** 31).astype(np.int32)


class CamState:
    """The reference's camMeta entries plus the camera optimiser state
    (trainer.js:306-313): double R, t, f0, f, g, b per camera; camM / camV
    (numCams + 1) x 8; camStep; logfScale."""

    def __init__(self, cams):
        self.cams = []
        for c in cams:
            f = float(np.float32(c["f"]))
            self.cams.append({"R": [float(np.float32(x)) for x in c["R"]],
                              "t": [float(np.float32(x)) for x in c["t"]],
                              "f0": f, "f": f,
                              "g": float(np.float32(c.get("g", 0.0))), "b": float(np.float32(c.get("b", 0.0))),
                              "cx": c["cx"], "cy": c["cy"], "w": c["w"], "h": c["h"]})
        nr = len(cams) + 1
        self.camM = [0.0] * (nr * 8)
        self.camV = [0.0] * (nr * 8)
        self.camStep = 0
        self.logfScale = 0.0

    def cam_dict(self, ci, override=None):
        """Camera fields for oracle.make_cam (doubles; make_cam stores f32)."""
        m = dict(self.cams[ci])
        if override:
            m.update(override)
        return m

    def apply(self, g, it, horizon, scene_radius, holdout=-1, exp_comp=False):
        """_applyCamGrads (trainer.js:489-535) on gradient rows g
        ((numCams + 1) x 8 int32, x64 fixed point)."""
        g = np.asarray(g, np.int32).reshape(-1)
        self.camStep += 1
        t = self.camStep
        decay = math.pow(0.02, min(1, it / (0.75 * horizon)))
        rotLr = 2e-4 * decay
        trnLr = 2e-4 * scene_radius * decay
        focLr = 1e-4 * decay
        M, V = self.camM, self.camV

        def step(row, slot, lr):
            j = row * 8 + slot
            grad = int(g[j]) / 64
            M[j] = B1 * M[j] + (1 - B1) * grad
            V[j] = B2 * V[j] + (1 - B2) * grad * grad
            mh = M[j] / (1 - math.pow(B1, t))
            vh = V[j] / (1 - math.pow(B2, t))
            return lr * mh / (math.sqrt(vh) + EPS)

        expLr = 5e-3
        nc = len(self.cams)
        for r in range(1, nc):          # camera 0 pinned (gauge + exposure anchor)
            if r == holdout:
                continue
            m = self.cams[r]
            dw = [-step(r, 0, rotLr), -step(r, 1, rotLr), -step(r, 2, rotLr)]
            m["R"] = m3mul(rodrigues(dw), m["R"])
            m["t"] = [m["t"][0] - step(r, 3, trnLr), m["t"][1] - step(r, 4, trnLr), m["t"][2] - step(r, 5, trnLr)]
            if exp_comp:
                m["g"] = _clamp(m["g"] - step(r, 6, expLr), -0.1, 0.1)
                m["b"] = _clamp(m["b"] - step(r, 7, expLr), -0.05, 0.05)
        self.logfScale = _clamp(self.logfScale - step(nc, 0, focLr), -0.3, 0.3)
        for m in self.cams:
            m["f"] = m["f0"] * math.exp(self.logfScale)


def refine_pose(R, t, eval_grad, steps, scene_radius, exp_comp=False):
    """evalCamPsnrRefined's loop (trainer.js:583-609). eval_grad(R, t, g, b)
    returns the camera's 8 int32 gradient slots of one evaluation pass.
    Returns the refined (R, t, g, b)."""
    R, t = list(R), list(t)
    g = b = 0.0
    m, v = [0.0] * 8, [0.0] * 8
    rotLr, trnLr, expLr = 3e-4, 3e-4 * scene_radius, 5e-3
    for s in range(1, steps + 1):
        cg = eval_grad(R, t, g, b)
        upd = [0.0] * 8
        for k in range(8):
            grad = int(cg[k]) / 64
            m[k] = B1 * m[k] + (1 - B1) * grad
            v[k] = B2 * v[k] + (1 - B2) * grad * grad
            mh = m[k] / (1 - math.pow(B1, s))
            vh = v[k] / (1 - math.pow(B2, s))
            upd[k] = (rotLr if k < 3 else (trnLr if k < 6 else expLr)) * mh / (math.sqrt(vh) + EPS)
        R = m3mul(rodrigues([-upd[0], -upd[1], -upd[2]]), R)
        t = [t[0] - upd[3], t[1] - upd[4], t[2] - upd[5]]
        if exp_comp:
            g = _clamp(g - upd[6], -0.1, 0.1)
            b = _clamp(b - upd[7], -0.05, 0.05)
    return R, t, g, b
