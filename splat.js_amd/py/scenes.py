"""Seeded synthetic scenes for the trainer step.

* `gradcheck_rig` restates the reference's finite-difference rig
  (src/gs/gradcheck.js:93-130): 160 splats from the LCG
  s = s*1664525 + 1013904223 (seed 123456789), a 64x64 sinusoidal target,
  one f=60 camera at the origin, strict cutoffs.
* `sh_rig_coeffs` restates gradCheckSH's SH coefficients (gradcheck.js:198-201).
* `gaussians` draws the distributions the benchmark configurations name:
  means uniform in a box, log-scales N(mu, sd), uniformly random unit
  quaternions, opacity logits N(0, 1), DC colour U(0,1) stored as a logit
  (the reference keeps DC as a sigmoid logit, shaders.js:19-22), SH rest
  N(0, 0.05) channel-major.
* `ring_cameras` places cameras on a circle looking at the origin with the
  reference's y-down look-at convention (src/synthetic.js:276-291).

Random streams come from numpy's PCG64 with stated seeds.
"""
import math

import numpy as np

STRIDE = 16


def sh_k(deg):
    return (deg + 1) * (deg + 1) - 1 if deg > 0 else 0


class _Lcg:
    """gradcheck.js:97: s = (imul(s, 1664525) + 1013904223) >>> 0; s / 2^32."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFF

    def __call__(self):
        self.s = (self.s * 1664525 + 1013904223) & 0xFFFFFFFF
        return self.s / 4294967296.0


def gradcheck_rig():
    """Returns (params n x 16 f32, cams, images, W, H, radius, trainer opts)."""
    n = 160
    rnd = _Lcg(123456789)
    data = np.zeros((n, STRIDE), np.float32)
    for i in range(n):
        d = data[i]
        d[0] = (rnd() - 0.5) * 2.4
        d[1] = (rnd() - 0.5) * 2.4
        d[2] = 2 + rnd() * 2
        d[3] = math.log(0.04 + rnd() * 0.1)
        d[4] = math.log(0.04 + rnd() * 0.1)
        d[5] = math.log(0.04 + rnd() * 0.1)
        d[6] = rnd() * 2 - 1
        d[7] = rnd() * 2 - 1
        d[8] = rnd() * 2 - 1
        d[9] = rnd() * 2 - 1
        d[10] = (rnd() - 0.5) * 4
        d[11] = (rnd() - 0.5) * 4
        d[12] = (rnd() - 0.5) * 4
        d[13] = math.log(0.4 / 0.6) + rnd() * 2 - 1
    W = H = 64
    yy, xx = np.meshgrid(np.arange(H, dtype=np.float64), np.arange(W, dtype=np.float64), indexing="ij")
    rgb = np.zeros((H, W, 3), np.float32)
    rgb[..., 0] = 0.5 + 0.4 * np.sin(xx * 0.31) * np.cos(yy * 0.17)
    rgb[..., 1] = 0.5 + 0.4 * np.sin(xx * 0.11 + 1) * np.cos(yy * 0.23 + 2)
    rgb[..., 2] = 0.5 + 0.4 * np.sin((xx + yy) * 0.19)
    images = [{"tw": W, "th": H, "rgb": rgb.reshape(-1)}]
    cams = [{"imgIdx": 0, "R": [1, 0, 0, 0, 1, 0, 0, 0, 1], "t": [0, 0, 0], "f": 60,
             "cx": W / 2, "cy": H / 2, "w": W, "h": H}]
    opts = {"eCut": 9.0, "aMin": 1e-4, "radClamp": 10.0, "anisoReg": 0.0}
    return data, cams, images, W, H, 1.5, opts


def sh_rig_coeffs(n, deg):
    """gradcheck.js:198-201: (rnd() - 0.5) * 0.5 from LCG seed 424242."""
    K = sh_k(deg)
    rnd = _Lcg(424242)
    out = np.zeros(n * 3 * K, np.float32)
    for i in range(out.size):
        out[i] = (rnd() - 0.5) * 0.5
    return out.reshape(n, 3 * K)


def _logit(u):
    u = np.clip(u, 1e-3, 1 - 1e-3)
    return np.log(u / (1 - u))


def gaussians(n, seed, box_lo, box_hi, ls_mu=-4.0, ls_sd=0.5, sh_deg=2):
    """Returns (params n x 16 f32, sh n x 3K f32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros((n, STRIDE), np.float32)
    lo = np.asarray(box_lo, np.float64)
    hi = np.asarray(box_hi, np.float64)
    p[:, 0:3] = lo + (hi - lo) * rng.random((n, 3))
    p[:, 3:6] = rng.normal(ls_mu, ls_sd, (n, 3))
    q = rng.normal(0.0, 1.0, (n, 4))
    q /= np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    p[:, 6:10] = q
    p[:, 10:13] = _logit(rng.random((n, 3)))
    p[:, 13] = rng.normal(0.0, 1.0, n)
    K = sh_k(sh_deg)
    sh = rng.normal(0.0, 0.05, (n, 3 * K)).astype(np.float32) if K else None
    return p, sh


def look_at(eye, center=(0.0, 0.0, 0.0)):
    """synthetic.js:276-291: world up is -y; rows of R are the camera axes."""
    eye = np.asarray(eye, np.float64)
    c = np.asarray(center, np.float64)
    up = np.array([0.0, -1.0, 0.0])
    zc = c - eye
    zc /= np.linalg.norm(zc)
    xc = np.cross(zc, up)
    xc /= np.linalg.norm(xc)
    yc = np.cross(zc, xc)
    R = np.stack([xc, yc, zc])
    t = -R @ eye
    return [float(v) for v in R.reshape(-1)], [float(v) for v in t]


def ring_cameras(count, radius, W, H, f, height=0.0):
    cams = []
    for k in range(count):
        a = 2 * math.pi * k / count
        eye = (radius * math.sin(a), height, -radius * math.cos(a))
        R, t = look_at(eye)
        cams.append({"imgIdx": k, "R": R, "t": t, "f": f, "cx": W / 2, "cy": H / 2, "w": W, "h": H})
    return cams


def config1(seed=0):
    """20k Gaussians, one 320x240 f=300 camera at the origin (identity pose)."""
    p, sh = gaussians(20000, seed, (-0.5, -0.5, 2.0), (0.5, 0.5, 6.0))
    cam = {"imgIdx": 0, "R": [1, 0, 0, 0, 1, 0, 0, 0, 1], "t": [0, 0, 0], "f": 300,
           "cx": 160, "cy": 120, "w": 320, "h": 240}
    return p, sh, [cam]


def config2(n=1_000_000, seed=0, views=64, W=1600, H=900, f=1100.0, radius=8.0):
    """1M Gaussians in a 10x10x10 box, ring of views at radius 8, 1600x900."""
    p, sh = gaussians(n, seed, (-5, -5, -5), (5, 5, 5))
    return p, sh, ring_cameras(views, radius, W, H, f)


# config 4: eight views of different sizes in one trainer
CONFIG4_SIZES = [(800, 450), (960, 540), (1024, 576), (1280, 720), (1366, 768), (1600, 900), (1760, 990),
                 (1920, 1080)]


def config4(n=1_000_000, seed=0, radius=8.0, sizes=None):
    """Config 2's Gaussians seen from 8 ring views of sizes 800x450 .. 1920x1080
    (focal 1100 * W / 1600, principal point at the centre)."""
    p, sh = gaussians(n, seed, (-5, -5, -5), (5, 5, 5))
    sizes = sizes or CONFIG4_SIZES
    cams = []
    for k, (W, H) in enumerate(sizes):
        a = 2 * math.pi * k / len(sizes)
        R, t = look_at((radius * math.sin(a), 0.0, -radius * math.cos(a)))
        cams.append({"imgIdx": k, "R": R, "t": t, "f": 1100.0 * W / 1600.0, "cx": W / 2, "cy": H / 2,
                     "w": W, "h": H})
    return p, sh, cams


def config5(n=3_000_000, seed=0, views=4, W=1920, H=1080, radius=8.0):
    """Stress: 3M Gaussians, SH degree 3 (45 coefficients), log-scales
    N(-3, 1) (large splats, ~250M keys per view), 4 views at 1920x1080."""
    p, sh = gaussians(n, seed, (-5, -5, -5), (5, 5, 5), ls_mu=-3.0, ls_sd=1.0, sh_deg=3)
    return p, sh, ring_cameras(views, radius, W, H, 1100.0 * W / 1600.0)
