"""ctypes wrapper over liboracle.so, the CPU restatement of the splat.js
trainer step.

TEST INFRASTRUCTURE ONLY. Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, for expected results. The product never
imports it.

Each stage follows src/gs/shaders.js (see the C sources for per-function
line citations). `prec` selects the single-precision twin ("f32", bit-exact
target for the HIP kernels) or the double-precision reference ("f64").
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


class OrcCam(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("t", C.c_float * 3), ("nearp", C.c_float),
                ("f", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("gain", C.c_float), ("bias", C.c_float), ("W", C.c_int), ("H", C.c_int),
                ("bg", C.c_float * 3), ("shActiveDeg", C.c_int), ("camPos", C.c_float * 3),
                ("tile", C.c_int)]


class OrcOpts(C.Structure):
    _fields_ = [("eCut", C.c_float), ("aMin", C.c_float), ("radClamp", C.c_float),
                ("anisoReg", C.c_float), ("shDeg", C.c_int)]


class OrcAdam(C.Structure):
    _fields_ = [("lr", C.c_float * 16), ("b1", C.c_float), ("b2", C.c_float), ("eps", C.c_float),
                ("rbc1", C.c_float), ("rbc2", C.c_float), ("minLS", C.c_float), ("maxLS", C.c_float),
                ("maxLogit", C.c_float), ("reg", C.c_float), ("shLr", C.c_float)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        for name, args, res in [
            ("orc_cam_pos", [P], None),
            ("orc_expf", [C.c_float], C.c_float),
            ("orc_expf_raster", [C.c_float], C.c_float),
            ("orc_logf", [C.c_float], C.c_float),
            ("orc_count_keys", [C.c_int64, P], C.c_int64),
            ("orc_sh_basis_f64", [C.c_int, P, P, P], None),
            ("orc_pack_targets", [C.c_int64, P, P], None),
            ("orc_adam_f32", [P, C.c_int64, P, P, P, P], None),
            ("orc_sh_adam_f32", [P, C.c_int64, P, P, P, P], None),
            ("orc_loss_l1ssim_f64", [P, P, P, C.c_double, P, P, P], None),
            ("orc_adam_mixed_f32", [P, C.c_int64, P, P, P, P, C.c_uint32], None),
            ("orc_sh_adam_mixed_f32", [P, C.c_int64, C.c_int, P, P, P, P, C.c_uint32], None),
            ("orc_set_threads", [C.c_int], None),
            ("orc_sh_adam_h16_f32", [P, C.c_int64, C.c_int, P, P, P, P, C.c_uint32, C.c_int], None),
            ("orc_f2h_rne", [C.c_int64, P, P], None),
            ("orc_f2h_sr", [C.c_int64, P, P, P], None),
            ("orc_f2h_up", [C.c_int64, P, P], None),
            ("orc_h2f", [C.c_int64, P, P], None),
            ("orc_f2bf_rne", [C.c_int64, P, P], None),
            ("orc_f2bf_sr", [C.c_int64, P, P, P], None),
            ("orc_bf2f", [C.c_int64, P, P], None),
            ("orc_sr_bits", [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32], C.c_uint32),
        ]:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        for s in ("f32", "f64"):
            for name, args, res in [
                ("orc_project_", [P, P, C.c_int64, P, P, P, P, P], None),
                ("orc_bin_sort_", [C.c_int64, P, P, P, C.c_int, C.c_int, P, P, P], C.c_int64),
                ("orc_render_fwd_", [P, P, P, P, P, P, P, P], None),
                ("orc_loss_charb_", [P, P, P, P, P, P, P, P], None),
                ("orc_render_bwd_", [P, P, P, P, P, P, P, P, P, P, P], None),
                ("orc_chain_", [P, P, C.c_int64, P, P, P, P, P, P, P, P, P], None),
            ]:
                fn = getattr(L, name + s)
                fn.argtypes = args
                fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def make_cam(cam, shActiveDeg=0, bg=(0.0, 0.0, 0.0), tile=16):
    """cam: dict with R (9), t (3), f, cx, cy, w, h, optional g (log gain), b.
    tile: tile side in pixels (the reference's 16, or 8 or 32)."""
    c = OrcCam()
    for k in range(9):
        c.R[k] = float(cam["R"][k])
    for k in range(3):
        c.t[k] = float(cam["t"][k])
        c.bg[k] = float(bg[k])
    c.nearp = 0.05
    c.f, c.cx, c.cy = float(cam["f"]), float(cam["cx"]), float(cam["cy"])
    c.gain = float(np.float32(math.exp(cam.get("g", 0.0))))   # libm exp, as the trainer's host code
    c.bias = float(cam.get("b", 0.0))
    c.W, c.H = int(cam["w"]), int(cam["h"])
    c.shActiveDeg = int(shActiveDeg)
    c.tile = int(tile)
    lib().orc_cam_pos(C.byref(c))
    return c


def make_opts(eCut=4.5, aMin=0.0039, radClamp=1.0, anisoReg=0.005, shDeg=2):
    o = OrcOpts()
    o.eCut, o.aMin, o.radClamp, o.anisoReg, o.shDeg = eCut, aMin, radClamp, anisoReg, shDeg
    return o


def sh_k(deg):
    return (deg + 1) * (deg + 1) - 1 if deg > 0 else 0


class Stage:
    """Runs the reference step stage by stage on one camera."""

    def __init__(self, opts, prec="f32"):
        self.o = opts
        self.prec = prec
        self.dt = np.float32 if prec == "f32" else np.float64
        self.L = lib()

    def fn(self, name):
        return getattr(self.L, name + self.prec)

    def project(self, cam, params, sh):
        n = params.shape[0]
        params = np.ascontiguousarray(params, np.float32)
        sh = None if sh is None else np.ascontiguousarray(sh, np.float32)
        proj = np.zeros((n, 16), self.dt)
        rect = np.zeros((n, 4), np.int32)
        cnt = np.zeros(n, np.uint32)
        self.fn("orc_project_")(C.byref(self.o), C.byref(cam), n, _p(params), _p(sh),
                                _p(proj), _p(rect), _p(cnt))
        return proj, rect, cnt

    def bin_sort(self, cam, proj, rect, cnt):
        n = proj.shape[0]
        T = cam.tile or 16
        tilesX = (cam.W + T - 1) // T
        numTiles = tilesX * ((cam.H + T - 1) // T)
        K = int(self.L.orc_count_keys(n, _p(cnt)))
        ids = np.zeros(max(K, 1), np.uint32)
        keys = np.zeros(max(K, 1), np.uint64)
        ts = np.zeros(numTiles + 1, np.uint32)
        self.fn("orc_bin_sort_")(n, _p(proj), _p(rect), _p(cnt), tilesX, numTiles,
                                 _p(ids), _p(ts), _p(keys))
        return ids[:K], ts, keys[:K]

    def render_fwd(self, cam, proj, ids, ts):
        npx = cam.W * cam.H
        img = np.zeros((npx, 4), self.dt)
        T = np.zeros(npx, self.dt)
        end = np.zeros(npx, np.uint32)
        ids = np.ascontiguousarray(ids if len(ids) else np.zeros(1, np.uint32))
        self.fn("orc_render_fwd_")(C.byref(self.o), C.byref(cam), _p(proj), _p(ids), _p(ts),
                                   _p(img), _p(T), _p(end))
        return img, T, end

    def loss_charb(self, cam, img, tgt, cam_grads=False):
        """cam_grads: also return the exposure gradients (log gain, bias) as
        x64 fixed-point wrapping i32 sums and as plain f64 sums."""
        npx = cam.W * cam.H
        gC = np.zeros((npx, 3), self.dt)
        valid = np.zeros(npx, np.uint8)
        stats = np.zeros(3, np.float64)
        efix = np.zeros(2, np.int32) if cam_grads else None
        esum = np.zeros(2, np.float64) if cam_grads else None
        self.fn("orc_loss_charb_")(C.byref(cam), _p(img), _p(np.ascontiguousarray(tgt, np.uint32)),
                                   _p(gC), _p(valid), _p(stats), _p(efix), _p(esum))
        if cam_grads:
            return gC, valid, stats, efix, esum
        return gC, valid, stats

    def loss_l1ssim(self, cam, img, tgt, lam=0.2):
        assert self.prec == "f64"
        npx = cam.W * cam.H
        gC = np.zeros((npx, 3), np.float64)
        valid = np.zeros(npx, np.uint8)
        stats = np.zeros(3, np.float64)
        self.L.orc_loss_l1ssim_f64(C.byref(cam), _p(np.ascontiguousarray(img, np.float64)),
                                   _p(np.ascontiguousarray(tgt, np.uint32)), lam,
                                   _p(gC), _p(valid), _p(stats))
        return gC, valid, stats

    def render_bwd(self, cam, proj, ids, ts, T, end, gC, valid, fixed=True):
        n = proj.shape[0]
        gfix = np.zeros((n, 10), np.int32) if fixed else None
        gsum = None if fixed else np.zeros((n, 10), np.float64)
        ids = np.ascontiguousarray(ids if len(ids) else np.zeros(1, np.uint32))
        self.fn("orc_render_bwd_")(C.byref(self.o), C.byref(cam), _p(proj), _p(ids), _p(ts),
                                   _p(T), _p(end), _p(np.ascontiguousarray(gC, self.dt)),
                                   _p(valid), _p(gfix), _p(gsum))
        return gfix if fixed else gsum

    def chain(self, cam, params, sh, proj, gfix=None, gsum=None, cam_grads=False):
        """cam_grads: also return the camera gradients (rotation 3,
        translation 3, log-focal 1) as x64 fixed-point wrapping i32 sums and
        as plain f64 sums."""
        n = params.shape[0]
        K = sh_k(self.o.shDeg)
        gradF = np.zeros((n, 16), self.dt)
        shGrad = np.zeros((n, max(3 * K, 1)), self.dt)
        cfix = np.zeros(7, np.int32) if cam_grads else None
        csum = np.zeros(7, np.float64) if cam_grads else None
        self.fn("orc_chain_")(C.byref(self.o), C.byref(cam), n,
                              _p(np.ascontiguousarray(params, np.float32)),
                              _p(None if sh is None else np.ascontiguousarray(sh, np.float32)),
                              _p(proj), _p(gfix), _p(gsum), _p(gradF), _p(shGrad), _p(cfix), _p(csum))
        if cam_grads:
            return gradF, shGrad[:, :3 * K], cfix, csum
        return gradF, shGrad[:, :3 * K]


SHARED_SORT = 2048     # shaders.js:33
SCAN_CHUNK = 64        # shaders.js:325


def padded_size(c):
    """shaders.js:329-332: segments over SHARED_SORT take a power-of-two slot."""
    c = int(c)
    return c if c <= SHARED_SORT else 1 << (c - 1).bit_length()


def tile_budget(tile_cnt, cap):
    """The reference's budgeted segment scan (shaders.js:321-370): thread li
    of 256 owns tiles [64 li, 64 li + 64); its chunk starts at the exclusive
    prefix of the padded sizes of earlier chunks (no drops applied); inside
    the chunk a tile whose padded size takes the running total past the
    budget is dropped (counted even when empty). Returns (kept count per
    tile, dropped tiles)."""
    cnt = [int(c) for c in tile_cnt]
    T = len(cnt)
    nch = (T + SCAN_CHUNK - 1) // SCAN_CHUNK
    sums = [sum(padded_size(c) for c in cnt[k * SCAN_CHUNK:(k + 1) * SCAN_CHUNK]) for k in range(nch)]
    kept = [0] * T
    dropped = 0
    pre = 0
    for k in range(nch):
        acc = pre
        for t in range(k * SCAN_CHUNK, min(T, (k + 1) * SCAN_CHUNK)):
            p = padded_size(cnt[t])
            if acc + p > cap:
                p = 0
                dropped += 1
            acc += p
            kept[t] = cnt[t] if p else 0
        pre += sums[k]
    return np.array(kept, np.int64), dropped


def apply_tile_budget(ids, keys, ts, cap):
    """A bin_sort result with the budget applied: the keys of dropped tiles
    removed, tile starts recomputed. Returns (ids, keys, ts, dropped)."""
    cnt = np.diff(ts.astype(np.int64))
    kept, dropped = tile_budget(cnt, cap)
    keep = np.repeat(kept == cnt, cnt)
    ts2 = np.zeros(len(ts), np.uint32)
    ts2[1:] = np.cumsum(kept)
    return ids[keep], keys[keep], ts2, dropped


def pack_targets(rgb):
    """float RGB (npx x 3, r < 0 = invalid) -> packed RGBA8, as trainer.js:128-139."""
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros(rgb.shape[0], np.uint32)
    lib().orc_pack_targets(rgb.shape[0], _p(rgb), _p(out))
    return out


def adam_hp(lr, b1=0.9, b2=0.999, eps=1e-15, t=1, minLS=-20.0, maxLS=2.0, maxLogit=8.0,
            reg=0.01, shLr=1.25e-4):
    """Bias corrections are computed in double and rounded to f32 once,
    exactly as the product's host code does."""
    h = OrcAdam()
    for k in range(16):
        h.lr[k] = float(lr[k])
    b1f, b2f = float(np.float32(b1)), float(np.float32(b2))
    h.b1, h.b2, h.eps = b1f, b2f, eps
    h.rbc1 = float(np.float32(1.0 / (1.0 - b1f ** t)))
    h.rbc2 = float(np.float32(1.0 / (1.0 - b2f ** t)))
    h.minLS, h.maxLS, h.maxLogit, h.reg, h.shLr = minLS, maxLS, maxLogit, reg, shLr
    return h


def adam(hp, params, grad, m, v):
    n = params.shape[0]
    lib().orc_adam_f32(C.byref(hp), n, _p(params), _p(np.ascontiguousarray(grad, np.float32)),
                       _p(m), _p(v))


def sh_adam(hp, sh, grad, m, v):
    lib().orc_sh_adam_f32(C.byref(hp), sh.size, _p(sh), _p(np.ascontiguousarray(grad, np.float32)),
                          _p(m), _p(v))


def adam_mixed(hp, params, grad, m16, v16, step):
    """Adam with bf16 moments (uint16 arrays, n x 16), stochastic rounding."""
    n = params.shape[0]
    lib().orc_adam_mixed_f32(C.byref(hp), n, _p(params), _p(np.ascontiguousarray(grad, np.float32)),
                             _p(m16), _p(v16), step)


def sh_adam_mixed(hp, sh16, grad, m16, v16, step):
    """SH Adam with bf16 coefficients and moments (uint16 arrays, n x 3K)."""
    n, nc = sh16.shape
    lib().orc_sh_adam_mixed_f32(C.byref(hp), n, nc, _p(sh16), _p(np.ascontiguousarray(grad, np.float32)),
                                _p(m16), _p(v16), step)


def set_threads(n):
    """Threads of the restatement's per-Gaussian and per-pixel loops (OpenMP).
    1, the default, keeps them serial, as the tests expect; above 1 the
    fixed-point sums stay exact and double sums change order."""
    lib().orc_set_threads(int(n))


def sh_adam_h16(hp, sh16, grad, m, v, step, fmt):
    """SH Adam with 16-bit coefficients (uint16, n x 3K; fmt 1 bf16, 2 fp16)
    and f32 moments, stochastic rounding of the coefficient."""
    n, nc = sh16.shape
    lib().orc_sh_adam_h16_f32(C.byref(hp), n, nc, _p(sh16), _p(np.ascontiguousarray(grad, np.float32)),
                              _p(m), _p(v), step, fmt)


def f2h_rne(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f2h_rne(x.size, _p(x), _p(out))
    return out


def f2h_sr(x, rnd):
    x = np.ascontiguousarray(x, np.float32)
    rnd = np.ascontiguousarray(rnd, np.uint32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f2h_sr(x.size, _p(x), _p(rnd), _p(out))
    return out


def f2h_up(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f2h_up(x.size, _p(x), _p(out))
    return out


def h2f(b):
    b = np.ascontiguousarray(b, np.uint16)
    out = np.zeros(b.shape, np.float32)
    lib().orc_h2f(b.size, _p(b), _p(out))
    return out


def f2bf_rne(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f2bf_rne(x.size, _p(x), _p(out))
    return out


def f2bf_sr(x, rnd):
    x = np.ascontiguousarray(x, np.float32)
    rnd = np.ascontiguousarray(rnd, np.uint32)
    out = np.zeros(x.shape, np.uint16)
    lib().orc_f2bf_sr(x.size, _p(x), _p(rnd), _p(out))
    return out


def bf2f(b):
    b = np.ascontiguousarray(b, np.uint16)
    out = np.zeros(b.shape, np.float32)
    lib().orc_bf2f(b.size, _p(b), _p(out))
    return out


def loss_of(stage, cam, params, sh, tgt):
    """Training loss (Charbonnier sum) of one forward pass."""
    proj, rect, cnt = stage.project(cam, params, sh)
    ids, ts, _ = stage.bin_sort(cam, proj, rect, cnt)
    img, T, end = stage.render_fwd(cam, proj, ids, ts)
    _, _, stats = stage.loss_charb(cam, img, tgt)
    return stats[1]


def analytic_grads(stage, cam, params, sh, tgt, fixed=False):
    proj, rect, cnt = stage.project(cam, params, sh)
    ids, ts, _ = stage.bin_sort(cam, proj, rect, cnt)
    img, T, end = stage.render_fwd(cam, proj, ids, ts)
    gC, valid, stats = stage.loss_charb(cam, img, tgt)
    g = stage.render_bwd(cam, proj, ids, ts, T, end, gC, valid, fixed=fixed)
    if fixed:
        return stage.chain(cam, params, sh, proj, gfix=g)
    return stage.chain(cam, params, sh, proj, gsum=g)
