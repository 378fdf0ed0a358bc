"""Python binding of the C ABI (include/splat_gs.h) with the reference's
GSTrainer method surface (src/gs/trainer.js): create / setup / stepOnce /
readLoss / renderView / readGaussians / evalCamPsnr, plus the inspection
entry points the tests and the benchmark use.

The library must be built (splat.js_amd/build.py); loading fails loudly
otherwise. There is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(HERE), "libsplatgs.so")

GS_LOSS_CHARBONNIER, GS_LOSS_L1_DSSIM = 0, 1
GS_ACCUM_FIXED, GS_ACCUM_F32 = 0, 1
GS_SORT_DEPTH_THEN_TILE, GS_SORT_GLOBAL64, GS_SORT_TILE_SEGMENTS = 0, 1, 2
GS_PREC_F32, GS_PREC_MIXED, GS_PREC_BF16, GS_PREC_FP16 = 0, 1, 2, 3
GS_VIEWS_PERSISTENT, GS_VIEWS_OVERLAP = 0, 1


class GsOpts(C.Structure):
    _fields_ = [("size", C.c_int32), ("eCut", C.c_float), ("aMin", C.c_float), ("radClamp", C.c_float),
                ("shDeg", C.c_int32), ("anisoReg", C.c_float), ("opacityReg", C.c_float),
                ("shLr", C.c_float), ("minScale", C.c_float), ("maxIters", C.c_int32),
                ("maxSplats", C.c_int64), ("capMult", C.c_float), ("entriesCap", C.c_int64),
                ("tileGrad", C.c_int32), ("posLr", C.c_float), ("loss", C.c_int32),
                ("ssimLambda", C.c_float), ("gradAccum", C.c_int32), ("sortMode", C.c_int32),
                ("seed", C.c_uint32), ("device", C.c_int32), ("precision", C.c_int32),
                ("camOpt", C.c_int32), ("expComp", C.c_int32), ("viewsPerStep", C.c_int32),
                ("graph", C.c_int32), ("viewsLaunch", C.c_int32), ("tile", C.c_int32), ("batch", C.c_int32)]


class GsCamera(C.Structure):
    _fields_ = [("R", C.c_float * 9), ("t", C.c_float * 3), ("f", C.c_float), ("cx", C.c_float),
                ("cy", C.c_float), ("w", C.c_int32), ("h", C.c_int32), ("imgIdx", C.c_int32),
                ("g", C.c_float), ("b", C.c_float)]


class GsCamState(C.Structure):
    _fields_ = [("R", C.c_double * 9), ("t", C.c_double * 3), ("f", C.c_double), ("g", C.c_double),
                ("b", C.c_double)]


class GsImage(C.Structure):
    _fields_ = [("tw", C.c_int32), ("th", C.c_int32), ("rgb", C.c_void_p)]


class GsInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("cap", C.c_int64), ("iter", C.c_int64), ("shDeg", C.c_int32),
                ("shK", C.c_int32), ("stride", C.c_int32), ("numCams", C.c_int32),
                ("lastCam", C.c_int32), ("holdout", C.c_int32), ("entriesCap", C.c_int64), ("maxTiles", C.c_int32),
                ("horizon", C.c_float), ("tile", C.c_int32), ("batch", C.c_int32)]


class GsStepStats(C.Structure):
    _fields_ = [("lastKeys", C.c_int64), ("lastVisible", C.c_int64), ("overflowKeys", C.c_int64)]


class GsReadbackResult(C.Structure):
    _fields_ = [("value", C.c_double), ("has", C.c_int32), ("overflow", C.c_int64)]


class GsDebugOut(C.Structure):
    _fields_ = [(name, C.c_void_p) for name in
                ("rec", "depthKey", "rect", "cnt", "ids", "keys", "numKeys", "tileRanges", "img", "T",
                 "end", "gC", "accum", "gradF", "shGrad", "stats")]


GS_TUNE_MEASURED, GS_TUNE_TOO_MANY_TILES, GS_TUNE_MORE_DROPS = 0, 1, 2


class GsTuneEntry(C.Structure):
    _fields_ = [("tile", C.c_int32), ("batch", C.c_int32), ("status", C.c_int32), ("pad", C.c_int32),
                ("ms", C.c_double), ("keys", C.c_int64), ("dropped", C.c_int64)]


class GsKernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("total_ms", C.c_double), ("launches", C.c_int64),
                ("bytes", C.c_double)]


# every function the header declares, with its ctypes signature
SIGNATURES = {
    "gs_default_opts": ([C.c_void_p], C.c_int),
    "gs_create": ([C.c_void_p, C.c_void_p], C.c_int),
    "gs_destroy": ([C.c_void_p], None),
    "gs_setup": ([C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                  C.c_int32, C.c_int32, C.c_int32, C.c_float], C.c_int),
    "gs_step": ([C.c_void_p, C.c_int32], C.c_int),
    "gs_read_loss": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "gs_render": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p], C.c_int),
    "gs_read_gaussians": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "gs_eval_cam_psnr": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_eval_cam_psnr_refined": ([C.c_void_p, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
    "gs_get_camera": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_read_cam_grads": ([C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "gs_set_holdout": ([C.c_void_p, C.c_int32], C.c_int),
    "gs_set_excluded": ([C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "gs_set_iter": ([C.c_void_p, C.c_int64], C.c_int),
    "gs_get_info": ([C.c_void_p, C.c_void_p], C.c_int),
    "gs_read_step_stats": ([C.c_void_p, C.c_void_p], C.c_int),
    "gs_read_loss_async": ([C.c_void_p, C.c_void_p], C.c_int),
    "gs_eval_cam_psnr_async": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_read_gaussians_async": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "gs_pending_wait": ([C.c_void_p, C.c_void_p], C.c_int),
    "gs_pending_error": ([C.c_void_p], C.c_char_p),
    "gs_pending_free": ([C.c_void_p], None),
    "gs_sync": ([C.c_void_p], C.c_int),
    "gs_refine": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "gs_step_batch": ([C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "gs_step_views": ([C.c_void_p, C.c_void_p, C.c_int32], C.c_int),
    "gs_fd_grad64": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_step_views_debug": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_export_ply": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "gs_export_ply_noalloc": ([C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p], C.c_int),
    "gs_last_error": ([C.c_void_p], C.c_char_p),
    "gs_step_debug": ([C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_profile": ([C.c_void_p, C.c_int32], C.c_int),
    "gs_profile_read": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p], C.c_int),
    "gs_stream": ([C.c_void_p], C.c_void_p),
    "gs_set_raster_shape": ([C.c_void_p, C.c_int32, C.c_int32], C.c_int),
    "gs_autotune": ([C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p], C.c_int),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python splat.js_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class GsError(RuntimeError):
    pass


class GsRefineOpts(C.Structure):
    _fields_ = [("size", C.c_int32), ("seed", C.c_uint32), ("growRate", C.c_float), ("growUntil", C.c_double),
                ("growLimit", C.c_int64)]


class GsRefineResult(C.Structure):
    _fields_ = [("moved", C.c_int64), ("grown", C.c_int64), ("n", C.c_int64), ("maxTile", C.c_int64),
                ("overflow", C.c_int64)]


def default_opts(**kw):
    o = GsOpts()
    lib().gs_default_opts(C.byref(o))
    names = {f[0] for f in GsOpts._fields_}
    for k, v in kw.items():
        if k not in names:
            raise KeyError(f"unknown trainer option {k}")
        setattr(o, k, v)
    return o


def make_camera(c):
    k = GsCamera()
    for i in range(9):
        k.R[i] = float(c["R"][i])
    for i in range(3):
        k.t[i] = float(c["t"][i])
    k.f, k.cx, k.cy = float(c["f"]), float(c["cx"]), float(c["cy"])
    k.w, k.h = int(c.get("w", 0)), int(c.get("h", 0))
    k.imgIdx = int(c.get("imgIdx", 0))
    k.g, k.b = float(c.get("g", 0.0)), float(c.get("b", 0.0))
    return k


class GSTrainer:
    """Mirror of the reference GSTrainer over the HIP library."""

    def __init__(self, **opts):
        self.opts = default_opts(**opts)
        h = C.c_void_p()
        rc = lib().gs_create(C.byref(self.opts), C.byref(h))
        if rc == -1:
            raise GsError("gs_create: an option is out of range (GS_E_ARG)")
        if rc != 0:
            raise GsError(f"gs_create failed ({rc}): no usable HIP device?")
        self._h = h
        self.shDeg = self.opts.shDeg
        self.shK = (self.shDeg + 1) ** 2 - 1 if self.shDeg > 0 else 0
        self.stride = 16

    @classmethod
    def create(cls, **opts):
        return cls(**opts)

    def _ck(self, rc):
        if rc != 0:
            raise GsError(lib().gs_last_error(self._h).decode() or f"error {rc}")

    def destroy(self):
        if self._h:
            lib().gs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # trainer.js:100
    def setup(self, gaussians, cams, images, maxViewW, maxViewH, sceneRadius):
        data = np.ascontiguousarray(gaussians["data"], np.float32).reshape(-1, 16)
        n = int(gaussians.get("n", data.shape[0]))
        if data.shape[0] < n:
            raise GsError("setup: gaussians.data shorter than n * 16")
        sh = gaussians.get("sh")
        if sh is not None:
            sh = np.ascontiguousarray(sh, np.float32).reshape(-1)
            if sh.size != n * 3 * self.shK:
                raise GsError(f"setup: gaussians.sh holds {sh.size} floats; SH degree {self.shDeg} "
                              f"needs n * 3 * {self.shK} = {n * 3 * self.shK}")
            sh = sh.reshape(n, -1) if self.shK else None
        self._cams = (GsCamera * len(cams))(*[make_camera(c) for c in cams])
        self._rgb = [np.ascontiguousarray(im["rgb"], np.float32).reshape(-1) for im in images]
        ims = (GsImage * len(images))()
        for i, im in enumerate(images):
            ims[i].tw, ims[i].th = int(im["tw"]), int(im["th"])
            ims[i].rgb = self._rgb[i].ctypes.data
        self._ck(lib().gs_setup(self._h, _p(data), n, _p(sh), self._cams, len(cams), ims, len(images),
                                int(maxViewW), int(maxViewH), float(sceneRadius)))
        self.camMeta = [dict(c, w=images[c["imgIdx"]]["tw"], h=images[c["imgIdx"]]["th"]) for c in cams]
        off = 0
        for m in self.camMeta:
            m["offset"] = off
            off += m["w"] * m["h"]

    def stepOnce(self, cam=-1):
        self._ck(lib().gs_step(self._h, int(cam)))

    step_once = stepOnce

    def stepBatch(self, cams):
        """len(cams) steps in one call, as that many stepOnce(cams[k]) calls."""
        a = np.ascontiguousarray(cams, np.int32)
        self._ck(lib().gs_step_batch(self._h, _p(a), int(a.size)))

    def stepViews(self, cams):
        """One batched step over len(cams) views (negative entries pick with
        the seeded RNG): summed gradients, one Adam step."""
        a = np.ascontiguousarray(cams, np.int32)
        self._ck(lib().gs_step_views(self._h, _p(a), int(a.size)))

    def step_views_debug(self, cams, want=("img", "T", "end", "gC", "accum", "tileRanges")):
        """stepViews without the graph, returning each view's intermediates
        and (in the first entry) the summed gradients before Adam."""
        info = self.info()
        n = info.n
        a = np.ascontiguousarray(cams, np.int32)
        outs, keep = [], []
        arr = (GsDebugOut * a.size)()
        for k, ci in enumerate(a):
            m = self.camMeta[int(ci)] if ci >= 0 else None
            npx = m["w"] * m["h"] if m else 0
            T = info.tile
            nt = ((m["w"] + T - 1) // T) * ((m["h"] + T - 1) // T) if m else 0
            shapes = {"img": ((npx, 4), np.float32), "T": ((npx,), np.float32), "end": ((npx,), np.uint32),
                      "gC": ((npx, 4), np.float32), "accum": ((n, 10), np.int32),
                      "tileRanges": ((nt, 2), np.uint32)}
            o = {key: np.zeros(*shapes[key]) for key in want}
            if k == 0:
                o["gradF"] = np.zeros((n, 16), np.float32)
                o["shGrad"] = np.zeros((n, max(3 * self.shK, 1)), np.float32)
                o["stats"] = np.zeros(4, np.float64)
            nk = np.zeros(1, np.int64)
            for key, val in o.items():
                setattr(arr[k], key, val.ctypes.data)
            arr[k].numKeys = nk.ctypes.data
            outs.append((o, nk))
        self._ck(lib().gs_step_views_debug(self._h, _p(a), int(a.size), arr))
        res = []
        for o, nk in outs:
            o["numKeys"] = int(nk[0])
            if "shGrad" in o:
                o["shGrad"] = o["shGrad"][:, :3 * self.shK]
            res.append(o)
        return res

    def fdGrad64(self, cam, gauss, slots, steps):
        """Central FP64 finite differences of camera `cam`'s training loss
        with respect to (Gaussian, slot) pairs with steps h, on the GPU."""
        g = np.ascontiguousarray(gauss, np.int64)
        sl = np.ascontiguousarray(slots, np.int32)
        h = np.ascontiguousarray(steps, np.float64)
        out = np.zeros(g.size, np.float64)
        self._ck(lib().gs_fd_grad64(self._h, int(cam), _p(g), _p(sl), _p(h), int(g.size), _p(out)))
        return out

    def setRasterShape(self, tile, batch):
        """Raster tile side (8, 16, 32) and LDS batch (64, 128, 256)."""
        self._ck(lib().gs_set_raster_shape(self._h, int(tile), int(batch)))

    def autotune(self, cams, reps=5):
        """Times the 9 raster shapes on cameras `cams` (one batched pass
        when there are several) and sets the fastest eligible one. Returns
        one dict per shape: tile, batch, status, ms, keys, dropped."""
        a = np.ascontiguousarray(cams, np.int32)
        out = (GsTuneEntry * 9)()
        self._ck(lib().gs_autotune(self._h, _p(a), int(a.size), int(reps), out))
        return [{"tile": e.tile, "batch": e.batch, "status": e.status, "ms": e.ms, "keys": e.keys,
                 "dropped": e.dropped} for e in out]

    def readLoss(self):
        mse = C.c_double()
        has = C.c_int32()
        self._ck(lib().gs_read_loss(self._h, C.byref(mse), C.byref(has)))
        return mse.value if has.value else None

    def renderView(self, cam, trainMode=0, offset=0):
        k = make_camera(cam)
        out = np.zeros((k.h, k.w, 4), np.float32)
        self._ck(lib().gs_render(self._h, C.byref(k), int(trainMode), int(offset), _p(out)))
        return out

    def renderTrainCam(self, ci):
        m = self.camMeta[ci]
        return self.renderView(m, 0, m["offset"])

    def readGaussians(self):
        info = self.info()
        data = np.zeros((info.n, 16), np.float32)
        sh = np.zeros((info.n, 3 * self.shK), np.float32) if self.shK else None
        self._ck(lib().gs_read_gaussians(self._h, _p(data), _p(sh)))
        return {"data": data, "n": info.n, "sh": sh, "shK": self.shK}

    def evalCamPsnr(self, ci):
        v = C.c_double()
        self._ck(lib().gs_eval_cam_psnr(self._h, int(ci), C.byref(v)))
        return v.value

    def evalCamPsnrRefined(self, ci, steps=200):
        v = C.c_double()
        self._ck(lib().gs_eval_cam_psnr_refined(self._h, int(ci), int(steps), C.byref(v)))
        return v.value

    def camState(self, ci):
        """The current pose, focal and exposure of camera ci (doubles): the
        fields of trainer.camMeta[ci] that camOpt updates."""
        st = GsCamState()
        self._ck(lib().gs_get_camera(self._h, int(ci), C.byref(st)))
        return {"R": list(st.R), "t": list(st.t), "f": st.f, "g": st.g, "b": st.b}

    def readCamGrads(self):
        """(numCams + 1) x 8 int32 camera gradient rows (x64 fixed point)."""
        rows = self.info().numCams + 1
        out = np.zeros((rows, 8), np.int32)
        self._ck(lib().gs_read_cam_grads(self._h, _p(out), out.size))
        return out

    @property
    def holdout(self):
        return self.info().holdout

    @holdout.setter
    def holdout(self, ci):
        self._ck(lib().gs_set_holdout(self._h, int(ci)))

    def set_excluded(self, cams):
        a = np.ascontiguousarray(list(cams), np.int32)
        self._ck(lib().gs_set_excluded(self._h, _p(a) if len(a) else None, len(a)))

    @property
    def iter(self):
        return self.info().iter

    @iter.setter
    def iter(self, v):
        self._ck(lib().gs_set_iter(self._h, int(v)))

    @property
    def n(self):
        return self.info().n

    def info(self):
        """Host fields (never waits for the GPU)."""
        i = GsInfo()
        self._ck(lib().gs_get_info(self._h, C.byref(i)))
        return i

    def step_stats(self):
        """The last raster's key and visibility counts and the cumulative
        entry-budget drops (waits for the GPU)."""
        s = GsStepStats()
        self._ck(lib().gs_read_step_stats(self._h, C.byref(s)))
        return s

    def read_loss_async(self):
        """Enqueues the loss readback; returns a Pending whose wait() gives
        (mse or None, cumulative budget drops)."""
        p = C.c_void_p()
        self._ck(lib().gs_read_loss_async(self._h, C.byref(p)))
        return Pending(p, lambda r: (r.value if r.has else None, r.overflow))

    def eval_cam_psnr_async(self, ci):
        p = C.c_void_p()
        self._ck(lib().gs_eval_cam_psnr_async(self._h, int(ci), C.byref(p)))
        return Pending(p, lambda r: r.value)

    def read_gaussians_async(self):
        info = self.info()
        data = np.zeros((info.n, 16), np.float32)
        sh = np.zeros((info.n, 3 * self.shK), np.float32) if self.shK else None
        p = C.c_void_p()
        self._ck(lib().gs_read_gaussians_async(self._h, _p(data), _p(sh), C.byref(p)))
        return Pending(p, lambda r: {"data": data, "n": info.n, "sh": sh, "shK": self.shK}, keep=(data, sh))

    def sync(self):
        self._ck(lib().gs_sync(self._h))

    def exportPly(self, bake=True, noalloc=False):
        """Standard 3DGS binary PLY bytes (ply.js gaussiansToPly), with the
        Mip opacity bake of Session.exportPlyBlob when bake. noalloc selects
        gs_export_ply_noalloc (no device allocation; same bytes)."""
        fn = lib().gs_export_ply_noalloc if noalloc else lib().gs_export_ply
        size = C.c_int64()
        self._ck(fn(self._h, 1 if bake else 0, None, 0, C.byref(size)))
        buf = (C.c_uint8 * size.value)()
        self._ck(fn(self._h, 1 if bake else 0, buf, size.value, C.byref(size)))
        return bytes(buf)

    def refine(self, seed=0, growRate=-1.0, growUntil=-1.0, growLimit=0):
        """trainer.js:666-794 with the reference's mulberry32 seeded by `seed`;
        negative growRate / growUntil and growLimit 0 select the reference's
        defaults (0.15, 0.75 * horizon, cap)."""
        o = GsRefineOpts(C.sizeof(GsRefineOpts), int(seed) & 0xFFFFFFFF, float(growRate), float(growUntil),
                         int(growLimit))
        r = GsRefineResult()
        self._ck(lib().gs_refine(self._h, C.byref(o), C.byref(r)))
        return {"moved": r.moved, "grown": r.grown, "n": r.n, "maxTile": r.maxTile, "overflow": r.overflow}

    def stream(self):
        return lib().gs_stream(self._h)

    # ---- inspection ----
    def step_debug(self, cam, want=("rec", "depthKey", "rect", "cnt", "ids", "keys", "tileRanges", "img", "T",
                                    "end", "gC", "accum", "gradF", "shGrad", "stats")):
        info = self.info()
        n = info.n
        m = self.camMeta[cam]
        npx = m["w"] * m["h"]
        T = info.tile
        nt = ((m["w"] + T - 1) // T) * ((m["h"] + T - 1) // T)
        cap = info.entriesCap
        shapes = {"rec": ((n, 16), np.float32), "depthKey": ((n,), np.uint32), "rect": ((n, 4), np.uint32),
                  "cnt": ((n,), np.uint32), "ids": ((cap,), np.uint32), "keys": ((cap,), np.uint64),
                  "tileRanges": ((nt, 2), np.uint32), "img": ((npx, 4), np.float32), "T": ((npx,), np.float32),
                  "end": ((npx,), np.uint32), "gC": ((npx, 4), np.float32), "accum": ((n, 10), np.int32),
                  "gradF": ((n, 16), np.float32), "shGrad": ((n, max(3 * self.shK, 1)), np.float32),
                  "stats": ((4,), np.float64)}
        out = {k: np.zeros(*shapes[k]) for k in want}
        nk = np.zeros(1, np.int64)
        d = GsDebugOut()
        for k, a in out.items():
            setattr(d, k, a.ctypes.data)
        d.numKeys = nk.ctypes.data
        self._ck(lib().gs_step_debug(self._h, int(cam), C.byref(d)))
        K = int(nk[0])
        for k in ("ids", "keys"):
            if k in out:
                out[k] = out[k][:K]
        out["numKeys"] = K
        if "shGrad" in out:
            out["shGrad"] = out["shGrad"][:, :3 * self.shK]
        return out

    def profile(self, enable=True):
        self._ck(lib().gs_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        buf = (GsKernelTime * 64)()
        cnt = C.c_int32()
        self._ck(lib().gs_profile_read(self._h, buf, 64, C.byref(cnt)))
        return {buf[i].name.decode(): {"ms": buf[i].total_ms, "launches": buf[i].launches, "bytes": buf[i].bytes}
                for i in range(min(cnt.value, 64))}


class Pending:
    """An enqueued readback (gs_*_async); wait() blocks until it lands."""

    def __init__(self, handle, finish, keep=()):
        self._p = handle
        self._finish = finish
        self._keep = keep     # the destination arrays stay alive until the wait

    def wait(self):
        if self._p is None:
            raise GsError("readback already waited")
        r = GsReadbackResult()
        rc = lib().gs_pending_wait(self._p, C.byref(r))
        msg = lib().gs_pending_error(self._p).decode()
        lib().gs_pending_free(self._p)
        self._p = None
        if rc != 0:
            raise GsError(msg or f"error {rc}")
        return self._finish(r)

    def __del__(self):
        if getattr(self, "_p", None) is not None:
            lib().gs_pending_free(self._p)


def createTrainer(**opts):
    """src/index.js:56-58"""
    return GSTrainer.create(**opts)
