"""Shared test helpers: oracle-side training steps with the same host
arithmetic the trainer uses (trainer.js:320-342, :409-424)."""
import math

import numpy as np

import camopt as CO
import oracle as O
import scenes


def f32(x):
    return float(np.float32(x))


def adam_for_iter(it, radius, posLr=0.0, horizon=60000, minScale=1e-4, opacityReg=0.01, shLr=1.25e-4, t=None):
    """AdamHP for the step at iteration `it` (after iter++; the position-lr
    schedule) whose Adam step is t (bias correction; t = it unless batched
    steps ran)."""
    r = f32(radius)
    base = f32(posLr) if posLr > 0 else 3e-4 * r    # gs_opts.posLr is an f32 field
    pos = f32(base * 0.01 ** min(1.0, it / (0.75 * horizon)))
    lr = [pos] * 3 + [5e-3] * 3 + [1e-3] * 4 + [1.5e-2] * 3 + [2.5e-2, 0, 0]
    return O.adam_hp(lr, t=it if t is None else t, minLS=f32(math.log(r * f32(minScale))),
                     maxLS=f32(math.log(r * 0.05)), maxLogit=8.0, reg=f32(opacityReg), shLr=f32(shLr))


class OracleTrainer:
    """The f32 restatement driven like GSTrainer.stepOnce on a fixed camera."""

    def __init__(self, params, sh, cams, images, radius, opts=None, shDeg=2, fixed=True, precision="f32",
                 tile=16):
        opts = dict(opts or {})
        self.o = O.make_opts(eCut=opts.get("eCut", 4.5), aMin=opts.get("aMin", 0.0039),
                             radClamp=opts.get("radClamp", 1.0), anisoReg=opts.get("anisoReg", 0.005),
                             shDeg=shDeg)
        self.opts = opts
        self.tile = tile      # raster tile side in pixels
        self.st = O.Stage(self.o, "f32")
        self.P = np.array(params, np.float32).reshape(-1, 16)
        n = self.P.shape[0]
        self.K = O.sh_k(shDeg)
        self.SH = (np.zeros((n, 3 * self.K), np.float32) if sh is None
                   else np.array(sh, np.float32).reshape(n, 3 * self.K))
        self.M = np.zeros_like(self.P)
        self.V = np.zeros_like(self.P)
        self.SM = np.zeros_like(self.SH)
        self.SV = np.zeros_like(self.SH)
        self.cams = cams
        self.radius = radius
        self.shDeg = shDeg
        self.fixed = fixed
        self.iter = 0       # also sets the Adam step counter, as gs_set_iter does
        self.tgt = []
        for c in cams:
            im = images[c["imgIdx"]]
            self.tgt.append(O.pack_targets(np.asarray(im["rgb"], np.float32).reshape(-1, 3)))
        self.cam_deg = [shDeg] * len(cams)
        # camera optimiser (opts camOpt / expComp): double camera state and
        # the (numCams + 1) x 8 gradient rows, wrapped to i32 when read
        self.camopt = CO.CamState(cams) if opts.get("camOpt") else None
        self.camGrad = np.zeros((len(cams) + 1, 8), np.int64)
        self.holdout = -1
        # precision modes: "mixed" keeps the SH coefficients and all moments in
        # bf16 (uint16 bits); "bf16" / "fp16" keep the SH coefficients in 16
        # bits and round the raster record's colours to nearest even
        self.mixed = precision == "mixed"
        self.fmt16 = {"bf16": 1, "fp16": 2}.get(precision, 0)
        if self.fmt16:
            self.SH16 = self._enc_rne(self.SH)
            self.SH = self._dec(self.SH16)
        if self.mixed:
            self.SH16 = O.f2bf_rne(self.SH)
            self.SH = O.bf2f(self.SH16)
            self.M16 = np.zeros(self.P.shape, np.uint16)
            self.V16 = np.zeros(self.P.shape, np.uint16)
            self.SM16 = np.zeros(self.SH.shape, np.uint16)
            self.SV16 = np.zeros(self.SH.shape, np.uint16)

    @property
    def iter(self):
        return self._iter

    @iter.setter
    def iter(self, v):
        self._iter = v
        self.adamT = v

    def _hp(self):
        return adam_for_iter(self._iter, self.radius, posLr=self.opts.get("posLr", 0.0),
                             horizon=self.opts.get("maxIters", 60000), opacityReg=self.opts.get("opacityReg", 0.01),
                             shLr=self.opts.get("shLr", 1.25e-4), t=self.adamT)

    def _adam(self, gradF, shG):
        hp = self._hp()
        if self.mixed:
            O.adam_mixed(hp, self.P, gradF, self.M16, self.V16, self.adamT)
            if self.K:
                O.sh_adam_mixed(hp, self.SH16, shG, self.SM16, self.SV16, self.adamT)
                self.SH = O.bf2f(self.SH16)
        elif self.fmt16:
            O.adam(hp, self.P, gradF, self.M, self.V)
            if self.K:
                O.sh_adam_h16(hp, self.SH16, shG, self.SM, self.SV, self.adamT, self.fmt16)
                self.SH = self._dec(self.SH16)
        else:
            O.adam(hp, self.P, gradF, self.M, self.V)
            if self.K:
                O.sh_adam(hp, self.SH, shG, self.SM, self.SV)

    def step_views(self, cis):
        """gs_step_views: every view from the same parameters, each view's
        sums chained through its camera, the gradients added in view order
        (the first view in which a Gaussian has keys gives the value, later
        ones add to it in f32), one Adam step; iter advances by the view
        count and the Adam step counter by one."""
        deg = min(self.shDeg, self._iter // 1000)
        st = self.st
        n = self.P.shape[0]
        gradF = np.zeros((n, 16), np.float32)
        shG = np.zeros((n, 3 * self.K), np.float32)
        seen = np.zeros(n, bool)
        per = []
        for ci in cis:
            cam = O.make_cam(self.cams[ci], shActiveDeg=deg, tile=self.tile)
            proj, rect, cnt = st.project(cam, self.P, self.SH if self.K else None)
            proj = self._round_record(proj)
            ids, ts, keys = st.bin_sort(cam, proj, rect, cnt)
            img, T, end = st.render_fwd(cam, proj, ids, ts)
            gC, valid, stats = st.loss_charb(cam, img, self.tgt[ci])[:3]
            g = st.render_bwd(cam, proj, ids, ts, T, end, gC, valid, fixed=self.fixed)
            sh = self.SH if self.K else None
            ch = (st.chain(cam, self.P, sh, proj, gfix=g) if self.fixed
                  else st.chain(cam, self.P, sh, proj, gsum=g))
            gv, sv = ch[0].astype(np.float32), ch[1].astype(np.float32)
            mask = cnt > 0
            first, later = mask & ~seen, mask & seen
            gradF[first] = gv[first]
            gradF[later] = gradF[later] + gv[later]
            if self.K:
                shG[first] = sv[first]
                shG[later] = shG[later] + sv[later]
            seen |= mask
            per.append({"ids": ids, "ts": ts, "img": img, "T": T, "end": end, "gC": gC, "valid": valid,
                        "stats": stats, "accum": g, "cnt": cnt})
        self._iter += len(cis)
        self.adamT += 1
        self._adam(gradF, shG)
        return per, gradF, shG

    def _enc_rne(self, x):
        return O.f2h_rne(x) if self.fmt16 == 2 else O.f2bf_rne(x)

    def _dec(self, b):
        return O.h2f(b) if self.fmt16 == 2 else O.bf2f(b)

    def _round_record(self, proj):
        """Record colours stored in 16 bits (rounded to nearest even)."""
        if self.fmt16:
            proj[:, 8:11] = self._dec(self._enc_rne(proj[:, 8:11].astype(np.float32)))
        return proj

    def cam_grads(self):
        return CO.wrap_i32(self.camGrad)

    def step(self, ci):
        deg = min(self.shDeg, self.iter // 1000)
        self.cam_deg[ci] = deg
        cd = self.camopt.cam_dict(ci) if self.camopt else self.cams[ci]
        cam = O.make_cam(cd, shActiveDeg=deg, tile=self.tile)
        self._iter += 1
        self.adamT += 1
        st = self.st
        co = self.camopt is not None
        proj, rect, cnt = st.project(cam, self.P, self.SH if self.K else None)
        proj = self._round_record(proj)
        ids, ts, keys = st.bin_sort(cam, proj, rect, cnt)
        img, T, end = st.render_fwd(cam, proj, ids, ts)
        lo = st.loss_charb(cam, img, self.tgt[ci], cam_grads=co)
        gC, valid, stats = lo[:3]
        g = st.render_bwd(cam, proj, ids, ts, T, end, gC, valid, fixed=self.fixed)
        sh = self.SH if self.K else None
        ch = (st.chain(cam, self.P, sh, proj, gfix=g, cam_grads=co) if self.fixed
              else st.chain(cam, self.P, sh, proj, gsum=g, cam_grads=co))
        gradF, shG = ch[:2]
        if co:
            self.camGrad[ci, 6:8] += lo[3]
            self.camGrad[ci, 0:6] += ch[2][:6]
            self.camGrad[len(self.cams), 0] += ch[2][6]
        pre = {"proj": proj, "rect": rect, "cnt": cnt, "ids": ids, "ts": ts, "keys": keys, "img": img, "T": T,
               "end": end, "gC": gC, "valid": valid, "stats": stats, "accum": g, "gradF": gradF, "shGrad": shG}
        self._adam(gradF, shG)
        if co:
            pre["camGrad"] = self.cam_grads()
            if self.iter > 1500 and self.iter % 25 == 0:      # trainer.js:468-470
                self.camopt.apply(self.cam_grads(), self.iter, self.opts.get("maxIters", 60000), f32(self.radius),
                                  holdout=self.holdout, exp_comp=bool(self.opts.get("expComp")))
                self.camGrad[:] = 0
        return pre

    def eval_pass(self, ci, override=None):
        """_evalPass (trainer.js:542-573): training-path pass on camera ci at
        an optional pose override, no update. Returns (psnr, the camera's 8
        gradient slots); all rows are drained afterwards."""
        src = self.camopt.cam_dict(ci, override) if self.camopt else dict(self.cams[ci], **(override or {}))
        cam = O.make_cam(src, shActiveDeg=self.shDeg, tile=self.tile)
        st = self.st
        proj, rect, cnt = st.project(cam, self.P, self.SH if self.K else None)
        proj = self._round_record(proj)
        ids, ts, _ = st.bin_sort(cam, proj, rect, cnt)
        img, T, end = st.render_fwd(cam, proj, ids, ts)
        gC, valid, stats, efix, _ = st.loss_charb(cam, img, self.tgt[ci], cam_grads=True)
        g = st.render_bwd(cam, proj, ids, ts, T, end, gC, valid, fixed=self.fixed)
        sh = self.SH if self.K else None
        ch = (st.chain(cam, self.P, sh, proj, gfix=g, cam_grads=True) if self.fixed
              else st.chain(cam, self.P, sh, proj, gsum=g, cam_grads=True))
        row = self.camGrad[ci].copy()
        row[6:8] += efix
        row[0:6] += ch[2][:6]
        self.camGrad[:] = 0
        mse = stats[0] / (max(1.0, stats[2]) * 3.0)
        psnr = -10.0 * math.log10(mse) if mse > 0 else math.inf
        return psnr, CO.wrap_i32(row)


def rig_setup(sh_deg=0):
    """The reference gradcheck rig as trainer inputs."""
    params, cams, images, W, H, radius, opts = scenes.gradcheck_rig()
    sh = scenes.sh_rig_coeffs(params.shape[0], sh_deg) if sh_deg > 0 else None
    return params, sh, cams, images, W, H, radius, opts


def config1_setup(n=20000, seed=0, target_seed=1):
    """Config 1 with its target rendered (f32 restatement) from a second set."""
    p, sh, cams = scenes.config1(seed)
    p, sh = p[:n], sh[:n]
    gt_p, gt_sh, _ = scenes.config1(target_seed)
    o = O.make_opts(shDeg=2)
    st = O.Stage(o, "f32")
    cam = O.make_cam(cams[0], shActiveDeg=2)
    proj, rect, cnt = st.project(cam, gt_p, gt_sh)
    ids, ts, _ = st.bin_sort(cam, proj, rect, cnt)
    img, _, _ = st.render_fwd(cam, proj, ids, ts)
    rgb = np.clip(img[:, :3], 0, 1).astype(np.float32)
    images = [{"tw": 320, "th": 240, "rgb": rgb.reshape(-1)}]
    return p, sh, cams, images, 320, 240, 1.0
