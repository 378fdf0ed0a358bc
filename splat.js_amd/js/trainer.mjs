// trainer.mjs — GSTrainer over the MI355X HIP trainer (N-API addon).
//
// Keeps the method surface of the reference's WebGPU GSTrainer
// (src/gs/trainer.js) so Session-style callers can drop it in:
//   GSTrainer.create(opts) / setup(gaussians, cams, images, maxViewW,
//   maxViewH, sceneRadius) / stepOnce() / readLoss() / renderView(cam, out,
//   trainMode, offset) / renderTrainCam(ci, out) / readGaussians() /
//   evalCamPsnr(ci) / evalCamPsnrRefined(ci, steps), the fields callers read
//   (iter, n, cap, camMeta (kept current under opts.camOpt),
//   holdout, excluded, lastCam, horizon, shK, stride, tileGrad, opts) and
//   device.queue.onSubmittedWorkDone().
// Differences: renderView writes into a caller Float32Array (w*h*4) instead
// of a canvas context (Node has no canvas); refine(rng) seeds the reference's
// mulberry32 from one draw of rng. Syntax stays ES2019 (no ?? or ?.) for
// Node 12.
//
// As in the reference, stepOnce only enqueues, the fields callers read every
// frame (iter, n, cap, lastCam, holdout, entryOverflowTiles) are host state
// and never wait for the GPU, and readLoss / readGaussians / evalCamPsnr
// return Promises that settle once their copy lands (the addon waits on a
// worker thread). A GPU error (code GS_E_HIP) settles device.lost with
// reason 'unknown', the signal the reference's watchLost turns into onLost
// (src/gpu/context.js:43-56); destroy() settles it with 'destroyed'.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const addon = require('../addon/gs_napi.node');

export const STRIDE = 16;
export const TILE = 16;
export const LOSS = { charbonnier: 0, l1_dssim: 1 };
export const ACCUM = { fixed: 0, f32: 1 };
export const SORT = { depthThenTile: 0, global64: 1, tileSegments: 2 };
export const PRECISION = { f32: 0, mixed: 1, bf16: 2, fp16: 3 };

const shRestCoefs = (deg) => (deg + 1) * (deg + 1) - 1;

export class GSTrainer {
  /** trainer.js:13-16 */
  static async create(opts = {}) {
    return new GSTrainer(opts);
  }

  constructor(opts = {}) {
    if (!addon.loaded()) throw new Error(addon.loadError());
    this.opts = opts;
    const o = Object.assign({}, opts);
    if (typeof o.loss === 'string') o.loss = LOSS[o.loss];
    if (typeof o.gradAccum === 'string') o.gradAccum = ACCUM[o.gradAccum];
    if (typeof o.sortMode === 'string') o.sortMode = SORT[o.sortMode];
    if (typeof o.precision === 'string') o.precision = PRECISION[o.precision];
    this._h = addon.create(o);
    this.entryOverflowTiles = 0;   // trainer.js:627, updated by readLoss
    this.tileGrad = opts.tileGrad === undefined ? true : opts.tileGrad;
    this.shDeg = opts.shDeg === undefined ? 2 : opts.shDeg;
    this.shK = this.shDeg > 0 ? shRestCoefs(this.shDeg) : 0;
    this.stride = STRIDE;
    this.pixelsSeen = 0;
    this.lastRefine = 0;
    this.growLimit = 0;   // trainer.js:777: soft ceiling below cap (0 = cap)
    this.excluded = null;
    const self = this;
    let settleLost = null;
    this._lostSettled = false;
    this._settleLost = (info) => {
      if (self._lostSettled) return;
      self._lostSettled = true;
      settleLost(info);
    };
    // session.js:520, :542 wait on the queue between frames
    this.device = {
      queue: { onSubmittedWorkDone() { return self._guard(() => { addon.sync(self._h); }); } },
      lost: new Promise((resolve) => { settleLost = resolve; }),
      destroy() { self.destroy(); },
    };
  }

  // Runs an addon call; a device error settles device.lost before it
  // propagates. Returns a Promise for Promise results.
  _guard(fn) {
    const onErr = (e) => {
      if (e && e.code === 'GS_E_HIP') this._settleLost({ reason: 'unknown', message: e.message });
      throw e;
    };
    let r;
    try {
      r = fn();
    } catch (e) {
      onErr(e);
    }
    return Promise.resolve(r).catch(onErr);
  }

  _call(fn) {
    try {
      return fn();
    } catch (e) {
      if (e && e.code === 'GS_E_HIP') this._settleLost({ reason: 'unknown', message: e.message });
      throw e;
    }
  }

  /** trainer.js:100-355 */
  setup(gaussians, cams, images, maxViewW, maxViewH, sceneRadius) {
    const sh = gaussians.sh || null;
    this._call(() => addon.setup(this._h, gaussians.data, gaussians.n, sh, cams, images, maxViewW | 0, maxViewH | 0,
                                 sceneRadius));
    this.cams = cams;
    this.sceneRadius = sceneRadius;
    let total = 0;
    this.camMeta = cams.map((c) => {
      const im = images[c.imgIdx];
      const meta = Object.assign({}, c, { w: im.tw, h: im.th, offset: total });
      total += im.tw * im.th;
      return meta;
    });
    for (const m of this.camMeta) m.f0 = m.f;
    this.horizon = this.opts.maxIters === undefined ? 60000 : this.opts.maxIters;
    this._iter = 0;   // host-side copy of iter: stepOnce reads it without a device sync
  }

  // host fields of the trainer: no GPU wait
  get iter() { return addon.info(this._h).iter; }
  set iter(v) { addon.setIter(this._h, v); this._iter = v; }
  get n() { return addon.info(this._h).n; }
  get cap() { return addon.info(this._h).cap; }
  get lastCam() { return addon.info(this._h).lastCam; }
  get holdout() { return addon.info(this._h).holdout; }
  set holdout(ci) { addon.setHoldout(this._h, ci); }

  /** trainer.js:398-471; `excluded` is a Set of camera indices as in :403 */
  stepOnce(cam = -1) {
    if (this.excluded !== this._excludedSent) {
      addon.setExcluded(this._h, this.excluded ? Array.from(this.excluded) : []);
      this._excludedSent = this.excluded;
    }
    this._call(() => addon.step(this._h, cam));
    this._iter++;
    // trainer.js:468-470, :533-535: a camera Adam step just moved the poses,
    // the shared focal and the exposures; mirror them into camMeta
    if (this.opts.camOpt && this._iter > 1500 && this._iter % 25 === 0) this._syncCamMeta();
  }

  /** Batched step (no reference counterpart): the views `cams` (-1 picks
   *  with the seeded RNG) rasterized forward and backward in one launch
   *  each, their gradients added, one Adam step; iter advances by the view
   *  count. Needs opts.viewsPerStep >= cams.length. Enqueues only. */
  stepViews(cams) {
    if (this.excluded !== this._excludedSent) {
      addon.setExcluded(this._h, this.excluded ? Array.from(this.excluded) : []);
      this._excludedSent = this.excluded;
    }
    this._call(() => addon.stepViews(this._h, cams));
    this._iter += cams.length;
  }

  /** Raster shape (no reference counterpart): tile side 8, 16 or 32 pixels
   *  and 64, 128 or 256 keys staged in LDS per round (opts.tile, opts.batch) */
  setRasterShape(tile, batch) {
    this._call(() => addon.setRasterShape(this._h, tile, batch));
  }

  /** Autotuner (no reference counterpart): times the 9 raster shapes on
   *  the views `cams` (one batched pass when there are several) and keeps
   *  the fastest eligible one; returns one entry per shape
   *  {tile, batch, status, ms, keys, dropped}. Synchronises. */
  autotune(cams, reps = 5) {
    return this._call(() => addon.autotune(this._h, cams, reps));
  }

  _syncCamMeta() {
    for (let ci = 0; ci < this.camMeta.length; ci++) {
      const c = addon.getCamera(this._h, ci);
      Object.assign(this.camMeta[ci], { R: c.R, t: c.t, f: c.f, g: c.g, b: c.b });
    }
  }

  /** trainer.js:615-631: MSE per valid pixel-channel since the last call
   *  (null when no pixel was seen); also updates entryOverflowTiles */
  readLoss() {
    return this._guard(() => addon.readLossAsync(this._h)).then((r) => {
      this.entryOverflowTiles = r.overflow;
      return r.mse;
    });
  }

  /** trainer.js:634-655: renders into `out` (Float32Array w*h*4) */
  renderView(camParams, out, trainMode = 0, offset = 0) {
    const buf = out || new Float32Array(camParams.w * camParams.h * 4);
    this._call(() => addon.render(this._h, camParams, trainMode, offset, buf));
    return buf;
  }

  /** trainer.js:657-660 */
  renderTrainCam(ci, out) {
    const meta = this.camMeta[ci];
    return this.renderView(meta, out, 0, meta.offset);
  }

  /** trainer.js:797-817 */
  readGaussians() {
    const n = this.n;
    const data = new Float32Array(n * STRIDE);
    const sh = this.shK ? new Float32Array(n * this.shK * 3) : null;
    return this._guard(() => addon.readGaussiansAsync(this._h, data, sh)).then(() => ({ data, n, sh, shK: this.shK }));
  }

  /** trainer.js:577-579 */
  evalCamPsnr(ci) {
    this.pixelsSeen = 0;
    return this._guard(() => addon.evalCamPsnrAsync(this._h, ci));
  }

  /** trainer.js:583-612: PSNR after test-time refinement of camera ci's
   *  pose (Gaussians frozen); the stored camera is unchanged */
  evalCamPsnrRefined(ci, steps = 200) {
    this.pixelsSeen = 0;
    return this._guard(() => addon.evalCamPsnrRefined(this._h, ci, steps));
  }

  /** the camera gradient rows (bufCamGrad, trainer.js:173) as an Int32Array */
  readCamGrads() {
    const out = new Int32Array((this.camMeta.length + 1) * 8);
    addon.readCamGrads(this._h, out);
    return out;
  }

  /** trainer.js:666-794. The plan's random stream is the reference's
   *  makeRng (mulberry32) seeded from one draw of `rng`; the reference
   *  consumes `rng` itself, so the same call draws different (equally
   *  distributed) numbers. growRate / growUntil come from opts as in the
   *  reference; growLimit is the trainer field. */
  async refine(rng = Math.random) {
    const o = this.opts;
    const r = this._call(() => addon.refine(this._h, {
      seed: Math.floor(rng() * 4294967296) >>> 0,
      growRate: o.growRate === undefined ? -1 : o.growRate,
      growUntil: o.growUntil === undefined ? -1 : o.growUntil,
      growLimit: this.growLimit || 0,
    }));
    return { moved: r.moved, grown: r.grown, n: r.n, maxTile: r.maxTile, overflow: r.overflow };
  }

  /** Binary 3DGS PLY of the current Gaussians (src/io/ply.js
   *  gaussiansToPly), built on the device; bake applies the Mip opacity
   *  compensation as Session.exportPlyBlob does (session.js:637-645). */
  exportPly(bake = true) { return addon.exportPly(this._h, !!bake); }

  info() { return addon.info(this._h); }

  /** the last raster's key and visible counts, and the budget drops (waits for the GPU) */
  stepStats() { return this._call(() => addon.stepStats(this._h)); }

  /** frees the device state now (device.destroy(), gradcheck.js:184) */
  destroy() {
    addon.destroy(this._h);
    this._settleLost({ reason: 'destroyed', message: 'GSTrainer.destroy()' });
  }
}
