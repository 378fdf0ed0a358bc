/* splat_gs.h — C ABI of the MI355X-native 3DGS trainer step.
 *
 * Drop-in boundary for the reference's GSTrainer (src/gs/trainer.js), reached
 * through createTrainer() (src/index.js:56-58). Plain pointers and sizes only.
 * A Node.js N-API addon (splat.js_amd/addon) binds these functions for the
 * JavaScript shim that keeps the GSTrainer method surface; Python binds them
 * with ctypes (splat.js_amd/py/gsabi.py).
 *
 * Every function returns 0 on success and a negative GS_E* code on failure;
 * gs_last_error() then holds a message (the JS shim throws it as an Error,
 * as the reference throws for oversize targets, trainer.js:122-126).
 *
 * Threading: one trainer is used from one host thread at a time. Work is
 * enqueued on the trainer's own HIP stream; gs_step never waits for the GPU
 * (the reference's queue.submit, trainer.js:458). The plain readbacks wait
 * for the stream; their *_async forms return a pending readback whose wait
 * may run on another thread.
 */
#ifndef SPLAT_GS_H
#define SPLAT_GS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_ABI_VERSION 1

/* opaque trainer handle: one per GSTrainer instance */
typedef struct gs_trainer gs_trainer_t;

enum {
  GS_OK = 0,
  GS_E_ARG = -1,       /* bad argument or call order */
  GS_E_HIP = -2,       /* HIP runtime error (device lost, launch failure) */
  GS_E_NOMEM = -3,     /* device allocation failed */
  GS_E_LIMIT = -4,     /* a size limit was exceeded (targets, view size, more than 2^28 Gaussians) */
};

enum { GS_LOSS_CHARBONNIER = 0, GS_LOSS_L1_DSSIM = 1 };
enum { GS_ACCUM_FIXED = 0, GS_ACCUM_F32 = 1 };
enum { GS_SORT_DEPTH_THEN_TILE = 0, GS_SORT_GLOBAL64 = 1, GS_SORT_TILE_SEGMENTS = 2 };
/* Storage precision. Blending, transmittance, gradients and all arithmetic
 * are f32 in every mode; stored values are decoded exactly.
 *   GS_PREC_F32:   everything f32 (the reference's layout, shaders.js:14-29)
 *   GS_PREC_MIXED: SH coefficients and all Adam moments in bf16
 *   GS_PREC_BF16:  SH coefficients and the raster record's colours in bf16,
 *                  Adam moments f32
 *   GS_PREC_FP16:  the same in IEEE fp16
 * Coefficient and moment stores round stochastically; record colours and
 * the setup upload round to nearest even. */
enum { GS_PREC_F32 = 0, GS_PREC_MIXED = 1, GS_PREC_BF16 = 2, GS_PREC_FP16 = 3 };
enum { GS_VIEWS_PERSISTENT = 0, GS_VIEWS_OVERLAP = 1 };

/* Trainer options: the reference's opts object (trainer.js:18-36, :45-49,
 * :74, :105-107, :302, :319, :338-341), plus the MI355X build's own knobs. */
typedef struct gs_opts {
  int32_t size;        /* sizeof(gs_opts); set by gs_default_opts */
  float eCut;          /* Gaussian exponent cutoff, default 4.5 (shaders.js:44); must lie in [0, 80] */
  float aMin;          /* minimum alpha, default 0.0039 (shaders.js:45) */
  float radClamp;      /* binning radius clamp, fraction of width, default 1 (shaders.js:46-53) */
  int32_t shDeg;       /* SH degree 0..3, default 2 (trainer.js:32) */
  float anisoReg;      /* anisotropy regulariser, default 0.005 (trainer.js:74) */
  float opacityReg;    /* opacity regulariser, default 0.01 (trainer.js:341) */
  float shLr;          /* SH learning rate, default 1.25e-4 (trainer.js:302) */
  float minScale;      /* min scale as a fraction of scene radius, default 1e-4 (trainer.js:338) */
  int32_t maxIters;    /* schedule horizon, default 60000 (trainer.js:319) */
  int64_t maxSplats;   /* capacity ceiling; 0 (default) = unset: capacity at most 600000
                          (trainer.js:107) */
  float capMult;       /* capacity = n * capMult, default 4 (trainer.js:106) */
  int64_t entriesCap;  /* (tile, Gaussian) key budget; 0 = 12M, or max(12M, 24*maxSplats) when
                          maxSplats is set (trainer.js:48-49) */
  int32_t tileGrad;    /* accepted for compatibility (trainer.js:26); both orders give identical sums */
  float posLr;         /* base position lr; 0 = 3e-4 * sceneRadius (trainer.js:322) */
  int32_t loss;        /* GS_LOSS_CHARBONNIER (the reference's loss) or GS_LOSS_L1_DSSIM */
  float ssimLambda;    /* D-SSIM weight for GS_LOSS_L1_DSSIM, default 0.2 */
  int32_t gradAccum;   /* GS_ACCUM_FIXED (the reference's x16384 i32 sums) or GS_ACCUM_F32 */
  int32_t sortMode;    /* key ordering strategy, all giving the identical permutation:
                          GS_SORT_TILE_SEGMENTS (default): per-tile segments sorted in LDS, with
                          the reference's entry budget (whole tiles dropped, shaders.js:321-370);
                          the two radix strategies drop the keys past the budget in sort order */
  uint32_t seed;       /* camera-pick RNG seed (the reference uses Math.random, trainer.js:344) */
  int32_t device;      /* HIP device ordinal, default 0 */
  int32_t precision;   /* GS_PREC_F32 (default), GS_PREC_MIXED, GS_PREC_BF16 or GS_PREC_FP16
                          (storage formats above) */
  int32_t camOpt;      /* camera pose + shared focal optimisation, default 0 (trainer.js:468):
                          per-camera gradients each step (shaders.js:614-616, :834-858), one
                          camera Adam step every 25 iterations after 1500 (trainer.js:476-537) */
  int32_t expComp;     /* with camOpt, also per-camera exposure gain/bias, default 0
                          (trainer.js:511); the exposure gradients come from the Charbonnier loss,
                          so with GS_LOSS_L1_DSSIM they stay zero */
  int32_t viewsPerStep;/* most views one gs_step_views call may batch, 1..16, default 1; setup
                          sizes one set of per-view buffers for each */
  int32_t graph;       /* gs_step_views replays a captured hipGraph when its shape repeats
                          (default 1) */
  int32_t viewsLaunch; /* rasterization of a gs_step_views step: GS_VIEWS_PERSISTENT (the
                          default: one persistent forward and one persistent backward launch over
                          all views' tiles) or GS_VIEWS_OVERLAP (each view's forward and backward on
                          a second stream as soon as its binning is done, beside the binning of the
                          next views; measured slower on MI355X); identical results */
  int32_t tile;        /* raster tile side in pixels: 16 (default; 0 means 16; the reference's
                          TILE, shaders.js:32), 8 or 32. The tile grid decides which (tile,
                          Gaussian) keys exist and which whole tiles the entry budget drops; the
                          restatement at the same tile size gives the same results bit for bit */
  int32_t batch;       /* keys staged in LDS per raster round: 256 (default; 0 means 256), 128 or
                          64; changes no result */
} gs_opts;

/* One training camera (trainer.js:95-100 cams[], :357-372): R row-major
 * world-to-camera, t, focal f, principal point, size, target image index,
 * exposure log-gain g and bias b. */
typedef struct gs_camera {
  float R[9];
  float t[3];
  float f, cx, cy;
  int32_t w, h;
  int32_t imgIdx;
  float g, b;
} gs_camera;

/* A camera's current state in double precision, the reference's camMeta
 * entry (trainer.js:306, :518-533): pose, focal (original focal times the
 * shared exp(logfScale)) and exposure log-gain g and bias b. */
typedef struct gs_cam_state {
  double R[9];
  double t[3];
  double f, g, b;
} gs_cam_state;

/* One target image (trainer.js:98 images[]): tw x th RGB floats in [0, 1],
 * r < 0 marks an invalid pixel (trainer.js:133). */
typedef struct gs_image {
  int32_t tw, th;
  const float* rgb;
} gs_image;

/* Fields the reference's callers read off the trainer object
 * (session.js:263-277, :479-497, :572-604; trainer.js:27, :102-107, :319).
 * All of them are host state: gs_get_info never waits for the GPU, so a
 * caller may read them every frame while steps are queued. */
typedef struct gs_info {
  int64_t n, cap, iter;
  int32_t shDeg, shK, stride, numCams;
  int32_t lastCam, holdout;
  int64_t entriesCap;
  int32_t maxTiles;
  float horizon;
  int32_t tile, batch; /* raster shape in use (gs_opts.tile / .batch, gs_set_raster_shape) */
} gs_info;

/* Device-side figures of the last raster (synchronises). */
typedef struct gs_step_stats {
  int64_t lastKeys;          /* (tile, Gaussian) keys of the last raster (of all its views when it
                                was a gs_step_views step) */
  int64_t lastVisible;       /* Gaussians with at least one key in the last raster (summed over
                                the views of a gs_step_views step) */
  int64_t overflowKeys;      /* entry-budget drops of the training steps, cumulative since setup:
                                tiles with the default sort (the reference's entryOverflowTiles,
                                trainer.js:627), keys with the radix sorts */
} gs_step_stats;

/* fills defaults (the reference's values) */
int gs_default_opts(gs_opts* o);

/* GSTrainer.create(opts) (trainer.js:13-16): creates the HIP device context. */
int gs_create(const gs_opts* o, gs_trainer_t** out);
/* device.destroy() (gradcheck.js:184) */
void gs_destroy(gs_trainer_t* h);

/* GSTrainer.setup(gaussians, cams, images, maxViewW, maxViewH, sceneRadius)
 * (trainer.js:100-355). params: n x 16 f32 in the reference layout
 * (shaders.js:14-16); sh: n x 3K f32 channel-major (K of the trainer's
 * shDeg; the bindings reject any other length) or NULL (zeros).
 * Copies everything; the caller keeps ownership. */
int gs_setup(gs_trainer_t* h, const float* params, int64_t n, const float* sh,
             const gs_camera* cams, int32_t ncams, const gs_image* images, int32_t nimages,
             int32_t maxViewW, int32_t maxViewH, float sceneRadius);

/* GSTrainer.stepOnce() (trainer.js:398-471). cam < 0 picks a camera with the
 * seeded RNG, skipping the holdout and excluded ones (trainer.js:400-406);
 * cam >= 0 trains on that camera. Enqueues only. */
int gs_step(gs_trainer_t* h, int32_t cam);

/* Batched step (no counterpart in the reference, whose stepOnce trains one
 * view): `count` views (cams[k] < 0 picks with the seeded RNG, as gs_step
 * does) are projected and binned one by one, then rasterized forward (with
 * the Charbonnier loss) in one persistent launch over all their tiles and
 * backward in another; each view's screen-space sums are chained through its
 * own camera, the views' gradients are added in view order, and one Adam
 * step follows. The iteration counter (band ramp, position-lr schedule)
 * advances by `count`; the Adam step counter (bias correction) by one. With
 * opts.graph the whole step is captured once per shape (view sizes, n) and
 * replayed as a hipGraph; the per-step values (cameras, Adam constants) come
 * from device memory that a copy before each step rewrites. count <=
 * opts.viewsPerStep; needs the Charbonnier loss, the tile-segment sort and
 * camOpt off (a view of more than 16384 tiles, past the tile-segment
 * binning's LDS tables, is binned by the depth-then-tile radix path inside
 * the same step). Enqueues only. */
int gs_step_views(gs_trainer_t* h, const int32_t* cams, int32_t count);

/* `count` training steps enqueued in one call, exactly as `count` gs_step
 * calls with cams[k] (k-th step; < 0 picks with the seeded RNG). No
 * counterpart in the reference, whose host submits one step per stepOnce. */
int gs_step_batch(gs_trainer_t* h, const int32_t* cams, int32_t count);

/* GSTrainer.readLoss() (trainer.js:615-631): mean squared error per valid
 * pixel-channel since the last call, then resets. *has = 0 is the
 * reference's null (no pixels seen). Synchronises; gs_read_loss_async is
 * the form that does not. */
int gs_read_loss(gs_trainer_t* h, double* mse, int32_t* has);

/* ---- asynchronous readbacks (the reference's copy-to-staging plus
 * mapAsync, trainer.js:615-631, :797-817, :542-573) ----
 * A *_async call enqueues the device work and the copy into pinned host
 * memory on the trainer's stream, updates the trainer's host state as the
 * reference does at that point (readLoss resets its pixel window), and
 * returns at once with a pending readback. gs_pending_wait blocks until the
 * copy has landed and finishes the result; it touches no trainer state, so
 * it may run on another thread (the addon's napi_async_work) while the
 * trainer's own thread keeps enqueueing steps. A pending readback stays
 * valid after gs_destroy of its trainer. */
typedef struct gs_pending gs_pending_t;

typedef struct gs_readback_result {
  double value;       /* readLoss: MSE; evalCamPsnr: PSNR */
  int32_t has;        /* readLoss: 0 is the reference's null (no pixel seen); otherwise 1 */
  int64_t overflow;   /* readLoss: entry-budget drops of the training steps, cumulative
                         (the reference's entryOverflowTiles after this readLoss) */
} gs_readback_result;

int gs_read_loss_async(gs_trainer_t* h, gs_pending_t** out);
/* evalCamPsnr(ci); the training loss window is reset at enqueue (trainer.js:570-571) */
int gs_eval_cam_psnr_async(gs_trainer_t* h, int32_t ci, gs_pending_t** out);
/* readGaussians(): params n x 16 and, when shDeg > 0, sh n x 3K are written
 * by gs_pending_wait (either may be NULL); n is the count at enqueue. The
 * caller keeps both arrays alive until the wait returns. */
int gs_read_gaussians_async(gs_trainer_t* h, float* params, float* sh, gs_pending_t** out);
/* waits for the readback; 0 or GS_E_HIP (message in gs_pending_error) */
int gs_pending_wait(gs_pending_t* p, gs_readback_result* out);
const char* gs_pending_error(gs_pending_t* p);
void gs_pending_free(gs_pending_t* p);

/* GSTrainer.renderView(camParams, ctx, trainMode, offset) (trainer.js:634-655):
 * renders into a caller buffer of w*h*4 f32 (C + T*bg, 1 - T) instead of a
 * canvas. trainMode = 1 also accumulates the loss against the target pixels
 * at `offset`, without gradients. */
int gs_render(gs_trainer_t* h, const gs_camera* cam, int32_t trainMode, int64_t offset,
              float* rgba);

/* GSTrainer.readGaussians() (trainer.js:797-817): params n x 16 and, when
 * shDeg > 0, sh n x 3K (either may be NULL). */
int gs_read_gaussians(gs_trainer_t* h, float* params, float* sh);

/* GSTrainer.evalCamPsnr(ci) (trainer.js:542-579). */
int gs_eval_cam_psnr(gs_trainer_t* h, int32_t ci, double* psnr);

/* GSTrainer.evalCamPsnrRefined(ci, steps) (trainer.js:583-612): `steps`
 * Adam steps on camera ci's pose alone (rotation lr 3e-4, translation
 * 3e-4 * sceneRadius; exposure from zero, lr 5e-3, when expComp), each from
 * one training-path pass with the Gaussians frozen, then the PSNR at the
 * refined pose. The stored camera is not changed. Synchronous. */
int gs_eval_cam_psnr_refined(gs_trainer_t* h, int32_t ci, int32_t steps, double* psnr);

/* trainer.camMeta[ci] (trainer.js:306): the camera's current pose, focal and
 * exposure, which camOpt updates. */
int gs_get_camera(gs_trainer_t* h, int32_t ci, gs_cam_state* out);

/* trainer.holdout = ci (trainer.js:309, :396-397); -1 clears. */
int gs_set_holdout(gs_trainer_t* h, int32_t ci);
/* trainer.excluded = new Set(cams) (trainer.js:403) */
int gs_set_excluded(gs_trainer_t* h, const int32_t* cams, int32_t count);
/* trainer.iter = iter: the reference's callers assign it (session.js resume) */
int gs_set_iter(gs_trainer_t* h, int64_t iter);

/* host fields only; never waits for the GPU */
int gs_get_info(gs_trainer_t* h, gs_info* out);
/* the last raster's key and visibility counts and the cumulative budget drops */
int gs_read_step_stats(gs_trainer_t* h, gs_step_stats* out);

/* Standard 3DGS binary PLY of the current Gaussians (the reference's
 * gaussiansToPly, src/io/ply.js:47-100: x y z, zero normals, f_dc, f_rest
 * channel-major, pre-sigmoid opacity, log-scales, unit quaternion w x y z).
 * bake != 0 first applies bakeOpacityCompensation (ply.js:12-37) as
 * Session.exportPlyBlob does (session.js:637-645): focal of camera 0, the
 * positions of all training cameras. The rows are built on the device in
 * f64 (the reference's JS numbers) and stored as f32. With buf == NULL only
 * *size is set; otherwise cap must be at least that size. */
int gs_export_ply(gs_trainer_t* h, int32_t bake, void* buf, int64_t cap, int64_t* size);

/* gs_export_ply without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments and file bytes as gs_export_ply; the rows are
 * built in chunks of at most 16 MiB in a fixed device array of the library
 * (about 16 MiB per GPU, held until the process ends) and copied into buf
 * chunk by chunk. Returns GS_E_LIMIT for more than 4096 cameras (the array
 * holds 4096 camera centres for the opacity bake). Calls that overlap on one
 * GPU, from any trainers or threads, share that array and race. */
int gs_export_ply_noalloc(gs_trainer_t* h, int32_t bake, void* buf, int64_t cap, int64_t* size);

/* GSTrainer.refine(rng) (trainer.js:666-794), MCMC-lite: relocate dead
 * splats (opacity < 0.02, at most ceil(5% n)) onto jittered clones of
 * donors (opacity > 0.4), then grow by ceil(growRate * n) (until growUntil,
 * at most growLimit) with 70% splits of the largest donors. The random
 * stream is the reference's makeRng (mulberry32, src/sfm/geometry.js:750-758)
 * seeded with `seed`; the reference's caller passes Math.random.
 * Only opacities and log-scales (16 B per Gaussian) come to the host, which
 * replays the reference's sequential choices; the copies, jitter and moment
 * resets are applied on the device. Synchronous. */
typedef struct gs_refine_opts {
  int32_t size;          /* sizeof(gs_refine_opts) */
  uint32_t seed;
  float growRate;        /* < 0: the reference's default 0.15 (trainer.js:779) */
  double growUntil;      /* < 0: 0.75 * horizon (trainer.js:778) */
  int64_t growLimit;     /* <= 0: cap (trainer.growLimit, trainer.js:777) */
} gs_refine_opts;

typedef struct gs_refine_result {
  int64_t moved, grown, n;
  int64_t maxTile;       /* most keys in one tile of the last raster */
  int64_t overflow;      /* entry-budget drops, cumulative (as gs_info.overflowKeys) */
} gs_refine_result;

int gs_refine(gs_trainer_t* h, const gs_refine_opts* o, gs_refine_result* out);

/* device.queue.onSubmittedWorkDone() (session.js:520, :542) */
int gs_sync(gs_trainer_t* h);

/* message of the last failing call on this trainer ("" if none) */
const char* gs_last_error(gs_trainer_t* h);

/* ---- inspection (no reference counterpart: the parity tests and the
 * benchmark read the step's intermediate buffers through these) ---- */

/* Host buffers for one step's intermediates; NULL members are skipped. */
typedef struct gs_debug_out {
  float* rec;           /* n x 16 raster records (mx,my,A,B, C,opa*comp,ex,ey, r,g,b,opa, comp,cnorm,0,0) */
  uint32_t* depthKey;   /* n, f32 depth bits or 0xffffffff when culled */
  uint32_t* rect;       /* n x 4 (tx0, tx1, ty0, ty1) */
  uint32_t* cnt;        /* n, tiles covered (0 = culled) */
  uint32_t* ids;        /* keys, Gaussian id per sorted key (capacity entriesCap) */
  uint64_t* keys;       /* keys, (tile << 32) | depth bits per sorted key */
  int64_t* numKeys;     /* 1 */
  uint32_t* tileRanges; /* numTiles x 2 (start, end) */
  float* img;           /* w*h*4 */
  float* T;             /* w*h final transmittance */
  uint32_t* end;        /* w*h one past the last processed key */
  float* gC;            /* w*h*4 dL/dC (xyz) and valid flag (w) */
  int32_t* accum;       /* n x 10 screen-space gradient sums (i32 fixed point or f32 bits) */
  float* gradF;         /* n x 16 parameter gradients (before Adam) */
  float* shGrad;        /* n x 3K */
  double* stats;        /* 4: sum |err|^2, loss, valid pixels, dropped keys */
} gs_debug_out;

/* One gs_step on camera `cam` that also copies the intermediates out
 * (synchronises). */
int gs_step_debug(gs_trainer_t* h, int32_t cam, gs_debug_out* out);

/* One gs_step_views without the graph, copying each view's intermediates
 * into outs[k] (img, T, end, gC, accum, tileRanges, numKeys); outs[0].gradF
 * and .shGrad receive the summed gradients before Adam. Synchronises. */
int gs_step_views_debug(gs_trainer_t* h, const int32_t* cams, int32_t count, gs_debug_out* outs);

/* The camera gradient rows accumulated since the last camera Adam step or
 * evaluation pass (trainer.js:173, bufCamGrad): (numCams + 1) x 8 i32 in
 * x64 fixed point, [dw x,y,z, dt x,y,z, d log gain, d bias] per camera and
 * d log f in slot 0 of the last row. count >= (numCams + 1) * 8. */
int gs_read_cam_grads(gs_trainer_t* h, int32_t* out, int32_t count);

/* Gradient check (the reference's gradCheck, gradcheck.js:266-356): central
 * finite differences in FP64 of the training loss of camera `cam` (the
 * Charbonnier sum over its valid pixels, at the current parameters and SH
 * band ramp) with respect to `count` parameters: entry k moves slot
 * slots[k] (0..13) of Gaussian gauss[k] by +-steps[k], and out[k] =
 * (L(+h) - L(-h)) / (2h). Runs on the GPU: the current raster's sorted keys
 * and records for every other Gaussian, the perturbed Gaussian's record in
 * FP64 at its key's place in the current sort order (as the analytic backward
 * assumes), the tiles of its rectangle plus one tile of margin. A Gaussian
 * without keys in the view (culled) gives 0. The parameters are not changed. No counterpart in the reference (whose gradCheck reruns
 * whole forward passes); synchronises. */
int gs_fd_grad64(gs_trainer_t* h, int32_t cam, const int64_t* gauss, const int32_t* slots, const double* steps,
                 int32_t count, double* out);

/* Raster shape: tile side 8, 16 or 32 pixels and 64, 128 or 256 keys staged
 * in LDS per round (as gs_opts.tile and .batch), from the next step or
 * render on. No counterpart in the reference (16-pixel tiles, shaders.js:32,
 * and no LDS staging). */
int gs_set_raster_shape(gs_trainer_t* h, int32_t tile, int32_t batch);

/* One shape of an autotuning sweep. */
typedef struct gs_tune_entry {
  int32_t tile, batch;
  int32_t status;       /* GS_TUNE_MEASURED or GS_TUNE_MORE_DROPS (the entry budget drops
                           more keys than at 16-pixel tiles: faster only by rendering
                           less). GS_TUNE_TOO_MANY_TILES keeps its value but is no longer
                           returned: every shape is measured */
  int32_t pad;
  double ms;            /* mean time of one pass (status GS_TUNE_MEASURED) */
  int64_t keys;         /* sorted keys of one pass, all views */
  int64_t dropped;      /* keys the entry budget dropped in one pass, all views */
} gs_tune_entry;
enum { GS_TUNE_MEASURED = 0, GS_TUNE_TOO_MANY_TILES = 1, GS_TUNE_MORE_DROPS = 2 };

/* Autotuner: for each of the 9 shapes (tile-major: tiles 8, 16, 32, batches
 * 64, 128, 256 within each), times `reps` passes (after one warm-up pass) of
 * the training raster of cameras cams[0..count): projection, binning, the
 * forward pass with the loss and the backward pass, as gs_step (count 1) or
 * gs_step_views (count > 1, at most opts.viewsPerStep) run them, without the
 * update. Parameters, optimizer state, iteration count and loss window are
 * left as they were. Writes the 9 entries to out and sets the fastest
 * measured shape (the current one if none is). No counterpart in the
 * reference; synchronises. */
int gs_autotune(gs_trainer_t* h, const int32_t* cams, int32_t count, int32_t reps, gs_tune_entry* out);

/* Per-kernel timing over subsequent steps, measured with HIP events on the
 * trainer's stream. enable = 0 stops and keeps the totals. */
typedef struct gs_kernel_time {
  char name[32];
  double total_ms;
  int64_t launches;
  double bytes;         /* algorithmic HBM bytes summed over the launches */
} gs_kernel_time;
int gs_profile(gs_trainer_t* h, int32_t enable);
int gs_profile_read(gs_trainer_t* h, gs_kernel_time* out, int32_t max, int32_t* count);

/* the trainer's HIP stream (hipStream_t), for callers that time it */
void* gs_stream(gs_trainer_t* h);

#ifdef __cplusplus
}
#endif
#endif
