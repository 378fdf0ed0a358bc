// gs_trainer.cpp — host side of the trainer: the C ABI in include/splat_gs.h.
//
// Mirrors GSTrainer (src/gs/trainer.js): buffer setup and target packing
// (:100-355), the per-step host sequencing (:398-471), readLoss (:615-631),
// renderView (:634-655), readGaussians (:797-817), evalCamPsnr (:542-579).
// All device work goes to one HIP stream owned by the trainer; the step
// enqueues only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/splat_gs.h"
#include "gs_kernels.h"

using namespace gs;

namespace {

constexpr int kSortBlocks = 512;    // radix grid (measured best of 256/512/1024/2048 at 1M Gaussians, 3M keys)
constexpr int64_t kDefaultEntries = 12000000;   // shaders.js:34 ENTRIES_CAP
// most device memory the batched binning's per-view copies may take
constexpr size_t kBatchedBinBytes = size_t(24) << 30;
constexpr int64_t kTargetLimitBytes = int64_t(64) << 30;
constexpr int kRing = 64;           // pinned ViewBlocks: the host may run this many batched steps ahead

struct CamMeta {
  gs_camera c;      // f32 view of the camera, as the kernels take it (trainer.js:357-372)
  int64_t offset;   // first target pixel
  int activeDeg;    // SH degree in this camera's uniform (trainer.js:367, :411)
  // the reference's camMeta entry in JS numbers: pose, focal (f0 the
  // original) and exposure, which the camera optimizer updates
  // (trainer.js:306, :518-533)
  double R[9], t[3];
  double f0, f, g, b;
};

struct ProfSlot {
  std::string name;
  double ms = 0.0;
  int64_t launches = 0;
  double bytes = 0.0;
};

struct PendingEv {
  int slot;
  hipEvent_t a, b;
  double bytes;
  bool ownsA;   // false: a is the previous stage's end event (returned with it)
};

uint32_t mulberry32(uint32_t& a) {
  // src/sfm/geometry.js:751-759
  a += 0x6D2B79F5u;
  uint32_t t = (a ^ (a >> 15)) * (1u | a);
  t = (t + ((t ^ (t >> 7)) * (61u | t))) ^ t;
  return t ^ (t >> 14);
}

// Math.round then Math.min(255, .) then ToInt32 (trainer.js:134-137)
int32_t js_round255(double x) {
  if (std::isnan(x)) return 0;
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1.0;
  if (r > 255.0) r = 255.0;
  if (std::isinf(r)) return 0;
  double m = std::fmod(r, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  return (int32_t)(uint32_t)m;
}

}  // namespace

struct gs_trainer {
  gs_opts o{};
  hipStream_t st = nullptr;
  std::string err;
  bool ready = false;

  int64_t n = 0, cap = 0;
  int shDeg = 0, shK = 0, shC = 0, shP4 = 0, shQ8 = 0;
  int prec = 0;         // gs_opts.precision (storage formats: Prec in gs_device.h)
  int shFmt = 0;        // SH coefficients: 0 f32 planes, 1 bf16, 2 fp16 (16-bit planes)
  bool mom16 = false;   // bf16 Adam moments (GS_PREC_MIXED)
  int recFmt = 0;       // raster record colours: 0 f32 (64 B records), 1 bf16, 2 fp16 (48 B)
  int64_t entriesCap = 0;
  int maxTiles = 0;      // tiles of the largest view at the smallest tile size (8): every shape fits
  int64_t maxPix = 0;
  RasterShape shape{kDefaultTile, kDefaultBatch};   // raster tile side and LDS batch

  std::vector<void*> allocs;
  float4* P = nullptr;
  void *M = nullptr, *V = nullptr, *SH = nullptr, *SHM = nullptr, *SHV = nullptr;   // f32 or bf16 planes
  float4* rec = nullptr;
  uint32_t *depthKey = nullptr, *cnt = nullptr;
  uint2* rect = nullptr;
  void* accum = nullptr;
  uint32_t *dkA = nullptr, *dvA = nullptr, *dkB = nullptr, *dvB = nullptr;
  void *tkA = nullptr, *tkB = nullptr;
  uint32_t *tvA = nullptr, *tvB = nullptr;
  SortScratch sort{};
  uint32_t* dN = nullptr;
  uint32_t* numKeys = nullptr;
  uint32_t* numVisible = nullptr;
  uint64_t* totalKeys = nullptr;
  uint2* ranges = nullptr;
  uint32_t* tileOrder = nullptr;
  float4* img = nullptr;
  float* Tfin = nullptr;
  uint32_t* endIdx = nullptr;
  float4* gC = nullptr;
  float* ssimScratch = nullptr;  // 9 x maxPix SSIM partials (L1 + D-SSIM loss only)
  uint32_t* targets = nullptr;
  int64_t totalPx = 0;
  double* stats = nullptr;      // training stats [4]
  double* evalStats = nullptr;  // eval/view stats [4]
  uint32_t* camGrad = nullptr;  // (numCams + 1) x 8 i32 camera gradient rows (trainer.js:173)

  std::vector<CamMeta> cams;
  // camera optimizer state (trainer.js:308-313)
  std::vector<double> camM, camV;
  int64_t camStep = 0;
  double logfScale = 0.0;
  int64_t iter = 0;
  int64_t adamT = 0;      // Adam steps (bias correction); equals iter unless batched steps ran
  int64_t pixelsSeen = 0;
  float sceneRadius = 1.0f;
  double basePosLr = 0.0;
  double horizon = 60000.0;
  int holdout = -1;
  std::vector<char> excluded;
  uint32_t rng = 0;
  int lastCam = -1;
  int lastNumTiles = 0;   // tiles of the last raster (refine's maxTile)
  bool lastSegPath = false;   // the last raster binned by tile segments (raw tile counts in lastTileCnt)
  const uint32_t* lastTileCnt = nullptr;   // its per-tile counts before the budget (tile segments)
  const uint2* lastRanges = nullptr;       // its tile ranges
  BinOut lastBin{};

  // batched views (gs_step_views): buffer slot k holds view k's records,
  // counts, gradient sums, sorted ids, tile ranges and order, and pixel
  // buffers; slot 0 is the single-view set above
  struct ViewSlot {
    float4* rec; uint32_t* cnt; void* accum; uint32_t* ids; uint2* ranges; uint32_t* tileOrder;
    float4* img; float* T; uint32_t* end; float4* gC;
    // the batched binning's own copies (slot 0: the single-view buffers)
    uint32_t* depthKey; uint2* rect; uint32_t* chunkKeys; uint32_t* blockStart; uint32_t* table;
    uint32_t* visPart; uint32_t* tileCnt; uint64_t* keys; uint64_t* tmp; uint64_t* totalKeys;
  };
  bool batchedBin = false;         // every view has its own binning buffers: one launch per stage
  std::vector<ViewSlot> vslots;
  ViewBlock* dVB = nullptr;        // device copy of the step's ViewBlock
  ViewBlock* hVB = nullptr;        // pinned ring of kRing host ViewBlocks
  std::vector<hipEvent_t> ringEv;  // copy-done event of each ring entry
  int ringPos = 0;
  uint32_t* dQueue = nullptr;      // work counter of the persistent raster launch
  uint32_t* viewStats = nullptr;   // kMaxViews key counts, then kMaxViews visible counts
  hipStream_t st2 = nullptr;       // the overlapped form's raster stream
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  hipEvent_t evBin[kMaxViews] = {};   // view k binned
  int lastViews = 0;               // views of the last raster when it was a batched step (else 0)
  int viewBlocks = 0;              // blocks of the persistent launches
  hipGraphExec_t gExec = nullptr;  // captured batched step and the shape it was captured for
  std::vector<int64_t> gKey;
  std::vector<int64_t> lastEagerKey;   // shape of the last eager batched step

  bool prof = false;
  std::vector<ProfSlot> slots;
  std::vector<PendingEv> pending;
  hipEvent_t profLast = nullptr;   // end event of the last profiled stage of this raster/step
  std::vector<hipEvent_t> evPool;
};

// One asynchronous readback: pinned host copies, filled on the trainer's
// stream, and an event behind them. It owns everything it touches, so it
// outlives its trainer and its wait can run on any thread.
struct gs_pending {
  enum Kind { kLoss, kPsnr, kGaussians } kind = kLoss;
  int device = 0;
  hipEvent_t ev = nullptr;
  double* stats = nullptr;     // pinned [4] (loss, psnr)
  int64_t px = 0;              // pixels the loss window saw (loss)
  void* hP = nullptr;          // pinned parameter planes (gaussians)
  void* hSH = nullptr;         // pinned SH planes (gaussians)
  float* dstP = nullptr;
  float* dstSH = nullptr;
  int64_t n = 0, cap = 0;
  int shC = 0, shP4 = 0, shQ8 = 0;
  int shFmt = 0;
  std::string err;
};

namespace {

int fail(gs_trainer* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  return code;
}

int hipfail(gs_trainer* h, const char* what, hipError_t e) {
  return fail(h, GS_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCK(expr)                                   \
  do {                                                \
    hipError_t e_ = (expr);                           \
    if (e_ != hipSuccess) return hipfail(h, #expr, e_); \
  } while (0)

template <typename T>
int dalloc(gs_trainer* h, T** p, size_t bytes) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(h, GS_E_NOMEM, "hipMalloc(" + std::to_string(bytes) + " B) failed: " + hipGetErrorString(e));
  }
  h->allocs.push_back(q);
  *p = (T*)q;
  return GS_OK;
}

void free_all(gs_trainer* h) {
  if (h->gExec) {
    (void)hipGraphExecDestroy(h->gExec);
    h->gExec = nullptr;
  }
  h->gKey.clear();
  h->lastEagerKey.clear();
  for (hipEvent_t e : h->ringEv) {
    (void)hipEventSynchronize(e);
    (void)hipEventDestroy(e);
  }
  h->ringEv.clear();
  if (h->hVB) {
    (void)hipHostFree(h->hVB);
    h->hVB = nullptr;
  }
  for (void* p : h->allocs) (void)hipFree(p);
  h->allocs.clear();
  h->vslots.clear();
  h->ready = false;
}

hipEvent_t ev_get(gs_trainer* h) {
  if (!h->evPool.empty()) {
    hipEvent_t e = h->evPool.back();
    h->evPool.pop_back();
    return e;
  }
  // Timing events only: without the system-scope fence a record costs no
  // cache writeback and invalidation, which otherwise idles the GPU for
  // ~10 us at every stage boundary of a profiled step.
  hipEvent_t e = nullptr;
  (void)hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

int prof_slot(gs_trainer* h, const char* name) {
  for (size_t k = 0; k < h->slots.size(); k++)
    if (h->slots[k].name == name) return (int)k;
  h->slots.push_back(ProfSlot{name});
  return (int)h->slots.size() - 1;
}

// Brackets one stage with events on the trainer's stream when profiling.
// Consecutive stages of one step share their boundary: a stage starts at the
// previous stage's end event (nothing is enqueued between them), so a step
// records one event per stage plus one, not two per stage.
struct Scope {
  gs_trainer* h;
  int slot = -1;
  hipEvent_t a = nullptr;
  bool ownsA = true;
  double bytes;
  Scope(gs_trainer* h_, const char* name, double bytes_) : h(h_), bytes(bytes_) {
    if (!h->prof) return;
    slot = prof_slot(h, name);
    if (h->profLast) {
      a = h->profLast;
      ownsA = false;
    } else {
      a = ev_get(h);
      (void)hipEventRecord(a, h->st);
    }
  }
  ~Scope() {
    if (slot < 0) return;
    hipEvent_t b = ev_get(h);
    (void)hipEventRecord(b, h->st);
    h->pending.push_back(PendingEv{slot, a, b, bytes, ownsA});
    h->profLast = b;
  }
};

void prof_collect(gs_trainer* h) {
  if (h->pending.empty()) return;
  (void)hipStreamSynchronize(h->st);
  for (auto& p : h->pending) {
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    h->slots[p.slot].ms += ms;
    h->slots[p.slot].launches += 1;
    h->slots[p.slot].bytes += p.bytes;
    if (p.ownsA) h->evPool.push_back(p.a);
    h->evPool.push_back(p.b);
  }
  h->pending.clear();
  h->profLast = nullptr;
}

constexpr int kMinTile = 8;   // smallest raster tile side

int tiles_of(int w, int hgt, int tile) { return ((w + tile - 1) / tile) * ((hgt + tile - 1) / tile); }

Cam make_cam(const gs_trainer* h, const gs_camera& c, int trainMode, int64_t offset, int activeDeg) {
  Cam k{};
  for (int i = 0; i < 9; i++) k.R[i] = c.R[i];
  for (int i = 0; i < 3; i++) { k.t[i] = c.t[i]; k.bg[i] = 0.0f; }
  k.nearp = 0.05f;                       // trainer.js:362
  k.f = c.f; k.cx = c.cx; k.cy = c.cy;
  k.gain = (float)std::exp((double)c.g); // trainer.js:363
  k.bias = c.b;
  // shaders.js:121-126, in f32
  const float* R = k.R;
  const float* t = k.t;
  k.camPos[0] = -((R[0] * t[0] + R[3] * t[1]) + R[6] * t[2]);
  k.camPos[1] = -((R[1] * t[0] + R[4] * t[1]) + R[7] * t[2]);
  k.camPos[2] = -((R[2] * t[0] + R[5] * t[1]) + R[8] * t[2]);
  k.W = c.w; k.H = c.h;
  k.tile = h->shape.tile;
  k.tilesX = (c.w + k.tile - 1) / k.tile;
  k.tilesY = (c.h + k.tile - 1) / k.tile;
  k.numTiles = k.tilesX * k.tilesY;
  const int a = (activeDeg + 1) * (activeDeg + 1) - 1;
  k.shActiveK = std::min(a, h->shK);
  k.tgtOffset = offset;
  (void)trainMode;
  return k;
}

// the training uniform of a camera: its f32 fields plus the exposure of
// its JS-number state (trainer.js:363, :368: gain = exp(g) and bias b, each
// rounded once to f32)
Cam make_cam_meta(const gs_trainer* h, const CamMeta& m, int activeDeg) {
  Cam k = make_cam(h, m.c, 1, m.offset, activeDeg);
  k.gain = (float)std::exp(m.g);
  k.bias = (float)m.b;
  return k;
}

// the f32 camera fields from the JS-number state (Float32Array stores)
void refresh_cam(CamMeta& m) {
  for (int k = 0; k < 9; k++) m.c.R[k] = (float)m.R[k];
  for (int k = 0; k < 3; k++) m.c.t[k] = (float)m.t[k];
  m.c.f = (float)m.f;
  m.c.g = (float)m.g;
  m.c.b = (float)m.b;
}

// ---- camera optimizer host math (src/sfm/geometry.js:108-114, :150-161) ----
void m3mul(const double* A, const double* B, double* C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[i * 3 + j] = (A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j]) + A[i * 3 + 2] * B[6 + j];
}

void rodrigues(const double* w, double* R) {
  const double th = std::sqrt((w[0] * w[0] + w[1] * w[1]) + w[2] * w[2]);
  if (th < 1e-12) {
    const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(R, I, sizeof I);
    return;
  }
  const double k[3] = {w[0] / th, w[1] / th, w[2] / th};
  const double c = std::cos(th), s = std::sin(th), v = 1 - c;
  R[0] = c + k[0] * k[0] * v;        R[1] = k[0] * k[1] * v - k[2] * s; R[2] = k[0] * k[2] * v + k[1] * s;
  R[3] = k[1] * k[0] * v + k[2] * s; R[4] = c + k[1] * k[1] * v;        R[5] = k[1] * k[2] * v - k[0] * s;
  R[6] = k[2] * k[0] * v - k[1] * s; R[7] = k[2] * k[1] * v + k[0] * s; R[8] = c + k[2] * k[2] * v;
}

// R <- rodrigues(-u[0..2]) R, t <- t - u[3..5] (trainer.js:517-523, :604-605)
void apply_pose_step(const double* u, double* R, double* t) {
  const double dw[3] = {-u[0], -u[1], -u[2]};
  double Rw[9], Rn[9];
  rodrigues(dw, Rw);
  m3mul(Rw, R, Rn);
  memcpy(R, Rn, sizeof Rn);
  for (int k = 0; k < 3; k++) t[k] = t[k] - u[3 + k];
}

double clampd(double x, double lo, double hi) { return std::max(lo, std::min(hi, x)); }

AdamHP make_hp(const gs_trainer* h) {
  // trainer.js:320-342, :417-424
  AdamHP a{};
  const double posLr = h->basePosLr * std::pow(0.01, std::min(1.0, (double)h->iter / (0.75 * h->horizon)));
  for (int k = 0; k < 3; k++) a.lr[k] = (float)posLr;
  for (int k = 3; k < 6; k++) a.lr[k] = 5e-3f;
  for (int k = 6; k < 10; k++) a.lr[k] = 1e-3f;
  for (int k = 10; k < 13; k++) a.lr[k] = 1.5e-2f;
  a.lr[13] = 2.5e-2f;
  a.lr[14] = a.lr[15] = 0.0f;
  a.b1 = 0.9f; a.b2 = 0.999f; a.eps = 1e-15f;
  a.rbc1 = (float)(1.0 / (1.0 - std::pow((double)0.9f, (double)h->adamT)));
  a.rbc2 = (float)(1.0 / (1.0 - std::pow((double)0.999f, (double)h->adamT)));
  a.step = (uint32_t)h->adamT;
  const double r = h->sceneRadius;
  a.minLS = (float)std::log(r * (double)h->o.minScale);
  a.maxLS = (float)std::log(r * 0.05);
  a.maxLogit = 8.0f;
  a.reg = h->o.opacityReg;
  a.shLr = h->o.shLr;
  a.anisoReg = h->o.anisoReg;
  return a;
}

// The Charbonnier loss's dL/dC = gain * err / sqrt(err^2 + d^2) is at most
// the exposure gain in magnitude (shaders.js:608-610).
bool charb_grad_bounded(const Cam& cam) { return std::fabs(cam.gain) <= 6.0e4f; }

Cut make_cut(const gs_trainer* h) {
  Cut c;
  c.eCut = h->o.eCut;
  c.aMin = h->o.aMin;
  c.radClamp = h->o.radClamp;
  return c;
}

// The depth-then-tile radix path (views of more than kMaxSegTiles tiles)
// ends its tile sort in a.tvA after an even number of passes and in a.tvB
// after an odd one; these set the two so that the sorted values land in
// `ids`, with `scratch` the other buffer.
void radix_values_into(BinArgs& a, uint32_t* ids, uint32_t* scratch) {
  const int passes = (tile_bits(a.numTiles) + 7) / 8;
  a.tvA = passes % 2 ? scratch : ids;
  a.tvB = passes % 2 ? ids : scratch;
}

// ---- one raster: project -> bin/sort -> forward (+ loss) ----
int run_raster(gs_trainer* h, const Cam& cam, int loss, double* statsBuf, bool withBwd = false) {
  const int64_t n = h->n;
  h->lastViews = 0;
  h->profLast = nullptr;   // a step or render starts its own first stage
  {
    Scope s(h, "project", (double)n * (56.0 + 4.0 * h->shC + 48.0 + 16.0));
    // the tile-segment binning balances its blocks by the keys per chunk of
    // Gaussians, which the projection sums on the way
    const bool segPath = h->o.sortMode == GS_SORT_TILE_SEGMENTS && cam.numTiles <= kMaxSegTiles;
    h->lastSegPath = segPath;
    h->lastTileCnt = h->sort.tileCnt;
    h->lastRanges = h->ranges;
    launch_project(h->st, h->shDeg, h->prec, n, h->cap, cam, make_cut(h), h->P, h->SH, h->rec, h->depthKey,
                   h->rect, h->cnt, segPath ? h->sort.chunkKeys : nullptr);
  }
  {
    BinArgs a{};
    a.mode = h->o.sortMode;
    a.tile = cam.tile;
    a.tilesX = cam.tilesX;
    a.numTiles = cam.numTiles;
    h->lastNumTiles = cam.numTiles;
    a.n = n;
    a.cap = (uint64_t)h->entriesCap;
    a.dN = h->dN;
    a.cnt = h->cnt;
    a.rect = h->rect;
    a.depthKey = h->depthKey;
    a.rec = h->rec;
    a.recFmt = h->recFmt;
    a.dkA = h->dkA; a.dvA = h->dvA; a.dkB = h->dkB; a.dvB = h->dvB;
    a.tkA = h->tkA; a.tkB = h->tkB; a.tvA = h->tvA; a.tvB = h->tvB;
    a.numKeys = h->numKeys;
    a.numVisible = h->numVisible;
    a.totalKeys = h->totalKeys;
    a.ranges = h->ranges;
    a.tileOrder = h->tileOrder;
    a.stats = statsBuf;
    a.scratch = h->sort;
    Scope s(h, "bin_sort", 0.0);
    launch_bin_sort(h->st, a, h->lastBin);
  }
  RasterBufs b{};
  b.ranges = h->ranges;
  b.tileOrder = h->lastBin.tileOrder;
  b.ids = h->lastBin.ids;
  b.rec = h->rec;
  b.recFmt = h->recFmt;
  b.img = h->img;
  b.T = h->Tfin;
  b.end = h->endIdx;
  b.gC = h->gC;
  b.tgt = h->targets;
  b.stats = statsBuf;
  if (withBwd) {
    Scope s(h, "raster_fwd_bwd", 0.0);
    launch_raster_fwd_bwd(h->st, h->shape, cam, make_cut(h), b, h->o.gradAccum, h->accum, charb_grad_bounded(cam));
  } else {
    Scope s(h, loss ? "raster_fwd_loss" : "raster_fwd", 0.0);
    launch_raster_fwd(h->st, h->shape, cam, make_cut(h), b, loss);
  }
  HIPCK(hipGetLastError());
  return GS_OK;
}

CamPoseBufs cam_pose_bufs(gs_trainer* h, int ci) {
  CamPoseBufs b{};
  b.n = h->n;
  b.cap = h->cap;
  b.P = h->P;
  b.accum = h->accum;
  b.cnt = h->cnt;
  b.row = h->camGrad + 8 * ci;
  b.focal = h->camGrad + 8 * h->cams.size();
  return b;
}

// GSTrainer._applyCamGrads (trainer.js:476-537): reads and clears the
// gradient rows, then one Adam step (JS numbers) on every camera pose but
// camera 0 (the gauge anchor) and the holdout, on the exposure when expComp
// is set, and on the shared log-focal. The reference applies the step when
// its asynchronous readback resolves, so the next few steps may still use
// the old uniforms; here it applies before the next step is enqueued.
int apply_cam_grads(gs_trainer* h) {
  const int nc = (int)h->cams.size();
  std::vector<int32_t> g((size_t)(nc + 1) * 8);
  HIPCK(hipMemcpyAsync(g.data(), h->camGrad, g.size() * 4, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipMemsetAsync(h->camGrad, 0, g.size() * 4, h->st));
  HIPCK(hipStreamSynchronize(h->st));
  h->camStep++;
  const double t = (double)h->camStep;
  const double decay = std::pow(0.02, std::min(1.0, (double)h->iter / (0.75 * h->horizon)));
  const double rotLr = 2e-4 * decay;
  const double trnLr = 2e-4 * (double)h->sceneRadius * decay;
  const double focLr = 1e-4 * decay;
  const double b1 = 0.9, b2 = 0.99, eps = 1e-15;
  auto step = [&](int row, int slot, double lr) {
    const size_t j = (size_t)row * 8 + slot;
    const double grad = g[j] / 64.0;
    h->camM[j] = b1 * h->camM[j] + (1 - b1) * grad;
    h->camV[j] = b2 * h->camV[j] + (1 - b2) * grad * grad;
    const double mh = h->camM[j] / (1 - std::pow(b1, t));
    const double vh = h->camV[j] / (1 - std::pow(b2, t));
    return lr * mh / (std::sqrt(vh) + eps);
  };
  const double expLr = 5e-3;
  for (int r = 1; r < nc; r++) {
    if (r == h->holdout) continue;
    CamMeta& m = h->cams[r];
    double u[6];
    for (int k = 0; k < 3; k++) u[k] = step(r, k, rotLr);
    for (int k = 3; k < 6; k++) u[k] = step(r, k, trnLr);
    apply_pose_step(u, m.R, m.t);
    if (h->o.expComp) {
      m.g = clampd(m.g - step(r, 6, expLr), -0.1, 0.1);
      m.b = clampd(m.b - step(r, 7, expLr), -0.05, 0.05);
    }
  }
  h->logfScale = clampd(h->logfScale - step(nc, 0, focLr), -0.3, 0.3);
  for (CamMeta& m : h->cams) {
    m.f = m.f0 * std::exp(h->logfScale);
    refresh_cam(m);
    m.activeDeg = h->shDeg;   // the rebuilt uniforms carry the full degree (trainer.js:369, :535)
  }
  return GS_OK;
}

// GSTrainer._evalPass (trainer.js:542-573) on camera ci at pose (R, t) and
// exposure (g, b): the training raster, loss and backward plus the camera
// terms of the chain, without any parameter update. Returns the view's PSNR
// and its camera gradient row (everything accumulated in the row, as the
// reference reads the buffer), then drains all rows and resets the training
// loss window, as the reference does.
int eval_pass(gs_trainer* h, int ci, const double* R, const double* t, double g, double b, double* psnr,
              int32_t* grad8) {
  CamMeta m = h->cams[ci];
  memcpy(m.R, R, sizeof m.R);
  memcpy(m.t, t, sizeof m.t);
  m.g = g;
  m.b = b;
  refresh_cam(m);
  const Cam cam = make_cam_meta(h, m, h->shDeg);   // _camUniform: full SH degree (trainer.js:369)
  HIPCK(hipMemsetAsync(h->evalStats, 0, 32, h->st));
  int rc = run_raster(h, cam, 1, h->evalStats, true);   // forward, loss and backward
  if (rc) return rc;
  launch_cam_exposure(h->st, cam, h->img, h->targets, h->camGrad + 8 * ci);
  launch_cam_pose(h->st, h->o.gradAccum, true, cam, cam_pose_bufs(h, ci));
  HIPCK(hipGetLastError());
  double s[4];
  int32_t row[8];
  const size_t rows = h->cams.size() + 1;
  HIPCK(hipMemcpyAsync(s, h->evalStats, 32, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipMemcpyAsync(row, h->camGrad + 8 * ci, 32, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipMemsetAsync(h->camGrad, 0, rows * 32, h->st));
  HIPCK(hipMemsetAsync(h->stats, 0, 24, h->st));   // the loss window; the budget drops keep counting
  HIPCK(hipStreamSynchronize(h->st));
  h->pixelsSeen = 0;
  if (grad8) memcpy(grad8, row, sizeof row);
  if (psnr) {
    const double mse = s[0] / (std::max(1.0, s[2]) * 3.0);
    *psnr = mse > 0 ? -10.0 * std::log10(mse) : INFINITY;
  }
  return GS_OK;
}

int pick_camera(gs_trainer* h) {
  // trainer.js:400-406 (Math.random replaced by a seeded mulberry32)
  const int nc = (int)h->cams.size();
  int ci = (int)((double)mulberry32(h->rng) / 4294967296.0 * nc);
  if (ci >= nc) ci = nc - 1;
  for (int tries = 0; tries < 16; tries++) {
    const bool bad = ci == h->holdout || (!h->excluded.empty() && h->excluded[ci]);
    if (!bad) break;
    ci = (ci + 1) % nc;
  }
  return ci;
}

int step_impl(gs_trainer* h, int ci, gs_debug_out* dbg) {
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (ci < 0) ci = pick_camera(h);
  if (ci >= (int)h->cams.size()) return fail(h, GS_E_ARG, "camera index out of range");
  CamMeta& meta = h->cams[ci];
  h->lastCam = ci;
  // trainer.js:409-412: one SH band per 1000 iterations
  meta.activeDeg = (int)std::min<int64_t>(h->shDeg, h->iter / 1000);
  const Cam cam = make_cam_meta(h, meta, meta.activeDeg);
  h->iter++;
  h->adamT++;
  h->pixelsSeen += (int64_t)meta.c.w * meta.c.h;
  const AdamHP hp = make_hp(h);

  const bool ssim = h->o.loss == GS_LOSS_L1_DSSIM;
  // the Charbonnier loss is per pixel, so its backward runs in the forward's launch
  int rc = run_raster(h, cam, ssim ? 2 : 1, h->stats, !ssim);
  if (rc) return rc;
  if (ssim) {
    RasterBufs lb{};
    lb.img = h->img;
    lb.gC = h->gC;
    lb.tgt = h->targets;
    lb.stats = h->stats;
    Scope sc(h, "loss_l1ssim", (double)cam.W * cam.H * 116.0);
    launch_loss_l1ssim(h->st, cam, h->o.ssimLambda, lb, h->ssimScratch);
  }
  if (h->o.camOpt && !ssim) {
    // exposure terms of the Charbonnier loss (shaders.js:614-616)
    Scope sc(h, "cam_exposure", (double)cam.W * cam.H * 20.0);
    launch_cam_exposure(h->st, cam, h->img, h->targets, h->camGrad + 8 * ci);
  }
  HIPCK(hipGetLastError());
  const int64_t npx = (int64_t)cam.W * cam.H;
  if (dbg) {
    HIPCK(hipStreamSynchronize(h->st));
    uint32_t nk = 0;
    HIPCK(hipMemcpy(&nk, h->numKeys, 4, hipMemcpyDeviceToHost));
    if (dbg->rec && h->recFmt == 0) {
      HIPCK(hipMemcpy(dbg->rec, h->rec, 64 * h->n, hipMemcpyDeviceToHost));
    } else if (dbg->rec) {
      // 48 B records with 16-bit colours, written out in the 16-float layout:
      // (mx, my, A, B, C, opa*comp, L, 0, r, g, b, opa, comp, -0.5 cnorm, 1 - opa, 0)
      // with the colours and L decoded
      std::vector<float> r((size_t)12 * h->n);
      HIPCK(hipMemcpy(r.data(), h->rec, 48 * h->n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < h->n; i++) {
        const float* q = &r[12 * (size_t)i];
        float* o = dbg->rec + 16 * i;
        uint32_t rg, bl;
        memcpy(&rg, &q[6], 4);
        memcpy(&bl, &q[7], 4);
        auto dec = [&](uint32_t b) { return h->recFmt == 2 ? gs_h2f(b & 0xffffu) : gs_bf2f(b); };
        const float v[16] = {q[0], q[1], q[2], q[3], q[4], q[5], gs_h2f(bl >> 16), 0.0f,
                             dec(rg), dec(rg >> 16), dec(bl), q[11], q[8], q[9], q[10], 0.0f};
        memcpy(o, v, sizeof v);
      }
    }
    std::vector<uint32_t> dk(h->n);
    HIPCK(hipMemcpy(dk.data(), h->depthKey, 4 * h->n, hipMemcpyDeviceToHost));
    if (dbg->depthKey) memcpy(dbg->depthKey, dk.data(), 4 * h->n);
    if (dbg->rect) {
      std::vector<uint2> r(h->n);
      HIPCK(hipMemcpy(r.data(), h->rect, 8 * h->n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < h->n; i++) {
        dbg->rect[4 * i + 0] = r[i].x & 0xffff; dbg->rect[4 * i + 1] = r[i].x >> 16;
        dbg->rect[4 * i + 2] = r[i].y & 0xffff; dbg->rect[4 * i + 3] = r[i].y >> 16;
      }
    }
    if (dbg->cnt) HIPCK(hipMemcpy(dbg->cnt, h->cnt, 4 * h->n, hipMemcpyDeviceToHost));
    if (dbg->numKeys) *dbg->numKeys = nk;
    if (dbg->ids || dbg->keys) {
      std::vector<uint32_t> ids(nk ? nk : 1);
      if (nk) HIPCK(hipMemcpy(ids.data(), h->lastBin.ids, 4ull * nk, hipMemcpyDeviceToHost));
      for (auto& v : ids) v &= kIdMask;   // block masks ride in the top bits
      if (dbg->ids) memcpy(dbg->ids, ids.data(), 4ull * nk);
      if (dbg->keys && nk) {
        if (h->lastBin.keyBytes == 8) {
          HIPCK(hipMemcpy(dbg->keys, h->lastBin.keys, 8ull * nk, hipMemcpyDeviceToHost));
        } else if (h->lastBin.keyBytes == 0) {
          // tile segments: a key's tile is the range that holds it
          std::vector<uint2> rg(cam.numTiles);
          HIPCK(hipMemcpy(rg.data(), h->ranges, 8ull * cam.numTiles, hipMemcpyDeviceToHost));
          for (int t = 0; t < cam.numTiles; t++)
            for (uint32_t k = rg[t].x; k < rg[t].y && k < nk; k++) dbg->keys[k] = ((uint64_t)t << 32) | dk[ids[k]];
        } else {
          std::vector<uint32_t> tk(nk);
          HIPCK(hipMemcpy(tk.data(), h->lastBin.keys, 4ull * nk, hipMemcpyDeviceToHost));
          for (uint32_t k = 0; k < nk; k++) dbg->keys[k] = ((uint64_t)tk[k] << 32) | dk[ids[k]];
        }
      }
    }
    if (dbg->tileRanges) HIPCK(hipMemcpy(dbg->tileRanges, h->ranges, 8ull * cam.numTiles, hipMemcpyDeviceToHost));
    if (dbg->img) HIPCK(hipMemcpy(dbg->img, h->img, 16 * npx, hipMemcpyDeviceToHost));
    if (dbg->T) HIPCK(hipMemcpy(dbg->T, h->Tfin, 4 * npx, hipMemcpyDeviceToHost));
    if (dbg->end) HIPCK(hipMemcpy(dbg->end, h->endIdx, 4 * npx, hipMemcpyDeviceToHost));
    if (dbg->gC) HIPCK(hipMemcpy(dbg->gC, h->gC, 16 * npx, hipMemcpyDeviceToHost));
  }
  RasterBufs b{};
  b.ranges = h->ranges;
  b.tileOrder = h->lastBin.tileOrder;
  b.ids = h->lastBin.ids;
  b.rec = h->rec;
  b.recFmt = h->recFmt;
  b.T = h->Tfin;
  b.end = h->endIdx;
  b.gC = h->gC;
  if (ssim) {
    Scope s(h, "raster_bwd", 0.0);
    launch_raster_bwd(h->st, h->shape, cam, make_cut(h), b, h->o.gradAccum, h->accum, false);
  }
  HIPCK(hipGetLastError());
  float* dGF = nullptr;
  float* dSG = nullptr;
  if (dbg) {
    HIPCK(hipStreamSynchronize(h->st));
    if (dbg->accum) {
      std::vector<int32_t> a(12 * h->n);
      HIPCK(hipMemcpy(a.data(), h->accum, 48 * h->n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < h->n; i++)
        for (int k = 0; k < 10; k++) dbg->accum[10 * i + k] = a[12 * i + k];
    }
    if (dbg->gradF || dbg->shGrad) {
      HIPCK(hipMalloc(&dGF, 64 * h->n));
      HIPCK(hipMalloc(&dSG, 4 * (h->shC ? h->shC : 1) * h->n));
    }
  }
  if (h->o.camOpt) {
    // pose and focal terms of the chain (shaders.js:834-858), read before
    // the update pass clears the accumulators
    Scope sc(h, "cam_pose", (double)h->n * (4.0 + 32.0 + 64.0));
    launch_cam_pose(h->st, h->o.gradAccum, false, cam, cam_pose_bufs(h, ci));
  }
  UpdateBufs u{};
  u.n = h->n;
  u.cap = h->cap;
  u.P = h->P; u.M = h->M; u.V = h->V;
  u.SH = h->SH; u.SHM = h->SHM; u.SHV = h->SHV;
  u.accum = h->accum;
  u.cnt = h->cnt;
  u.dbgGradF = dGF;
  u.dbgShGrad = dSG;
  {
    Scope s(h, "update", (double)h->n * (100.0 + 6 * 64.0 + 6 * 4.0 * h->shC));
    launch_update(h->st, h->shDeg, h->o.gradAccum, h->prec, cam, hp, u);
  }
  HIPCK(hipGetLastError());
  if (dbg) {
    HIPCK(hipStreamSynchronize(h->st));
    if (dGF) {
      if (dbg->gradF) HIPCK(hipMemcpy(dbg->gradF, dGF, 64 * h->n, hipMemcpyDeviceToHost));
      if (dbg->shGrad && h->shC) HIPCK(hipMemcpy(dbg->shGrad, dSG, 4ull * h->shC * h->n, hipMemcpyDeviceToHost));
      (void)hipFree(dGF);
      (void)hipFree(dSG);
    }
    if (dbg->stats) HIPCK(hipMemcpy(dbg->stats, h->stats, 32, hipMemcpyDeviceToHost));
  }
  // trainer.js:468-470: one camera Adam step every 25 iterations after 1500
  if (h->o.camOpt && h->iter > 1500 && h->iter % 25 == 0) return apply_cam_grads(h);
  return GS_OK;
}

// ---- batched views (gs_step_views) ----

// The launches of one batched step. Every per-step value (cameras, Adam
// constants, the work offsets) comes from the device ViewBlock, so the same
// launches serve eager steps and a captured graph; what the host fixes here
// is the shape: n, each view's tile grid and the colour-clamp choice.
// dbg (eager only): copy the views' intermediates between the stages.
// statsBuf: the loss window and budget drops (the views' RasterBufs.stats
// must point there too); update false: no chain and Adam (the gradient sums
// stay in the views' accumulators).
int enqueue_views(gs_trainer* h, const ViewBlock& vb, gs_debug_out* dbg, float* dGF, float* dSG, double* statsBuf,
                  bool update) {
  const int V = vb.count;
  const int64_t n = h->n;
  const Cut cut = make_cut(h);
  // Overlapped form: each view's forward and backward run on the second
  // stream once its binning is done, beside the next views' binning; the
  // debug and profiling passes use the persistent form (one stream, so the
  // stages can be read and timed one by one).
  const bool overlap = h->o.viewsLaunch == GS_VIEWS_OVERLAP && !dbg && !h->prof;
  bool bounded = true;
  for (int k = 0; k < V; k++) bounded = bounded && charb_grad_bounded(vb.cams[k]);
  if (overlap) {
    HIPCK(hipEventRecord(h->evFork, h->st));
    HIPCK(hipStreamWaitEvent(h->st2, h->evFork, 0));
  } else {
    launch_fill_u32(h->st, h->dQueue, 1, 0u);
  }
  if (h->batchedBin && !overlap) {
    // every view's projection, then each binning stage of all views, one launch each
    {
      Scope sc(h, "project_views", (double)n * (56.0 + 4.0 * h->shC) + (double)V * n * (48.0 + 16.0));
      launch_project_views(h->st, h->shDeg, h->prec, n, h->cap, h->dVB, V, cut, h->P, h->SH);
    }
    {
      Scope sc(h, "bin_sort_views", 0.0);
      launch_bin_sort_views(h->st, (const SegJob*)((const char*)h->dVB + offsetof(ViewBlock, seg)), vb.seg, V);
      // views above the segment path's tiles, one after another by the
      // depth-then-tile radix path into their own slots (its key buffers are
      // slot 0's segment keys, free again once the launches above are done)
      for (int k = 0; k < V; k++) {
        if (!vb.seg[k].skip) continue;
        const gs_trainer::ViewSlot& sl = h->vslots[k];
        BinArgs a{};
        a.mode = GS_SORT_TILE_SEGMENTS;
        a.tile = vb.cams[k].tile;
        a.tilesX = vb.cams[k].tilesX;
        a.numTiles = vb.cams[k].numTiles;
        a.n = n;
        a.cap = (uint64_t)h->entriesCap;
        a.dN = h->dN;
        a.cnt = sl.cnt;
        a.rect = sl.rect;
        a.depthKey = sl.depthKey;
        a.rec = sl.rec;
        a.recFmt = h->recFmt;
        a.dkA = h->dkA; a.dvA = h->dvA; a.dkB = h->dkB; a.dvB = h->dvB;
        a.tkA = h->tkA; a.tkB = h->tkB;
        radix_values_into(a, sl.ids, h->tvB);
        a.numKeys = h->viewStats + k;
        a.numVisible = h->viewStats + kMaxViews + k;
        a.totalKeys = sl.totalKeys;
        a.ranges = sl.ranges;
        a.tileOrder = sl.tileOrder;
        a.stats = statsBuf;
        a.scratch = h->sort;
        BinOut out{};
        launch_bin_sort(h->st, a, out);
      }
    }
    if (dbg) {
      HIPCK(hipStreamSynchronize(h->st));
      for (int k = 0; k < V; k++) {
        uint32_t nk = 0;
        HIPCK(hipMemcpy(&nk, h->viewStats + k, 4, hipMemcpyDeviceToHost));
        if (dbg[k].numKeys) *dbg[k].numKeys = nk;
        if (dbg[k].tileRanges)
          HIPCK(hipMemcpy(dbg[k].tileRanges, h->vslots[k].ranges, 8ull * vb.cams[k].numTiles, hipMemcpyDeviceToHost));
      }
    }
  }
  for (int k = 0; k < V && !(h->batchedBin && !overlap); k++) {
    const gs_trainer::ViewSlot& sl = h->vslots[k];
    const Cam& cam = vb.cams[k];
    {
      Scope sc(h, "project", (double)n * (56.0 + 4.0 * h->shC + 48.0 + 16.0));
      launch_project_view(h->st, h->shDeg, h->prec, n, h->cap, h->dVB, k, cut, h->P, h->SH, sl.rec, h->depthKey,
                          h->rect, sl.cnt, cam.numTiles <= kMaxSegTiles ? h->sort.chunkKeys : nullptr);
    }
    BinArgs a{};
    a.mode = GS_SORT_TILE_SEGMENTS;
    a.tile = cam.tile;
    a.tilesX = cam.tilesX;
    a.numTiles = cam.numTiles;
    a.n = n;
    a.cap = (uint64_t)h->entriesCap;
    a.dN = h->dN;
    a.cnt = sl.cnt;
    a.rect = h->rect;
    a.depthKey = h->depthKey;
    a.rec = sl.rec;
    a.recFmt = h->recFmt;
    a.dkA = h->dkA; a.dvA = h->dvA; a.dkB = h->dkB; a.dvB = h->dvB;
    a.tkA = h->tkA; a.tkB = h->tkB; a.tvA = sl.ids; a.tvB = h->tvB;
    // above the segment path's tiles launch_bin_sort takes the radix path
    if (cam.numTiles > kMaxSegTiles) radix_values_into(a, sl.ids, h->tvB);
    a.numKeys = h->viewStats + k;
    a.numVisible = h->viewStats + kMaxViews + k;
    a.totalKeys = h->totalKeys;
    a.ranges = sl.ranges;
    a.tileOrder = sl.tileOrder;
    a.stats = statsBuf;
    a.scratch = h->sort;
    {
      Scope sc(h, "bin_sort", 0.0);
      BinOut out{};
      launch_bin_sort(h->st, a, out);
    }
    if (dbg) {
      HIPCK(hipStreamSynchronize(h->st));
      uint32_t nk = 0;
      HIPCK(hipMemcpy(&nk, h->viewStats + k, 4, hipMemcpyDeviceToHost));
      if (dbg[k].numKeys) *dbg[k].numKeys = nk;
      if (dbg[k].tileRanges)
        HIPCK(hipMemcpy(dbg[k].tileRanges, sl.ranges, 8ull * cam.numTiles, hipMemcpyDeviceToHost));
    }
    if (overlap) {
      HIPCK(hipEventRecord(h->evBin[k], h->st));
      HIPCK(hipStreamWaitEvent(h->st2, h->evBin[k], 0));
      launch_raster_fwd_bwd_view(h->st2, h->shape, h->dVB, k, cam.numTiles, cut, h->o.gradAccum, bounded);
    }
  }
  h->lastNumTiles = vb.cams[V - 1].numTiles;
  h->lastSegPath = vb.cams[V - 1].numTiles <= kMaxSegTiles;
  // the last view's raw tile counts: its own slot's with batched binning,
  // the shared buffer (which it binned last) otherwise
  h->lastTileCnt = h->batchedBin && !overlap ? h->vslots[V - 1].tileCnt : h->sort.tileCnt;
  h->lastRanges = h->vslots[V - 1].ranges;
  h->lastViews = V;
  if (overlap) {
    HIPCK(hipEventRecord(h->evJoin, h->st2));
    HIPCK(hipStreamWaitEvent(h->st, h->evJoin, 0));
  } else {
    // each tile's backward right after its forward, in one persistent launch
    Scope sc(h, "raster_fwd_bwd_views", 0.0);
    launch_raster_fwd_bwd_views(h->st, h->shape, h->dVB, cut, h->o.gradAccum, bounded, h->dQueue, h->viewBlocks);
  }
  if (dbg) {
    HIPCK(hipStreamSynchronize(h->st));
    for (int k = 0; k < V; k++) {
      const gs_trainer::ViewSlot& sl = h->vslots[k];
      const int64_t npx = (int64_t)vb.cams[k].W * vb.cams[k].H;
      if (dbg[k].img) HIPCK(hipMemcpy(dbg[k].img, sl.img, 16 * npx, hipMemcpyDeviceToHost));
      if (dbg[k].T) HIPCK(hipMemcpy(dbg[k].T, sl.T, 4 * npx, hipMemcpyDeviceToHost));
      if (dbg[k].end) HIPCK(hipMemcpy(dbg[k].end, sl.end, 4 * npx, hipMemcpyDeviceToHost));
      if (dbg[k].gC) HIPCK(hipMemcpy(dbg[k].gC, sl.gC, 16 * npx, hipMemcpyDeviceToHost));
    }
  }
  if (dbg) {
    HIPCK(hipStreamSynchronize(h->st));
    for (int k = 0; k < V; k++) {
      if (!dbg[k].accum) continue;
      std::vector<int32_t> acc(12 * h->n);
      HIPCK(hipMemcpy(acc.data(), h->vslots[k].accum, 48 * h->n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < h->n; i++)
        for (int q = 0; q < 10; q++) dbg[k].accum[10 * i + q] = acc[12 * i + q];
    }
  }
  if (update) {
    UpdateBufs u{};
    u.n = n;
    u.cap = h->cap;
    u.P = h->P; u.M = h->M; u.V = h->V;
    u.SH = h->SH; u.SHM = h->SHM; u.SHV = h->SHV;
    u.dbgGradF = dGF;
    u.dbgShGrad = dSG;
    Scope sc(h, "update_views", (double)n * (100.0 + 6 * 64.0 + 6 * 4.0 * h->shC));
    launch_update_views(h->st, h->shDeg, h->o.gradAccum, h->prec, h->dVB, V, u);
  }
  HIPCK(hipGetLastError());
  return GS_OK;
}

// Fills vb's views from cameras ci[0..count) at SH degree activeDeg (the
// raster statistics to statsBuf) and appends the step's shape to key.
int fill_views(gs_trainer* h, const int* ci, int count, int activeDeg, double* statsBuf, ViewBlock& vb,
               std::vector<int64_t>& key) {
  memset(&vb, 0, sizeof vb);
  vb.count = count;
  uint32_t off = 0;
  for (int k = 0; k < count; k++) {
    CamMeta& meta = h->cams[ci[k]];
    meta.activeDeg = activeDeg;
    vb.cams[k] = make_cam_meta(h, meta, activeDeg);
    const gs_trainer::ViewSlot& sl = h->vslots[k];
    RasterBufs& b = vb.bufs[k];
    b.recFmt = h->recFmt;
    b.ranges = sl.ranges;
    b.tileOrder = sl.tileOrder;
    b.ids = sl.ids;
    b.rec = sl.rec;
    b.img = sl.img;
    b.T = sl.T;
    b.end = sl.end;
    b.gC = sl.gC;
    b.tgt = h->targets;
    b.stats = statsBuf;
    vb.accum[k] = sl.accum;
    vb.cnt[k] = sl.cnt;
    if (h->batchedBin) {
      SegJob& j = vb.seg[k];
      j.n = h->n;
      j.cap = (uint64_t)h->entriesCap;
      j.nb = seg_blocks(h->n);
      j.tilesX = vb.cams[k].tilesX;
      j.tile = vb.cams[k].tile;
      j.numTiles = vb.cams[k].numTiles;
      // above the segment path's tiles the view bins by the radix path
      // (enqueue_views), which needs no chunk sums from the projection
      j.skip = vb.cams[k].numTiles > kMaxSegTiles;
      j.recFmt = h->recFmt;
      j.rec = sl.rec;
      j.depthKey = sl.depthKey;
      j.rect = sl.rect;
      j.cnt = sl.cnt;
      j.chunkKeys = j.skip ? nullptr : sl.chunkKeys;
      j.blockStart = sl.blockStart;
      j.table = sl.table;
      j.visPart = sl.visPart;
      j.tileCnt = sl.tileCnt;
      j.ranges = sl.ranges;
      j.tileOrder = sl.tileOrder;
      j.numKeys = h->viewStats + k;
      j.numVisible = h->viewStats + kMaxViews + k;
      j.totalKeys = sl.totalKeys;
      j.stats = statsBuf;
      j.keys = sl.keys;
      j.tmp = sl.tmp;
      j.ids = sl.ids;
    }
    vb.off[k] = off;
    off += (uint32_t)vb.cams[k].numTiles;
    key.push_back(vb.cams[k].W);
    key.push_back(vb.cams[k].H);
    if (!charb_grad_bounded(vb.cams[k])) key[2] = 1;
  }
  vb.off[count] = off;
  return GS_OK;
}

int step_views_impl(gs_trainer* h, const int32_t* cams, int count, gs_debug_out* dbg) {
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (count < 1 || count > (int)h->vslots.size())
    return fail(h, GS_E_ARG, "stepViews: " + std::to_string(count) + " views; opts.viewsPerStep allows 1.." +
                                 std::to_string(h->vslots.size()));
  if (h->vslots.size() < 2 || !h->dVB)
    return fail(h, GS_E_ARG, "stepViews needs opts.viewsPerStep > 1 at create");
  if (h->o.loss != GS_LOSS_CHARBONNIER || h->o.sortMode != GS_SORT_TILE_SEGMENTS || h->o.camOpt)
    return fail(h, GS_E_ARG, "stepViews needs the Charbonnier loss, the tile-segment sort and camOpt off");
  // the view list, picked as `count` stepOnce calls would pick it
  int ci[kMaxViews];
  for (int k = 0; k < count; k++) {
    ci[k] = cams && cams[k] >= 0 ? cams[k] : pick_camera(h);
    if (ci[k] >= (int)h->cams.size()) return fail(h, GS_E_ARG, "stepViews: camera index out of range");
  }
  // the ring entry this step fills: its previous copy must have landed
  ViewBlock& vb = h->hVB[h->ringPos];
  hipEvent_t ev = h->ringEv[(size_t)h->ringPos];
  HIPCK(hipEventSynchronize(ev));
  const int activeDeg = (int)std::min<int64_t>(h->shDeg, h->iter / 1000);   // trainer.js:409-412
  std::vector<int64_t> key = {count, h->n, 0, h->shape.tile, h->shape.batch};
  int frc = fill_views(h, ci, count, activeDeg, h->stats, vb, key);
  if (frc) return frc;
  h->iter += count;
  h->adamT += 1;
  for (int k = 0; k < count; k++) h->pixelsSeen += (int64_t)vb.cams[k].W * vb.cams[k].H;
  h->lastCam = ci[count - 1];
  vb.hp = make_hp(h);
  HIPCK(hipMemcpyAsync(h->dVB, &vb, sizeof vb, hipMemcpyHostToDevice, h->st));
  HIPCK(hipEventRecord(ev, h->st));
  h->ringPos = (h->ringPos + 1) % kRing;
  h->profLast = nullptr;

  float* dGF = nullptr;
  float* dSG = nullptr;
  if (dbg && (dbg[0].gradF || dbg[0].shGrad)) {
    HIPCK(hipMalloc(&dGF, 64 * h->n));
    HIPCK(hipMalloc(&dSG, 4 * (h->shC ? h->shC : 1) * h->n));
  }
  int rc = GS_OK;
  // A shape is captured on its second step in a row: the first runs eagerly
  // (one-time host set-up of the launches happens outside any capture).
  const bool useGraph = h->o.graph && !dbg && !h->prof;
  if (useGraph && h->gExec && key == h->gKey) {
    HIPCK(hipGraphLaunch(h->gExec, h->st));
  } else if (useGraph && key == h->lastEagerKey) {
    if (h->gExec) {
      (void)hipGraphExecDestroy(h->gExec);
      h->gExec = nullptr;
    }
    HIPCK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    rc = enqueue_views(h, vb, nullptr, nullptr, nullptr, h->stats, true);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(h->st, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hipfail(h, "stepViews: graph capture", e);
    const hipError_t ei = hipGraphInstantiate(&h->gExec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) { h->gExec = nullptr; return hipfail(h, "stepViews: graph instantiate", ei); }
    h->gKey = key;
    HIPCK(hipGraphLaunch(h->gExec, h->st));
  } else {
    rc = enqueue_views(h, vb, dbg, dGF, dSG, h->stats, true);
    h->lastEagerKey = key;
  }
  if (dbg && rc == GS_OK) {
    HIPCK(hipStreamSynchronize(h->st));
    if (dGF && dbg[0].gradF) HIPCK(hipMemcpy(dbg[0].gradF, dGF, 64 * h->n, hipMemcpyDeviceToHost));
    if (dSG && dbg[0].shGrad && h->shC)
      HIPCK(hipMemcpy(dbg[0].shGrad, dSG, 4ull * h->shC * h->n, hipMemcpyDeviceToHost));
    if (dbg[0].stats) HIPCK(hipMemcpy(dbg[0].stats, h->stats, 32, hipMemcpyDeviceToHost));
  }
  if (dGF) (void)hipFree(dGF);
  if (dSG) (void)hipFree(dSG);
  return rc;
}

void repack_to_planes(const float* src, int64_t n, int width, int planes, int64_t cap,
                      std::vector<float>& dst) {
  // src: n x width (row-major); dst: planes x cap x 4
  dst.assign((size_t)planes * cap * 4, 0.0f);
  for (int64_t i = 0; i < n; i++)
    for (int j = 0; j < width; j++) dst[((size_t)(j / 4) * cap + i) * 4 + (j % 4)] = src[i * width + j];
}

// the no-allocation export's device scratch, looked up per device (below)
int scratch_init(int device);

}  // namespace

extern "C" {

int gs_default_opts(gs_opts* o) {
  if (!o) return GS_E_ARG;
  memset(o, 0, sizeof(*o));
  o->size = (int32_t)sizeof(gs_opts);
  o->eCut = 4.5f;
  o->aMin = 0.0039f;
  o->radClamp = 1.0f;
  o->shDeg = 2;
  o->anisoReg = 0.005f;
  o->opacityReg = 0.01f;
  o->shLr = 1.25e-4f;
  o->minScale = 1e-4f;
  o->maxIters = 60000;
  o->maxSplats = 0;   // unset: capacity at most 600000, budget 12M
  o->capMult = 4.0f;
  o->entriesCap = 0;
  o->tileGrad = 1;
  o->viewsPerStep = 1;
  o->graph = 1;
  o->viewsLaunch = GS_VIEWS_PERSISTENT;
  o->tile = kDefaultTile;
  o->batch = kDefaultBatch;
  o->posLr = 0.0f;
  o->loss = GS_LOSS_CHARBONNIER;
  o->ssimLambda = 0.2f;
  o->gradAccum = GS_ACCUM_FIXED;
  o->sortMode = GS_SORT_TILE_SEGMENTS;
  o->seed = 0;
  o->device = 0;
  o->precision = GS_PREC_F32;
  return GS_OK;
}

int gs_create(const gs_opts* o, gs_trainer_t** out) {
  if (!out) return GS_E_ARG;
  *out = nullptr;
  gs_trainer* h = new gs_trainer();
  gs_default_opts(&h->o);
  if (o) {
    if (o->size != (int32_t)sizeof(gs_opts)) {
      delete h;
      return GS_E_ARG;
    }
    h->o = *o;
  }
  if (h->o.shDeg < 0 || h->o.shDeg > 3 || h->o.sortMode < 0 || h->o.sortMode > 2 ||
      h->o.gradAccum < 0 || h->o.gradAccum > 1 || h->o.loss < 0 || h->o.loss > 1 ||
      h->o.precision < 0 || h->o.precision > 3 ||
      h->o.viewsPerStep < 1 || h->o.viewsPerStep > kMaxViews ||
      h->o.viewsLaunch < 0 || h->o.viewsLaunch > 1 ||
      !raster_shape_ok(h->o.tile ? h->o.tile : kDefaultTile, h->o.batch ? h->o.batch : kDefaultBatch) ||
      !(h->o.eCut >= 0.0f && h->o.eCut <= 80.0f)) {
    delete h;
    return GS_E_ARG;
  }
  h->shape = RasterShape{h->o.tile ? h->o.tile : kDefaultTile, h->o.batch ? h->o.batch : kDefaultBatch};
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || h->o.device >= ndev) {
    (void)hipGetLastError();
    delete h;
    return GS_E_HIP;
  }
  if (hipSetDevice(h->o.device) != hipSuccess || hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return GS_E_HIP;
  }
  // this device's address of the no-allocation export's scratch array
  if (scratch_init(h->o.device) != GS_OK) {
    (void)hipStreamDestroy(h->st);
    delete h;
    return GS_E_HIP;
  }
  // batched views: the second stream and the events that order it
  if (h->o.viewsPerStep > 1) {
    bool ok = hipStreamCreateWithFlags(&h->st2, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&h->evFork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&h->evJoin, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < kMaxViews && ok; k++)
      ok = hipEventCreateWithFlags(&h->evBin[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
      gs_destroy(h);
      return GS_E_HIP;
    }
  }
  h->shDeg = h->o.shDeg;
  h->shK = h->shDeg > 0 ? (h->shDeg + 1) * (h->shDeg + 1) - 1 : 0;
  h->shC = 3 * h->shK;
  h->shP4 = (h->shC + 3) / 4;
  h->shQ8 = (h->shC + 7) / 8;
  h->prec = h->o.precision;
  h->shFmt = h->prec == GS_PREC_F32 ? 0 : (h->prec == GS_PREC_FP16 ? 2 : 1);
  h->mom16 = h->prec == GS_PREC_MIXED;
  h->recFmt = h->prec == GS_PREC_BF16 ? 1 : (h->prec == GS_PREC_FP16 ? 2 : 0);
  h->rng = h->o.seed;
  *out = h;
  return GS_OK;
}

void gs_destroy(gs_trainer_t* h) {
  if (!h) return;
  (void)hipSetDevice(h->o.device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->st2) (void)hipStreamSynchronize(h->st2);
  free_all(h);
  for (auto& p : h->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : h->evPool) (void)hipEventDestroy(e);
  if (h->st2) (void)hipStreamDestroy(h->st2);
  if (h->evFork) (void)hipEventDestroy(h->evFork);
  if (h->evJoin) (void)hipEventDestroy(h->evJoin);
  for (hipEvent_t e : h->evBin)
    if (e) (void)hipEventDestroy(e);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
}

int gs_setup(gs_trainer_t* h, const float* params, int64_t n, const float* sh, const gs_camera* cams,
             int32_t ncams, const gs_image* images, int32_t nimages, int32_t maxViewW,
             int32_t maxViewH, float sceneRadius) {
  if (!h) return GS_E_ARG;
  if (!params || n <= 0 || !cams || ncams <= 0 || !images || nimages <= 0)
    return fail(h, GS_E_ARG, "setup: empty gaussians, cameras or images");
  HIPCK(hipSetDevice(h->o.device));
  HIPCK(hipStreamSynchronize(h->st));
  free_all(h);
  h->n = n;
  // trainer.js:105-107; capacity never below n
  const int64_t maxSplats = h->o.maxSplats > 0 ? h->o.maxSplats : 600000;
  h->cap = std::max<int64_t>(n, std::min<int64_t>(std::max<int64_t>((int64_t)std::floor(n * (double)h->o.capMult), n), maxSplats));
  if (h->cap > (int64_t)kIdMask + 1) return fail(h, GS_E_LIMIT, "setup: more than 2^28 Gaussians");
  // trainer.js:48-49: the budget follows maxSplats only when the caller set it
  h->entriesCap = h->o.entriesCap > 0 ? h->o.entriesCap
                  : h->o.maxSplats > 0 ? std::max<int64_t>(kDefaultEntries, 24 * h->o.maxSplats)
                                       : kDefaultEntries;
  h->entriesCap = std::min<int64_t>(h->entriesCap, 0xffffffffll);

  // targets (trainer.js:115-139)
  h->cams.clear();
  int64_t total = 0;
  for (int i = 0; i < ncams; i++) {
    const gs_camera& c = cams[i];
    if (c.imgIdx < 0 || c.imgIdx >= nimages) return fail(h, GS_E_ARG, "setup: camera imgIdx out of range");
    const gs_image& im = images[c.imgIdx];
    CamMeta m;
    m.c = c;
    m.c.w = im.tw;
    m.c.h = im.th;
    m.offset = total;
    m.activeDeg = h->shDeg;
    for (int k = 0; k < 9; k++) m.R[k] = c.R[k];
    for (int k = 0; k < 3; k++) m.t[k] = c.t[k];
    m.f0 = m.f = c.f;   // trainer.js:306
    m.g = c.g;
    m.b = c.b;
    total += (int64_t)im.tw * im.th;
    h->cams.push_back(m);
  }
  if (total * 4 > kTargetLimitBytes) {
    char buf[256];
    snprintf(buf, sizeof buf, "training targets (%.0fMB) exceed the device binding limit (%.0fMB) — reduce image count or resolution",
             total * 4 / 1e6, kTargetLimitBytes / 1e6);
    return fail(h, GS_E_LIMIT, buf);
  }
  h->totalPx = total;
  std::vector<uint32_t> tgt((size_t)std::max<int64_t>(total, 4));
  for (const CamMeta& m : h->cams) {
    const float* rgb = images[m.c.imgIdx].rgb;
    const int64_t np = (int64_t)m.c.w * m.c.h;
    for (int64_t p = 0; p < np; p++) {
      const double r = rgb[3 * p];
      if (r < 0) { tgt[m.offset + p] = 0; continue; }
      const uint32_t vb = (uint32_t)js_round255((double)rgb[3 * p + 2] * 255.0) << 16;
      const uint32_t vg = (uint32_t)js_round255((double)rgb[3 * p + 1] * 255.0) << 8;
      const uint32_t vr = (uint32_t)js_round255(r * 255.0);
      tgt[m.offset + p] = (255u << 24) | vb | vg | vr;
    }
  }
  h->maxPix = (int64_t)maxViewW * maxViewH;
  h->maxTiles = tiles_of(maxViewW, maxViewH, kMinTile);
  for (const CamMeta& m : h->cams) {
    h->maxPix = std::max<int64_t>(h->maxPix, (int64_t)m.c.w * m.c.h);
    h->maxTiles = std::max(h->maxTiles, tiles_of(m.c.w, m.c.h, kMinTile));
  }

  const int64_t cap = h->cap;
  const size_t pb = (size_t)cap * 64;
  int rc = 0;
  rc |= dalloc(h, &h->P, pb);
  const size_t mb = h->mom16 ? (size_t)cap * 32 : pb;   // 16 moments: 2 bf16 planes or 4 f32 planes
  rc |= dalloc(h, &h->M, mb);
  rc |= dalloc(h, &h->V, mb);
  // SH coefficients in f32 or 16-bit planes; SH moments in 16-bit planes only with mom16
  const size_t sb = (size_t)cap * 16 * std::max(h->shFmt ? h->shQ8 : h->shP4, 1);
  const size_t smb = (size_t)cap * 16 * std::max(h->mom16 ? h->shQ8 : h->shP4, 1);
  rc |= dalloc(h, &h->SH, sb);
  rc |= dalloc(h, &h->SHM, smb);
  rc |= dalloc(h, &h->SHV, smb);
  rc |= dalloc(h, &h->rec, (size_t)cap * 16 * rec_stride(h->recFmt));
  rc |= dalloc(h, &h->depthKey, (size_t)cap * 4);
  rc |= dalloc(h, &h->cnt, (size_t)cap * 4);
  rc |= dalloc(h, &h->rect, (size_t)cap * 8);
  rc |= dalloc(h, &h->accum, (size_t)cap * 48);
  rc |= dalloc(h, &h->dkA, (size_t)cap * 4);
  rc |= dalloc(h, &h->dvA, (size_t)cap * 4);
  rc |= dalloc(h, &h->dkB, (size_t)cap * 4);
  rc |= dalloc(h, &h->dvB, (size_t)cap * 4);
  const size_t kb = h->o.sortMode == GS_SORT_DEPTH_THEN_TILE ? 4 : 8;
  rc |= dalloc(h, &h->tkA, (size_t)h->entriesCap * kb);
  rc |= dalloc(h, &h->tkB, (size_t)h->entriesCap * kb);
  rc |= dalloc(h, &h->tvA, (size_t)h->entriesCap * 4);
  rc |= dalloc(h, &h->tvB, (size_t)h->entriesCap * 4);
  h->sort.G = kSortBlocks;
  rc |= dalloc(h, &h->sort.hist, (size_t)256 * kSortBlocks * 4);
  rc |= dalloc(h, &h->sort.totals, 256 * 4);
  const size_t gc = (size_t)((cap + 255) / 256);
  rc |= dalloc(h, &h->sort.partials, gc * 8);
  rc |= dalloc(h, &h->sort.vpartials, gc * 4);
  rc |= dalloc(h, &h->sort.cntS, (size_t)cap * 4);
  rc |= dalloc(h, &h->sort.rectS, (size_t)cap * 8);
  rc |= dalloc(h, &h->sort.dkS, (size_t)cap * 4);
  rc |= dalloc(h, &h->sort.nDepth, 4);
  rc |= dalloc(h, &h->sort.tileCnt, (size_t)h->maxTiles * 4);
  if (h->o.sortMode == GS_SORT_TILE_SEGMENTS) {
    // rows = ceil(n / seg_group(n)) <= min(kSegMaxRows, ceil(cap / kSegMinGroup)) for every n <= cap
    const int64_t rows = std::min<int64_t>(kSegMaxRows, (cap + kSegMinGroup - 1) / kSegMinGroup);
    // (the segment path takes views of at most kMaxSegTiles tiles)
    rc |= dalloc(h, &h->sort.segTable, (size_t)rows * std::min(h->maxTiles, kMaxSegTiles) * 4);
    rc |= dalloc(h, &h->sort.visPart, (size_t)rows * 4);
    rc |= dalloc(h, &h->sort.blockStart, (size_t)(rows + 1) * 4);
    const size_t chunkBytes = (size_t)((cap + kSegChunk - 1) / kSegChunk) * 4;
    rc |= dalloc(h, &h->sort.chunkKeys, chunkBytes);
    // the projection adds to these sums and the scan clears them after use
    if (rc == 0) HIPCK(hipMemset(h->sort.chunkKeys, 0, chunkBytes));
  }
  rc |= dalloc(h, &h->dN, 4);
  rc |= dalloc(h, &h->numKeys, 4);
  rc |= dalloc(h, &h->numVisible, 4);
  rc |= dalloc(h, &h->totalKeys, 8);
  rc |= dalloc(h, &h->ranges, (size_t)h->maxTiles * 8);
  rc |= dalloc(h, &h->tileOrder, (size_t)h->maxTiles * 4);
  rc |= dalloc(h, &h->img, (size_t)h->maxPix * 16);
  rc |= dalloc(h, &h->Tfin, (size_t)h->maxPix * 4);
  rc |= dalloc(h, &h->endIdx, (size_t)h->maxPix * 4);
  rc |= dalloc(h, &h->gC, (size_t)h->maxPix * 16);
  if (h->o.loss == GS_LOSS_L1_DSSIM) rc |= dalloc(h, &h->ssimScratch, (size_t)h->maxPix * 36);
  rc |= dalloc(h, &h->targets, tgt.size() * 4);
  rc |= dalloc(h, &h->stats, 32);
  rc |= dalloc(h, &h->evalStats, 32);
  rc |= dalloc(h, &h->camGrad, (size_t)(ncams + 1) * 32);
  // batched views: slot 0 is the single-view set, slots 1.. their own copies
  h->vslots.assign((size_t)h->o.viewsPerStep, gs_trainer::ViewSlot{});
  h->vslots[0] = {h->rec, h->cnt, h->accum, h->tvA, h->ranges, h->tileOrder, h->img, h->Tfin, h->endIdx, h->gC,
                  h->depthKey, h->rect, h->sort.chunkKeys, h->sort.blockStart, h->sort.segTable, h->sort.visPart,
                  h->sort.tileCnt, (uint64_t*)h->tkA, (uint64_t*)h->tkB, h->totalKeys};
  for (size_t k = 1; k < h->vslots.size(); k++) {
    gs_trainer::ViewSlot& v = h->vslots[k];
    rc |= dalloc(h, &v.rec, (size_t)cap * 16 * rec_stride(h->recFmt));
    rc |= dalloc(h, &v.cnt, (size_t)cap * 4);
    rc |= dalloc(h, &v.accum, (size_t)cap * 48);
    rc |= dalloc(h, &v.ids, (size_t)h->entriesCap * 4);
    rc |= dalloc(h, &v.ranges, (size_t)h->maxTiles * 8);
    rc |= dalloc(h, &v.tileOrder, (size_t)h->maxTiles * 4);
    rc |= dalloc(h, &v.img, (size_t)h->maxPix * 16);
    rc |= dalloc(h, &v.T, (size_t)h->maxPix * 4);
    rc |= dalloc(h, &v.end, (size_t)h->maxPix * 4);
    rc |= dalloc(h, &v.gC, (size_t)h->maxPix * 16);
  }
  if (h->o.viewsPerStep > 1) {
    rc |= dalloc(h, &h->dVB, sizeof(ViewBlock));
    rc |= dalloc(h, &h->dQueue, 4);
    rc |= dalloc(h, &h->viewStats, 8 * kMaxViews);
  }
  // Batched binning: each further view gets its own binning buffers, so
  // all views' binning stages run as one launch each. Skipped when those
  // copies would pass kBatchedBinBytes (then the views bin one after
  // another through the shared buffers).
  if (rc == 0 && h->o.viewsPerStep > 1 && h->o.sortMode == GS_SORT_TILE_SEGMENTS) {
    const int64_t rows = std::min<int64_t>(kSegMaxRows, (cap + kSegMinGroup - 1) / kSegMinGroup);
    const size_t segTiles = (size_t)std::min(h->maxTiles, kMaxSegTiles);
    const size_t chunkBytes = (size_t)((cap + kSegChunk - 1) / kSegChunk) * 4;
    const size_t per = (size_t)cap * 12 + chunkBytes + (size_t)(2 * rows + 1) * 4 + (size_t)rows * segTiles * 4 +
                       (size_t)h->maxTiles * 4 + (size_t)h->entriesCap * 16 + 8;
    if (per * (h->vslots.size() - 1) <= kBatchedBinBytes) {
      for (size_t k = 1; k < h->vslots.size() && rc == 0; k++) {
        gs_trainer::ViewSlot& v = h->vslots[k];
        rc |= dalloc(h, &v.depthKey, (size_t)cap * 4);
        rc |= dalloc(h, &v.rect, (size_t)cap * 8);
        rc |= dalloc(h, &v.chunkKeys, chunkBytes);
        rc |= dalloc(h, &v.blockStart, (size_t)(rows + 1) * 4);
        rc |= dalloc(h, &v.table, (size_t)rows * segTiles * 4);
        rc |= dalloc(h, &v.visPart, (size_t)rows * 4);
        rc |= dalloc(h, &v.tileCnt, (size_t)h->maxTiles * 4);
        rc |= dalloc(h, &v.keys, (size_t)h->entriesCap * 8);
        rc |= dalloc(h, &v.tmp, (size_t)h->entriesCap * 8);
        rc |= dalloc(h, &v.totalKeys, 8);
        // the projection adds to these sums and the scan clears them after use
        if (rc == 0) HIPCK(hipMemset(v.chunkKeys, 0, chunkBytes));
      }
      h->batchedBin = rc == 0;
    }
  }
  if (rc) { free_all(h); return rc < 0 ? rc : GS_E_NOMEM; }
  if (h->o.viewsPerStep > 1) {
    hipError_t e = hipHostMalloc((void**)&h->hVB, sizeof(ViewBlock) * kRing, hipHostMallocDefault);
    for (int k = 0; k < kRing && e == hipSuccess; k++) {
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) h->ringEv.push_back(ev);
    }
    if (e != hipSuccess) {
      free_all(h);
      return hipfail(h, "setup: batched-view staging", e);
    }
    h->ringPos = 0;
    h->viewBlocks = raster_views_blocks(h->o.device, h->shape);
  }

  std::vector<float> planes;
  repack_to_planes(params, n, 16, 4, cap, planes);
  HIPCK(hipMemcpy(h->P, planes.data(), pb, hipMemcpyHostToDevice));
  if (h->shC && !h->shFmt) {
    if (sh) repack_to_planes(sh, n, h->shC, h->shP4, cap, planes);
    else planes.assign((size_t)h->shP4 * cap * 4, 0.0f);
    HIPCK(hipMemcpy(h->SH, planes.data(), sb, hipMemcpyHostToDevice));
  } else if (h->shC) {
    // 16-bit planes: coefficient j in plane j/8, half j%8; f32 rounded to nearest even
    std::vector<uint16_t> b16((size_t)h->shQ8 * cap * 8, 0);
    if (sh)
      for (int64_t i = 0; i < n; i++)
        for (int j = 0; j < h->shC; j++) {
          const float x = sh[i * h->shC + j];
          b16[((size_t)(j / 8) * cap + i) * 8 + (j % 8)] =
              (uint16_t)(h->shFmt == 2 ? gs_f2h_rne(x) : gs_f2bf_rne(x));
        }
    HIPCK(hipMemcpy(h->SH, b16.data(), sb, hipMemcpyHostToDevice));
  }
  HIPCK(hipMemset(h->M, 0, mb));
  HIPCK(hipMemset(h->V, 0, mb));
  HIPCK(hipMemset(h->SHM, 0, smb));
  HIPCK(hipMemset(h->SHV, 0, smb));
  HIPCK(hipMemset(h->accum, 0, (size_t)cap * 48));
  HIPCK(hipMemset(h->stats, 0, 32));
  HIPCK(hipMemset(h->evalStats, 0, 32));
  HIPCK(hipMemset(h->camGrad, 0, (size_t)(ncams + 1) * 32));
  for (size_t k = 1; k < h->vslots.size(); k++) HIPCK(hipMemset(h->vslots[k].accum, 0, (size_t)cap * 48));
  const uint32_t n32 = (uint32_t)n;
  HIPCK(hipMemcpy(h->dN, &n32, 4, hipMemcpyHostToDevice));
  HIPCK(hipMemcpy(h->targets, tgt.data(), tgt.size() * 4, hipMemcpyHostToDevice));

  h->sceneRadius = sceneRadius;
  h->horizon = h->o.maxIters > 0 ? h->o.maxIters : 60000;
  h->basePosLr = h->o.posLr > 0 ? (double)h->o.posLr : 3e-4 * (double)sceneRadius;
  h->holdout = -1;
  h->excluded.assign(ncams, 0);
  h->camM.assign((size_t)(ncams + 1) * 8, 0.0);
  h->camV.assign((size_t)(ncams + 1) * 8, 0.0);
  h->camStep = 0;
  h->logfScale = 0.0;
  h->iter = 0;
  h->adamT = 0;
  h->pixelsSeen = 0;
  h->ready = true;
  return GS_OK;
}

int gs_step(gs_trainer_t* h, int32_t cam) {
  if (!h) return GS_E_ARG;
  return step_impl(h, cam, nullptr);
}

int gs_step_views(gs_trainer_t* h, const int32_t* cams, int32_t count) {
  if (!h) return GS_E_ARG;
  HIPCK(hipSetDevice(h->o.device));
  return step_views_impl(h, cams, count, nullptr);
}

int gs_step_views_debug(gs_trainer_t* h, const int32_t* cams, int32_t count, gs_debug_out* outs) {
  if (!h || !outs) return GS_E_ARG;
  HIPCK(hipSetDevice(h->o.device));
  return step_views_impl(h, cams, count, outs);
}

namespace {

void set_shape(gs_trainer* h, const RasterShape& sh) {
  if (sh.tile == h->shape.tile && sh.batch == h->shape.batch) return;
  h->shape = sh;
  if (h->dVB) h->viewBlocks = raster_views_blocks(h->o.device, sh);
}

// One autotuning pass over cameras ci[0..count) at the current shape: the
// training raster and backward of gs_step (count 1) or gs_step_views
// (count > 1) without the update; statistics go to evalStats, the gradient
// sums stay in the accumulators.
int tune_pass(gs_trainer* h, const int* ci, int count, int activeDeg) {
  if (count == 1) {
    const Cam cam = make_cam_meta(h, h->cams[ci[0]], activeDeg);
    const bool ssim = h->o.loss == GS_LOSS_L1_DSSIM;
    // as the step runs it: Charbonnier's backward inside the forward's launch
    int rc = run_raster(h, cam, ssim ? 2 : 1, h->evalStats, !ssim);
    if (rc) return rc;
    if (!ssim) return GS_OK;
    {
      RasterBufs lb{};
      lb.img = h->img;
      lb.gC = h->gC;
      lb.tgt = h->targets;
      lb.stats = h->evalStats;
      launch_loss_l1ssim(h->st, cam, h->o.ssimLambda, lb, h->ssimScratch);
    }
    RasterBufs b{};
    b.ranges = h->ranges;
    b.tileOrder = h->lastBin.tileOrder;
    b.ids = h->lastBin.ids;
    b.rec = h->rec;
    b.recFmt = h->recFmt;
    b.T = h->Tfin;
    b.end = h->endIdx;
    b.gC = h->gC;
    launch_raster_bwd(h->st, h->shape, cam, make_cut(h), b, h->o.gradAccum, h->accum, false);
    HIPCK(hipGetLastError());
    return GS_OK;
  }
  ViewBlock& vb = h->hVB[h->ringPos];
  hipEvent_t ev = h->ringEv[(size_t)h->ringPos];
  HIPCK(hipEventSynchronize(ev));
  std::vector<int64_t> key = {count, h->n, 0, h->shape.tile, h->shape.batch};
  int rc = fill_views(h, ci, count, activeDeg, h->evalStats, vb, key);
  if (rc) return rc;
  vb.hp = make_hp(h);
  HIPCK(hipMemcpyAsync(h->dVB, &vb, sizeof vb, hipMemcpyHostToDevice, h->st));
  HIPCK(hipEventRecord(ev, h->st));
  h->ringPos = (h->ringPos + 1) % kRing;
  return enqueue_views(h, vb, nullptr, nullptr, nullptr, h->evalStats, false);
}

}  // namespace

int gs_set_raster_shape(gs_trainer_t* h, int32_t tile, int32_t batch) {
  if (!h) return GS_E_ARG;
  if (!raster_shape_ok(tile, batch))
    return fail(h, GS_E_ARG, "raster shape: tile 8, 16 or 32 and batch 64, 128 or 256");
  HIPCK(hipSetDevice(h->o.device));
  set_shape(h, RasterShape{tile, batch});
  return GS_OK;
}

int gs_autotune(gs_trainer_t* h, const int32_t* cams, int32_t count, int32_t reps, gs_tune_entry* out) {
  if (!h || !cams || !out || count < 1 || reps < 1) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  HIPCK(hipSetDevice(h->o.device));
  int ci[kMaxViews];
  if (count > 1) {
    if (count > (int)h->vslots.size() || !h->dVB)
      return fail(h, GS_E_ARG, "autotune: " + std::to_string(count) + " views; opts.viewsPerStep allows 1.." +
                                   std::to_string(std::max<size_t>(1, h->vslots.size())));
    if (h->o.loss != GS_LOSS_CHARBONNIER || h->o.sortMode != GS_SORT_TILE_SEGMENTS || h->o.camOpt)
      return fail(h, GS_E_ARG, "autotune of batched views needs the Charbonnier loss, the tile-segment sort "
                               "and camOpt off");
  }
  for (int k = 0; k < count; k++) {
    if (cams[k] < 0 || cams[k] >= (int)h->cams.size()) return fail(h, GS_E_ARG, "autotune: camera index out of range");
    ci[k] = cams[k];
  }
  const RasterShape start = h->shape;
  const int activeDeg = (int)std::min<int64_t>(h->shDeg, h->iter / 1000);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  int rc = GS_OK;
  const int tiles[3] = {8, 16, 32}, batches[3] = {64, 128, 256};
  for (int i = 0; i < 9; i++) {
    memset(&out[i], 0, sizeof out[i]);
    out[i].tile = tiles[i / 3];
    out[i].batch = batches[i % 3];
  }
  for (int i = 0; i < 9 && rc == GS_OK; i++) {
    gs_tune_entry& en = out[i];
    set_shape(h, RasterShape{en.tile, en.batch});
    // warm-up pass: also gives the pass's keys and budget drops
    if (hipMemsetAsync(h->evalStats, 0, 32, h->st) != hipSuccess) { rc = hipfail(h, "autotune", hipGetLastError()); break; }
    rc = tune_pass(h, ci, count, activeDeg);
    if (rc) break;
    double st4[4];
    uint32_t nk[kMaxViews] = {};
    HIPCK(hipMemcpyAsync(st4, h->evalStats, 32, hipMemcpyDeviceToHost, h->st));
    if (count == 1) HIPCK(hipMemcpyAsync(nk, h->numKeys, 4, hipMemcpyDeviceToHost, h->st));
    else HIPCK(hipMemcpyAsync(nk, h->viewStats, 4ull * count, hipMemcpyDeviceToHost, h->st));
    HIPCK(hipStreamSynchronize(h->st));
    en.dropped = (int64_t)st4[3];
    for (int k = 0; k < count; k++) en.keys += nk[k];
    HIPCK(hipEventRecord(e0, h->st));
    for (int r = 0; r < reps && rc == GS_OK; r++) rc = tune_pass(h, ci, count, activeDeg);
    if (rc) break;
    HIPCK(hipEventRecord(e1, h->st));
    HIPCK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCK(hipEventElapsedTime(&ms, e0, e1));
    en.ms = (double)ms / reps;
    en.status = GS_TUNE_MEASURED;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the gradient sums of the passes and their statistics are discarded
  launch_fill_u32(h->st, (uint32_t*)h->accum, 12 * h->cap, 0u);
  for (const gs_trainer::ViewSlot& sl : h->vslots) launch_fill_u32(h->st, (uint32_t*)sl.accum, 12 * h->cap, 0u);
  launch_fill_u32(h->st, (uint32_t*)h->evalStats, 8, 0u);
  HIPCK(hipStreamSynchronize(h->st));
  if (rc) {
    set_shape(h, start);
    return rc;
  }
  // eligible: measured, and dropping no more keys than the reference's 16-pixel tiles
  const int64_t drops16 = out[5].status == GS_TUNE_MEASURED ? out[5].dropped : INT64_MAX;
  int best = -1;
  for (int i = 0; i < 9; i++) {
    if (out[i].status != GS_TUNE_MEASURED) continue;
    if (out[i].dropped > drops16) {
      out[i].status = GS_TUNE_MORE_DROPS;
      continue;
    }
    if (best < 0 || out[i].ms < out[best].ms) best = i;
  }
  set_shape(h, best >= 0 ? RasterShape{out[best].tile, out[best].batch} : start);
  return GS_OK;
}

int gs_step_batch(gs_trainer_t* h, const int32_t* cams, int32_t count) {
  if (!h || count < 0 || (count > 0 && !cams)) return GS_E_ARG;
  for (int32_t k = 0; k < count; k++) {
    const int rc = step_impl(h, cams[k], nullptr);
    if (rc != GS_OK) return rc;
  }
  return GS_OK;
}

int gs_step_debug(gs_trainer_t* h, int32_t cam, gs_debug_out* out) {
  if (!h || !out) return GS_E_ARG;
  return step_impl(h, cam, out);
}

// Runs a *_async entry point, waits and frees: the synchronous form.
static int wait_pending(gs_trainer* h, int rc, gs_pending* p, gs_readback_result* r) {
  if (rc != GS_OK) return rc;
  rc = gs_pending_wait(p, r);
  if (rc != GS_OK) fail(h, rc, p->err);
  gs_pending_free(p);
  return rc;
}

int gs_read_loss(gs_trainer_t* h, double* mse, int32_t* has) {
  if (!h || !mse || !has) return GS_E_ARG;
  gs_pending* p = nullptr;
  gs_readback_result r{};
  const int rc = wait_pending(h, gs_read_loss_async(h, &p), p, &r);
  if (rc != GS_OK) return rc;
  *mse = r.value;
  *has = r.has;
  return GS_OK;
}

int gs_render(gs_trainer_t* h, const gs_camera* c, int32_t trainMode, int64_t offset, float* rgba) {
  if (!h || !c || !rgba) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if ((int64_t)c->w * c->h > h->maxPix || tiles_of(c->w, c->h, h->shape.tile) > h->maxTiles)
    return fail(h, GS_E_LIMIT, "renderView: view larger than maxViewW x maxViewH");
  if (c->w <= 0 || c->h <= 0) return fail(h, GS_E_ARG, "renderView: empty view");
  if (trainMode && (offset < 0 || offset + (int64_t)c->w * c->h > h->totalPx))
    return fail(h, GS_E_ARG, "renderView: target offset outside the packed targets");
  const Cam cam = make_cam(h, *c, trainMode, offset, h->shDeg);
  int rc = run_raster(h, cam, trainMode ? 1 : 0, h->evalStats);
  if (rc) return rc;
  HIPCK(hipMemcpyAsync(rgba, h->img, 16ull * c->w * c->h, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipStreamSynchronize(h->st));
  return GS_OK;
}

int gs_eval_cam_psnr(gs_trainer_t* h, int32_t ci, double* psnr) {
  if (!h || !psnr) return GS_E_ARG;
  gs_pending* p = nullptr;
  gs_readback_result r{};
  const int rc = wait_pending(h, gs_eval_cam_psnr_async(h, ci, &p), p, &r);
  if (rc != GS_OK) return rc;
  *psnr = r.value;
  return GS_OK;
}

int gs_eval_cam_psnr_refined(gs_trainer_t* h, int32_t ci, int32_t steps, double* psnr) {
  // trainer.js:583-612: test-time refinement of one camera's pose (and
  // exposure with expComp) against its photo, Gaussians frozen
  if (!h || !psnr || steps < 0) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (ci < 0 || ci >= (int)h->cams.size()) return fail(h, GS_E_ARG, "camera index out of range");
  HIPCK(hipSetDevice(h->o.device));
  double R[9], t[3];
  memcpy(R, h->cams[ci].R, sizeof R);
  memcpy(t, h->cams[ci].t, sizeof t);
  double g = 0.0, b = 0.0;
  double m[8] = {0}, v[8] = {0};
  const double b1 = 0.9, b2 = 0.99, eps = 1e-15;
  const double rotLr = 3e-4, trnLr = 3e-4 * (double)h->sceneRadius, expLr = 5e-3;
  for (int s = 1; s <= steps; s++) {
    int32_t cg[8];
    int rc = eval_pass(h, ci, R, t, g, b, nullptr, cg);
    if (rc) return rc;
    double upd[8];
    for (int k = 0; k < 8; k++) {
      const double grad = cg[k] / 64.0;
      m[k] = b1 * m[k] + (1 - b1) * grad;
      v[k] = b2 * v[k] + (1 - b2) * grad * grad;
      const double mh = m[k] / (1 - std::pow(b1, (double)s));
      const double vh = v[k] / (1 - std::pow(b2, (double)s));
      upd[k] = (k < 3 ? rotLr : (k < 6 ? trnLr : expLr)) * mh / (std::sqrt(vh) + eps);
    }
    apply_pose_step(upd, R, t);
    if (h->o.expComp) {
      g = clampd(g - upd[6], -0.1, 0.1);
      b = clampd(b - upd[7], -0.05, 0.05);
    }
  }
  return eval_pass(h, ci, R, t, g, b, psnr, nullptr);
}

int gs_get_camera(gs_trainer_t* h, int32_t ci, gs_cam_state* out) {
  if (!h || !out) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (ci < 0 || ci >= (int)h->cams.size()) return fail(h, GS_E_ARG, "camera index out of range");
  const CamMeta& m = h->cams[ci];
  memcpy(out->R, m.R, sizeof out->R);
  memcpy(out->t, m.t, sizeof out->t);
  out->f = m.f;
  out->g = m.g;
  out->b = m.b;
  return GS_OK;
}

int gs_read_cam_grads(gs_trainer_t* h, int32_t* out, int32_t count) {
  if (!h || !out) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  const int64_t total = (int64_t)(h->cams.size() + 1) * 8;
  if (count < total) return fail(h, GS_E_ARG, "readCamGrads: buffer smaller than (numCams + 1) * 8");
  HIPCK(hipMemcpyAsync(out, h->camGrad, total * 4, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipStreamSynchronize(h->st));
  return GS_OK;
}

int gs_read_gaussians(gs_trainer_t* h, float* params, float* sh) {
  if (!h) return GS_E_ARG;
  gs_pending* p = nullptr;
  gs_readback_result r{};
  return wait_pending(h, gs_read_gaussians_async(h, params, sh, &p), p, &r);
}

// ---- asynchronous readbacks ----

static int pending_new(gs_trainer* h, gs_pending::Kind kind, gs_pending** out) {
  gs_pending* p = new gs_pending();
  p->kind = kind;
  p->device = h->o.device;
  hipError_t e = hipEventCreateWithFlags(&p->ev, hipEventDisableTiming);
  if (e == hipSuccess && kind != gs_pending::kGaussians) e = hipHostMalloc((void**)&p->stats, 32, hipHostMallocDefault);
  if (e != hipSuccess) {
    gs_pending_free(p);
    return hipfail(h, "readback setup", e);
  }
  *out = p;
  return GS_OK;
}

int gs_read_loss_async(gs_trainer_t* h, gs_pending_t** out) {
  if (!h || !out) return GS_E_ARG;
  *out = nullptr;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  HIPCK(hipSetDevice(h->o.device));
  gs_pending* p = nullptr;
  int rc = pending_new(h, gs_pending::kLoss, &p);
  if (rc != GS_OK) return rc;
  // copy the window's stats, then reset the window; stats[3] (the budget
  // drops) keeps counting across windows (trainer.js:621-627)
  hipError_t e = hipMemcpyAsync(p->stats, h->stats, 32, hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipMemsetAsync(h->stats, 0, 24, h->st);
  if (e == hipSuccess) e = hipEventRecord(p->ev, h->st);
  if (e != hipSuccess) {
    gs_pending_free(p);
    return hipfail(h, "readLoss", e);
  }
  p->px = h->pixelsSeen;
  h->pixelsSeen = 0;
  *out = p;
  return GS_OK;
}

int gs_eval_cam_psnr_async(gs_trainer_t* h, int32_t ci, gs_pending_t** out) {
  if (!h || !out) return GS_E_ARG;
  *out = nullptr;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (ci < 0 || ci >= (int)h->cams.size()) return fail(h, GS_E_ARG, "camera index out of range");
  HIPCK(hipSetDevice(h->o.device));
  const CamMeta& m = h->cams[ci];
  // the reference evaluates with the camera's training uniform (trainer.js:547)
  const Cam cam = make_cam_meta(h, m, m.activeDeg);
  HIPCK(hipMemsetAsync(h->evalStats, 0, 32, h->st));
  int rc = run_raster(h, cam, 1, h->evalStats);
  if (rc) return rc;
  gs_pending* p = nullptr;
  rc = pending_new(h, gs_pending::kPsnr, &p);
  if (rc != GS_OK) return rc;
  hipError_t e = hipMemcpyAsync(p->stats, h->evalStats, 32, hipMemcpyDeviceToHost, h->st);
  // trainer.js:562, :570-571: the camera gradient rows are drained and the
  // training loss window is reset too
  if (e == hipSuccess) e = hipMemsetAsync(h->camGrad, 0, (h->cams.size() + 1) * 32, h->st);
  if (e == hipSuccess) e = hipMemsetAsync(h->stats, 0, 24, h->st);
  if (e == hipSuccess) e = hipEventRecord(p->ev, h->st);
  if (e != hipSuccess) {
    gs_pending_free(p);
    return hipfail(h, "evalCamPsnr", e);
  }
  h->pixelsSeen = 0;
  *out = p;
  return GS_OK;
}

int gs_read_gaussians_async(gs_trainer_t* h, float* params, float* sh, gs_pending_t** out) {
  if (!h || !out) return GS_E_ARG;
  *out = nullptr;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  HIPCK(hipSetDevice(h->o.device));
  gs_pending* p = nullptr;
  int rc = pending_new(h, gs_pending::kGaussians, &p);
  if (rc != GS_OK) return rc;
  p->n = h->n;
  p->cap = h->cap;
  p->shC = h->shC;
  p->shP4 = h->shP4;
  p->shQ8 = h->shQ8;
  p->shFmt = h->shFmt;
  p->dstP = params;
  p->dstSH = h->shC ? sh : nullptr;
  hipError_t e = hipSuccess;
  if (params) {
    const size_t bytes = (size_t)h->cap * 64;
    e = hipHostMalloc(&p->hP, bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemcpyAsync(p->hP, h->P, bytes, hipMemcpyDeviceToHost, h->st);
  }
  if (e == hipSuccess && p->dstSH) {
    const size_t bytes = (size_t)h->cap * 16 * (h->shFmt ? h->shQ8 : h->shP4);
    e = hipHostMalloc(&p->hSH, bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemcpyAsync(p->hSH, h->SH, bytes, hipMemcpyDeviceToHost, h->st);
  }
  if (e == hipSuccess) e = hipEventRecord(p->ev, h->st);
  if (e != hipSuccess) {
    gs_pending_free(p);
    return hipfail(h, "readGaussians", e);
  }
  *out = p;
  return GS_OK;
}

int gs_pending_wait(gs_pending_t* p, gs_readback_result* out) {
  if (!p || !out) return GS_E_ARG;
  memset(out, 0, sizeof(*out));
  hipError_t e = hipSetDevice(p->device);
  if (e == hipSuccess) e = hipEventSynchronize(p->ev);
  if (e != hipSuccess) {
    p->err = std::string("readback: ") + hipGetErrorString(e);
    return GS_E_HIP;
  }
  const double* s = p->stats;
  switch (p->kind) {
    case gs_pending::kLoss:
      // trainer.js:621-630
      out->overflow = (int64_t)s[3];
      if (p->px != 0 && s[2] != 0.0) {
        out->has = 1;
        out->value = s[0] / (s[2] * 3.0);
      }
      break;
    case gs_pending::kPsnr: {
      // trainer.js:565-572
      const double mse = s[0] / (std::max(1.0, s[2]) * 3.0);
      out->has = 1;
      out->value = mse > 0 ? -10.0 * std::log10(mse) : INFINITY;
      break;
    }
    case gs_pending::kGaussians: {
      out->has = 1;
      const int64_t n = p->n, cap = p->cap;
      if (p->dstP) {
        const float* pl = (const float*)p->hP;
        for (int64_t i = 0; i < n; i++)
          for (int j = 0; j < 16; j++) p->dstP[i * 16 + j] = pl[((size_t)(j / 4) * cap + i) * 4 + (j % 4)];
      }
      if (p->dstSH && !p->shFmt) {
        const float* pl = (const float*)p->hSH;
        for (int64_t i = 0; i < n; i++)
          for (int j = 0; j < p->shC; j++) p->dstSH[i * p->shC + j] = pl[((size_t)(j / 4) * cap + i) * 4 + (j % 4)];
      } else if (p->dstSH) {
        const uint16_t* b16 = (const uint16_t*)p->hSH;
        for (int64_t i = 0; i < n; i++)
          for (int j = 0; j < p->shC; j++)
            p->dstSH[i * p->shC + j] = p->shFmt == 2 ? gs_h2f(b16[((size_t)(j / 8) * cap + i) * 8 + (j % 8)])
                                                      : gs_bf2f(b16[((size_t)(j / 8) * cap + i) * 8 + (j % 8)]);
      }
      break;
    }
  }
  return GS_OK;
}

const char* gs_pending_error(gs_pending_t* p) { return p ? p->err.c_str() : "null readback"; }

void gs_pending_free(gs_pending_t* p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->ev) {
    (void)hipEventSynchronize(p->ev);   // a copy in flight still targets the pinned buffers
    (void)hipEventDestroy(p->ev);
  }
  if (p->stats) (void)hipHostFree(p->stats);
  if (p->hP) (void)hipHostFree(p->hP);
  if (p->hSH) (void)hipHostFree(p->hSH);
  delete p;
}

int gs_set_holdout(gs_trainer_t* h, int32_t ci) {
  if (!h) return GS_E_ARG;
  h->holdout = ci;
  return GS_OK;
}

int gs_set_excluded(gs_trainer_t* h, const int32_t* c, int32_t count) {
  if (!h || (count > 0 && !c)) return GS_E_ARG;
  std::fill(h->excluded.begin(), h->excluded.end(), 0);
  for (int i = 0; i < count; i++)
    if (c[i] >= 0 && c[i] < (int)h->excluded.size()) h->excluded[c[i]] = 1;
  return GS_OK;
}

int gs_set_iter(gs_trainer_t* h, int64_t iter) {
  if (!h || iter < 0) return GS_E_ARG;
  h->iter = iter;
  h->adamT = iter;   // the reference's Adam t is its iteration counter
  return GS_OK;
}

int gs_get_info(gs_trainer_t* h, gs_info* out) {
  if (!h || !out) return GS_E_ARG;
  memset(out, 0, sizeof(*out));
  out->n = h->n;
  out->cap = h->cap;
  out->iter = h->iter;
  out->shDeg = h->shDeg;
  out->shK = h->shK;
  out->stride = 16;
  out->numCams = (int32_t)h->cams.size();
  out->lastCam = h->lastCam;
  out->holdout = h->holdout;
  out->entriesCap = h->entriesCap;
  out->maxTiles = h->maxTiles;
  out->horizon = (float)h->horizon;
  out->tile = h->shape.tile;
  out->batch = h->shape.batch;
  return GS_OK;
}

int gs_read_step_stats(gs_trainer_t* h, gs_step_stats* out) {
  if (!h || !out) return GS_E_ARG;
  memset(out, 0, sizeof(*out));
  if (!h->ready) return GS_OK;
  HIPCK(hipSetDevice(h->o.device));
  uint32_t nk = 0, nv = 0;
  uint32_t vs[2 * kMaxViews];
  double s[4];
  HIPCK(hipMemcpyAsync(&nk, h->numKeys, 4, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipMemcpyAsync(&nv, h->numVisible, 4, hipMemcpyDeviceToHost, h->st));
  if (h->lastViews > 0) HIPCK(hipMemcpyAsync(vs, h->viewStats, sizeof vs, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipMemcpyAsync(s, h->stats, 32, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipStreamSynchronize(h->st));
  out->lastKeys = nk;
  out->lastVisible = nv;
  if (h->lastViews > 0) {   // a batched step: all its views
    out->lastKeys = out->lastVisible = 0;
    for (int k = 0; k < h->lastViews; k++) {
      out->lastKeys += vs[k];
      out->lastVisible += vs[kMaxViews + k];
    }
  }
  out->overflowKeys = (int64_t)s[3];
  return GS_OK;
}

// ---- refine (trainer.js:666-794) ----
namespace {

struct JsRng {   // makeRng (src/sfm/geometry.js:750-758): mulberry32, doubles in [0, 1)
  uint32_t a;
  double operator()() { return (double)mulberry32(a) / 4294967296.0; }
};

}  // namespace

int gs_refine(gs_trainer_t* h, const gs_refine_opts* ro, gs_refine_result* out) {
  if (!h || !ro || !out || ro->size != (int32_t)sizeof(gs_refine_opts)) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  memset(out, 0, sizeof(*out));
  HIPCK(hipSetDevice(h->o.device));
  HIPCK(hipStreamSynchronize(h->st));
  const int64_t n = h->n;

  // telemetry (trainer.js:780-792): most keys in one tile of the last
  // raster, counted before the entry budget (the reference reads the raw
  // per-tile counts), and the cumulative entry-budget drops
  if (h->lastNumTiles > 0 && h->lastSegPath) {
    std::vector<uint32_t> tc((size_t)h->lastNumTiles);
    HIPCK(hipMemcpy(tc.data(), h->lastTileCnt, tc.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t c : tc) out->maxTile = std::max<int64_t>(out->maxTile, c);
  } else if (h->lastNumTiles > 0) {
    // the radix strategies hold no per-tile totals; their ranges lack only
    // keys past the budget
    std::vector<uint32_t> rg(2 * (size_t)h->lastNumTiles);
    HIPCK(hipMemcpy(rg.data(), h->lastRanges, rg.size() * 4, hipMemcpyDeviceToHost));
    for (int t = 0; t < h->lastNumTiles; t++) out->maxTile = std::max<int64_t>(out->maxTile, rg[2 * t + 1] - rg[2 * t]);
  }
  {
    double s[4];
    HIPCK(hipMemcpy(s, h->stats, 32, hipMemcpyDeviceToHost));
    out->overflow = (int64_t)s[3];
  }

  // opacity logit and log-scales of every Gaussian (16 B each)
  std::vector<float> q(4 * (size_t)n);
  float4* dq = nullptr;
  HIPCK(hipMalloc(&dq, (size_t)std::max<int64_t>(n, 1) * 16));
  launch_refine_gather(h->st, n, h->cap, h->P, dq);
  hipError_t e = hipMemcpyAsync(q.data(), dq, q.size() * 4, hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  (void)hipFree(dq);
  if (e != hipSuccess) return hipfail(h, "refine gather", e);

  // dead and donor sets in index order (trainer.js:687-694)
  std::vector<int64_t> dead, donors;
  for (int64_t i = 0; i < n; i++) {
    const double o = 1.0 / (1.0 + std::exp(-(double)q[4 * i]));
    if (o < 0.02) dead.push_back(i);
    else if (o > 0.4) donors.push_back(i);
  }
  out->n = n;
  if (donors.size() < 16) return GS_OK;
  const int64_t moveCap = (int64_t)std::ceil((double)n * 0.05);
  if ((int64_t)dead.size() > moveCap) dead.resize((size_t)moveCap);

  JsRng rng{ro->seed};
  auto gauss = [&]() {
    const double u = std::max(1e-9, rng()), w = rng();
    return std::sqrt(-2.0 * std::log(u)) * std::cos(2.0 * M_PI * w);
  };
  auto meanLogScale = [&](int64_t i) {
    return ((double)q[4 * i + 1] + (double)q[4 * i + 2] + (double)q[4 * i + 3]) / 3.0;
  };
  // largest donors first, stable (Array.prototype.sort), top max(16, |donors|/4)
  std::vector<int64_t> bigDonors(donors);
  std::stable_sort(bigDonors.begin(), bigDonors.end(),
                   [&](int64_t a, int64_t b) { return meanLogScale(b) - meanLogScale(a) < 0.0; });
  bigDonors.resize(std::min(bigDonors.size(), std::max<size_t>(16, donors.size() >> 2)));

  std::vector<RefineOp> ops;
  const double shrink = std::log(1.0 / 1.6), cloneLs = std::log(0.85);
  const float cloneOpa = (float)std::log(0.25 / 0.75);
  auto spawnAt = [&](int64_t target, bool allowSplit) {
    const bool doSplit = allowSplit && bigDonors.size() > 16 && rng() < 0.7;
    int64_t don;
    if (doSplit) {
      const size_t di = (size_t)(int32_t)(rng() * (double)bigDonors.size());
      don = bigDonors[di];
      bigDonors[di] = bigDonors.back();
      bigDonors.pop_back();   // a splat splits at most once per call
    } else {
      don = donors[(size_t)(int32_t)(rng() * (double)donors.size())];
    }
    RefineOp op{};
    op.target = target;
    op.donor = don;
    const double s = std::exp(meanLogScale(don));
    for (int c = 0; c < 3; c++) op.jit[c] = gauss() * s * 0.7;
    for (int c = 0; c < 3; c++) op.ls[c] = q[4 * don + 1 + c];
    if (doSplit) {
      for (int c = 0; c < 3; c++) {
        op.ls[c] = (float)((double)op.ls[c] + shrink);
        q[4 * don + 1 + c] = (float)((double)q[4 * don + 1 + c] + shrink);
      }
      op.opa = q[4 * don];
      op.split = 1;
    } else {
      for (int c = 0; c < 3; c++) op.ls[c] = (float)((double)op.ls[c] + cloneLs);
      op.opa = cloneOpa;
    }
    ops.push_back(op);
  };
  for (int64_t i : dead) spawnAt(i, false);
  const int64_t limit = std::min<int64_t>(h->cap, ro->growLimit > 0 ? ro->growLimit : h->cap);
  const double growUntil = ro->growUntil >= 0.0 ? ro->growUntil : 0.75 * h->horizon;
  const double growRate = ro->growRate >= 0.0f ? (double)ro->growRate : 0.15;
  const int64_t grown = (double)h->iter < growUntil
      ? std::max<int64_t>(0, std::min<int64_t>((int64_t)std::ceil((double)n * growRate), limit - n)) : 0;
  for (int64_t k = 0; k < grown; k++) spawnAt(n + k, true);

  if (!ops.empty()) {
    RefineOp* dops = nullptr;
    HIPCK(hipMalloc(&dops, ops.size() * sizeof(RefineOp)));
    e = hipMemcpy(dops, ops.data(), ops.size() * sizeof(RefineOp), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      launch_refine_apply(h->st, h->shDeg, h->prec, h->cap, dops, (int64_t)ops.size(), h->P, h->M, h->V, h->SH,
                          h->SHM, h->SHV);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(h->st);
    }
    (void)hipFree(dops);
    if (e != hipSuccess) return hipfail(h, "refine apply", e);
  }
  if (grown > 0) {
    h->n = n + grown;
    const uint32_t n32 = (uint32_t)h->n;
    HIPCK(hipMemcpy(h->dN, &n32, 4, hipMemcpyHostToDevice));
  }
  out->moved = (int64_t)dead.size();
  out->grown = grown;
  out->n = h->n;
  return GS_OK;
}

// ---- PLY export (src/io/ply.js; session.js:637-645) ----
}  // extern "C"

namespace {

// The PLY header of gaussiansToPly (src/io/ply.js:47-70) for n rows with shC
// rest coefficients; written to dst when it is not null. Returns its length.
// Builds nothing on the heap (the no-allocation export shares it).
int64_t ply_header(char* dst, int shC, int64_t n) {
  int64_t len = 0;
  auto put = [&](const char* text) {
    const size_t k = strlen(text);
    if (dst) memcpy(dst + len, text, k);
    len += (int64_t)k;
  };
  char line[64];
  put("ply\nformat binary_little_endian 1.0\n");
  snprintf(line, sizeof line, "element vertex %lld\n", (long long)n);
  put(line);
  const char* head[] = {"x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"};
  for (const char* p : head) {
    snprintf(line, sizeof line, "property float %s\n", p);
    put(line);
  }
  for (int k = 0; k < shC; k++) {
    snprintf(line, sizeof line, "property float f_rest_%d\n", k);
    put(line);
  }
  const char* tail[] = {"opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"};
  for (const char* p : tail) {
    snprintf(line, sizeof line, "property float %s\n", p);
    put(line);
  }
  put("end_header\n");
  return len;
}

// camera centre -(R^T t) in JS numbers, stored through a Float32Array
// (session.js:32-38, :642)
void cam_centre(const CamMeta& k, float* out) {
  for (int a = 0; a < 3; a++)
    out[a] = (float)(-((k.R[a] * k.t[0] + k.R[3 + a] * k.t[1]) + k.R[6 + a] * k.t[2]));
}

// ---- device scratch of the no-allocation export (gs_export_ply_noalloc) ----
// A fixed array per GPU, held until the process ends: camera centres for the
// opacity bake (at most kNoAllocMaxCams), then one chunk of PLY rows. The
// host never passes `scratch` itself to a kernel or copy (in host code the
// name is a stand-in, not a device address); it uses the address that
// hipGetSymbolAddress returned for the current device.
constexpr int kNoAllocMaxCams = 4096;
constexpr int64_t kNoAllocCamFloats = 3 * kNoAllocMaxCams;          // 48 KiB
constexpr int64_t kNoAllocRowFloats = int64_t(4) << 20;             // 16 MiB of rows
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kNoAllocCamFloats + kNoAllocRowFloats];

std::mutex g_scratchMu;
float* g_scratchPtr[kMaxDevices] = {};

// Looks up this device's copy of `scratch` once per device; the device must
// be current. Called by gs_create.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return GS_E_ARG;
  std::lock_guard<std::mutex> lk(g_scratchMu);
  if (g_scratchPtr[device]) return GS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return GS_E_HIP;
  }
  g_scratchPtr[device] = (float*)ptr;
  return GS_OK;
}

float* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratchMu);
  return g_scratchPtr[device];
}

}  // namespace

extern "C" {

int gs_export_ply(gs_trainer_t* h, int32_t bake, void* buf, int64_t cap, int64_t* size) {
  if (!h || !size) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  const int64_t n = h->n;
  const int W = 17 + h->shC;
  const int64_t hdrLen = ply_header(nullptr, h->shC, n);
  const int64_t body = n * (int64_t)W * 4;
  *size = hdrLen + body;
  if (!buf) return GS_OK;
  if (cap < *size) return fail(h, GS_E_ARG, "exportPly: buffer smaller than the file");
  HIPCK(hipSetDevice(h->o.device));
  const int nc = (int)h->cams.size();
  std::vector<float> cp(3 * (size_t)nc);
  for (int c = 0; c < nc; c++) cam_centre(h->cams[c], &cp[3 * (size_t)c]);
  float* dCam = nullptr;
  float* dOut = nullptr;
  HIPCK(hipMalloc(&dCam, cp.size() * 4 + 4));
  hipError_t e = hipMalloc(&dOut, (size_t)std::max<int64_t>(body, 4));
  if (e == hipSuccess) e = hipMemcpy(dCam, cp.data(), cp.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    launch_ply_rows(h->st, h->shDeg, h->prec, 0, n, h->cap, h->P, h->SH, bake ? 1 : 0, h->cams[0].f, dCam,
                    nc, dOut);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  ply_header((char*)buf, h->shC, n);
  if (e == hipSuccess) e = hipMemcpy((char*)buf + hdrLen, dOut, (size_t)body, hipMemcpyDeviceToHost);
  (void)hipFree(dCam);
  if (dOut) (void)hipFree(dOut);
  if (e != hipSuccess) return hipfail(h, "exportPly", e);
  return GS_OK;
}

int gs_export_ply_noalloc(gs_trainer_t* h, int32_t bake, void* buf, int64_t cap, int64_t* size) {
  if (!h || !size) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  const int nc = (int)h->cams.size();
  if (nc > kNoAllocMaxCams)
    return fail(h, GS_E_LIMIT, "exportPlyNoAlloc: " + std::to_string(nc) + " cameras; the device scratch holds " +
                                   std::to_string(kNoAllocMaxCams) + " camera centres");
  const int64_t n = h->n;
  const int W = 17 + h->shC;
  const int64_t hdrLen = ply_header(nullptr, h->shC, n);
  const int64_t body = n * (int64_t)W * 4;
  *size = hdrLen + body;
  if (!buf) return GS_OK;
  if (cap < *size) return fail(h, GS_E_ARG, "exportPlyNoAlloc: buffer smaller than the file");
  HIPCK(hipSetDevice(h->o.device));
  float* base = scratch_ptr(h->o.device);
  if (!base) return fail(h, GS_E_HIP, "exportPlyNoAlloc: no device scratch for this device");
  float* dCam = base;
  float* dRows = base + kNoAllocCamFloats;
  // Every row of a chunk and every centre the bake reads is written before
  // it is read, so nothing carries over from an earlier call.
  if (bake) {
    float cp[3 * 256];
    for (int c0 = 0; c0 < nc; c0 += 256) {
      const int m = std::min(256, nc - c0);
      for (int c = 0; c < m; c++) cam_centre(h->cams[c0 + c], &cp[3 * c]);
      HIPCK(hipMemcpyAsync(dCam + 3 * c0, cp, (size_t)m * 12, hipMemcpyHostToDevice, h->st));
      HIPCK(hipStreamSynchronize(h->st));   // cp is reused by the next group
    }
  }
  ply_header((char*)buf, h->shC, n);
  const int64_t rowsPerChunk = kNoAllocRowFloats / W;
  for (int64_t first = 0; first < n; first += rowsPerChunk) {
    const int64_t count = std::min(rowsPerChunk, n - first);
    launch_ply_rows(h->st, h->shDeg, h->prec, first, count, h->cap, h->P, h->SH, bake ? 1 : 0, h->cams[0].f,
                    dCam, nc, dRows);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync((char*)buf + hdrLen + first * (int64_t)W * 4, dRows, (size_t)(count * W * 4),
                         hipMemcpyDeviceToHost, h->st));
    HIPCK(hipStreamSynchronize(h->st));   // the next chunk overwrites the rows
  }
  return GS_OK;
}

int gs_fd_grad64(gs_trainer_t* h, int32_t ci, const int64_t* gauss, const int32_t* slots, const double* steps,
                 int32_t count, double* out) {
  if (!h || count < 0 || (count > 0 && (!gauss || !slots || !steps || !out))) return GS_E_ARG;
  if (!h->ready) return fail(h, GS_E_ARG, "setup() has not been called");
  if (ci < 0 || ci >= (int)h->cams.size()) return fail(h, GS_E_ARG, "fdGrad: camera index out of range");
  if (count == 0) return GS_OK;
  HIPCK(hipSetDevice(h->o.device));
  for (int k = 0; k < count; k++) {
    if (gauss[k] < 0 || gauss[k] >= h->n || slots[k] < 0 || slots[k] > 13 || !(steps[k] > 0.0))
      return fail(h, GS_E_ARG, "fdGrad: sample " + std::to_string(k) + " out of range");
  }
  // the raster of the current parameters, with the training uniform
  CamMeta& meta = h->cams[ci];
  const int activeDeg = (int)std::min<int64_t>(h->shDeg, h->iter / 1000);
  const Cam cam = make_cam_meta(h, meta, activeDeg);
  HIPCK(hipMemsetAsync(h->evalStats, 0, 32, h->st));
  int rc = run_raster(h, cam, 1, h->evalStats);
  if (rc) return rc;
  // the samples' tile rectangles, one tile of margin each side
  std::vector<uint2> rects((size_t)count);
  for (int k = 0; k < count; k++)
    HIPCK(hipMemcpyAsync(&rects[k], h->rect + gauss[k], 8, hipMemcpyDeviceToHost, h->st));
  HIPCK(hipStreamSynchronize(h->st));
  std::vector<uint2> items;
  std::vector<FdSample> smp((size_t)count);
  for (int k = 0; k < count; k++) {
    smp[k] = FdSample{gauss[k], slots[k], 0, steps[k], rects[k]};
    const int tx0 = std::max<int>(0, (int)(rects[k].x & 0xffff) - 1);
    const int tx1 = std::min<int>(cam.tilesX - 1, (int)(rects[k].x >> 16) + 1);
    const int ty0 = std::max<int>(0, (int)(rects[k].y & 0xffff) - 1);
    const int ty1 = std::min<int>(cam.tilesY - 1, (int)(rects[k].y >> 16) + 1);
    for (int ty = ty0; ty <= ty1; ty++)
      for (int tx = tx0; tx <= tx1; tx++) items.push_back(make_uint2((uint32_t)k, (uint32_t)(ty * cam.tilesX + tx)));
  }
  FdSample* dS = nullptr;
  uint2* dI = nullptr;
  double* dO = nullptr;
  hipError_t e = hipMalloc(&dS, smp.size() * sizeof(FdSample));
  if (e == hipSuccess) e = hipMalloc(&dI, std::max<size_t>(items.size(), 1) * sizeof(uint2));
  if (e == hipSuccess) e = hipMalloc(&dO, (size_t)count * 8);
  if (e == hipSuccess) e = hipMemcpy(dS, smp.data(), smp.size() * sizeof(FdSample), hipMemcpyHostToDevice);
  if (e == hipSuccess && !items.empty())
    e = hipMemcpy(dI, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemsetAsync(dO, 0, (size_t)count * 8, h->st);
  if (e == hipSuccess) {
    FdArgs a{};
    a.cap = h->cap;
    a.P = h->P;
    a.SH = h->SH;
    a.rec = h->rec;
    a.recFmt = h->recFmt;
    a.depthKey = h->depthKey;
    a.ids = h->lastBin.ids;
    a.ranges = h->ranges;
    a.tgt = h->targets;
    a.samples = dS;
    a.items = dI;
    a.out = dO;
    // at most 2^31 - 1 blocks per launch
    const int64_t kChunk = int64_t(1) << 30;
    for (int64_t i0 = 0; i0 < (int64_t)items.size() && e == hipSuccess; i0 += kChunk) {
      a.items = dI + i0;
      launch_fd64(h->st, h->shDeg, h->shFmt, std::min<int64_t>(kChunk, (int64_t)items.size() - i0), cam, make_cut(h), a);
      e = hipGetLastError();
    }
  }
  std::vector<double> sums((size_t)count, 0.0);
  if (e == hipSuccess) e = hipMemcpyAsync(sums.data(), dO, (size_t)count * 8, hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  if (dS) (void)hipFree(dS);
  if (dI) (void)hipFree(dI);
  if (dO) (void)hipFree(dO);
  if (e != hipSuccess) return hipfail(h, "fdGrad", e);
  for (int k = 0; k < count; k++) out[k] = sums[k] / (2.0 * steps[k]);
  return GS_OK;
}

int gs_sync(gs_trainer_t* h) {
  if (!h) return GS_E_ARG;
  HIPCK(hipStreamSynchronize(h->st));
  return GS_OK;
}

const char* gs_last_error(gs_trainer_t* h) {
  return h ? h->err.c_str() : "null trainer";
}

int gs_profile(gs_trainer_t* h, int32_t enable) {
  if (!h) return GS_E_ARG;
  if (enable && !h->prof) {
    h->slots.clear();
  }
  if (!enable) prof_collect(h);
  h->prof = enable != 0;
  return GS_OK;
}

int gs_profile_read(gs_trainer_t* h, gs_kernel_time* out, int32_t max, int32_t* count) {
  if (!h || !count) return GS_E_ARG;
  prof_collect(h);
  *count = (int32_t)h->slots.size();
  for (int k = 0; k < (int)h->slots.size() && k < max; k++) {
    memset(&out[k], 0, sizeof(gs_kernel_time));
    snprintf(out[k].name, sizeof out[k].name, "%s", h->slots[k].name.c_str());
    out[k].total_ms = h->slots[k].ms;
    out[k].launches = h->slots[k].launches;
    out[k].bytes = h->slots[k].bytes;
  }
  return GS_OK;
}

void* gs_stream(gs_trainer_t* h) { return h ? (void*)h->st : nullptr; }

}  // extern "C"
