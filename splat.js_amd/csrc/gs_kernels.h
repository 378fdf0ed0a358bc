// gs_kernels.h — host-side launchers of the trainer kernels (internal).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_device.h"

namespace gs {

// Sorted key values: Gaussian id in bits 0-27, the 4-bit mask of the tile's
// 8x8 pixel blocks that may take the splat (gs_cover_tile_mask) in bits 28-31.
constexpr int kMaskShift = 28;
constexpr uint32_t kIdMask = (1u << kMaskShift) - 1u;

struct SortScratch {
  int G;                  // blocks per radix pass (fixed grid)
  uint32_t* hist;         // 256 x G
  uint32_t* totals;       // 256
  uint64_t* partials;     // ceil(cap / 256): key count per 256 Gaussians
  uint32_t* vpartials;    // ceil(cap / 256): visible Gaussians per 256
  uint32_t* cntS;         // cap: key counts in sort order
  uint2* rectS;           // cap: tile rectangles in sort order
  uint32_t* dkS;          // cap: depth bits in id order (global64 only)
  uint32_t* nDepth;       // 1: Gaussians left after the first depth pass drops the culled ones
  uint32_t* tileCnt;      // maxTiles: keys per tile (tile segments)
  uint32_t* segTable;     // rows x maxTiles: per-block key counts, then offsets (tile segments)
  uint32_t* visPart;      // rows: Gaussians with keys per count block (tile segments)
  uint32_t* chunkKeys;    // ceil(cap / kSegChunk): keys per chunk of Gaussians, summed by the
                          // projection, cleared by the scan (tile segments)
  uint32_t* blockStart;   // rows + 1: first Gaussian of each count/scatter block (tile segments)
};

// Most tiles the tile-segment strategy handles (LDS histograms of u32; the
// reference's scan covers 256 x 64 tiles, shaders.js:325); larger views use
// the depth-then-tile radix sort.
constexpr int kMaxSegTiles = 16384;
// Gaussians per count/scatter block of the tile-segment strategy: at least
// kSegMinGroup, so that the table has ceil(n / seg_group(n)) <=
// min(kSegMaxRows, ceil(n / kSegMinGroup)) rows; from 256 * kSegMinGroup
// Gaussians up, n / kSegBlocks, one 1024-thread block per CU of MI355X's 256
constexpr int64_t kSegMinGroup = 2048;
constexpr int64_t kSegMaxRows = 1024;
constexpr int64_t kSegBlocks = 256;
// Gaussians per chunk of the block balancing: count/scatter blocks take whole
// chunks, as many as give each block an equal share of the work (keys plus
// one per Gaussian)
constexpr int64_t kSegChunk = 256;
int64_t seg_group(int64_t n);
int seg_blocks(int64_t n);   // count/scatter blocks of the tile-segment binning

struct BinArgs {
  int mode;               // 0 depth-then-tile, 1 global64, 2 tile segments
  int tilesX, numTiles;
  int tile;               // tile size in pixels (8, 16 or 32)
  int64_t n;              // Gaussians
  uint64_t cap;           // key budget
  const uint32_t* dN;     // device: number of Gaussians
  const uint32_t* cnt;
  const uint2* rect;
  const uint32_t* depthKey;
  const float4* rec;      // raster records: the emit pass derives each key's block mask
  int recFmt;             // record colours: 0 f32 (64 B records), 1 bf16, 2 fp16 (48 B)
  uint32_t *dkA, *dvA, *dkB, *dvB;   // n-sized depth-sort buffers
  void *tkA, *tkB;                    // key buffers (4 or 8 B per key)
  uint32_t *tvA, *tvB;                // id buffers
  uint32_t* numKeys;      // device: min(total, cap)
  uint32_t* numVisible;   // device: Gaussians with at least one key
  uint64_t* totalKeys;    // device: keys before the budget
  uint2* ranges;          // numTiles
  uint32_t* tileOrder;    // numTiles: raster launch order, most keys first
  double* stats;          // [3] += dropped keys
  SortScratch scratch;
};

// One view's tile-segment binning: its inputs (the projection's outputs),
// its scratch and its outputs. The single-view path passes one by value;
// the batched step keeps one per view in device memory and runs every
// binning stage of all views in one launch (the view from the block index).
struct SegJob {
  int64_t n;              // Gaussians
  uint64_t cap;           // key budget
  int nb;                 // count/scatter blocks (ceil(n / seg_group(n)))
  int tilesX, tile, numTiles;
  int recFmt;
  int skip;               // batched only: a view of more than kMaxSegTiles tiles, binned by the
                          // depth-then-tile radix path instead (every stage's blocks leave at once)
  float4* rec;            // raster records (the projection's output)
  uint32_t* depthKey;
  uint2* rect;
  uint32_t* cnt;
  uint32_t* chunkKeys;    // keys per chunk of Gaussians (the projection adds, the scan clears)
  uint32_t* blockStart;   // nb + 1
  uint32_t* table;        // nb x numTiles
  uint32_t* visPart;      // nb
  uint32_t* tileCnt;      // numTiles
  uint2* ranges;          // numTiles
  uint32_t* tileOrder;    // numTiles
  uint32_t* numKeys;
  uint32_t* numVisible;
  uint64_t* totalKeys;
  double* stats;          // [3] += dropped keys (atomically: views run together)
  uint64_t* keys;         // cap: the scattered keys
  uint64_t* tmp;          // cap: merge scratch of segments over 1024 keys
  uint32_t* ids;          // cap: the raster values in sorted order
};

// a SegJob read from device memory, every pointer loaded by gs_global
__device__ __forceinline__ SegJob global_job(const SegJob& s) {
  SegJob r = s;
  r.rec = gs_global(s.rec);
  r.depthKey = gs_global(s.depthKey);
  r.rect = gs_global(s.rect);
  r.cnt = gs_global(s.cnt);
  r.chunkKeys = gs_global(s.chunkKeys);
  r.blockStart = gs_global(s.blockStart);
  r.table = gs_global(s.table);
  r.visPart = gs_global(s.visPart);
  r.tileCnt = gs_global(s.tileCnt);
  r.ranges = gs_global(s.ranges);
  r.tileOrder = gs_global(s.tileOrder);
  r.numKeys = gs_global(s.numKeys);
  r.numVisible = gs_global(s.numVisible);
  r.totalKeys = gs_global(s.totalKeys);
  r.stats = gs_global(s.stats);
  r.keys = gs_global(s.keys);
  r.tmp = gs_global(s.tmp);
  r.ids = gs_global(s.ids);
  return r;
}

// All views' binning stages, one launch each; hJobs: the host copy of the
// device array dJobs (for the launch shapes). Jobs with `skip` set are left
// out (their blocks return at once).
void launch_bin_sort_views(hipStream_t st, const SegJob* dJobs, const SegJob* hJobs, int V);

struct BinOut {
  const void* keys;       // null for tile segments (a key's tile is the range holding it)
  const uint32_t* ids;
  const uint32_t* tileOrder;
  int keyBytes;           // 8 (tile << 32 | depth), 4 (tile), 0 (none)
};

struct RasterBufs {
  int recFmt;             // record colours: 0 f32 (64 B records), 1 bf16, 2 fp16 (48 B)
  const uint2* ranges;
  const uint32_t* tileOrder;   // block b rasterizes tile tileOrder[b]
  const uint32_t* ids;
  const float4* rec;
  float4* img;
  float* T;
  uint32_t* end;
  float4* gC;             // dL/dC, w = valid
  const uint32_t* tgt;    // packed RGBA8 targets (all views)
  double* stats;          // [0] sum |err|^2, [1] loss, [2] valid pixels
};

// a RasterBufs read from device memory, every pointer loaded by gs_global
__device__ __forceinline__ RasterBufs global_bufs(const RasterBufs& s) {
  RasterBufs r;
  r.recFmt = s.recFmt;
  r.ranges = gs_global(s.ranges);
  r.tileOrder = gs_global(s.tileOrder);
  r.ids = gs_global(s.ids);
  r.rec = gs_global(s.rec);
  r.img = gs_global(s.img);
  r.T = gs_global(s.T);
  r.end = gs_global(s.end);
  r.gC = gs_global(s.gC);
  r.tgt = gs_global(s.tgt);
  r.stats = gs_global(s.stats);
  return r;
}

struct UpdateBufs {
  int64_t n, cap;
  float4* P;              // 4 float4 planes
  void *M, *V;            // 4 float4 planes, or 2 bf16 uint4 planes (precision mode 1)
  void *SH, *SHM, *SHV;   // ShDims::P4 float4 planes, or ShDims::Q8 16-bit planes (SH in modes 1-3,
                          // SH moments in mode 1)
  void* accum;            // n x 12 (10 used) i32 or f32
  const uint32_t* cnt;
  float* dbgGradF;        // optional n x 16
  float* dbgShGrad;       // optional n x 3K
};

// ---- batched views (several training views per step) ----
constexpr int kMaxViews = 16;

// One batched step's per-view data, in device memory, rewritten before each
// step: the views' camera constants, raster buffers and gradient sums, the
// work-item offsets of the persistent raster launches and the step's Adam
// constants.
struct ViewBlock {
  int count;                       // views V <= kMaxViews
  SegJob seg[kMaxViews];           // each view's binning (batched binning only)
  uint32_t off[kMaxViews + 1];     // first work item (tile) of each view; off[V] = all
  Cam cams[kMaxViews];
  RasterBufs bufs[kMaxViews];
  void* accum[kMaxViews];          // screen-space gradient sums of each view
  const uint32_t* cnt[kMaxViews];  // tiles covered per Gaussian in each view (0 = culled)
  AdamHP hp;
};

void launch_fill_u32(hipStream_t st, uint32_t* p, int64_t count, uint32_t v);

// chunkKeys (may be null): += each chunk of kSegChunk Gaussians' key count
void launch_project(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const Cam& cam,
                    const Cut& cut, const float4* P, const void* SH, float4* rec, uint32_t* depthKey,
                    uint2* rect, uint32_t* cnt, uint32_t* chunkKeys);

int tile_bits(int numTiles);
void launch_bin_sort(hipStream_t st, const BinArgs& a, BinOut& out);

// Shape of the raster launches: tile side in pixels (8, 16 or 32; the
// camera's tile grid, Cam::tile, must use the same) and keys staged in LDS
// per round (64, 128 or 256).
struct RasterShape {
  int tile, batch;
};
bool raster_shape_ok(int tile, int batch);

// loss: 0 none (view render), 1 Charbonnier fused into the forward pass,
// 2 forward only, loss computed by a separate kernel
void launch_raster_fwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                       int loss);
void launch_loss_l1ssim(hipStream_t st, const Cam& cam, float lambda, const RasterBufs& b,
                        float* scratch);
// boundedColourGrad: every |dL/dC| * 16384 is far below 1e9 (the
// Charbonnier loss), so the colour terms skip the fixed-point clamp
void launch_raster_bwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                       int accumMode, void* accum, bool boundedColourGrad);
// forward with the Charbonnier loss and backward of a training step in one
// launch (the same results as launch_raster_fwd with loss 1, then
// launch_raster_bwd)
void launch_raster_fwd_bwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                           int accumMode, void* accum, bool boundedColourGrad);

void launch_update(hipStream_t st, int deg, int accumMode, int prec, const Cam& cam, const AdamHP& hp,
                   const UpdateBufs& u);

// Batched views: the projection of view v with its camera from vb->cams[v];
// one persistent forward and one persistent backward launch over every
// view's tiles (queue: a counter zeroed before each launch; blocks from
// raster_views_blocks); one update that chains each view's sums through
// its camera, adds the views' gradients and takes one Adam step (vb->hp).
// every view's projection in one launch, into vb->seg[v]'s buffers
void launch_project_views(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const ViewBlock* vb, int V,
                          const Cut& cut, const float4* P, const void* SH);
void launch_project_view(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const ViewBlock* vb, int v,
                         const Cut& cut, const float4* P, const void* SH, float4* rec, uint32_t* depthKey,
                         uint2* rect, uint32_t* cnt, uint32_t* chunkKeys);
// forward with the Charbonnier loss and backward of every view's tiles in one
// persistent launch (blocks: raster_views_blocks)
void launch_raster_fwd_bwd_views(hipStream_t st, const RasterShape& sh, const ViewBlock* vb, const Cut& cut,
                                 int accumMode, bool boundedColourGrad, uint32_t* queue, int blocks);
int raster_views_blocks(int device, const RasterShape& sh);
// view v alone (numTiles blocks, camera and buffers from vb), for the
// overlapped form of the batched step: forward with the loss and backward
void launch_raster_fwd_bwd_view(hipStream_t st, const RasterShape& sh, const ViewBlock* vb, int v, int numTiles,
                                const Cut& cut, int accumMode, bool boundedColourGrad);
// V: the views in vb (sizes the kernel's LDS)
void launch_update_views(hipStream_t st, int deg, int accumMode, int prec, const ViewBlock* vb, int V,
                         const UpdateBufs& u);

// ---- camera pose, focal and exposure gradients (camOpt) ----
struct CamPoseBufs {
  int64_t n, cap;
  const float4* P;        // 4 float4 planes
  void* accum;            // n x 12 screen-space sums (i32 or f32)
  const uint32_t* cnt;    // tiles covered (0 = culled)
  uint32_t* row;          // this camera's gradient row (slots 0-5)
  uint32_t* focal;        // the shared log-focal slot
};

// row: this camera's gradient row; slots 6 (log gain) and 7 (bias) are added to
void launch_cam_exposure(hipStream_t st, const Cam& cam, const float4* img, const uint32_t* tgt, uint32_t* row);
// clear: also zero the screen-space accumulators (evaluation passes)
void launch_cam_pose(hipStream_t st, int accumMode, bool clear, const Cam& cam, const CamPoseBufs& b);

// ---- refine (MCMC-lite relocation and growth) ----

// One spawn of the refine plan: target row gets the donor's row with its
// position jittered (f32(f64(donor pos) + jit)), the planned log-scales and
// opacity logit, slots 14-15 zeroed, fresh Adam moments and the donor's SH
// with fresh SH moments. A split also gives the donor the same log-scales
// and clears the donor's moments and SH moments.
struct RefineOp {
  int64_t target, donor;
  double jit[3];
  float ls[3];
  float opa;
  int32_t split;
  int32_t pad;
};

// q[i] = (opacity logit, log-scale x, y, z) for i < n
void launch_refine_gather(hipStream_t st, int64_t n, int64_t cap, const float4* P, float4* q);
void launch_refine_apply(hipStream_t st, int deg, int prec, int64_t cap, const RefineOp* ops, int64_t count,
                         float4* P, void* M, void* V, void* SH, void* SHM, void* SHV);

// ---- FP64 finite-difference gradient check (gs_fd_grad64) ----
struct FdSample {
  int64_t g;      // Gaussian
  int32_t slot;   // parameter slot 0..13
  int32_t pad;
  double h;       // step
  uint2 rect;     // its tile rectangle in the current raster (packed as the projection's)
};

struct FdArgs {
  int64_t cap;
  const float4* P;          // parameter planes
  const void* SH;           // SH planes (storage of the trainer's precision mode)
  const float4* rec;        // the current raster's records
  int recFmt;
  const uint32_t* depthKey; // f32 depth bits per Gaussian
  const uint32_t* ids;      // sorted key values of the current raster
  const uint2* ranges;      // per tile
  const uint32_t* tgt;      // packed targets
  const FdSample* samples;
  const uint2* items;       // (sample, tile) per block
  double* out;              // per sample: sum over pixels of L(+h) - L(-h) (zeroed by the caller)
};

// One block per item; cam is the view of the current raster.
void launch_fd64(hipStream_t st, int deg, int shFmt, int64_t items, const Cam& cam, const Cut& cut,
                 const FdArgs& a);

// ---- PLY export rows (gs_export_ply) ----
// out: rows first .. first + count - 1, each (17 + 3K) f32, from out[0];
// camPos: nc x 3 f32 (bake only)
void launch_ply_rows(hipStream_t st, int deg, int prec, int64_t first, int64_t count, int64_t cap,
                     const float4* P, const void* SH, int bake, double f, const float* camPos, int nc,
                     float* out);

}  // namespace gs
