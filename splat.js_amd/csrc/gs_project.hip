// gs_project.hip — per-Gaussian projection (the reference's `project` pass,
// shaders.js:228-315): EWA 3D->2D covariance, Mip-Splatting opacity
// compensation, SH colour, opacity-aware radius, tile rectangle.
//
// Output per visible Gaussian:
//   rec            raster record, 64 B with f32 colours: (mx, my, A, B),
//                  (C, opa*comp, L, 0), (r, g, b, opa), (comp, -0.5 cnorm,
//                  1 - opa, 0); 48 B with 16-bit colours (precision modes
//                  bf16 and fp16, rounded to nearest even): (mx, my, A, B),
//                  (C, opa*comp, r | g << 16, b | L16 << 16), (comp,
//                  -0.5 cnorm, 1 - opa, opa), with L16 the smallest fp16 at or
//                  above L. L (gs_cover.h) bounds the exponent at which a pixel
//                  can still take the splat (min(eCut, ln(opa*comp / A_MIN))
//                  plus rounding slack); the raster passes use it only to
//                  skip whole waves, never to change a pixel's value, so any
//                  larger bound is as correct.
//   depthKey[i]    f32 depth bits (positive, so they order like the depths);
//                  0xffffffff when culled, which sorts culled Gaussians last
//   rect[i]        (tx0 | tx1 << 16, ty0 | ty1 << 16)
//   cnt[i]         tiles covered, 0 when culled
// Reads 56-64 B of params + 4*3K B of SH (2*3K B when stored in 16 bits),
// writes 64 B (48 B) of record and 16 B of key/rect/count per visible Gaussian.
#include "gs_kernels.h"

namespace gs {

// A Gaussian's projection inputs that no camera changes: parameters,
// shape, opacity and colour sigmoids (shaders.js:252, :277-280) and the SH
// coefficients. One view at a time they are computed inside project_one for
// the Gaussians that pass the culls; the batched launch computes them once
// per Gaussian for all of its views.
template <int kDeg>
struct ProjFixed {
  float p[16];
  GeomShape sp;
  float opa, sig[3];
  float sh[ShDims<kDeg>::S > 0 ? ShDims<kDeg>::S : 1];
};

template <int kDeg, int kPrec>
__device__ __forceinline__ void proj_fixed(int64_t cap, const float4* __restrict__ P, const void* __restrict__ SH,
                                           int64_t i, ProjFixed<kDeg>& pf) {
  load_params(P, cap, i, pf.p);
  compute_shape(pf.p, pf.sp);
  pf.opa = gs_sigmoid_clamped(pf.p[13]);
  for (int c = 0; c < 3; c++) pf.sig[c] = gs_sigmoid_clamped(pf.p[10 + c]);
  if constexpr (kDeg > 0) load_sh_any<kDeg, Prec<kPrec>::sh>(SH, cap, i, pf.sh);
}

// kPre: pf holds proj_fixed's values; otherwise project_one loads and
// computes them itself, the parameters first and the rest only for a
// Gaussian that passes the culls. Same values either way.
template <int kDeg, int kPrec, bool kPre>
__device__ __forceinline__ uint32_t project_one(int64_t cap, const Cam& cam, const Cut& cut,
                                                const float4* __restrict__ P, const void* __restrict__ SH,
                                                ProjFixed<kDeg>& pf, float4* __restrict__ rec,
                                                uint32_t* __restrict__ depthKey, uint2* __restrict__ rect,
                                                uint32_t* __restrict__ cnt, int64_t i) {
  if constexpr (!kPre) load_params(P, cap, i, pf.p);
  const float* p = pf.p;
  Geom g;
  if constexpr (kPre) compute_geom(cam, p, pf.sp, g);
  else compute_geom(cam, p, g);
  uint32_t key = 0xffffffffu, c = 0;
  uint2 rc = make_uint2(0, 0);
  while (g.ok) {  // single pass; `break` = culled (shaders.js:243-266)
    const float detV = fmaxf(g.va * g.vc - g.vb * g.vb, 0.0f);
    const float ad = g.va + 0.3f;
    const float cd = g.vc + 0.3f;
    const float detVd = ad * cd - g.vb * g.vb;
    if (detVd < 1e-8f) break;
    const float comp = sqrtf(fmaxf(detV / detVd, 0.0f));
    const float opa = kPre ? pf.opa : gs_sigmoid_clamped(p[13]);
    if (opa * comp < cut.aMin) break;
    const float f = cam.f;
    const float mx = f * g.pc[0] / g.pc[2] + cam.cx;
    const float my = f * g.pc[1] / g.pc[2] + cam.cy;
    const float mid = 0.5f * (ad + cd);
    const float disc = sqrtf(fmaxf(mid * mid - detVd, 0.0f));
    const float eMax = fminf(cut.eCut, gs_logf(fmaxf(opa * comp / cut.aMin, 1.0001f)));
    const float W = (float)cam.W, H = (float)cam.H;
    const float rad = fminf(sqrtf(2.0f * eMax) * sqrtf(mid + disc), cut.radClamp * W);
    if (mx + rad < 0.0f || my + rad < 0.0f || mx - rad > W || my - rad > H) break;
    const float inv = 1.0f / detVd;
    float col[3];
    for (int k = 0; k < 3; k++) col[k] = kPre ? pf.sig[k] : gs_sigmoid_clamped(p[10 + k]);
    if constexpr (kDeg > 0) {
      constexpr int K = ShDims<kDeg>::K;
      float v[3];
      view_dir(cam, p, v);
      float Y[15];
      sh_basis<kDeg, false>(v, Y, nullptr);
      if constexpr (!kPre) load_sh_any<kDeg, Prec<kPrec>::sh>(SH, cap, i, pf.sh);
      const float* sh = pf.sh;
      const int aK = cam.shActiveK;
#pragma unroll
      for (int k = 0; k < K; k++) {
        if (k < aK) {
          col[0] = col[0] + sh[k] * Y[k];
          col[1] = col[1] + sh[K + k] * Y[k];
          col[2] = col[2] + sh[2 * K + k] * Y[k];
        }
      }
      col[0] = fmaxf(col[0], 0.0f);
      col[1] = fmaxf(col[1], 0.0f);
      col[2] = fmaxf(col[2], 0.0f);
    }
    const float cnorm = 1.0f / (1.0f + conic_lmax(g.va, g.vb, g.vc));
    if constexpr (Prec<kPrec>::rec == 0) {
      rec[4 * i + 0] = make_float4(mx, my, cd * inv, (-g.vb) * inv);
      rec[4 * i + 1] = make_float4(ad * inv, opa * comp, gs_cover_level(eMax), 0.0f);
      rec[4 * i + 2] = make_float4(col[0], col[1], col[2], opa);
      rec[4 * i + 3] = make_float4(comp, -0.5f * cnorm, 1.0f - opa, 0.0f);
    } else {
      constexpr int F = Prec<kPrec>::rec;
      const uint32_t rg = gs_f_to_16_rne<F>(col[0]) | (gs_f_to_16_rne<F>(col[1]) << 16);
      const uint32_t bl = gs_f_to_16_rne<F>(col[2]) | (gs_f2h_up(gs_cover_level(eMax)) << 16);
      rec[3 * i + 0] = make_float4(mx, my, cd * inv, (-g.vb) * inv);
      rec[3 * i + 1] = make_float4(ad * inv, opa * comp, __uint_as_float(rg), __uint_as_float(bl));
      rec[3 * i + 2] = make_float4(comp, -0.5f * cnorm, 1.0f - opa, opa);
    }
    key = gs_f2u(g.pc[2]);
    // tile = 2^k: x * (1 / tile) rounds exactly as x / tile (the same real
    // value, one rounding), so the reference's division is kept bit for bit
    const float rTF = 1.0f / (float)cam.tile;
    const int tx0 = (int)clampf(floorf((mx - rad) * rTF), 0.0f, (float)(cam.tilesX - 1));
    const int tx1 = (int)clampf(floorf((mx + rad) * rTF), 0.0f, (float)(cam.tilesX - 1));
    const int ty0 = (int)clampf(floorf((my - rad) * rTF), 0.0f, (float)(cam.tilesY - 1));
    const int ty1 = (int)clampf(floorf((my + rad) * rTF), 0.0f, (float)(cam.tilesY - 1));
    rc = make_uint2((uint32_t)tx0 | ((uint32_t)tx1 << 16), (uint32_t)ty0 | ((uint32_t)ty1 << 16));
    c = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
    break;
  }
  depthKey[i] = key;
  rect[i] = rc;
  cnt[i] = c;
  return c;
}

// += the wave's keys to its chunk's count (the wave's 64 Gaussians lie in one
// chunk: kSegChunk is a multiple of 64); every lane of the wave takes part
__device__ __forceinline__ void add_chunk_keys(uint32_t* __restrict__ chunkKeys, int64_t i, uint32_t c) {
  if (!chunkKeys) return;
  uint32_t s = c;
  for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o);
  const int64_t i0 = i - (threadIdx.x & 63);
  if ((threadIdx.x & 63) == 0 && s != 0u) atomicAdd(&chunkKeys[i0 / kSegChunk], s);
}

template <int kDeg, int kPrec>
__device__ __forceinline__ void project_thread(int64_t n, int64_t cap, const Cam& cam, const Cut& cut,
                                               const float4* __restrict__ P, const void* __restrict__ SH,
                                               float4* __restrict__ rec, uint32_t* __restrict__ depthKey,
                                               uint2* __restrict__ rect, uint32_t* __restrict__ cnt,
                                               uint32_t* __restrict__ chunkKeys) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  ProjFixed<kDeg> pf;
  const uint32_t c =
      i < n ? project_one<kDeg, kPrec, false>(cap, cam, cut, P, SH, pf, rec, depthKey, rect, cnt, i) : 0u;
  add_chunk_keys(chunkKeys, i, c);
}

template <int kDeg, int kPrec>
__global__ __launch_bounds__(kBlock) void k_project(int64_t n, int64_t cap, Cam cam, Cut cut,
                                                    const float4* __restrict__ P,
                                                    const void* __restrict__ SH,
                                                    float4* __restrict__ rec,
                                                    uint32_t* __restrict__ depthKey,
                                                    uint2* __restrict__ rect,
                                                    uint32_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ chunkKeys) {
  project_thread<kDeg, kPrec>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys);
}

// view v of a batched step: the camera constants come from device memory
template <int kDeg, int kPrec>
__global__ __launch_bounds__(kBlock) void k_project_view(int64_t n, int64_t cap, const ViewBlock* __restrict__ vb,
                                                         int v, Cut cut, const float4* __restrict__ P,
                                                         const void* __restrict__ SH, float4* __restrict__ rec,
                                                         uint32_t* __restrict__ depthKey, uint2* __restrict__ rect,
                                                         uint32_t* __restrict__ cnt,
                                                         uint32_t* __restrict__ chunkKeys) {
  const Cam cam = vb->cams[v];
  project_thread<kDeg, kPrec>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys);
}

// Waves per SIMD the batched projection is compiled for: up to SH degree 2
// it fits 72 VGPRs, 7 waves (left free it took 79, 6 waves, and 1.7% more
// time); at degree 3 the 48 coefficients it keeps for all views need ~108.
template <int kDeg>
constexpr int kProjViewsWaves = kDeg <= 2 ? 7 : 4;

// every view of a batched step in one launch: one thread per Gaussian loads
// its parameters and SH and computes the camera-independent terms once,
// then projects them through views 0..V-1 in turn (cameras from vb->cams,
// buffers from vb->seg)
template <int kDeg, int kPrec>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kProjViewsWaves<kDeg>))) void k_project_views(int64_t n, int64_t cap, const ViewBlock* __restrict__ vb,
                                                          int V, Cut cut, const float4* __restrict__ P,
                                                          const void* __restrict__ SH) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  ProjFixed<kDeg> pf;
  if (i < n) proj_fixed<kDeg, kPrec>(cap, P, SH, i, pf);
  for (int v = 0; v < V; v++) {
    const Cam& cam = vb->cams[v];
    const SegJob& j = vb->seg[v];
    const uint32_t c = i < n ? project_one<kDeg, kPrec, true>(cap, cam, cut, P, SH, pf, gs_global(j.rec),
                                                              gs_global(j.depthKey), gs_global(j.rect),
                                                              gs_global(j.cnt), i)
                             : 0u;
    add_chunk_keys(gs_global(j.chunkKeys), i, c);
  }
}

template <int kPrec>
static void launch_project_views_t(hipStream_t st, int deg, int64_t n, int64_t cap, const ViewBlock* vb, int V,
                                   const Cut& cut, const float4* P, const void* SH) {
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  switch (deg) {
    case 0: k_project_views<0, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, V, cut, P, SH); break;
    case 1: k_project_views<1, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, V, cut, P, SH); break;
    case 2: k_project_views<2, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, V, cut, P, SH); break;
    default: k_project_views<3, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, V, cut, P, SH); break;
  }
}

void launch_project_views(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const ViewBlock* vb, int V,
                          const Cut& cut, const float4* P, const void* SH) {
  if (n <= 0 || V <= 0) return;
  switch (prec) {
    case 0: launch_project_views_t<0>(st, deg, n, cap, vb, V, cut, P, SH); break;
    case 1: launch_project_views_t<1>(st, deg, n, cap, vb, V, cut, P, SH); break;
    case 2: launch_project_views_t<2>(st, deg, n, cap, vb, V, cut, P, SH); break;
    default: launch_project_views_t<3>(st, deg, n, cap, vb, V, cut, P, SH); break;
  }
}

template <int kPrec>
static void launch_project_t(hipStream_t st, int deg, int64_t n, int64_t cap, const Cam& cam, const Cut& cut,
                             const float4* P, const void* SH, float4* rec, uint32_t* depthKey, uint2* rect,
                             uint32_t* cnt, uint32_t* chunkKeys) {
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  switch (deg) {
    case 0: k_project<0, kPrec><<<grid, kBlock, 0, st>>>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 1: k_project<1, kPrec><<<grid, kBlock, 0, st>>>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 2: k_project<2, kPrec><<<grid, kBlock, 0, st>>>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    default: k_project<3, kPrec><<<grid, kBlock, 0, st>>>(n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
  }
}

template <int kPrec>
static void launch_project_view_t(hipStream_t st, int deg, int64_t n, int64_t cap, const ViewBlock* vb, int v,
                                  const Cut& cut, const float4* P, const void* SH, float4* rec, uint32_t* depthKey,
                                  uint2* rect, uint32_t* cnt, uint32_t* chunkKeys) {
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  switch (deg) {
    case 0: k_project_view<0, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 1: k_project_view<1, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 2: k_project_view<2, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    default: k_project_view<3, kPrec><<<grid, kBlock, 0, st>>>(n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
  }
}

void launch_project_view(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const ViewBlock* vb, int v,
                         const Cut& cut, const float4* P, const void* SH, float4* rec, uint32_t* depthKey,
                         uint2* rect, uint32_t* cnt, uint32_t* chunkKeys) {
  if (n <= 0) return;
  switch (prec) {
    case 0: launch_project_view_t<0>(st, deg, n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 1: launch_project_view_t<1>(st, deg, n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 2: launch_project_view_t<2>(st, deg, n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    default: launch_project_view_t<3>(st, deg, n, cap, vb, v, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
  }
}

void launch_project(hipStream_t st, int deg, int prec, int64_t n, int64_t cap, const Cam& cam,
                    const Cut& cut, const float4* P, const void* SH, float4* rec, uint32_t* depthKey,
                    uint2* rect, uint32_t* cnt, uint32_t* chunkKeys) {
  if (n <= 0) return;
  switch (prec) {
    case 0: launch_project_t<0>(st, deg, n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 1: launch_project_t<1>(st, deg, n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    case 2: launch_project_t<2>(st, deg, n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
    default: launch_project_t<3>(st, deg, n, cap, cam, cut, P, SH, rec, depthKey, rect, cnt, chunkKeys); break;
  }
}

}  // namespace gs
