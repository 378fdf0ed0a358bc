// gs_raster.hip — per-tile alpha compositing, forward and backward
// (the reference's `render` pass, shaders.js:511-702).
//
// One 256-thread workgroup (4 waves) per 16x16 tile, one thread per pixel,
// each wave an 8x8 block. The tile's sorted keys are consumed in batches of
// 256: every thread loads one raster record into LDS, then every pixel walks
// the batch from LDS (broadcast reads). A wave skips, as a whole, every splat
// whose alpha >= A_MIN ellipse misses its 8x8 block (the pixels it skips
// could not have taken the splat). The forward stops the whole tile once
// every pixel has reached T < 1e-4.
//
// Forward (shaders.js:556-585) writes the image, the final transmittance T
// and `end` (one past the last key the pixel processed); in training it also
// evaluates the Charbonnier loss per pixel (shaders.js:589-618) and stores
// dL/dC, so the backward pass needs no second loss read.
//
// Backward (shaders.js:620-700) walks each tile from the largest `end` of
// its pixels down to the first key, recomputing each pixel's transmittance
// in front of a splat from the stored final T (Tb = Ta/(1-a), through one
// reciprocal that also scales the colour-behind sum S).
// Per key, the 10 screen-space gradient terms of the wave's 64 pixels are
// summed with a transpose reduction (13 lane shuffles instead of 60) and
// added to the Gaussian's accumulator with one atomic per term per wave:
//   GS_ACCUM_FIXED: each pixel's term is converted to i32 first,
//                   i32(clamp(v * 16384, -1e9, 1e9)), exactly as the
//                   reference's atomicAdd on atomic<i32> (shaders.js:529-541);
//                   integer sums commute, so the result is deterministic and
//                   bit-identical to any other summation order.
//   GS_ACCUM_F32:   plain f32 sums (atomic add order is not deterministic).
#include "gs_kernels.h"

namespace gs {

__device__ __forceinline__ int tile_of_block(const Cam& cam) { return (int)blockIdx.x; }

// Each wave owns an 8x8 pixel block of the tile (wave w: columns 8*(w&1),
// rows 8*(w>>1)); lane l is pixel (l & 7, l >> 3) of that block.
struct PixelMap {
  int x, y;        // pixel
  float bx0, by0;  // centre of the block's first pixel
};
__device__ __forceinline__ PixelMap pixel_of_thread(const Cam& cam, int tile, int tid) {
  const int tx = tile % cam.tilesX, ty = tile / cam.tilesX;
  const int w = tid >> 6, l = tid & 63;
  const int bx = tx * kTile + (w & 1) * 8, by = ty * kTile + (w >> 1) * 8;
  PixelMap m;
  m.x = bx + (l & 7);
  m.y = by + (l >> 3);
  m.bx0 = (float)bx + 0.5f;
  m.by0 = (float)by + 0.5f;
  return m;
}

// true when the splat's alpha >= A_MIN ellipse (bounding box mx +- ex,
// my +- ey) cannot reach any pixel centre of the wave's 8x8 block
__device__ __forceinline__ bool wave_misses(float mx, float my, float ex, float ey, const PixelMap& m) {
  return mx + ex < m.bx0 || mx - ex > m.bx0 + 7.0f || my + ey < m.by0 || my - ey > m.by0 + 7.0f;
}

template <int kLoss>
__global__ __launch_bounds__(kBlock) void k_raster_fwd(Cam cam, Cut cut, RasterBufs b) {
  __shared__ float4 sA[kBlock];   // mx, my, A, B
  __shared__ float4 sB[kBlock];   // C, opa*comp, ex, ey
  __shared__ float4 sC[kBlock];   // r, g, b, opa
  __shared__ double sStat[3][kBlock / 64];
  const int tid = threadIdx.x;
  const int tile = tile_of_block(cam);
  const PixelMap pm = pixel_of_thread(cam, tile, tid);
  const int x = pm.x, y = pm.y;
  const bool inside = x < cam.W && y < cam.H;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const uint2 rg = b.ranges[tile];
  const uint32_t s = rg.x, e = rg.y;
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  uint32_t end = s;
  bool done = !inside;
  for (uint32_t base = s; base < e; base += kBlock) {
    if (__syncthreads_count(!done) == 0) break;
    const uint32_t k = base + tid;
    if (k < e) {
      const uint32_t id = b.ids[k];
      sA[tid] = b.rec[4 * id + 0];
      sB[tid] = b.rec[4 * id + 1];
      sC[tid] = b.rec[4 * id + 2];
    }
    __syncthreads();
    if (!done) {
      const int cnt = (int)min((uint32_t)kBlock, e - base);
      int j = 0;
      for (; j < cnt; j++) {
        const float4 a = sA[j];
        const float4 bb = sB[j];
        if (wave_misses(a.x, a.y, bb.z, bb.w, pm)) continue;
        const float dx = px - a.x, dy = py - a.y;
        const float t2 = fmaf(bb.x * dy, dy, (a.z * dx) * dx);
        const float ee = fmaxf(fmaf(0.5f, t2, (a.w * dx) * dy), 0.0f);
        if (ee > cut.eCut) continue;
        const float araw = bb.y * gs_expf(-ee);
        const float alpha = fminf(0.99f, araw);
        if (alpha < cut.aMin) continue;
        const float4 c4 = sC[j];
        const float w = T * alpha;
        C0 = fmaf(w, c4.x, C0);
        C1 = fmaf(w, c4.y, C1);
        C2 = fmaf(w, c4.z, C2);
        T = T * (1.0f - alpha);
        if (T < 1e-4f) { done = true; break; }
      }
      end = done ? base + (uint32_t)j + 1u : base + (uint32_t)cnt;
    }
  }
  double st0 = 0.0, st1 = 0.0, st2 = 0.0;
  if (inside) {
    const int64_t pi = (int64_t)y * cam.W + x;
    const float o0 = C0 + T * cam.bg[0], o1 = C1 + T * cam.bg[1], o2 = C2 + T * cam.bg[2];
    b.img[pi] = make_float4(o0, o1, o2, 1.0f - T);
    if (kLoss) {
      b.T[pi] = T;
      b.end[pi] = end;
    }
    if (kLoss == 1) {
      const uint32_t pk = b.tgt[cam.tgtOffset + pi];
      float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if ((pk >> 24) != 0u) {
        const float o[3] = {o0, o1, o2};
        float err[3], root[3], gg[3];
        for (int c = 0; c < 3; c++) {
          const float tc = (float)((pk >> (8 * c)) & 0xffu) / 255.0f;
          err[c] = (cam.gain * o[c] + cam.bias) - tc;
          root[c] = sqrtf(err[c] * err[c] + 0.0009f);
          gg[c] = cam.gain * (err[c] / root[c]);
        }
        g = make_float4(gg[0], gg[1], gg[2], 1.0f);
        st0 = (double)dot3(err, err);
        st1 = (double)(((root[0] + root[1]) + root[2]) - 0.09f);
        st2 = 1.0;
      }
      b.gC[pi] = g;
    }
  }
  if (kLoss == 1) {
    for (int o = 32; o > 0; o >>= 1) {
      st0 += __shfl_xor(st0, o);
      st1 += __shfl_xor(st1, o);
      st2 += __shfl_xor(st2, o);
    }
    const int w = tid >> 6;
    if ((tid & 63) == 0) { sStat[0][w] = st0; sStat[1][w] = st1; sStat[2][w] = st2; }
    __syncthreads();
    if (tid < 3) {
      const double v = sStat[tid][0] + sStat[tid][1] + sStat[tid][2] + sStat[tid][3];
      if (v != 0.0) atomicAdd(&b.stats[tid], v);
    }
  }
}

// ---- wave transpose reduction of 10 values ----
// Step masks 32, 16, 8, 4 halve the value count (10 -> 5 -> 3 -> 2 -> 1) by
// swapping halves with the partner lane; masks 2 and 1 finish the sum. Lane
// l ends with the 64-lane total of slot `sid` (or sid < 0: padding).
template <typename T>
__device__ __forceinline__ T shfl_x(T v, int m);
template <>
__device__ __forceinline__ uint32_t shfl_x<uint32_t>(uint32_t v, int m) {
  return (uint32_t)__shfl_xor((int)v, m);
}
template <>
__device__ __forceinline__ float shfl_x<float>(float v, int m) {
  return __shfl_xor(v, m);
}

template <typename T, int N>
__device__ __forceinline__ void halve(const T* v, const int* sid, T* out, int* osid, int mask) {
  constexpr int H = (N + 1) / 2;
  const bool hi = (__lane_id() & mask) != 0;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const bool hasHi = i + H < N;
    const T lo_v = v[i];
    const T hi_v = hasHi ? v[i + H] : (T)0;
    const T keep = hi ? hi_v : lo_v;
    const T send = hi ? lo_v : hi_v;
    out[i] = keep + shfl_x<T>(send, mask);
    osid[i] = hi ? (hasHi ? sid[i + H] : -1) : sid[i];
  }
}

template <typename T>
__device__ __forceinline__ T wave_reduce10(const T* v, int& slot) {
  int s0[10];
#pragma unroll
  for (int i = 0; i < 10; i++) s0[i] = i;
  T a[5], b3[3], c[2], d[1];
  int sa[5], sb[3], sc[2], sd[1];
  halve<T, 10>(v, s0, a, sa, 32);
  halve<T, 5>(a, sa, b3, sb, 16);
  halve<T, 3>(b3, sb, c, sc, 8);
  halve<T, 2>(c, sc, d, sd, 4);
  T r = d[0];
  r = r + shfl_x<T>(r, 2);
  r = r + shfl_x<T>(r, 1);
  slot = sd[0];
  return r;
}

template <int kAccum>
__global__ __launch_bounds__(kBlock) void k_raster_bwd(Cam cam, Cut cut, RasterBufs b, void* accum) {
  __shared__ float4 sA[kBlock];   // mx, my, A, B
  __shared__ float4 sB[kBlock];   // C, opa*comp, ex, ey
  __shared__ float4 sC[kBlock];   // r, g, b, opa
  __shared__ float2 sD[kBlock];   // comp, cnorm
  __shared__ uint32_t sId[kBlock];
  __shared__ uint32_t sEnd;
  const int tid = threadIdx.x;
  const int tile = tile_of_block(cam);
  const PixelMap pm = pixel_of_thread(cam, tile, tid);
  const int x = pm.x, y = pm.y;
  const bool inside = x < cam.W && y < cam.H;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const uint2 rg = b.ranges[tile];
  const uint32_t s = rg.x;
  const int64_t pi = (int64_t)y * cam.W + x;
  float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float Ta = 0.0f;
  uint32_t myEnd = s;
  if (inside) {
    g = b.gC[pi];
    if (g.w != 0.0f) {
      Ta = b.T[pi];
      myEnd = b.end[pi];
    }
  }
  const bool active = inside && g.w != 0.0f;
  if (tid == 0) sEnd = s;
  __syncthreads();
  if (myEnd > s) atomicMax(&sEnd, myEnd);
  __syncthreads();
  const uint32_t endMax = sEnd;
  float S0 = cam.bg[0] * Ta, S1 = cam.bg[1] * Ta, S2 = cam.bg[2] * Ta;
  const float gC[3] = {g.x, g.y, g.z};
  for (uint32_t hi = endMax; hi > s; ) {
    const uint32_t lo = hi - s > (uint32_t)kBlock ? hi - kBlock : s;
    const int n = (int)(hi - lo);
    __syncthreads();
    if (tid < n) {
      const uint32_t id = b.ids[lo + tid];
      sId[tid] = id;
      sA[tid] = b.rec[4 * id + 0];
      sB[tid] = b.rec[4 * id + 1];
      sC[tid] = b.rec[4 * id + 2];
      const float4 d = b.rec[4 * id + 3];
      sD[tid] = make_float2(d.x, d.y);
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; j--) {
      const float4 a = sA[j];
      const float4 bb = sB[j];
      if (wave_misses(a.x, a.y, bb.z, bb.w, pm)) continue;
      const uint32_t k = lo + (uint32_t)j;
      bool contrib = false;
      float v[10];
#pragma unroll
      for (int q = 0; q < 10; q++) v[q] = 0.0f;
      if (active && k < myEnd) {
        const float dx = px - a.x, dy = py - a.y;
        const float cA = a.z, cB = a.w, cC = bb.x;
        const float t2 = fmaf(cC * dy, dy, (cA * dx) * dx);
        const float e = fmaxf(fmaf(0.5f, t2, (cB * dx) * dy), 0.0f);
        if (e <= cut.eCut) {
          const float G = gs_expf(-e);
          const float araw = bb.y * G;
          const float alpha = fminf(0.99f, araw);
          if (alpha >= cut.aMin) {
            contrib = true;
            const float4 c4 = sC[j];
            const float2 d2 = sD[j];
            const float col[3] = {c4.x, c4.y, c4.z};
            const float opa = c4.w, cnorm = d2.y;
            // Tb = Ta / (1 - alpha), S / (1 - alpha) through one reciprocal
            const float rom = 1.0f / (1.0f - alpha);
            const float Tb = Ta * rom;
            const float aT = alpha * Tb;
            const float u[3] = {col[0] * Tb - S0 * rom, col[1] * Tb - S1 * rom, col[2] * Tb - S2 * rom};
            float galpha = dot3(gC, u);
            if (araw > 0.99f) galpha = 0.0f;
            const float ga = galpha * araw;
            v[0] = ga * (cA * dx + cB * dy);
            v[1] = ga * (cB * dx + cC * dy);
            v[2] = (-ga) * 0.5f * dx * dx * cnorm;
            v[3] = (-ga) * dx * dy * cnorm;
            v[4] = (-ga) * 0.5f * dy * dy * cnorm;
            v[5] = galpha * opa * G;
            v[6] = ga * (1.0f - opa);
            v[7] = gC[0] * aT;
            v[8] = gC[1] * aT;
            v[9] = gC[2] * aT;
            S0 = S0 + col[0] * alpha * Tb;
            S1 = S1 + col[1] * alpha * Tb;
            S2 = S2 + col[2] * alpha * Tb;
            Ta = Tb;
          }
        }
      }
      if (__ballot(contrib) == 0ull) continue;
      const uint32_t gid = sId[j];
      int slot;
      if (kAccum == 0) {
        uint32_t iv[10];
#pragma unroll
        for (int q = 0; q < 10; q++) {
          const float cl = fminf(fmaxf(v[q] * kFixed, -1.0e9f), 1.0e9f);
          iv[q] = contrib ? (uint32_t)(int32_t)cl : 0u;
        }
        const uint32_t r = wave_reduce10<uint32_t>(iv, slot);
        if ((__lane_id() & 3) == 0 && slot >= 0 && r != 0u)
          atomicAdd((uint32_t*)accum + (int64_t)gid * 12 + slot, r);
      } else {
        const float r = wave_reduce10<float>(v, slot);
        if ((__lane_id() & 3) == 0 && slot >= 0 && r != 0.0f)
          unsafeAtomicAdd((float*)accum + (int64_t)gid * 12 + slot, r);
      }
    }
    hi = lo;
  }
}

void launch_raster_fwd(hipStream_t st, const Cam& cam, const Cut& cut, const RasterBufs& b, int loss) {
  const dim3 grid((unsigned)cam.numTiles);
  if (loss == 1) k_raster_fwd<1><<<grid, kBlock, 0, st>>>(cam, cut, b);
  else if (loss == 2) k_raster_fwd<2><<<grid, kBlock, 0, st>>>(cam, cut, b);
  else k_raster_fwd<0><<<grid, kBlock, 0, st>>>(cam, cut, b);
}

void launch_raster_bwd(hipStream_t st, const Cam& cam, const Cut& cut, const RasterBufs& b,
                       int accumMode, void* accum) {
  const dim3 grid((unsigned)cam.numTiles);
  if (accumMode == 0) k_raster_bwd<0><<<grid, kBlock, 0, st>>>(cam, cut, b, accum);
  else k_raster_bwd<1><<<grid, kBlock, 0, st>>>(cam, cut, b, accum);
}

}  // namespace gs
