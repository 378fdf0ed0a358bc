// gs_raster.hip — per-tile alpha compositing, forward and backward
// (the reference's `render` pass, shaders.js:511-702).
//
// One workgroup per tile, one thread per pixel, each wave an 8x8 block: by
// default 16x16 tiles (256 threads, 4 waves) and batches of 256 keys; tiles
// of 8 or 32 pixels and batches of 64 or 128 are the other shapes (Shape
// below). The tile's sorted keys are consumed in batches: the waves load
// the batch's raster records into LDS (lane l of a wave one record of each
// 64-key chunk it stages), then every pixel walks the batch from LDS
// (broadcast reads). Each key carries the mask of the tile's quadrants whose
// pixels may take the splat (gs_cover.h, set by the binning pass): a wave
// visits only the splats of its quadrant's bit, a splat with an empty mask
// is not even loaded, and a wave leaves a splat early once no lane is inside
// the cutoff or takes it. The batch size changes no result (each pixel still
// walks its keys in order); the tile size changes which keys exist (the
// tile rectangles and the entry budget are per tile), as in the restatement
// run at the same tile size. The forward stops the whole tile once
// every pixel has reached T < 1e-4. Both passes evaluate alpha with
// gs_expf_raster (gs_math.h), as the restatement's raster passes do.
//
// Forward (shaders.js:556-585) writes the image, the final transmittance T
// and `end` (one past the last key the pixel processed); in training it also
// evaluates the Charbonnier loss per pixel (shaders.js:589-618) and stores
// dL/dC, so the backward pass needs no second loss read.
//
// Backward (shaders.js:620-700) walks each tile from the largest `end` of
// its pixels down to the first key (each wave from the largest `end` of its
// own pixels), recomputing each pixel's transmittance
// in front of a splat from the stored final T (Tb = Ta/(1-a), through one
// reciprocal that also scales the colour-behind sum S).
// Per key, the 10 screen-space gradient terms of the wave's 64 pixels are
// summed with a transpose reduction of lane exchanges (below) and added to
// the Gaussian's accumulator with one atomic per term per wave:
//   GS_ACCUM_FIXED: each pixel's term is converted to i32 first,
//                   i32(clamp(v * 16384, -1e9, 1e9)), exactly as the
//                   reference's atomicAdd on atomic<i32> (shaders.js:529-541);
//                   integer sums commute, so the result is deterministic and
//                   bit-identical to any other summation order. (Under the
//                   Charbonnier loss the three colour terms cannot reach the
//                   clamp and convert without it.)
//   GS_ACCUM_F32:   plain f32 sums (atomic add order is not deterministic).
//
// Training under the Charbonnier loss runs both passes in one launch: the
// loss gradient of a pixel needs only that pixel, so a block runs its tile's
// backward right after the forward (k_raster_fwd_bwd, and the persistent
// k_raster_fwd_bwd_views over a batched step's views). The two passes share
// one LDS array, and the backward stages the tile's records again while
// they are still in L2 from the forward. Renders, and the L1 + D-SSIM loss
// (whose gradient needs neighbouring pixels), run the passes apart.
#include <type_traits>

#include "gs_kernels.h"

namespace gs {

__device__ __forceinline__ int tile_of_block(const RasterBufs& b) { return (int)b.tileOrder[blockIdx.x]; }

// Shape of a raster launch: kT x kT pixel tiles, one thread per pixel, in
// waves of 8x8 pixel blocks (wave w: block column w % (kT/8), block row
// w / (kT/8); lane l is pixel (l & 7, l >> 3) of its block), and kB keys
// staged in LDS per round, in chunks of 64 keys (one ballot word each).
// Each key carries 4 cover bits, one per quadrant of the tile (kT/2 x kT/2
// pixels, bit 2 row + column): a wave tests its quadrant's bit (at kT = 16
// the quadrant is the wave's own block; at kT = 32 four waves share one; at
// kT = 8 the one wave spans all four and tests whether any bit is set).
template <int kT, int kB>
struct Shape {
  static constexpr int T = kT, B = kB;
  static constexpr int kThreads = kT * kT;
  static constexpr int kWaves = kThreads / 64;
  static constexpr int kChunks = kB / 64;
  static constexpr int kBW = kT / 8;            // 8x8 blocks per tile row
  static constexpr int kMasks = kT == 8 ? 1 : 4;
  // occupancy the persistent forward-and-backward launch is compiled for: 8
  // waves per SIMD (at most 64 VGPRs; left free, the compiler takes up to 80
  // there and the grid, sized by occupancy, shrinks); 8-pixel tiles with
  // 128- or 256-key batches are held below that by their LDS anyway
  static constexpr int kBwdWavesPerEU = (kT == 8 && kB > 64) ? 1 : 8;
  static_assert((kT == 8 || kT == 16 || kT == 32) && (kB == 64 || kB == 128 || kB == 256), "raster shape");
  __device__ static int mask_slot(int w) {
    if constexpr (kT == 8) {
      return 0;
    } else {
      constexpr int h = kBW / 2;   // blocks per quadrant row
      return 2 * ((w / kBW) / h) + (w % kBW) / h;
    }
  }
};

// LDS of the raster passes, carved from one array that the kernel declares,
// so a launch that runs both passes holds the larger layout, not both.
// Forward: records sA, sB, sC (16 B each per key), the chunk masks, the loss
// partials; backward: records sA-sD, the chunk masks, the block's largest
// end. Sizes in float4.
template <class S>
struct RasterLds {
  static constexpr int kMaskBytes = S::kChunks * S::kMasks * 8;
  static constexpr int kFwd = 3 * S::B + (kMaskBytes + 3 * S::kWaves * 8 + 15) / 16;
  static constexpr int kBwd = 4 * S::B + (kMaskBytes + 4 + 15) / 16;
  static constexpr int kBoth = kFwd > kBwd ? kFwd : kBwd;
};

struct PixelMap {
  int x, y;        // pixel
};

// What the forward with the loss leaves for the backward at a thread's
// pixel: dL/dC (w = 1 for a valid target pixel, 0 otherwise, and 0 outside
// the image), the final T and `end`. Handed over in registers when both
// passes run in one launch.
struct PixelState {
  float4 g;
  float T;
  uint32_t end;
};
template <class S>
__device__ __forceinline__ PixelMap pixel_of_thread(const Cam& cam, int tile, int tid) {
  const int tx = tile % cam.tilesX, ty = tile / cam.tilesX;
  const int w = tid >> 6, l = tid & 63;
  PixelMap m;
  m.x = tx * S::T + (w % S::kBW) * 8 + (l & 7);
  m.y = ty * S::T + (w / S::kBW) * 8 + (l >> 3);
  return m;
}

// Stores the staged chunk's ballots of the cover bits the waves test (lane
// l of the calling wave holds key 64 c + l of the batch; keys past the
// segment end have cover 0 and contribute no bits).
template <class S>
__device__ __forceinline__ void store_masks(uint32_t cover, uint64_t* m) {
  const bool lead = (threadIdx.x & 63) == 0;
  if constexpr (S::kMasks == 1) {
    const uint64_t bal = wave_ballot(cover != 0u);
    if (lead) m[0] = bal;
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint64_t bal = wave_ballot(((cover >> q) & 1u) != 0u);
      if (lead) m[q] = bal;
    }
  }
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Byte offset j * size of a staged record as a VGPR value the compiler
// cannot recompute: every LDS read of record j then shares this one address
// register (the compiler otherwise moves the uniform offset into a fresh
// VGPR before each group of reads).
__device__ __forceinline__ uint32_t lds_offset(uint32_t bytes) {
  uint32_t v;
  asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(bytes));
  return v;
}

template <typename T>
__device__ __forceinline__ const T& lds_at(const void* base, uint32_t off) {
  return *(const T*)((const char*)base + off);
}

// The exponent clamped to [0, 87] in one v_med3_f32 (fminf(fmaxf()) takes
// two): max(e, 0) is the reference's (shaders.js:565); the upper clamp only
// keeps the exp's argument in range, since e > eCut (<= 80) is never taken.
// e is finite (bounded conic and offsets), where med3 equals the pair.
__device__ __forceinline__ float clamp_e(float e) { return __builtin_amdgcn_fmed3f(e, 0.0f, 87.0f); }

// 1/x for x in [2^-9, 1]: hardware reciprocal estimate plus one Newton step
// with fused residual. Equal to IEEE 1.0f/x for all 75,497,473 floats of that
// range (exhaustive check on gfx950); callers guarantee the range.
__device__ __forceinline__ float rcp_exact_unit(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  const float e = fmaf(-x, r, 1.0f);
  return fmaf(r, e, r);
}

// The record's colours as f32: stored as f32 (recFmt 0: (r, g, b, opa) is
// the third float4) or as bf16 / fp16 pairs in the second float4's z, w
// (recFmt 1, 2), which decode exactly.
__device__ __forceinline__ float4 record_colour(int recFmt, const float4& r1, float w) {
  const uint32_t rg = __float_as_uint(r1.z), bl = __float_as_uint(r1.w);
  if (recFmt == 2)
    return make_float4(dec16<2>(rg), dec16<2>(rg >> 16), dec16<2>(bl), w);
  return make_float4(dec16<1>(rg), dec16<1>(rg >> 16), dec16<1>(bl), w);
}

// One tile of the forward pass, by the whole 256-thread block. The block's
// threads must all call it (it synchronises), and nothing else may use the
// block's LDS until every thread has returned.
template <class S, int kLoss>
__device__ __forceinline__ void raster_fwd_tile(const Cam& cam, const Cut& cut, const RasterBufs& b, int tile,
                                                float4* lds, PixelState* out = nullptr) {
  float4* const sA = lds;              // mx, my, A, B
  float4* const sB = lds + S::B;       // C, opa*comp, L, 0
  float4* const sC = lds + 2 * S::B;   // r, g, b, opa
  uint64_t (*const sMask)[S::kMasks] = reinterpret_cast<uint64_t (*)[S::kMasks]>(lds + 3 * S::B);
  double (*const sStat)[S::kWaves] = reinterpret_cast<double (*)[S::kWaves]>(&sMask[S::kChunks][0]);
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int mslot = S::mask_slot(wv);
  const PixelMap pm = pixel_of_thread<S>(cam, tile, tid);
  const int x = pm.x, y = pm.y;
  const bool inside = x < cam.W && y < cam.H;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  const uint2 rg = b.ranges[tile];
  const uint32_t s = __builtin_amdgcn_readfirstlane(rg.x), e = __builtin_amdgcn_readfirstlane(rg.y);
  if (out) *out = PixelState{make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, s};
  float T = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f;
  uint32_t end = s;
  bool done = !inside;
  for (uint32_t base = s; base < e; base += S::B) {
    if (__syncthreads_count(!done) == 0) break;
    // wave w stages chunks w, w + kWaves, ... (lane l: key 64 c + l)
    for (int c = wv; c < S::kChunks; c += S::kWaves) {
      const int j = 64 * c + lane;
      const uint32_t k = base + (uint32_t)j;
      uint32_t cover = 0;
      if (k < e) {
        const uint32_t v = b.ids[k];
        cover = v >> kMaskShift;
        if (cover != 0u) {           // splats no block of the tile can take stay unloaded
          const uint32_t id = v & kIdMask;
          if (b.recFmt == 0) {
            sA[j] = b.rec[4 * id + 0];
            sB[j] = b.rec[4 * id + 1];
            sC[j] = b.rec[4 * id + 2];
          } else {
            const float4 r1 = b.rec[3 * id + 1];
            sA[j] = b.rec[3 * id + 0];
            sB[j] = r1;
            sC[j] = record_colour(b.recFmt, r1, 0.0f);
          }
        }
      }
      store_masks<S>(cover, sMask[c]);
    }
    __syncthreads();
    const uint32_t cnt = min((uint32_t)S::B, e - base);
    // lanes still compositing; a lane is done once outside the image or
    // once T < 1e-4, so the vote is kept current with one compare per splat
    uint64_t run = wave_ballot(!done);
    if (run != 0ull) {
      uint32_t term = 0xffffffffu;  // batch index of this pixel's last splat
      for (int c = 0; c < S::kChunks; c++) {
        uint64_t m = uniform64(sMask[c][mslot]);
        while (m) {
          const int j = c * 64 + (int)__builtin_ctzll(m);
          m &= m - 1ull;
          // Straight-line for all lanes; only lanes that take the splat
          // (not done, inside the cutoff, alpha >= A_MIN) update their state.
          const uint32_t jo = lds_offset((uint32_t)j * 16u);
          const float4 a = lds_at<float4>(sA, jo);
          const float2 bb = lds_at<float2>(sB, jo);   // C, opa*comp
          asm volatile("" ::"v"(bb.y));                // one 8-byte read
          const float dx = px - a.x, dy = py - a.y;
          const float t2 = fmaf(bb.x * dy, dy, (a.z * dx) * dx);
          // clamped to [0, 87]: an exponent above 87 > eCut is never taken,
          // and the exp below needs its argument in [-87, 0]
          const float ee = clamp_e(fmaf(0.5f, t2, (a.w * dx) * dy));
          const bool live = !done && ee <= cut.eCut;
          if ((run & wave_ballot(ee <= cut.eCut)) != 0ull) {
            const float araw = bb.y * gs_expf_raster(-ee);
            const float alpha = fminf(0.99f, araw);
            if (live && alpha >= cut.aMin) {
              const float4 c4 = lds_at<float4>(sC, jo);
              const float w = T * alpha;
              C0 = fmaf(w, c4.x, C0);
              C1 = fmaf(w, c4.y, C1);
              C2 = fmaf(w, c4.z, C2);
              T = T * (1.0f - alpha);
            }
            // T only falls here, so a lane reaches T < 1e-4 on this splat
            // exactly when it is still running and now below; that happens
            // once per pixel, so its bookkeeping sits behind a scalar branch
            const bool fin = T < 1e-4f;
            const uint64_t left = run & ~wave_ballot(fin);
            if (left != run) {
              if (fin && !done) { done = true; term = (uint32_t)j; }
              run = left;
              if (run == 0ull) { m = 0ull; c = S::kChunks; }
            }
          }
        }
      }
      if (term != 0xffffffffu) end = base + term + 1u;
      else if (!done) end = base + cnt;
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
      if (out) {
        out->T = T;
        out->end = end;
      }
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
      if (out) out->g = g;
    }
  }
  if (kLoss == 1) {
    for (int o = 32; o > 0; o >>= 1) {
      st0 += __shfl_xor(st0, o);
      st1 += __shfl_xor(st1, o);
      st2 += __shfl_xor(st2, o);
    }
    if ((tid & 63) == 0) { sStat[0][wv] = st0; sStat[1][wv] = st1; sStat[2][wv] = st2; }
    __syncthreads();
    if (tid < 3) {
      double v = sStat[tid][0];
      for (int w = 1; w < S::kWaves; w++) v += sStat[tid][w];
      if (v != 0.0) atomicAdd(&b.stats[tid], v);
    }
  }
}

// ---- wave transpose reduction of a visit's 10 values (VALU lane exchanges) ----
// Six halving steps, one per lane-index bit, in this order: l ^ 15 (row_mirror,
// across bit 3), l ^ 7 (row_half_mirror, bit 2), l ^ 2 and l ^ 1 (quad_perm),
// l ^ 16 (v_permlane16_swap), l ^ 32 (v_permlane32_swap). In each a lane
// keeps one half of its values and adds its partner's copy of that half (an
// odd count is padded with a zero), so after the last step each slot's
// 64-lane total sits in exactly one lane and every other lane holds a
// padding zero. The slot each lane ends with is measured once per kernel by
// running the network on constants.
//
// The first two steps separate the lower and upper half by 4-lane bank
// (lanes 0-7 / 8-15 of a row, then 0-3, 8-11 / 4-7, 12-15), so they run as
// two DPP adds per output, each writing one half through the DPP bank mask:
// no selects. They run on each visit's terms as they are made (10 -> 5 -> 3
// values per lane); the last four steps run once per two visits on their
// 3 + 3 values (keep/send selects and one DPP add per output for the quad
// steps; the cross-row swaps on one value each). The asm blocks start with
// two wait states for the DPP read of a value the previous VALU instruction
// wrote, and end with two for a DPP read of their results that follows.
// Every step runs with all 64 lanes active (the visit loop branches only on
// wave-uniform votes, and every wave of a block is full): a DPP add whose
// source lane were inactive would leave its output unwritten.
template <int kCtrl>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, kCtrl, 0xf, 0xf, false);
}

template <typename T>
__device__ __forceinline__ uint32_t as_u(T v) { return __builtin_bit_cast(uint32_t, v); }
template <typename T>
__device__ __forceinline__ T from_u(uint32_t v) { return __builtin_bit_cast(T, v); }

template <typename T, int N, bool k32>
__device__ __forceinline__ void swap_halve(const T* v, T* out) {
  constexpr int H = (N + 1) / 2;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const uint32_t lo = as_u(v[i]);
    const uint32_t hi = i + H < N ? as_u(v[i + H]) : as_u((T)0);
    uint32_t a, bq;
    if (k32) {
      const auto r = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
      a = r[0];
      bq = r[1];
    } else {
      const auto r = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
      a = r[0];
      bq = r[1];
    }
    out[i] = from_u<T>(a) + from_u<T>(bq);
  }
}

template <typename T, int N, int kCtrl>
__device__ __forceinline__ void dpp_halve(const T* v, T* out, bool upper) {
  constexpr int H = (N + 1) / 2;
#pragma unroll
  for (int i = 0; i < H; i++) {
    const T lo = v[i];
    const T hi = i + H < N ? v[i + H] : (T)0;
    const T keep = upper ? hi : lo;
    const T send = upper ? lo : hi;
    out[i] = keep + from_u<T>(dpp<kCtrl>(as_u(send)));
  }
}

// Row step 1 (l ^ 15) on one visit's 10 values: lanes 0-7 of each row keep
// values 0-4, lanes 8-15 values 5-9.
#define GS_DPP_M(OP, O, LO, HI, CTRL, MLO, MHI) \
  OP " " O ", " LO ", " LO " " CTRL " row_mask:0xf bank_mask:" MLO "\n\t" \
  OP " " O ", " HI ", " HI " " CTRL " row_mask:0xf bank_mask:" MHI "\n\t"
#define GS_ROW_STEP1(OP)                                              \
  "s_nop 1\n\t"                                                      \
  GS_DPP_M(OP, "%0", "%5", "%10", "row_mirror", "0x3", "0xc")         \
  GS_DPP_M(OP, "%1", "%6", "%11", "row_mirror", "0x3", "0xc")         \
  GS_DPP_M(OP, "%2", "%7", "%12", "row_mirror", "0x3", "0xc")         \
  GS_DPP_M(OP, "%3", "%8", "%13", "row_mirror", "0x3", "0xc")         \
  GS_DPP_M(OP, "%4", "%9", "%14", "row_mirror", "0x3", "0xc")         \
  "s_nop 1"
// Row step 2 (l ^ 7) on 5 values: lanes 0-3 and 8-11 keep values 0-2,
// lanes 4-7 and 12-15 values 3, 4 and a zero.
#define GS_ROW_STEP2(OP)                                              \
  "s_nop 1\n\t"                                                      \
  GS_DPP_M(OP, "%0", "%3", "%6", "row_half_mirror", "0x5", "0xa")     \
  GS_DPP_M(OP, "%1", "%4", "%7", "row_half_mirror", "0x5", "0xa")     \
  "v_mov_b32 %2, 0\n\t"                                              \
  OP " %2, %5, %5 row_half_mirror row_mask:0xf bank_mask:0x5\n\t"    \
  "s_nop 1"

template <typename T>
__device__ __forceinline__ void visit_halve(const T* v, T* h3) {
  static_assert(sizeof(T) == 4, "32-bit terms");
  T h5[5];
  if constexpr (std::is_same<T, float>::value) {
    asm volatile(GS_ROW_STEP1("v_add_f32_dpp")
                 : "=&v"(h5[0]), "=&v"(h5[1]), "=&v"(h5[2]), "=&v"(h5[3]), "=&v"(h5[4])
                 : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]),
                   "v"(v[5]), "v"(v[6]), "v"(v[7]), "v"(v[8]), "v"(v[9]));
    asm volatile(GS_ROW_STEP2("v_add_f32_dpp")
                 : "=&v"(h3[0]), "=&v"(h3[1]), "=&v"(h3[2])
                 : "v"(h5[0]), "v"(h5[1]), "v"(h5[2]), "v"(h5[3]), "v"(h5[4]));
  } else {
    asm volatile(GS_ROW_STEP1("v_add_u32_dpp")
                 : "=&v"(h5[0]), "=&v"(h5[1]), "=&v"(h5[2]), "=&v"(h5[3]), "=&v"(h5[4])
                 : "v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]),
                   "v"(v[5]), "v"(v[6]), "v"(v[7]), "v"(v[8]), "v"(v[9]));
    asm volatile(GS_ROW_STEP2("v_add_u32_dpp")
                 : "=&v"(h3[0]), "=&v"(h3[1]), "=&v"(h3[2])
                 : "v"(h5[0]), "v"(h5[1]), "v"(h5[2]), "v"(h5[3]), "v"(h5[4]));
  }
}
#undef GS_ROW_STEP2
#undef GS_ROW_STEP1
#undef GS_DPP_M

// The last four steps on two visits' row-halved values (the waiting one's
// in p, the new one's in h): 20 wave totals, slots 0-9 the waiting visit's.
template <typename T>
__device__ __forceinline__ T pair_reduce(const T* p, const T* h) {
  const uint32_t lane = __lane_id();
  const T v[6] = {p[0], p[1], p[2], h[0], h[1], h[2]};
  T c[3], d[2], e[1], g[1];
  dpp_halve<T, 6, 0x4E>(v, c, (lane & 2) != 0);
  dpp_halve<T, 3, 0xB1>(c, d, (lane & 1) != 0);
  swap_halve<T, 2, false>(d, e);
  swap_halve<T, 1, true>(e, g);
  return g[0];
}

// The last four steps on one visit's row-halved values (a visit left
// without a partner at the end of the tile).
template <typename T>
__device__ __forceinline__ T single_reduce(const T* h) {
  const uint32_t lane = __lane_id();
  T d[2], e[1], f[1], g[1];
  dpp_halve<T, 3, 0x4E>(h, d, (lane & 2) != 0);
  dpp_halve<T, 2, 0xB1>(d, e, (lane & 1) != 0);
  swap_halve<T, 1, false>(e, f);
  swap_halve<T, 1, true>(f, g);
  return g[0];
}

// slot whose 64-lane total this lane holds after pair_reduce (0-9: the
// waiting visit's terms, 10-19: the new visit's; -1: none)
__device__ __forceinline__ int pair_slot() {
  uint32_t va[10], vb[10], ha[3], hb[3];
#pragma unroll
  for (int i = 0; i < 10; i++) {
    va[i] = (uint32_t)(i + 1);
    vb[i] = (uint32_t)(i + 11);
  }
  visit_halve<uint32_t>(va, ha);
  visit_halve<uint32_t>(vb, hb);
  const int s = (int)(pair_reduce<uint32_t>(ha, hb) / 64u) - 1;
  return s >= 0 && s < 20 ? s : -1;
}

// slot whose 64-lane total this lane holds after single_reduce (-1: none)
__device__ __forceinline__ int single_slot() {
  uint32_t v[10], h[3];
#pragma unroll
  for (int i = 0; i < 10; i++) v[i] = (uint32_t)(i + 1);
  visit_halve<uint32_t>(v, h);
  const int s = (int)(single_reduce<uint32_t>(h) / 64u) - 1;
  return s >= 0 && s < 10 ? s : -1;
}

// Adds one wave total to slot q of Gaussian g's accumulator row (the lanes
// holding a total; zeros skipped).
template <int kAccum, typename T>
__device__ __forceinline__ void flush_sum(void* accum, uint32_t g, int q, T r, bool flusher) {
  if constexpr (kAccum == 0) {
    if (flusher && r != 0u) atomicAdd((uint32_t*)accum + (int64_t)g * 12 + q, r);
  } else {
    if (flusher && r != 0.0f) unsafeAtomicAdd((float*)accum + (int64_t)g * 12 + q, r);
  }
}

// kClampColour false: dL/dC is bounded (the Charbonnier loss: |dL/dC| <=
// gain * |err| / sqrt(err^2 + d^2) <= gain), so the colour terms
// 16384 dL/dC * alpha * T (alpha * T <= 1) stay far inside the +-1e9 clamp
// and convert without it.
// One tile of the backward pass, by the whole block (as raster_fwd_tile).
template <class S, int kAccum, bool kClampColour>
__device__ __forceinline__ void raster_bwd_tile(const Cam& cam, const Cut& cut, const RasterBufs& b, void* accum,
                                                int tile, float4* lds, const PixelState* in = nullptr) {
  // four 16-byte records per staged splat, all at index j: one LDS address
  float4* const sA = lds;              // mx, my, A, B
  float4* const sB = lds + S::B;       // C, opa*comp, L, 0
  float4* const sC = lds + 2 * S::B;   // r, g, b, opa
  float4* const sD = lds + 3 * S::B;   // -0.5 cnorm, 1 - opa, id (bits), 0
  uint64_t (*const sMask)[S::kMasks] = reinterpret_cast<uint64_t (*)[S::kMasks]>(lds + 4 * S::B);
  uint32_t& sEnd = *reinterpret_cast<uint32_t*>(&sMask[S::kChunks][0]);
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int mslot = S::mask_slot(wv);
  const PixelMap pm = pixel_of_thread<S>(cam, tile, tid);
  const int x = pm.x, y = pm.y;
  const bool inside = x < cam.W && y < cam.H;
  const float px = (float)x + 0.5f, py = (float)y + 0.5f;
  // the tile's range and the block's largest end are the same in every
  // lane; read through readfirstlane so the compiler keeps the key indices
  // and the visit masks derived from them in scalar registers (as LDS and
  // global loads they would count as per-lane values, and every mask
  // operation of the visit loop would run on the VALU)
  const uint2 rg = b.ranges[tile];
  const uint32_t s = __builtin_amdgcn_readfirstlane(rg.x);
  const int64_t pi = (int64_t)y * cam.W + x;
  const int mySlot = single_slot();
  const bool flusher = mySlot >= 0;
  const int mySlot20 = pair_slot();
  const bool flusher20 = mySlot20 >= 0;
  // A visit's terms, halved twice inside each row as they are made (3
  // values per lane), wait for the next visit's, and the two finish in one
  // pair_reduce (slots 0-9 the waiting visit's, 10-19 the new one's); the
  // sums are added to the accumulators in a different order than visited,
  // which changes nothing (integer sums; f32 atomics have no fixed order
  // anyway).
  using AccT = typename std::conditional<kAccum == 0, uint32_t, float>::type;
  AccT pend[3];
  uint32_t pendGid = 0;
  bool hasPend = false;
  float4 g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float Ta = 0.0f;
  uint32_t myEnd = s;
  if (in) {   // the forward of this launch left them in registers
    g = in->g;
    if (g.w != 0.0f) {
      Ta = in->T;
      myEnd = in->end;
    }
  } else if (inside) {
    g = b.gC[pi];
    if (g.w != 0.0f) {
      Ta = b.T[pi];
      myEnd = b.end[pi];
    }
  }
  const bool active = inside && g.w != 0.0f;
  const uint64_t activeMask = wave_ballot(active);
  if (tid == 0) sEnd = s;
  __syncthreads();
  if (myEnd > s) atomicMax(&sEnd, myEnd);
  __syncthreads();
  const uint32_t endMax = __builtin_amdgcn_readfirstlane(sEnd);
  // this wave's largest end: keys at or past it are past every lane's end,
  // so the wave leaves them out of its visit set (they would add nothing)
  uint32_t we = myEnd;
  for (int o = 32; o > 0; o >>= 1) we = max(we, (uint32_t)__shfl_xor((int)we, o));
  const uint32_t waveEnd = __builtin_amdgcn_readfirstlane(we);
  const float gC[3] = {g.x, g.y, g.z};
  // GS_ACCUM_FIXED: the x16384 fixed-point scale rides on dL/dC from the
  // start, so on every product with it (dot(gC, c), dot(gC, S), galpha) and
  // on every term, which carries galpha or gC. Scaling by a power of two is
  // exact, so each scaled value equals the value times 16384 bit for bit;
  // the only exceptions are values whose unscaled form is subnormal, and
  // terms that small truncate to zero either way.
  constexpr float kS = kAccum == 0 ? kFixed : 1.0f;
  const float gCs[3] = {gC[0] * kS, gC[1] * kS, gC[2] * kS};
  // 16384 dot(gC, S) of the colour-behind sum S, which enters only through it
  float GS = fmaf(gCs[2], cam.bg[2], fmaf(gCs[1], cam.bg[1], gCs[0] * cam.bg[0])) * Ta;
  for (uint32_t hi = endMax; hi > s; ) {
    const uint32_t lo = hi - s > (uint32_t)S::B ? hi - S::B : s;
    const int n = (int)(hi - lo);
    __syncthreads();
    for (int c = wv; c < S::kChunks; c += S::kWaves) {
      const int j = 64 * c + lane;
      uint32_t cover = 0;
      if (j < n) {
        const uint32_t v = b.ids[lo + (uint32_t)j];
        cover = v >> kMaskShift;
        if (cover != 0u) {
          const uint32_t id = v & kIdMask;
          if (b.recFmt == 0) {
            sA[j] = b.rec[4 * id + 0];
            sB[j] = b.rec[4 * id + 1];
            sC[j] = b.rec[4 * id + 2];
            const float4 d = b.rec[4 * id + 3];
            sD[j] = make_float4(d.y, d.z, __uint_as_float(id), 0.0f);
          } else {
            const float4 r1 = b.rec[3 * id + 1];
            const float4 d = b.rec[3 * id + 2];   // comp, -0.5 cnorm, 1 - opa, opa
            sA[j] = b.rec[3 * id + 0];
            sB[j] = r1;
            sC[j] = record_colour(b.recFmt, r1, d.w);
            sD[j] = make_float4(d.y, d.z, __uint_as_float(id), 0.0f);
          }
        }
      }
      store_masks<S>(cover, sMask[c]);
    }
    __syncthreads();
    for (int c = S::kChunks - 1; c >= 0; c--) {
      uint64_t m = uniform64(sMask[c][mslot]);
      const uint32_t g0 = lo + 64u * (uint32_t)c;   // key of bit 0
      if (waveEnd <= g0) m = 0ull;
      else if (waveEnd - g0 < 64u) m &= (1ull << (waveEnd - g0)) - 1ull;
      while (m) {
        const int bit = 63 - (int)__builtin_clzll(m);
        m &= ~(1ull << bit);
        const int j = c * 64 + bit;
        const uint32_t k = lo + (uint32_t)j;
        // Straight-line for all 64 lanes; lanes that do not take the splat
        // (outside the image, past their `end`, outside the cutoff, alpha
        // below A_MIN) keep their state and contribute zero terms. The wave
        // leaves as soon as no lane is inside the cutoff (before the exp) or
        // no lane takes the splat.
        const uint32_t jo = lds_offset((uint32_t)j * 16u);
        const float4 a = lds_at<float4>(sA, jo);
        const float2 bb = lds_at<float2>(sB, jo);   // C, opa*comp
        // keep both in this one 8-byte read (otherwise the compiler sinks
        // the second word into the next block with its own address move)
        asm volatile("" ::"v"(bb.y));
        const float dx = px - a.x, dy = py - a.y;
        const float cA = a.z, cB = a.w, cC = bb.x;
        const float t2 = fmaf(cC * dy, dy, (cA * dx) * dx);
        // clamped to [0, 87] as in the forward pass
        const float e = clamp_e(fmaf(0.5f, t2, (cB * dx) * dy));
        // votes of single compares, combined in scalar registers (a vote
        // of a combined condition costs two VALU instructions)
        const bool live = active && k < myEnd && e <= cut.eCut;
        const uint64_t liveMask = activeMask & wave_ballot(k < myEnd) & wave_ballot(e <= cut.eCut);
        if (liveMask == 0ull) continue;
        const float G = gs_expf_raster(-e);
        const float araw = bb.y * G;
        const float alpha = fminf(0.99f, araw);
        const bool contrib = live && alpha >= cut.aMin;
        if ((liveMask & wave_ballot(alpha >= cut.aMin)) == 0ull) continue;
        float v[10];
        uint32_t gid;
        {
          const float4 c4 = lds_at<float4>(sC, jo);
          const float4 d2 = lds_at<float4>(sD, jo);   // -0.5 cnorm, 1 - opa, id
          asm volatile("" ::"v"(d2.z));
          gid = __float_as_uint(d2.z);
          // Tb = Ta / (1 - alpha), S / (1 - alpha) through one reciprocal.
          // 1 - alpha lies in [0.0099, 1] (alpha in [0, 0.99]); on [2^-9, 1]
          // v_rcp_f32 refined by one FMA Newton step equals the correctly
          // rounded 1/x for every float (checked exhaustively on gfx950).
          const float rom = rcp_exact_unit(1.0f - alpha);
          const float Tb = Ta * rom;
          const float gcol = fmaf(gCs[2], c4.z, fmaf(gCs[1], c4.y, gCs[0] * c4.x));   // scaled
          // Every term is a product with galpha or aT; zeroing those two for
          // the lanes that do not take the splat zeroes all ten terms (all
          // other factors are finite), and leaves the taking lanes' bits as is.
          const float galphaS = (contrib && !(araw > 0.99f)) ? fmaf(gcol, Tb, -(GS * rom)) : 0.0f;
          const float aT = contrib ? alpha * Tb : 0.0f;
          const float ga = galphaS * araw;
          const float gq = ga * d2.x;
          const float gqdx = gq * dx;
          v[0] = ga * (cA * dx + cB * dy);
          v[1] = ga * (cB * dx + cC * dy);
          v[2] = gqdx * dx;
          v[3] = 2.0f * (gqdx * dy);        // the middle conic term is 2 gq dx dy
          v[4] = (gq * dy) * dy;
          v[5] = galphaS * c4.w * G;
          v[6] = ga * d2.y;
          v[7] = gCs[0] * aT;
          v[8] = gCs[1] * aT;
          v[9] = gCs[2] * aT;
          GS = fmaf(gcol, aT, GS);
          Ta = contrib ? Tb : Ta;
        }
        AccT cur[10];
        if constexpr (kAccum == 0) {
          // i32(clamp(v * 16384, -1e9, 1e9)) per pixel; v already carries the scale
#pragma unroll
          for (int q = 0; q < 10; q++)
            cur[q] = (q >= 7 && !kClampColour) ? (uint32_t)(int32_t)v[q]
                                                : (uint32_t)(int32_t)fminf(fmaxf(v[q], -1.0e9f), 1.0e9f);
        } else {
#pragma unroll
          for (int q = 0; q < 10; q++) cur[q] = v[q];
        }
        AccT h[3];
        visit_halve<AccT>(cur, h);
        if (hasPend) {
          const AccT r = pair_reduce<AccT>(pend, h);
          const uint32_t g2 = mySlot20 < 10 ? pendGid : gid;
          flush_sum<kAccum>(accum, g2, mySlot20 % 10, r, flusher20);
        }
        // copied on both paths (after a pair it is never read), so the
        // waiting values need no merge of two register sets at the join
#pragma unroll
        for (int q = 0; q < 3; q++) pend[q] = h[q];
        pendGid = gid;
        hasPend = !hasPend;
      }
    }
    hi = lo;
  }
  if (hasPend) flush_sum<kAccum>(accum, pendGid, mySlot, single_reduce<AccT>(pend), flusher);
}

template <class S, int kLoss>
__global__ __launch_bounds__(S::kThreads) void k_raster_fwd(Cam cam, Cut cut, RasterBufs b) {
  __shared__ float4 lds[RasterLds<S>::kFwd];
  raster_fwd_tile<S, kLoss>(cam, cut, b, tile_of_block(b), lds);
}

template <class S, int kAccum, bool kClampColour>
__global__ __launch_bounds__(S::kThreads) void k_raster_bwd(Cam cam, Cut cut, RasterBufs b, void* accum) {
  __shared__ float4 lds[RasterLds<S>::kBwd];
  raster_bwd_tile<S, kAccum, kClampColour>(cam, cut, b, accum, tile_of_block(b), lds);
}

// A training step's forward with the Charbonnier loss and its backward in
// one launch: the loss gradient of a pixel needs only that pixel, so each
// block runs the backward of its tile right after the forward (each thread
// keeps its pixel's T, end and dL/dC in registers; they are stored too,
// for the debug readbacks).
template <class S, int kAccum, bool kClampColour>
__global__ __launch_bounds__(S::kThreads) void k_raster_fwd_bwd(Cam cam, Cut cut, RasterBufs b, void* accum) {
  __shared__ float4 lds[RasterLds<S>::kBoth];   // the two passes' layouts share it
  const int tile = tile_of_block(b);
  PixelState ps;
  raster_fwd_tile<S, 1>(cam, cut, b, tile, lds, &ps);
  __syncthreads();   // the forward's last LDS reads are done
  raster_bwd_tile<S, kAccum, kClampColour>(cam, cut, b, accum, tile, lds, &ps);
}

// ---- several views in one persistent launch ----
// Work item w in [0, off[V]) is tile tileOrder_v[w - off[v]] of the view v
// with off[v] <= w < off[v + 1]: each view's tiles in its own heaviest-first
// order, the views one after another. Each block takes the next item from
// the queue counter (zeroed before the launch) until the items run out, so
// every block leaves once the counter passes off[V].
struct ViewItem {
  int v;
  int tile;
};

__device__ __forceinline__ bool next_item(const ViewBlock* __restrict__ vb, uint32_t* queue, ViewItem& it) {
  __shared__ uint32_t sWork;
  __syncthreads();   // the previous item's LDS reads are done
  if (threadIdx.x == 0) sWork = atomicAdd(queue, 1u);
  __syncthreads();
  const uint32_t w = __builtin_amdgcn_readfirstlane(sWork);
  const int V = vb->count;
  if (w >= vb->off[V]) return false;
  int v = 0;
  while (w >= vb->off[v + 1]) v++;
  it.v = v;
  it.tile = (int)gs_global(vb->bufs[v].tileOrder)[w - vb->off[v]];
  return true;
}

// Forward with the Charbonnier loss and backward of every view's tiles in
// one persistent launch: each work item's tile runs both passes back to
// back, as k_raster_fwd_bwd does for one view.
template <class S, int kAccum, bool kClampColour>
__global__ __launch_bounds__(S::kThreads) __attribute__((amdgpu_waves_per_eu(S::kBwdWavesPerEU))) void
k_raster_fwd_bwd_views(const ViewBlock* __restrict__ vb, Cut cut, uint32_t* queue) {
  __shared__ float4 lds[RasterLds<S>::kBoth];
  ViewItem it;
  while (next_item(vb, queue, it)) {
    // read in place (the block is read-only here): copies held every camera
    // and buffer field in scalar registers across the whole tile and spilled
    // some of them. The buffers' pointers stay generic (flat accesses): a
    // copy with each loaded as a global pointer (global_bufs) took 0.2-0.4%
    // longer.
    const Cam& cam = vb->cams[it.v];
    const RasterBufs& b = vb->bufs[it.v];
    PixelState ps;
    raster_fwd_tile<S, 1>(cam, cut, b, it.tile, lds, &ps);
    __syncthreads();   // the forward's last LDS reads are done
    raster_bwd_tile<S, kAccum, kClampColour>(cam, cut, b, gs_global(vb->accum[it.v]), it.tile, lds, &ps);
  }
}

// one view of a batched step, one block per tile (camera and buffers from
// device memory), forward with the loss and backward in one launch
template <class S, int kAccum, bool kClampColour>
__global__ __launch_bounds__(S::kThreads) void k_raster_fwd_bwd_view(const ViewBlock* __restrict__ vb, int v,
                                                                     Cut cut) {
  const Cam cam = vb->cams[v];
  const RasterBufs b = global_bufs(vb->bufs[v]);
  __shared__ float4 lds[RasterLds<S>::kBoth];
  const int tile = tile_of_block(b);
  PixelState ps;
  raster_fwd_tile<S, 1>(cam, cut, b, tile, lds, &ps);
  __syncthreads();
  raster_bwd_tile<S, kAccum, kClampColour>(cam, cut, b, gs_global(vb->accum[v]), tile, lds, &ps);
}

// f(Shape<tile, batch>{}) for a supported shape (checked by the callers)
template <int kT, typename F>
static void with_batch(int batch, F&& f) {
  if (batch == 64) f(Shape<kT, 64>{});
  else if (batch == 128) f(Shape<kT, 128>{});
  else f(Shape<kT, 256>{});
}

template <typename F>
static void with_shape(const RasterShape& sh, F&& f) {
  if (sh.tile == 8) with_batch<8>(sh.batch, f);
  else if (sh.tile == 32) with_batch<32>(sh.batch, f);
  else with_batch<16>(sh.batch, f);
}

bool raster_shape_ok(int tile, int batch) {
  return (tile == 8 || tile == 16 || tile == 32) && (batch == 64 || batch == 128 || batch == 256);
}

void launch_raster_fwd_bwd_view(hipStream_t st, const RasterShape& sh, const ViewBlock* vb, int v, int numTiles,
                                const Cut& cut, int accumMode, bool boundedColourGrad) {
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (accumMode != 0) k_raster_fwd_bwd_view<S, 1, true><<<numTiles, S::kThreads, 0, st>>>(vb, v, cut);
    else if (boundedColourGrad) k_raster_fwd_bwd_view<S, 0, false><<<numTiles, S::kThreads, 0, st>>>(vb, v, cut);
    else k_raster_fwd_bwd_view<S, 0, true><<<numTiles, S::kThreads, 0, st>>>(vb, v, cut);
  });
}

void launch_raster_fwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                       int loss) {
  const dim3 grid((unsigned)cam.numTiles);
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (loss == 1) k_raster_fwd<S, 1><<<grid, S::kThreads, 0, st>>>(cam, cut, b);
    else if (loss == 2) k_raster_fwd<S, 2><<<grid, S::kThreads, 0, st>>>(cam, cut, b);
    else k_raster_fwd<S, 0><<<grid, S::kThreads, 0, st>>>(cam, cut, b);
  });
}

void launch_raster_bwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                       int accumMode, void* accum, bool boundedColourGrad) {
  const dim3 grid((unsigned)cam.numTiles);
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (accumMode != 0) k_raster_bwd<S, 1, true><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
    else if (boundedColourGrad) k_raster_bwd<S, 0, false><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
    else k_raster_bwd<S, 0, true><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
  });
}

void launch_raster_fwd_bwd(hipStream_t st, const RasterShape& sh, const Cam& cam, const Cut& cut, const RasterBufs& b,
                           int accumMode, void* accum, bool boundedColourGrad) {
  const dim3 grid((unsigned)cam.numTiles);
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (accumMode != 0) k_raster_fwd_bwd<S, 1, true><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
    else if (boundedColourGrad) k_raster_fwd_bwd<S, 0, false><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
    else k_raster_fwd_bwd<S, 0, true><<<grid, S::kThreads, 0, st>>>(cam, cut, b, accum);
  });
}

void launch_raster_fwd_bwd_views(hipStream_t st, const RasterShape& sh, const ViewBlock* vb, const Cut& cut,
                                 int accumMode, bool boundedColourGrad, uint32_t* queue, int blocks) {
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (accumMode != 0) k_raster_fwd_bwd_views<S, 1, true><<<blocks, S::kThreads, 0, st>>>(vb, cut, queue);
    else if (boundedColourGrad) k_raster_fwd_bwd_views<S, 0, false><<<blocks, S::kThreads, 0, st>>>(vb, cut, queue);
    else k_raster_fwd_bwd_views<S, 0, true><<<blocks, S::kThreads, 0, st>>>(vb, cut, queue);
  });
}

// blocks of the persistent launches of one shape: as many as fit on the
// device at once
int raster_views_blocks(int device, const RasterShape& sh) {
  int cus = 0, per = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0) cus = 256;
  with_shape(sh, [&](auto shape) {
    using S = decltype(shape);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_raster_fwd_bwd_views<S, 0, false>, S::kThreads, 0) !=
        hipSuccess)
      per = 1;
  });
  if (per < 1) per = 1;
  return cus * per;
}

}  // namespace gs
