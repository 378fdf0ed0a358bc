// gs_sort.hip — global binning: one key per (Gaussian, covered tile),
// ordered by (tile, depth, Gaussian id), the total order the reference's
// per-tile bitonic sort produces (shaders.js:418-500, tie-break :454).
//
// Two strategies, identical output:
//  * depth-then-tile (default): stable LSD radix sort of the n Gaussians by
//    depth bits (emitted in id order, so ties keep id order), then keys are
//    emitted in that order and stably sorted by tile id only
//    (ceil(tileBits/8) passes). Sorting Gaussians instead of keys moves the
//    32 depth bits off the long key array.
//  * global64: keys (tile << 32 | depth bits) emitted in id order and sorted
//    by a stable LSD radix sort over all 32 + tileBits bits.
//
// Every radix pass (8-bit digit) is reduce-then-scan, with no inter-block
// waiting, over a fixed grid of G blocks, each owning a contiguous chunk
// processed in rounds of 2048 keys (8 per thread, wave-striped so every
// load is coalesced):
//   hist    per-wave digit counts in LDS; each wave groups its 64 keys by
//           digit with 8 ballots, and one lane per group adds the group size;
//           the block's 256 counts go to a digit-major table hist[d][block]
//   scan    one block per digit: exclusive scan of the digit's row (a
//           contiguous row) and the digit total
//   scatter digit bases from an in-block scan of the 256 totals;
//           wave-local stable ranks from the same ballot groups, one
//           block-wide digit prefix per round, keys staged in LDS in
//           (digit, position) order, then written out in digit runs by
//           consecutive threads
// Key counts that only exist on the device (the emitted key total) are read
// from device memory by every kernel; the grids are fixed. No pass uses
// global atomics.
#include "gs_kernels.h"

namespace gs {

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}

// mask of active lanes whose 8-bit digit equals this lane's
__device__ __forceinline__ uint64_t match_digit8(uint32_t d, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = wave_ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

constexpr int kItems = 8;
constexpr int kRound = kBlock * kItems;   // keys per block round

// block b's contiguous chunk [lo, hi) of N elements split over G blocks,
// rounded to `align` elements
__device__ __forceinline__ void chunk_range(uint32_t N, int G, uint32_t align, uint32_t& lo, uint32_t& hi) {
  const uint32_t per = (uint32_t)(((uint64_t)N + G - 1) / G);
  const uint32_t chunk = (per + align - 1u) / align * align;
  const uint64_t l = (uint64_t)blockIdx.x * chunk;
  lo = (uint32_t)(l < N ? l : N);
  const uint64_t h = (uint64_t)lo + chunk;
  hi = (uint32_t)(h < N ? h : N);
}

__global__ void k_fill_u32(uint32_t* p, int64_t count, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = v;
}

template <typename KeyT>
__device__ __forceinline__ uint32_t digit_of(KeyT k, int shift) {
  return (uint32_t)(k >> shift) & 0xffu;
}

template <typename KeyT>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const KeyT* __restrict__ keys,
                                                       const uint32_t* __restrict__ dN, int shift,
                                                       int G, uint32_t* __restrict__ hist) {
  __shared__ uint32_t wc[4][256];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  wc[0][tid] = wc[1][tid] = wc[2][tid] = wc[3][tid] = 0;
  __syncthreads();
  uint32_t lo, hi;
  chunk_range(*dN, G, kRound, lo, hi);
  for (uint32_t r = lo; r < hi; r += kRound) {
    const uint32_t base = r + (uint32_t)w * 64u * kItems;
    uint32_t d[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const uint32_t idx = base + (uint32_t)i * 64u + lane;
      d[i] = idx < hi ? digit_of(keys[idx], shift) : 0x100u;
    }
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const bool ok = d[i] < 0x100u;
      const uint64_t act = wave_ballot(ok);
      if (act == 0ull) break;
      const uint64_t m = match_digit8(d[i] & 0xffu, act);
      // one lane per digit group: distinct addresses within an instruction
      if (ok && (m & lanemask_lt()) == 0ull) atomicAdd(&wc[w][d[i]], (uint32_t)__popcll(m));
    }
  }
  __syncthreads();
  hist[(int64_t)tid * G + blockIdx.x] = wc[0][tid] + wc[1][tid] + wc[2][tid] + wc[3][tid];
}

// one block per digit: exclusive scan of hist[d][0..G) in place, and the
// digit's total into totals[d] (the scatter adds the digit bases)
__global__ __launch_bounds__(kBlock) void k_radix_scan(uint32_t* __restrict__ hist, int G,
                                                       uint32_t* __restrict__ totals) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int d = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t carry = 0u;
  uint32_t* row = hist + (int64_t)d * G;
  for (int b0 = 0; b0 < G; b0 += kBlock) {
    const int b = b0 + tid;
    const uint32_t v = b < G ? row[b] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    __syncthreads();
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wpre = 0;
    for (int k = 0; k < w; k++) wpre += wsum[k];
    if (b < G) row[b] = carry + wpre + x - v;
    carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
  }
  if (tid == 0) totals[d] = carry;
}

template <typename KeyT, bool kWriteKeys>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const KeyT* __restrict__ keysIn,
                                                          const uint32_t* __restrict__ valsIn,
                                                          const uint32_t* __restrict__ dN,
                                                          int shift, int G,
                                                          const uint32_t* __restrict__ scan,
                                                          KeyT* __restrict__ keysOut,
                                                          uint32_t* __restrict__ valsOut,
                                                          const uint32_t* __restrict__ totals) {
  __shared__ uint32_t wc[4][256];
  __shared__ uint32_t wpre[4][256];
  __shared__ uint32_t gbase[256];
  __shared__ uint32_t lstart[256];
  __shared__ uint32_t wsum[4];
  __shared__ KeyT sk[kRound];
  __shared__ uint32_t sv[kRound];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  uint32_t lo, hi;
  chunk_range(*dN, G, kRound, lo, hi);
  {
    // digit bases: exclusive scan of the digit totals
    const uint32_t t = totals[tid];
    uint32_t x = t;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wp = 0;
    for (int k = 0; k < w; k++) wp += wsum[k];
    gbase[tid] = wp + x - t + scan[(int64_t)tid * G + blockIdx.x];
  }
  wc[0][tid] = wc[1][tid] = wc[2][tid] = wc[3][tid] = 0;
  __syncthreads();
  for (uint32_t r = lo; r < hi; r += kRound) {
    // 1. load, then rank within the wave (stable: item-major, then lane)
    const uint32_t base = r + (uint32_t)w * 64u * kItems;
    KeyT key[kItems];
    uint32_t val[kItems], rank[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const uint32_t idx = base + (uint32_t)i * 64u + lane;
      const bool ok = idx < hi;
      key[i] = ok ? keysIn[idx] : (KeyT)0;
      val[i] = ok ? (valsIn ? valsIn[idx] : idx) : 0u;
      rank[i] = ok ? 0u : 0xffffffffu;
    }
    // ranks: all groups' counter updates are issued back to back (the LDS
    // applies one wave's operations in order), then each lane fetches its
    // group leader's pre-add count
    uint64_t mm[kItems];
    uint32_t old[kItems];
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const bool ok = rank[i] == 0u;
      mm[i] = match_digit8(digit_of(key[i], shift), wave_ballot(ok)) ;
      old[i] = 0u;
      if (ok && (mm[i] & lanemask_lt()) == 0ull)
        old[i] = atomicAdd(&wc[w][digit_of(key[i], shift)], (uint32_t)__popcll(mm[i]));
    }
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const bool ok = rank[i] == 0u;
      const int leader = mm[i] ? (int)__builtin_ctzll(mm[i]) : 0;
      const uint32_t b = (uint32_t)__shfl((int)old[i], leader);
      if (ok) rank[i] = b + (uint32_t)__popcll(mm[i] & lanemask_lt());
    }
    __syncthreads();
    // 2. per digit: wave prefixes, round total, local start (block scan)
    const uint32_t c0 = wc[0][tid], c1 = wc[1][tid], c2 = wc[2][tid], c3 = wc[3][tid];
    const uint32_t tot = c0 + c1 + c2 + c3;
    wpre[0][tid] = 0u;
    wpre[1][tid] = c0;
    wpre[2][tid] = c0 + c1;
    wpre[3][tid] = c0 + c1 + c2;
    wc[0][tid] = wc[1][tid] = wc[2][tid] = wc[3][tid] = 0;
    uint32_t x = tot;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t wp = 0;
    for (int k = 0; k < w; k++) wp += wsum[k];
    const uint32_t myStart = wp + x - tot;
    lstart[tid] = myStart;
    __syncthreads();
    // 3. stage in local (digit, rank) order
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      if (rank[i] != 0xffffffffu) {
        const uint32_t d = digit_of(key[i], shift);
        const uint32_t p = lstart[d] + wpre[w][d] + rank[i];
        sk[p] = key[i];
        sv[p] = val[i];
      }
    }
    __syncthreads();
    // 4. write out: consecutive threads write consecutive slots of a digit run
    const uint32_t cnt = min((uint32_t)kRound, hi - r);
    for (uint32_t j = tid; j < cnt; j += kBlock) {
      const KeyT k = sk[j];
      const uint32_t d = digit_of(k, shift);
      const uint32_t g = gbase[d] + (j - lstart[d]);
      if (kWriteKeys) keysOut[g] = k;
      valsOut[g] = sv[j];
    }
    __syncthreads();
    gbase[tid] += tot;
  }
}

// ---- emission offsets: exclusive scan of cnt[] in sorted order ----

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t& total, uint32_t* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0;
  for (int k = 0; k < w; k++) pre += wsum[k];
  total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
  return pre + x - v;
}

// One block per 256 Gaussians in (depth) order `order` (nullptr: id
// order): gathers each one's key count and tile rectangle once, stores them
// in sort order for the emission pass, and writes the block's key total and
// visible count.
__global__ __launch_bounds__(kBlock) void k_count_partials(const uint32_t* __restrict__ order,
                                                           const uint32_t* __restrict__ cnt,
                                                           const uint2* __restrict__ rect,
                                                           const uint32_t* __restrict__ depthKey,
                                                           const uint32_t* __restrict__ dN,
                                                           uint32_t* __restrict__ cntS,
                                                           uint2* __restrict__ rectS,
                                                           uint32_t* __restrict__ dkS,
                                                           uint64_t* __restrict__ partials,
                                                           uint32_t* __restrict__ vpartials) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t wvis[4];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const uint32_t N = *dN;
  const uint32_t r = blockIdx.x * kBlock + tid;
  uint32_t c = 0;
  if (r < N) {
    const uint32_t id = order ? order[r] : r;
    c = cnt[id];
    cntS[r] = c;
    rectS[r] = c ? rect[id] : make_uint2(0, 0);
    if (dkS) dkS[r] = c ? depthKey[id] : 0u;
  }
  uint32_t acc = c, vis = c != 0u;
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o);
    vis += __shfl_xor(vis, o);
  }
  if (lane == 0) { wsum[w] = acc; wvis[w] = vis; }
  __syncthreads();
  if (tid == 0) {
    partials[blockIdx.x] = (uint64_t)wsum[0] + wsum[1] + wsum[2] + wsum[3];
    vpartials[blockIdx.x] = wvis[0] + wvis[1] + wvis[2] + wvis[3];
  }
}

// single block: exclusive scan of the G partials; numKeys = min(total, cap)
__global__ __launch_bounds__(1024) void k_count_top(uint64_t* __restrict__ partials,
                                                    const uint32_t* __restrict__ vpartials, int G,
                                                    uint64_t cap, uint32_t* __restrict__ numKeys,
                                                    uint64_t* __restrict__ totalKeys,
                                                    uint32_t* __restrict__ numVisible,
                                                    double* __restrict__ stats) {
  __shared__ uint64_t wsum[16];
  __shared__ uint32_t wvis[16];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int per = (G + 1023) / 1024;
  uint64_t mine = 0;
  uint32_t vis = 0;
  for (int k = 0; k < per; k++) {
    const int b = tid * per + k;
    if (b < G) { mine += partials[b]; vis += vpartials[b]; }
  }
  for (int o = 32; o > 0; o >>= 1) vis += __shfl_xor(vis, o);
  if (lane == 0) wvis[w] = vis;
  uint64_t x = mine;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint64_t pre = 0;
  for (int k = 0; k < w; k++) pre += wsum[k];
  uint64_t run = pre + x - mine;
  for (int k = 0; k < per; k++) {
    const int b = tid * per + k;
    if (b < G) {
      const uint64_t v = partials[b];
      partials[b] = run;
      run += v;
    }
  }
  if (tid == 0) {
    uint32_t v = 0;
    for (int k = 0; k < 16; k++) v += wvis[k];
    *numVisible = v;
  }
  if (tid == 1023) {
    const uint64_t total = pre + x;
    *totalKeys = total;
    *numKeys = (uint32_t)(total < cap ? total : cap);
    if (total > cap) stats[3] += (double)(total - cap);
  }
}

// Emits one key per covered tile, cooperatively: for every round of 256
// Gaussians (in sort order) the block scans their key counts, then each
// thread produces the keys at consecutive output positions, finding its
// Gaussian by binary search over the round's offsets in LDS. Writes are
// coalesced and the work per thread is balanced whatever the splat sizes.
// kGlobal64: Gaussians in id order, key = (tile << 32) | depth bits.
// Otherwise Gaussians in depth order, key = tile. Keys past the budget `cap`
// are dropped (counted by k_count_top).
template <bool kGlobal64>
__global__ __launch_bounds__(kBlock) void k_emit(const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ cntS,
                                                 const uint2* __restrict__ rectS,
                                                 const uint32_t* __restrict__ dkS,
                                                 const uint32_t* __restrict__ dN,
                                                 const uint64_t* __restrict__ partials, int tilesX,
                                                 uint64_t cap, const float4* __restrict__ rec,
                                                 void* __restrict__ keysOut,
                                                 uint32_t* __restrict__ valsOut) {
  __shared__ uint32_t wsum[4];
  __shared__ uint32_t sOff[kBlock];
  __shared__ uint2 sRect[kBlock];
  __shared__ uint32_t sId[kBlock];
  __shared__ uint32_t sDk[kBlock];
  __shared__ gs_cover_splat sG[kBlock];
  const int tid = threadIdx.x;
  const uint32_t N = *dN;
  const uint32_t r = blockIdx.x * kBlock + tid;
  const bool ok = r < N;
  const uint32_t c = ok ? cntS[r] : 0u;
  uint32_t tot;
  const uint32_t ex = block_excl_scan(c, tot, wsum);
  sOff[tid] = ex;
  sRect[tid] = ok ? rectS[r] : make_uint2(0, 0);
  const uint32_t gidr = ok ? (kGlobal64 ? r : order[r]) : 0u;
  sId[tid] = gidr;
  if (kGlobal64) sDk[tid] = ok ? dkS[r] : 0u;
  if (c != 0u) {
    const float4 g0 = rec[4 * (size_t)gidr + 0];
    const float4 g1 = rec[4 * (size_t)gidr + 1];
    sG[tid] = gs_cover_prepare(g0.x, g0.y, g0.z, g0.w, g1.x, g1.z);
  }
  __syncthreads();
  const uint64_t run = partials[blockIdx.x];
  for (uint32_t q = tid; q < tot; q += kBlock) {
    int g = 0;
#pragma unroll
    for (int step = kBlock / 2; step > 0; step >>= 1)
      if (sOff[g + step] <= q) g += step;
    const uint64_t pos = run + q;
    if (pos < cap) {
      const uint2 rc = sRect[g];
      const uint32_t tx0 = rc.x & 0xffffu, tx1 = rc.x >> 16, ty0 = rc.y & 0xffffu;
      const uint32_t wdt = tx1 - tx0 + 1u;
      const uint32_t li = q - sOff[g];
      const uint32_t dy = li / wdt;
      const uint32_t tx = tx0 + (li - dy * wdt), ty = ty0 + dy;
      const uint32_t tile = ty * (uint32_t)tilesX + tx;
      if (kGlobal64) ((uint64_t*)keysOut)[pos] = ((uint64_t)tile << 32) | sDk[g];
      else ((uint32_t*)keysOut)[pos] = tile;
      const gs_cover_splat sg = sG[g];
      const uint32_t mask = gs_cover_tile_mask(&sg, (int)tx, (int)ty);
      valsOut[pos] = sId[g] | (mask << kMaskShift);
    }
  }
}

// Tile ranges from the sorted keys: ranges[t] = (first key, one past the
// last). A tile with no keys gets (o, o), o = number of keys of lower tiles,
// as the reference's exclusive scan gives it (shaders.js:366-371); the
// threads at key boundaries fill the empty tiles in between, and the thread
// holding the last key fills the tiles after it.
template <typename KeyT>
__global__ void k_tile_ranges(const KeyT* __restrict__ keys, const uint32_t* __restrict__ dN,
                              int shift, int numTiles, uint2* __restrict__ ranges) {
  const uint32_t N = *dN;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < N; k += gridDim.x * blockDim.x) {
    const uint32_t t = (uint32_t)(keys[k] >> shift);
    const uint32_t tp = k == 0 ? 0xffffffffu : (uint32_t)(keys[k - 1] >> shift);
    if (tp != t) {
      ranges[t].x = k;
      for (uint32_t e = tp + 1u; e < t; e++) ranges[e] = make_uint2(k, k);  // tp + 1 wraps to 0 for k = 0
    }
    if (k == N - 1) {
      ranges[t].y = N;
      for (uint32_t e = t + 1u; e < (uint32_t)numTiles; e++) ranges[e] = make_uint2(N, N);
    } else if ((uint32_t)(keys[k + 1] >> shift) != t) {
      ranges[t].y = k + 1;
    }
  }
}

// Raster launch order: tiles by descending key count (a 256-bucket
// pseudo-logarithmic counting sort; order within a bucket is arbitrary).
// Launching the heaviest tiles first keeps the last partial wave of
// workgroups short. The order changes no result: tiles are independent and
// the fixed-point gradient sums are integer.
__device__ __forceinline__ uint32_t tile_bucket(uint2 r) {
  const uint32_t c = r.y - r.x;
  if (c < 16u) return c;
  const int e = 31 - (int)__builtin_clz(c);
  return min(255u, (uint32_t)(e - 3) * 16u + ((c >> (e - 4)) & 15u));
}

__global__ __launch_bounds__(1024) void k_tile_order(const uint2* __restrict__ ranges, int numTiles,
                                                     uint32_t* __restrict__ order) {
  __shared__ uint32_t cnt[256];
  const int tid = threadIdx.x;
  if (tid < 256) cnt[tid] = 0u;
  __syncthreads();
  for (int t = tid; t < numTiles; t += 1024) atomicAdd(&cnt[tile_bucket(ranges[t])], 1u);
  __syncthreads();
  if (tid == 0) {   // exclusive offsets, highest bucket first
    uint32_t run = 0;
    for (int b = 255; b >= 0; b--) {
      const uint32_t c = cnt[b];
      cnt[b] = run;
      run += c;
    }
  }
  __syncthreads();
  for (int t = tid; t < numTiles; t += 1024) order[atomicAdd(&cnt[tile_bucket(ranges[t])], 1u)] = (uint32_t)t;
}

// ---- host launchers ----

void launch_fill_u32(hipStream_t st, uint32_t* p, int64_t count, uint32_t v) {
  if (count <= 0) return;
  k_fill_u32<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(p, count, v);
}

// one pass: histogram, per-digit scan (with digit totals), stable scatter
template <typename KeyT>
static void radix_pass(hipStream_t st, const SortScratch& s, const KeyT* kin, const uint32_t* vin,
                       KeyT* kout, uint32_t* vout, const uint32_t* dN, int shift) {
  k_radix_hist<KeyT><<<s.G, kBlock, 0, st>>>(kin, dN, shift, s.G, s.hist);
  k_radix_scan<<<256, kBlock, 0, st>>>(s.hist, s.G, s.totals);
  k_radix_scatter<KeyT, true><<<s.G, kBlock, 0, st>>>(kin, vin, dN, shift, s.G, s.hist, kout, vout, s.totals);
}

// Stable LSD sort of (key, val) pairs by key bits [bit0, bit1), 8 per pass.
// Pass p reads the previous pass's output (the input for p = 0; vals = iota
// when vin is null) and writes (k1, v1) for even p, (k2, v2) for odd p. The
// input may alias (k2, v2): it is consumed by the first pass.
template <typename KeyT>
static void radix_sort(hipStream_t st, const SortScratch& s, const KeyT* kin, const uint32_t* vin,
                       KeyT* k1, uint32_t* v1, KeyT* k2, uint32_t* v2, const uint32_t* dN,
                       int bit0, int bit1, const KeyT** kres, const uint32_t** vres) {
  const KeyT* ki = kin;
  const uint32_t* vi = vin;
  bool odd = false;
  for (int b = bit0; b < bit1; b += 8) {
    KeyT* ko = odd ? k2 : k1;
    uint32_t* vo = odd ? v2 : v1;
    radix_pass<KeyT>(st, s, ki, vi, ko, vo, dN, b);
    ki = ko;
    vi = vo;
    odd = !odd;
  }
  *kres = ki;
  *vres = vi;
}

int tile_bits(int numTiles) {
  int b = 0;
  while ((1 << b) < numTiles) b++;
  return b < 1 ? 1 : b;
}

void launch_bin_sort(hipStream_t st, const BinArgs& a, BinOut& out) {
  const SortScratch& s = a.scratch;
  const int tb = tile_bits(a.numTiles);
  const int tpasses = (tb + 7) / 8;
  const int Gc = (int)((a.n + kBlock - 1) / kBlock);
  if (a.mode == 0) {
    // 1) stable sort of the Gaussians by depth bits, values = ids
    const uint32_t* dk;
    const uint32_t* order;
    radix_sort<uint32_t>(st, s, a.depthKey, nullptr, a.dkA, a.dvA, a.dkB, a.dvB, a.dN, 0, 32, &dk, &order);
    // 2) per-Gaussian key counts in depth order -> offsets -> one key per tile
    k_count_partials<<<Gc, kBlock, 0, st>>>(order, a.cnt, a.rect, a.depthKey, a.dN, s.cntS, s.rectS, nullptr,
                                            s.partials, s.vpartials);
    k_count_top<<<1, 1024, 0, st>>>(s.partials, s.vpartials, Gc, a.cap, a.numKeys, a.totalKeys, a.numVisible, a.stats);
    k_emit<false><<<Gc, kBlock, 0, st>>>(order, s.cntS, s.rectS, nullptr, a.dN, s.partials, a.tilesX, a.cap,
                                         a.rec, a.tkA, a.tvA);
    // 3) stable sort by tile id
    const uint32_t* tk;
    const uint32_t* tv;
    radix_sort<uint32_t>(st, s, (const uint32_t*)a.tkA, a.tvA, (uint32_t*)a.tkB, a.tvB,
                         (uint32_t*)a.tkA, a.tvA, a.numKeys, 0, 8 * tpasses, &tk, &tv);
    out.keys = tk;
    out.ids = tv;
    out.keyBytes = 4;
    launch_fill_u32(st, (uint32_t*)a.ranges, 2 * (int64_t)a.numTiles, 0u);
    k_tile_ranges<uint32_t><<<1024, 256, 0, st>>>(tk, a.numKeys, 0, a.numTiles, a.ranges);
    k_tile_order<<<1, 1024, 0, st>>>(a.ranges, a.numTiles, a.tileOrder);
  } else {
    k_count_partials<<<Gc, kBlock, 0, st>>>(nullptr, a.cnt, a.rect, a.depthKey, a.dN, s.cntS, s.rectS, s.dkS,
                                            s.partials, s.vpartials);
    k_count_top<<<1, 1024, 0, st>>>(s.partials, s.vpartials, Gc, a.cap, a.numKeys, a.totalKeys, a.numVisible, a.stats);
    k_emit<true><<<Gc, kBlock, 0, st>>>(nullptr, s.cntS, s.rectS, s.dkS, a.dN, s.partials, a.tilesX, a.cap,
                                        a.rec, a.tkA, a.tvA);
    const uint64_t* tk;
    const uint32_t* tv;
    radix_sort<uint64_t>(st, s, (const uint64_t*)a.tkA, a.tvA, (uint64_t*)a.tkB, a.tvB,
                         (uint64_t*)a.tkA, a.tvA, a.numKeys, 0, 32 + 8 * tpasses, &tk, &tv);
    out.keys = tk;
    out.ids = tv;
    out.keyBytes = 8;
    launch_fill_u32(st, (uint32_t*)a.ranges, 2 * (int64_t)a.numTiles, 0u);
    k_tile_ranges<uint64_t><<<1024, 256, 0, st>>>(tk, a.numKeys, 32, a.numTiles, a.ranges);
    k_tile_order<<<1, 1024, 0, st>>>(a.ranges, a.numTiles, a.tileOrder);
  }
  out.tileOrder = a.tileOrder;
}

}  // namespace gs
