// gs_sort.hip — global binning: one key per (Gaussian, covered tile),
// ordered by (tile, depth, Gaussian id), the total order the reference's
// per-tile bitonic sort produces (shaders.js:418-500, tie-break :454).
//
// Three strategies, identical output: tile segments (the default, below the
// radix passes in this file), and two radix strategies:
//  * depth-then-tile (also the fallback for views above kMaxSegTiles tiles):
//    stable LSD radix sort of the n Gaussians by
//    depth bits (emitted in id order, so ties keep id order), then keys are
//    emitted in that order and stably sorted by tile id only
//    (ceil(tileBits/8) passes). Sorting Gaussians instead of keys moves the
//    32 depth bits off the long key array. The first depth pass drops the
//    culled Gaussians (key 0xffffffff, which would sort last and emit no
//    keys) and counts the rest on the device, so the later passes, the count
//    and the emission run over the visible Gaussians only.
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

// kFilter: keys equal to ~0 (culled Gaussians) are left out
template <typename KeyT, bool kFilter>
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
      d[i] = 0x100u;
      if (idx < hi) {
        const KeyT k = keys[idx];
        if (!kFilter || k != (KeyT)~(KeyT)0) d[i] = digit_of(k, shift);
      }
    }
#pragma unroll
    for (int i = 0; i < kItems; i++) {
      const bool ok = d[i] < 0x100u;
      const uint64_t act = wave_ballot(ok);
      if (!kFilter && act == 0ull) break;   // unfiltered, only the chunk's tail is empty
      if (act == 0ull) continue;
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

// kFilter: keys equal to ~0 are dropped, and block 0 stores the number of
// keys kept in *nOut
template <typename KeyT, bool kWriteKeys, bool kFilter>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const KeyT* __restrict__ keysIn,
                                                          const uint32_t* __restrict__ valsIn,
                                                          const uint32_t* __restrict__ dN,
                                                          int shift, int G,
                                                          const uint32_t* __restrict__ scan,
                                                          KeyT* __restrict__ keysOut,
                                                          uint32_t* __restrict__ valsOut,
                                                          const uint32_t* __restrict__ totals,
                                                          uint32_t* __restrict__ nOut) {
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
    if (kFilter && blockIdx.x == 0 && tid == 0) *nOut = wsum[0] + wsum[1] + wsum[2] + wsum[3];
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
      bool ok = idx < hi;
      key[i] = ok ? keysIn[idx] : (KeyT)0;
      if (kFilter) ok = ok && key[i] != (KeyT)~(KeyT)0;
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
    const uint32_t kept = wsum[0] + wsum[1] + wsum[2] + wsum[3];   // keys of this round staged below
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
    for (uint32_t j = tid; j < kept; j += kBlock) {
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

// Cover-test constants of Gaussian i's raster record (gs_project.hip): the
// conic and mean from its first float4, C and the cover level L from its
// second (L in the high half of w as an fp16 upper bound when the record
// holds 16-bit colours).
__device__ __forceinline__ gs_cover_splat record_cover(const float4* __restrict__ rec, int recFmt, size_t i) {
  const size_t rs = (size_t)rec_stride(recFmt);
  const float4 g0 = rec[rs * i + 0];
  const float4 g1 = rec[rs * i + 1];
  const float L = recFmt == 0 ? g1.z : gs_h2f(__float_as_uint(g1.w) >> 16);
  return gs_cover_prepare(g0.x, g0.y, g0.z, g0.w, g1.x, L);
}

template <bool kGlobal64>
__global__ __launch_bounds__(kBlock) void k_emit(const uint32_t* __restrict__ order,
                                                 const uint32_t* __restrict__ cntS,
                                                 const uint2* __restrict__ rectS,
                                                 const uint32_t* __restrict__ dkS,
                                                 const uint32_t* __restrict__ dN,
                                                 const uint64_t* __restrict__ partials, int tilesX, int tileSize,
                                                 uint64_t cap, const float4* __restrict__ rec, int recFmt,
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
  if (c != 0u) sG[tid] = record_cover(rec, recFmt, gidr);
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
      const uint32_t mask = gs_cover_tile_mask(&sg, (int)tx, (int)ty, tileSize);
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

// ---- tile segments (default strategy) ----
//
// The reference's own structure (shaders.js:318-500): per-tile key counts, a
// scan that lays out one segment per tile under the entry budget, a scatter
// of every (Gaussian, tile) key into its tile's segment, and a sort of each
// segment. Keys are u64 (depth bits << 32 | id << 4 | block mask): within a
// tile no two keys share an id, so sorting by the whole word orders a
// segment by (depth, id), the reference's order (shaders.js:454), whatever
// order the scatter left it in.
//
// No global atomics: blocks of kSegGroup Gaussians count their keys per tile
// in an LDS histogram and write it as one row of a [block][tile] table; a
// column scan turns the table into each block's offset inside every tile's
// segment (and the tile totals); the scatter then places keys with LDS
// counters only. A segment of up to 1024 keys sorts in one block: 256-key
// runs by single waves (bitonic through lane shuffles), then up to two LDS
// merge levels; longer segments sort 1024-key runs that way and merge them
// pairwise in global memory.

constexpr int kSegThreads = 1024;     // threads of the count and scatter blocks
// per-tile LDS arrays (histogram, counters, scan input) are sized to the
// view's tile count at launch, so that small views fit more blocks per CU
extern __shared__ __align__(16) uint32_t seg_dyn[];
constexpr uint32_t kShSort = 2048;    // shaders.js:33 SHARED_SORT: longer segments count pow2-padded
constexpr int kScanChunk = 64;        // tiles per scan thread (shaders.js:325)

__device__ __forceinline__ uint64_t seg_key(uint32_t depth, uint32_t id, uint32_t mask) {
  return ((uint64_t)depth << 32) | ((uint64_t)id << 4) | (uint64_t)mask;
}
__device__ __forceinline__ uint32_t seg_val(uint64_t k) {
  return (uint32_t)((k >> 4) & kIdMask) | ((uint32_t)(k & 0xfu) << kMaskShift);
}

// the binning job of this block: the single view's (by value), or view v
// of the batched array
template <bool kJobs>
__device__ __forceinline__ SegJob seg_job(const SegJob& arg, const SegJob* __restrict__ jobs, int v) {
  if constexpr (kJobs) return global_job(jobs[v]);
  else return arg;
}

// exclusive scan over a block of kT threads (kT / 64 waves)
template <int kT>
__device__ __forceinline__ uint32_t block_excl_scan_t(uint32_t v, uint32_t& total, uint32_t* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  __syncthreads();
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < kT / 64; k++) {
    const uint32_t ws = wsum[k];
    pre += k < w ? ws : 0u;
    tot += ws;
  }
  total = tot;
  return pre + x - v;
}

// Gaussians [lo, hi) of a segment block, one round of kSegThreads at a time:
// per round the keys are spread over the threads (block scan of the counts;
// each Gaussian writes its index over its key slots in an owner table, or,
// for a round of more than kOwnerCap keys, each key finds its Gaussian by a
// binary search of the offsets), and f(g, r0, tile, tx, ty) runs once per
// key, g the round-local index. `prep` stages per-Gaussian data: prep.load
// issues its global reads, prep.store puts them in LDS.
//
// The loads are pipelined one round ahead (a block holds only 16 waves per
// CU, too few to hide a load's latency otherwise): the next round's counts
// and rectangles are issued at the start of a round, its per-Gaussian data
// before this round's keys, and the first round's before `init` (the
// block's per-tile setup).
constexpr uint32_t kOwnerCap = 8192;

struct SegRound {
  uint32_t* sOff;
  uint2* sRect;
  uint32_t* wsum;
  uint16_t* owner;   // kOwnerCap
};

template <typename Prep, typename Init, typename F>
__device__ __forceinline__ void seg_rounds(int64_t lo, int64_t hi, const uint32_t* __restrict__ cnt,
                                           const uint2* __restrict__ rect, int tilesX, const SegRound& sr,
                                           const Prep& prep, Init init, F f) {
  const int tid = threadIdx.x;
  int64_t i = lo + tid;
  uint32_t c = i < hi ? cnt[i] : 0u;
  uint2 rc = i < hi ? rect[i] : make_uint2(0, 0);   // read whatever c is: no dependent load
  typename Prep::State ps = prep.load(i, c);
  init();
  for (int64_t r0 = lo; r0 < hi; r0 += kSegThreads) {
    const int64_t in = i + kSegThreads;
    const uint32_t cn = in < hi ? cnt[in] : 0u;
    const uint2 rn = in < hi ? rect[in] : make_uint2(0, 0);
    uint32_t tot;
    const uint32_t ex = block_excl_scan_t<kSegThreads>(c, tot, sr.wsum);
    sr.sOff[tid] = ex;
    sr.sRect[tid] = c ? rc : make_uint2(0, 0);
    prep.store(tid, c, ps);
    const bool owned = tot <= kOwnerCap;
    if (owned)
      for (uint32_t k = 0; k < c; k++) sr.owner[ex + k] = (uint16_t)tid;
    __syncthreads();
    ps = prep.load(in, cn);
    for (uint32_t q = tid; q < tot; q += kSegThreads) {
      int g = 0;
      if (owned) {
        g = sr.owner[q];
      } else {
#pragma unroll
        for (int step = kSegThreads / 2; step > 0; step >>= 1)
          if (sr.sOff[g + step] <= q) g += step;
      }
      const uint2 rq = sr.sRect[g];
      const uint32_t tx0 = rq.x & 0xffffu, tx1 = rq.x >> 16, ty0 = rq.y & 0xffffu;
      const uint32_t wdt = tx1 - tx0 + 1u;
      const uint32_t li = q - sr.sOff[g];
      const uint32_t dy = li / wdt;
      const uint32_t tx = tx0 + (li - dy * wdt), ty = ty0 + dy;
      f(g, r0, ty * (uint32_t)tilesX + tx, tx, ty);
    }
    __syncthreads();
    i = in;
    c = cn;
    rc = rn;
  }
}

// count pass: no per-Gaussian data beyond the count; tallies the Gaussians
// with keys
struct CountPrep {
  uint32_t* vis;
  struct State {};
  __device__ State load(int64_t, uint32_t) const { return {}; }
  __device__ void store(int, uint32_t c, const State&) const { *vis += c != 0u; }
};

// Row b of the [block][tile] table: the keys of Gaussians [b G, (b+1) G) per
// tile; visPart[b]: how many of them have keys.
// This block's Gaussians: the chunks of kSegChunk Gaussians whose cost
// prefix p (over the chunks before them) satisfies b K <= p nb < (b + 1) K,
// K the total, so each of the nb blocks takes about K / nb of the cost. A
// chunk costs its keys plus one per Gaussian: a round of Gaussians costs
// its loads, scan and barriers even when they have no keys, so key shares
// alone would hand a block a long run of off-view Gaussians. The first
// block starts at 0 and the last ends at n, and block b's end is block
// b + 1's start by the same formula: the blocks partition [0, n) whatever
// the chunk sums hold (they only set the balance). Returns [lo, hi).
// Logical index of this count/scatter block. Workgroups are dispatched
// round-robin over MI355X's 8 XCDs, so logical blocks 0..nb/8-1 run on the
// first XCD, the next nb/8 on the second, and so on: each tile segment's
// runs from neighbouring logical blocks (adjacent in memory) are written
// from one XCD. Measured ~0.5-1% faster than the identity order at the same
// state. Any bijection partitions the work the same way.
__device__ __forceinline__ int seg_logical_block(int nb) {
  const int i = (int)blockIdx.x;
  return nb % 8 == 0 ? (i % 8) * (nb / 8) + i / 8 : i;
}

__device__ __forceinline__ int2 seg_block_chunks(int64_t n, const uint32_t* __restrict__ chunkKeys, int nb,
                                                 uint64_t* ws64, uint32_t (*wsc)[kSegThreads / 64]) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int64_t nChunks = (n + kSegChunk - 1) / kSegChunk;
  const int64_t m = (nChunks + kSegThreads - 1) / kSegThreads;   // chunks per thread, in order
  const int64_t c0 = (int64_t)tid * m, c1 = c0 + m < nChunks ? c0 + m : nChunks;
  uint64_t part = 0;
  for (int64_t c = c0; c < c1; c++) part += chunkKeys[c] + (uint64_t)kSegChunk;
  uint64_t x = part;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws64[w] = x;
  __syncthreads();
  uint64_t pre = 0, total = 0;
  for (int k = 0; k < kSegThreads / 64; k++) {
    const uint64_t v = ws64[k];
    if (k < w) pre += v;
    total += v;
  }
  pre += x - part;
  const uint64_t b = (uint64_t)seg_logical_block(nb);
  const uint64_t tLo = b * total, tHi = (b + 1) * total;
  uint32_t nLo = 0, nHi = 0;          // chunks with prefix * nb below each bound
  for (int64_t c = c0; c < c1; c++) {
    const uint64_t pn = pre * (uint64_t)nb;
    nLo += pn < tLo;
    nHi += pn < tHi;
    pre += chunkKeys[c] + (uint64_t)kSegChunk;
  }
  for (int o = 32; o > 0; o >>= 1) {
    nLo += (uint32_t)__shfl_xor((int)nLo, o);
    nHi += (uint32_t)__shfl_xor((int)nHi, o);
  }
  if (lane == 0) { wsc[0][w] = nLo; wsc[1][w] = nHi; }
  __syncthreads();
  uint32_t sLo = 0, sHi = 0;
  for (int k = 0; k < kSegThreads / 64; k++) { sLo += wsc[0][k]; sHi += wsc[1][k]; }
  __syncthreads();   // ws64 and wsc are free again
  const int64_t lo = (int64_t)sLo * kSegChunk;
  const int64_t hi = b == (uint64_t)(nb - 1) ? n : (int64_t)sHi * kSegChunk;
  return make_int2((int)(lo < n ? lo : n), (int)(hi < n ? hi : n));
}

template <bool kJobs>
__global__ __launch_bounds__(kSegThreads) void k_seg_count(SegJob arg, const SegJob* __restrict__ jobs) {
  const SegJob j = seg_job<kJobs>(arg, jobs, blockIdx.y);
  if (kJobs && j.skip) return;
  const int64_t n = j.n;
  const int nb = j.nb, tilesX = j.tilesX, numTiles = j.numTiles;
  const uint32_t* __restrict__ cnt = j.cnt;
  const uint2* __restrict__ rect = j.rect;
  const uint32_t* __restrict__ chunkKeys = j.chunkKeys;
  uint32_t* __restrict__ blockStart = j.blockStart;
  uint32_t* __restrict__ table = j.table;
  uint32_t* __restrict__ visPart = j.visPart;
  uint32_t* hist = seg_dyn;   // numTiles
  __shared__ uint32_t sOff[kSegThreads];
  __shared__ uint2 sRect[kSegThreads];
  __shared__ uint32_t wsum[kSegThreads / 64];
  __shared__ uint16_t owner[kOwnerCap];
  __shared__ uint64_t ws64[kSegThreads / 64];
  __shared__ uint32_t wsc[2][kSegThreads / 64];
  const int tid = threadIdx.x;
  const int2 r = seg_block_chunks(n, chunkKeys, nb, ws64, wsc);
  const int lb = seg_logical_block(nb);
  const int64_t lo = r.x, hi = r.y;
  if (tid == 0) {
    blockStart[lb] = (uint32_t)lo;
    if (lb == nb - 1) blockStart[nb] = (uint32_t)hi;
  }
  uint32_t vis = 0;
  const SegRound sr{sOff, sRect, wsum, owner};
  seg_rounds(lo, hi, cnt, rect, tilesX, sr, CountPrep{&vis},
             [&]() { for (int t = tid; t < numTiles; t += kSegThreads) hist[t] = 0u; },
             [&](int, int64_t, uint32_t tile, uint32_t, uint32_t) { atomicAdd(&hist[tile], 1u); });
  uint32_t tot;
  block_excl_scan_t<kSegThreads>(vis, tot, wsum);
  if (tid == 0) visPart[lb] = tot;
  __syncthreads();
  uint32_t* row = table + (int64_t)lb * numTiles;
  for (int t = tid; t < numTiles; t += kSegThreads) row[t] = hist[t];
}

// Column scan of the table: table[b][t] becomes the exclusive prefix over
// blocks, tileCnt[t] the column total. Blocks of 256 threads handle 16
// tiles, 16 threads per tile (a wave reads 16 consecutive tiles of 4 rows),
// each thread a sixteenth of the rows.
template <bool kJobs>
__global__ __launch_bounds__(kBlock) void k_seg_colscan(SegJob arg, const SegJob* __restrict__ jobs) {
  const SegJob j = seg_job<kJobs>(arg, jobs, blockIdx.y);
  if (kJobs && j.skip) return;
  uint32_t* __restrict__ table = j.table;
  const int nb = j.nb, numTiles = j.numTiles;
  uint32_t* __restrict__ tileCnt = j.tileCnt;
  __shared__ uint32_t part[16][16];
  const int tl = threadIdx.x & 15, q = threadIdx.x >> 4;
  const int t = blockIdx.x * 16 + tl;
  const int per = (nb + 15) / 16;
  const int b0 = q * per, b1 = b0 + per < nb ? b0 + per : nb;
  // up to kRegRows rows per thread (tables of up to 256 rows: 1M Gaussians)
  // stay in registers: one batch of independent loads, no second read
  constexpr int kRegRows = 16;
  uint32_t v[kRegRows];
  uint32_t sum = 0;
  if (per <= kRegRows) {
#pragma unroll
    for (int k = 0; k < kRegRows; k++) {
      const int b = b0 + k;
      v[k] = (t < numTiles && b < b1) ? table[(int64_t)b * numTiles + t] : 0u;
      sum += v[k];
    }
  } else if (t < numTiles) {
    for (int b = b0; b < b1; b++) sum += table[(int64_t)b * numTiles + t];
  }
  part[q][tl] = sum;
  __syncthreads();
  if (t >= numTiles) return;
  uint32_t run = 0;
  for (int k = 0; k < q; k++) run += part[k][tl];
  if (per <= kRegRows) {
#pragma unroll
    for (int k = 0; k < kRegRows; k++) {
      const int b = b0 + k;
      if (b < b1) {
        table[(int64_t)b * numTiles + t] = run;
        run += v[k];
      }
    }
  } else {
    for (int b = b0; b < b1; b++) {
      uint32_t* p = table + (int64_t)b * numTiles + t;
      const uint32_t x = *p;
      *p = run;
      run += x;
    }
  }
  if (q == 15) tileCnt[t] = run;
}

__device__ __forceinline__ uint32_t seg_padded(uint32_t c) {
  // shaders.js:329-332: segments over SHARED_SORT take a power-of-two slot
  return c <= kShSort ? c : 1u << (32 - __clz((int)(c - 1u)));
}

// Single block of 1024 threads: the reference's budgeted segment scan
// (shaders.js:321-370), the raster launch order and the visible count.
//
// The reference's thread li owns the chunk of tiles [64 li, 64 li + 64). The
// chunk starts at the unbudgeted exclusive prefix of the padded sizes of
// the chunks before it; within the chunk, a tile whose padded size would
// take the running total past the budget is dropped (and counted, as the
// reference counts it, even when empty). Here four threads share a chunk,
// 16 tiles each: a chunk whose unbudgeted end stays within the budget drops
// nothing, so each quarter starts at its unbudgeted prefix; a chunk that
// may overflow is walked in order by its first thread, which leaves each
// tile's kept count in LDS. Kept tiles get compact, unpadded segments:
// ranges[t] = (start, start + count). Then the tiles are ordered by
// descending key count for the raster launch (as k_tile_order).
constexpr int kScanThreads = 1024;
constexpr int kScanQuarter = kScanChunk / 4;   // tiles per thread

template <bool kJobs>
__global__ __launch_bounds__(kScanThreads) void k_seg_scan(SegJob arg, const SegJob* __restrict__ jobs) {
  const SegJob j = seg_job<kJobs>(arg, jobs, blockIdx.y);
  if (kJobs && j.skip) return;
  const uint32_t* __restrict__ tileCnt = j.tileCnt;
  const int numTiles = j.numTiles, nb = j.nb;
  const uint64_t cap = j.cap;
  const uint32_t* __restrict__ visPart = j.visPart;
  uint2* __restrict__ ranges = j.ranges;
  uint32_t* __restrict__ order = j.tileOrder;
  uint32_t* __restrict__ chunkKeys = j.chunkKeys;
  const int64_t nChunks = (j.n + kSegChunk - 1) / kSegChunk;
  // the count pass has read the chunk key sums: clear them for the next
  // projection to fill
  for (int64_t c = threadIdx.x; c < nChunks; c += kScanThreads) chunkKeys[c] = 0u;
  // tile t at sc[t + t / 16] (numTiles + numTiles / 16 + 1 words): coalesced
  // fill, conflict-free walks (stride 17 between threads)
  uint32_t* sc = seg_dyn;
  __shared__ uint64_t wsA[kScanThreads / 64];
  __shared__ uint64_t wsB[kScanThreads / 64];
  __shared__ uint32_t wsC[kScanThreads / 64];
  __shared__ uint32_t wsV[kScanThreads / 64];
  __shared__ uint32_t bucket[256];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int t = tid; t < numTiles; t += kScanThreads) sc[t + t / kScanQuarter] = tileCnt[t];
  if (tid < 256) bucket[tid] = 0u;
  __syncthreads();
  const int t0 = tid * kScanQuarter;
  uint32_t* mine = sc + tid * (kScanQuarter + 1);
  const int nt = numTiles - t0 >= kScanQuarter ? kScanQuarter : (numTiles - t0 > 0 ? numTiles - t0 : 0);
  uint64_t pad = 0, raw = 0;
  for (int k = 0; k < nt; k++) {
    const uint32_t c = mine[k];
    pad += seg_padded(c);
    raw += c;
  }
  // unbudgeted exclusive prefix (tile order), the chunk's start and end
  uint64_t x = pad;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsA[w] = x;
  __syncthreads();
  uint64_t pre = 0;
  for (int k = 0; k < w; k++) pre += wsA[k];
  const uint64_t acc0 = pre + x - pad;                // this quarter's unbudgeted start
  const uint64_t q0 = __shfl(acc0, lane & ~3);        // the chunk's start
  const uint64_t qend = __shfl(acc0 + pad, lane | 3); // the chunk's unbudgeted end
  // an overflowing chunk: its first thread walks its 64 tiles in order and
  // rewrites their counts as kept counts (0 = dropped)
  uint32_t dropped = 0;
  if (qend > cap && (tid & 3) == 0) {
    uint64_t acc = q0;
    for (int k = 0; k < kScanChunk; k++) {
      const int t = t0 + k;
      if (t >= numTiles) break;
      uint32_t* cp = sc + t + t / kScanQuarter;
      const uint32_t c = *cp;
      uint64_t p = seg_padded(c);
      if (acc + p > cap) { p = 0; dropped++; }
      acc += p;
      *cp = p ? c : 0u;
    }
  }
  __syncthreads();
  uint64_t kept = 0;
  for (int k = 0; k < nt; k++) kept += mine[k];
  // compact starts
  uint64_t y = kept;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t z = __shfl_up(y, o);
    if (lane >= o) y += z;
  }
  uint64_t rsum = raw;
  uint32_t dsum = dropped;
  uint32_t vsum = tid < nb ? visPart[tid] : 0u;   // nb <= kSegMaxRows = kScanThreads
  for (int o = 32; o > 0; o >>= 1) {
    rsum += __shfl_xor(rsum, o);
    dsum += __shfl_xor(dsum, o);
    vsum += __shfl_xor(vsum, o);
  }
  if (lane == 63) wsA[w] = y;   // wsA's earlier reads all precede the barrier above
  if (lane == 0) { wsB[w] = rsum; wsC[w] = dsum; wsV[w] = vsum; }
  __syncthreads();
  uint64_t kpre = 0;
  for (int k = 0; k < w; k++) kpre += wsA[k];
  uint64_t start = kpre + y - kept;
  for (int k = 0; k < nt; k++) {
    const uint32_t len = mine[k];
    const uint2 rg = make_uint2((uint32_t)start, (uint32_t)(start + len));
    ranges[t0 + k] = rg;
    atomicAdd(&bucket[tile_bucket(rg)], 1u);
    start += len;
  }
  if (tid == kScanThreads - 1) *j.numKeys = (uint32_t)start;
  if (tid == 0) {
    uint64_t r = 0;
    uint32_t d = 0, v = 0;
    for (int k = 0; k < kScanThreads / 64; k++) { r += wsB[k]; d += wsC[k]; v += wsV[k]; }
    *j.totalKeys = r;
    if (d) atomicAdd(&j.stats[3], (double)d);   // the views' scans run together
    *j.numVisible = v;
  }
  __syncthreads();
  // exclusive bucket offsets, heaviest bucket first: threads 0..255 scan
  // bucket 255 - tid (four waves)
  uint32_t bc = 0, bx = 0;
  if (tid < 256) {
    bc = bucket[255 - tid];
    bx = bc;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t z = __shfl_up(bx, o);
      if (lane >= o) bx += z;
    }
    if (lane == 63) wsC[w] = bx;
  }
  __syncthreads();
  if (tid < 256) {
    uint32_t bp = 0;
    for (int k = 0; k < w; k++) bp += wsC[k];
    bucket[255 - tid] = bp + bx - bc;
  }
  __syncthreads();
  for (int t = tid; t < numTiles; t += kScanThreads) {
    const uint32_t len = sc[t + t / kScanQuarter];
    order[atomicAdd(&bucket[tile_bucket(make_uint2(0u, len))], 1u)] = (uint32_t)t;
  }
}

// scatter pass: depth key and cover-test constants of each Gaussian with keys
struct ScatterPrep {
  const uint32_t* __restrict__ depthKey;
  const float4* __restrict__ rec;
  int recFmt;
  uint32_t* sDk;
  gs_cover_splat* sG;
  struct State {
    uint32_t dk;
    float4 g0, g1;
  };
  __device__ State load(int64_t i, uint32_t c) const {
    State s{0u, make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
    if (c != 0u) {
      const int64_t rs = rec_stride(recFmt);
      s.dk = depthKey[i];
      s.g0 = rec[rs * i + 0];
      s.g1 = rec[rs * i + 1];
    }
    return s;
  }
  __device__ void store(int k, uint32_t c, const State& s) const {
    sDk[k] = s.dk;
    if (c != 0u) {
      const float L = recFmt == 0 ? s.g1.z : gs_h2f(__float_as_uint(s.g1.w) >> 16);
      sG[k] = gs_cover_prepare(s.g0.x, s.g0.y, s.g0.z, s.g0.w, s.g1.x, L);
    }
  }
};

// Every (Gaussian, kept tile) key into its tile's segment: each tile's LDS
// counter starts at the tile's segment start plus this block's offset in
// the tile (the column scan), so a key's slot is one LDS atomic; the tile's
// block mask (gs_cover.h) rides in the low 4 bits.
template <bool kJobs>
__global__ __launch_bounds__(kSegThreads) void k_seg_scatter(SegJob arg, const SegJob* __restrict__ jobs) {
  const SegJob j = seg_job<kJobs>(arg, jobs, blockIdx.y);
  if (kJobs && j.skip) return;
  const uint32_t* __restrict__ blockStart = j.blockStart;
  const uint32_t* __restrict__ cnt = j.cnt;
  const uint2* __restrict__ rect = j.rect;
  const uint32_t* __restrict__ depthKey = j.depthKey;
  const float4* __restrict__ rec = j.rec;
  const int recFmt = j.recFmt, tilesX = j.tilesX, tileSize = j.tile, numTiles = j.numTiles;
  const uint32_t* __restrict__ table = j.table;
  const uint2* __restrict__ ranges = j.ranges;
  uint64_t* __restrict__ keys = j.keys;
  uint32_t* cur = seg_dyn;                            // numTiles: next slot of each tile
  uint8_t* dropped = (uint8_t*)(seg_dyn + numTiles);  // numTiles: tile dropped by the budget
  __shared__ uint32_t sOff[kSegThreads];
  __shared__ uint2 sRect[kSegThreads];
  __shared__ uint32_t sDk[kSegThreads];
  __shared__ gs_cover_splat sG[kSegThreads];
  __shared__ uint32_t wsum[kSegThreads / 64];
  __shared__ uint16_t owner[kOwnerCap];
  const int tid = threadIdx.x;
  const int lb = seg_logical_block(j.nb);
  const uint32_t* row = table + (int64_t)lb * numTiles;
  const int64_t lo = blockStart[lb], hi = blockStart[lb + 1];   // the count pass's blocks
  const SegRound sr{sOff, sRect, wsum, owner};
  seg_rounds(lo, hi, cnt, rect, tilesX, sr, ScatterPrep{depthKey, rec, recFmt, sDk, sG},
             [&]() {
#pragma unroll 4
               for (int t = tid; t < numTiles; t += kSegThreads) {
                 const uint2 rg = ranges[t];
                 cur[t] = rg.x + row[t];
                 dropped[t] = rg.y == rg.x;
               }
             },
             [&](int g, int64_t r0, uint32_t tile, uint32_t tx, uint32_t ty) {
               const uint32_t pos = atomicAdd(&cur[tile], 1u);
               if (dropped[tile]) return;   // tile dropped by the budget (shaders.js:406-407)
               const gs_cover_splat sg = sG[g];
               const uint32_t mask = gs_cover_tile_mask(&sg, (int)tx, (int)ty, tileSize);
               keys[pos] = seg_key(sDk[g], (uint32_t)(r0 + g), mask);
             });
}

// x from lane ^ m (m = 1, 2: DPP quad permutes on the VALU; otherwise a
// ds_bpermute), for 64-bit x
template <int kM>
__device__ __forceinline__ uint64_t lane_xor_u64(uint64_t x) {
  const int lo = (int)(uint32_t)x, hi = (int)(uint32_t)(x >> 32);
  int a, b;
  if constexpr (kM == 1) {
    a = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xf, 0xf, false);   // quad_perm [1,0,3,2]
    b = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xf, 0xf, false);
  } else if constexpr (kM == 2) {
    a = __builtin_amdgcn_update_dpp(0, lo, 0x4E, 0xf, 0xf, false);   // quad_perm [2,3,0,1]
    b = __builtin_amdgcn_update_dpp(0, hi, 0x4E, 0xf, 0xf, false);
  } else {
    a = __shfl_xor(lo, kM);
    b = __shfl_xor(hi, kM);
  }
  return ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
}

// One compare-exchange stage (k, j) of the bitonic network over 256 keys,
// element i = 4 lane + r in v[r]: partners at distance j < 4 are in the same
// lane, others in lane ^ (j / 4).
template <uint32_t k, uint32_t j>
__device__ __forceinline__ void bitonic_stage(uint64_t (&v)[4], uint32_t lane) {
  uint64_t o[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    if constexpr (j < 4) o[r] = v[r ^ (int)j];
    else o[r] = lane_xor_u64<(int)(j / 4)>(v[r]);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t i = 4u * lane + (uint32_t)r;
    const bool keepMin = ((i & j) == 0u) == ((i & k) == 0u);
    const bool lt = v[r] < o[r];
    v[r] = keepMin == lt ? v[r] : o[r];
  }
}

template <uint32_t k, uint32_t j>
__device__ __forceinline__ void bitonic_merge_steps(uint64_t (&v)[4], uint32_t lane) {
  bitonic_stage<k, j>(v, lane);
  if constexpr (j > 1) bitonic_merge_steps<k, j / 2>(v, lane);
}

template <uint32_t k>
__device__ __forceinline__ void bitonic_blocks(uint64_t (&v)[4], uint32_t lane) {
  bitonic_merge_steps<k, k / 2>(v, lane);
  if constexpr (k < 256) bitonic_blocks<k * 2>(v, lane);
}

// Ascending bitonic sort of 256 keys within one wave, keys 4 lane .. 4 lane
// + 3 in v[0..3]: no LDS traffic except the exchanges at lane distances of
// 4 and more, no barriers.
__device__ __forceinline__ void wave_sort256(uint64_t (&v)[4]) {
  bitonic_blocks<2>(v, __lane_id());
}

// Sorts m <= kChunk keys: R = ceil(m / 256) runs of 256 sorted by waves
// (the last padded with ~0), then merged pairwise in LDS (merge path, P
// outputs per thread per level; with three runs the last pair of the first
// level has no partner and is copied). Returns the LDS buffer holding the
// sorted keys (sorted[0, m)).
constexpr uint32_t kChunk = 4 * 256;

__device__ uint64_t* seg_sort_chunk(const uint64_t* __restrict__ src, uint32_t m, uint64_t* sa, uint64_t* sb) {
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  // runs of 256 keys: as many as m needs (a segment of 513-768 keys sorts
  // three runs, not four padded ones)
  const uint32_t R = (m + 255u) / 256u;
  // src is device memory: read through a global pointer (in this
  // out-of-line function a plain one compiles to flat loads)
  const __attribute__((address_space(1))) uint64_t* gsrc = (const __attribute__((address_space(1))) uint64_t*)src;
  if (w < R) {
    uint64_t v[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const uint32_t i = 256u * w + 4u * lane + (uint32_t)r;
      v[r] = i < m ? gsrc[i] : ~0ull;
    }
    wave_sort256(v);
#pragma unroll
    for (int r = 0; r < 4; r++) sa[256u * w + 4u * lane + (uint32_t)r] = v[r];
  }
  __syncthreads();
  uint64_t* in = sa;
  uint64_t* out = sb;
  const uint32_t N = 256u * R;
  // outputs per thread: a power of two, so a thread's outputs never straddle
  // two merged pairs (2W >= 512); with three runs 192 threads write 4 each
  const uint32_t P = R == 3u ? 4u : R;
  for (uint32_t W = 256u; W < N; W <<= 1) {
    // thread tid writes outputs [tid P, tid P + P) of N; the pair at base
    // merges A = [base, base + na) and B = [base + na, base + na + nb), the
    // last pair possibly shorter (nb = 0: A is copied)
    const uint32_t o = tid * P;
    if (o < N) {
      const uint32_t base = o / (2u * W) * (2u * W);
      const uint32_t na = N - base < W ? N - base : W;
      const uint32_t nb = N - base - na < W ? N - base - na : W;
      const uint64_t* A = in + base;
      const uint64_t* B = A + na;
      const uint32_t d = o - base;
      uint32_t lo = d > nb ? d - nb : 0u, hi = d < na ? d : na;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (A[mid] < B[d - mid - 1u]) lo = mid + 1u;
        else hi = mid;
      }
      uint32_t ia = lo, ib = d - lo;
      for (uint32_t q = 0; q < P; q++) {
        const bool takeA = ib >= nb || (ia < na && A[ia] < B[ib]);
        out[o + q] = takeA ? A[ia++] : B[ib++];
      }
    }
    __syncthreads();
    uint64_t* t = in;
    in = out;
    out = t;
  }
  return in;
}

// out[0, na + nb) = merge of the sorted runs A and B (keys are distinct):
// each thread finds where its 8 outputs start by a merge-path search.
__device__ void seg_merge(const uint64_t* A, uint32_t na, const uint64_t* B, uint32_t nb, uint64_t* out) {
  const uint32_t N = na + nb;
  for (uint32_t o = threadIdx.x * 8u; o < N; o += kBlock * 8u) {
    uint32_t lo = o > nb ? o - nb : 0u, hi = o < na ? o : na;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (A[mid] < B[o - mid - 1u]) lo = mid + 1u;
      else hi = mid;
    }
    uint32_t ia = lo, ib = o - lo;
    const uint32_t e = o + 8u < N ? o + 8u : N;
    for (uint32_t q = o; q < e; q++) {
      const bool takeA = ib >= nb || (ia < na && A[ia] < B[ib]);
      out[q] = takeA ? A[ia++] : B[ib++];
    }
  }
}

// One block per tile: sorts the tile's segment and writes the raster values
// (id | mask << 28) in order. Segments over kChunk keys sort kChunk-key runs
// and merge them pairwise in global memory (keys <-> tmp). Blocks take the
// tiles in the raster launch order (most keys first), so the few segments
// that need global merges start early instead of trailing the launch.
// Batched: view blockIdx.x, tile rank blockIdx.y, so the launch starts every
// view's heaviest tiles first.
template <bool kJobs>
__global__ __launch_bounds__(kBlock) void k_seg_sort(SegJob arg, const SegJob* __restrict__ jobs) {
  const SegJob j = seg_job<kJobs>(arg, jobs, kJobs ? blockIdx.x : 0);
  if (kJobs && j.skip) return;
  const uint32_t rank = kJobs ? blockIdx.y : blockIdx.x;
  if (kJobs && rank >= (uint32_t)j.numTiles) return;
  uint64_t* __restrict__ keys = j.keys;
  uint64_t* __restrict__ tmp = j.tmp;
  uint32_t* __restrict__ vals = j.ids;
  __shared__ uint64_t sa[kChunk];
  __shared__ uint64_t sb[kChunk];
  const uint2 r = j.ranges[j.tileOrder[rank]];
  const uint32_t s = r.x, len = r.y - r.x;
  const uint32_t tid = threadIdx.x;
  if (len == 0u) return;
  if (len == 1u) {
    if (tid == 0) vals[s] = seg_val(keys[s]);
    return;
  }
  if (len <= kChunk) {
    const uint64_t* sorted = seg_sort_chunk(keys + s, len, sa, sb);
    for (uint32_t i = tid; i < len; i += kBlock) vals[s + i] = seg_val(sorted[i]);
    return;
  }
  for (uint32_t c0 = 0; c0 < len; c0 += kChunk) {
    const uint32_t m = len - c0 < kChunk ? len - c0 : kChunk;
    const uint64_t* sorted = seg_sort_chunk(keys + s + c0, m, sa, sb);
    for (uint32_t i = tid; i < m; i += kBlock) keys[s + c0 + i] = sorted[i];
    __syncthreads();
  }
  uint64_t* src = keys + s;
  uint64_t* dst = tmp + s;
  for (uint32_t wdt = kChunk; wdt < len; wdt <<= 1) {
    for (uint32_t a = 0; a < len; a += 2u * wdt) {
      const uint32_t na = len - a < wdt ? len - a : wdt;
      const uint32_t nb = len - a - na < wdt ? len - a - na : wdt;
      seg_merge(src + a, na, src + a + na, nb, dst + a);
    }
    __syncthreads();
    uint64_t* t = src;
    src = dst;
    dst = t;
  }
  for (uint32_t i = tid; i < len; i += kBlock) vals[s + i] = seg_val(src[i]);
}

// one pass: histogram, per-digit scan (with digit totals), stable scatter
// nOut != nullptr: drop keys equal to ~0 and store the count kept there
template <typename KeyT>
static void radix_pass(hipStream_t st, const SortScratch& s, const KeyT* kin, const uint32_t* vin,
                       KeyT* kout, uint32_t* vout, const uint32_t* dN, int shift, uint32_t* nOut) {
  if (nOut) k_radix_hist<KeyT, true><<<s.G, kBlock, 0, st>>>(kin, dN, shift, s.G, s.hist);
  else k_radix_hist<KeyT, false><<<s.G, kBlock, 0, st>>>(kin, dN, shift, s.G, s.hist);
  k_radix_scan<<<256, kBlock, 0, st>>>(s.hist, s.G, s.totals);
  if (nOut)
    k_radix_scatter<KeyT, true, true><<<s.G, kBlock, 0, st>>>(kin, vin, dN, shift, s.G, s.hist, kout, vout, s.totals,
                                                             nOut);
  else
    k_radix_scatter<KeyT, true, false><<<s.G, kBlock, 0, st>>>(kin, vin, dN, shift, s.G, s.hist, kout, vout,
                                                              s.totals, nullptr);
}

// Stable LSD sort of (key, val) pairs by key bits [bit0, bit1), 8 per pass.
// Pass p reads the previous pass's output (the input for p = 0; vals = iota
// when vin is null) and writes (k1, v1) for even p, (k2, v2) for odd p. The
// input may alias (k2, v2): it is consumed by the first pass. nKept !=
// nullptr: the first pass drops keys equal to ~0, stores the number kept in
// *nKept, and the later passes sort that many.
template <typename KeyT>
static void radix_sort(hipStream_t st, const SortScratch& s, const KeyT* kin, const uint32_t* vin,
                       KeyT* k1, uint32_t* v1, KeyT* k2, uint32_t* v2, const uint32_t* dN,
                       int bit0, int bit1, const KeyT** kres, const uint32_t** vres, uint32_t* nKept = nullptr) {
  const KeyT* ki = kin;
  const uint32_t* vi = vin;
  bool odd = false;
  for (int b = bit0; b < bit1; b += 8) {
    KeyT* ko = odd ? k2 : k1;
    uint32_t* vo = odd ? v2 : v1;
    const bool first = b == bit0;
    radix_pass<KeyT>(st, s, ki, vi, ko, vo, first || !nKept ? dN : nKept, b, first ? nKept : nullptr);
    ki = ko;
    vi = vo;
    odd = !odd;
  }
  *kres = ki;
  *vres = vi;
}

int64_t seg_group(int64_t n) {
  // Gaussians per count/scatter block: n / kSegBlocks (one block per CU),
  // at least kSegMinGroup; the blocks take whole chunks by cost, so the
  // group only sets their number. kSegBlocks <= kSegMaxRows keeps the table
  // within its rows.
  static_assert(kSegBlocks <= kSegMaxRows, "block count above the table's rows");
  const int64_t g = (n + kSegBlocks - 1) / kSegBlocks;
  return g > kSegMinGroup ? g : kSegMinGroup;
}

int tile_bits(int numTiles) {
  int b = 0;
  while ((1 << b) < numTiles) b++;
  return b < 1 ? 1 : b;
}

int seg_blocks(int64_t n) {
  const int64_t G = seg_group(n);
  return (int)((n + G - 1) / G);
}

// The five tile-segment stages for V jobs (kJobs: from the device array
// jobs, view from the block index; else the one job j by value); maxTiles:
// the most tiles of any job. Every job has the same Gaussian count (nb).
template <bool kJobs>
static void seg_launches(hipStream_t st, const SegJob& j, const SegJob* jobs, int V, int maxTiles) {
  static const bool ldsRaised = [] {
    // the per-tile LDS arrays at kMaxSegTiles: 5 B per tile in the scatter
    const int dyn = (kMaxSegTiles + kMaxSegTiles / kScanQuarter + 1) * 5;
    (void)hipFuncSetAttribute((const void*)k_seg_count<kJobs>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    (void)hipFuncSetAttribute((const void*)k_seg_scan<kJobs>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    (void)hipFuncSetAttribute((const void*)k_seg_scatter<kJobs>, hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
    return true;
  }();
  (void)ldsRaised;
  const unsigned v = (unsigned)V;
  k_seg_count<kJobs><<<dim3((unsigned)j.nb, v), kSegThreads, (size_t)maxTiles * 4, st>>>(j, jobs);
  k_seg_colscan<kJobs><<<dim3((unsigned)((maxTiles + 15) / 16), v), kBlock, 0, st>>>(j, jobs);
  k_seg_scan<kJobs><<<dim3(1u, v), kScanThreads, (size_t)(maxTiles + maxTiles / kScanQuarter + 1) * 4, st>>>(j, jobs);
  k_seg_scatter<kJobs><<<dim3((unsigned)j.nb, v), kSegThreads, (size_t)maxTiles * 5, st>>>(j, jobs);
  if (kJobs) k_seg_sort<kJobs><<<dim3(v, (unsigned)maxTiles), kBlock, 0, st>>>(j, jobs);
  else k_seg_sort<kJobs><<<dim3((unsigned)maxTiles), kBlock, 0, st>>>(j, jobs);
}

void launch_bin_sort_views(hipStream_t st, const SegJob* dJobs, const SegJob* hJobs, int V) {
  int maxTiles = 0;
  for (int k = 0; k < V; k++)
    if (!hJobs[k].skip) maxTiles = maxTiles > hJobs[k].numTiles ? maxTiles : hJobs[k].numTiles;
  if (maxTiles > 0) seg_launches<true>(st, hJobs[0], dJobs, V, maxTiles);
}

void launch_bin_sort(hipStream_t st, const BinArgs& a, BinOut& out) {
  const SortScratch& s = a.scratch;
  const int tb = tile_bits(a.numTiles);
  const int tpasses = (tb + 7) / 8;
  const int Gc = (int)((a.n + kBlock - 1) / kBlock);
  if (a.mode == 2 && a.numTiles <= kMaxSegTiles) {
    // tile segments: count table, column scan, budgeted scan, scatter, per-tile sort
    SegJob j{};
    j.n = a.n;
    j.cap = a.cap;
    j.nb = seg_blocks(a.n);
    j.tilesX = a.tilesX;
    j.tile = a.tile;
    j.numTiles = a.numTiles;
    j.recFmt = a.recFmt;
    j.rec = const_cast<float4*>(a.rec);
    j.depthKey = const_cast<uint32_t*>(a.depthKey);
    j.rect = const_cast<uint2*>(a.rect);
    j.cnt = const_cast<uint32_t*>(a.cnt);
    j.chunkKeys = s.chunkKeys;
    j.blockStart = s.blockStart;
    j.table = s.segTable;
    j.visPart = s.visPart;
    j.tileCnt = s.tileCnt;
    j.ranges = a.ranges;
    j.tileOrder = a.tileOrder;
    j.numKeys = a.numKeys;
    j.numVisible = a.numVisible;
    j.totalKeys = a.totalKeys;
    j.stats = a.stats;
    j.keys = (uint64_t*)a.tkA;
    j.tmp = (uint64_t*)a.tkB;
    j.ids = a.tvA;
    seg_launches<false>(st, j, nullptr, 1, j.numTiles);
    out.keys = nullptr;
    out.ids = a.tvA;
    out.keyBytes = 0;
    out.tileOrder = a.tileOrder;
    return;
  }
  if (a.mode == 0 || a.mode == 2) {   // mode 2 lands here above kMaxSegTiles tiles
    // 1) stable sort of the Gaussians by depth bits, values = ids
    const uint32_t* dk;
    const uint32_t* order;
    radix_sort<uint32_t>(st, s, a.depthKey, nullptr, a.dkA, a.dvA, a.dkB, a.dvB, a.dN, 0, 32, &dk, &order,
                         s.nDepth);
    // 2) per-Gaussian key counts in depth order -> offsets -> one key per tile
    k_count_partials<<<Gc, kBlock, 0, st>>>(order, a.cnt, a.rect, a.depthKey, s.nDepth, s.cntS, s.rectS, nullptr,
                                            s.partials, s.vpartials);
    k_count_top<<<1, 1024, 0, st>>>(s.partials, s.vpartials, Gc, a.cap, a.numKeys, a.totalKeys, a.numVisible, a.stats);
    k_emit<false><<<Gc, kBlock, 0, st>>>(order, s.cntS, s.rectS, nullptr, s.nDepth, s.partials, a.tilesX, a.tile, a.cap,
                                         a.rec, a.recFmt, a.tkA, a.tvA);
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
    k_emit<true><<<Gc, kBlock, 0, st>>>(nullptr, s.cntS, s.rectS, s.dkS, a.dN, s.partials, a.tilesX, a.tile, a.cap,
                                        a.rec, a.recFmt, a.tkA, a.tvA);
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
