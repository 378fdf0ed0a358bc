// gs_refine.hip — device side of the MCMC-lite refine (trainer.js:666-794).
//
// The reference reads the whole param/moment/SH state back to the host
// (~0.5 KB per Gaussian at SH degree 2), edits it there and writes it all
// back. Here only the opacity logit and log-scales come back (16 B per
// Gaussian, k_refine_gather): the host replays the reference's sequential
// random choices on those, and k_refine_apply performs the planned spawns in
// place, one thread per spawn.
//
// The spawns of one plan never conflict: targets are distinct, and no target
// is a donor (targets are dead Gaussians or new slots past n, donors are
// live). A split rewrites only its donor's log-scales (with values the host
// already used for every later clone of that donor) and zeroes the donor's
// moments, fields no other spawn reads, and a donor splits at most once.
#include "gs_kernels.h"

namespace gs {

__global__ __launch_bounds__(kBlock) void k_refine_gather(int64_t n, int64_t cap, const float4* __restrict__ P,
                                                         float4* __restrict__ q) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float4 a = P[i], b = P[cap + i], d = P[3 * cap + i];
  q[i] = make_float4(d.y, a.w, b.x, b.y);   // slot 13, slots 3-5
}

template <int kDeg, int kPrec>
__global__ __launch_bounds__(kBlock) void k_refine_apply(int64_t cap, const RefineOp* __restrict__ ops,
                                                        int64_t count, float4* __restrict__ P, void* M, void* V,
                                                        void* SH, void* SHM, void* SHV) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= count) return;
  const RefineOp op = ops[k];
  const int64_t t = op.target, d = op.donor;
  float p[16];
  load_params(P, cap, d, p);
  float o[16];
  // params[bi + c] = params[bd + c] + gauss() * s * 0.7, in f64, stored as f32
  for (int c = 0; c < 3; c++) o[c] = (float)((double)p[c] + op.jit[c]);
  for (int c = 0; c < 3; c++) o[3 + c] = op.ls[c];
  for (int c = 6; c <= 12; c++) o[c] = p[c];
  o[13] = op.opa;
  o[14] = 0.0f;
  o[15] = 0.0f;
  for (int c = 0; c < 4; c++) P[c * cap + t] = make_float4(o[4 * c], o[4 * c + 1], o[4 * c + 2], o[4 * c + 3]);
  if (op.split) {
    P[d] = make_float4(p[0], p[1], p[2], op.ls[0]);
    P[cap + d] = make_float4(op.ls[1], op.ls[2], p[6], p[7]);
  }
  // fresh moments for the target (and for a split donor)
  const int64_t who[2] = {t, d};
  for (int r = 0; r < (op.split ? 2 : 1); r++) {
    const int64_t g = who[r];
    if constexpr (Prec<kPrec>::mom16) {
      const uint4 z = make_uint4(0u, 0u, 0u, 0u);
      ((uint4*)M)[g] = z; ((uint4*)M)[cap + g] = z;
      ((uint4*)V)[g] = z; ((uint4*)V)[cap + g] = z;
    } else {
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      for (int c = 0; c < 4; c++) { ((float4*)M)[c * cap + g] = z; ((float4*)V)[c * cap + g] = z; }
    }
  }
  if constexpr (kDeg > 0) {
    // the clone inherits the donor's view dependence, with fresh moments
    constexpr int Q8 = ShDims<kDeg>::Q8;
    constexpr int P4 = ShDims<kDeg>::P4;
    if constexpr (Prec<kPrec>::sh != 0) {
      for (int c = 0; c < Q8; c++) ((uint4*)SH)[c * cap + t] = ((const uint4*)SH)[c * cap + d];
    } else {
      for (int c = 0; c < P4; c++) ((float4*)SH)[c * cap + t] = ((const float4*)SH)[c * cap + d];
    }
    if constexpr (Prec<kPrec>::mom16) {
      const uint4 z = make_uint4(0u, 0u, 0u, 0u);
      for (int c = 0; c < Q8; c++) {
        ((uint4*)SHM)[c * cap + t] = z;
        ((uint4*)SHV)[c * cap + t] = z;
        if (op.split) { ((uint4*)SHM)[c * cap + d] = z; ((uint4*)SHV)[c * cap + d] = z; }
      }
    } else {
      const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      for (int c = 0; c < P4; c++) {
        ((float4*)SHM)[c * cap + t] = z;
        ((float4*)SHV)[c * cap + t] = z;
        if (op.split) { ((float4*)SHM)[c * cap + d] = z; ((float4*)SHV)[c * cap + d] = z; }
      }
    }
  }
}

void launch_refine_gather(hipStream_t st, int64_t n, int64_t cap, const float4* P, float4* q) {
  if (n <= 0) return;
  k_refine_gather<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, st>>>(n, cap, P, q);
}

template <int kPrec>
static void apply_prec(hipStream_t st, int deg, int64_t cap, const RefineOp* ops, int64_t count, float4* P,
                       void* M, void* V, void* SH, void* SHM, void* SHV) {
  const dim3 grid((unsigned)((count + kBlock - 1) / kBlock));
  switch (deg) {
    case 0: k_refine_apply<0, kPrec><<<grid, kBlock, 0, st>>>(cap, ops, count, P, M, V, SH, SHM, SHV); break;
    case 1: k_refine_apply<1, kPrec><<<grid, kBlock, 0, st>>>(cap, ops, count, P, M, V, SH, SHM, SHV); break;
    case 2: k_refine_apply<2, kPrec><<<grid, kBlock, 0, st>>>(cap, ops, count, P, M, V, SH, SHM, SHV); break;
    default: k_refine_apply<3, kPrec><<<grid, kBlock, 0, st>>>(cap, ops, count, P, M, V, SH, SHM, SHV); break;
  }
}

void launch_refine_apply(hipStream_t st, int deg, int prec, int64_t cap, const RefineOp* ops, int64_t count,
                         float4* P, void* M, void* V, void* SH, void* SHM, void* SHV) {
  if (count <= 0) return;
  // SH and moment storage: modes 2 and 3 move the same planes
  if (prec == 0) apply_prec<0>(st, deg, cap, ops, count, P, M, V, SH, SHM, SHV);
  else if (prec == 1) apply_prec<1>(st, deg, cap, ops, count, P, M, V, SH, SHM, SHV);
  else apply_prec<2>(st, deg, cap, ops, count, P, M, V, SH, SHM, SHV);
}

}  // namespace gs
