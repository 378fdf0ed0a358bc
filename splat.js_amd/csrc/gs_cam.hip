// gs_cam.hip — camera pose, shared focal and exposure gradients (the
// reference's camAdd terms: shaders.js:614-616 in render, :834-858 in chain).
//
// Each term is converted to i32 per contribution exactly as the reference's
// camAdd does, i32(clamp(v * 64, -1e9, 1e9)) (shaders.js:158, :522-524,
// :717-719), and summed with wrapping adds. Integer sums commute, so the
// per-camera totals do not depend on the order of contributions: each block
// sums its threads' terms in registers and LDS and adds one atomic per slot,
// and the result is the same as one atomic per contribution.
//
// Gradient rows (8 i32 per camera, trainer.js:173): [dw x,y,z, dt x,y,z,
// d log gain, d bias]; the row after the last camera holds d log f in slot 0.
//
// Both kernels are grid-stride loops over a fixed grid, so the atomics per
// launch stay bounded (at most kCamBlocks per slot).
#include "gs_kernels.h"

namespace gs {

constexpr int kCamBlocks = 1024;

__device__ __forceinline__ uint32_t cam_fix(float v) {
  return (uint32_t)(int32_t)clampf(v * 64.0f, -1.0e9f, 1.0e9f);
}

__device__ __forceinline__ void cross3(const float* a, const float* b, float* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// Wrapping block sum of N per-thread values; slot k goes to *dst[k].
template <int N>
__device__ __forceinline__ void block_add(uint32_t (&v)[N], uint32_t* const (&dst)[N]) {
  __shared__ uint32_t part[kBlock / 64][N];
#pragma unroll
  for (int k = 0; k < N; k++)
    for (int o = 32; o > 0; o >>= 1) v[k] += (uint32_t)__shfl_xor((int)v[k], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < N; k++) part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x < N) {
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; q++) s += part[q][threadIdx.x];
    if (s != 0u) atomicAdd(dst[threadIdx.x], s);
  }
}

// Exposure terms of the loss (shaders.js:600-616), recomputed from the
// stored image exactly as the forward pass evaluates them: per valid pixel
// d/d(log gain) = dot(eg, C) * gain and d/d(bias) = eg.x + eg.y + eg.z,
// eg = err / sqrt(err^2 + 0.03^2), err = gain * C + bias - target.
__global__ __launch_bounds__(kBlock) void k_cam_exposure(Cam cam, const float4* __restrict__ img,
                                                         const uint32_t* __restrict__ tgt, uint32_t* row) {
  const int64_t npx = (int64_t)cam.W * cam.H;
  uint32_t acc[2] = {0u, 0u};
  for (int64_t pi = (int64_t)blockIdx.x * kBlock + threadIdx.x; pi < npx; pi += (int64_t)gridDim.x * kBlock) {
    const uint32_t pk = tgt[cam.tgtOffset + pi];
    if ((pk >> 24) == 0u) continue;
    const float4 o4 = img[pi];
    const float o[3] = {o4.x, o4.y, o4.z};
    float eg[3];
    for (int c = 0; c < 3; c++) {
      const float tc = (float)((pk >> (8 * c)) & 0xffu) / 255.0f;
      const float err = (cam.gain * o[c] + cam.bias) - tc;
      const float root = sqrtf(err * err + 0.0009f);
      eg[c] = err / root;
    }
    acc[0] += cam_fix(dot3(eg, o) * cam.gain);
    acc[1] += cam_fix((eg[0] + eg[1]) + eg[2]);
  }
  uint32_t* const dst[2] = {row + 6, row + 7};
  block_add<2>(acc, dst);
}

// Pose and focal terms of the chain (shaders.js:834-858) for p_c =
// exp(w^) R p + t: dL/dt = dpc, dL/dw = (p_c - t) x dpc + sum_k W_k x dW_k
// (W_k the k-th column of R, dW the 2D-Jacobian gradient through T = J W),
// dL/dlog f = f dL/df over the mean and Jacobian paths. kClear also zeroes
// the screen-space accumulators (an evaluation pass, which runs no update).
template <int kAccum, bool kClear>
__global__ __launch_bounds__(kBlock) void k_cam_pose(Cam cam, CamPoseBufs b) {
  uint32_t acc[7] = {0u, 0u, 0u, 0u, 0u, 0u, 0u};
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < b.n; i += (int64_t)gridDim.x * kBlock) {
    int4* ac = (int4*)b.accum + 3 * i;
    const int4 a0 = ac[0], a1 = ac[1];
    if (kClear) {
      const int4 z = make_int4(0, 0, 0, 0);
      ac[0] = z; ac[1] = z; ac[2] = z;
    }
    if (b.cnt[i] == 0u) continue;
    const int raw[6] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y};
    float gp[6];
#pragma unroll
    for (int k = 0; k < 6; k++) gp[k] = kAccum == 0 ? (float)raw[k] / kFixed : __int_as_float(raw[k]);
    float p[16];
    load_params(b.P, b.cap, i, p);
    Geom g;
    compute_geom(cam, p, g);
    if (!g.ok) continue;
    const float cden = 1.0f + conic_lmax(g.va, g.vb, g.vc);
    gp[2] = gp[2] * cden;
    gp[3] = gp[3] * cden;
    gp[4] = gp[4] * cden;
    const float f = cam.f;
    const float iz = 1.0f / g.pc[2];
    ChainScreen cs;
    chain_screen(cam, g, gp, iz, cs);

    const float rp[3] = {g.pc[0] - cam.t[0], g.pc[1] - cam.t[1], g.pc[2] - cam.t[2]};
    float dw[3], sum[3];
    cross3(rp, cs.dpc, dw);
    const float fiz = f * iz;
    const float bx = (((-f) * g.pc[0]) * iz) * iz;
    const float by = (((-f) * g.pc[1]) * iz) * iz;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float wc[3] = {cam.R[k], cam.R[3 + k], cam.R[6 + k]};
      const float dv[3] = {fiz * cs.dT0[k], fiz * cs.dT1[k], bx * cs.dT0[k] + by * cs.dT1[k]};
      float c[3];
      cross3(wc, dv, c);
      for (int a = 0; a < 3; a++) sum[a] = k == 0 ? c[a] : sum[a] + c[a];
    }
    const float dlogf = (((((gp[0] * f) * g.pc[0]) * iz + ((gp[1] * f) * g.pc[1]) * iz)
                         + (cs.dJ00 + cs.dJ11) * fiz) + cs.dJ02 * bx) + cs.dJ12 * by;
#pragma unroll
    for (int a = 0; a < 3; a++) {
      acc[a] += cam_fix(dw[a] + sum[a]);
      acc[3 + a] += cam_fix(cs.dpc[a]);
    }
    acc[6] += cam_fix(dlogf);
  }
  uint32_t* const dst[7] = {b.row, b.row + 1, b.row + 2, b.row + 3, b.row + 4, b.row + 5, b.focal};
  block_add<7>(acc, dst);
}

static unsigned cam_grid(int64_t items) {
  const int64_t g = (items + kBlock - 1) / kBlock;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, kCamBlocks));
}

void launch_cam_exposure(hipStream_t st, const Cam& cam, const float4* img, const uint32_t* tgt, uint32_t* row) {
  const int64_t npx = (int64_t)cam.W * cam.H;
  if (npx <= 0) return;
  k_cam_exposure<<<cam_grid(npx), kBlock, 0, st>>>(cam, img, tgt, row);
}

void launch_cam_pose(hipStream_t st, int accumMode, bool clear, const Cam& cam, const CamPoseBufs& b) {
  if (b.n <= 0) return;
  const dim3 grid(cam_grid(b.n));
  if (accumMode == 0) {
    if (clear) k_cam_pose<0, true><<<grid, kBlock, 0, st>>>(cam, b);
    else k_cam_pose<0, false><<<grid, kBlock, 0, st>>>(cam, b);
  } else {
    if (clear) k_cam_pose<1, true><<<grid, kBlock, 0, st>>>(cam, b);
    else k_cam_pose<1, false><<<grid, kBlock, 0, st>>>(cam, b);
  }
}

}  // namespace gs
