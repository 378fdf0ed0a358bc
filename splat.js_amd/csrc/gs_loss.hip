// gs_loss.hip — L1 + lambda * D-SSIM loss (textbook; the reference trains
// with Charbonnier only, shaders.js:11-12).
//
//   L = sum over valid pixels p and channels c of
//       (1 - lambda) |x - y| + lambda (1 - ssim_c(p))
//   x = gain * C + bias (rendered, exposure-adjusted), y = target
//   ssim (Wang et al. 2004): 11x11 Gaussian window, sigma 1.5, C1 = 0.01^2,
//   C2 = 0.03^2, zero padding outside the image. An invalid target pixel
//   (undistortion border, alpha byte 0) enters every window as zero in both
//   images, like the padding, so the loss does not depend on what is
//   rendered there and its dL/dC is zero, as its valid flag (gC.w = 0) tells
//   the backward pass.
//
// Two LDS-tiled passes over 16x16 pixel tiles with a 5-pixel halo:
//   k_ssim_fwd  blurs x, y, x^2, y^2, xy (separable: 26 rows x 16 columns
//               horizontally, then 16 x 16 vertically), forms ssim and its
//               partials A = dS/dmu_x - 2 mu_x dS/dvar_x - mu_y dS/dcov,
//               B = dS/dvar_x, Cm = dS/dcov (zero at invalid pixels), and
//               accumulates the loss.
//   k_ssim_bwd  blurs A, B, Cm back (the window is symmetric) and forms
//               dL/dx = (1 - lambda) sign(x - y) - lambda (A~ + 2 x B~ + y Cm~),
//               dL/dC = gain * dL/dx.
// Both are HBM-bound (36 B/px in, 36 B/px out for the first; 36 + 28 in,
// 16 out for the second) with the halo re-reads served by L2.
#include <cmath>

#include "gs_kernels.h"

namespace gs {

// The loss kernels' own image tiling (independent of the raster's tiles):
// 16 x 16 pixels per block, with a 5-pixel halo for the 11-tap windows.
constexpr int kTile = 16;
constexpr int kHalo = 5;
constexpr int kReg = kTile + 2 * kHalo;   // 26

__device__ __forceinline__ int ssim_tiles_x(const Cam& cam) { return (cam.W + kTile - 1) / kTile; }

struct SsimWin {
  float g[11];
};

__device__ __forceinline__ void load_xy(const Cam& cam, const RasterBufs& b, int gx, int gy, float* x,
                                        float* y, bool& valid) {
  valid = false;
  if (gx < 0 || gy < 0 || gx >= cam.W || gy >= cam.H) {
    for (int c = 0; c < 3; c++) { x[c] = 0.0f; y[c] = 0.0f; }
    return;
  }
  const int64_t pi = (int64_t)gy * cam.W + gx;
  const uint32_t pk = b.tgt[cam.tgtOffset + pi];
  valid = (pk >> 24) != 0u;
  if (!valid) {
    for (int c = 0; c < 3; c++) { x[c] = 0.0f; y[c] = 0.0f; }
    return;
  }
  const float4 o = b.img[pi];
  const float oc[3] = {o.x, o.y, o.z};
  for (int c = 0; c < 3; c++) {
    x[c] = cam.gain * oc[c] + cam.bias;
    y[c] = (float)((pk >> (8 * c)) & 0xffu) / 255.0f;
  }
}

__global__ __launch_bounds__(kBlock) void k_ssim_fwd(Cam cam, SsimWin win, float lambda, RasterBufs b,
                                                     float* __restrict__ part) {
  __shared__ float sX[3][kReg][kReg];
  __shared__ float sY[3][kReg][kReg];
  __shared__ float sH[15][kReg][kTile];   // horizontal blurs of x, y, xx, yy, xy per channel
  __shared__ double sStat[3][kBlock / 64];
  const int tid = threadIdx.x;
  const int tile = blockIdx.x;
  const int tx = tile % ssim_tiles_x(cam), ty = tile / ssim_tiles_x(cam);
  const int ox = tx * kTile - kHalo, oy = ty * kTile - kHalo;
  for (int k = tid; k < kReg * kReg; k += kBlock) {
    const int r = k / kReg, c = k % kReg;
    float x[3], y[3];
    bool v;
    load_xy(cam, b, ox + c, oy + r, x, y, v);
    for (int ch = 0; ch < 3; ch++) { sX[ch][r][c] = x[ch]; sY[ch][r][c] = y[ch]; }
  }
  __syncthreads();
  for (int k = tid; k < kReg * kTile; k += kBlock) {
    const int r = k / kTile, c = k % kTile;
    for (int ch = 0; ch < 3; ch++) {
      float hx = 0.f, hy = 0.f, hxx = 0.f, hyy = 0.f, hxy = 0.f;
#pragma unroll
      for (int q = 0; q < 11; q++) {
        const float xv = sX[ch][r][c + q], yv = sY[ch][r][c + q], w = win.g[q];
        hx = fmaf(w, xv, hx);
        hy = fmaf(w, yv, hy);
        hxx = fmaf(w, xv * xv, hxx);
        hyy = fmaf(w, yv * yv, hyy);
        hxy = fmaf(w, xv * yv, hxy);
      }
      sH[5 * ch + 0][r][c] = hx; sH[5 * ch + 1][r][c] = hy; sH[5 * ch + 2][r][c] = hxx;
      sH[5 * ch + 3][r][c] = hyy; sH[5 * ch + 4][r][c] = hxy;
    }
  }
  __syncthreads();
  const int lx = tid & 15, ly = tid >> 4;
  const int x = tx * kTile + lx, y = ty * kTile + ly;
  double st0 = 0.0, st1 = 0.0, st2 = 0.0;
  if (x < cam.W && y < cam.H) {
    const int64_t pi = (int64_t)y * cam.W + x;
    const int64_t npx = (int64_t)cam.W * cam.H;
    const bool valid = (b.tgt[cam.tgtOffset + pi] >> 24) != 0u;
    const float C1 = 0.0001f, C2 = 0.0009f;
    for (int ch = 0; ch < 3; ch++) {
      float m[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int q = 0; q < 11; q++) {
        const float w = win.g[q];
        for (int s = 0; s < 5; s++) m[s] = fmaf(w, sH[5 * ch + s][ly + q][lx], m[s]);
      }
      const float ux = m[0], uy = m[1];
      const float vx = m[2] - ux * ux, vy = m[3] - uy * uy, cxy = m[4] - ux * uy;
      const float N1 = 2.f * ux * uy + C1, N2 = 2.f * cxy + C2;
      const float D1 = ux * ux + uy * uy + C1, D2 = vx + vy + C2;
      const float S = (N1 * N2) / (D1 * D2);
      float A = 0.f, Bv = 0.f, Cv = 0.f;
      if (valid) {
        const float xv = sX[ch][ly + kHalo][lx + kHalo], yv = sY[ch][ly + kHalo][lx + kHalo];
        const float err = xv - yv;
        st0 += (double)(err * err);
        st1 += (double)((1.f - lambda) * fabsf(err) + lambda * (1.f - S));
        const float dmx = (2.f * uy * N2) / (D1 * D2) - S * (2.f * ux) / D1;
        const float dvx = -S / D2;
        const float dcxy = 2.f * N1 / (D1 * D2);
        A = dmx - 2.f * ux * dvx - uy * dcxy;
        Bv = dvx;
        Cv = dcxy;
      }
      part[(int64_t)(3 * ch + 0) * npx + pi] = A;
      part[(int64_t)(3 * ch + 1) * npx + pi] = Bv;
      part[(int64_t)(3 * ch + 2) * npx + pi] = Cv;
    }
    if (valid) st2 = 1.0;
  }
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

__global__ __launch_bounds__(kBlock) void k_ssim_bwd(Cam cam, SsimWin win, float lambda, RasterBufs b,
                                                     const float* __restrict__ part) {
  __shared__ float sP[9][kReg][kReg];
  __shared__ float sH[9][kReg][kTile];
  const int tid = threadIdx.x;
  const int tile = blockIdx.x;
  const int tx = tile % ssim_tiles_x(cam), ty = tile / ssim_tiles_x(cam);
  const int ox = tx * kTile - kHalo, oy = ty * kTile - kHalo;
  const int64_t npx = (int64_t)cam.W * cam.H;
  for (int k = tid; k < kReg * kReg; k += kBlock) {
    const int r = k / kReg, c = k % kReg;
    const int gx = ox + c, gy = oy + r;
    const bool in = gx >= 0 && gy >= 0 && gx < cam.W && gy < cam.H;
    const int64_t pi = (int64_t)gy * cam.W + gx;
    for (int s = 0; s < 9; s++) sP[s][r][c] = in ? part[(int64_t)s * npx + pi] : 0.0f;
  }
  __syncthreads();
  for (int k = tid; k < kReg * kTile; k += kBlock) {
    const int r = k / kTile, c = k % kTile;
    for (int s = 0; s < 9; s++) {
      float h = 0.f;
#pragma unroll
      for (int q = 0; q < 11; q++) h = fmaf(win.g[q], sP[s][r][c + q], h);
      sH[s][r][c] = h;
    }
  }
  __syncthreads();
  const int lx = tid & 15, ly = tid >> 4;
  const int x = tx * kTile + lx, y = ty * kTile + ly;
  if (x >= cam.W || y >= cam.H) return;
  const int64_t pi = (int64_t)y * cam.W + x;
  float xs[3], ys[3];
  bool valid;
  load_xy(cam, b, x, y, xs, ys, valid);
  float g[3] = {0.f, 0.f, 0.f};
  for (int ch = 0; ch < 3 && valid; ch++) {
    float bl[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < 11; q++)
      for (int s = 0; s < 3; s++) bl[s] = fmaf(win.g[q], sH[3 * ch + s][ly + q][lx], bl[s]);
    const float dssim = bl[0] + 2.f * xs[ch] * bl[1] + ys[ch] * bl[2];
    const float err = xs[ch] - ys[ch];
    const float d = -lambda * dssim + (1.f - lambda) * (err > 0.f ? 1.f : (err < 0.f ? -1.f : 0.f));
    g[ch] = cam.gain * d;
  }
  b.gC[pi] = make_float4(g[0], g[1], g[2], valid ? 1.0f : 0.0f);
}

void launch_loss_l1ssim(hipStream_t st, const Cam& cam, float lambda, const RasterBufs& b, float* scratch) {
  SsimWin w;
  double g[11], s = 0.0;
  for (int i = 0; i < 11; i++) { g[i] = std::exp(-((i - 5) * (i - 5)) / (2.0 * 1.5 * 1.5)); s += g[i]; }
  for (int i = 0; i < 11; i++) w.g[i] = (float)(g[i] / s);
  const dim3 grid((unsigned)(((cam.W + kTile - 1) / kTile) * ((cam.H + kTile - 1) / kTile)));
  k_ssim_fwd<<<grid, kBlock, 0, st>>>(cam, w, lambda, b, scratch);
  k_ssim_bwd<<<grid, kBlock, 0, st>>>(cam, w, lambda, b, scratch);
}

}  // namespace gs
