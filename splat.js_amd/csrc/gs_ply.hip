// gs_ply.hip — rows of the standard 3DGS PLY export (src/io/ply.js), one
// thread per Gaussian, written straight into the file body in device memory.
//
// The reference computes in JS numbers (f64) over Float32Array inputs and
// stores f32: the same here. Optional bake (bakeOpacityCompensation,
// ply.js:12-37): the Mip-Splatting compensation of the splat's mean 2D size
// at its nearest training camera, multiplied into the opacity.
#include "gs_kernels.h"

namespace gs {

// Rows first .. first + count - 1; row i goes to out + (i - first) * W.
template <int kDeg, int kPrec>
__global__ __launch_bounds__(kBlock) void k_ply_rows(int64_t first, int64_t count, int64_t cap,
                                                    const float4* __restrict__ P,
                                                    const void* __restrict__ SH, int bake, double f,
                                                    const float* __restrict__ camPos, int nc,
                                                    float* __restrict__ out) {
  constexpr int NC = ShDims<kDeg>::C;
  constexpr int W = 17 + NC;   // floats per row
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r >= count) return;
  const int64_t i = first + r;
  float p[16];
  load_params(P, cap, i, p);
  float opaLogit = p[13];
  if (bake) {
    double z2min = INFINITY;
    for (int c = 0; c < nc; c++) {
      const double dx = (double)p[0] - (double)camPos[3 * c];
      const double dy = (double)p[1] - (double)camPos[3 * c + 1];
      const double dz = (double)p[2] - (double)camPos[3 * c + 2];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < z2min) z2min = d2;
    }
    const double z = fmax(1e-3, sqrt(z2min));
    const double sMean = exp(((double)p[3] + (double)p[4] + (double)p[5]) / 3.0);
    const double s2d = f * sMean / z;
    const double comp = (s2d * s2d) / (s2d * s2d + 0.3);
    const double opa = comp / (1.0 + exp(-(double)p[13]));
    const double cl = fmin(1.0 - 1e-6, fmax(1e-6, opa));
    opaLogit = (float)log(cl / (1.0 - cl));
  }
  float* o = out + (int64_t)W * r;
  o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
  o[3] = 0.0f; o[4] = 0.0f; o[5] = 0.0f;
  const double C0 = 0.28209479177387814;
  for (int c = 0; c < 3; c++) o[6 + c] = (float)((1.0 / (1.0 + exp(-(double)p[10 + c])) - 0.5) / C0);
  if constexpr (kDeg > 0) {
    float sh[ShDims<kDeg>::S];
    load_sh_any<kDeg, Prec<kPrec>::sh>(SH, cap, i, sh);
    for (int k = 0; k < NC; k++) o[9 + k] = sh[k];
  }
  float* q = o + 9 + NC;
  q[0] = opaLogit;
  q[1] = p[3]; q[2] = p[4]; q[3] = p[5];
  double qw = p[6], qx = p[7], qy = p[8], qz = p[9];
  const double ql = sqrt(qw * qw + qx * qx + qy * qy + qz * qz);   // Math.hypot
  if (ql < 1e-6) { qw = 1.0; qx = qy = qz = 0.0; }
  else { qw /= ql; qx /= ql; qy /= ql; qz /= ql; }
  q[4] = (float)qw; q[5] = (float)qx; q[6] = (float)qy; q[7] = (float)qz;
}

template <int kPrec>
static void ply_prec(hipStream_t st, int deg, int64_t first, int64_t count, int64_t cap, const float4* P,
                     const void* SH, int bake, double f, const float* camPos, int nc, float* out) {
  const dim3 grid((unsigned)((count + kBlock - 1) / kBlock));
  switch (deg) {
    case 0: k_ply_rows<0, kPrec><<<grid, kBlock, 0, st>>>(first, count, cap, P, SH, bake, f, camPos, nc, out); break;
    case 1: k_ply_rows<1, kPrec><<<grid, kBlock, 0, st>>>(first, count, cap, P, SH, bake, f, camPos, nc, out); break;
    case 2: k_ply_rows<2, kPrec><<<grid, kBlock, 0, st>>>(first, count, cap, P, SH, bake, f, camPos, nc, out); break;
    default: k_ply_rows<3, kPrec><<<grid, kBlock, 0, st>>>(first, count, cap, P, SH, bake, f, camPos, nc, out); break;
  }
}

void launch_ply_rows(hipStream_t st, int deg, int prec, int64_t first, int64_t count, int64_t cap,
                     const float4* P, const void* SH, int bake, double f, const float* camPos, int nc,
                     float* out) {
  if (count <= 0) return;
  // only the SH storage matters here: f32, bf16 (modes 1 and 2) or fp16
  if (prec == 0) ply_prec<0>(st, deg, first, count, cap, P, SH, bake, f, camPos, nc, out);
  else if (prec == 3) ply_prec<3>(st, deg, first, count, cap, P, SH, bake, f, camPos, nc, out);
  else ply_prec<1>(st, deg, first, count, cap, P, SH, bake, f, camPos, nc, out);
}

}  // namespace gs
