// gs_device.h — types and per-Gaussian math shared by the trainer kernels.
//
// Every expression here is written in the association the reference's WGSL
// spells out (a*b*c = (a*b)*c, dot = (x + y) + z), and all HIP sources are
// compiled with -ffp-contract=off, so single-precision results are
// reproducible on the host bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gs_math.h"
#include "gs_cover.h"

namespace gs {

// Raster tile: kDefaultTile x kDefaultTile pixels (the reference's TILE,
// shaders.js:32, and its 16 x 16 raster workgroups, :544) and kDefaultBatch
// keys staged in LDS per round (the reference reads each splat from global
// memory); a trainer may choose tiles of 8, 16 or 32 and batches of 64, 128
// or 256 instead.
constexpr int kDefaultTile = 16;
constexpr int kDefaultBatch = 256;
constexpr int kBlock = 256;
constexpr float kFixed = 16384.0f;   // shaders.js:35-39 fixed-point scale

// Per-view constants (the reference's CAM_STRUCT uniform, shaders.js:140-151).
struct Cam {
  float R[9];
  float t[3];
  float nearp;
  float f, cx, cy;
  float gain, bias;
  float bg[3];
  float camPos[3];
  int W, H, tilesX, tilesY, numTiles;
  int tile;                 // tile side in pixels (8, 16 or 32)
  int shActiveK;            // active SH rest coefficients (<= K)
  int64_t tgtOffset;        // first target pixel of this view
};

struct Cut {
  float eCut, aMin, radClamp;
};

// Adam hyperparameters for one step (the reference's AdamU / SHA uniforms,
// shaders.js:952-1036, filled as trainer.js:315-342, :417-429 do).
struct AdamHP {
  float lr[16];
  float b1, b2, eps;
  float rbc1, rbc2;   // 1 / (1 - b1^t), 1 / (1 - b2^t): bias corrections as multipliers
  float minLS, maxLS, maxLogit, reg;
  float shLr;
  float anisoReg;
  uint32_t step;   // Adam step t: seeds the stochastic rounding of bf16 stores
};

// Parameters live in four float4 planes (plane c = slots 4c..4c+3 of every
// Gaussian), so every per-Gaussian load is a fully coalesced 16-byte access.
__device__ __forceinline__ void load_params(const float4* __restrict__ P, int64_t cap, int64_t i,
                                            float* p) {
  const float4 a = P[i], b = P[cap + i], c = P[2 * cap + i], d = P[3 * cap + i];
  p[0] = a.x; p[1] = a.y; p[2] = a.z; p[3] = a.w;
  p[4] = b.x; p[5] = b.y; p[6] = b.z; p[7] = b.w;
  p[8] = c.x; p[9] = c.y; p[10] = c.z; p[11] = c.w;
  p[12] = d.x; p[13] = d.y; p[14] = d.z; p[15] = d.w;
}

// A pointer field that a kernel reads from device memory (a batched step's
// per-view block or binning jobs), loaded as a pointer to global memory.
// Loaded as a plain pointer, the compiler cannot tell global memory from LDS
// or scratch and emits flat loads and stores, which count against the LDS
// and scalar-load wait counter as well: every LDS or scalar-load wait then
// also waits for the global loads in flight. Every such field holds device
// (or pinned host) memory.
template <typename T>
__device__ __forceinline__ T* gs_global(T* const& field) {
  const void* at = &field;
  return (T*)*static_cast<__attribute__((address_space(1))) T* const*>(at);
}

// Wave vote on a lane predicate. The builtin takes the bool itself; HIP's
// __ballot(int) converts it to an int and back, which costs two VALU
// instructions per vote.
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

__device__ __forceinline__ float clampf(float x, float lo, float hi) {
  return fminf(fmaxf(x, lo), hi);
}

__device__ __forceinline__ float dot3(const float* a, const float* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// The camera-independent part of a Gaussian's geometry: unit quaternion,
// rotation rows, scales and the 3D covariance Sigma = (R S)(R S)^T.
struct GeomShape {
  float q[4], r0[3], r1[3], r2[3], s[3];
  float s00, s01, s02, s11, s12, s22;
};

struct Geom : GeomShape {
  bool ok;
  float pc[3], t0[3], t1[3];
  float va, vb, vc;
};

// shaders.js:183-201 (the part of computeGeom no camera enters)
__device__ __forceinline__ void compute_shape(const float* P, GeomShape& g) {
  float q[4] = {P[6], P[7], P[8], P[9]};
  const float ql = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  if (ql < 1e-6f) { q[0] = 1.0f; q[1] = 0.0f; q[2] = 0.0f; q[3] = 0.0f; }
  else { q[0] = q[0] / ql; q[1] = q[1] / ql; q[2] = q[2] / ql; q[3] = q[3] / ql; }
  g.q[0] = q[0]; g.q[1] = q[1]; g.q[2] = q[2]; g.q[3] = q[3];
  const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  g.r0[0] = 1.0f - 2.0f * (qy * qy + qz * qz);
  g.r0[1] = 2.0f * (qx * qy - qw * qz);
  g.r0[2] = 2.0f * (qx * qz + qw * qy);
  g.r1[0] = 2.0f * (qx * qy + qw * qz);
  g.r1[1] = 1.0f - 2.0f * (qx * qx + qz * qz);
  g.r1[2] = 2.0f * (qy * qz - qw * qx);
  g.r2[0] = 2.0f * (qx * qz - qw * qy);
  g.r2[1] = 2.0f * (qy * qz + qw * qx);
  g.r2[2] = 1.0f - 2.0f * (qx * qx + qy * qy);

  for (int k = 0; k < 3; k++) g.s[k] = gs_expf_mid(clampf(P[3 + k], -12.0f, 6.0f));

  float m0[3], m1[3], m2[3];
  for (int k = 0; k < 3; k++) {
    m0[k] = g.r0[k] * g.s[k];
    m1[k] = g.r1[k] * g.s[k];
    m2[k] = g.r2[k] * g.s[k];
  }
  g.s00 = dot3(m0, m0); g.s01 = dot3(m0, m1); g.s02 = dot3(m0, m2);
  g.s11 = dot3(m1, m1); g.s12 = dot3(m1, m2); g.s22 = dot3(m2, m2);
}

// camera-space position (shaders.js:179-181); false when behind the near plane
__device__ __forceinline__ bool geom_position(const Cam& cam, const float* P, Geom& g) {
  g.ok = false;
  const float p[3] = {P[0], P[1], P[2]};
  g.pc[0] = dot3(cam.R + 0, p) + cam.t[0];
  g.pc[1] = dot3(cam.R + 3, p) + cam.t[1];
  g.pc[2] = dot3(cam.R + 6, p) + cam.t[2];
  return !(g.pc[2] < cam.nearp);
}

// projection Jacobian rows T = J W and the 2D covariance V = T Sigma T^T
// (shaders.js:203-222), from g.pc and the shape already in g
__device__ __forceinline__ void geom_project(const Cam& cam, Geom& g) {
  const float f = cam.f;
  const float iz = 1.0f / g.pc[2];
  const float a = f * iz;
  const float bx = ((-f) * g.pc[0] * iz) * iz;
  const float by = ((-f) * g.pc[1] * iz) * iz;
  for (int k = 0; k < 3; k++) {
    g.t0[k] = a * cam.R[0 + k] + bx * cam.R[6 + k];
    g.t1[k] = a * cam.R[3 + k] + by * cam.R[6 + k];
  }
  float st0[3], st1[3];
  st0[0] = (g.s00 * g.t0[0] + g.s01 * g.t0[1]) + g.s02 * g.t0[2];
  st0[1] = (g.s01 * g.t0[0] + g.s11 * g.t0[1]) + g.s12 * g.t0[2];
  st0[2] = (g.s02 * g.t0[0] + g.s12 * g.t0[1]) + g.s22 * g.t0[2];
  st1[0] = (g.s00 * g.t1[0] + g.s01 * g.t1[1]) + g.s02 * g.t1[2];
  st1[1] = (g.s01 * g.t1[0] + g.s11 * g.t1[1]) + g.s12 * g.t1[2];
  st1[2] = (g.s02 * g.t1[0] + g.s12 * g.t1[1]) + g.s22 * g.t1[2];
  g.va = fmaxf(dot3(g.t0, st0), 0.0f);
  g.vb = dot3(g.t1, st0);
  g.vc = fmaxf(dot3(g.t1, st1), 0.0f);
  g.ok = true;
}

// shaders.js:176-224 computeGeom
__device__ __forceinline__ void compute_geom(const Cam& cam, const float* P, Geom& g) {
  if (!geom_position(cam, P, g)) return;
  compute_shape(P, g);
  geom_project(cam, g);
}

// the same from a shape computed once for several cameras (identical values)
__device__ __forceinline__ void compute_geom(const Cam& cam, const float* P, const GeomShape& sp, Geom& g) {
  if (!geom_position(cam, P, g)) return;
  static_cast<GeomShape&>(g) = sp;
  geom_project(cam, g);
}

// shaders.js:674-678 (render) / :736-740 (chain): largest eigenvalue of the
// dilated 2D covariance; conic gradients are accumulated times 1/(1+lmax).
__device__ __forceinline__ float conic_lmax(float va, float vb, float vc) {
  const float cva = va + 0.3f, cvc = vc + 0.3f;
  const float cmid = 0.5f * (cva + cvc);
  return cmid + sqrtf(fmaxf(cmid * cmid - (cva * cvc - vb * vb), 0.0f));
}

// Screen-space gradients to the 2D-covariance, Jacobian and camera-space
// position gradients (shaders.js:759-827). gp[2..4] already carry the
// 1 + lmax factor. Shared by the update pass and the camera-gradient pass.
struct ChainScreen {
  float dS00, dS01, dS02, dS11, dS12, dS22;   // dL/dSigma (symmetric)
  float dT0[3], dT1[3];                       // dL/dT rows
  float dJ00, dJ02, dJ11, dJ12;               // dL/dJ
  float dpc[3];                               // dL/d(camera-space position)
};

__device__ __forceinline__ void chain_screen(const Cam& cam, const Geom& g, const float* gp, float iz,
                                             ChainScreen& c) {
  const float f = cam.f;
  // conic -> dilated 2D covariance (:759-777)
  const float ad = g.va + 0.3f, cd = g.vc + 0.3f;
  const float detVd = ad * cd - g.vb * g.vb;
  const float inv = 1.0f / detVd;
  const float kA = cd * inv, kB = (-g.vb) * inv, kC = ad * inv;
  const float g11 = gp[2], g12 = 0.5f * gp[3], g22 = gp[4];
  const float m100 = g11 * kA + g12 * kB, m101 = g11 * kB + g12 * kC;
  const float m110 = g12 * kA + g22 * kB, m111 = g12 * kB + g22 * kC;
  const float p00 = kA * m100 + kB * m110, p01 = kA * m101 + kB * m111;
  const float p10 = kB * m100 + kC * m110, p11 = kB * m101 + kC * m111;
  float gva = -p00, gvb = -(p01 + p10), gvc = -p11;

  // comp -> raw 2D covariance (:780-788)
  const float detV = fmaxf(g.va * g.vc - g.vb * g.vb, 0.0f);
  const float comp = sqrtf(fmaxf(detV / detVd, 0.0f));
  if (comp > 1e-4f) {
    const float gcomp = gp[5];
    const float denom = 2.0f * comp * detVd;
    gva = gva + gcomp * (g.vc - comp * comp * cd) / denom;
    gvc = gvc + gcomp * (g.va - comp * comp * ad) / denom;
    gvb = gvb + gcomp * (2.0f * g.vb * (comp * comp - 1.0f)) / denom;
  }

  // V = T Sigma T^T backward (:791-810)
  const float h11 = gva, h12 = 0.5f * gvb, h22 = gvc;
  const float* t0 = g.t0;
  const float* t1 = g.t1;
  float u0[3], u1[3];
  for (int k = 0; k < 3; k++) { u0[k] = h11 * t0[k] + h12 * t1[k]; u1[k] = h12 * t0[k] + h22 * t1[k]; }
  c.dS00 = t0[0] * u0[0] + t1[0] * u1[0];
  c.dS01 = 0.5f * ((t0[0] * u0[1] + t1[0] * u1[1]) + (t0[1] * u0[0] + t1[1] * u1[0]));
  c.dS02 = 0.5f * ((t0[0] * u0[2] + t1[0] * u1[2]) + (t0[2] * u0[0] + t1[2] * u1[0]));
  c.dS11 = t0[1] * u0[1] + t1[1] * u1[1];
  c.dS12 = 0.5f * ((t0[1] * u0[2] + t1[1] * u1[2]) + (t0[2] * u0[1] + t1[2] * u1[1]));
  c.dS22 = t0[2] * u0[2] + t1[2] * u1[2];
  float sT0[3], sT1[3];
  sT0[0] = (g.s00 * t0[0] + g.s01 * t0[1]) + g.s02 * t0[2];
  sT0[1] = (g.s01 * t0[0] + g.s11 * t0[1]) + g.s12 * t0[2];
  sT0[2] = (g.s02 * t0[0] + g.s12 * t0[1]) + g.s22 * t0[2];
  sT1[0] = (g.s00 * t1[0] + g.s01 * t1[1]) + g.s02 * t1[2];
  sT1[1] = (g.s01 * t1[0] + g.s11 * t1[1]) + g.s12 * t1[2];
  sT1[2] = (g.s02 * t1[0] + g.s12 * t1[1]) + g.s22 * t1[2];
  for (int k = 0; k < 3; k++) {
    c.dT0[k] = 2.0f * (h11 * sT0[k] + h12 * sT1[k]);
    c.dT1[k] = 2.0f * (h12 * sT0[k] + h22 * sT1[k]);
  }

  // T = J W (:813-822) and the mean path (:825-827)
  c.dJ00 = dot3(c.dT0, cam.R + 0);
  c.dJ02 = dot3(c.dT0, cam.R + 6);
  c.dJ11 = dot3(c.dT1, cam.R + 3);
  c.dJ12 = dot3(c.dT1, cam.R + 6);
  float* dpc = c.dpc;
  dpc[0] = 0.0f; dpc[1] = 0.0f; dpc[2] = 0.0f;
  dpc[0] = dpc[0] + c.dJ02 * ((-f) * iz * iz);
  dpc[1] = dpc[1] + c.dJ12 * ((-f) * iz * iz);
  dpc[2] = dpc[2] + ((c.dJ00 + c.dJ11) * ((-f) * iz * iz)
         + c.dJ02 * (2.0f * f * g.pc[0] * iz * iz * iz))
         + c.dJ12 * (2.0f * f * g.pc[1] * iz * iz * iz);
  dpc[0] = dpc[0] + gp[0] * f * iz;
  dpc[1] = dpc[1] + gp[1] * f * iz;
  dpc[2] = dpc[2] + (-f) * (gp[0] * g.pc[0] + gp[1] * g.pc[1]) * iz * iz;
}

// shaders.js:56-111 real SH basis (degrees 1..3) and dY/dv; D is K x 3.
// Y[14] follows shaders.js:102 (x * (xx - yy)).
template <int kDeg, bool kGrad>
__device__ __forceinline__ void sh_basis(const float* v, float* Y, float* D) {
  const float x = v[0], y = v[1], z = v[2];
  const float xx = x * x, yy = y * y, zz = z * z;
  const float C1 = 0.4886025119029199f;
  const float C20 = 1.0925484305920792f, C21 = -1.0925484305920792f, C22 = 0.31539156525252005f,
              C23 = -1.0925484305920792f, C24 = 0.5462742152960396f;
  const float C30 = -0.5900435899266435f, C31 = 2.890611442640554f, C32 = -0.4570457994644658f,
              C33 = 0.3731763325901154f, C34 = -0.4570457994644658f, C35 = 1.445305721320277f,
              C36 = -0.5900435899266435f;
  Y[0] = (-C1) * y; Y[1] = C1 * z; Y[2] = (-C1) * x;
  if (kGrad) {
    D[0] = 0.0f; D[1] = -C1; D[2] = 0.0f;
    D[3] = 0.0f; D[4] = 0.0f; D[5] = C1;
    D[6] = -C1; D[7] = 0.0f; D[8] = 0.0f;
  }
  if (kDeg >= 2) {
    Y[3] = C20 * x * y;
    Y[4] = C21 * y * z;
    Y[5] = C22 * ((2.0f * zz - xx) - yy);
    Y[6] = C23 * x * z;
    Y[7] = C24 * (xx - yy);
    if (kGrad) {
      D[9] = C20 * y; D[10] = C20 * x; D[11] = 0.0f;
      D[12] = 0.0f; D[13] = C21 * z; D[14] = C21 * y;
      D[15] = C22 * (-2.0f * x); D[16] = C22 * (-2.0f * y); D[17] = C22 * (4.0f * z);
      D[18] = C23 * z; D[19] = 0.0f; D[20] = C23 * x;
      D[21] = C24 * (2.0f * x); D[22] = C24 * (-2.0f * y); D[23] = 0.0f;
    }
  }
  if (kDeg >= 3) {
    Y[8] = C30 * y * (3.0f * xx - yy);
    Y[9] = C31 * x * y * z;
    Y[10] = C32 * y * ((4.0f * zz - xx) - yy);
    Y[11] = C33 * z * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
    Y[12] = C34 * x * ((4.0f * zz - xx) - yy);
    Y[13] = C35 * z * (xx - yy);
    Y[14] = C36 * x * (xx - yy);
    if (kGrad) {
      D[24] = C30 * (6.0f * x * y); D[25] = C30 * (3.0f * xx - 3.0f * yy); D[26] = 0.0f;
      D[27] = C31 * (y * z); D[28] = C31 * (x * z); D[29] = C31 * (x * y);
      D[30] = C32 * (-2.0f * x * y); D[31] = C32 * ((4.0f * zz - xx) - 3.0f * yy); D[32] = C32 * (8.0f * y * z);
      D[33] = C33 * (-6.0f * x * z); D[34] = C33 * (-6.0f * y * z); D[35] = C33 * ((6.0f * zz - 3.0f * xx) - 3.0f * yy);
      D[36] = C34 * ((4.0f * zz - 3.0f * xx) - yy); D[37] = C34 * (-2.0f * x * y); D[38] = C34 * (8.0f * x * z);
      D[39] = C35 * (2.0f * x * z); D[40] = C35 * (-2.0f * y * z); D[41] = C35 * (xx - yy);
      D[42] = C36 * ((3.0f * xx) - yy); D[43] = C36 * (-2.0f * x * y); D[44] = 0.0f;
    }
  }
}

template <int kDeg>
struct ShDims {
  static constexpr int K = kDeg > 0 ? (kDeg + 1) * (kDeg + 1) - 1 : 0;
  static constexpr int C = 3 * K;               // coefficients per Gaussian
  static constexpr int P4 = (C + 3) / 4;        // float4 planes (f32 storage)
  static constexpr int Q8 = (C + 7) / 8;        // uint4 planes of 8 bf16 (mixed storage)
  static constexpr int S = 8 * Q8 > 4 * P4 ? 8 * Q8 : 4 * P4;   // register copy size
};

// view direction from the camera (shaders.js:284-285, :914-916)
__device__ __forceinline__ float view_dir(const Cam& cam, const float* P, float* v) {
  const float un[3] = {P[0] - cam.camPos[0], P[1] - cam.camPos[1], P[2] - cam.camPos[2]};
  const float ul = fmaxf(sqrtf(dot3(un, un)), 1e-9f);
  v[0] = un[0] / ul; v[1] = un[1] / ul; v[2] = un[2] / ul;
  return ul;
}

// Gaussian i's SH coefficient c lives in plane c/4, lane c%4.
template <int kDeg>
__device__ __forceinline__ void load_sh(const float4* __restrict__ SH, int64_t cap, int64_t i,
                                        float* sh) {
#pragma unroll
  for (int c = 0; c < ShDims<kDeg>::P4; c++) {
    const float4 v = SH[c * cap + i];
    sh[4 * c + 0] = v.x; sh[4 * c + 1] = v.y; sh[4 * c + 2] = v.z; sh[4 * c + 3] = v.w;
  }
}

// 16-bit planes (bf16 or fp16): value c of a Gaussian lives in uint4 plane
// c/8, half c%8 (even halves in the low 16 bits of each 32-bit word).
// kFmt 1 bf16, 2 fp16; both decode exactly.
template <int kFmt>
__device__ __forceinline__ float dec16(uint32_t b) {
  if constexpr (kFmt == 2) return (float)__builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu));
  else return gs_bf2f(b);
}

template <int kFmt>
__device__ __forceinline__ void unpack16x8(uint4 w, float* o) {
  o[0] = dec16<kFmt>(w.x); o[1] = dec16<kFmt>(w.x >> 16);
  o[2] = dec16<kFmt>(w.y); o[3] = dec16<kFmt>(w.y >> 16);
  o[4] = dec16<kFmt>(w.z); o[5] = dec16<kFmt>(w.z >> 16);
  o[6] = dec16<kFmt>(w.w); o[7] = dec16<kFmt>(w.w >> 16);
}

__device__ __forceinline__ void unpack_bf8(uint4 w, float* o) { unpack16x8<1>(w, o); }

__device__ __forceinline__ uint4 pack_bf8(const uint32_t* b) {
  return make_uint4(b[0] | (b[1] << 16), b[2] | (b[3] << 16), b[4] | (b[5] << 16), b[6] | (b[7] << 16));
}

// Storage of each precision mode (gs_opts.precision):
//   0 f32   everything f32
//   1 mixed bf16 SH coefficients and bf16 Adam moments (parameters and SH)
//   2 bf16  bf16 SH coefficients and raster-record colours; f32 moments
//   3 fp16  fp16 SH coefficients and raster-record colours; f32 moments
// Parameters, transmittance, gradients and all arithmetic are f32 in every
// mode; stored values are decoded exactly.
template <int kPrec>
struct Prec {
  static constexpr int sh = kPrec == 0 ? 0 : (kPrec == 3 ? 2 : 1);   // 0 f32 planes, 1 bf16, 2 fp16
  static constexpr bool mom16 = kPrec == 1;                           // bf16 moments
  static constexpr int rec = kPrec >= 2 ? kPrec - 1 : 0;              // record colours: 0 f32, 1 bf16, 2 fp16
};

// float4s per raster record: 4 (f32 colours) or 3 (16-bit colours)
constexpr int rec_stride(int recFmt) { return recFmt == 0 ? 4 : 3; }

template <int kDeg, int kFmt>
__device__ __forceinline__ void load_sh16(const uint4* __restrict__ SH, int64_t cap, int64_t i, float* sh) {
#pragma unroll
  for (int c = 0; c < ShDims<kDeg>::Q8; c++) unpack16x8<kFmt>(SH[c * cap + i], sh + 8 * c);
}

// SH coefficients from any storage (kSh 0: f32 planes, 1: bf16, 2: fp16)
template <int kDeg, int kSh>
__device__ __forceinline__ void load_sh_any(const void* __restrict__ SH, int64_t cap, int64_t i, float* sh) {
  if constexpr (kSh != 0) load_sh16<kDeg, kSh>((const uint4*)SH, cap, i, sh);
  else load_sh<kDeg>((const float4*)SH, cap, i, sh);
}

}  // namespace gs
