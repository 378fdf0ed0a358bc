/* gs_oracle.c — CPU restatement of the splat.js 3DGS trainer step.
 *
 * TEST INFRASTRUCTURE ONLY: it computes the expected results. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it; the
 * product (splat.js_amd/) never links or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off, no fast-math, so the
 * single-precision instantiation rounds exactly as written).
 */
#include "gs_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "../splat.js_amd/csrc/gs_math.h"

/* Threads of the per-Gaussian and per-pixel loops (orc_set_threads). 1, the
 * default, keeps every loop serial: the tests' expected values come from
 * that. Above 1 (the CPU baseline), the loops run under OpenMP; fixed-point
 * sums stay exact (integer atomics), double sums change order. */
static int orc_threads = 1;

void orc_set_threads(int n) {
  orc_threads = n < 1 ? 1 : n;
  omp_set_num_threads(orc_threads);
}

/* ---------------- single precision: the pinned-order twin ---------------- */
#define REAL float
#define SUFFIX _f32
#define EXP gs_expf
#define EXP_RASTER gs_expf_raster
#define LOG gs_logf
#define SQRT sqrtf
#define FMA fmaf
#define FMAX fmaxf
#define FMIN fminf
#define FLOOR floorf
#include "gs_oracle_body.h"
#undef REAL
#undef SUFFIX
#undef EXP
#undef EXP_RASTER
#undef LOG
#undef SQRT
#undef FMA
#undef FMAX
#undef FMIN
#undef FLOOR

/* ---------------- double precision: the mathematical reference ----------- */
#define REAL double
#define SUFFIX _f64
#define EXP exp
#define EXP_RASTER exp
#define LOG log
#define SQRT sqrt
#define FMA fma
#define FMAX fmax
#define FMIN fmin
#define FLOOR floor
#include "gs_oracle_body.h"

float orc_expf(float x) { return gs_expf(x); }
float orc_expf_raster(float x) { return gs_expf_raster(x); }
float orc_logf(float x) { return gs_logf(x); }

void orc_sh_basis_f64(int deg, const double v[3], double* Y, double* dY) {
  sh_basis_f64(deg, v, Y, dY);
}

/* shaders.js:121-126 camPosWorld, evaluated in f32 */
void orc_cam_pos(orc_cam* c) {
  const float* R = c->R; const float* t = c->t;
  c->camPos[0] = -((R[0] * t[0] + R[3] * t[1]) + R[6] * t[2]);
  c->camPos[1] = -((R[1] * t[0] + R[4] * t[1]) + R[7] * t[2]);
  c->camPos[2] = -((R[2] * t[0] + R[5] * t[1]) + R[8] * t[2]);
}

int64_t orc_count_keys(int64_t n, const uint32_t* cnt) {
  int64_t K = 0;
  for (int64_t i = 0; i < n; i++) K += cnt[i];
  return K;
}

/* shaders.js:984-1036 ADAM_SRC */
void orc_adam_f32(const orc_adam* hp, int64_t n, float* P, const float* G,
                  float* m, float* v) {
#pragma omp parallel for schedule(static) if (orc_threads > 1)
  for (int64_t j = 0; j < n * 16; j++) {
    const int slot = (int)(j % 16);
    const float lr = hp->lr[slot];
    if (lr == 0.0f) continue;
    float g = G[j];
    if (!(fabsf(g) < 1e18f)) g = 0.0f;
    if (slot == 13) {
      const float sg = gs_sigmoid_clamped(P[j]);
      g = g + hp->reg * sg * (1.0f - sg);
    }
    const float mm = hp->b1 * m[j] + (1.0f - hp->b1) * g;
    const float vv = hp->b2 * v[j] + (1.0f - hp->b2) * g * g;
    m[j] = mm; v[j] = vv;
    const float mh = mm * hp->rbc1;
    const float vh = vv * hp->rbc2;
    float p = P[j] - lr * mh / (sqrtf(vh) + hp->eps);
    if (slot >= 3 && slot <= 5) p = fminf(fmaxf(p, hp->minLS), hp->maxLS);
    if (slot >= 10 && slot <= 13) p = fminf(fmaxf(p, -hp->maxLogit), hp->maxLogit);
    P[j] = p;
  }
}

/* shaders.js:952-982 SH_ADAM_SRC */
void orc_sh_adam_f32(const orc_adam* hp, int64_t count, float* sh, const float* G,
                     float* m, float* v) {
#pragma omp parallel for schedule(static) if (orc_threads > 1)
  for (int64_t j = 0; j < count; j++) {
    float g = G[j];
    if (!(fabsf(g) < 1e18f)) g = 0.0f;
    const float mm = hp->b1 * m[j] + (1.0f - hp->b1) * g;
    const float vv = hp->b2 * v[j] + (1.0f - hp->b2) * g * g;
    m[j] = mm; v[j] = vv;
    const float mh = mm * hp->rbc1;
    const float vh = vv * hp->rbc2;
    sh[j] = sh[j] - hp->shLr * mh / (sqrtf(vh) + hp->eps);
  }
}

/* Mixed precision: the same Adam, with the moments (and the SH coefficients)
 * held in bf16. Each stored value is rounded stochastically with the bits
 * gs_sr_bits(gs_sr_seed(Gaussian, step), 64 * stream + slot), streams 0/1 for
 * the parameter moments m/v and 2/3/4 for the SH coefficient and its m/v, as
 * the GPU update does. Slots Adam skips (lr 0) keep their bits. */
void orc_adam_mixed_f32(const orc_adam* hp, int64_t n, float* P, const float* G,
                        uint16_t* m, uint16_t* v, uint32_t step) {
  for (int64_t i = 0; i < n; i++)
    for (int slot = 0; slot < 16; slot++) {
      const int64_t j = 16 * i + slot;
      const uint32_t seed = gs_sr_seed((uint64_t)i, step);
      float mf = gs_bf2f(m[j]), vf = gs_bf2f(v[j]);
      const float lr = hp->lr[slot];
      if (lr != 0.0f) {
        float g = G[j];
        if (!(fabsf(g) < 1e18f)) g = 0.0f;
        if (slot == 13) {
          const float sg = gs_sigmoid_clamped(P[j]);
          g = g + hp->reg * sg * (1.0f - sg);
        }
        const float mm = hp->b1 * mf + (1.0f - hp->b1) * g;
        const float vv = hp->b2 * vf + (1.0f - hp->b2) * g * g;
        mf = mm; vf = vv;
        const float mh = mm * hp->rbc1;
        const float vh = vv * hp->rbc2;
        float p = P[j] - lr * mh / (sqrtf(vh) + hp->eps);
        if (slot >= 3 && slot <= 5) p = fminf(fmaxf(p, hp->minLS), hp->maxLS);
        if (slot >= 10 && slot <= 13) p = fminf(fmaxf(p, -hp->maxLogit), hp->maxLogit);
        P[j] = p;
      }
      m[j] = (uint16_t)gs_f2bf_sr(mf, gs_sr_bits(seed, 0u * 64u + (uint32_t)slot));
      v[j] = (uint16_t)gs_f2bf_sr(vf, gs_sr_bits(seed, 1u * 64u + (uint32_t)slot));
    }
}

void orc_sh_adam_mixed_f32(const orc_adam* hp, int64_t n, int nc, uint16_t* sh, const float* G,
                           uint16_t* m, uint16_t* v, uint32_t step) {
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < nc; c++) {
      const int64_t j = (int64_t)nc * i + c;
      const uint32_t seed = gs_sr_seed((uint64_t)i, step);
      float g = G[j];
      if (!(fabsf(g) < 1e18f)) g = 0.0f;
      const float mm = hp->b1 * gs_bf2f(m[j]) + (1.0f - hp->b1) * g;
      const float vv = hp->b2 * gs_bf2f(v[j]) + (1.0f - hp->b2) * g * g;
      const float mh = mm * hp->rbc1;
      const float vh = vv * hp->rbc2;
      const float s = gs_bf2f(sh[j]) - hp->shLr * mh / (sqrtf(vh) + hp->eps);
      sh[j] = (uint16_t)gs_f2bf_sr(s, gs_sr_bits(seed, 2u * 64u + (uint32_t)c));
      m[j] = (uint16_t)gs_f2bf_sr(mm, gs_sr_bits(seed, 3u * 64u + (uint32_t)c));
      v[j] = (uint16_t)gs_f2bf_sr(vv, gs_sr_bits(seed, 4u * 64u + (uint32_t)c));
    }
}

/* Precision modes bf16 / fp16 (fmt 1 / 2): the SH coefficients held in 16
 * bits, the moments in f32. Plain SH Adam (orc_sh_adam_f32's arithmetic);
 * the new coefficient is stored with stochastic rounding, bits
 * gs_sr_bits(gs_sr_seed(Gaussian, step), 2 * 64 + slot) as in mode mixed. */
void orc_sh_adam_h16_f32(const orc_adam* hp, int64_t n, int nc, uint16_t* sh, const float* G,
                         float* m, float* v, uint32_t step, int fmt) {
  for (int64_t i = 0; i < n; i++)
    for (int c = 0; c < nc; c++) {
      const int64_t j = (int64_t)nc * i + c;
      const uint32_t seed = gs_sr_seed((uint64_t)i, step);
      float g = G[j];
      if (!(fabsf(g) < 1e18f)) g = 0.0f;
      const float mm = hp->b1 * m[j] + (1.0f - hp->b1) * g;
      const float vv = hp->b2 * v[j] + (1.0f - hp->b2) * g * g;
      m[j] = mm; v[j] = vv;
      const float mh = mm * hp->rbc1;
      const float vh = vv * hp->rbc2;
      const float cur = fmt == 2 ? gs_h2f(sh[j]) : gs_bf2f(sh[j]);
      const float s = cur - hp->shLr * mh / (sqrtf(vh) + hp->eps);
      const uint32_t r = gs_sr_bits(seed, 2u * 64u + (uint32_t)c);
      sh[j] = (uint16_t)(fmt == 2 ? gs_f2h_sr(s, r) : gs_f2bf_sr(s, r));
    }
}

void orc_f2h_rne(int64_t count, const float* src, uint16_t* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = (uint16_t)gs_f2h_rne(src[j]);
}

void orc_f2h_sr(int64_t count, const float* src, const uint32_t* rnd, uint16_t* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = (uint16_t)gs_f2h_sr(src[j], rnd[j]);
}

void orc_f2h_up(int64_t count, const float* src, uint16_t* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = (uint16_t)gs_f2h_up(src[j]);
}

void orc_h2f(int64_t count, const uint16_t* src, float* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = gs_h2f(src[j]);
}

void orc_f2bf_rne(int64_t count, const float* src, uint16_t* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = (uint16_t)gs_f2bf_rne(src[j]);
}

void orc_f2bf_sr(int64_t count, const float* src, const uint32_t* rnd, uint16_t* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = (uint16_t)gs_f2bf_sr(src[j], rnd[j]);
}

void orc_bf2f(int64_t count, const uint16_t* src, float* dst) {
  for (int64_t j = 0; j < count; j++) dst[j] = gs_bf2f(src[j]);
}

uint32_t orc_sr_bits(uint64_t gaussian, uint32_t slot, uint32_t step, uint32_t stream) {
  return gs_sr_bits(gs_sr_seed(gaussian, step), 64u * stream + slot);
}

/* JS Math.round (ties toward +inf) followed by ToInt32 */
static int32_t js_round_to_int32(double x) {
  if (x != x || isinf(x)) return 0;
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  /* ToInt32: modulo 2^32 */
  double m = fmod(r, 4294967296.0);
  if (m < 0) m += 4294967296.0;
  uint32_t u = (uint32_t)m;
  return (int32_t)u;
}

static int32_t js_min255(double x) {
  /* Math.min(255, Math.round(x)) then ToInt32 via the shift */
  if (x != x) return 0;
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  if (r > 255.0) r = 255.0;
  return js_round_to_int32(r);
}

/* trainer.js:128-139 */
void orc_pack_targets(int64_t npx, const float* rgb, uint32_t* out) {
  for (int64_t p = 0; p < npx; p++) {
    const double r = rgb[3 * p], g = rgb[3 * p + 1], b = rgb[3 * p + 2];
    if (r < 0) { out[p] = 0; continue; }
    const uint32_t vb = (uint32_t)js_min255(b * 255.0) << 16;
    const uint32_t vg = (uint32_t)js_min255(g * 255.0) << 8;
    const uint32_t vr = (uint32_t)js_min255(r * 255.0);
    out[p] = (255u << 24) | vb | vg | vr;
  }
}

/* ---------------- L1 + D-SSIM (textbook; absent from the reference) ------
 * Wang et al. 2004 SSIM with an 11x11 Gaussian window (sigma 1.5), computed
 * per channel with zero padding, as the INRIA 3DGS training loss uses it. */
static void ssim_window(double* g) {
  double s = 0;
  for (int i = 0; i < 11; i++) { g[i] = exp(-((i - 5) * (i - 5)) / (2.0 * 1.5 * 1.5)); s += g[i]; }
  for (int i = 0; i < 11; i++) g[i] /= s;
}

/* separable zero-padded correlation with the 11-tap window */
static void blur(int W, int H, const double* in, double* out, double* tmp, const double* g) {
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double s = 0;
      for (int k = -5; k <= 5; k++) { int xx = x + k; if (xx >= 0 && xx < W) s += g[k + 5] * in[(int64_t)y * W + xx]; }
      tmp[(int64_t)y * W + x] = s;
    }
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      double s = 0;
      for (int k = -5; k <= 5; k++) { int yy = y + k; if (yy >= 0 && yy < H) s += g[k + 5] * tmp[(int64_t)yy * W + x]; }
      out[(int64_t)y * W + x] = s;
    }
}

void orc_loss_l1ssim_f64(const orc_cam* cam, const double* img, const uint32_t* tgt,
                         double lambda, double* gC, uint8_t* valid, double* stats) {
  const int W = cam->W, H = cam->H;
  const int64_t np = (int64_t)W * H;
  const double C1 = 0.01 * 0.01, C2 = 0.03 * 0.03;
  double g[11];
  ssim_window(g);
  double* X = (double*)malloc(sizeof(double) * np * 13);
  double *Y = X + np, *mx = Y + np, *my = mx + np, *sxx = my + np, *syy = sxx + np, *sxy = syy + np;
  double *tmp = sxy + np, *A = tmp + np, *B = A + np, *Cm = B + np, *t1 = Cm + np, *t2 = t1 + np;
  stats[0] = stats[1] = stats[2] = 0;
  for (int64_t p = 0; p < np; p++) {
    valid[p] = (tgt[p] >> 24) != 0;
    if (valid[p]) stats[2] += 1;
    for (int c = 0; c < 3; c++) gC[3 * p + c] = 0;
  }
  for (int c = 0; c < 3; c++) {
    /* an invalid target pixel (undistortion border) enters every window as
     * zero in both images, like the padding outside the image, so the loss
     * does not depend on what is rendered there */
    for (int64_t p = 0; p < np; p++) {
      X[p] = valid[p] ? (double)cam->gain * img[4 * p + c] + (double)cam->bias : 0.0;
      Y[p] = valid[p] ? (double)((tgt[p] >> (8 * c)) & 0xffu) / 255.0 : 0.0;
    }
    blur(W, H, X, mx, tmp, g);
    blur(W, H, Y, my, tmp, g);
    for (int64_t p = 0; p < np; p++) t1[p] = X[p] * X[p];
    blur(W, H, t1, sxx, tmp, g);
    for (int64_t p = 0; p < np; p++) t1[p] = Y[p] * Y[p];
    blur(W, H, t1, syy, tmp, g);
    for (int64_t p = 0; p < np; p++) t1[p] = X[p] * Y[p];
    blur(W, H, t1, sxy, tmp, g);
    for (int64_t p = 0; p < np; p++) {
      const double ux = mx[p], uy = my[p];
      const double vx = sxx[p] - ux * ux, vy = syy[p] - uy * uy, cxy = sxy[p] - ux * uy;
      const double N1 = 2 * ux * uy + C1, N2 = 2 * cxy + C2;
      const double D1 = ux * ux + uy * uy + C1, D2 = vx + vy + C2;
      const double S = (N1 * N2) / (D1 * D2);
      if (valid[p]) {
        const double err = X[p] - Y[p];
        const double sq = err * err;
        stats[0] += sq;
        stats[1] += (1.0 - lambda) * fabs(err) + lambda * (1.0 - S);
        const double dS_dmx = (2 * uy * N2) / (D1 * D2) - S * (2 * ux) / D1;
        const double dS_dvx = -S / D2;
        const double dS_dcxy = 2 * N1 / (D1 * D2);
        A[p] = dS_dmx - 2 * ux * dS_dvx - uy * dS_dcxy;
        B[p] = dS_dvx;
        Cm[p] = dS_dcxy;
      } else {
        A[p] = B[p] = Cm[p] = 0;
      }
    }
    /* the window is symmetric, so correlation and convolution coincide */
    blur(W, H, A, t1, tmp, g);
    blur(W, H, B, t2, tmp, g);
    blur(W, H, Cm, A, tmp, g);
    for (int64_t p = 0; p < np; p++) {
      if (!valid[p]) continue;   /* constant zero in the loss: no gradient */
      const double dssim = t1[p] + 2.0 * X[p] * t2[p] + Y[p] * A[p];
      const double err = X[p] - Y[p];
      const double d = -lambda * dssim + (1.0 - lambda) * (err > 0 ? 1.0 : (err < 0 ? -1.0 : 0.0));
      gC[3 * p + c] = (double)cam->gain * d;
    }
  }
  free(X);
}
