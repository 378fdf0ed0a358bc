// gs_cover.h — conservative "can a pixel of this rectangle take the splat"
// test, used to skip (splat, 8x8 pixel block) pairs in the rasterizer.
//
// A pixel takes a splat only if its exponent
//     e = 0.5 * (A dx^2 + C dy^2) + B dx dy,   dx = px - mx, dy = py - my
// (as the raster kernels evaluate it in f32) is at most L, the splat's
// cover level (an upper bound on min(eCut, ln(opa*comp / A_MIN)), stored by
// the projection). For the rectangle of pixel centres [u0,u1] x [w0,w1]
// (offsets from the mean) the test computes the exact minimum of the
// quadratic over the rectangle: 0 when the mean lies inside, else the
// smallest of the four edge minima (on an edge the quadratic is a 1-D
// parabola, minimised at the clamped vertex). The conic is positive
// definite (A > 0, C > 0), so the minimum over a box that misses the
// origin lies on its boundary.
//
// Rounding. With S = 0.5 (A du^2 + C dw^2) + |B| du dw, du = max|u|,
// dw = max|w| over the rectangle, the f32 exponent of any pixel centre in
// the rectangle differs from the exact quadratic by at most about 6 ulp * S
// (two roundings per product, one per sum, plus the rounding of dx, dy),
// and so does the f32 minimum computed here (the clamped vertex is off by
// O(ulp), which changes the parabola by O(ulp^2)). The rectangle is
// rejected only when its minimum exceeds L + 2e-6 * S (about 33 ulp * S),
// so every rejected rectangle holds only pixels whose f32 exponent exceeds
// L: skipping it changes no pixel. For ordinary splats S is a few tens and
// the margin is negligible; for needle-like splats (large S) the test
// loosens by itself instead of dropping pixels.
//
// Included from C (tests) and from HIP device code.
#ifndef GS_COVER_H
#define GS_COVER_H

#include "gs_math.h"

// Rounding margin per unit of S (see above); a test build may override it.
#ifndef GS_COVER_MARGIN
#define GS_COVER_MARGIN 2.0e-6f
#endif

GS_HD float gs_cover_q(float A, float B, float C, float u, float w) {
  return 0.5f * (A * u * u + C * w * w) + B * u * w;
}

// clamp(v, lo, hi) for lo <= hi and finite v (the only uses below). On the
// device one v_med3_f32: fminf/fmaxf of values computed in another basic
// block each cost an extra canonicalizing instruction per operand.
GS_HD float gs_cover_clamp(float v, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_fmed3f(v, lo, hi);
#else
  return fminf(fmaxf(v, lo), hi);
#endif
}

// 1 when some pixel centre of [u0,u1] x [w0,w1] (offsets from the mean,
// u0 <= u1, w0 <= w1) may have e <= L; 0 only when none can.
GS_HD int gs_cover_rect(float A, float B, float C, float L, float u0, float u1, float w0, float w1) {
  if (u0 <= 0.0f && u1 >= 0.0f && w0 <= 0.0f && w1 >= 0.0f) return 1;
  if (!(A > 0.0f && C > 0.0f)) return 1;   // not positive definite: never reject
  const float kC = -B / C, kA = -B / A;
  float q = gs_cover_q(A, B, C, u0, gs_cover_clamp(kC * u0, w0, w1));
  q = fminf(q, gs_cover_q(A, B, C, u1, gs_cover_clamp(kC * u1, w0, w1)));
  q = fminf(q, gs_cover_q(A, B, C, gs_cover_clamp(kA * w0, u0, u1), w0));
  q = fminf(q, gs_cover_q(A, B, C, gs_cover_clamp(kA * w1, u0, u1), w1));
  const float du = fmaxf(fabsf(u0), fabsf(u1)), dw = fmaxf(fabsf(w0), fabsf(w1));
  const float S = 0.5f * (A * du * du + C * dw * dw) + fabsf(B) * du * dw;
  return !(q > L + GS_COVER_MARGIN * S);
}

// Per-splat constants of the tile test (gs_cover_tile_mask).
typedef struct {
  float mx, my, hA, B, hC, L;   // hA = A/2, hC = C/2
  float kA, kC;                 // -B/A, -B/C: edge-vertex slopes
  int pd;                       // conic positive definite
} gs_cover_splat;

GS_HD gs_cover_splat gs_cover_prepare(float mx, float my, float A, float B, float C, float L) {
  gs_cover_splat s;
  s.mx = mx; s.my = my; s.hA = 0.5f * A; s.B = B; s.hC = 0.5f * C; s.L = L;
  s.pd = A > 0.0f && C > 0.0f;
  s.kA = s.pd ? -B / A : 0.0f;
  s.kC = s.pd ? -B / C : 0.0f;
  return s;
}

// Bit w (w = 0..3: column w & 1, row w >> 1 of the tile's 2x2 quadrants)
// set when some pixel centre of that quadrant of the T x T tile (tx, ty)
// may take the splat (T = 8, 16 or 32; quadrants of T/2 pixels; at T = 16
// they are the 8x8 blocks the rasterizer's waves own). The same test as
// gs_cover_rect on each quadrant, arranged so the four share their edge
// lines: u[0..3] = x offsets of pixel-centre columns 0, h - 1, h, 2h - 1
// (h = T/2), w[0..3] likewise for rows; quadrant (i, j) spans u[2i..2i+1] x
// w[2j..2j+1]. A vertical edge line u is evaluated at the clamped vertex
// kC*u of each block row, q = hA u^2 + wc (hC wc + B u), and symmetrically
// for horizontal lines; the f32 error of this form is also a few ulp of S.
// Branch-free, so the whole test is one basic block on the device (a
// conic that is not positive definite keeps every block: kA = kC = 0 there
// and the mask computed for it is discarded).
GS_HD uint32_t gs_cover_tile_mask(const gs_cover_splat* s, int tx, int ty, int T) {
  const float x0 = (float)(tx * T) + 0.5f, y0 = (float)(ty * T) + 0.5f;
  const float h1 = (float)(T / 2 - 1), h2 = (float)(T / 2), h3 = (float)(T - 1);
  const float u[4] = {x0 - s->mx, (x0 + h1) - s->mx, (x0 + h2) - s->mx, (x0 + h3) - s->mx};
  const float w[4] = {y0 - s->my, (y0 + h1) - s->my, (y0 + h2) - s->my, (y0 + h3) - s->my};
  float qv[4][2], qh[4][2];   // [line][block row / column]
  for (int a = 0; a < 4; a++) {
    const float au = s->hA * u[a] * u[a], bu = s->B * u[a], tu = s->kC * u[a];
    const float aw = s->hC * w[a] * w[a], bw = s->B * w[a], tw = s->kA * w[a];
    for (int j = 0; j < 2; j++) {
      const float wc = gs_cover_clamp(tu, w[2 * j], w[2 * j + 1]);
      qv[a][j] = fmaf(wc, fmaf(s->hC, wc, bu), au);
      const float uc = gs_cover_clamp(tw, u[2 * j], u[2 * j + 1]);
      qh[a][j] = fmaf(uc, fmaf(s->hA, uc, bw), aw);
    }
  }
  const float aB = fabsf(s->B);
  uint32_t m = 0;
  for (int bj = 0; bj < 2; bj++)
    for (int bi = 0; bi < 2; bi++) {
      const float u0 = u[2 * bi], u1 = u[2 * bi + 1], w0 = w[2 * bj], w1 = w[2 * bj + 1];
      const int inside = (u0 <= 0.0f) & (u1 >= 0.0f) & (w0 <= 0.0f) & (w1 >= 0.0f);
      const float q = fminf(fminf(qv[2 * bi][bj], qv[2 * bi + 1][bj]), fminf(qh[2 * bj][bi], qh[2 * bj + 1][bi]));
      const float du = fmaxf(fabsf(u0), fabsf(u1)), dw = fmaxf(fabsf(w0), fabsf(w1));
      const float S = fmaf(s->hA * du, du, fmaf(s->hC * dw, dw, aB * du * dw));
      const int near = !(q > s->L + GS_COVER_MARGIN * S);
      m |= (uint32_t)(inside | near) << (2 * bj + bi);
    }
  return s->pd ? m : 0xfu;
}

// Cover level stored with each splat: an upper bound on the largest f32
// exponent a pixel can take the splat at. eMax = min(eCut, ln(opa*comp /
// A_MIN)) is computed in f32 (division, gs_logf: a few ulp), and alpha =
// min(0.99, opa*comp * gs_expf_raster(-e)) >= A_MIN needs e <=
// ln(opa*comp/A_MIN) up to the 2 ulp of gs_expf_raster; the relative and
// absolute 1e-5 cover both.
GS_HD float gs_cover_level(float eMax) { return eMax * 1.00001f + 1.0e-5f; }

#endif  // GS_COVER_H
