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

GS_HD float gs_cover_clamp(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

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

// Bit w (w = 0..3: column w & 1, row w >> 1 of 8x8 blocks) set when some
// pixel centre of that block of 16x16 tile (tx, ty) may take the splat.
GS_HD uint32_t gs_cover_tile_mask(float mx, float my, float A, float B, float C, float L, int tx, int ty) {
  const float x0 = (float)(tx * 16) + 0.5f, y0 = (float)(ty * 16) + 0.5f;
  uint32_t m = 0;
  for (int w = 0; w < 4; w++) {
    const float bx = x0 + (float)((w & 1) * 8), by = y0 + (float)((w >> 1) * 8);
    if (gs_cover_rect(A, B, C, L, bx - mx, (bx + 7.0f) - mx, by - my, (by + 7.0f) - my)) m |= 1u << w;
  }
  return m;
}

// Cover level stored with each splat: an upper bound on the largest f32
// exponent a pixel can take the splat at. eMax = min(eCut, ln(opa*comp /
// A_MIN)) is computed in f32 (division, gs_logf: a few ulp), and alpha =
// min(0.99, opa*comp * gs_expf(-e)) >= A_MIN needs e <= ln(opa*comp/A_MIN)
// up to the ~1 ulp of gs_expf; the relative and absolute 1e-5 cover both.
GS_HD float gs_cover_level(float eMax) { return eMax * 1.00001f + 1.0e-5f; }

#endif  // GS_COVER_H
