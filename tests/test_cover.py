"""The rasterizer's wave-skip test (splat.js_amd/csrc/gs_cover.h) against
brute force: for random splats — conics built the way the projection builds
them (covariance + 0.3 px dilation, inverted in f32), with many needle-like,
diagonal and huge ones — and 8x8 pixel blocks placed around the ellipse
boundary, every pixel centre's exponent is evaluated with the kernels' exact
f32 operations. A block the test rejects must hold no pixel with e <= L
(else skipping it would change an image); the test should also reject most
blocks that indeed hold none (else it saves nothing)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <math.h>
#include "gs_cover.h"

static uint64_t s = 0x9E3779B97F4A7C15ull;
static double urand(void) {            /* xorshift64*, [0,1) */
  s ^= s >> 12; s ^= s << 25; s ^= s >> 27;
  return (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

/* the raster kernels' per-pixel exponent, operation for operation */
static float pixel_e(float px, float py, float mx, float my, float A, float B, float C) {
  const float dx = px - mx, dy = py - my;
  const float t2 = fmaf(C * dy, dy, (A * dx) * dx);
  return fmaxf(fmaf(0.5f, t2, (B * dx) * dy), 0.0f);
}

int main(int argc, char** argv) {
  const long cases = atol(argv[1]);
  long blocks = 0, unsafe = 0, takes = 0, empty = 0, empty_rejected = 0;
  for (long it = 0; it < cases; it++) {
    /* covariance: log-uniform diagonal, correlation up to 1 - 1e-8 */
    const float va = (float)pow(10.0, -4.0 + 10.0 * urand());
    const float vc = (float)pow(10.0, -4.0 + 10.0 * urand());
    const double rho = (urand() < 0.5 ? -1.0 : 1.0) * (1.0 - pow(10.0, -8.0 * urand()));
    const float vb = (float)(rho * sqrt((double)va * (double)vc) * (urand() < 0.2 ? 0.0 : 1.0));
    const float ad = va + 0.3f, cd = vc + 0.3f;
    const float detVd = ad * cd - vb * vb;
    if (detVd < 1e-8f) continue;
    const float inv = 1.0f / detVd;
    const float A = cd * inv, B = (-vb) * inv, C = ad * inv;
    const float L = (float)pow(10.0, -4.0 + 6.0 * urand());       /* 1e-4 .. 100 */
    const float mx = (float)(-200.0 + 2400.0 * urand()), my = (float)(-200.0 + 1400.0 * urand());
    /* blocks near the ellipse: offsets up to ~1.3x its extent along x and y */
    const double exx = sqrt(2.0 * L * (double)ad) * 1.3 + 12.0, eyy = sqrt(2.0 * L * (double)cd) * 1.3 + 12.0;
    for (int b = 0; b < 8; b++) {
      const float bx = (float)floor(mx + (2.0 * urand() - 1.0) * exx) + 0.5f;
      const float by = (float)floor(my + (2.0 * urand() - 1.0) * eyy) + 0.5f;
      float emin = INFINITY;
      for (int j = 0; j < 8; j++)
        for (int i = 0; i < 8; i++)
          emin = fminf(emin, pixel_e(bx + (float)i, by + (float)j, mx, my, A, B, C));
      /* every other block is borderline: L set to its smallest pixel exponent */
      const float Lb = (b & 1) ? L : emin;
      if (!(Lb > 0.0f) || Lb == INFINITY) continue;
      const int any = emin <= Lb;
      const int hit = gs_cover_rect(A, B, C, Lb, bx - mx, (bx + 7.0f) - mx, by - my, (by + 7.0f) - my);
      blocks++;
      if (any) { takes++; if (!hit) unsafe++; }
      else { empty++; if (!hit) empty_rejected++; }
    }
  }
  /* the tile form used by the binning pass: T x T tiles (T = 8, 16, 32 in
     turn), 4 quadrant bits */
  long tblocks = 0, tunsafe = 0, ttakes = 0, tempty = 0, tempty_rejected = 0;
  for (long it = 0; it < cases / 4; it++) {
    const int T = 8 << (it % 3), hq = T / 2;
    const float va = (float)pow(10.0, -4.0 + 10.0 * urand());
    const float vc = (float)pow(10.0, -4.0 + 10.0 * urand());
    const double rho = (urand() < 0.5 ? -1.0 : 1.0) * (1.0 - pow(10.0, -8.0 * urand()));
    const float vb = (float)(rho * sqrt((double)va * (double)vc) * (urand() < 0.2 ? 0.0 : 1.0));
    const float ad = va + 0.3f, cd = vc + 0.3f;
    const float detVd = ad * cd - vb * vb;
    if (detVd < 1e-8f) continue;
    const float inv = 1.0f / detVd;
    const float A = cd * inv, B = (-vb) * inv, C = ad * inv;
    const float L0 = (float)pow(10.0, -4.0 + 6.0 * urand());
    const float mx = (float)(-200.0 + 2400.0 * urand()), my = (float)(-200.0 + 1400.0 * urand());
    const double exx = sqrt(2.0 * L0 * (double)ad) * 1.3 + 24.0, eyy = sqrt(2.0 * L0 * (double)cd) * 1.3 + 24.0;
    for (int b = 0; b < 4; b++) {
      const int tx = (int)floor((mx + (2.0 * urand() - 1.0) * exx) / (double)T);
      const int ty = (int)floor((my + (2.0 * urand() - 1.0) * eyy) / (double)T);
      float emin[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
      for (int j = 0; j < T; j++)
        for (int i = 0; i < T; i++) {
          const float e = pixel_e((float)(tx * T + i) + 0.5f, (float)(ty * T + j) + 0.5f, mx, my, A, B, C);
          const int blk = (i / hq) + 2 * (j / hq);
          emin[blk] = fminf(emin[blk], e);
        }
      /* every other tile is borderline: L = smallest exponent of one block */
      const float L = (b & 1) ? L0 : emin[(int)(urand() * 4.0) & 3];
      if (!(L > 0.0f) || L == INFINITY) continue;
      const gs_cover_splat sp = gs_cover_prepare(mx, my, A, B, C, L);
      const uint32_t m = gs_cover_tile_mask(&sp, tx, ty, T);
      for (int blk = 0; blk < 4; blk++) {
        tblocks++;
        const int any = emin[blk] <= L, hit = (m >> blk) & 1u;
        if (any) { ttakes++; if (!hit) tunsafe++; }
        else { tempty++; if (!hit) tempty_rejected++; }
      }
    }
  }
  printf("%ld %ld %ld %ld %ld %ld %ld %ld %ld %ld\n", blocks, unsafe, takes, empty, empty_rejected,
         tblocks, tunsafe, ttakes, tempty, tempty_rejected);
  return 0;
}
"""


def _run(tmp_path, cases, defines=()):
    src = tmp_path / "cover.c"
    src.write_text(PROBE)
    exe = tmp_path / ("cover" + "".join(d.replace("=", "_") for d in defines))
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", *[f"-D{d}" for d in defines],
                           "-I", os.path.join(ROOT, "splat.js_amd", "csrc"), "-o", str(exe), str(src), "-lm"])
    return map(int, subprocess.check_output([str(exe), str(cases)], text=True).split())


def test_cover_never_rejects_a_taking_block(tmp_path):
    r = list(_run(tmp_path, 400000))
    for label, (blocks, unsafe, takes, empty, empty_rejected) in (("rect", r[:5]), ("tile mask", r[5:])):
        assert blocks > 1_000_000 and takes > 200_000 and empty > 200_000, label
        assert unsafe == 0, f"{label}: {unsafe} blocks with a taking pixel were rejected"
        # tight: rejects nearly every block that no pixel centre takes
        assert empty_rejected >= 0.90 * empty, (label, empty_rejected, empty)


def test_cover_probe_catches_a_missing_rounding_margin(tmp_path):
    """Without the margin, needle-like splats lose pixels: the probe sees it
    in both forms."""
    r = list(_run(tmp_path, 400000, ["GS_COVER_MARGIN=0.0f"]))
    assert r[1] > 0 and r[6] > 0, r
