/* gs_oracle.h — CPU restatement of the splat.js 3DGS trainer step.
 *
 * TEST INFRASTRUCTURE ONLY. This library computes the expected results for
 * the HIP implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it. The product never links it, and nothing in the
 * product falls back to it.
 *
 * Every function exists twice: *_f32 (single precision, in a pinned operation
 * order that the HIP kernels reproduce bit for bit: same formulas, same
 * association, fused multiply-adds at the same places, the portable exp/log of
 * splat.js_amd/csrc/gs_math.h) and *_f64 (double precision, libm exp/log: the
 * mathematical reference that finite differences pin).
 *
 * Layouts are the reference's (src/gs/shaders.js:14-29):
 *   params  n x 16 f32 [0-2 pos, 3-5 logScale, 6-9 quat (w,x,y,z) raw,
 *                       10-12 colour logits, 13 opacity logit, 14-15 pad]
 *   sh      n x 3K f32, channel-major (K red, K green, K blue)
 *   proj    n x 16     [mx, my, depth, conicA, conicB, conicC, comp, opacity,
 *                       r, g, b, visible, Va, Vb, Vc, radius]
 *   grad2d  n x 10     [dMeanX, dMeanY, dConicA*, dConicB*, dConicC*, dComp,
 *                       dLogitOpacity, dR, dG, dB]  (* = times cnorm)
 */
#ifndef GS_ORACLE_H
#define GS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  float R[9];          /* world-to-camera rotation, row-major */
  float t[3];
  float nearp;         /* 0.05 in the reference (trainer.js:362) */
  float f, cx, cy;
  float gain, bias;    /* exposure: gain = exp(g), bias = b */
  int W, H;
  float bg[3];
  int shActiveDeg;     /* runtime SH degree (ramps one band per 1000 iters) */
  float camPos[3];     /* -(R^T t), see orc_cam_pos */
  int tile;            /* tile side in pixels (8, 16 or 32); 0 means the reference's 16
                          (shaders.js:32) */
} orc_cam;

/* the camera's tile side */
static inline int orc_tile(const orc_cam* cam) { return cam->tile > 0 ? cam->tile : 16; }

typedef struct {
  float eCut, aMin, radClamp, anisoReg;
  int shDeg;           /* compiled degree: K = (deg+1)^2 - 1 rest coefficients */
} orc_opts;

typedef struct {
  float lr[16];
  float b1, b2, eps;
  float rbc1, rbc2;    /* 1 / (1 - b1^t), 1 / (1 - b2^t): bias corrections as multipliers */
  float minLS, maxLS, maxLogit, reg;
  float shLr;
} orc_adam;

/* camera position in world space, computed in f32 as the reference does
 * (shaders.js:121-126) */
void orc_cam_pos(orc_cam* cam);

/* portable exp/log (gs_math.h), exported so tests can pin them against libm */
float orc_expf(float x);
float orc_expf_raster(float x);
float orc_logf(float x);

/* R3/R4: projection. rect = n x 4 int32 (tx0, tx1, ty0, ty1), cnt = tiles
 * covered (0 when culled). Culled Gaussians have proj[11] = 0. */
void orc_project_f32(const orc_opts* o, const orc_cam* cam, int64_t n,
                     const float* params, const float* sh,
                     float* proj, int32_t* rect, uint32_t* cnt);
void orc_project_f64(const orc_opts* o, const orc_cam* cam, int64_t n,
                     const float* params, const float* sh,
                     double* proj, int32_t* rect, uint32_t* cnt);

/* SH basis values and gradients at unit direction v (K = 15 max) */
void orc_sh_basis_f64(int deg, const double v[3], double* Y, double* dY);

/* R6-R8: binning + sort. Orders every (Gaussian, covered tile) pair by
 * (tile, depth, id) by brute force (qsort). ids = n_keys u32,
 * tile_start = numTiles + 1 offsets. keys (optional) = (tile << 32) | f32
 * depth bits. Returns the number of keys. */
int64_t orc_bin_sort_f32(int64_t n, const float* proj, const int32_t* rect,
                         const uint32_t* cnt, int tilesX, int numTiles,
                         uint32_t* ids, uint32_t* tile_start, uint64_t* keys);
int64_t orc_bin_sort_f64(int64_t n, const double* proj, const int32_t* rect,
                         const uint32_t* cnt, int tilesX, int numTiles,
                         uint32_t* ids, uint32_t* tile_start, uint64_t* keys);
int64_t orc_count_keys(int64_t n, const uint32_t* cnt);

/* R9: forward composite. img = W*H*4 (C + T*bg, 1 - T), Tfin = final T,
 * end = one past the last processed entry (absolute index into ids). */
void orc_render_fwd_f32(const orc_opts* o, const orc_cam* cam, const float* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        float* img, float* Tfin, uint32_t* end);
void orc_render_fwd_f64(const orc_opts* o, const orc_cam* cam, const double* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        double* img, double* Tfin, uint32_t* end);

/* R10: Charbonnier loss (delta 0.03). tgt = packed RGBA8 for this camera's
 * pixels (alpha 0 = invalid). gC = W*H*3 dL/dC (0 for invalid pixels),
 * valid = W*H flags. stats[0] = sum |err|^2, [1] = loss, [2] = valid px.
 * expFix (2 x i32, x64 fixed point, wrapping) / expSum (2 x f64), either
 * may be NULL: exposure gradients d/d(log gain), d/d(bias) summed over the
 * valid pixels. */
void orc_loss_charb_f32(const orc_cam* cam, const float* img, const uint32_t* tgt,
                        float* gC, uint8_t* valid, double* stats, int32_t* expFix, double* expSum);
void orc_loss_charb_f64(const orc_cam* cam, const double* img, const uint32_t* tgt,
                        double* gC, uint8_t* valid, double* stats, int32_t* expFix, double* expSum);

/* L1 + lambda * D-SSIM (textbook; the reference has no SSIM). Loss summed
 * over valid pixel-channels: (1-lambda)|x - y| + lambda (1 - ssim(p, c)),
 * ssim with an 11x11 Gaussian window (sigma 1.5), C1 = 0.01^2, C2 = 0.03^2,
 * zero padding. x = gain*C + bias. gC = dL/dC. */
void orc_loss_l1ssim_f64(const orc_cam* cam, const double* img, const uint32_t* tgt,
                         double lambda, double* gC, uint8_t* valid, double* stats);

/* R11: backward composite. Exactly one of gfix (n x 10 int32, fixed point
 * x16384 with the reference's clamp+truncate, wrapping sums) and gsum
 * (n x 10, plain sums) is non-NULL. */
void orc_render_bwd_f32(const orc_opts* o, const orc_cam* cam, const float* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        const float* Tfin, const uint32_t* end,
                        const float* gC, const uint8_t* valid,
                        int32_t* gfix, double* gsum);
void orc_render_bwd_f64(const orc_opts* o, const orc_cam* cam, const double* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        const double* Tfin, const uint32_t* end,
                        const double* gC, const uint8_t* valid,
                        int32_t* gfix, double* gsum);

/* R12: chain to parameter gradients. Input: gfix (decoded /16384 as the
 * reference does) or gsum. Output gradF n x 16, shGrad n x 3K. camFix
 * (7 x i32, x64 fixed point, wrapping) / camSum (7 x f64), either may be
 * NULL: camera rotation (3), translation (3) and log-focal gradients summed
 * over the Gaussians. */
void orc_chain_f32(const orc_opts* o, const orc_cam* cam, int64_t n,
                   const float* params, const float* sh, const float* proj,
                   const int32_t* gfix, const double* gsum,
                   float* gradF, float* shGrad, int32_t* camFix, double* camSum);
void orc_chain_f64(const orc_opts* o, const orc_cam* cam, int64_t n,
                   const float* params, const float* sh, const double* proj,
                   const int32_t* gfix, const double* gsum,
                   double* gradF, double* shGrad, int32_t* camFix, double* camSum);

/* R13/R14: Adam over params (n x 16) and SH (n x 3K), in place. */
void orc_adam_f32(const orc_adam* hp, int64_t n, float* params, const float* gradF,
                  float* m, float* v);
void orc_sh_adam_f32(const orc_adam* hp, int64_t count, float* sh, const float* g,
                     float* m, float* v);
/* threads of the per-Gaussian and per-pixel loops; 1 (default) is serial */
void orc_set_threads(int n);

/* precision modes bf16 / fp16 (fmt 1 / 2): 16-bit SH coefficients, f32 moments */
void orc_sh_adam_h16_f32(const orc_adam* hp, int64_t n, int nc, uint16_t* sh, const float* g,
                         float* m, float* v, uint32_t step, int fmt);
void orc_f2h_rne(int64_t count, const float* src, uint16_t* dst);
void orc_f2h_sr(int64_t count, const float* src, const uint32_t* rnd, uint16_t* dst);
void orc_f2h_up(int64_t count, const float* src, uint16_t* dst);
void orc_h2f(int64_t count, const uint16_t* src, float* dst);
/* mixed precision: bf16 moments and SH coefficients, stochastic rounding */
void orc_adam_mixed_f32(const orc_adam* hp, int64_t n, float* params, const float* gradF,
                        uint16_t* m, uint16_t* v, uint32_t step);
void orc_sh_adam_mixed_f32(const orc_adam* hp, int64_t n, int nc, uint16_t* sh, const float* g,
                           uint16_t* m, uint16_t* v, uint32_t step);
void orc_f2bf_rne(int64_t count, const float* src, uint16_t* dst);
void orc_f2bf_sr(int64_t count, const float* src, const uint32_t* rnd, uint16_t* dst);
void orc_bf2f(int64_t count, const uint16_t* src, float* dst);
uint32_t orc_sr_bits(uint64_t gaussian, uint32_t slot, uint32_t step, uint32_t stream);

/* R16: pack float RGB (r < 0 = invalid) into RGBA8 exactly as
 * trainer.js:128-139 does (JS Math.round, ToInt32 shifts). */
void orc_pack_targets(int64_t npx, const float* rgb, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
