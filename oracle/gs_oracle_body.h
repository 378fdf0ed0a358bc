/* gs_oracle_body.h — precision-generic body of the CPU restatement.
 * TEST INFRASTRUCTURE ONLY (see gs_oracle.h).
 *
 * Included twice by gs_oracle.c with REAL = float / double. Each function
 * cites the reference lines it restates. Association follows the WGSL text
 * left to right (a*b*c = (a*b)*c, dot = (x + y) + z); FMA() marks the few
 * places where the pinned order uses one fused multiply-add (the HIP kernels
 * use fmaf at exactly these places). */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SUFFIX)
#define GEOM CAT(geom, SUFFIX)
#define KEYREC CAT(keyrec, SUFFIX)
#define RL(x) ((REAL)(x))

typedef struct {
  int ok;
  REAL pc[3], q[4], r0[3], r1[3], r2[3], s[3], t0[3], t1[3];
  REAL va, vb, vc, s00, s01, s02, s11, s12, s22;
} GEOM;

static REAL FN(clampr)(REAL x, REAL lo, REAL hi) { return FMIN(FMAX(x, lo), hi); }
static REAL FN(sigm)(REAL x) { return RL(1.0) / (RL(1.0) + EXP(-FN(clampr)(x, RL(-9.0), RL(9.0)))); }
static REAL FN(dot3)(const REAL* a, const REAL* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static void FN(cross3)(const REAL* a, const REAL* b, REAL* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
/* camAdd (shaders.js:522-524, :717-719): i32(clamp(v * 64, -1e9, 1e9)) added
 * with wrapping; dsum (optional) takes the unscaled value */
static void FN(cam_add)(int32_t* fix, double* dsum, int slot, REAL v) {
  if (fix) {
    const REAL fv = v * RL(64.0);
    const REAL cl = FMIN(FMAX(fv, RL(-1.0e9)), RL(1.0e9));
    fix[slot] = (int32_t)((uint32_t)fix[slot] + (uint32_t)(int32_t)cl);
  }
  if (dsum) dsum[slot] += (double)v;
}

/* shaders.js:176-224 computeGeom */
static void FN(compute_geom)(const orc_cam* cam, const float* P, GEOM* g) {
  REAL R[9], t[3];
  for (int k = 0; k < 9; k++) R[k] = cam->R[k];
  for (int k = 0; k < 3; k++) t[k] = cam->t[k];
  g->ok = 0;
  const REAL p[3] = {P[0], P[1], P[2]};
  g->pc[0] = FN(dot3)(R + 0, p) + t[0];
  g->pc[1] = FN(dot3)(R + 3, p) + t[1];
  g->pc[2] = FN(dot3)(R + 6, p) + t[2];
  if (g->pc[2] < RL(cam->nearp)) return;

  REAL q[4] = {P[6], P[7], P[8], P[9]};
  const REAL ql = SQRT(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
  if (ql < RL(1e-6)) { q[0] = 1; q[1] = 0; q[2] = 0; q[3] = 0; }
  else { for (int k = 0; k < 4; k++) q[k] = q[k] / ql; }
  for (int k = 0; k < 4; k++) g->q[k] = q[k];
  const REAL qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  g->r0[0] = RL(1.0) - RL(2.0) * (qy * qy + qz * qz);
  g->r0[1] = RL(2.0) * (qx * qy - qw * qz);
  g->r0[2] = RL(2.0) * (qx * qz + qw * qy);
  g->r1[0] = RL(2.0) * (qx * qy + qw * qz);
  g->r1[1] = RL(1.0) - RL(2.0) * (qx * qx + qz * qz);
  g->r1[2] = RL(2.0) * (qy * qz - qw * qx);
  g->r2[0] = RL(2.0) * (qx * qz - qw * qy);
  g->r2[1] = RL(2.0) * (qy * qz + qw * qx);
  g->r2[2] = RL(1.0) - RL(2.0) * (qx * qx + qy * qy);

  for (int k = 0; k < 3; k++) g->s[k] = EXP(FN(clampr)(RL(P[3 + k]), RL(-12.0), RL(6.0)));

  REAL m0[3], m1[3], m2[3];
  for (int k = 0; k < 3; k++) { m0[k] = g->r0[k] * g->s[k]; m1[k] = g->r1[k] * g->s[k]; m2[k] = g->r2[k] * g->s[k]; }
  g->s00 = FN(dot3)(m0, m0); g->s01 = FN(dot3)(m0, m1); g->s02 = FN(dot3)(m0, m2);
  g->s11 = FN(dot3)(m1, m1); g->s12 = FN(dot3)(m1, m2); g->s22 = FN(dot3)(m2, m2);

  const REAL f = cam->f;
  const REAL iz = RL(1.0) / g->pc[2];
  const REAL a = f * iz;
  const REAL bx = ((-f) * g->pc[0] * iz) * iz;
  const REAL by = ((-f) * g->pc[1] * iz) * iz;
  for (int k = 0; k < 3; k++) {
    g->t0[k] = a * R[0 + k] + bx * R[6 + k];
    g->t1[k] = a * R[3 + k] + by * R[6 + k];
  }
  REAL st0[3], st1[3];
  st0[0] = (g->s00 * g->t0[0] + g->s01 * g->t0[1]) + g->s02 * g->t0[2];
  st0[1] = (g->s01 * g->t0[0] + g->s11 * g->t0[1]) + g->s12 * g->t0[2];
  st0[2] = (g->s02 * g->t0[0] + g->s12 * g->t0[1]) + g->s22 * g->t0[2];
  st1[0] = (g->s00 * g->t1[0] + g->s01 * g->t1[1]) + g->s02 * g->t1[2];
  st1[1] = (g->s01 * g->t1[0] + g->s11 * g->t1[1]) + g->s12 * g->t1[2];
  st1[2] = (g->s02 * g->t1[0] + g->s12 * g->t1[1]) + g->s22 * g->t1[2];
  g->va = FMAX(FN(dot3)(g->t0, st0), RL(0.0));
  g->vb = FN(dot3)(g->t1, st0);
  g->vc = FMAX(FN(dot3)(g->t1, st1), RL(0.0));
  g->ok = 1;
}

/* shaders.js:69-111 shBasis / shBasisGrad; constants :60-64 */
static void FN(sh_basis)(int deg, const REAL* v, REAL* Y, REAL* D /* K x 3, may be NULL */) {
  const REAL x = v[0], y = v[1], z = v[2];
  const REAL xx = x * x, yy = y * y, zz = z * z;
  const REAL C1 = RL(0.4886025119029199);
  const REAL C20 = RL(1.0925484305920792), C21 = RL(-1.0925484305920792), C22 = RL(0.31539156525252005),
             C23 = RL(-1.0925484305920792), C24 = RL(0.5462742152960396);
  const REAL C30 = RL(-0.5900435899266435), C31 = RL(2.890611442640554), C32 = RL(-0.4570457994644658),
             C33 = RL(0.3731763325901154), C34 = RL(-0.4570457994644658), C35 = RL(1.445305721320277),
             C36 = RL(-0.5900435899266435);
  Y[0] = (-C1) * y; Y[1] = C1 * z; Y[2] = (-C1) * x;
  if (D) {
    D[0] = 0; D[1] = -C1; D[2] = 0;
    D[3] = 0; D[4] = 0; D[5] = C1;
    D[6] = -C1; D[7] = 0; D[8] = 0;
  }
  if (deg >= 2) {
    Y[3] = C20 * x * y;
    Y[4] = C21 * y * z;
    Y[5] = C22 * ((RL(2.0) * zz - xx) - yy);
    Y[6] = C23 * x * z;
    Y[7] = C24 * (xx - yy);
    if (D) {
      D[9] = C20 * y; D[10] = C20 * x; D[11] = 0;
      D[12] = 0; D[13] = C21 * z; D[14] = C21 * y;
      D[15] = C22 * (RL(-2.0) * x); D[16] = C22 * (RL(-2.0) * y); D[17] = C22 * (RL(4.0) * z);
      D[18] = C23 * z; D[19] = 0; D[20] = C23 * x;
      D[21] = C24 * (RL(2.0) * x); D[22] = C24 * (RL(-2.0) * y); D[23] = 0;
    }
  }
  if (deg >= 3) {
    Y[8] = C30 * y * (RL(3.0) * xx - yy);
    Y[9] = C31 * x * y * z;
    Y[10] = C32 * y * ((RL(4.0) * zz - xx) - yy);
    Y[11] = C33 * z * ((RL(2.0) * zz - RL(3.0) * xx) - RL(3.0) * yy);
    Y[12] = C34 * x * ((RL(4.0) * zz - xx) - yy);
    Y[13] = C35 * z * (xx - yy);
    Y[14] = C36 * x * (xx - yy);
    if (D) {
      D[24] = C30 * (RL(6.0) * x * y); D[25] = C30 * (RL(3.0) * xx - RL(3.0) * yy); D[26] = 0;
      D[27] = C31 * (y * z); D[28] = C31 * (x * z); D[29] = C31 * (x * y);
      D[30] = C32 * (RL(-2.0) * x * y); D[31] = C32 * ((RL(4.0) * zz - xx) - RL(3.0) * yy); D[32] = C32 * (RL(8.0) * y * z);
      D[33] = C33 * (RL(-6.0) * x * z); D[34] = C33 * (RL(-6.0) * y * z); D[35] = C33 * ((RL(6.0) * zz - RL(3.0) * xx) - RL(3.0) * yy);
      D[36] = C34 * ((RL(4.0) * zz - RL(3.0) * xx) - yy); D[37] = C34 * (RL(-2.0) * x * y); D[38] = C34 * (RL(8.0) * x * z);
      D[39] = C35 * (RL(2.0) * x * z); D[40] = C35 * (RL(-2.0) * y * z); D[41] = C35 * (xx - yy);
      D[42] = C36 * ((RL(3.0) * xx) - yy); D[43] = C36 * (RL(-2.0) * x * y); D[44] = 0;
    }
  }
}

static int FN(sh_active_k)(int compiledDeg, int activeDeg) {
  /* shaders.js:117-120 */
  const int K = (compiledDeg + 1) * (compiledDeg + 1) - 1;
  const int a = (activeDeg + 1) * (activeDeg + 1) - 1;
  return a < K ? a : K;
}

/* shaders.js:228-315 makeProjectSrc */
void FN(orc_project)(const orc_opts* o, const orc_cam* cam, int64_t n,
                     const float* params, const float* sh,
                     REAL* proj, int32_t* rect, uint32_t* cnt) {
  const int K = o->shDeg > 0 ? (o->shDeg + 1) * (o->shDeg + 1) - 1 : 0;
  const REAL E_CUT = o->eCut, A_MIN = o->aMin, RADCL = o->radClamp;
  const int TS = orc_tile(cam);
  const int tilesX = (cam->W + TS - 1) / TS;
  const int tilesY = (cam->H + TS - 1) / TS;
  const REAL W = (REAL)cam->W, H = (REAL)cam->H;
#pragma omp parallel for schedule(static) if (orc_threads > 1)
  for (int64_t i = 0; i < n; i++) {
    const float* P = params + 16 * i;
    REAL* pr = proj + 16 * i;
    for (int k = 0; k < 16; k++) pr[k] = 0;
    rect[4 * i + 0] = rect[4 * i + 1] = rect[4 * i + 2] = rect[4 * i + 3] = 0;
    cnt[i] = 0;
    GEOM g;
    FN(compute_geom)(cam, P, &g);
    if (!g.ok) continue;
    const REAL detV = FMAX(g.va * g.vc - g.vb * g.vb, RL(0.0));
    const REAL ad = g.va + RL(0.3);
    const REAL cd = g.vc + RL(0.3);
    const REAL detVd = ad * cd - g.vb * g.vb;
    if (detVd < RL(1e-8)) continue;
    const REAL comp = SQRT(FMAX(detV / detVd, RL(0.0)));
    const REAL opa = FN(sigm)(RL(P[13]));
    if (opa * comp < A_MIN) continue;
    const REAL f = cam->f;
    const REAL mx = f * g.pc[0] / g.pc[2] + RL(cam->cx);
    const REAL my = f * g.pc[1] / g.pc[2] + RL(cam->cy);
    const REAL mid = RL(0.5) * (ad + cd);
    const REAL disc = SQRT(FMAX(mid * mid - detVd, RL(0.0)));
    const REAL eMax = FMIN(E_CUT, LOG(FMAX(opa * comp / A_MIN, RL(1.0001))));
    const REAL rad = FMIN(SQRT(RL(2.0) * eMax) * SQRT(mid + disc), RADCL * W);
    if (mx + rad < RL(0.0) || my + rad < RL(0.0) || mx - rad > W || my - rad > H) continue;
    const REAL inv = RL(1.0) / detVd;
    pr[0] = mx; pr[1] = my; pr[2] = g.pc[2];
    pr[3] = cd * inv; pr[4] = (-g.vb) * inv; pr[5] = ad * inv;
    pr[6] = comp; pr[7] = opa;
    REAL col[3] = {FN(sigm)(RL(P[10])), FN(sigm)(RL(P[11])), FN(sigm)(RL(P[12]))};
    if (K > 0) {
      /* shaders.js:281-293 */
      const REAL un[3] = {RL(P[0]) - RL(cam->camPos[0]), RL(P[1]) - RL(cam->camPos[1]), RL(P[2]) - RL(cam->camPos[2])};
      const REAL ul = FMAX(SQRT(FN(dot3)(un, un)), RL(1e-9));
      const REAL v[3] = {un[0] / ul, un[1] / ul, un[2] / ul};
      REAL Y[15];
      FN(sh_basis)(o->shDeg, v, Y, 0);
      const int aK = FN(sh_active_k)(o->shDeg, cam->shActiveDeg);
      const float* S = sh + (int64_t)3 * K * i;
      for (int k = 0; k < aK; k++) {
        col[0] = col[0] + RL(S[k]) * Y[k];
        col[1] = col[1] + RL(S[K + k]) * Y[k];
        col[2] = col[2] + RL(S[2 * K + k]) * Y[k];
      }
      for (int c = 0; c < 3; c++) col[c] = FMAX(col[c], RL(0.0));
    }
    pr[8] = col[0]; pr[9] = col[1]; pr[10] = col[2];
    pr[11] = 1;
    pr[12] = g.va; pr[13] = g.vb; pr[14] = g.vc; pr[15] = rad;
    const REAL TF = (REAL)TS;   /* a power of two: the divisions are exact */
    const int tx0 = (int)FN(clampr)(FLOOR((mx - rad) / TF), RL(0.0), (REAL)(tilesX - 1));
    const int tx1 = (int)FN(clampr)(FLOOR((mx + rad) / TF), RL(0.0), (REAL)(tilesX - 1));
    const int ty0 = (int)FN(clampr)(FLOOR((my - rad) / TF), RL(0.0), (REAL)(tilesY - 1));
    const int ty1 = (int)FN(clampr)(FLOOR((my + rad) / TF), RL(0.0), (REAL)(tilesY - 1));
    rect[4 * i + 0] = tx0; rect[4 * i + 1] = tx1; rect[4 * i + 2] = ty0; rect[4 * i + 3] = ty1;
    cnt[i] = (uint32_t)((tx1 - tx0 + 1) * (ty1 - ty0 + 1));
  }
}

/* shaders.js:378-500: the reference scatters (depth bits, id) into tile
 * segments and bitonic-sorts each by (key, id). The same total order is
 * (tile, depth, id); positive depths order like their f32 bits. */
typedef struct { uint32_t tile; REAL depth; uint32_t id; } KEYREC;
static int FN(key_cmp)(const void* pa, const void* pb) {
  const KEYREC* a = (const KEYREC*)pa; const KEYREC* b = (const KEYREC*)pb;
  if (a->tile != b->tile) return a->tile < b->tile ? -1 : 1;
  if (a->depth != b->depth) return a->depth < b->depth ? -1 : 1;
  if (a->id != b->id) return a->id < b->id ? -1 : 1;
  return 0;
}

int64_t FN(orc_bin_sort)(int64_t n, const REAL* proj, const int32_t* rect,
                         const uint32_t* cnt, int tilesX, int numTiles,
                         uint32_t* ids, uint32_t* tile_start, uint64_t* keys) {
  const int64_t K = orc_count_keys(n, cnt);
  KEYREC* recs = (KEYREC*)malloc(sizeof(KEYREC) * (size_t)(K > 0 ? K : 1));
  int64_t w = 0;
  for (int64_t i = 0; i < n; i++) {
    if (!cnt[i]) continue;
    const int32_t* r = rect + 4 * i;
    for (int ty = r[2]; ty <= r[3]; ty++)
      for (int tx = r[0]; tx <= r[1]; tx++) {
        recs[w].tile = (uint32_t)(ty * tilesX + tx);
        recs[w].depth = proj[16 * i + 2];
        recs[w].id = (uint32_t)i;
        w++;
      }
  }
  qsort(recs, (size_t)K, sizeof(KEYREC), FN(key_cmp));
  for (int t = 0; t <= numTiles; t++) tile_start[t] = 0;
  for (int64_t k = 0; k < K; k++) {
    ids[k] = recs[k].id;
    tile_start[recs[k].tile + 1]++;
    if (keys) {
      const float df = (float)recs[k].depth;
      uint32_t db; memcpy(&db, &df, 4);
      keys[k] = ((uint64_t)recs[k].tile << 32) | db;
    }
  }
  for (int t = 0; t < numTiles; t++) tile_start[t + 1] += tile_start[t];
  free(recs);
  return K;
}

/* shaders.js:544-585 forward (per pixel, front to back) */
void FN(orc_render_fwd)(const orc_opts* o, const orc_cam* cam, const REAL* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        REAL* img, REAL* Tfin, uint32_t* endOut) {
  const REAL E_CUT = o->eCut, A_MIN = o->aMin;
  const int TS = orc_tile(cam);
  const int W = cam->W, H = cam->H, tilesX = (W + TS - 1) / TS;
#pragma omp parallel for collapse(2) schedule(dynamic, 64) if (orc_threads > 1)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const int tile = (y / TS) * tilesX + (x / TS);
      const uint32_t segS = tile_start[tile], segE = tile_start[tile + 1];
      const REAL px = (REAL)x + RL(0.5), py = (REAL)y + RL(0.5);
      REAL T = 1, C[3] = {0, 0, 0};
      uint32_t end = segS;
      for (uint32_t k = segS; k < segE; k++) {
        end = k + 1;
        const REAL* pr = proj + 16 * (int64_t)ids[k];
        const REAL dx = px - pr[0], dy = py - pr[1];
        const REAL t2 = FMA(pr[5] * dy, dy, (pr[3] * dx) * dx);
        const REAL e = FMAX(FMA(RL(0.5), t2, (pr[4] * dx) * dy), RL(0.0));
        if (e > E_CUT) continue;
        const REAL araw = pr[7] * pr[6] * EXP_RASTER(-e);
        const REAL alpha = FMIN(RL(0.99), araw);
        if (alpha < A_MIN) continue;
        const REAL w = T * alpha;
        for (int c = 0; c < 3; c++) C[c] = FMA(w, pr[8 + c], C[c]);
        T = T * (RL(1.0) - alpha);
        if (T < RL(1e-4)) break;
      }
      const int64_t pi = (int64_t)y * W + x;
      for (int c = 0; c < 3; c++) img[4 * pi + c] = C[c] + T * RL(cam->bg[c]);
      img[4 * pi + 3] = RL(1.0) - T;
      Tfin[pi] = T;
      endOut[pi] = end;
    }
}

/* shaders.js:589-618 loss (Charbonnier, delta 0.03). expFix / expSum
 * (optional, 2 each) take the exposure gradients of :614-616: slot 0
 * d/d(log gain) = dot(eg, C) * gain, slot 1 d/d(bias) = sum(eg), with
 * eg = err / root per channel. */
void FN(orc_loss_charb)(const orc_cam* cam, const REAL* img, const uint32_t* tgt,
                        REAL* gC, uint8_t* valid, double* stats, int32_t* expFix, double* expSum) {
  const int64_t npx = (int64_t)cam->W * cam->H;
  const REAL gain = cam->gain, bias = cam->bias;
  stats[0] = stats[1] = stats[2] = 0;
  for (int64_t pi = 0; pi < npx; pi++) {
    const uint32_t pk = tgt[pi];
    gC[3 * pi] = gC[3 * pi + 1] = gC[3 * pi + 2] = 0;
    valid[pi] = 0;
    if ((pk >> 24) == 0) continue;
    valid[pi] = 1;
    REAL err[3], root[3], eg[3];
    for (int c = 0; c < 3; c++) {
      const REAL tc = (REAL)((pk >> (8 * c)) & 0xffu) / RL(255.0);
      err[c] = (gain * img[4 * pi + c] + bias) - tc;
      root[c] = SQRT(err[c] * err[c] + RL(0.0009));
      eg[c] = err[c] / root[c];
      gC[3 * pi + c] = gain * eg[c];
    }
    if (expFix || expSum) {
      const REAL col[3] = {img[4 * pi], img[4 * pi + 1], img[4 * pi + 2]};
      FN(cam_add)(expFix, expSum, 0, FN(dot3)(eg, col) * gain);
      FN(cam_add)(expFix, expSum, 1, (eg[0] + eg[1]) + eg[2]);
    }
    stats[0] += (double)FN(dot3)(err, err);
    stats[1] += (double)(((root[0] + root[1]) + root[2]) - RL(0.09));
    stats[2] += 1;
  }
}

/* shaders.js:620-700 backward (per pixel, back to front from `end`) */
void FN(orc_render_bwd)(const orc_opts* o, const orc_cam* cam, const REAL* proj,
                        const uint32_t* ids, const uint32_t* tile_start,
                        const REAL* Tfin, const uint32_t* endIn,
                        const REAL* gCin, const uint8_t* valid,
                        int32_t* gfix, double* gsum) {
  const REAL E_CUT = o->eCut, A_MIN = o->aMin;
  const int TS = orc_tile(cam);
  const int W = cam->W, H = cam->H, tilesX = (W + TS - 1) / TS;
  const int par = orc_threads > 1;
#pragma omp parallel for collapse(2) schedule(dynamic, 64) if (par)
  for (int y = 0; y < H; y++)
    for (int x = 0; x < W; x++) {
      const int64_t pi = (int64_t)y * W + x;
      if (!valid[pi]) continue;
      const int tile = (y / TS) * tilesX + (x / TS);
      const uint32_t segS = tile_start[tile];
      const REAL px = (REAL)x + RL(0.5), py = (REAL)y + RL(0.5);
      const REAL gC[3] = {gCin[3 * pi], gCin[3 * pi + 1], gCin[3 * pi + 2]};
      /* The colour-behind sum S (shaders.js:627, :686) enters only through
       * dot(gC, S), so that dot product GS is carried instead of S:
       * GS starts at dot(gC, bg) * T_final and grows by dot(gC, c) * alpha * Tb. */
      REAL Ta = Tfin[pi];
      const REAL gbg = FMA(gC[2], RL(cam->bg[2]), FMA(gC[1], RL(cam->bg[1]), gC[0] * RL(cam->bg[0])));
      REAL GS = gbg * Ta;
      for (uint32_t kk = endIn[pi]; kk > segS; kk--) {
        const uint32_t gid = ids[kk - 1];
        const REAL* pr = proj + 16 * (int64_t)gid;
        const REAL dx = px - pr[0], dy = py - pr[1];
        const REAL cA = pr[3], cB = pr[4], cC = pr[5];
        const REAL t2 = FMA(cC * dy, dy, (cA * dx) * dx);
        const REAL e = FMAX(FMA(RL(0.5), t2, (cB * dx) * dy), RL(0.0));
        if (!(e <= E_CUT)) continue;
        const REAL comp = pr[6], opa = pr[7];
        const REAL G = EXP_RASTER(-e);
        const REAL araw = opa * comp * G;
        const REAL alpha = FMIN(RL(0.99), araw);
        if (!(alpha >= A_MIN)) continue;
        const REAL c[3] = {pr[8], pr[9], pr[10]};
        /* Tb = Ta / (1 - alpha) and S / (1 - alpha) (shaders.js:656, :660),
         * both through one reciprocal of (1 - alpha);
         * galpha = dot(gC, c Tb - S/(1-alpha)) = dot(gC, c) Tb - GS/(1-alpha) */
        const REAL rom = RL(1.0) / (RL(1.0) - alpha);
        const REAL Tb = Ta * rom;
        const REAL aT = alpha * Tb;
        const REAL gcol = FMA(gC[2], c[2], FMA(gC[1], c[1], gC[0] * c[0]));
        REAL galpha = FMA(gcol, Tb, -(GS * rom));
        if (araw > RL(0.99)) galpha = 0;
        const REAL ga = galpha * araw;
        /* cnorm: shaders.js:674-678 */
        const REAL cva = pr[12] + RL(0.3), cvc = pr[14] + RL(0.3);
        const REAL cmid = RL(0.5) * (cva + cvc);
        const REAL lmax = cmid + SQRT(FMAX(cmid * cmid - (cva * cvc - pr[13] * pr[13]), RL(0.0)));
        const REAL cnorm = RL(1.0) / (RL(1.0) + lmax);
        REAL v[10];
        v[0] = ga * (cA * dx + cB * dy);
        v[1] = ga * (cB * dx + cC * dy);
        /* -ga * 0.5 * d.x * d.x * cnorm etc. (shaders.js:679-681) through the
         * shared factor gq = ga * (-0.5 cnorm); the middle term is 2 gq dx dy
         * (the doubling is exact) */
        const REAL gq = ga * (RL(-0.5) * cnorm);
        const REAL gqdx = gq * dx;
        v[2] = gqdx * dx;
        v[3] = RL(2.0) * (gqdx * dy);
        v[4] = (gq * dy) * dy;
        v[5] = galpha * opa * G;
        v[6] = ga * (RL(1.0) - opa);
        v[7] = gC[0] * aT; v[8] = gC[1] * aT; v[9] = gC[2] * aT;
        if (gfix) {
          /* shaders.js:529-541: i32(clamp(v * 16384, -1e9, 1e9)), wrapping sums */
          for (int s = 0; s < 10; s++) {
            const float fv = (float)v[s] * 16384.0f;
            const float cl = fminf(fmaxf(fv, -1.0e9f), 1.0e9f);
            const int32_t iv = (int32_t)cl;
            uint32_t* dst = (uint32_t*)&gfix[10 * (int64_t)gid + s];
            if (par) {
#pragma omp atomic
              *dst += (uint32_t)iv;
            } else {
              *dst += (uint32_t)iv;
            }
          }
        } else {
          for (int s = 0; s < 10; s++) {
            double* dst = &gsum[10 * (int64_t)gid + s];
            if (par) {
#pragma omp atomic
              *dst += (double)v[s];
            } else {
              *dst += (double)v[s];
            }
          }
        }
        GS = FMA(gcol, aT, GS);
        Ta = Tb;
      }
    }
}

/* shaders.js:722-948 makeChainSrc. camFix / camSum (optional, 7 each) take
 * the camera terms of :834-858: rotation dw (0-2), translation dt = dpc (3-5),
 * shared log-focal (6). */
void FN(orc_chain)(const orc_opts* o, const orc_cam* cam, int64_t n,
                   const float* params, const float* sh, const REAL* proj,
                   const int32_t* gfix, const double* gsum,
                   REAL* gradF, REAL* shGrad, int32_t* camFix, double* camSum) {
  const int K = o->shDeg > 0 ? (o->shDeg + 1) * (o->shDeg + 1) - 1 : 0;
  const REAL AREG = o->anisoReg;
  REAL R[9];
  for (int k = 0; k < 9; k++) R[k] = cam->R[k];
  for (int64_t i = 0; i < n; i++) {
    const float* P = params + 16 * i;
    const REAL* pr = proj + 16 * i;
    REAL* gF = gradF + 16 * i;
    REAL gp[10];
    for (int k = 0; k < 10; k++) {
      if (gfix) gp[k] = (REAL)gfix[10 * i + k] / RL(16384.0);
      else gp[k] = (REAL)gsum[10 * i + k];
    }
    {
      const REAL cva = pr[12] + RL(0.3), cvc = pr[14] + RL(0.3);
      const REAL cmid = RL(0.5) * (cva + cvc);
      const REAL lmax = cmid + SQRT(FMAX(cmid * cmid - (cva * cvc - pr[13] * pr[13]), RL(0.0)));
      const REAL cden = RL(1.0) + lmax;
      gp[2] = gp[2] * cden; gp[3] = gp[3] * cden; gp[4] = gp[4] * cden;
    }
    for (int k = 0; k < 16; k++) gF[k] = 0;
    if (K > 0) for (int k = 0; k < 3 * K; k++) shGrad[3 * K * i + k] = 0;
    if (pr[11] <= RL(0.0)) continue;
    GEOM g;
    FN(compute_geom)(cam, P, &g);
    if (!g.ok) continue;
    const REAL f = cam->f;
    const REAL z = g.pc[2];
    const REAL iz = RL(1.0) / z;

    /* conic -> dilated 2D covariance (:759-777) */
    const REAL ad = g.va + RL(0.3), cd = g.vc + RL(0.3);
    const REAL detVd = ad * cd - g.vb * g.vb;
    const REAL inv = RL(1.0) / detVd;
    const REAL kA = cd * inv, kB = (-g.vb) * inv, kC = ad * inv;
    const REAL g11 = gp[2], g12 = RL(0.5) * gp[3], g22 = gp[4];
    const REAL m100 = g11 * kA + g12 * kB, m101 = g11 * kB + g12 * kC;
    const REAL m110 = g12 * kA + g22 * kB, m111 = g12 * kB + g22 * kC;
    const REAL p00 = kA * m100 + kB * m110, p01 = kA * m101 + kB * m111;
    const REAL p10 = kB * m100 + kC * m110, p11 = kB * m101 + kC * m111;
    REAL gva = -p00, gvb = -(p01 + p10), gvc = -p11;

    /* comp -> raw 2D covariance (:780-788) */
    const REAL detV = FMAX(g.va * g.vc - g.vb * g.vb, RL(0.0));
    const REAL comp = SQRT(FMAX(detV / detVd, RL(0.0)));
    if (comp > RL(1e-4)) {
      const REAL gcomp = gp[5];
      const REAL denom = RL(2.0) * comp * detVd;
      gva = gva + gcomp * (g.vc - comp * comp * cd) / denom;
      gvc = gvc + gcomp * (g.va - comp * comp * ad) / denom;
      gvb = gvb + gcomp * (RL(2.0) * g.vb * (comp * comp - RL(1.0))) / denom;
    }

    /* V = T Sigma T^T backward (:791-810) */
    const REAL h11 = gva, h12 = RL(0.5) * gvb, h22 = gvc;
    REAL u0[3], u1[3];
    for (int k = 0; k < 3; k++) { u0[k] = h11 * g.t0[k] + h12 * g.t1[k]; u1[k] = h12 * g.t0[k] + h22 * g.t1[k]; }
    const REAL* t0 = g.t0; const REAL* t1 = g.t1;
    const REAL dS00 = t0[0] * u0[0] + t1[0] * u1[0];
    const REAL dS01 = RL(0.5) * ((t0[0] * u0[1] + t1[0] * u1[1]) + (t0[1] * u0[0] + t1[1] * u1[0]));
    const REAL dS02 = RL(0.5) * ((t0[0] * u0[2] + t1[0] * u1[2]) + (t0[2] * u0[0] + t1[2] * u1[0]));
    const REAL dS11 = t0[1] * u0[1] + t1[1] * u1[1];
    const REAL dS12 = RL(0.5) * ((t0[1] * u0[2] + t1[1] * u1[2]) + (t0[2] * u0[1] + t1[2] * u1[1]));
    const REAL dS22 = t0[2] * u0[2] + t1[2] * u1[2];
    REAL sT0[3], sT1[3];
    sT0[0] = (g.s00 * t0[0] + g.s01 * t0[1]) + g.s02 * t0[2];
    sT0[1] = (g.s01 * t0[0] + g.s11 * t0[1]) + g.s12 * t0[2];
    sT0[2] = (g.s02 * t0[0] + g.s12 * t0[1]) + g.s22 * t0[2];
    sT1[0] = (g.s00 * t1[0] + g.s01 * t1[1]) + g.s02 * t1[2];
    sT1[1] = (g.s01 * t1[0] + g.s11 * t1[1]) + g.s12 * t1[2];
    sT1[2] = (g.s02 * t1[0] + g.s12 * t1[1]) + g.s22 * t1[2];
    REAL dT0[3], dT1[3];
    for (int k = 0; k < 3; k++) {
      dT0[k] = RL(2.0) * (h11 * sT0[k] + h12 * sT1[k]);
      dT1[k] = RL(2.0) * (h12 * sT0[k] + h22 * sT1[k]);
    }

    /* T = J W (:813-822) */
    const REAL dJ00 = FN(dot3)(dT0, R + 0);
    const REAL dJ02 = FN(dot3)(dT0, R + 6);
    const REAL dJ11 = FN(dot3)(dT1, R + 3);
    const REAL dJ12 = FN(dot3)(dT1, R + 6);
    REAL dpc[3] = {0, 0, 0};
    dpc[0] = dpc[0] + dJ02 * ((-f) * iz * iz);
    dpc[1] = dpc[1] + dJ12 * ((-f) * iz * iz);
    dpc[2] = dpc[2] + ((dJ00 + dJ11) * ((-f) * iz * iz)
           + dJ02 * (RL(2.0) * f * g.pc[0] * iz * iz * iz))
           + dJ12 * (RL(2.0) * f * g.pc[1] * iz * iz * iz);
    /* mean path (:825-827) */
    dpc[0] = dpc[0] + gp[0] * f * iz;
    dpc[1] = dpc[1] + gp[1] * f * iz;
    dpc[2] = dpc[2] + (-f) * (gp[0] * g.pc[0] + gp[1] * g.pc[1]) * iz * iz;
    /* camera pose (p_c = exp(w^) R p + t) and shared log-focal (:834-858) */
    if (camFix || camSum) {
      const REAL rp[3] = {g.pc[0] - RL(cam->t[0]), g.pc[1] - RL(cam->t[1]), g.pc[2] - RL(cam->t[2])};
      REAL dw[3], cs[3] = {0, 0, 0};
      FN(cross3)(rp, dpc, dw);
      const REAL fiz = f * iz;
      const REAL bx = (((-f) * g.pc[0]) * iz) * iz;
      const REAL by = (((-f) * g.pc[1]) * iz) * iz;
      for (int k = 0; k < 3; k++) {
        /* column k of W against (dW0[k], dW1[k], dW2[k]) */
        const REAL wc[3] = {R[k], R[3 + k], R[6 + k]};
        const REAL dv[3] = {fiz * dT0[k], fiz * dT1[k], bx * dT0[k] + by * dT1[k]};
        REAL c[3];
        FN(cross3)(wc, dv, c);
        for (int a = 0; a < 3; a++) cs[a] = k == 0 ? c[a] : cs[a] + c[a];
      }
      for (int a = 0; a < 3; a++) dw[a] = dw[a] + cs[a];
      const REAL dlogf = (((((gp[0] * f) * g.pc[0]) * iz + ((gp[1] * f) * g.pc[1]) * iz)
                          + (dJ00 + dJ11) * fiz) + dJ02 * bx) + dJ12 * by;
      for (int a = 0; a < 3; a++) FN(cam_add)(camFix, camSum, a, dw[a]);
      for (int a = 0; a < 3; a++) FN(cam_add)(camFix, camSum, 3 + a, dpc[a]);
      FN(cam_add)(camFix, camSum, 6, dlogf);
    }
    /* W^T dpc (:830-832) */
    gF[0] = (R[0] * dpc[0] + R[3] * dpc[1]) + R[6] * dpc[2];
    gF[1] = (R[1] * dpc[0] + R[4] * dpc[1]) + R[7] * dpc[2];
    gF[2] = (R[2] * dpc[0] + R[5] * dpc[1]) + R[8] * dpc[2];

    /* Sigma = M M^T backward (:861-884) */
    REAL m0[3], m1[3], m2[3], dM0[3], dM1[3], dM2[3];
    for (int k = 0; k < 3; k++) { m0[k] = g.r0[k] * g.s[k]; m1[k] = g.r1[k] * g.s[k]; m2[k] = g.r2[k] * g.s[k]; }
    for (int k = 0; k < 3; k++) {
      dM0[k] = RL(2.0) * ((dS00 * m0[k] + dS01 * m1[k]) + dS02 * m2[k]);
      dM1[k] = RL(2.0) * ((dS01 * m0[k] + dS11 * m1[k]) + dS12 * m2[k]);
      dM2[k] = RL(2.0) * ((dS02 * m0[k] + dS12 * m1[k]) + dS22 * m2[k]);
    }
    REAL dsv[3];
    for (int k = 0; k < 3; k++) dsv[k] = (dM0[k] * g.r0[k] + dM1[k] * g.r1[k]) + dM2[k] * g.r2[k];
    REAL ls[3];
    for (int k = 0; k < 3; k++) ls[k] = FN(clampr)(RL(P[3 + k]), RL(-12.0), RL(6.0));
    const REAL mls = ((ls[0] + ls[1]) + ls[2]) / RL(3.0);
    for (int k = 0; k < 3; k++) gF[3 + k] = dsv[k] * g.s[k] + AREG * (ls[k] - mls);

    /* quaternion backward (:887-904) */
    REAL dR0[3], dR1[3], dR2[3];
    for (int k = 0; k < 3; k++) { dR0[k] = dM0[k] * g.s[k]; dR1[k] = dM1[k] * g.s[k]; dR2[k] = dM2[k] * g.s[k]; }
    const REAL qw = g.q[0], qx = g.q[1], qy = g.q[2], qz = g.q[3];
    const REAL gw = RL(2.0) * (((((dR0[1] * (-qz) + dR0[2] * qy) + dR1[0] * qz) + dR1[2] * (-qx)) + dR2[0] * (-qy)) + dR2[1] * qx);
    const REAL gx = RL(2.0) * (((((((dR0[1] * qy + dR0[2] * qz) + dR1[0] * qy) + dR1[1] * (RL(-2.0) * qx)) + dR1[2] * (-qw)) + dR2[0] * qz) + dR2[1] * qw) + dR2[2] * (RL(-2.0) * qx));
    const REAL gy = RL(2.0) * (((((((dR0[0] * (RL(-2.0) * qy) + dR0[1] * qx) + dR0[2] * qw) + dR1[0] * qx) + dR1[2] * qz) + dR2[0] * (-qw)) + dR2[1] * qz) + dR2[2] * (RL(-2.0) * qy));
    const REAL gz = RL(2.0) * (((((((dR0[0] * (RL(-2.0) * qz) + dR0[1] * (-qw)) + dR0[2] * qx) + dR1[0] * qw) + dR1[1] * (RL(-2.0) * qz)) + dR1[2] * qy) + dR2[0] * qx) + dR2[1] * qy);
    const REAL gq[4] = {gw, gx, gy, gz};
    const REAL qr[4] = {P[6], P[7], P[8], P[9]};
    const REAL ql = FMAX(SQRT(((qr[0] * qr[0] + qr[1] * qr[1]) + qr[2] * qr[2]) + qr[3] * qr[3]), RL(1e-6));
    const REAL qdg = ((g.q[0] * gq[0] + g.q[1] * gq[1]) + g.q[2] * gq[2]) + g.q[3] * gq[3];
    for (int k = 0; k < 4; k++) gF[6 + k] = (gq[k] - g.q[k] * qdg) / ql;

    /* colour: DC sigmoid chain + SH bands + position via view dir (:907-947) */
    const REAL sCol[3] = {FN(sigm)(RL(P[10])), FN(sigm)(RL(P[11])), FN(sigm)(RL(P[12]))};
    REAL dRGB[3] = {gp[7], gp[8], gp[9]};
    if (K > 0) {
      const REAL un[3] = {RL(P[0]) - RL(cam->camPos[0]), RL(P[1]) - RL(cam->camPos[1]), RL(P[2]) - RL(cam->camPos[2])};
      const REAL ulen = FMAX(SQRT(FN(dot3)(un, un)), RL(1e-9));
      const REAL v[3] = {un[0] / ulen, un[1] / ulen, un[2] / ulen};
      REAL Y[15], D[45];
      FN(sh_basis)(o->shDeg, v, Y, D);
      const int aK = FN(sh_active_k)(o->shDeg, cam->shActiveDeg);
      const float* S = sh + (int64_t)3 * K * i;
      REAL col[3] = {sCol[0], sCol[1], sCol[2]};
      for (int k = 0; k < aK; k++) {
        col[0] = col[0] + RL(S[k]) * Y[k];
        col[1] = col[1] + RL(S[K + k]) * Y[k];
        col[2] = col[2] + RL(S[2 * K + k]) * Y[k];
      }
      for (int c = 0; c < 3; c++) dRGB[c] = dRGB[c] * (col[c] > RL(0.0) ? RL(1.0) : RL(0.0));
      REAL gv[3] = {0, 0, 0};
      REAL* sG = shGrad + (int64_t)3 * K * i;
      for (int k = 0; k < aK; k++) {
        sG[k] = dRGB[0] * Y[k];
        sG[K + k] = dRGB[1] * Y[k];
        sG[2 * K + k] = dRGB[2] * Y[k];
        const REAL w = (dRGB[0] * RL(S[k]) + dRGB[1] * RL(S[K + k])) + dRGB[2] * RL(S[2 * K + k]);
        for (int c = 0; c < 3; c++) gv[c] = gv[c] + w * D[3 * k + c];
      }
      const REAL vg = FN(dot3)(v, gv);
      for (int c = 0; c < 3; c++) gF[c] = gF[c] + (gv[c] - v[c] * vg) / ulen;
    }
    for (int c = 0; c < 3; c++) gF[10 + c] = dRGB[c] * sCol[c] * (RL(1.0) - sCol[c]);
    gF[13] = gp[6];
  }
}

#undef CAT2
#undef CAT
#undef FN
#undef GEOM
#undef KEYREC
#undef RL
