// gs_update.hip — fused per-Gaussian update: the reference's `chain`
// (shaders.js:704-948), `adam` (:984-1036) and `sh-adam` (:952-982) passes in
// one kernel.
//
// Per Gaussian it reads the screen-space gradient sums (and clears them, as
// chain does at :727-732), recomputes the projection geometry, chains the
// gradient to the 14 parameters and the SH coefficients in registers, and
// applies Adam to params/m/v and sh/m/v. The parameter gradients never
// touch HBM (the reference writes and re-reads gradF and shGrad), so the pass
// moves params, SH and both moment buffers once in each direction plus 48 B
// of accumulators: ~1.06 KB per Gaussian at SH degree 2, HBM-bound.
//
// Invisible Gaussians still take the Adam step with a zero gradient (plus the
// opacity regulariser), like the reference, whose Adam runs over all n.
//
// Precision modes (Prec in gs_device.h): in mode 1 the SH coefficients and
// all Adam moments are stored in bf16 (exact decode, stochastic rounding on
// store), which takes the pass from ~1.06 KB to ~0.64 KB per Gaussian at SH
// degree 2; in modes 2 and 3 only the SH coefficients are stored in 16 bits
// (bf16 or fp16, stochastic rounding) and the moments stay f32 (~0.92 KB).
// Parameters, gradients and all arithmetic stay f32.
#include "gs_kernels.h"

namespace gs {

// Screen-space gradient sums of Gaussian i: loaded, decoded, and cleared in
// place as the reference's chain clears them (shaders.js:727-732). The
// steps are separate so that a batched update can load the next view's sums
// while it chains the current view's.
struct RawSums {
  int4 a0, a1, a2;
};

__device__ __forceinline__ RawSums load_sums(const void* accum, int64_t i) {
  const int4* acc = (const int4*)accum + 3 * i;
  return RawSums{acc[0], acc[1], acc[2]};
}

__device__ __forceinline__ void clear_sums(void* accum, int64_t i) {
  int4* acc = (int4*)accum + 3 * i;
  const int4 z = make_int4(0, 0, 0, 0);
  acc[0] = z; acc[1] = z; acc[2] = z;
}

template <int kAccum>
__device__ __forceinline__ void decode_sums(const RawSums& r, float* gp) {
  const int raw[10] = {r.a0.x, r.a0.y, r.a0.z, r.a0.w, r.a1.x, r.a1.y, r.a1.z, r.a1.w, r.a2.x, r.a2.y};
#pragma unroll
  for (int k = 0; k < 10; k++) gp[k] = kAccum == 0 ? (float)raw[k] / kFixed : __int_as_float(raw[k]);
}

template <int kAccum>
__device__ __forceinline__ void take_sums(void* accum, int64_t i, float* gp) {
  const RawSums r = load_sums(accum, i);
  clear_sums(accum, i);
  decode_sums<kAccum>(r, gp);
}

// Terms of the chain that no camera enters, computed once per Gaussian (a
// batched step chains up to kMaxViews views with them): the shape, the
// anisotropy regulariser's gradient (shaders.js:873-884), the quaternion
// length (:899) and the colour sigmoids (:907-910).
struct ChainFixed {
  GeomShape sp;
  float aniso[3];   // anisoReg * (clamped log-scale - their mean)
  float ql;         // max(|q|, 1e-6)
  float sCol[3];
};

__device__ __forceinline__ void chain_fixed(const AdamHP& hp, const float* p, ChainFixed& cf) {
  compute_shape(p, cf.sp);
  float ls[3];
  for (int k = 0; k < 3; k++) ls[k] = clampf(p[3 + k], -12.0f, 6.0f);
  const float mls = ((ls[0] + ls[1]) + ls[2]) / 3.0f;
  for (int k = 0; k < 3; k++) cf.aniso[k] = hp.anisoReg * (ls[k] - mls);
  cf.ql = fmaxf(sqrtf(((p[6] * p[6] + p[7] * p[7]) + p[8] * p[8]) + p[9] * p[9]), 1e-6f);
  for (int c = 0; c < 3; c++) cf.sCol[c] = gs_sigmoid_clamped(p[10 + c]);
}

// The reference's chain (shaders.js:735-947) for one view, in two halves
// that a batched update runs as separate passes over its views. Both take
// the decoded screen-space sums gp[10] of a Gaussian with keys in the view
// and add the view's gradients to gF (14 parameter slots) and shG (SH
// coefficients), or write them there when `first` (the step's first view
// with keys): each value is the view's own gradient first, so the sums
// round as the views' gradients added in view order.

// Colour half (:906-947): SH coefficient gradients into shG, colour-logit
// gradients into gF[10..12]; dpos gets the view-direction term of the
// position gradient (SH degree > 0), which chain_geometry adds to the
// projection's term before it adds the sum to gF[0..2].
template <int kDeg>
__device__ __forceinline__ void chain_colour(const Cam& cam, const float* sCol, const float* p, const float* sh,
                                             const float* gp, bool first, float* gF, float* shG, float* dpos) {
  constexpr int K = ShDims<kDeg>::K;
  auto put = [first](float& d, float v) { d = first ? v : d + v; };
  Geom g;
  if (!geom_position(cam, p, g)) {   // a zero gradient (a Gaussian behind the near plane has no keys)
    for (int c = 0; c < 3; c++) put(gF[10 + c], 0.0f);
    for (int k = 0; k < 3 * K; k++) put(shG[k], 0.0f);
    return;
  }
  float dRGB[3] = {gp[7], gp[8], gp[9]};
  if constexpr (kDeg > 0) {
    float v[3];
    const float ulen = view_dir(cam, p, v);
    float Y[15], D[45];
    sh_basis<kDeg, true>(v, Y, D);
    const int aK = cam.shActiveK;
    float col[3] = {sCol[0], sCol[1], sCol[2]};
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (k < aK) {
        col[0] = col[0] + sh[k] * Y[k];
        col[1] = col[1] + sh[K + k] * Y[k];
        col[2] = col[2] + sh[2 * K + k] * Y[k];
      }
    }
    for (int c = 0; c < 3; c++) dRGB[c] = dRGB[c] * (col[c] > 0.0f ? 1.0f : 0.0f);
    float gv[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < K; k++) {
      if (k < aK) {
        put(shG[k], dRGB[0] * Y[k]);
        put(shG[K + k], dRGB[1] * Y[k]);
        put(shG[2 * K + k], dRGB[2] * Y[k]);
        const float w = (dRGB[0] * sh[k] + dRGB[1] * sh[K + k]) + dRGB[2] * sh[2 * K + k];
        for (int c = 0; c < 3; c++) gv[c] = gv[c] + w * D[3 * k + c];
      } else {
        put(shG[k], 0.0f);
        put(shG[K + k], 0.0f);
        put(shG[2 * K + k], 0.0f);
      }
    }
    const float vg = dot3(v, gv);
    for (int c = 0; c < 3; c++) dpos[c] = (gv[c] - v[c] * vg) / ulen;
  }
  for (int c = 0; c < 3; c++) put(gF[10 + c], dRGB[c] * sCol[c] * (1.0f - sCol[c]));
}

// Geometry half (:735-904): position, log-scale and rotation gradients into
// gF[0..9] (the position's with chain_colour's view-direction term dpos),
// the opacity gradient into gF[13].
template <int kDeg>
__device__ __forceinline__ void chain_geometry(const Cam& cam, const ChainFixed& cf, const float* p, float* gp,
                                               bool first, float* gF, const float* dpos) {
  auto put = [first](float& d, float v) { d = first ? v : d + v; };
  Geom g;
  compute_geom(cam, p, cf.sp, g);
  if (!g.ok) {   // a zero gradient (see chain_colour)
    for (int k = 0; k < 10; k++) put(gF[k], 0.0f);
    put(gF[13], 0.0f);
    return;
  }
  const float cden = 1.0f + conic_lmax(g.va, g.vb, g.vc);
  gp[2] = gp[2] * cden;
  gp[3] = gp[3] * cden;
  gp[4] = gp[4] * cden;
  const float iz = 1.0f / g.pc[2];

  // screen space -> Sigma, T, J and the camera-space position (:759-827)
  ChainScreen cs;
  chain_screen(cam, g, gp, iz, cs);
  const float dS00 = cs.dS00, dS01 = cs.dS01, dS02 = cs.dS02;
  const float dS11 = cs.dS11, dS12 = cs.dS12, dS22 = cs.dS22;
  const float* dpc = cs.dpc;
  const float* R = cam.R;
  float pos[3];
  pos[0] = (R[0] * dpc[0] + R[3] * dpc[1]) + R[6] * dpc[2];
  pos[1] = (R[1] * dpc[0] + R[4] * dpc[1]) + R[7] * dpc[2];
  pos[2] = (R[2] * dpc[0] + R[5] * dpc[1]) + R[8] * dpc[2];
  if constexpr (kDeg > 0)
    for (int c = 0; c < 3; c++) pos[c] = pos[c] + dpos[c];   // (:914-935)
  for (int c = 0; c < 3; c++) put(gF[c], pos[c]);

  // Sigma = M M^T backward (:861-884)
  float m0[3], m1[3], m2[3], dM0[3], dM1[3], dM2[3];
  for (int k = 0; k < 3; k++) { m0[k] = g.r0[k] * g.s[k]; m1[k] = g.r1[k] * g.s[k]; m2[k] = g.r2[k] * g.s[k]; }
  for (int k = 0; k < 3; k++) {
    dM0[k] = 2.0f * ((dS00 * m0[k] + dS01 * m1[k]) + dS02 * m2[k]);
    dM1[k] = 2.0f * ((dS01 * m0[k] + dS11 * m1[k]) + dS12 * m2[k]);
    dM2[k] = 2.0f * ((dS02 * m0[k] + dS12 * m1[k]) + dS22 * m2[k]);
  }
  for (int k = 0; k < 3; k++) {
    const float dsv = (dM0[k] * g.r0[k] + dM1[k] * g.r1[k]) + dM2[k] * g.r2[k];
    put(gF[3 + k], dsv * g.s[k] + cf.aniso[k]);
  }

  // quaternion backward (:887-904)
  float dR0[3], dR1[3], dR2[3];
  for (int k = 0; k < 3; k++) { dR0[k] = dM0[k] * g.s[k]; dR1[k] = dM1[k] * g.s[k]; dR2[k] = dM2[k] * g.s[k]; }
  const float qw = g.q[0], qx = g.q[1], qy = g.q[2], qz = g.q[3];
  const float gw = 2.0f * (((((dR0[1] * (-qz) + dR0[2] * qy) + dR1[0] * qz) + dR1[2] * (-qx)) + dR2[0] * (-qy)) + dR2[1] * qx);
  const float gx = 2.0f * (((((((dR0[1] * qy + dR0[2] * qz) + dR1[0] * qy) + dR1[1] * (-2.0f * qx)) + dR1[2] * (-qw)) + dR2[0] * qz) + dR2[1] * qw) + dR2[2] * (-2.0f * qx));
  const float gy = 2.0f * (((((((dR0[0] * (-2.0f * qy) + dR0[1] * qx) + dR0[2] * qw) + dR1[0] * qx) + dR1[2] * qz) + dR2[0] * (-qw)) + dR2[1] * qz) + dR2[2] * (-2.0f * qy));
  const float gz = 2.0f * (((((((dR0[0] * (-2.0f * qz) + dR0[1] * (-qw)) + dR0[2] * qx) + dR1[0] * qw) + dR1[1] * (-2.0f * qz)) + dR1[2] * qy) + dR2[0] * qx) + dR2[1] * qy);
  const float gq[4] = {gw, gx, gy, gz};
  const float qdg = ((g.q[0] * gq[0] + g.q[1] * gq[1]) + g.q[2] * gq[2]) + g.q[3] * gq[3];
  for (int k = 0; k < 4; k++) put(gF[6 + k], (gq[k] - g.q[k] * qdg) / cf.ql);
  put(gF[13], gp[6]);
}

// Adam over one Gaussian's SH coefficients (the reference's sh-adam pass)
// with gradients shG, in the storage formats of precision mode kPrec; sh
// holds the current values.
template <int kDeg, int kPrec>
__device__ __forceinline__ void adam_sh(const AdamHP& hp, const UpdateBufs& u, int64_t i, const float* sh,
                                        const float* shG) {
  constexpr int NC = ShDims<kDeg>::C;
  constexpr int P4 = ShDims<kDeg>::P4;
  constexpr int Q8 = ShDims<kDeg>::Q8;
  const int64_t cap = u.cap;
  // ---- Adam over the SH coefficients (shaders.js:963-981) ----
  if constexpr (kDeg > 0 && Prec<kPrec>::mom16) {
    // bf16 coefficients and moments (streams 2: coefficient, 3: m, 4: v);
    // the padding halves of the last plane stay zero
    const uint4* SHM16 = (const uint4*)u.SHM;
    const uint4* SHV16 = (const uint4*)u.SHV;
    const uint32_t seed = gs_sr_seed((uint64_t)i, hp.step);
#pragma unroll
    for (int c = 0; c < Q8; c++) {
      float ms[8], vs[8];
      unpack_bf8(SHM16[c * cap + i], ms);
      unpack_bf8(SHV16[c * cap + i], vs);
      // random bits of values 8c .. 8c + 7 of streams 2, 3, 4: one hash per pair
      uint32_t hs[4], hm[4], hv[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        hs[q] = gs_sr_pair(seed, 2u * 32u + 4u * (uint32_t)c + (uint32_t)q);
        hm[q] = gs_sr_pair(seed, 3u * 32u + 4u * (uint32_t)c + (uint32_t)q);
        hv[q] = gs_sr_pair(seed, 4u * 32u + 4u * (uint32_t)c + (uint32_t)q);
      }
      uint32_t bs[8], bm[8], bv[8];
#pragma unroll
      for (int l = 0; l < 8; l++) {
        const int j = 8 * c + l;
        float ss = 0.0f, mm = 0.0f, vn = 0.0f;
        if (j < NC) {
          float gr = shG[j];
          if (!(fabsf(gr) < 1e18f)) gr = 0.0f;
          mm = hp.b1 * ms[l] + (1.0f - hp.b1) * gr;
          vn = hp.b2 * vs[l] + (1.0f - hp.b2) * gr * gr;
          const float mh = mm * hp.rbc1;
          const float vh = vn * hp.rbc2;
          ss = sh[j] - hp.shLr * mh / (sqrtf(vh) + hp.eps);
        }
        const int sh16 = (l & 1) ? 16 : 0;   // value j = 8c + l: half l & 1 of pair l / 2
        bs[l] = gs_f2bf_sr(ss, hs[l >> 1] >> sh16);
        bm[l] = gs_f2bf_sr(mm, hm[l >> 1] >> sh16);
        bv[l] = gs_f2bf_sr(vn, hv[l >> 1] >> sh16);
      }
      ((uint4*)u.SH)[c * cap + i] = pack_bf8(bs);
      ((uint4*)u.SHM)[c * cap + i] = pack_bf8(bm);
      ((uint4*)u.SHV)[c * cap + i] = pack_bf8(bv);
    }
  } else if constexpr (kDeg > 0 && Prec<kPrec>::sh != 0) {
    // 16-bit coefficients (stream 2, as above), f32 moments
    constexpr int F = Prec<kPrec>::sh;
    float4* SHM = (float4*)u.SHM;
    float4* SHV = (float4*)u.SHV;
    float ss[8 * Q8];
#pragma unroll
    for (int j = 0; j < 8 * Q8; j++) ss[j] = 0.0f;
#pragma unroll
    for (int c = 0; c < P4; c++) {
      const float4 m4 = SHM[c * cap + i];
      const float4 v4 = SHV[c * cap + i];
      float ms[4] = {m4.x, m4.y, m4.z, m4.w};
      float vs[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const int j = 4 * c + l;
        if (j >= NC) break;
        float gr = shG[j];
        if (!(fabsf(gr) < 1e18f)) gr = 0.0f;
        const float mm = hp.b1 * ms[l] + (1.0f - hp.b1) * gr;
        const float vn = hp.b2 * vs[l] + (1.0f - hp.b2) * gr * gr;
        ms[l] = mm;
        vs[l] = vn;
        const float mh = mm * hp.rbc1;
        const float vh = vn * hp.rbc2;
        ss[j] = sh[j] - hp.shLr * mh / (sqrtf(vh) + hp.eps);
      }
      SHM[c * cap + i] = make_float4(ms[0], ms[1], ms[2], ms[3]);
      SHV[c * cap + i] = make_float4(vs[0], vs[1], vs[2], vs[3]);
    }
    const uint32_t seed = gs_sr_seed((uint64_t)i, hp.step);
#pragma unroll
    for (int c = 0; c < Q8; c++) {
      uint32_t b[8];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint32_t h = gs_sr_pair(seed, 2u * 32u + 4u * (uint32_t)c + (uint32_t)q);
        b[2 * q] = gs_f_to_16_sr<F>(ss[8 * c + 2 * q], h);
        b[2 * q + 1] = gs_f_to_16_sr<F>(ss[8 * c + 2 * q + 1], h >> 16);
      }
      ((uint4*)u.SH)[c * cap + i] = pack_bf8(b);
    }
  } else if constexpr (kDeg > 0) {
    float4* SH = (float4*)u.SH;
    float4* SHM = (float4*)u.SHM;
    float4* SHV = (float4*)u.SHV;
#pragma unroll
    for (int c = 0; c < P4; c++) {
      const float4 m4 = SHM[c * cap + i];
      const float4 v4 = SHV[c * cap + i];
      float ms[4] = {m4.x, m4.y, m4.z, m4.w};
      float vs[4] = {v4.x, v4.y, v4.z, v4.w};
      float ss[4] = {sh[4 * c], sh[4 * c + 1], sh[4 * c + 2], sh[4 * c + 3]};
#pragma unroll
      for (int l = 0; l < 4; l++) {
        const int j = 4 * c + l;
        if (j >= NC) break;
        float gr = shG[j];
        if (!(fabsf(gr) < 1e18f)) gr = 0.0f;
        const float mm = hp.b1 * ms[l] + (1.0f - hp.b1) * gr;
        const float vn = hp.b2 * vs[l] + (1.0f - hp.b2) * gr * gr;
        ms[l] = mm;
        vs[l] = vn;
        const float mh = mm * hp.rbc1;
        const float vh = vn * hp.rbc2;
        ss[l] = ss[l] - hp.shLr * mh / (sqrtf(vh) + hp.eps);
      }
      SH[c * cap + i] = make_float4(ss[0], ss[1], ss[2], ss[3]);
      SHM[c * cap + i] = make_float4(ms[0], ms[1], ms[2], ms[3]);
      SHV[c * cap + i] = make_float4(vs[0], vs[1], vs[2], vs[3]);
    }
  }
}

// Adam over one Gaussian's 16 parameter slots (the reference's adam pass)
// with gradients gF; p holds the current values and takes the new ones.
template <int kPrec>
__device__ __forceinline__ void adam_params(const AdamHP& hp, const UpdateBufs& u, int64_t i, float* p,
                                            const float* gF) {
  const int64_t cap = u.cap;
  // ---- Adam over the 16 param slots (shaders.js:1000-1035) ----
  // One storage plane at a time (4 f32 or 8 bf16 slots), so only that
  // plane's moments are live.
  auto adam_slot = [&](int j, float& mj, float& vj) {
    const float lr = hp.lr[j];
    if (j >= 14 || lr == 0.0f) return;
    float gr = gF[j];
    if (!(fabsf(gr) < 1e18f)) gr = 0.0f;
    if (j == 13) {
      const float sg = gs_sigmoid_clamped(p[j]);
      gr = gr + hp.reg * sg * (1.0f - sg);
    }
    const float mm = hp.b1 * mj + (1.0f - hp.b1) * gr;
    const float vn = hp.b2 * vj + (1.0f - hp.b2) * gr * gr;
    mj = mm;
    vj = vn;
    const float mh = mm * hp.rbc1;
    const float vh = vn * hp.rbc2;
    float pn = p[j] - lr * mh / (sqrtf(vh) + hp.eps);
    if (j >= 3 && j <= 5) pn = clampf(pn, hp.minLS, hp.maxLS);
    if (j >= 10 && j <= 13) pn = clampf(pn, -hp.maxLogit, hp.maxLogit);
    p[j] = pn;
  };
  if constexpr (Prec<kPrec>::mom16) {
    // moments stored in bf16 with stochastic rounding (streams 0: m, 1: v);
    // slots Adam skips re-encode to their own bits
    uint4* M16 = (uint4*)u.M;
    uint4* V16 = (uint4*)u.V;
    const uint32_t seed = gs_sr_seed((uint64_t)i, hp.step);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      float m[8], vv[8];
      unpack_bf8(M16[h * cap + i], m);
      unpack_bf8(V16[h * cap + i], vv);
      // random bits of slots 8h .. 8h + 7 of streams 0, 1: one hash per pair
      uint32_t hm[4], hv[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        hm[q] = gs_sr_pair(seed, 0u * 32u + 4u * (uint32_t)h + (uint32_t)q);
        hv[q] = gs_sr_pair(seed, 1u * 32u + 4u * (uint32_t)h + (uint32_t)q);
      }
      uint32_t bm[8], bv[8];
#pragma unroll
      for (int l = 0; l < 8; l++) {
        const int j = 8 * h + l;
        adam_slot(j, m[l], vv[l]);
        const int sh16 = (l & 1) ? 16 : 0;
        bm[l] = gs_f2bf_sr(m[l], hm[l >> 1] >> sh16);
        bv[l] = gs_f2bf_sr(vv[l], hv[l >> 1] >> sh16);
      }
      M16[h * cap + i] = pack_bf8(bm);
      V16[h * cap + i] = pack_bf8(bv);
#pragma unroll
      for (int c = 2 * h; c < 2 * h + 2; c++)
        u.P[c * cap + i] = make_float4(p[4 * c], p[4 * c + 1], p[4 * c + 2], p[4 * c + 3]);
    }
  } else {
    float4* M = (float4*)u.M;
    float4* V = (float4*)u.V;
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const float4 m4 = M[c * cap + i], v4 = V[c * cap + i];
      float m[4] = {m4.x, m4.y, m4.z, m4.w};
      float vv[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
      for (int l = 0; l < 4; l++) adam_slot(4 * c + l, m[l], vv[l]);
      u.P[c * cap + i] = make_float4(p[4 * c], p[4 * c + 1], p[4 * c + 2], p[4 * c + 3]);
      M[c * cap + i] = make_float4(m[0], m[1], m[2], m[3]);
      V[c * cap + i] = make_float4(vv[0], vv[1], vv[2], vv[3]);
    }
  }
}

template <int kDeg, int kAccum, bool kDebug, int kPrec>
__global__ __launch_bounds__(kBlock) void k_update(Cam cam, AdamHP hp, UpdateBufs u) {
  constexpr int NC = ShDims<kDeg>::C;
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= u.n) return;
  const int64_t cap = u.cap;

  // Only Gaussians with keys in this step can have nonzero sums (the
  // backward adds to keyed Gaussians only, and every pass leaves the sums
  // cleared), so a culled Gaussian's sums are zero and stay untouched.
  const bool keyed = u.cnt[i] != 0u;
  float p[16];
  load_params(u.P, cap, i, p);
  float gF[14];
#pragma unroll
  for (int k = 0; k < 14; k++) gF[k] = 0.0f;
  float gp[10];
  if (keyed) take_sums<kAccum>(u.accum, i, gp);
  // the colour half of the chain and the SH Adam step first, so the SH
  // registers are free during the geometry half (as in k_update_views)
  float dpos[3] = {0.0f, 0.0f, 0.0f};
  {
    float sh[ShDims<kDeg>::S > 0 ? ShDims<kDeg>::S : 1];
    if constexpr (kDeg > 0) load_sh_any<kDeg, Prec<kPrec>::sh>(u.SH, cap, i, sh);
    float shG[NC > 0 ? NC : 1];
#pragma unroll
    for (int k = 0; k < NC; k++) shG[k] = 0.0f;
    if (keyed) {
      float sCol[3];
      for (int c = 0; c < 3; c++) sCol[c] = gs_sigmoid_clamped(p[10 + c]);
      chain_colour<kDeg>(cam, sCol, p, sh, gp, true, gF, shG, dpos);
    }
    if (kDebug)
      for (int k = 0; k < NC; k++) u.dbgShGrad[(int64_t)NC * i + k] = shG[k];
    adam_sh<kDeg, kPrec>(hp, u, i, sh, shG);
  }
  if (keyed) {
    ChainFixed cf;
    chain_fixed(hp, p, cf);
    chain_geometry<kDeg>(cam, cf, p, gp, true, gF, dpos);
  }
  if (kDebug) {
    for (int k = 0; k < 14; k++) u.dbgGradF[16 * i + k] = gF[k];
    u.dbgGradF[16 * i + 14] = 0.0f;
    u.dbgGradF[16 * i + 15] = 0.0f;
  }
  adam_params<kPrec>(hp, u, i, p, gF);
}

// A batched step's update: each view's sums chained through that view's
// camera, the views' gradients added in view order (the first view with
// keys gives the value, later ones add to it), then one Adam step with the
// step's constants (vb->hp). Gaussians without keys in a view skip it.
//
// Two passes over the views keep the registers of one half of the chain
// live at a time: the colour pass accumulates the SH and colour-logit
// gradients and ends with the SH Adam step; the geometry pass then
// accumulates the position, scale, rotation and opacity gradients and ends
// with the parameter Adam step. Each view's view-direction term of the
// position gradient goes from the colour pass to the geometry pass through
// LDS (dposLds: 3 floats per view and thread, V * 3 * kBlock floats per
// block). In one pass the chain held 244 VGPRs at SH degree 2 and more than
// 256 at degree 3 (1 wave per SIMD); in two, 164 and ~230. The
// camera-independent terms are computed once for all views, the views' key
// counts are read up front, and the geometry pass loads the next view's
// sums while it chains the current one.
template <int kDeg, int kAccum, bool kDebug, int kPrec>
__global__ __launch_bounds__(kBlock) void k_update_views(const ViewBlock* __restrict__ vb, UpdateBufs u) {
  constexpr int NC = ShDims<kDeg>::C;
  extern __shared__ float dposLds[];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= u.n) return;
  const int64_t cap = u.cap;
  const AdamHP hp = vb->hp;
  const int V = vb->count;
  uint32_t keyed = 0;   // bit v: Gaussian i has keys in view v
  for (int v = 0; v < V; v++) keyed |= (gs_global(vb->cnt[v])[i] != 0u ? 1u : 0u) << v;
  // the colour pass reads only the position and the colour logits; the
  // geometry pass loads the parameters again, so the other slots hold no
  // registers meanwhile
  float p[16];
  load_params(u.P, cap, i, p);
  float gF[14];
#pragma unroll
  for (int k = 0; k < 14; k++) gF[k] = 0.0f;
  float sCol[3];
  if (keyed)
    for (int c = 0; c < 3; c++) sCol[c] = gs_sigmoid_clamped(p[10 + c]);

  // ---- colour pass: SH and colour-logit gradients, then the SH Adam step
  {
    float sh[ShDims<kDeg>::S > 0 ? ShDims<kDeg>::S : 1];
    if constexpr (kDeg > 0) load_sh_any<kDeg, Prec<kPrec>::sh>(u.SH, cap, i, sh);
    float shG[NC > 0 ? NC : 1];
#pragma unroll
    for (int k = 0; k < NC; k++) shG[k] = 0.0f;
    bool first = true;
    for (int v = 0; v < V; v++) {
      if (!((keyed >> v) & 1u)) continue;
      // sums 7..9: the last word of the second int4 and the third int4's first two
      const int4* acc = (const int4*)gs_global(vb->accum[v]) + 3 * i;
      const int4 a1 = acc[1], a2 = acc[2];
      float gp[10];
      const int raw[3] = {a1.w, a2.x, a2.y};
      for (int k = 0; k < 3; k++) gp[7 + k] = kAccum == 0 ? (float)raw[k] / kFixed : __int_as_float(raw[k]);
      float dpos[3] = {0.0f, 0.0f, 0.0f};
      chain_colour<kDeg>(vb->cams[v], sCol, p, sh, gp, first, gF, shG, dpos);
      if constexpr (kDeg > 0)
        for (int c = 0; c < 3; c++) dposLds[(3 * v + c) * kBlock + threadIdx.x] = dpos[c];
      first = false;
    }
    if (kDebug)
      for (int k = 0; k < NC; k++) u.dbgShGrad[(int64_t)NC * i + k] = shG[k];
    adam_sh<kDeg, kPrec>(hp, u, i, sh, shG);
  }

  // ---- geometry pass: position, scale, rotation and opacity gradients,
  // then the parameter Adam step
  load_params(u.P, cap, i, p);
  ChainFixed cf;
  if (keyed) chain_fixed(hp, p, cf);
  // view v's sums in cur; view v + 1's load is in flight during view v's chain
  RawSums cur{}, nxt{};
  if (keyed & 1u) cur = load_sums(gs_global(vb->accum[0]), i);
  bool first = true;
  for (int v = 0; v < V; v++) {
    if (v + 1 < V && ((keyed >> (v + 1)) & 1u)) nxt = load_sums(gs_global(vb->accum[v + 1]), i);
    if ((keyed >> v) & 1u) {
      clear_sums(gs_global(vb->accum[v]), i);
      float gp[10];
      decode_sums<kAccum>(cur, gp);
      float dpos[3] = {0.0f, 0.0f, 0.0f};
      if constexpr (kDeg > 0)
        for (int c = 0; c < 3; c++) dpos[c] = dposLds[(3 * v + c) * kBlock + threadIdx.x];
      chain_geometry<kDeg>(vb->cams[v], cf, p, gp, first, gF, dpos);
      first = false;
    }
    cur = nxt;
  }
  if (kDebug) {
    for (int k = 0; k < 14; k++) u.dbgGradF[16 * i + k] = gF[k];
    u.dbgGradF[16 * i + 14] = 0.0f;
    u.dbgGradF[16 * i + 15] = 0.0f;
  }
  adam_params<kPrec>(hp, u, i, p, gF);
}

template <int kDeg, int kPrec>
static void launch_views_deg(hipStream_t st, int accumMode, bool dbg, const ViewBlock* vb, int V,
                             const UpdateBufs& u) {
  const dim3 grid((unsigned)((u.n + kBlock - 1) / kBlock));
  const size_t lds = kDeg > 0 ? (size_t)V * 3 * kBlock * sizeof(float) : 0;   // dposLds
  if (accumMode == 0) {
    if (dbg) k_update_views<kDeg, 0, true, kPrec><<<grid, kBlock, lds, st>>>(vb, u);
    else k_update_views<kDeg, 0, false, kPrec><<<grid, kBlock, lds, st>>>(vb, u);
  } else {
    if (dbg) k_update_views<kDeg, 1, true, kPrec><<<grid, kBlock, lds, st>>>(vb, u);
    else k_update_views<kDeg, 1, false, kPrec><<<grid, kBlock, lds, st>>>(vb, u);
  }
}

template <int kPrec>
static void launch_views_prec(hipStream_t st, int deg, int accumMode, bool dbg, const ViewBlock* vb, int V,
                              const UpdateBufs& u) {
  switch (deg) {
    case 0: launch_views_deg<0, kPrec>(st, accumMode, dbg, vb, V, u); break;
    case 1: launch_views_deg<1, kPrec>(st, accumMode, dbg, vb, V, u); break;
    case 2: launch_views_deg<2, kPrec>(st, accumMode, dbg, vb, V, u); break;
    default: launch_views_deg<3, kPrec>(st, accumMode, dbg, vb, V, u); break;
  }
}

void launch_update_views(hipStream_t st, int deg, int accumMode, int prec, const ViewBlock* vb, int V,
                         const UpdateBufs& u) {
  if (u.n <= 0) return;
  const bool dbg = u.dbgGradF != nullptr;
  switch (prec) {
    case 0: launch_views_prec<0>(st, deg, accumMode, dbg, vb, V, u); break;
    case 1: launch_views_prec<1>(st, deg, accumMode, dbg, vb, V, u); break;
    case 2: launch_views_prec<2>(st, deg, accumMode, dbg, vb, V, u); break;
    default: launch_views_prec<3>(st, deg, accumMode, dbg, vb, V, u); break;
  }
}

template <int kDeg, int kPrec>
static void launch_deg(hipStream_t st, int accumMode, bool dbg, const Cam& cam, const AdamHP& hp,
                       const UpdateBufs& u) {
  const dim3 grid((unsigned)((u.n + kBlock - 1) / kBlock));
  if (accumMode == 0) {
    if (dbg) k_update<kDeg, 0, true, kPrec><<<grid, kBlock, 0, st>>>(cam, hp, u);
    else k_update<kDeg, 0, false, kPrec><<<grid, kBlock, 0, st>>>(cam, hp, u);
  } else {
    if (dbg) k_update<kDeg, 1, true, kPrec><<<grid, kBlock, 0, st>>>(cam, hp, u);
    else k_update<kDeg, 1, false, kPrec><<<grid, kBlock, 0, st>>>(cam, hp, u);
  }
}

template <int kPrec>
static void launch_prec(hipStream_t st, int deg, int accumMode, bool dbg, const Cam& cam, const AdamHP& hp,
                        const UpdateBufs& u) {
  switch (deg) {
    case 0: launch_deg<0, kPrec>(st, accumMode, dbg, cam, hp, u); break;
    case 1: launch_deg<1, kPrec>(st, accumMode, dbg, cam, hp, u); break;
    case 2: launch_deg<2, kPrec>(st, accumMode, dbg, cam, hp, u); break;
    default: launch_deg<3, kPrec>(st, accumMode, dbg, cam, hp, u); break;
  }
}

void launch_update(hipStream_t st, int deg, int accumMode, int prec, const Cam& cam, const AdamHP& hp,
                   const UpdateBufs& u) {
  if (u.n <= 0) return;
  const bool dbg = u.dbgGradF != nullptr;
  switch (prec) {
    case 0: launch_prec<0>(st, deg, accumMode, dbg, cam, hp, u); break;
    case 1: launch_prec<1>(st, deg, accumMode, dbg, cam, hp, u); break;
    case 2: launch_prec<2>(st, deg, accumMode, dbg, cam, hp, u); break;
    default: launch_prec<3>(st, deg, accumMode, dbg, cam, hp, u); break;
  }
}

}  // namespace gs
