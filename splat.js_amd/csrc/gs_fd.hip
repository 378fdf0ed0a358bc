// gs_fd.hip — central finite differences of the training loss in FP64, for
// checking the analytic gradients (the reference's gradCheck procedure,
// gradcheck.js:266-356: perturb one parameter by +-h, rerun the forward pass,
// (L+ - L-) / 2h). This is the only FP64 arithmetic on the GPU.
//
// Only the perturbed Gaussian's footprint changes, so each sample evaluates
// just the tiles of its rectangle (one tile of margin each side): per pixel,
// the tile's sorted key list is composited front to back twice, with the
// Gaussian's FP64 record for +h and for -h (the other Gaussians keep the
// stored f32 records of the current raster), and the pixel's Charbonnier
// loss difference is summed. The sort order is the current raster's: the
// Gaussian stays at its key's place in the tiles it has keys in (the
// analytic backward also treats the order as fixed; a depth step that
// crossed other splats would otherwise add the jumps of the reordering),
// and goes in by its perturbed depth in the margin tiles. Per-pixel rules are
// the forward pass's (shaders.js:556-585): e = max(0, 1/2 (A dx^2 + C dy^2) +
// B dx dy), skip e > eCut, alpha = min(0.99, opa comp e^-e), skip alpha <
// A_MIN, stop after the splat that takes T below 1e-4; loss per valid pixel
// and channel sqrt(err^2 + 0.03^2) - 0.03 (shaders.js:605-611).
#include "gs_kernels.h"

namespace gs {

struct Rec64 {
  double mx, my, A, B, C, oc, c0, c1, c2, depth;
  int ok;
};

__device__ __forceinline__ double sigm64(double x) {
  x = fmin(fmax(x, -9.0), 9.0);
  return 1.0 / (1.0 + exp(-x));
}

// real SH basis of degrees 1..deg at unit v (shaders.js:56-138, including its
// Y[14] = C36 x (x^2 - y^2))
__device__ void sh_basis64(int deg, const double* v, double* Y) {
  const double x = v[0], y = v[1], z = v[2];
  const double xx = x * x, yy = y * y, zz = z * z;
  const double C1 = 0.4886025119029199;
  Y[0] = -C1 * y; Y[1] = C1 * z; Y[2] = -C1 * x;
  if (deg >= 2) {
    Y[3] = 1.0925484305920792 * x * y;
    Y[4] = -1.0925484305920792 * y * z;
    Y[5] = 0.31539156525252005 * (2.0 * zz - xx - yy);
    Y[6] = -1.0925484305920792 * x * z;
    Y[7] = 0.5462742152960396 * (xx - yy);
  }
  if (deg >= 3) {
    Y[8] = -0.5900435899266435 * y * (3.0 * xx - yy);
    Y[9] = 2.890611442640554 * x * y * z;
    Y[10] = -0.4570457994644658 * y * (4.0 * zz - xx - yy);
    Y[11] = 0.3731763325901154 * z * (2.0 * zz - 3.0 * xx - 3.0 * yy);
    Y[12] = -0.4570457994644658 * x * (4.0 * zz - xx - yy);
    Y[13] = 1.445305721320277 * z * (xx - yy);
    Y[14] = -0.5900435899266435 * x * (xx - yy);
  }
}

// The forward-pass record of one Gaussian in FP64 (computeGeom and project,
// shaders.js:176-315), from its 16 parameters and SH coefficients.
__device__ Rec64 project64(const Cam& cam, const Cut& cut, const double* P, const double* sh, int deg) {
  Rec64 r{};
  r.ok = 0;
  double R[9], t[3];
  for (int k = 0; k < 9; k++) R[k] = cam.R[k];
  for (int k = 0; k < 3; k++) t[k] = cam.t[k];
  double pc[3];
  for (int a = 0; a < 3; a++) pc[a] = R[3 * a] * P[0] + R[3 * a + 1] * P[1] + R[3 * a + 2] * P[2] + t[a];
  if (pc[2] < (double)cam.nearp) return r;
  double q[4] = {P[6], P[7], P[8], P[9]};
  const double ql = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (ql < 1e-6) { q[0] = 1.0; q[1] = q[2] = q[3] = 0.0; }
  else for (int k = 0; k < 4; k++) q[k] /= ql;
  const double qw = q[0], qx = q[1], qy = q[2], qz = q[3];
  const double Rq[9] = {1.0 - 2.0 * (qy * qy + qz * qz), 2.0 * (qx * qy - qw * qz), 2.0 * (qx * qz + qw * qy),
                        2.0 * (qx * qy + qw * qz), 1.0 - 2.0 * (qx * qx + qz * qz), 2.0 * (qy * qz - qw * qx),
                        2.0 * (qx * qz - qw * qy), 2.0 * (qy * qz + qw * qx), 1.0 - 2.0 * (qx * qx + qy * qy)};
  double s[3];
  for (int k = 0; k < 3; k++) s[k] = exp(fmin(fmax(P[3 + k], -12.0), 6.0));
  double M[9], S[9];
  for (int a = 0; a < 3; a++)
    for (int k = 0; k < 3; k++) M[3 * a + k] = Rq[3 * a + k] * s[k];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) S[3 * a + b] = M[3 * a] * M[3 * b] + M[3 * a + 1] * M[3 * b + 1] + M[3 * a + 2] * M[3 * b + 2];
  const double f = cam.f, iz = 1.0 / pc[2];
  const double a0 = f * iz, bx = -f * pc[0] * iz * iz, by = -f * pc[1] * iz * iz;
  double t0[3], t1[3];
  for (int k = 0; k < 3; k++) {
    t0[k] = a0 * R[k] + bx * R[6 + k];
    t1[k] = a0 * R[3 + k] + by * R[6 + k];
  }
  double st0[3], st1[3];
  for (int a = 0; a < 3; a++) {
    st0[a] = S[3 * a] * t0[0] + S[3 * a + 1] * t0[1] + S[3 * a + 2] * t0[2];
    st1[a] = S[3 * a] * t1[0] + S[3 * a + 1] * t1[1] + S[3 * a + 2] * t1[2];
  }
  const double va = fmax(t0[0] * st0[0] + t0[1] * st0[1] + t0[2] * st0[2], 0.0);
  const double vb = t1[0] * st0[0] + t1[1] * st0[1] + t1[2] * st0[2];
  const double vc = fmax(t1[0] * st1[0] + t1[1] * st1[1] + t1[2] * st1[2], 0.0);
  const double detV = fmax(va * vc - vb * vb, 0.0);
  const double ad = va + 0.3, cd = vc + 0.3;
  const double detVd = ad * cd - vb * vb;
  if (detVd < 1e-8) return r;
  const double comp = sqrt(fmax(detV / detVd, 0.0));
  const double opa = sigm64(P[13]);
  if (opa * comp < (double)cut.aMin) return r;
  r.mx = f * pc[0] / pc[2] + (double)cam.cx;
  r.my = f * pc[1] / pc[2] + (double)cam.cy;
  r.A = cd / detVd;
  r.B = -vb / detVd;
  r.C = ad / detVd;
  r.oc = opa * comp;
  double col[3] = {sigm64(P[10]), sigm64(P[11]), sigm64(P[12])};
  if (deg > 0) {
    const int K = (deg + 1) * (deg + 1) - 1;
    double un[3];
    for (int k = 0; k < 3; k++) un[k] = P[k] - (double)cam.camPos[k];
    const double ul = fmax(sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]), 1e-9);
    const double v[3] = {un[0] / ul, un[1] / ul, un[2] / ul};
    double Y[15];
    sh_basis64(deg, v, Y);
    for (int k = 0; k < K && k < cam.shActiveK; k++)
      for (int c = 0; c < 3; c++) col[c] += sh[c * K + k] * Y[k];
    for (int c = 0; c < 3; c++) col[c] = fmax(col[c], 0.0);
  }
  r.c0 = col[0]; r.c1 = col[1]; r.c2 = col[2];
  r.depth = pc[2];
  r.ok = 1;
  return r;
}

// One splat into one pixel's composite (T, C, done), forward-pass rules.
__device__ __forceinline__ void blend64(const Cut& cut, double px, double py, double mx, double my, double A,
                                        double B, double C, double oc, double c0, double c1, double c2, double& T,
                                        double* Cc, bool& done) {
  if (done) return;
  const double dx = px - mx, dy = py - my;
  const double e = fmax(0.5 * (A * dx * dx + C * dy * dy) + B * dx * dy, 0.0);
  if (!(e <= (double)cut.eCut)) return;
  const double alpha = fmin(0.99, oc * exp(-e));
  if (!(alpha >= (double)cut.aMin)) return;
  const double w = T * alpha;
  Cc[0] += w * c0; Cc[1] += w * c1; Cc[2] += w * c2;
  T *= 1.0 - alpha;
  if (T < 1e-4) done = true;
}

__device__ __forceinline__ void blend_rec(const Cut& cut, double px, double py, const Rec64& r, double& T,
                                          double* Cc, bool& done) {
  if (r.ok) blend64(cut, px, py, r.mx, r.my, r.A, r.B, r.C, r.oc, r.c0, r.c1, r.c2, T, Cc, done);
}

__device__ __forceinline__ double charb64(const Cam& cam, const double* Cc, double T, uint32_t pk) {
  double L = 0.0;
  for (int c = 0; c < 3; c++) {
    const double o = Cc[c] + T * (double)cam.bg[c];
    const double tc = (double)((pk >> (8 * c)) & 0xffu) / 255.0;
    const double err = (double)cam.gain * o + (double)cam.bias - tc;
    L += sqrt(err * err + 0.0009) - 0.03;
  }
  return L;
}

// Work item: (sample, tile). Block = 256 threads over the tile's pixels (at
// 8-pixel tiles 64 of them have a pixel; at 32-pixel tiles each thread takes
// four in turn); it adds the sum over the tile's pixels of L(+h) - L(-h) to
// out[sample].
template <int kDeg, int kSh>
__global__ __launch_bounds__(kBlock) void k_fd64(Cam cam, Cut cut, FdArgs a) {
  constexpr int NC = ShDims<kDeg>::C;
  __shared__ Rec64 sG[2];
  __shared__ float4 sA[kBlock];   // mx, my, A, B
  __shared__ float4 sB[kBlock];   // C, opa*comp, depth, id bits
  __shared__ float4 sC[kBlock];   // r, g, b
  __shared__ double sSum[kBlock / 64];
  const uint2 item = a.items[blockIdx.x];
  const FdSample smp = a.samples[item.x];
  const int tile = (int)item.y;
  const int tid = threadIdx.x;
  if (tid < 2) {
    // this sample's Gaussian at +h (tid 0) and -h (tid 1), in FP64 from its stored values
    float p32[16];
    load_params(a.P, a.cap, smp.g, p32);
    double P[16];
    for (int k = 0; k < 16; k++) P[k] = p32[k];
    P[smp.slot] += tid == 0 ? smp.h : -smp.h;
    double sh[NC > 0 ? NC : 1];
    if constexpr (kDeg > 0) {
      float s32[ShDims<kDeg>::S];
      load_sh_any<kDeg, kSh>(a.SH, a.cap, smp.g, s32);
      for (int k = 0; k < NC; k++) sh[k] = s32[k];
    }
    sG[tid] = project64(cam, cut, P, sh, kDeg);
  }
  __syncthreads();
  const Rec64 gp = sG[0], gm = sG[1];
  const int T = cam.tile;
  const int tx = tile % cam.tilesX, ty = tile / cam.tilesX;
  const uint2 rg = a.ranges[tile];
  const uint32_t gid = (uint32_t)smp.g;
  // (uses smp.rect: the Gaussian's tile rectangle in the current raster)
  const bool keyed = tx >= (int)(smp.rect.x & 0xffffu) && tx <= (int)(smp.rect.x >> 16) &&
                     ty >= (int)(smp.rect.y & 0xffffu) && ty <= (int)(smp.rect.y >> 16);
  double d = 0.0;
  // the tile's T x T pixels, kBlock at a time (pixel q: column q % T, row q / T)
  for (int q0 = 0; q0 < T * T; q0 += kBlock) {
    const int q = q0 + tid;
    const int x = tx * T + q % T, y = ty * T + q / T;
    const bool inside = q < T * T && x < cam.W && y < cam.H;
    const int64_t pi = (int64_t)y * cam.W + x;
    const uint32_t pk = inside ? a.tgt[cam.tgtOffset + pi] : 0u;
    const bool valid = inside && (pk >> 24) != 0u;
    const double px = (double)x + 0.5, py = (double)y + 0.5;
    double Tp = 1.0, Tm = 1.0, Cp[3] = {0.0, 0.0, 0.0}, Cm[3] = {0.0, 0.0, 0.0};
    bool dp = !valid, dm = !valid, ip = !gp.ok, im = !gm.ok;   // done, g inserted
    for (uint32_t base = rg.x; base < rg.y; base += kBlock) {
      if (__syncthreads_count(!(dp && dm)) == 0) break;
      const uint32_t k = base + tid;
      if (k < rg.y) {
        const uint32_t id = a.ids[k] & kIdMask;
        const float4* rr = a.rec + (int64_t)rec_stride(a.recFmt) * id;
        const float4 r0 = rr[0], r1 = rr[1];
        float4 c;
        if (a.recFmt == 0) {
          c = rr[2];
        } else {
          const uint32_t cg = __float_as_uint(r1.z), cb = __float_as_uint(r1.w);
          c = a.recFmt == 2 ? make_float4(dec16<2>(cg), dec16<2>(cg >> 16), dec16<2>(cb), 0.0f)
                            : make_float4(dec16<1>(cg), dec16<1>(cg >> 16), dec16<1>(cb), 0.0f);
        }
        sA[tid] = r0;
        sB[tid] = make_float4(r1.x, r1.y, __uint_as_float(a.depthKey[id]), __uint_as_float(id));
        sC[tid] = c;
      }
      __syncthreads();
      const int cnt = (int)min((uint32_t)kBlock, rg.y - base);
      for (int j = 0; j < cnt && !(dp && dm); j++) {
        const float4 r0 = sA[j], r1 = sB[j], c = sC[j];
        const uint32_t id = __float_as_uint(r1.w);
        const double dk = (double)r1.z;
        if (id == gid) {   // the perturbed Gaussian at its key's place
          if (!ip) { blend_rec(cut, px, py, gp, Tp, Cp, dp); ip = true; }
          if (!im) { blend_rec(cut, px, py, gm, Tm, Cm, dm); im = true; }
          continue;
        }
        // in a margin tile: in front of the first key behind it in (depth, id) order
        if (!keyed && !ip && (gp.depth < dk || (gp.depth == dk && gid < id))) { blend_rec(cut, px, py, gp, Tp, Cp, dp); ip = true; }
        if (!keyed && !im && (gm.depth < dk || (gm.depth == dk && gid < id))) { blend_rec(cut, px, py, gm, Tm, Cm, dm); im = true; }
        blend64(cut, px, py, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, c.x, c.y, c.z, Tp, Cp, dp);
        blend64(cut, px, py, r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, c.x, c.y, c.z, Tm, Cm, dm);
      }
    }
    if (!ip) blend_rec(cut, px, py, gp, Tp, Cp, dp);
    if (!im) blend_rec(cut, px, py, gm, Tm, Cm, dm);
    if (valid) d += charb64(cam, Cp, Tp, pk) - charb64(cam, Cm, Tm, pk);
    __syncthreads();   // the next pixel group restages the records
  }
  for (int o = 32; o > 0; o >>= 1) d += __shfl_xor(d, o);
  if ((tid & 63) == 0) sSum[tid >> 6] = d;
  __syncthreads();
  if (tid == 0) {
    const double s = (sSum[0] + sSum[1]) + (sSum[2] + sSum[3]);
    if (s != 0.0) atomicAdd(&a.out[item.x], s);
  }
}

template <int kSh>
static void fd_sh(hipStream_t st, int deg, int blocks, const Cam& cam, const Cut& cut, const FdArgs& a) {
  switch (deg) {
    case 0: k_fd64<0, kSh><<<blocks, kBlock, 0, st>>>(cam, cut, a); break;
    case 1: k_fd64<1, kSh><<<blocks, kBlock, 0, st>>>(cam, cut, a); break;
    case 2: k_fd64<2, kSh><<<blocks, kBlock, 0, st>>>(cam, cut, a); break;
    default: k_fd64<3, kSh><<<blocks, kBlock, 0, st>>>(cam, cut, a); break;
  }
}

void launch_fd64(hipStream_t st, int deg, int shFmt, int64_t items, const Cam& cam, const Cut& cut,
                 const FdArgs& a) {
  if (items <= 0) return;
  const int blocks = (int)items;
  if (shFmt == 2) fd_sh<2>(st, deg, blocks, cam, cut, a);
  else if (shFmt == 1) fd_sh<1>(st, deg, blocks, cam, cut, a);
  else fd_sh<0>(st, deg, blocks, cam, cut, a);
}

}  // namespace gs
