// Stage (c): W(k-k') table and BSE matrix assembly.
//
// Reference: subroutines/coulomb_pot.f90:2-469 (the nine potentials),
// subroutines/bse_subs.f90:2-124 (matrizelbse), build loop
// subprograms/bse_solver-tool-diel.f90:431-458, geometry general_subs.f90:183-308.
//
// The reference evaluates V(k_I - k_J) for every matrix element.  k - k' is
// never folded (general_subs.f90:297-308), so V depends only on the integer
// grid difference Delta = (i-i', j-j', l-l') in [-(n-1), n-1]^3; the table
// holds one value per Delta and the assembly gathers from it.
#include "wtb_kernels.h"

namespace wtb {

__device__ __constant__ double kCic = -90.4755935668945312;   // -(0.0904756)*10**3 in REAL*4
__device__ __constant__ double kPi32b = 3.14159274101257324;

__device__ inline void cross3(const double* v1, const double* v2, double* vx) {
#pragma clang fp contract(off)
  vx[0] = (v1[1] * v2[2]) - (v1[2] * v2[1]);
  vx[1] = (v1[2] * v2[0]) - (v1[0] * v2[2]);
  vx[2] = (v1[0] * v2[1]) - (v1[1] * v2[0]);
}
__device__ inline double vsize3(const double* v) {
#pragma clang fp contract(off)
  return sqrt((v[0] * v[0]) + (v[1] * v[1]) + (v[2] * v[2]));
}

// Modified Bessel K_nu(x), nu = 0, 1, from K_nu(x) = int_0^inf exp(-x cosh t) cosh(nu t) dt
// with the trapezoidal rule (exponentially convergent for this analytic, even integrand).
__device__ double bessel_k(int nu, double x) {
  const double h = 0.125;
  double s = 0.5 * exp(-x);
  for (int k = 1; k < 4000; ++k) {
    double t = k * h;
    double ex = x * cosh(t);
    if (ex > 745.0) break;
    s += exp(-ex) * (nu == 0 ? 1.0 : cosh(t));
  }
  return h * s;
}

// integralJ0pot (coulomb_pot.f90:437-469): 100-point trapezoid of r J0(g r) log r
__device__ double integral_j0pot(double rmax, double gperp) {
#pragma clang fp contract(off)
  const int n = 100;
  const double eta = (double)1e-4f;
  double r1 = ((1.0 - 1.0) / (n - 1.0)) * rmax + eta;
  double r2 = ((2.0 - 1.0) / (n - 1.0)) * rmax + eta;
  double h = r2 - r1;
  double res = 0.0;
  double rprev = r1, fprev = r1 * j0(r1 * gperp) * log(r1);
  for (int i = 2; i <= n; ++i) {
    double r = ((i - 1.0) / (n - 1.0)) * rmax + eta;
    double f = r * j0(r * gperp) * log(r);
    res = res + (h / 2.0) * (fprev + f);
    fprev = f;
    rprev = r;
  }
  (void)rprev;
  return res;
}

// V for one k-difference dk, selected as matrizelbse does (bse_subs.f90:45-91)
__device__ double coulomb_dev(const CoulombArgs& ca, const double a[3][3], const double dk[3]) {
#pragma clang fp contract(off)
  const double nkt = (double)(ca.n1 * ca.n2 * ca.n3);
  const double modk = sqrt((dk[0] * dk[0] + dk[1] * dk[1]) + dk[2] * dk[2]);
  const double tolr = ca.ktol;
  double v12[3];
  cross3(a[0], a[1], v12);
  const double area = vsize3(v12);                                      // vcell2D
  const double vol = fabs(v12[0] * a[2][0] + v12[1] * a[2][1] + v12[2] * a[2][2]);  // vcell3D
  switch (ca.type) {
    case WTB_V2DRK: {                                                    // coulomb_pot.f90:276-340
      double vbz = 1.0 / (nkt * area);
      if (modk < tolr) return 0.0;
      double epar = ca.ediel[1], et = ca.ediel[0], eb = ca.ediel[2];
      double eta = sqrt(epar / ca.ez), kappa = sqrt(epar * ca.ez);
      double pb = (eb - kappa) / (eb + kappa), pt = (et - kappa) / (et + kappa);
      double aux1 = (1.0 - (pb * pt * exp(-2.0 * modk * eta * ca.lc))) * kappa;
      double aux2 = (1.0 - (pt * exp(-eta * modk * ca.lc))) * (1.0 - (pb * exp(-eta * modk * ca.lc)));
      double aux3 = ca.r0 * modk * exp(-modk * ca.w);
      double ew = (aux1 / aux2) + aux3;
      return (vbz * kCic) * exp(-modk * ca.w) * (1.0 / ew) * (1.0 / modk);
    }
    case WTB_V3D:                                                        // :55-85
    case WTB_V3DL: {                                                     // :89-122
      double ed = ca.type == WTB_V3D ? 1.0 : ca.ediel[1];
      double vbz = 1.0 / (nkt * vol);
      if (modk < tolr) return 0.0;
      return vbz * (kCic / ed) * (1.0 / (modk * modk));
    }
    case WTB_V2DK: {                                                     // :2-51
      double a0 = 0.5 * (vsize3(a[0]) + vsize3(a[1]));
      double r0 = ((ca.ediel[1] - 1.0) * ca.lc) / (ca.ediel[0] + ca.ediel[2]);
      double vbz = 1.0 / (nkt * area);
      double g1 = (double)(ca.n1 * ca.n2);
      double auxi = (2.0 * kPi32b * r0) / (a0 * sqrt(g1));
      double ed = (ca.ediel[0] + ca.ediel[2]) / 2.0;
      const double alpha1 = (double)1.76f, alpha2 = 1.0, alpha3 = 0.0;
      if (modk < tolr)
        return vbz * (kCic / ed) * (a0 * sqrt(g1) / (2.0 * kPi32b)) *
               (alpha1 + auxi * alpha2 + alpha3 * (auxi * auxi));
      return vbz * (kCic / ed) * (1.0 / (modk * (1 + (r0 * modk))));
    }
    case WTB_V2DT: {                                                     // :126-183
      double vbz = 1.0 / (nkt * vol);
      double gz = fabs(dk[2]);
      double gpar = sqrt((dk[0] * dk[0]) + (dk[1] * dk[1]));
      double rc = 0.5 * a[2][2];
      if (gpar < tolr && gz < tolr) return (vbz * kCic) * (-0.5 * rc * rc);
      if (gpar < tolr)
        return (vbz * kCic) * (1.0 / (modk * modk)) * (1.0 - cos(gz * rc) - (gz * rc * sin(gz * rc)));
      double aux1 = gz / gpar, aux2 = gpar * rc, aux3 = gz * rc;
      double aux4 = aux1 * sin(aux3), aux5 = cos(aux3);
      return (vbz * kCic) * (1.0 / (modk * modk)) * (1.0 + (exp(-aux2) * (aux4 - aux5)));
    }
    case WTB_V2DT2: {                                                    // :233-272
      double vbz = 1.0 / (nkt * vol);
      double qxy = sqrt((dk[0] * dk[0]) + (dk[1] * dk[1]));
      if (modk < tolr) return 0.0;
      return (vbz * kCic) * (1.0 / (modk * modk)) *
             (1.0 - exp(-0.5 * qxy * ca.lc) * cos(0.5 * ca.lc * dk[2]));
    }
    case WTB_V0DT: {                                                     // :187-228
      double vbz = 1.0 / (nkt * vol);
      double cr = fmin(fmin(vsize3(a[0]), vsize3(a[1])), vsize3(a[2]));
      cr = 0.5 * cr;
      if (modk < tolr) return (vbz * kCic) * (0.5) * cr * cr;
      return (vbz * kCic) * (1.0 / (modk * modk)) * (1.0 - cos(cr * modk));
    }
    case WTB_V2DOH: {                                                    // :344-377
      double vbz = 1.0 / (nkt * area);
      if (modk < tolr) return 0.0;
      return (vbz * kCic) * exp(-ca.w * modk) * (1.0 / (ca.ez * modk));
    }
    case WTB_V1DT: {                                                     // :381-432
      double vbz = 1.0 / (nkt * vol);
      double ax = fabs(dk[0]);
      double qyz = sqrt((dk[1] * dk[1]) + (dk[2] * dk[2]));
      if (ax < tolr && qyz >= tolr) return (vbz * kCic) * integral_j0pot(ca.lc, qyz);
      if (ax < tolr && qyz < tolr)
        return (vbz * kCic) * ((ca.lc * ca.lc) / 16.0) * (2.0 * log(ca.lc) - 1.0);
      double vj0 = j0(qyz * ca.lc), vj1 = j1(qyz * ca.lc);
      double vk0 = bessel_k(0, ax * ca.lc), vk1 = bessel_k(1, ax * ca.lc);
      double c1 = qyz * ca.lc * vj1 * vk0;
      double c2 = ax * ca.lc * vj0 * vk1;
      return (vbz * kCic) * (1.0 + c1 + c2);
    }
    default:
      return 0.0;
  }
}

// recvec with single-precision pi (mhkpack_subs.f90:409-432)
__device__ inline void recvec_b(const double a[3][3], double b[3][3]) {
#pragma clang fp contract(off)
  double v23[3], v31[3], v12[3];
  cross3(a[1], a[2], v23);
  cross3(a[2], a[0], v31);
  cross3(a[0], a[1], v12);
  double vol = fabs((a[0][0] * v23[0]) + (a[0][1] * v23[1]) + (a[0][2] * v23[2]));
  double f = (2.0 * kPi32b) / vol;
  for (int c = 0; c < 3; ++c) {
    b[0][c] = f * v23[c];
    b[1][c] = f * v31[c];
    b[2][c] = f * v12[c];
  }
}

__global__ void k_wtable(const double* __restrict__ rl, CoulombArgs ca, double* __restrict__ wt) {
#pragma clang fp contract(off)
  const int D1 = 2 * ca.n1 - 1, D2 = 2 * ca.n2 - 1, D3 = 2 * ca.n3 - 1;
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= D1 * D2 * D3) return;
  int t3 = idx % D3, t2 = (idx / D3) % D2, t1 = idx / (D2 * D3);
  double d1 = t1 - (ca.n1 - 1), d2 = t2 - (ca.n2 - 1), d3 = t3 - (ca.n3 - 1);
  double a[3][3], b[3][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) a[i][c] = rl[i + 3 * c];
  recvec_b(a, b);
  double dk[3];
  for (int c = 0; c < 3; ++c)
    dk[c] = (b[0][c] / (double)ca.n1) * d1 + (b[1][c] / (double)ca.n2) * d2 +
            (b[2][c] / (double)ca.n3) * d3;
  wt[idx] = coulomb_dev(ca, a, dk);
}

// ---------------------------------------------------------------------------
// Assembly.  One workgroup per lower-triangle tile of TK x TK k-points.  The
// band overlaps O_c[(k,c),(k',c')] = <c k|c' k'> and O_v likewise are complex
// GEMMs over the orbital index on MFMA (A = conj(U_rows), B = U_cols^T); the
// element is then W(Delta(k,k')) * O_c * O_v, diagonal (Ec - Ev) + W(0).
// Only the lower triangle (I >= J) is written: its elements are the conjugates
// of the reference's upper-triangle elements (bse_solver-tool-diel.f90:433-437).
// ---------------------------------------------------------------------------
struct BuildArgs {
  int nk, nc, nv, nw, nb, ncv, TK, n1, n2, n3, D2, D3, N, ntiles_side, ldh;
};

__device__ inline zc load_band(const zc* vec, const BuildArgs& ba, int k, int band, int m) {
  return vec[((size_t)k * ba.nb + band) * ba.nw + m];
}

// Overlap GEMM for one 16x16 block (complex), rows r -> (k0r + r/nbnd, band0 + r%nbnd)
// F32: the precision form, FP32 operands and accumulation on v_mfma_f32_16x16x4_f32
template <bool F32>
__device__ inline void overlap_block(const zc* __restrict__ vec, const BuildArgs& ba, int kr0, int kc0,
                                     int nbnd, int band0, int rows, int rb, int cb, zc* out, int ldo) {
  const int lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  int r = rb * 16 + lr;          // A row
  int c = cb * 16 + lr;          // B column
  int kr = kr0 + r / nbnd, br = band0 + r % nbnd;
  int kc = kc0 + c / nbnd, bc = band0 + c % nbnd;
  bool rv = r < rows && kr < ba.nk;
  bool cv = c < rows && kc < ba.nk;
  if constexpr (!F32) {
    d4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    for (int m0 = 0; m0 < ba.nw; m0 += 4) {
      int m = m0 + lk;
      zc a = (rv && m < ba.nw) ? load_band(vec, ba, kr, br, m) : mkz(0, 0);
      zc b = (cv && m < ba.nw) ? load_band(vec, ba, kc, bc, m) : mkz(0, 0);
      // conj(a) * b
      cr = mfma16x16x4(a.re, b.re, cr);
      cr = mfma16x16x4(a.im, b.im, cr);
      ci = mfma16x16x4(a.re, b.im, ci);
      ci = mfma16x16x4(-a.im, b.re, ci);
    }
    for (int q = 0; q < 4; ++q) {
      int orow = rb * 16 + lk + 4 * q, ocol = cb * 16 + lr;   // f64: row lk + 4q
      if (orow < rows && ocol < rows) out[orow * ldo + ocol] = mkz(cr[q], ci[q]);
    }
  } else {
    f4 cr = {0, 0, 0, 0}, ci = {0, 0, 0, 0};
    for (int m0 = 0; m0 < ba.nw; m0 += 4) {
      int m = m0 + lk;
      zc a = (rv && m < ba.nw) ? load_band(vec, ba, kr, br, m) : mkz(0, 0);
      zc b = (cv && m < ba.nw) ? load_band(vec, ba, kc, bc, m) : mkz(0, 0);
      const float ar = (float)a.re, ai = (float)a.im, br_ = (float)b.re, bi = (float)b.im;
      cr = mfma16x16x4f32(ar, br_, cr);
      cr = mfma16x16x4f32(ai, bi, cr);
      ci = mfma16x16x4f32(ar, bi, ci);
      ci = mfma16x16x4f32(-ai, br_, ci);
    }
    for (int q = 0; q < 4; ++q) {
      int orow = rb * 16 + 4 * lk + q, ocol = cb * 16 + lr;   // f32: row 4lk + q
      if (orow < rows && ocol < rows) out[orow * ldo + ocol] = mkz(cr[q], ci[q]);
    }
  }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_bse_build(BuildArgs ba, const double* __restrict__ wtab,
                                                   const double* __restrict__ eigv,
                                                   const zc* __restrict__ vec, zc* __restrict__ H,
                                                   zf* __restrict__ Hf, const double* __restrict__ fdeh,
                                                   const zc* __restrict__ xa, const double* __restrict__ dvq) {
  extern __shared__ unsigned char smem[];
  // lower-triangle tile index -> (tr, tc), tr >= tc
  int t = blockIdx.x;
  int tr = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((tr + 1) * (tr + 2) / 2 <= t) ++tr;
  while (tr * (tr + 1) / 2 > t) --tr;
  int tc = t - tr * (tr + 1) / 2;
  const int TK = ba.TK;
  const int kr0 = tr * TK, kc0 = tc * TK;
  const int rc_ = TK * ba.nc, rv_ = TK * ba.nv;
  zc* Oc = reinterpret_cast<zc*>(smem);
  zc* Ov = Oc + rc_ * rc_;
  const int wave = threadIdx.x >> 6;
  const int bc_ = (rc_ + 15) / 16, bv_ = (rv_ + 15) / 16;
  const int nblocks = bc_ * bc_ + bv_ * bv_;
  for (int blk = wave; blk < nblocks; blk += 4) {
    if (blk < bc_ * bc_)
      overlap_block<F32>(vec, ba, kr0, kc0, ba.nc, ba.nv, rc_, blk / bc_, blk % bc_, Oc, rc_);
    else {
      int b2 = blk - bc_ * bc_;
      overlap_block<F32>(vec, ba, kr0, kc0, ba.nv, 0, rv_, b2 / bv_, b2 % bv_, Ov, rv_);
    }
  }
  const int side = TK * ba.ncv;
  // per row and per column of the tile, once per workgroup: the W-table index splits as
  // RW[i] - CW[j] (it is linear in the two k-points' grid coordinates), and the overlaps'
  // row and column offsets; the element loop then needs no integer division
  int* RW = reinterpret_cast<int*>(Ov + rv_ * rv_);
  int* CW = RW + side;
  int* ROc = CW + side;
  int* COc = ROc + side;
  int* ROv = COc + side;
  int* COv = ROv + side;
  const int c0 = ba.n2 * ba.n3;
  for (int x = threadIdx.x; x < side; x += 256) {
    const int kl = x / ba.ncv, cx = (x % ba.ncv) / ba.nv, vx = x % ba.nv;
    const int kr = kr0 + kl, kc = kc0 + kl;
    RW[x] = ((kr / c0) * ba.D2 + (kr / ba.n3) % ba.n2) * ba.D3 + kr % ba.n3 +
            ((ba.n1 - 1) * ba.D2 + (ba.n2 - 1)) * ba.D3 + (ba.n3 - 1);
    CW[x] = ((kc / c0) * ba.D2 + (kc / ba.n3) % ba.n2) * ba.D3 + kc % ba.n3;
    ROc[x] = (kl * ba.nc + cx) * rc_;
    COc[x] = kl * ba.nc + cx;
    ROv[x] = (kl * ba.nv + vx) * rv_;
    COv[x] = kl * ba.nv + vx;
  }
  __syncthreads();
  const int I0 = kr0 * ba.ncv, J0 = kc0 * ba.ncv;
  const double vq = xa ? *dvq : 0.0;
  // element e = i + side j, i the row (consecutive threads, consecutive addresses)
  const int di = 256 % side, dj = 256 / side;
  int i = threadIdx.x % side, j = threadIdx.x / side;
  for (; j < side; i += di, j += dj) {
    if (i >= side) {
      i -= side;
      ++j;
      if (j >= side) break;
    }
    const int I = I0 + i, J = J0 + j;
    if (I >= ba.N || J >= ba.N || I < J) continue;
    const double wv = wtab[RW[i] - CW[j]];
    zc val;
    if (I == J) {
      const int kl = i / ba.ncv, ci = (i % ba.ncv) / ba.nv, vi = i % ba.nv;
      const int k = kr0 + kl;
      double ec = eigv[(size_t)k * ba.nb + ba.nv + ci], ev = eigv[(size_t)k * ba.nb + vi];
      if (fdeh) {
        // finite temperature: (Ec - Ev) + V(0) f_I (bse_subs_temp.f90:105-108); with finite Q
        // (Ec - Ev) + (V(0) - V(q) <c|v><v|c>) f_I (:300-310)
        zc kern = mkz(wv, 0.0);
        if (xa) kern = kern - (vq * xa[I]) * conj(xa[I]);
        val = mkz(ec - ev, 0.0) + kern * fdeh[J];
      } else {
        val = mkz((ec - ev) + wv, 0.0);
        // finite Q: - V(q) <c|v> <v|c> (bse_subs_kpath.f90:158-160)
        if (xa) val = val - (vq * xa[I]) * conj(xa[I]);
      }
    } else {
      zc oc = Oc[ROc[i] + COc[j]];
      zc ov = Ov[ROv[i] + COv[j]];
      val = (wv * oc) * ov;
      // finite Q: the reference's (J, I) element subtracts V(q) a_J conj(a_I)
      // (bse_subs_kpath.f90:178-191); conjugated: V(q) conj(a_J) a_I
      if (xa) val = val - (vq * conj(xa[J])) * xa[I];
      // finite temperature: the reference's element (J, I), J < I, is the kernel times f_J
      // (bse_subs_temp.f90:112-120, 312-327); this one is its conjugate
      if (fdeh) val = val * fdeh[J];
    }
    size_t off = (size_t)I + (size_t)ba.ldh * J;
    if constexpr (F32) stz(Hf[off], val);
    else H[off] = val;
  }
}

int launch_wtable(hipStream_t s, const double* d_rlat, const CoulombArgs& ca, double* d_wtab) {
  int tot = (2 * ca.n1 - 1) * (2 * ca.n2 - 1) * (2 * ca.n3 - 1);
  hipLaunchKernelGGL(k_wtable, dim3((tot + 127) / 128), dim3(128), 0, s, d_rlat, ca, d_wtab);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_bse_build(hipStream_t s, int nk, int nc, int nv, int nw, const int ng[3],
                     const double* d_wtab, const double* d_eigv, const zc* d_vec, zc* d_H, int ldh,
                     zf* d_Hf, const double* d_fdeh, const zc* d_xa, const double* d_vq) {
  BuildArgs ba;
  ba.nk = nk;
  ba.nc = nc;
  ba.nv = nv;
  ba.nw = nw;
  ba.nb = nc + nv;
  ba.ncv = nc * nv;
  int tk = 64 / ba.ncv;
  if (tk > 16) tk = 16;
  if (tk < 1) tk = 1;
  ba.TK = tk;
  ba.n1 = ng[0];
  ba.n2 = ng[1];
  ba.n3 = ng[2];
  ba.D2 = 2 * ng[1] - 1;
  ba.D3 = 2 * ng[2] - 1;
  ba.N = nk * nc * nv;
  ba.ldh = ldh;
  int side = (nk + tk - 1) / tk;
  ba.ntiles_side = side;
  int ntiles = side * (side + 1) / 2;
  // the two overlap blocks, then six int tables of the tile's side (rows and columns)
  size_t lds = sizeof(zc) * ((size_t)tk * nc * tk * nc + (size_t)tk * nv * tk * nv) +
               sizeof(int) * 6 * (size_t)tk * nc * nv;
  if (d_H)
    hipLaunchKernelGGL(k_bse_build<false>, dim3(ntiles), dim3(256), lds, s, ba, d_wtab, d_eigv, d_vec,
                       d_H, (zf*)nullptr, d_fdeh, d_xa, d_vq);
  if (d_Hf)
    hipLaunchKernelGGL(k_bse_build<true>, dim3(ntiles), dim3(256), lds, s, ba, d_wtab, d_eigv, d_vec,
                       (zc*)nullptr, d_Hf, d_fdeh, d_xa, d_vq);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// fermidisteh (general_subs.f90:606-635) with E_F = 0 for every basis state
// I = (k nc + c) nv + v, the order of the assembly
__global__ void k_fermi_factors(int N, int nc, int nv, const double* __restrict__ eigv, double beta,
                                double* __restrict__ fdeh) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= N) return;
  const int ncv = nc * nv, nb = nc + nv;
  const int k = I / ncv, c = (I % ncv) / nv, v = I % nv;
  const double ec = eigv[(size_t)k * nb + nv + c], ev = eigv[(size_t)k * nb + v];
  const double fv = 1.0 / (exp(ev * beta) + 1.0);
  const double fc = 1.0 / (exp(ec * beta) + 1.0);
  fdeh[I] = fv - fc;
}

int launch_fermi_factors(hipStream_t s, int nk, int nc, int nv, const double* d_eigv, double beta,
                         double* d_fdeh) {
  const int N = nk * nc * nv;
  hipLaunchKernelGGL(k_fermi_factors, dim3((N + 255) / 256), dim3(256), 0, s, N, nc, nv, d_eigv, beta, d_fdeh);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

__global__ void k_scale_rows(int nvec, int N, const zc* __restrict__ x, const double* __restrict__ f,
                             zc* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double fi = f[i];
  for (int q = 0; q < nvec; ++q) y[(size_t)q * N + i] = x[(size_t)q * N + i] * fi;
}

int launch_scale_rows(hipStream_t s, int nvec, int N, const zc* x, const double* f, zc* y) {
  hipLaunchKernelGGL(k_scale_rows, dim3((N + 255) / 256), dim3(256), 0, s, nvec, N, x, f, y);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb

namespace wtb {

// ---------------------------------------------------------------------------
// Finite-Q exciton bands (bsebnds, subprograms/bse_kpath-tool.f90:187-389): the band
// window of the conduction states comes from the k+q grid, the valence window from k.
// ---------------------------------------------------------------------------
__global__ void k_mix_q(int nk, int nc, int nv, int nw, const double* __restrict__ eigvq,
                        const zc* __restrict__ vecq, double* __restrict__ eigv, zc* __restrict__ vec) {
  const int nb = nc + nv;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;   // (k, conduction band c, orbital m)
  if (t >= nk * nc * nw) return;
  const int m = t % nw, c = (t / nw) % nc, k = t / (nw * nc);
  const size_t row = (size_t)k * nb + nv + c;
  vec[row * nw + m] = vecq[row * nw + m];
  if (m == 0) eigv[row] = eigvq[row];
}

int launch_mix_q(hipStream_t s, int nk, int nc, int nv, int nw, const double* d_eigvq, const zc* d_vecq,
                 double* d_eigv, zc* d_vec) {
  const int n = nk * nc * nw;
  hipLaunchKernelGGL(k_mix_q, dim3((n + 255) / 256), dim3(256), 0, s, nk, nc, nv, nw, d_eigvq, d_vecq, d_eigv,
                     d_vec);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// a[I] = sum_m conj(c_m) v_m for basis state I = (k nc + c) nv + v: prodintsq(conj(vbc1), vbv1)
// (bse_subs_kpath.f90:150-156), summed in orbital order
__global__ void k_exchange_vec(int nk, int nc, int nv, int nw, const zc* __restrict__ vec, zc* __restrict__ a) {
  const int N = nk * nc * nv;
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= N) return;
  const int ncv = nc * nv, nb = nc + nv;
  const int k = I / ncv, c = (I % ncv) / nv, v = I % nv;
  const zc* pc = vec + ((size_t)k * nb + nv + c) * nw;
  const zc* pv = vec + ((size_t)k * nb + v) * nw;
  zc acc = mkz(0.0, 0.0);
  for (int m = 0; m < nw; ++m) acc = acc + conj(pc[m]) * pv[m];
  a[I] = acc;
}

int launch_exchange_vec(hipStream_t s, int nk, int nc, int nv, int nw, const zc* d_vec, zc* d_a) {
  const int N = nk * nc * nv;
  hipLaunchKernelGGL(k_exchange_vec, dim3((N + 255) / 256), dim3(256), 0, s, nk, nc, nv, nw, d_vec, d_a);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// V(q) = the potential at k1 = q, k2 = 0 (bse_subs_kpath.f90:70-119); 0 when |q| = 0, where
// the reference takes the branch without the exchange term (:125-142)
__global__ void k_vq(const double* __restrict__ rl, CoulombArgs ca, const double* __restrict__ dq,
                     double* __restrict__ vq) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double dk[3] = {dq[1], dq[2], dq[3]};
  const double modq = sqrt((dk[0] * dk[0] + dk[1] * dk[1]) + dk[2] * dk[2]);
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) a[i][c] = rl[i + 3 * c];
  vq[0] = modq == 0.0 ? 0.0 : coulomb_dev(ca, a, dk);
}

int launch_vq(hipStream_t s, const double* d_rlat, const CoulombArgs& ca, const double* d_q, double* d_vq) {
  hipLaunchKernelGGL(k_vq, dim3(1), dim3(64), 0, s, d_rlat, ca, d_q, d_vq);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
