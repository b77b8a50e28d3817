// Stage (e): oscillator strengths and eps(omega).
//
// Reference: dielbsev/dielbsep (subroutines/bse_subs.f90:278-378) and their
// driver (subprograms/bse_solver-tool-diel.f90:540-602); the "(7F15.4)" write
// of bse_opt_diel.dat and its read-back (:607-620, bse_diel-tool.f90:117-123);
// optdiel/avgsme/imagrenorm (subroutines/optics.f90:160-244,302-390) in the
// bsedielraw loop (subprograms/bse_diel-tool.f90:111,135-201).
#include "wtb_kernels.h"

namespace wtb {

// G = Z^T (Q^T h) as 6 real vectors [x.re, x.im, y.re, y.im, z.re, z.im]; the
// activities are |t|^2 and Re(t_a conj(t_b)) with t = X^T h (no conjugate on X).
__global__ void k_activities(int N, const double* __restrict__ G, int dtfull, int cpol,
                             double* __restrict__ act) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  zc tx = mkz(G[n], G[(size_t)N + n]);
  zc ty = mkz(G[2 * (size_t)N + n], G[3 * (size_t)N + n]);
  zc tz = mkz(G[4 * (size_t)N + n], G[5 * (size_t)N + n]);
  auto sq = [](zc a) { return a.re * a.re + a.im * a.im; };
  auto cr = [](zc a, zc b) { return a.re * b.re + a.im * b.im; };   // Re(a conj(b))
  act[n] = sq(tx);
  act[(size_t)N + n] = sq(ty);
  act[2 * (size_t)N + n] = sq(tz);
  act[3 * (size_t)N + n] = dtfull ? cr(tx, ty) : 0.0;
  act[4 * (size_t)N + n] = dtfull ? cr(tx, tz) : 0.0;
  act[5 * (size_t)N + n] = dtfull ? cr(ty, tz) : 0.0;
  // hrsp = (hrx + i hry) * REAL(1./sqrt(2.)), hrsm = (hrx - i hry) * ... (:376-377)
  const double c = (double)(1.0f / sqrtf(2.0f));
  zc tp = mkz(tx.re - ty.im, tx.im + ty.re) * c;
  zc tm = mkz(tx.re + ty.im, tx.im - ty.re) * c;
  act[6 * (size_t)N + n] = cpol ? sq(tp) : 0.0;
  act[7 * (size_t)N + n] = cpol ? sq(tm) : 0.0;
}

// dielbseptemp (bse_subs_temp.f90:340-417) for the eight components
// (bse_solver-tool-diel-temp.f90:570-626): Re[(X^T h_a) conj(X^T (f h_b))], where
// G rows 0-5 hold X^T h (x, y, z as re, im) and rows 6-11 X^T (f h)
__global__ void k_activities_temp(int N, const double* __restrict__ G, int dtfull, int cpol,
                                  double* __restrict__ act) {
  int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  auto row = [&](int q) { return mkz(G[(size_t)(2 * q) * N + n], G[(size_t)(2 * q + 1) * N + n]); };
  const zc tx = row(0), ty = row(1), tz = row(2), fx = row(3), fy = row(4), fz = row(5);
  auto cr = [](zc a, zc b) { return a.re * b.re + a.im * b.im; };   // Re(a conj(b))
  act[n] = cr(tx, fx);
  act[(size_t)N + n] = cr(ty, fy);
  act[2 * (size_t)N + n] = cr(tz, fz);
  act[3 * (size_t)N + n] = dtfull ? cr(tx, fy) : 0.0;
  act[4 * (size_t)N + n] = dtfull ? cr(tx, fz) : 0.0;
  act[5 * (size_t)N + n] = dtfull ? cr(ty, fz) : 0.0;
  // hrsp, hrsm as at T = 0 (bse_solver-tool-diel-temp.f90:406-407); X^T is linear
  const double c = (double)(1.0f / sqrtf(2.0f));
  const zc tp = mkz(tx.re - ty.im, tx.im + ty.re) * c, tm = mkz(tx.re + ty.im, tx.im - ty.re) * c;
  const zc fp = mkz(fx.re - fy.im, fx.im + fy.re) * c, fm = mkz(fx.re + fy.im, fx.im - fy.re) * c;
  act[6 * (size_t)N + n] = cpol ? cr(tp, fp) : 0.0;
  act[7 * (size_t)N + n] = cpol ? cr(tm, fm) : 0.0;
}

// x -> the value a "(F15.4)" write and a list-directed read give back: the
// decimal rounding of the exact binary value to 4 places, then the nearest
// double of that decimal.
__device__ inline double f15_4(double x) {
  double y = x * 1e4;
  double err = fma(x, 1e4, -y);          // x*1e4 = y + err exactly
  double fl = floor(y);
  double r;
  if (y - fl == 0.5) r = err > 0 ? fl + 1.0 : (err < 0 ? fl : rint(y));
  else r = rint(y);
  return r / 1e4;
}

__global__ void k_quantize(int n, const double* __restrict__ src, double* __restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f15_4(src[i]);
}

// avgsme (optics.f90:302-324): sequential sum of successive differences, in the
// reference's order; the energies are staged through LDS in chunks
__global__ __launch_bounds__(256) void k_avgsme(int N, const double* __restrict__ E, double* __restrict__ out) {
  __shared__ double buf[4097];
  double s = 0.0, prev = 0.0;
  for (int c0 = 0; c0 < N; c0 += 4096) {
    const int m = min(4096, N - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += 256) buf[i] = E[c0 + i];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 0; i < m; ++i) {
        const int gi = c0 + i;
        if (gi > 0) s = s + buf[i] - prev;
        prev = buf[i];
      }
    }
  }
  if (threadIdx.x == 0) out[0] = 10.0 * (s / (double)N);
}

// optdiel for 4 photon energies per workgroup and all six components:
//   eps2 = caux sum f g/((w-E)^2+g^2),  eps1 = delta + caux sum f (E-w)/((w-E)^2+g^2)
// with g = sme + avgsme; partial sums reduced in a fixed tree order.
#define OMB 4
__global__ __launch_bounds__(256) void k_optdiel(int N, const double* __restrict__ E,
                                                 const double* __restrict__ f6, double caux,
                                                 double ebse0, double ebsef, int numbse, double sme,
                                                 const double* __restrict__ smeavg,
                                                 double* __restrict__ eps) {
#pragma clang fp contract(off)
  __shared__ double red[256];
  const int j0 = blockIdx.x * OMB;
  const double smef = smeavg[0] + sme;
  double om[OMB];
  for (int q = 0; q < OMB; ++q) {
    int j = j0 + q;
    om[q] = ebse0 + ((ebsef - ebse0) * (double)j) / ((double)numbse - 1.0);
  }
  double sr[OMB][6], si[OMB][6];
  for (int q = 0; q < OMB; ++q)
    for (int c = 0; c < 6; ++c) sr[q][c] = si[q][c] = 0.0;
  for (int n = threadIdx.x; n < N; n += 256) {
    double e = E[n];
    double f[6];
    for (int c = 0; c < 6; ++c) f[c] = f6[(size_t)c * N + n];
    for (int q = 0; q < OMB; ++q) {
      double dw = om[q] - e;
      double eaux = (dw * dw) + (smef * smef);
      double rr = (e - om[q]) / eaux;
      double ii = smef / eaux;
      for (int c = 0; c < 6; ++c) {
        sr[q][c] += f[c] * rr;
        si[q][c] += f[c] * ii;
      }
    }
  }
  for (int q = 0; q < OMB; ++q)
    for (int c = 0; c < 6; ++c)
      for (int part = 0; part < 2; ++part) {
        double v = part ? si[q][c] : sr[q][c];
        red[threadIdx.x] = v;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
          if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
          __syncthreads();
        }
        if (threadIdx.x == 0) {
          int j = j0 + q;
          if (j < numbse) {
            double out = part ? red[0] * caux : (c < 3 ? 1.0 : 0.0) + red[0] * caux;
            eps[j + (size_t)numbse * (part + 2 * c)] = out;
          }
        }
        __syncthreads();
      }
}

// imagrenorm (optics.f90:327-390) on eps2 of each component
__global__ __launch_bounds__(256) void k_imagrenorm(const double* __restrict__ E, double ebse0, double ebsef,
                                                    int numbse, double* __restrict__ eps) {
#pragma clang fp contract(off)
  __shared__ int sib;
  __shared__ double saux3;
  const double factor = (double)0.15f;
  const double aux = E[0] - factor;
  if (aux <= 0.0) return;
  if (threadIdx.x == 0) {
    int ib = 0;
    for (int j = 0; j < numbse; ++j) {
      double w = ebse0 + ((ebsef - ebse0) * (double)j) / ((double)numbse - 1.0);
      if (w <= aux) ib = j;
      else break;
    }
    sib = ib;
  }
  __syncthreads();
  const int ib = sib;
  for (int c = 0; c < 6; ++c) {
    double* im = eps + (size_t)numbse * (1 + 2 * c);
    if (threadIdx.x == 0) saux3 = (ib + 1 < numbse) ? im[ib + 1] : 0.0;
    __syncthreads();
    const double aux3 = saux3;
    for (int j = threadIdx.x; j < numbse; j += 256) {
      double v;
      if (ib + 1 >= numbse) v = 0.0;          // the reference would read past the grid; zero it all
      else if (j <= ib) v = 0.0;
      else {
        v = im[j] - aux3;
        v = v < 0.0 ? 0.0 : v;
      }
      im[j] = v;
    }
    __syncthreads();
  }
}

// t_i = sum_j X(j,i) h(j) for up to two vectors (dielbsev / dielbsep), one wave per column i
__global__ __launch_bounds__(256) void k_xth(int n, const zc* __restrict__ X, const zc* __restrict__ h1,
                                             const zc* __restrict__ h2, double* __restrict__ act) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;
  zc a = mkz(0, 0), b = mkz(0, 0);
  const zc* col = X + (size_t)i * n;
  for (int j = lane; j < n; j += 64) {
    zc x = col[j];
    a = fma_z(x, h1[j], a);
    if (h2) b = fma_z(x, h2[j], b);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a.re += __shfl_xor(a.re, off);
    a.im += __shfl_xor(a.im, off);
    b.re += __shfl_xor(b.re, off);
    b.im += __shfl_xor(b.im, off);
  }
  if (lane == 0) act[i] = h2 ? (a.re * b.re + a.im * b.im) : (a.re * a.re + a.im * a.im);
}

int launch_activities(hipStream_t s, int N, const double* d_G, int dtfull, int cpol, double* d_act) {
  hipLaunchKernelGGL(k_activities, dim3((N + 255) / 256), dim3(256), 0, s, N, d_G, dtfull, cpol, d_act);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_activities_temp(hipStream_t s, int N, const double* d_G, int dtfull, int cpol, double* d_act) {
  hipLaunchKernelGGL(k_activities_temp, dim3((N + 255) / 256), dim3(256), 0, s, N, d_G, dtfull, cpol, d_act);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_quantize(hipStream_t s, int n, const double* src, double* dst) {
  hipLaunchKernelGGL(k_quantize, dim3((n + 255) / 256), dim3(256), 0, s, n, src, dst);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// d_work: >= 1 double.  caux = fl32(180.90) / (REAL(nk) * vc) (optics.f90:188-196)
int launch_spectrum(hipStream_t s, int N, const double* d_e, const double* d_act6, double vc,
                    double nkf32, double ebse0, double ebsef, int numbse, double sme, int renorm,
                    double* d_work, double* d_eps) {
  const double gama = (double)180.90f;
  const double baux = nkf32 * vc;
  const double caux = gama / (baux * 1.0);
  hipLaunchKernelGGL(k_avgsme, dim3(1), dim3(256), 0, s, N, d_e, d_work);
  hipLaunchKernelGGL(k_optdiel, dim3((numbse + OMB - 1) / OMB), dim3(256), 0, s, N, d_e, d_act6, caux,
                     ebse0, ebsef, numbse, sme, d_work, d_eps);
  if (renorm) hipLaunchKernelGGL(k_imagrenorm, dim3(1), dim3(256), 0, s, d_e, ebse0, ebsef, numbse, d_eps);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_imagrenorm(hipStream_t s, const double* d_e, double ebse0, double ebsef, int numbse, double* d_eps) {
  hipLaunchKernelGGL(k_imagrenorm, dim3(1), dim3(256), 0, s, d_e, ebse0, ebsef, numbse, d_eps);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_xth(hipStream_t s, int n, const zc* X, const zc* h1, const zc* h2, double* act) {
  hipLaunchKernelGGL(k_xth, dim3((n + 3) / 4), dim3(256), 0, s, n, X, h1, h2, act);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb

namespace wtb {

// The single-particle table the reference writes beside the BSE results
// (bse_solver-tool-diel.f90:365-420; bse_solver-tool-diel-temp.f90:405-424): per basis
// state I = (k nc + c) nv + v, in REAL (float32) as vecres(I, 1..15): k (3), nocp, the
// conduction and valence band numbers, Ec - Ev, |h_x|^2, |h_y|^2, |h_z|^2,
// Re(h_x conj h_y), Re(h_x conj h_z), Re(h_y conj h_z), |h_sp|^2, |h_sm|^2; at finite
// temperature each of the last eight times f_v - f_c (REAL value times the double factor).
__global__ void k_sp_table(int N, int nk, int nc, int nv, const double* __restrict__ kpt,
                           const int* __restrict__ nocp, const double* __restrict__ eigv,
                           const zc* __restrict__ dip, const double* __restrict__ fdeh, float* __restrict__ tab) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= N) return;
  const int ncv = nc * nv, nb = nc + nv;
  const int k = I / ncv, c = (I % ncv) / nv, v = I % nv;
  const int no = nocp[k];
  const zc hx = dip[I], hy = dip[(size_t)N + I], hz = dip[2 * (size_t)N + I];
  // hrsp = (hrx + i hry) * REAL(1./sqrt(2.)), hrsm = (hrx - i hry) * ... (:376-377)
  const double s2 = (double)(1.0f / sqrtf(2.0f));
  const zc hp = mkz(hx.re - hy.im, hx.im + hy.re) * s2;
  const zc hm = mkz(hx.re + hy.im, hx.im - hy.re) * s2;
  auto re_ab = [](zc a, zc b) { return (a * conj(b)).re; };   // real(a*conjg(b))
  const double q[8] = {re_ab(hx, hx), re_ab(hy, hy), re_ab(hz, hz), re_ab(hx, hy),
                       re_ab(hx, hz), re_ab(hy, hz), re_ab(hp, hp), re_ab(hm, hm)};
  tab[I] = (float)kpt[k];
  tab[(size_t)N + I] = (float)kpt[(size_t)nk + k];
  tab[2 * (size_t)N + I] = (float)kpt[2 * (size_t)nk + k];
  tab[3 * (size_t)N + I] = (float)no;
  tab[4 * (size_t)N + I] = (float)(no + c + 1);
  tab[5 * (size_t)N + I] = (float)(no - nv + v + 1);
  tab[6 * (size_t)N + I] = (float)(eigv[(size_t)k * nb + nv + c] - eigv[(size_t)k * nb + v]);
  for (int j = 0; j < 8; ++j) {
    const float r = (float)q[j];
    tab[(size_t)(7 + j) * N + I] = fdeh ? (float)((double)r * fdeh[I]) : r;
  }
}

// Bubblem (general_subs.f90:405-440) on column 7: a stable ascending sort, here by rank
// (the states of lower energy, then those of equal energy and lower index)
__global__ void k_sp_sort(int N, const float* __restrict__ tab, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const float* key = tab + 6 * (size_t)N;
  const float ki = key[i];
  int rank = 0;
  for (int j = 0; j < N; ++j) {
    const float kj = key[j];
    rank += (kj < ki) || (kj == ki && j < i);
  }
  for (int c = 0; c < 15; ++c) out[(size_t)c * N + rank] = tab[(size_t)c * N + i];
}

int launch_sp_table(hipStream_t s, int N, int nk, int nc, int nv, const double* kpt, const int* nocp,
                    const double* eigv, const zc* dip, const double* fdeh, float* work, float* out) {
  hipLaunchKernelGGL(k_sp_table, dim3((N + 255) / 256), dim3(256), 0, s, N, nk, nc, nv, kpt, nocp, eigv, dip, fdeh,
                     work);
  hipLaunchKernelGGL(k_sp_sort, dim3((N + 255) / 256), dim3(256), 0, s, N, work, out);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// ---------------------------------------------------------------------------
// Post-processing of the spectra and the exciton lifetimes, element-wise.
// ---------------------------------------------------------------------------

// Optical constants of one (photon energy, component) from eps = e1 + i e2
// (subroutines/diel-pp-subs.f90:1-73, called by bseoptprop, subprograms/diel-pp-bse.f90:
// 89-140, and its spoptprop / -pol variants), in the reference's operation order and with
// its single-precision literals (sqrt(2.0) and hc = 19.746E-06 are REAL):
//   n = sqrt((|eps| + e1)/2), k = sqrt((|eps| - e1)/2), R = ((n-1)^2 + k^2)/((n+1)^2 + k^2),
//   alpha = sqrt(2) E (1/hc) sqrt(|eps| - e1), L = e2/|eps|^2.
// out[q][c][i], q = 0..4 for n, k, R, alpha, L.
__global__ __launch_bounds__(256) void k_optprop(int n, int nc, const double* __restrict__ en,
                                                 const double* __restrict__ er, const double* __restrict__ ei,
                                                 double* __restrict__ out) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (i >= n || c >= nc) return;
  const double e1 = er[(size_t)c * n + i], e2 = ei[(size_t)c * n + i];
  const double aux = sqrt((e1 * e1) + (e2 * e2));
  const double refr = sqrt((aux + e1) / 2.0);
  const double ext = sqrt((aux - e1) / 2.0);
  const double a1 = (refr - 1.0) * (refr - 1.0);
  const double a2 = (refr + 1.0) * (refr + 1.0);
  const double a3 = ext * ext;
  const double refl = (a1 + a3) / (a2 + a3);
  const double sqrt2f = (double)sqrtf(2.0f);
  const double hc = (double)19.746E-06f;
  const double absc = ((sqrt2f * en[i]) * (1.0 / hc)) * sqrt(aux - e1);
  const double eloss = e2 / ((e1 * e1) + (e2 * e2));
  const size_t plane = (size_t)nc * n, at = (size_t)c * n + i;
  out[at] = refr;
  out[plane + at] = ext;
  out[2 * plane + at] = refl;
  out[3 * plane + at] = absc;
  out[4 * plane + at] = eloss;
}

// exclft (subroutines/bse_subs.f90:472-530): radiative lifetime (s) of an exciton of
// energy E and oscillator strength f for sysdim 3 (PRL 95, 247402), 2 (Nano Lett. 15, 2794)
// or 1, with the REAL literals of the reference (cfe 7.297E-3, lspd 299.792E16, hbar
// 6.582E-16, pi = acos(-1.)) held in doubles.  nkd = dble(n1 n2 n3), vc2d = |a1 x a2|,
// a11 = rlat(1,1).
__device__ double exclft_dev(int sysdim, double nkd, double vc2d, double a11, double f, double E) {
#pragma clang fp contract(off)
  const double cfe = (double)7.297E-3f, lspd = (double)299.792E16f, hbar = (double)6.582E-16f;
  const double pi = (double)3.14159274101257324;
  if (sysdim == 3) {
    const double aux1 = (4.0 * cfe) * ((E * E) * E);
    const double aux2 = f / nkd;
    const double aux3 = (3.0 * (lspd * lspd)) * ((hbar * hbar) * hbar);
    const double l = (aux1 * aux2) / aux3;
    return 1.0 / l;
  } else if (sysdim == 2) {
    const double aux1 = (nkd * vc2d) * hbar;
    const double aux2 = (((8.0 * pi) * cfe) * E) * f;
    return aux1 / aux2;
  }
  const double aux2 = ((((2.0 * pi) * cfe) * E) * E) * f;
  const double aux1 = (((nkd * a11) * hbar) * hbar) * lspd;
  return aux1 / aux2;
}

// lifetime / lifetimepol (subprograms/exciton_lifetime.f90:2-92, 95-187): per exciton i
// of table fosc(n, 1 + ncomp) (energy, then the activities as bse_opt_diel.dat holds
// them: xx, yy, zz, ... or x, y, z, sp, sm), lft(n, 1 + nout): first the lifetime of the
// summed strength f_1 + f_2 + f_3, then one per component 1..nout.
__global__ __launch_bounds__(256) void k_lifetime(int n, int ncol, int nout, int sysdim, double nkd, double vc2d,
                                                  double a11, const double* __restrict__ fosc,
                                                  double* __restrict__ lft) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double E = fosc[i];
  const double fsum = (fosc[(size_t)n + i] + fosc[(size_t)2 * n + i]) + fosc[(size_t)3 * n + i];
  lft[i] = exclft_dev(sysdim, nkd, vc2d, a11, fsum, E);
  for (int c = 1; c <= nout; ++c) lft[(size_t)c * n + i] = exclft_dev(sysdim, nkd, vc2d, a11, fosc[(size_t)c * n + i], E);
}

int launch_optprop(hipStream_t s, int n, int nc, const double* en, const double* er, const double* ei, double* out) {
  hipLaunchKernelGGL(k_optprop, dim3((n + 255) / 256, nc), dim3(256), 0, s, n, nc, en, er, ei, out);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_lifetime(hipStream_t s, int n, int ncol, int nout, int sysdim, double nkd, double vc2d, double a11,
                    const double* fosc, double* lft) {
  hipLaunchKernelGGL(k_lifetime, dim3((n + 255) / 256), dim3(256), 0, s, n, ncol, nout, sysdim, nkd, vc2d, a11,
                     fosc, lft);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
