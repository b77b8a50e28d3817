// Kernels of the iterative form of stage (d)+(e): the lowest excitons by block
// Davidson on the full Hermitian BSE matrix, and eps(omega) by Lanczos
// continued fractions (Haydock) instead of all N oscillator strengths.
//
// eps(omega) (optdiel, subroutines/optics.f90:160-244) is a sum over excitons of
// f_n L(omega - E_n) with f_n = |A_n^T h|^2 (no conjugate on A, bse_subs.f90:336-378).
// conj(A_n) is an eigenvector of conj(H) with the same E_n, so
//   sum_n f_n / (E_n - z) = conj(h)^H (H - z)^-1 conj(h),
// the resolvent of H on the start vector conj(h): a Lanczos chain from conj(h)
// gives it as the continued fraction |h|^2 / (z - a0 - b1^2 / (z - a1 - ...)) with
// z = omega + i gamma.  Off-diagonal strengths Re(t_a conj(t_b)) follow from the
// chains of h_a + h_b and h_a - h_b (polarisation).
#include "wtb_kernels.h"

namespace wtb {

__device__ inline zf conjT(zf v) { v.im = -v.im; return v; }
__device__ inline zc conjT(zc v) { v.im = -v.im; return v; }

// ---- full Hermitian matrix from the lower triangle: 64 x 64 tiles through LDS ----------
template <class T>
__global__ __launch_bounds__(256) void k_sym_fill(int N, T* __restrict__ H, int ld) {
  __shared__ T tile[64][65];
  int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int R0 = ti * 64, C0 = tj * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int rl = e & 63, cl = e >> 6;
    const int r = R0 + rl, c = C0 + cl;
    if (r < N && c < N && r >= c) tile[cl][rl] = H[(size_t)r + (size_t)ld * c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int cl = e & 63, rl = e >> 6;
    const int r = R0 + rl, c = C0 + cl;
    if (r < N && c < N && r >= c) {
      T v = tile[cl][rl];
      if (r == c) {
        v.im = 0;
        H[(size_t)r + (size_t)ld * r] = v;
      } else {
        H[(size_t)c + (size_t)ld * r] = conjT(v);
      }
    }
  }
}

// ---- per basis state i: D_i = Re H_ii; over j != i: sum |H_ij|, sqrt(sum |H_ij|^2), max |H_ij|
template <class T>
__global__ __launch_bounds__(256) void k_row_stats(int N, const T* __restrict__ H, int ld, double* __restrict__ D,
                                                   float* __restrict__ st) {
  __shared__ double r1[256], r2[256], r3[256];
  const int i = blockIdx.x;
  const T* col = H + (size_t)ld * i;      // column i = conj(row i)
  double s1 = 0.0, s2 = 0.0, mx = 0.0;
  for (int j = threadIdx.x; j < N; j += 256) {
    if (j == i) continue;
    const zc v = ldz(col[j]);
    const double a2 = norm2(v), a = sqrt(a2);
    s1 += a;
    s2 += a2;
    mx = fmax(mx, a);
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  r3[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
      r3[threadIdx.x] = fmax(r3[threadIdx.x], r3[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    D[i] = ldz(col[i]).re;
    st[i] = (float)r1[0];
    st[(size_t)N + i] = (float)sqrt(r2[0]);
    st[2 * (size_t)N + i] = (float)r3[0];
  }
}

// ---- R(:,c) = HU(:,c) - theta_c U(:,c) and its norm, one workgroup per column --------
__global__ __launch_bounds__(256) void k_residual(int N, const zc* __restrict__ U, const zc* __restrict__ HU,
                                                  int ld, const double* __restrict__ theta, zc* __restrict__ R,
                                                  double* __restrict__ rn) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  const double th = theta[c];
  double s = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) {
    const size_t o = (size_t)i + (size_t)ld * c;
    const zc r = HU[o] - th * U[o];
    R[o] = r;
    s += norm2(r);
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rn[c] = sqrt(red[0]);
}

// ---- Rayleigh-quotient refinement of exciton c (one workgroup per column) -------------------
// theta = Re(x^H H x) / x^H x from the vector x and its FP64 product Hx; the energy lam[c] is
// kept in lam0[c] and replaced by theta; res[c] = |Hx - theta x| / |x|, the distance within
// which the FP64 matrix has an eigenvalue.  Sums in a fixed tree order.
__global__ __launch_bounds__(256) void k_rq_refine(int N, const zc* __restrict__ X, const zc* __restrict__ HX,
                                                   int ld, double* __restrict__ lam, double* __restrict__ lam0,
                                                   double* __restrict__ res) {
  __shared__ double ra[256], rb[256];
  const int c = blockIdx.x, t = threadIdx.x;
  double xx = 0.0, xhx = 0.0;
  for (int i = t; i < N; i += 256) {
    const size_t o = (size_t)i + (size_t)ld * c;
    const zc x = X[o], h = HX[o];
    xx += norm2(x);
    xhx += x.re * h.re + x.im * h.im;
  }
  ra[t] = xx;
  rb[t] = xhx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      ra[t] += ra[t + w];
      rb[t] += rb[t + w];
    }
    __syncthreads();
  }
  const double nx = ra[0], th = rb[0] / nx;
  __syncthreads();
  double r2 = 0.0;
  for (int i = t; i < N; i += 256) {
    const size_t o = (size_t)i + (size_t)ld * c;
    r2 += norm2(HX[o] - th * X[o]);
  }
  ra[t] = r2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) ra[t] += ra[t + w];
    __syncthreads();
  }
  if (t == 0) {
    lam0[c] = lam[c];
    lam[c] = th;
    res[c] = sqrt(ra[0] / nx);
  }
}

// ---- diagonal preconditioner: T(:,q) = R(:,idx q) / (D - theta_idx q), |den| >= dmin --------
__global__ __launch_bounds__(256) void k_precond_diag(int N, const int* __restrict__ idx, const zc* __restrict__ R,
                                                      int ld, const double* __restrict__ theta,
                                                      const double* __restrict__ D, double dmin,
                                                      zc* __restrict__ T, int ldt) {
  const int q = blockIdx.y, c = idx[q];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  double den = D[i] - theta[c];
  if (fabs(den) < dmin) den = den < 0.0 ? -dmin : dmin;
  T[(size_t)i + (size_t)ldt * q] = R[(size_t)i + (size_t)ld * c] * (1.0 / den);
}

// ---- Cholesky of a Gram matrix and the inverse of its factor (one workgroup) ----------
// G (b x b Hermitian, lower used) = L L^H, Linv = L^-1 (b x b, column-major).  A pivot
// below tol * max(diag) drops its column (keep = 0; L(:,k) = e_k, so the remaining
// columns of T Linv^H do not involve it).
__global__ __launch_bounds__(256) void k_chol_inv(int b, zc* __restrict__ G, int ld, zc* __restrict__ Linv,
                                                  int* __restrict__ keep, double tol) {
  __shared__ double s_piv, s_dmax;
  __shared__ int s_drop;
  const int tid = threadIdx.x;
  if (tid == 0) {
    double m = 0.0;
    for (int i = 0; i < b; ++i) m = fmax(m, G[(size_t)i + (size_t)ld * i].re);
    s_dmax = m;
  }
  __syncthreads();
  const double thr = tol * s_dmax;
  for (int k = 0; k < b; ++k) {
    if (tid == 0) {
      const double p = G[(size_t)k + (size_t)ld * k].re;
      s_drop = !(p > thr);
      s_piv = s_drop ? 1.0 : sqrt(p);
      keep[k] = s_drop ? 0 : 1;
    }
    __syncthreads();
    const double piv = s_piv;
    const int drop = s_drop;
    for (int i = k + tid; i < b; i += 256) {
      zc& g = G[(size_t)i + (size_t)ld * k];
      if (i == k) g = mkz(piv, 0.0);
      else g = drop ? mkz(0, 0) : g * (1.0 / piv);
    }
    __syncthreads();
    if (!drop) {
      const int m = b - k - 1;
      for (int e = tid; e < m * m; e += 256) {
        const int ii = k + 1 + e % m, jj = k + 1 + e / m;
        if (ii < jj) continue;
        const zc lik = G[(size_t)ii + (size_t)ld * k], ljk = G[(size_t)jj + (size_t)ld * k];
        G[(size_t)ii + (size_t)ld * jj] -= lik * conj(ljk);
      }
    }
    __syncthreads();
  }
  // Linv column c by forward substitution, one thread per column
  for (int c = tid; c < b; c += 256) {
    for (int i = 0; i < b; ++i) {
      zc acc = mkz(i == c ? 1.0 : 0.0, 0.0);
      if (i > c) {
        acc = mkz(0, 0);
        for (int j = c; j < i; ++j) acc -= G[(size_t)i + (size_t)ld * j] * Linv[(size_t)j + (size_t)b * c];
      }
      if (i < c) {
        Linv[(size_t)i + (size_t)b * c] = mkz(0, 0);
        continue;
      }
      const double d = G[(size_t)i + (size_t)ld * i].re;
      Linv[(size_t)i + (size_t)b * c] = acc * (1.0 / d);
    }
  }
}

// ---- blocked Cholesky and inverse factor (b <= 512) -------------------------------------
// k_chol_left: one workgroup of 1024 threads, left-looking by panels of CP (8 or 16) columns.  The
// panel (rows p..b-1) is staged in LDS beside the rows L(p:p+CP, 0:p) that update it; the
// update with the previous columns runs in four K quarters added in a fixed order, then
// the panel is factored in LDS (same pivot rule as k_chol_inv) and written back.
template <int CP>
__global__ __launch_bounds__(1024) void k_chol_left(int b, zc* __restrict__ G, int ld, int* __restrict__ keep,
                                                    double tol) {
  extern __shared__ zc sm[];
  zc* P = sm;                      // P[lr * CP + c] = (p + lr, p + c)
  zc* Lr = sm + (size_t)b * CP;    // Lr[k * CP + c] = L(p + c, k), k < p
  __shared__ double s_red[16];
  __shared__ double s_piv;
  __shared__ int s_drop;
  const int tid = threadIdx.x;
  double mx = 0.0;
  for (int i = tid; i < b; i += 1024) mx = fmax(mx, G[(size_t)i + (size_t)ld * i].re);
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off));
  if ((tid & 63) == 0) s_red[tid >> 6] = mx;
  __syncthreads();
  double dmax = 0.0;
  for (int w = 0; w < 16; ++w) dmax = fmax(dmax, s_red[w]);
  const double thr = tol * dmax;
  const int ks = tid >> 8, rl = tid & 255;
  for (int p = 0; p < b; p += CP) {
    const int pc = min(CP, b - p), nr = b - p;
    for (int e = tid; e < nr * pc; e += 1024) {
      const int c = e / nr, lr = e - c * nr;
      P[lr * CP + c] = (lr >= c) ? G[(size_t)(p + lr) + (size_t)ld * (p + c)] : mkz(0, 0);
    }
    for (int e = tid; e < p * pc; e += 1024) {
      const int k = e / pc, c = e - k * pc;
      Lr[k * CP + c] = G[(size_t)(p + c) + (size_t)ld * k];
    }
    __syncthreads();
    if (p > 0) {
      const int kq = (p + 3) / 4;
      const int kb = ks * kq, ke = min(p, kb + kq);
      for (int r0 = 0; r0 < nr; r0 += 256) {
        const int lr = r0 + rl;
        zc acc[CP];
#pragma unroll
        for (int c = 0; c < CP; ++c) acc[c] = mkz(0, 0);
        if (lr < nr) {
          // UL loads in flight (fewer for the wider panel: 1024 threads leave 128 registers
          // each), then their products in k order
          constexpr int UL = CP == 16 ? 2 : 4;
          int k = kb;
          for (; k + UL - 1 < ke; k += UL) {
            zc l[UL];
#pragma unroll
            for (int u = 0; u < UL; ++u) l[u] = G[(size_t)(p + lr) + (size_t)ld * (k + u)];
#pragma unroll
            for (int u = 0; u < UL; ++u)
#pragma unroll
              for (int c = 0; c < CP; ++c) acc[c] = fma_conj(Lr[(k + u) * CP + c], l[u], acc[c]);   // l conj(L(p+c,k))
          }
          for (; k < ke; ++k) {
            const zc l = G[(size_t)(p + lr) + (size_t)ld * k];
#pragma unroll
            for (int c = 0; c < CP; ++c) acc[c] = fma_conj(Lr[k * CP + c], l, acc[c]);
          }
        }
        for (int q = 0; q < 4; ++q) {
          if (ks == q && lr < nr) {
#pragma unroll
            for (int c = 0; c < CP; ++c)
              if (c < pc) P[lr * CP + c] = P[lr * CP + c] - acc[c];
          }
          __syncthreads();
        }
      }
    }
    for (int c = 0; c < pc; ++c) {
      if (tid == 0) {
        const double pv = P[c * CP + c].re;
        const int drop = !(pv > thr);
        s_drop = drop;
        s_piv = drop ? 1.0 : sqrt(pv);
        keep[p + c] = drop ? 0 : 1;
      }
      __syncthreads();
      const double piv = s_piv;
      const int drop = s_drop;
      for (int lr = c + tid; lr < nr; lr += 1024) {
        const zc g = P[lr * CP + c];
        P[lr * CP + c] = (lr == c) ? mkz(piv, 0.0) : (drop ? mkz(0, 0) : g * (1.0 / piv));
      }
      __syncthreads();
      const int w = pc - c - 1;   // panel columns right of c
      if (!drop && w > 0) {
        for (int e = tid; e < (nr - c - 1) * w; e += 1024) {
          const int lr = c + 1 + e / w, c2 = c + 1 + e % w;
          if (lr >= c2) P[lr * CP + c2] = P[lr * CP + c2] - P[lr * CP + c] * conj(P[c2 * CP + c]);
        }
      }
      __syncthreads();
    }
    for (int e = tid; e < nr * pc; e += 1024) {
      const int c = e / nr, lr = e - c * nr;
      if (lr >= c) G[(size_t)(p + lr) + (size_t)ld * (p + c)] = P[lr * CP + c];
    }
    __syncthreads();
  }
}

// k_trinv: columns 16 J .. 16 J + 15 of L^-1 (L the lower factor left in G), one workgroup
// per 16 columns: block forward substitution, 16 rows per step, X(:, J) kept in LDS.
__global__ __launch_bounds__(256) void k_trinv(int b, const zc* __restrict__ G, int ld, zc* __restrict__ Linv) {
  extern __shared__ zc X[];        // X[(i - c0) * 16 + c]
  __shared__ zc Lb[16][17];
  const int tid = threadIdx.x;
  const int r = tid >> 4, c = tid & 15;
  const int c0 = blockIdx.x * 16;
  const int nc = min(16, b - c0);
  for (int I0 = c0; I0 < b; I0 += 16) {
    const int ni = min(16, b - I0);
    zc acc = mkz(0, 0);
    if (r < ni && c < nc) {
      int k = c0;
      for (; k + 7 < I0; k += 8) {   // eight loads in flight, products in k order
        zc g8[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) g8[u] = G[(size_t)(I0 + r) + (size_t)ld * (k + u)];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + g8[u] * X[(k + u - c0) * 16 + c];
      }
      for (; k < I0; ++k) acc = acc + G[(size_t)(I0 + r) + (size_t)ld * k] * X[(k - c0) * 16 + c];
    }
    {
      const int rr = tid & 15, kk = tid >> 4;
      Lb[rr][kk] = (rr < ni && kk < ni && rr >= kk) ? G[(size_t)(I0 + rr) + (size_t)ld * (I0 + kk)] : mkz(0, 0);
    }
    __syncthreads();
    const zc rhs = mkz((I0 + r == c0 + c) ? 1.0 : 0.0, 0.0) - acc;
    for (int rr = 0; rr < ni; ++rr) {
      if (r == rr && c < nc) {
        zc v = rhs;
        for (int kk = 0; kk < rr; ++kk) v = v - Lb[rr][kk] * X[(I0 + kk - c0) * 16 + c];
        X[(I0 + rr - c0) * 16 + c] = v * (1.0 / Lb[rr][rr].re);
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < b * nc; e += 256) {
    const int cc = e / b, i = e - cc * b;
    Linv[(size_t)i + (size_t)b * (c0 + cc)] = (i >= c0) ? X[(i - c0) * 16 + cc] : mkz(0, 0);
  }
}

// ---- Lanczos (Haydock) chains, nq columns at once ---------------------------------------
// part 1: w = H v - b_j v_prev; partials of Re(v^H H v) and |w|^2 per workgroup
__global__ __launch_bounds__(256) void k_lz_a(int N, const zc* __restrict__ V, const zc* __restrict__ Vp,
                                              const zc* __restrict__ HV, const double* __restrict__ bcur,
                                              zc* __restrict__ W, double* __restrict__ part) {
  __shared__ double ra[256], rw[256];
  const int q = blockIdx.y;
  const size_t off = (size_t)N * q;
  const double bj = bcur[q];
  double sa = 0.0, sw = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += 256 * gridDim.x) {
    const zc v = V[off + i], hv = HV[off + i];
    sa += v.re * hv.re + v.im * hv.im;
    const zc w = hv - bj * Vp[off + i];
    W[off + i] = w;
    sw += norm2(w);
  }
  ra[threadIdx.x] = sa;
  rw[threadIdx.x] = sw;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      ra[threadIdx.x] += ra[threadIdx.x + s];
      rw[threadIdx.x] += rw[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((size_t)q * gridDim.x + blockIdx.x) * 2] = ra[0];
    part[((size_t)q * gridDim.x + blockIdx.x) * 2 + 1] = rw[0];
  }
}

// part 2: a_j = sum of the partials (every workgroup sums them in the same order;
// workgroup 0 records a_j); w' = w - a_j v into Vn, partials of |w'|^2
__global__ __launch_bounds__(256) void k_lz_b(int N, int j, int kmax, const zc* __restrict__ V,
                                              const zc* __restrict__ W, const double* __restrict__ part,
                                              int nblk, double* __restrict__ acoef, zc* __restrict__ Vn,
                                              double* __restrict__ part2) {
  __shared__ double s_a;
  __shared__ double rw[256];
  const int q = blockIdx.y;
  if (threadIdx.x == 0) {
    double a = 0.0;
    for (int k = 0; k < nblk; ++k) a += part[((size_t)q * nblk + k) * 2];
    s_a = a;
    if (blockIdx.x == 0) acoef[(size_t)q * kmax + j] = a;
  }
  __syncthreads();
  const double a = s_a;
  const size_t off = (size_t)N * q;
  double sw = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += 256 * gridDim.x) {
    const zc w = W[off + i] - a * V[off + i];
    Vn[off + i] = w;
    sw += norm2(w);
  }
  rw[threadIdx.x] = sw;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) rw[threadIdx.x] += rw[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part2[(size_t)q * gridDim.x + blockIdx.x] = rw[0];
}

// part 3: b_{j+1} = |w'|, v_{j+1} = w' / b_{j+1}
__global__ __launch_bounds__(256) void k_lz_c(int N, int j, int kmax, const double* __restrict__ part2, int nblk,
                                              double* __restrict__ bcoef, double* __restrict__ bcur,
                                              zc* __restrict__ Vn) {
  __shared__ double s_b;
  const int q = blockIdx.y;
  if (threadIdx.x == 0) {
    double w2 = 0.0;
    for (int k = 0; k < nblk; ++k) w2 += part2[(size_t)q * nblk + k];
    s_b = sqrt(w2);
    if (blockIdx.x == 0) {
      bcoef[(size_t)q * kmax + j] = s_b;
      bcur[q] = s_b;
    }
  }
  __syncthreads();
  const double b = s_b;
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  const size_t off = (size_t)N * q;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += 256 * gridDim.x) Vn[off + i] = Vn[off + i] * inv;
}

// start vectors conj(h), conj(h_a + h_b), conj(h_a - h_b) (normalised by the host-side
// norms computed in k_lz_norms)
// f (optional, finite temperature): the second vector of a pair is f h_b
__global__ __launch_bounds__(256) void k_lz_start(int N, int nq, const zc* __restrict__ dip,
                                                  const int* __restrict__ ca, const int* __restrict__ cb,
                                                  const double* __restrict__ sg, const double* __restrict__ f,
                                                  zc* __restrict__ V) {
  const int q = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N || q >= nq) return;
  zc h = dip[(size_t)ca[q] * N + i];
  if (cb[q] >= 0) {
    if (f) h = h + (sg[q] * f[i]) * dip[(size_t)cb[q] * N + i];
    else h = h + sg[q] * dip[(size_t)cb[q] * N + i];
  }
  V[(size_t)q * N + i] = conj(h);
}

__global__ __launch_bounds__(256) void k_lz_norms(int N, zc* __restrict__ V, double* __restrict__ hn2,
                                                  zc* __restrict__ Vprev, double* __restrict__ bcur) {
  __shared__ double red[256];
  const int q = blockIdx.x;
  zc* v = V + (size_t)N * q;
  double s = 0.0;
  for (int i = threadIdx.x; i < N; i += 256) s += norm2(v[i]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double n2 = red[0];
  const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 0.0;
  for (int i = threadIdx.x; i < N; i += 256) {
    v[i] = v[i] * inv;
    Vprev[(size_t)N * q + i] = mkz(0, 0);
  }
  if (threadIdx.x == 0) {
    hn2[q] = n2;
    bcur[q] = 0.0;
  }
}

// largest eigenvalue of the chain-0 Lanczos tridiagonal (a_0..a_{k-1}, b_1..b_{k-1}) by
// Sturm-count bisection; gamma = sme + 10 (E_max - E_1) / N as avgsme telescopes
__global__ void k_lz_gamma(int k, const double* __restrict__ acoef, const double* __restrict__ bcoef,
                           const double* __restrict__ E1, int N, double* __restrict__ smeavg) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < k; ++i) {
    const double r = (i > 0 ? fabs(bcoef[i - 1]) : 0.0) + (i + 1 < k ? fabs(bcoef[i]) : 0.0);
    lo = fmin(lo, acoef[i] - r);
    hi = fmax(hi, acoef[i] + r);
  }
  // count eigenvalues below x; the largest is where the count reaches k
  for (int it = 0; it < 200; ++it) {
    const double x = 0.5 * (lo + hi);
    if (x <= lo || x >= hi) break;
    int cnt = 0;
    double d = 1.0;
    for (int i = 0; i < k; ++i) {
      const double b2 = i > 0 ? bcoef[i - 1] * bcoef[i - 1] : 0.0;
      d = (acoef[i] - x) - (i > 0 ? b2 / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) ++cnt;
    }
    if (cnt >= k) hi = x; else lo = x;
  }
  smeavg[0] = 10.0 * ((hi - E1[0]) / (double)N);
  smeavg[1] = hi;
}

// eps(omega) from the chains: S_q(z) = -|h_q|^2 G_q(z), z = omega + i gamma; the tail past
// step k is closed by the square-root terminator of the last coefficients' mean.
__device__ inline zc zdiv(zc a, zc b) {
  const double d = b.re * b.re + b.im * b.im;
  return mkz((a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d);
}
__device__ inline zc zsqrt(zc a) {
  const double r = sqrt(sqrt(a.re * a.re + a.im * a.im));
  const double t = 0.5 * atan2(a.im, a.re);
  return mkz(r * cos(t), r * sin(t));
}
__global__ __launch_bounds__(256) void k_haydock_eps(int nq, int k, int kmax, const double* __restrict__ acoef,
                                                     const double* __restrict__ bcoef, const double* __restrict__ hn2,
                                                     double caux, double ebse0, double ebsef, int numbse,
                                                     double sme, const double* __restrict__ smeavg, int dtfull,
                                                     int temp, double* __restrict__ eps) {
  const int jw = blockIdx.x * 256 + threadIdx.x;
  if (jw >= numbse) return;
  const double om = ebse0 + ((ebsef - ebse0) * (double)jw) / ((double)numbse - 1.0);
  const double g = sme + smeavg[0];
  const zc z = mkz(om, g);
  const int nt = k / 4 > 8 ? 8 : (k / 4 > 0 ? k / 4 : 1);
  double S[12][2];
  for (int q = 0; q < nq; ++q) {
    const double* a = acoef + (size_t)q * kmax;
    const double* b = bcoef + (size_t)q * kmax;   // b[j] = b_{j+1}
    double am = 0.0, bm = 0.0;
    for (int i = k - nt; i < k; ++i) {
      am += a[i];
      bm += b[i - 1 >= 0 ? i - 1 : 0];
    }
    am /= nt;
    bm /= nt;
    // terminator t(z) = ((z-a) - sqrt((z-a)^2 - 4b^2)) / (2 b^2), the root with Im t <= 0
    zc t = mkz(0, 0);
    if (bm > 0.0) {
      const zc za = mkz(z.re - am, z.im);
      const zc disc = za * za - mkz(4.0 * bm * bm, 0.0);
      zc sq = zsqrt(disc);
      zc t1 = (za - sq) * (1.0 / (2.0 * bm * bm));
      zc t2 = (za + sq) * (1.0 / (2.0 * bm * bm));
      t = (t1.im <= 0.0) ? t1 : t2;
      if (t1.im <= 0.0 && t2.im <= 0.0) t = (norm2(t1) < norm2(t2)) ? t1 : t2;
    }
    // G = 1/(z - a_0 - b_1^2/(z - a_1 - ... - b_{k-1}^2 /(z - a_{k-1} - b_k^2 t)))
    zc G = t;
    for (int i = k - 1; i >= 0; --i) {
      const double bi = b[i];     // b_{i+1}
      const zc den = mkz(z.re - a[i], z.im) - (bi * bi) * G;
      G = zdiv(mkz(1, 0), den);
    }
    // S = sum f/(E - z) = -|h|^2 G;  eps1 part Re S, eps2 part Im S
    S[q][0] = -hn2[q] * G.re;
    S[q][1] = -hn2[q] * G.im;
  }
  for (int c = 0; c < 6; ++c) {
    double sr = 0.0, si = 0.0;
    if (temp) {
      // finite temperature: Re[t_a conj(u_b)] = (|t_a + u_b|^2 - |t_a - u_b|^2) / 4 with
      // t = X^T h, u = X^T (f h): chains of h_a + f h_b, h_a - f h_b, every component
      if (c < 3 || dtfull) {
        const int qp = 2 * c;
        sr = 0.25 * (S[qp][0] - S[qp + 1][0]);
        si = 0.25 * (S[qp][1] - S[qp + 1][1]);
      }
    } else if (c < 3) {
      sr = S[c][0];
      si = S[c][1];
    } else if (dtfull) {
      const int qp = 3 + 2 * (c - 3);   // chains of h_a + h_b, h_a - h_b
      sr = 0.25 * (S[qp][0] - S[qp + 1][0]);
      si = 0.25 * (S[qp][1] - S[qp + 1][1]);
    }
    eps[jw + (size_t)numbse * (2 * c)] = (c < 3 ? 1.0 : 0.0) + caux * sr;
    eps[jw + (size_t)numbse * (2 * c + 1)] = caux * si;
  }
}

// ---------------------------------------------------------------------------
int sym_fill_f32(hipStream_t s, int N, zf* H, int ld) {
  const int nt = (N + 63) / 64;
  hipLaunchKernelGGL(k_sym_fill<zf>, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, N, H, ld);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int sym_fill_f64(hipStream_t s, int N, zc* H, int ld) {
  const int nt = (N + 63) / 64;
  hipLaunchKernelGGL(k_sym_fill<zc>, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, N, H, ld);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int row_stats_f32(hipStream_t s, int N, const zf* H, int ld, double* D, float* st) {
  hipLaunchKernelGGL(k_row_stats<zf>, dim3(N), dim3(256), 0, s, N, H, ld, D, st);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int launch_residual(hipStream_t s, int N, int nc, const zc* U, const zc* HU, int ld, const double* theta,
                    zc* R, double* rn) {
  if (nc <= 0) return WTB_OK;
  hipLaunchKernelGGL(k_residual, dim3(nc), dim3(256), 0, s, N, U, HU, ld, theta, R, rn);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int launch_rq_refine(hipStream_t s, int N, int nc, const zc* X, const zc* HX, int ld, double* lam, double* lam0,
                     double* res) {
  if (nc <= 0) return WTB_OK;
  hipLaunchKernelGGL(k_rq_refine, dim3(nc), dim3(256), 0, s, N, X, HX, ld, lam, lam0, res);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int launch_precond_diag(hipStream_t s, int N, int nq, const int* idx, const zc* R, int ld,
                        const double* theta, const double* D, double dmin, zc* T, int ldt) {
  if (nq <= 0) return WTB_OK;
  hipLaunchKernelGGL(k_precond_diag, dim3((N + 255) / 256, nq), dim3(256), 0, s, N, idx, R, ld, theta, D,
                     dmin, T, ldt);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}
int launch_chol_inv(hipStream_t s, int b, zc* G, int ld, zc* Linv, int* keep, double tol) {
  if (b <= 0) return WTB_OK;
  if (b <= 512) {
    // panels of 16 columns while the panel and its L rows (2 b 16 complex) fit in LDS
    const size_t lds = sizeof(zc) * (size_t)b * 16;
    WTB_HIP(hipFuncSetAttribute((const void*)k_trinv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (b <= 288) {
      const size_t lds16 = 2 * lds;
      WTB_HIP(hipFuncSetAttribute((const void*)k_chol_left<16>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds16));
      hipLaunchKernelGGL(k_chol_left<16>, dim3(1), dim3(1024), lds16, s, b, G, ld, keep, tol);
    } else {
      WTB_HIP(hipFuncSetAttribute((const void*)k_chol_left<8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds));
      hipLaunchKernelGGL(k_chol_left<8>, dim3(1), dim3(1024), lds, s, b, G, ld, keep, tol);
    }
    hipLaunchKernelGGL(k_trinv, dim3((b + 15) / 16), dim3(256), lds, s, b, G, ld, Linv);
    WTB_LAUNCH_CHECK();
    return WTB_OK;
  }
  hipLaunchKernelGGL(k_chol_inv, dim3(1), dim3(256), 0, s, b, G, ld, Linv, keep, tol);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// at most 12 chains: 9 at T = 0 (x, y, z and the pairs of the cross terms), 12 at finite
// temperature (a pair for every component)
constexpr int kLzMaxChains = 12;

size_t haydock_work_bytes(int N, int steps) {
  const size_t nblk = 128, m = kLzMaxChains;
  return thin_work_bytes(N, kLzMaxChains) + 256 + sizeof(zc) * (size_t)N * m * 5 +
         sizeof(double) * (m * nblk * 3 + 2 * m * (size_t)(steps + 1) + 64) + sizeof(int) * 64 + 16 * 256;
}

namespace {
// the chains' buffers, carved from the workspace in one fixed order (haydock_work_bytes)
struct HaydockBufs {
  zc* Vb[3];
  zc *HV, *W, *hw;
  double *part, *part2, *acoef, *bcoef, *hn2, *bcur, *sg;
  int *ca, *cb;
};
constexpr int kLzBlocks = 128;
HaydockBufs haydock_carve(void* work, int N, int steps, int nq) {
  HaydockBufs b;
  char* w = static_cast<char*>(work);
  auto take = [&](size_t bytes) {
    char* q = w;
    w += (bytes + 255) & ~size_t(255);
    return q;
  };
  const size_t m = kLzMaxChains;
  for (int i = 0; i < 3; ++i) b.Vb[i] = (zc*)take(sizeof(zc) * (size_t)N * m);
  b.HV = (zc*)take(sizeof(zc) * (size_t)N * m);
  b.W = (zc*)take(sizeof(zc) * (size_t)N * m);
  b.part = (double*)take(sizeof(double) * m * kLzBlocks * 2);
  b.part2 = (double*)take(sizeof(double) * m * kLzBlocks);
  b.acoef = (double*)take(sizeof(double) * m * (size_t)(steps + 1));
  b.bcoef = (double*)take(sizeof(double) * m * (size_t)(steps + 1));
  b.hn2 = (double*)take(sizeof(double) * 16);
  b.bcur = (double*)take(sizeof(double) * 16);
  b.sg = (double*)take(sizeof(double) * 16);
  b.ca = (int*)take(sizeof(int) * 16);
  b.cb = (int*)take(sizeof(int) * 16);
  b.hw = (zc*)take(thin_work_bytes(N, kLzMaxChains));
  (void)nq;
  return b;
}
}  // namespace

int run_haydock_chains(hipStream_t s, int N, const zf* H, int ldh, const zc* dip, int dtfull, int steps,
                       void* work, const double* fdeh) {
  const int nq = fdeh ? (dtfull ? 12 : 6) : (dtfull ? 9 : 3);
  if (steps < 2 || N < 2) {
    set_error("haydock: needs N >= 2 and at least 2 steps");
    return WTB_ERR_ARG;
  }
  HaydockBufs b = haydock_carve(work, N, steps, nq);
  // chains: x, y, z, x+y, x-y, x+z, x-z, y+z, y-z (static: the copies are asynchronous)
  static const int hca[9] = {0, 1, 2, 0, 0, 0, 0, 1, 1};
  static const int hcb[9] = {-1, -1, -1, 1, 1, 2, 2, 2, 2};
  static const double hsg[9] = {0, 0, 0, 1, -1, 1, -1, 1, -1};
  // finite temperature: h_a + f h_b and h_a - f h_b for xx, yy, zz, xy, xz, yz
  static const int tca[12] = {0, 0, 1, 1, 2, 2, 0, 0, 0, 0, 1, 1};
  static const int tcb[12] = {0, 0, 1, 1, 2, 2, 1, 1, 2, 2, 2, 2};
  static const double tsg[12] = {1, -1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1};
  WTB_HIP(hipMemcpyAsync(b.ca, fdeh ? tca : hca, fdeh ? sizeof(tca) : sizeof(hca), hipMemcpyHostToDevice, s));
  WTB_HIP(hipMemcpyAsync(b.cb, fdeh ? tcb : hcb, fdeh ? sizeof(tcb) : sizeof(hcb), hipMemcpyHostToDevice, s));
  WTB_HIP(hipMemcpyAsync(b.sg, fdeh ? tsg : hsg, fdeh ? sizeof(tsg) : sizeof(hsg), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_lz_start, dim3((N + 255) / 256, nq), dim3(256), 0, s, N, nq, dip, b.ca, b.cb, b.sg, fdeh,
                     b.Vb[1]);
  hipLaunchKernelGGL(k_lz_norms, dim3(nq), dim3(256), 0, s, N, b.Vb[1], b.hn2, b.Vb[0], b.bcur);
  WTB_LAUNCH_CHECK();
  const int kmax = steps + 1, nblk = kLzBlocks;
  int ip = 0, ic = 1, in = 2;
  for (int j = 0; j < steps; ++j) {
    int rc = hmm_thin_f32(s, N, nq, H, ldh, b.Vb[ic], N, b.HV, N, b.hw);
    if (rc) return rc;
    hipLaunchKernelGGL(k_lz_a, dim3(nblk, nq), dim3(256), 0, s, N, b.Vb[ic], b.Vb[ip], b.HV, b.bcur, b.W, b.part);
    hipLaunchKernelGGL(k_lz_b, dim3(nblk, nq), dim3(256), 0, s, N, j, kmax, b.Vb[ic], b.W, b.part, nblk, b.acoef,
                       b.Vb[in], b.part2);
    hipLaunchKernelGGL(k_lz_c, dim3(nblk, nq), dim3(256), 0, s, N, j, kmax, b.part2, nblk, b.bcoef, b.bcur,
                       b.Vb[in]);
    const int t = ip;
    ip = ic;
    ic = in;
    in = t;
  }
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int run_haydock_eps(hipStream_t s, int N, int dtfull, int steps, const double* d_E1, double caux, double ebse0,
                    double ebsef, int numbse, double sme, int renorm, void* work, double* d_eps, double* d_smeavg,
                    double* h_coef, int temp) {
  const int nq = temp ? (dtfull ? 12 : 6) : (dtfull ? 9 : 3);
  HaydockBufs b = haydock_carve(work, N, steps, nq);
  const int kmax = steps + 1;
  hipLaunchKernelGGL(k_lz_gamma, dim3(1), dim3(1), 0, s, steps, b.acoef, b.bcoef, d_E1, N, d_smeavg);
  hipLaunchKernelGGL(k_haydock_eps, dim3((numbse + 255) / 256), dim3(256), 0, s, nq, steps, kmax, b.acoef,
                     b.bcoef, b.hn2, caux, ebse0, ebsef, numbse, sme, d_smeavg, dtfull, temp, d_eps);
  WTB_LAUNCH_CHECK();
  if (renorm) {
    int rc = launch_imagrenorm(s, d_E1, ebse0, ebsef, numbse, d_eps);
    if (rc) return rc;
  }
  if (h_coef) {
    WTB_HIP(hipMemcpyAsync(h_coef, b.acoef, sizeof(double) * nq * kmax, hipMemcpyDeviceToHost, s));
    WTB_HIP(hipMemcpyAsync(h_coef + (size_t)nq * kmax, b.bcoef, sizeof(double) * nq * kmax, hipMemcpyDeviceToHost,
                           s));
  }
  return WTB_OK;
}

int run_haydock(hipStream_t s, int N, const zf* H, int ldh, const zc* dip, int dtfull, int steps,
                const double* d_E1, double caux, double ebse0, double ebsef, int numbse, double sme,
                int renorm, void* work, double* d_eps, double* d_smeavg, double* h_coef) {
  int rc = run_haydock_chains(s, N, H, ldh, dip, dtfull, steps, work);
  if (rc) return rc;
  return run_haydock_eps(s, N, dtfull, steps, d_E1, caux, ebse0, ebsef, numbse, sme, renorm, work, d_eps, d_smeavg,
                         h_coef);
}

}  // namespace wtb
