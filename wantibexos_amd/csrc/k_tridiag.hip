// Stage (d1): blocked Householder reduction of a Hermitian matrix (lower
// triangle, column-major, ld) to a real symmetric tridiagonal (d, e) with the
// arithmetic of LAPACK ZHETRD/ZLATRD/ZLARFG/ZHER2K for UPLO = 'L' (the
// reference calls ZHEEV, bse_solver-tool-diel.f90:485-488, whose first step
// this is), applying Q^T to a few vectors g on the way (Q = H(1)...H(N-1)), so
// that X^T h = Z^T (Q^T h) never needs the eigenvectors X = Q Z.
//
// Two kernels per reduced column j (panel column jj = j - p):
//   rowstep(j)  rows >= j, 16 rows x 16 lanes per workgroup.  Every workgroup
//               first sums the per-row-block partials hemvstep(j-1) left and
//               forms reflector j-1's ZLARFG scalars and W^H v, V^H v, v^T g
//               itself; it then finishes the reflector (v; y = A22 v from the
//               unscaled GEMV by linearity; u = y - V (W^H v) - W (V^H v);
//               g <- g - tau conj(v) (v^T g)) and updates column j with the panel
//               (ZLATRD), leaving out the 2 Re(alpha') v term of W(:,jj-1) that
//               needs the global u^H v, of which it writes its partial.
//   hemvstep(j) lower 64x64 tiles of A22 (rows, cols > j), one per workgroup.
//               Every workgroup sums the u^H v partials into alpha', adds the
//               deferred term to column j and forms z = A22 x~ with x~ = column j
//               below row j+1 (the reflector before its scale is known).  Each
//               tile is read once and serves both A x~ (rows) and A^H x~ (cols).
//               Diagonal tiles also write x_new, W(:,jj-1), d_j and partials of
//               |x~|^2, W^H x~, V^H x~, x~^T g.
// Per panel: finish (rowstep without a column), W(:,last), and the trailing
// update A22 -= V W^H + W V^H on MFMA (lower tiles).
// Global reductions go through partials that the next kernel sums (every
// workgroup of it, in a fixed order): no atomics, no device fences, no waits.
#include "wtb_kernels.h"
#include <algorithm>

namespace wtb {

#define TB 64   // tile / row-block edge
constexpr int kTriBatchMax = 64;   // matrices per batched launch
constexpr int kScalLen = 2 + 32 + 32 + 16;
#define RPW 16  // rows per rowstep workgroup

template <class T>
struct TriArgs {
  T* H;
  int ld, N, nb;
  zc* W;       // nb x N (column c at W + c*N)
  zc* Vp;      // nb x N
  zc* u;       // N
  zc* tau;     // N
  double* d;
  double* e;
  double* pn;      // [nblk]           |x~|^2
  zc* puv;         // [N/RPW+1]        u^H v per rowstep workgroup
  zc* pt1;         // [nblk][nb]       W^H x~
  zc* pt2;         // [nblk][nb]       V^H x~
  zc* pvg;         // [nblk][nrhs]     x~^T g
  zc* ypart;       // [nblk][nblk][TB] z = A22 x~ partials
  zc* g;           // nrhs x N
  zc* xn;          // N: reduced column j below its diagonal (x_new)
  zc* xp;          // N: column j after the panel update without the deferred term
  zc* gs;          // [16] row j+1 of g as hemvstep(j) saw it (rowstep(j+1) updates that row)
  zc* scal;        // [2 + 32 + 32 + 16] reflector j-1's scalars for the batched rowstep:
                   // tau, s, W^H v, V^H v, v^T g
  int nrhs;
  int nblk;
};

__device__ inline zc wave_sum(zc v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off);
    v.im += __shfl_xor(v.im, off);
  }
  return v;
}

// Sum of each of 16 values over the 64 lanes of a wave (pv: a zc[16] in registers):
// every butterfly step halves the set a lane keeps; afterwards lane l holds the sum for
// index (l >> 2) & 15.  A macro, so that the array is never addressed through memory.
#define BUTTERFLY16(pv, lane)                                             \
  do {                                                                    \
    int n_ = 16;                                                          \
    for (int off_ = 32; off_ >= 4; off_ >>= 1) {                          \
      n_ >>= 1;                                                           \
      const bool up_ = ((lane) & off_) != 0;                              \
      for (int i_ = 0; i_ < n_; ++i_) {                                   \
        const zc send_ = up_ ? pv[i_] : pv[i_ + n_];                      \
        zc recv_;                                                         \
        recv_.re = __shfl_xor(send_.re, off_);                            \
        recv_.im = __shfl_xor(send_.im, off_);                            \
        pv[i_] = (up_ ? pv[i_ + n_] : pv[i_]) + recv_;                    \
      }                                                                   \
    }                                                                     \
    for (int off_ = 2; off_ > 0; off_ >>= 1) {                            \
      pv[0].re += __shfl_xor(pv[0].re, off_);                             \
      pv[0].im += __shfl_xor(pv[0].im, off_);                             \
    }                                                                     \
  } while (0)

// the same for 8 values: lane holds the sum for value (lane >> 3) & 7
#define BUTTERFLY8(pv, lane)                                              \
  do {                                                                    \
    int n_ = 8;                                                           \
    for (int off_ = 32; off_ >= 8; off_ >>= 1) {                          \
      n_ >>= 1;                                                           \
      const bool up_ = ((lane) & off_) != 0;                              \
      for (int i_ = 0; i_ < n_; ++i_) {                                   \
        const zc send_ = up_ ? pv[i_] : pv[i_ + n_];                      \
        zc recv_;                                                         \
        recv_.re = __shfl_xor(send_.re, off_);                            \
        recv_.im = __shfl_xor(send_.im, off_);                            \
        pv[i_] = (up_ ? pv[i_ + n_] : pv[i_]) + recv_;                    \
      }                                                                   \
    }                                                                     \
    for (int off_ = 4; off_ > 0; off_ >>= 1) {                            \
      pv[0].re += __shfl_xor(pv[0].re, off_);                             \
      pv[0].im += __shfl_xor(pv[0].im, off_);                             \
    }                                                                     \
  } while (0)

// the same on COMPLEX*8 values (FP32 sums over the 64 rows of a tile)
__device__ inline zf operator+(zf a, zf b) {
  zf c;
  c.re = a.re + b.re;
  c.im = a.im + b.im;
  return c;
}
#define BUTTERFLY16F(pv, lane)                                            \
  do {                                                                    \
    int n_ = 16;                                                          \
    for (int off_ = 32; off_ >= 4; off_ >>= 1) {                          \
      n_ >>= 1;                                                           \
      const bool up_ = ((lane) & off_) != 0;                              \
      for (int i_ = 0; i_ < n_; ++i_) {                                   \
        const zf send_ = up_ ? pv[i_] : pv[i_ + n_];                      \
        zf recv_;                                                         \
        recv_.re = __shfl_xor(send_.re, off_);                            \
        recv_.im = __shfl_xor(send_.im, off_);                            \
        pv[i_] = (up_ ? pv[i_ + n_] : pv[i_]) + recv_;                    \
      }                                                                   \
    }                                                                     \
    for (int off_ = 2; off_ > 0; off_ >>= 1) {                            \
      pv[0].re += __shfl_xor(pv[0].re, off_);                             \
      pv[0].im += __shfl_xor(pv[0].im, off_);                             \
    }                                                                     \
  } while (0)

// alpha' = -1/2 |tau|^2 sum_b puv[b] over the np partials of the previous rowstep, in a
// fixed order (every thread of the workgroup returns it)
__device__ inline zc alpha_from_puv(const zc* __restrict__ puv, int np, zc tau) {
  __shared__ zc red[4];
  zc acc = mkz(0, 0);
  for (int b = threadIdx.x; b < np; b += blockDim.x) acc += puv[b];
  acc = wave_sum(acc);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  zc t = mkz(0, 0);
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  const double t2 = norm2(tau);
  return mkz(-0.5 * t2 * t.re, -0.5 * t2 * t.im);
}

// Reflector jm's scalars from the partials of hemvstep(jm) (panel column jjh): ZLARFG on
// alpha = x_new(jm+1) with |x|^2 = sum |x~|^2, then W^H v = conj(W(jm+1,:)) + s W^H x~,
// V^H v likewise and v^T g, into the workgroup's shared arrays.  256 threads.
// The caller preloads (before any barrier) alpha = x_new(jm+1) in thread 0, W(jm+1, c) and
// V(jm+1, c) in thread c < jjh, and g(q, jm+1) as hemvstep(jm) saw it in thread 64 + q.
template <class T>
__device__ void reflector_scalars(const TriArgs<T>& a, int jm, int jjh, zc* sums, zc* st1, zc* st2, zc* svg,
                                  zc* s_tau, zc* s_s, bool write_global, zc alpha_pre, zc w_pre, zc v_pre,
                                  zc g_pre) {
  const int tid = threadIdx.x;
  const int N = a.N;
  const int b0h = (jm + 1) / TB;
  const int nq = 1 + 2 * jjh + a.nrhs;
  const int sub = tid & 3;
  for (int q0 = 0; q0 < nq; q0 += 64) {
    const int q = q0 + (tid >> 2);
    zc acc = mkz(0, 0);
    if (q == 0) {
      acc.re = strided_sum(a.pn, 1, b0h + sub, a.nblk, 4, 0.0);
    } else if (q < nq) {
      const zc* base;
      size_t stride;
      if (q <= jjh) {
        base = a.pt1 + (q - 1);
        stride = a.nb;
      } else if (q <= 2 * jjh) {
        base = a.pt2 + (q - 1 - jjh);
        stride = a.nb;
      } else {
        base = a.pvg + (q - 1 - 2 * jjh);
        stride = a.nrhs;
      }
      acc = strided_sum(base, stride, b0h + sub, a.nblk, 4, acc);
    }
    for (int off = 2; off > 0; off >>= 1) {
      acc.re += __shfl_xor(acc.re, off, 4);
      acc.im += __shfl_xor(acc.im, off, 4);
    }
    if (sub == 0 && q < nq) sums[q] = acc;
  }
  __syncthreads();
  if (tid == 0) {
    const double xn2 = sums[0].re;
    const zc alpha = alpha_pre;
    zc tau, sc;
    double beta;
    if (xn2 == 0.0 && alpha.im == 0.0) {
      tau = mkz(0, 0);
      sc = mkz(0, 0);
      beta = alpha.re;
    } else {
      beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2), alpha.re);
      tau = mkz((beta - alpha.re) / beta, -alpha.im / beta);
      const zc den = mkz(alpha.re - beta, alpha.im);
      const double dd = den.re * den.re + den.im * den.im;
      sc = mkz(den.re / dd, -den.im / dd);
    }
    *s_tau = tau;
    *s_s = sc;
    if (write_global) {
      a.tau[jm] = tau;
      a.e[jm] = beta;
    }
  }
  __syncthreads();
  const zc sc = *s_s;
  if (tid < jjh) {
    const int c = tid;
    st1[c] = conj(w_pre) + sc * sums[1 + c];
    st2[c] = conj(v_pre) + sc * sums[1 + jjh + c];
  }
  if (tid >= 64 && tid < 64 + a.nrhs) {
    const int q = tid - 64;
    svg[q] = g_pre + sc * sums[1 + 2 * jjh + q];
  }
  __syncthreads();
}

// ---- rowstep ------------------------------------------------------------------------
// Loads that do not depend on reflector j-1's scalars (the GEMV partials, the panel
// rows, column j) are issued first; the panel values serve both u and the column update.
// PRE: reflector j-1's scalars come from k_tri_scal (the batched form) instead of being
// formed by every workgroup; both run reflector_scalars on the same inputs
template <class T, bool PRE>
__device__ __forceinline__ void tri_rowstep(const TriArgs<T>& a, int bx, int j, int jj, int update_col) {
  __shared__ zc st1[32], st2[32], svg[16], sWj[32], sVj[32], sums[96];
  __shared__ zc su_j, s_tau, s_s;
  __shared__ zc suv[RPW];
  const int tid = threadIdx.x;
  const int row_l = tid >> 4, sub = tid & 15;
  const int N = a.N;
  const int jm = j - 1, jjm = jj - 1;
  const int bm = (jm + 1) / TB;                   // first row block of reflector jm's GEMV
  const int r = j + bx * RPW + row_l;
  const bool valid = r < N;
  auto ysum = [&](int row) -> zc {
    const zc* yp = a.ypart + ((size_t)(row / TB) * a.nblk) * TB + (row % TB);
    return strided_sum(yp, (size_t)TB, bm + sub, a.nblk, 16, mkz(0, 0));
  };
  // panel rows j and r, columns sub and sub + 16
  zc yj = mkz(0, 0), yr = mkz(0, 0), hjj = mkz(0, 0), hcol = mkz(0, 0);
  zc vj[2], wj[2], vr[2], wr[2];
  for (int k = 0; k < 2; ++k) vj[k] = wj[k] = vr[k] = wr[k] = mkz(0, 0);
  if (row_l == 0) {
    yj = ysum(j);
    hjj = ldz(a.H[(size_t)j + (size_t)a.ld * j]);
    for (int k = 0; k < 2; ++k) {
      const int c = sub + 16 * k;
      if (c < jjm) {
        vj[k] = a.Vp[(size_t)c * N + j];
        wj[k] = a.W[(size_t)c * N + j];
      }
    }
  }
  if (valid) {
    yr = ysum(r);
    hcol = ldz(a.H[(size_t)r + (size_t)a.ld * j]);   // A22(r, j): column j before its update
    for (int k = 0; k < 2; ++k) {
      const int c = sub + 16 * k;
      if (c < jjm) {
        vr[k] = a.Vp[(size_t)c * N + r];
        wr[k] = a.W[(size_t)c * N + r];
      }
    }
  }
  const zc xnr = (valid && r != j) ? a.xn[r] : mkz(0, 0);
  const zc gr = (valid && sub < a.nrhs) ? a.g[(size_t)sub * N + r] : mkz(0, 0);
  if constexpr (PRE) {
    if (tid == 0) {
      s_tau = a.scal[0];
      s_s = a.scal[1];
    }
    if (tid < jjm) {
      st1[tid] = a.scal[2 + tid];
      st2[tid] = a.scal[34 + tid];
    }
    if (tid >= 64 && tid < 64 + a.nrhs) svg[tid - 64] = a.scal[66 + tid - 64];
    __syncthreads();
  } else {
    // reflector jm's inputs, none of which depends on its scalars
    zc alpha_pre = mkz(0, 0), w_pre = mkz(0, 0), v_pre = mkz(0, 0), g_pre = mkz(0, 0);
    if (tid == 0) alpha_pre = a.xn[jm + 1];
    if (tid < jjm) {
      w_pre = a.W[(size_t)tid * N + j];
      v_pre = a.Vp[(size_t)tid * N + j];
    }
    if (tid >= 64 && tid < 64 + a.nrhs) g_pre = a.gs[tid - 64];
    reflector_scalars(a, jm, jjm, sums, st1, st2, svg, &s_tau, &s_s, bx == 0, alpha_pre, w_pre, v_pre,
                      g_pre);
  }
  const zc tau = s_tau, s = s_s;
  // s z(row) - V t1 - W t2, each row summed by its 16 lanes
  auto u_of = [&](zc ys, const zc* v2, const zc* w2) -> zc {
    zc acc = s * ys;
    for (int k = 0; k < 2; ++k) {
      const int c = sub + 16 * k;
      if (c < jjm) {
        acc = acc - v2[k] * st1[c];
        acc = acc - w2[k] * st2[c];
      }
    }
    for (int off = 8; off > 0; off >>= 1) {
      acc.re += __shfl_xor(acc.re, off, 16);
      acc.im += __shfl_xor(acc.im, off, 16);
    }
    return acc;
  };
  // u(j): every workgroup needs it for the column update; the same lanes publish
  // row j of the panel (W(j, jj-1) = tau u(j), V(j, jj-1) = 1) for it
  if (row_l == 0) {
    const zc up = u_of(yj, vj, wj);
    zc h = hjj;
    h.im = 0.0;                                    // A22 diagonal: real part (ZHEMV)
    const zc uj = h + up;
    if (sub == 0) su_j = uj;
    if (update_col) {
      for (int k = 0; k < 2; ++k) {
        const int c = sub + 16 * k;
        if (c < jjm) {
          sWj[c] = wj[k];
          sVj[c] = vj[k];
        } else if (c == jjm) {
          sWj[c] = tau * uj;
          sVj[c] = mkz(1, 0);
        }
      }
    }
  }
  __syncthreads();
  zc uu = mkz(0, 0), v = mkz(0, 0);
  if (valid) {
    const zc up = u_of(yr, vr, wr);
    uu = (r == j) ? su_j : hcol + up;
    v = (r == j) ? mkz(1, 0) : s * xnr;
    if (sub == 0) {
      a.Vp[(size_t)jjm * N + r] = v;
      a.u[r] = uu;
    }
    if (sub < a.nrhs) {
      const int q = sub;
      const zc gq = gr - (tau * conj(v)) * svg[q];
      a.g[(size_t)q * N + r] = gq;
    }
  }
  if (sub == 0) suv[row_l] = valid ? conj(uu) * v : mkz(0, 0);
  // column j with the panel (ZLATRD), W(:,jjm) taken as tau*u (alpha' part deferred)
  if (update_col && valid) {
    zc acc = mkz(0, 0);
    for (int k = 0; k < 2; ++k) {
      const int c = sub + 16 * k;
      if (c < jj) {
        const zc vc = (c == jjm) ? v : vr[k];
        const zc wc = (c == jjm) ? tau * uu : wr[k];
        acc = acc + vc * conj(sWj[c]) + wc * conj(sVj[c]);
      }
    }
    for (int off = 8; off > 0; off >>= 1) {
      acc.re += __shfl_xor(acc.re, off, 16);
      acc.im += __shfl_xor(acc.im, off, 16);
    }
    if (sub == 0) a.xp[r] = hcol - acc;
  }
  __syncthreads();
  if (tid == 0) {
    zc uv = mkz(0, 0);
    for (int i = 0; i < RPW; ++i) uv += suv[i];
    a.puv[bx] = uv;
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_rowstep(TriArgs<T> a, int j, int jj, int update_col) {
  tri_rowstep<T, false>(a, blockIdx.x, j, jj, update_col);
}

// reflector j-1's scalars once per matrix (the batched form's rowstep reads them)
template <class T>
__device__ __forceinline__ void tri_scal(const TriArgs<T>& a, int j, int jj) {
  __shared__ zc st1[32], st2[32], svg[16], sums[96];
  __shared__ zc s_tau, s_s;
  const int tid = threadIdx.x;
  const int N = a.N;
  const int jm = j - 1, jjm = jj - 1;
  zc alpha_pre = mkz(0, 0), w_pre = mkz(0, 0), v_pre = mkz(0, 0), g_pre = mkz(0, 0);
  if (tid == 0) alpha_pre = a.xn[jm + 1];
  if (tid < jjm) {
    w_pre = a.W[(size_t)tid * N + j];
    v_pre = a.Vp[(size_t)tid * N + j];
  }
  if (tid >= 64 && tid < 64 + a.nrhs) g_pre = a.gs[tid - 64];
  reflector_scalars(a, jm, jjm, sums, st1, st2, svg, &s_tau, &s_s, true, alpha_pre, w_pre, v_pre, g_pre);
  if (tid == 0) {
    a.scal[0] = s_tau;
    a.scal[1] = s_s;
  }
  if (tid < jjm) {
    a.scal[2 + tid] = st1[tid];
    a.scal[34 + tid] = st2[tid];
  }
  if (tid >= 64 && tid < 64 + a.nrhs) a.scal[66 + tid - 64] = svg[tid - 64];
}


// ---- hemvstep: 64 x 64 tiles ------------------------------------------------------------
// The diagonal tile's dot products over its 64 rows (quantity q: 0 = |x~|^2, 1..jj = W^H x~,
// jj+1..2jj = V^H x~, then x~^T g), 4 waves x EPQ quantities per pass: 8 for the COMPLEX*8
// form, 16 for FP64.  Pass 0 runs in the diagonal tile's workgroup; for the COMPLEX*8
// form later passes run in workgroups of their own (tri_hemv_quant), each forming x~ and
// W(:, jj-1) of its rows the same way, so that the diagonal tile's chain of dependent
// loads is one pass long (FP64: all passes in the diagonal tile, whose registers an extra
// path would overflow).
template <class T>
__host__ __device__ constexpr int hemv_epq() { return sizeof(T) == sizeof(zc) ? 16 : 8; }
template <class T>
__host__ __device__ inline int hemv_split_passes(int jj, int nrhs) {
  if (sizeof(T) == sizeof(zc)) return 1;
  const int nq = 1 + 2 * jj + nrhs;
  return (nq + 4 * hemv_epq<T>() - 1) / (4 * hemv_epq<T>());
}

// one pass of the dot products of row block bi: wave w takes quantities q0 = EPQ w + 4 EPQ
// pass .. + EPQ - 1, lane = row r (x~(r) = xr, W(r, jj-1) = wl), from the pass's operands
// loaded beforehand (hemv_quant_load)
template <class T>
__device__ __forceinline__ const zc* hemv_operand(const TriArgs<T>& a, int q, int jj, int nq, int r) {
  // the loaded operand: W(r, q-1), V(r, q-1-jj) or g(q-1-2jj, r); W(r, jj-1) comes from LDS
  const int N = a.N;
  if (q >= 1 && q < jj) return a.W + (size_t)(q - 1) * N + r;
  if (q > jj && q <= 2 * jj) return a.Vp + (size_t)(q - 1 - jj) * N + r;
  if (q > 2 * jj && q < nq) return a.g + (size_t)(q - 1 - 2 * jj) * N + r;
  return nullptr;
}
template <int EPQ>
struct HemvPV {
  zc v[EPQ];
};
template <class T>
__device__ __forceinline__ HemvPV<hemv_epq<T>()> hemv_quant_load(const TriArgs<T>& a, int q0, int jj, int nq, int r,
                                                                 bool rok) {
  HemvPV<hemv_epq<T>()> o;
#pragma unroll
  for (int k = 0; k < hemv_epq<T>(); ++k) {
    const zc* src = hemv_operand(a, q0 + k, jj, nq, r);
    o.v[k] = (src && rok) ? *src : mkz(0, 0);
  }
  return o;
}
template <class T>
__device__ __forceinline__ void hemv_quant_pass(const TriArgs<T>& a, HemvPV<hemv_epq<T>()> in, int q0, int bi,
                                                int jj, int nq, int r, bool rok, zc xr, zc wl, int lane) {
  constexpr int EPQ = hemv_epq<T>();
  zc pv[EPQ];
#pragma unroll
  for (int k = 0; k < EPQ; ++k) pv[k] = in.v[k];
  const double xr2 = norm2(xr);
#pragma unroll
  for (int k = 0; k < EPQ; ++k) {
    const int q = q0 + k;
    const zc val = (q == jj) ? wl : pv[k];
    zc v = (q <= 2 * jj) ? fma_conj(val, xr, mkz(0, 0)) : fma_z(xr, val, mkz(0, 0));
    if (q == 0) v = mkz(xr2, 0.0);
    pv[k] = (q < nq && rok) ? v : mkz(0, 0);
  }
  int q;
  bool head;
  if constexpr (EPQ == 8) {
    BUTTERFLY8(pv, lane);
    q = q0 + ((lane >> 3) & 7);
    head = (lane & 7) == 0;
  } else {
    BUTTERFLY16(pv, lane);
    q = q0 + ((lane >> 2) & 15);
    head = (lane & 3) == 0;
  }
  if (head && q < nq) {
    if (q == 0) a.pn[bi] = pv[0].re;
    else if (q <= jj) a.pt1[(size_t)bi * a.nb + (q - 1)] = pv[0];
    else if (q <= 2 * jj) a.pt2[(size_t)bi * a.nb + (q - 1 - jj)] = pv[0];
    else a.pvg[(size_t)bi * a.nrhs + (q - 1 - 2 * jj)] = pv[0];
  }
}

// 4 waves, wave w owns 16 columns, lane = row: every load moves one column segment.
// FP64 matrix: one tile per workgroup.  COMPLEX*8 matrix: the diagonal tiles one per
// workgroup, then the off-diagonal ones two per workgroup, both tiles' loads in flight
// together (a COMPLEX*8 tile alone is too little data to cover a workgroup's fixed
// costs); each tile's sums are formed exactly as for one tile.
// Waves per SIMD that the registers allow without spilling: 4 (127 registers for FP64,
// two tiles' loads for COMPLEX*8).
template <class T>
__device__ __forceinline__ void tri_hemv_tile(const TriArgs<T>& a, int ti, int tj, int j, int jj, int nprev) {
  constexpr int XRN = TB;                   // rows of x~ a workgroup needs
  constexpr int XT = XRN + TB;              // threads loading x~ (rows, then columns)
  __shared__ zc xR[XRN], xC[TB];
  __shared__ zf xCf[TB];                    // x~ of the tile's columns in FP32 (COMPLEX*8 form)
  __shared__ zc y1s[4][XRN], y2s[1][TB];
  __shared__ zc sWl[TB];
  __shared__ zc s_ap;
  const int N = a.N;
  const int b0 = (j + 1) / TB;
  const int bi = b0 + ti, bj = b0 + tj;
  const int tid = threadIdx.x;
  const int R0 = bi * TB, C0 = bj * TB;
  const bool diag = bi == bj;
  const bool first = bj == b0 && diag;      // the workgroup holding row block b0's diagonal
  const int lane = tid & 63, w = tid >> 6;
  const int r = R0 + lane;                  // the diagonal epilogue's row
  // the tile and the raw pieces of x_new first: none of them depends on alpha'.  The loads
  // are unconditional (masked-off elements read a clamped in-range address) and the mask is
  // applied when the values are used, so that all 16 stay in flight together.
  // Interior tiles (off the diagonal, inside the trailing matrix: the common case) need no
  // masks at all.
  const bool interior = !diag && C0 > j && R0 + TB <= N && C0 + TB <= N;
  zc hv[16];
  T hraw[16];
  unsigned okm = 0;
  if (interior) {
    const T* col = a.H + (size_t)r + (size_t)a.ld * (C0 + 16 * w);
#pragma unroll
    for (int k = 0; k < 16; ++k) hraw[k] = col[(size_t)a.ld * k];
  } else {
    const bool rvalid = r > j && r < N;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int cl = 16 * w + k, c = C0 + cl;
      const bool ok = rvalid && c > j && c < N && (!diag || lane >= cl);
      const size_t off = ok ? (size_t)r + (size_t)a.ld * c : 0;
      hraw[k] = a.H[off];
      okm |= (unsigned)ok << k;
    }
  }
  // x_new = x_p - 2 Re(alpha') v_{j-1} (or column j of H at a panel's first column), rows
  // >= j+1 (x~ leaves out row j+1); diagonal tiles also load what their epilogue writes:
  // u (for W(:,jj-1)), row j of x_p and v (for d_j) and row j+1 of g
  const int xrow = (tid < XRN) ? R0 + tid : C0 + tid - XRN;
  const bool xok = tid < XT && xrow >= j + 1 && xrow < N;
  zc x0 = mkz(0, 0), x1 = mkz(0, 0), uo = mkz(0, 0);
  if (xok) {
    if (jj == 0) {
      x0 = ldz(a.H[(size_t)xrow + (size_t)a.ld * j]);
    } else {
      x0 = a.xp[xrow];
      x1 = a.Vp[(size_t)(jj - 1) * N + xrow];
      if (diag && tid < TB) uo = a.u[xrow];
    }
  }
  const bool dj = first && tid == 255;
  const bool gsl = first && tid >= XT && tid < XT + a.nrhs;
  if (dj) {
    if (jj == 0) {
      x0 = ldz(a.H[(size_t)j + (size_t)a.ld * j]);
    } else {
      x0 = a.xp[j];
      x1 = a.Vp[(size_t)(jj - 1) * N + j];
    }
  }
  if (gsl) x0 = a.g[(size_t)(tid - XT) * N + (j + 1)];
  const zc taup = (jj > 0) ? a.tau[j - 1] : mkz(0, 0);
  if (jj > 0) {
    const zc al = alpha_from_puv(a.puv, nprev, taup);
    if (tid == 0) s_ap = al;
  } else if (tid == 0) {
    s_ap = mkz(0, 0);
  }
  __syncthreads();
  const zc ap = s_ap;
  const double sig = 2.0 * ap.re;
  const zc xv = x0 - sig * x1;                       // x_new(xrow) where xok
  if (tid < XT) {
    const zc xt = (xok && xrow >= j + 2) ? xv : mkz(0, 0);
    if (tid < XRN) {
      xR[tid] = xt;
    } else {
      xC[tid - XRN] = xt;
      if constexpr (std::is_same<T, zf>::value) stz(xCf[tid - XRN], xt);
    }
  }
  // W(r, jj-1) = tau' u(r) + alpha' v(r): written by diagonal tiles, used in their partials
  const zc wo = (jj > 0 && xok) ? taup * uo + ap * x1 : mkz(0, 0);
  if (diag && tid < TB) sWl[tid] = wo;
  __syncthreads();
  const zc xr = xR[lane];
  zc y1 = mkz(0, 0);
  if (interior) okm = 0xffffu;
  if constexpr (std::is_same<T, zf>::value) {
    // COMPLEX*8 matrix: a wave's 16-column A x~ sums and the A^H x~ terms with their sums
    // over the tile's 64 rows in FP32; every sum across waves and tiles stays FP64
    zf hf[16];
    zf xrf, y1f;
    xrf.re = (float)xr.re;
    xrf.im = (float)xr.im;
    y1f.re = 0.0f;
    y1f.im = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int cl = 16 * w + k;
      const bool on = (okm >> k) & 1u;
      const zf h = hraw[k];
      zf x;
      x.re = on ? h.re : 0.0f;
      x.im = (on && !(diag && lane == cl)) ? h.im : 0.0f;   // diagonal: real part (ZHEMV)
      const zf xc = xCf[cl];
      y1f.re = fmaf(x.re, xc.re, fmaf(-x.im, xc.im, y1f.re));
      y1f.im = fmaf(x.re, xc.im, fmaf(x.im, xc.re, y1f.im));
      zf t;
      t.re = h.re * xrf.re + h.im * xrf.im;               // conj(h) x~(row)
      t.im = h.re * xrf.im - h.im * xrf.re;
      const bool low = on && (!diag || lane > cl);         // strictly lower
      hf[k].re = low ? t.re : 0.0f;
      hf[k].im = low ? t.im : 0.0f;
    }
    y1s[w][lane] = ldz(y1f);
    BUTTERFLY16F(hf, lane);
    if ((lane & 3) == 0) y2s[0][16 * w + ((lane >> 2) & 15)] = ldz(hf[0]);
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int cl = 16 * w + k;
      zc x = ((okm >> k) & 1u) ? ldz(hraw[k]) : mkz(0, 0);
      if (diag && lane == cl) x.im = 0.0;                 // diagonal: real part (ZHEMV)
      y1 = fma_z(x, xC[cl], y1);
      hv[k] = (!diag || lane > cl) ? conj(x) * xr : mkz(0, 0);   // A^H x~ terms (strictly lower)
    }
    y1s[w][lane] = y1;
    // transpose-reduce the 16 column sums over the 64 lanes: halve the set per butterfly step
    BUTTERFLY16(hv, lane);
    // lane holds the sum for column 16w + ((lane >> 2) & 15)
    if ((lane & 3) == 0) y2s[0][16 * w + ((lane >> 2) & 15)] = hv[0];
  }
  // the first pass of the dot products (its operands are loaded here, before the barrier,
  // so that the loads overlap the tile's epilogue); later passes run in workgroups of
  // their own
  const int nq = 1 + 2 * jj + a.nrhs;
  const bool rok = r >= j + 2 && r < N;
  auto operand = [&](int q) -> const zc* {
    // the loaded operand: W(r, q-1), V(r, q-1-jj) or g(q-1-2jj, r); W(r, jj-1) comes from LDS
    if (q >= 1 && q < jj) return a.W + (size_t)(q - 1) * N + r;
    if (q > jj && q <= 2 * jj) return a.Vp + (size_t)(q - 1 - jj) * N + r;
    if (q > 2 * jj && q < nq) return a.g + (size_t)(q - 1 - 2 * jj) * N + r;
    return nullptr;
  };
  constexpr int EPQ = hemv_epq<T>();
  zc pv[EPQ];
  if (diag) {
#pragma unroll
    for (int k = 0; k < EPQ; ++k) {
      const zc* src = operand(EPQ * w + k);
      pv[k] = (src && rok) ? *src : mkz(0, 0);
    }
  }
  __syncthreads();
  zc* yp = a.ypart;
  if (tid < TB) {
    zc sm = y1s[0][tid] + y1s[1][tid] + y1s[2][tid] + y1s[3][tid];
    if (diag) sm += y2s[0][tid];
    yp[((size_t)bi * a.nblk + bj) * TB + tid] = sm;
  } else if (tid < 2 * TB && !diag) {
    yp[((size_t)bj * a.nblk + bi) * TB + (tid - TB)] = y2s[0][tid - TB];
  }
  if (!diag) return;
  // ---- diagonal tiles: x_new, W(:,jj-1), d_j, g(j+1), the first pass of dot products ----
  if (xok && tid < TB) {
    a.xn[xrow] = xv;
    if (jj > 0) a.W[(size_t)(jj - 1) * N + xrow] = wo;
  }
  if (dj) a.d[j] = xv.re;                          // row j may lie in the block before b0
  // row j+1 (block b0) of g, for the next rowstep's v^T g, which also updates that row
  if (gsl) a.gs[tid - XT] = x0;
  const zc wl = rok ? sWl[lane] : mkz(0, 0);
  if constexpr (sizeof(T) != sizeof(zc)) {
    HemvPV<EPQ> in;
#pragma unroll
    for (int k = 0; k < EPQ; ++k) in.v[k] = pv[k];
    hemv_quant_pass(a, in, EPQ * w, bi, jj, nq, r, rok, xr, wl, lane);
  } else {
    // FP64: every pass here (the operands of the first were loaded above)
    const double xr2 = norm2(xr);
    for (int q0 = EPQ * w; q0 < nq; q0 += 4 * EPQ) {
      if (q0 > EPQ * w) {
#pragma unroll
        for (int k = 0; k < EPQ; ++k) {
          const zc* src = operand(q0 + k);
          pv[k] = (src && rok) ? *src : mkz(0, 0);
        }
      }
#pragma unroll
      for (int k = 0; k < EPQ; ++k) {
        const int q = q0 + k;
        const zc val = (q == jj) ? wl : pv[k];
        zc v = (q <= 2 * jj) ? fma_conj(val, xr, mkz(0, 0)) : fma_z(xr, val, mkz(0, 0));
        if (q == 0) v = mkz(xr2, 0.0);
        pv[k] = (q < nq && rok) ? v : mkz(0, 0);
      }
      BUTTERFLY16(pv, lane);
      const int q = q0 + ((lane >> 2) & 15);
      if ((lane & 3) == 0 && q < nq) {
        if (q == 0) a.pn[bi] = pv[0].re;
        else if (q <= jj) a.pt1[(size_t)bi * a.nb + (q - 1)] = pv[0];
        else if (q <= 2 * jj) a.pt2[(size_t)bi * a.nb + (q - 1 - jj)] = pv[0];
        else a.pvg[(size_t)bi * a.nrhs + (q - 1 - 2 * jj)] = pv[0];
      }
    }
  }
}

// later passes of row block b0 + ti's dot products: x~ and W(:, jj-1) of its rows formed
// as the diagonal tile forms them
template <class T>
__device__ __forceinline__ void tri_hemv_quant(const TriArgs<T>& a, int ti, int pass, int j, int jj, int nprev) {
  __shared__ zc qxR[TB], qW[TB];
  __shared__ zc q_ap;
  const int N = a.N;
  const int bi = (j + 1) / TB + ti;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int R0 = bi * TB;
  const int xrow = R0 + tid;
  const bool xok = tid < TB && xrow >= j + 1 && xrow < N;
  zc x0 = mkz(0, 0), x1 = mkz(0, 0), uo = mkz(0, 0);
  if (xok) {
    if (jj == 0) {
      x0 = ldz(a.H[(size_t)xrow + (size_t)a.ld * j]);
    } else {
      x0 = a.xp[xrow];
      x1 = a.Vp[(size_t)(jj - 1) * N + xrow];
      uo = a.u[xrow];
    }
  }
  const int nq = 1 + 2 * jj + a.nrhs;
  constexpr int EPQ = hemv_epq<T>();
  const int q0 = EPQ * w + 4 * EPQ * pass;
  const int r = R0 + lane;
  const bool rok = r >= j + 2 && r < N;
  const HemvPV<EPQ> pv0 = hemv_quant_load(a, q0, jj, nq, r, rok);
  const zc taup = (jj > 0) ? a.tau[j - 1] : mkz(0, 0);
  if (jj > 0) {
    const zc al = alpha_from_puv(a.puv, nprev, taup);
    if (tid == 0) q_ap = al;
  } else if (tid == 0) {
    q_ap = mkz(0, 0);
  }
  __syncthreads();
  const zc ap = q_ap;
  const double sig = 2.0 * ap.re;
  const zc xv = x0 - sig * x1;
  if (tid < TB) {
    qxR[tid] = (xok && xrow >= j + 2) ? xv : mkz(0, 0);
    qW[tid] = (jj > 0 && xok) ? taup * uo + ap * x1 : mkz(0, 0);
  }
  __syncthreads();
  const zc xr = qxR[lane];
  const zc wl = rok ? qW[lane] : mkz(0, 0);
  hemv_quant_pass(a, pv0, q0, bi, jj, nq, r, rok, xr, wl, lane);
}

// lower-triangle tile index t -> (ti, tj), tj <= ti
__device__ inline void tri_tile(int t, int& ti, int& tj) {
  ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  tj = t - ti * (ti + 1) / 2;
}

// two off-diagonal tiles o0 = 2 pb, o1 = o0 + 1 (strictly-lower index o -> ti >= 1,
// tj < ti) of the COMPLEX*8 matrix: the off-diagonal branch of tri_hemv_tile for both
template <class T>
__device__ __forceinline__ void tri_hemv_pair(const TriArgs<T>& a, int pb, int noff, int j, int jj, int nprev) {
  __shared__ zc xR[2][TB];
  __shared__ zf xCf[2][TB];
  __shared__ zc y1s[2][4][TB], y2s[2][TB];
  __shared__ zc s_ap;
  const int N = a.N;
  const int b0 = (j + 1) / TB;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const bool two = 2 * pb + 1 < noff;
  int bi[2], bj[2];
  // strictly-lower index o: o = ti (ti - 1) / 2 + tj with 0 <= tj < ti
  for (int u = 0; u < 2; ++u) {
    const int o = min(2 * pb + u, noff - 1);
    int ti = (int)((sqrt(8.0 * o + 1.0) + 1.0) * 0.5);
    while (ti * (ti - 1) / 2 > o) --ti;
    while ((ti + 1) * ti / 2 <= o) ++ti;
    bi[u] = b0 + ti;
    bj[u] = b0 + (o - ti * (ti - 1) / 2);
  }
  T hraw[2][16];
  unsigned okm[2] = {0xffffu, 0xffffu};
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int R0 = bi[u] * TB, C0 = bj[u] * TB;
    const int r = R0 + lane;
    const bool interior = C0 > j && R0 + TB <= N && C0 + TB <= N;
    if (interior) {
      const T* col = a.H + (size_t)r + (size_t)a.ld * (C0 + 16 * w);
#pragma unroll
      for (int k = 0; k < 16; ++k) hraw[u][k] = col[(size_t)a.ld * k];
    } else {
      const bool rvalid = r > j && r < N;
      okm[u] = 0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int c = C0 + 16 * w + k;
        const bool ok = rvalid && c > j && c < N;
        const size_t off = ok ? (size_t)r + (size_t)a.ld * c : 0;
        hraw[u][k] = a.H[off];
        okm[u] |= (unsigned)ok << k;
      }
    }
  }
  // x~ of both tiles' rows and columns: thread tid loads row/column (tid & 127) of tile
  // tid >> 7
  const int ux = tid >> 7, lx = tid & 127;
  const int xrow = (lx < TB) ? bi[ux] * TB + lx : bj[ux] * TB + lx - TB;
  const bool xok = xrow >= j + 1 && xrow < N;
  zc x0 = mkz(0, 0), x1 = mkz(0, 0);
  if (xok) {
    if (jj == 0) {
      x0 = ldz(a.H[(size_t)xrow + (size_t)a.ld * j]);
    } else {
      x0 = a.xp[xrow];
      x1 = a.Vp[(size_t)(jj - 1) * N + xrow];
    }
  }
  const zc taup = (jj > 0) ? a.tau[j - 1] : mkz(0, 0);
  if (jj > 0) {
    const zc al = alpha_from_puv(a.puv, nprev, taup);
    if (tid == 0) s_ap = al;
  } else if (tid == 0) {
    s_ap = mkz(0, 0);
  }
  __syncthreads();
  const zc ap = s_ap;
  const double sig = 2.0 * ap.re;
  const zc xv = x0 - sig * x1;
  {
    const zc xt = (xok && xrow >= j + 2) ? xv : mkz(0, 0);
    if (lx < TB) xR[ux][lx] = xt;
    else stz(xCf[ux][lx - TB], xt);
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const zc xr = xR[u][lane];
    zf hf[16];
    zf xrf, y1f;
    xrf.re = (float)xr.re;
    xrf.im = (float)xr.im;
    y1f.re = 0.0f;
    y1f.im = 0.0f;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int cl = 16 * w + k;
      const bool on = (okm[u] >> k) & 1u;
      const zf h = hraw[u][k];
      zf x;
      x.re = on ? h.re : 0.0f;
      x.im = on ? h.im : 0.0f;
      const zf xc = xCf[u][cl];
      y1f.re = fmaf(x.re, xc.re, fmaf(-x.im, xc.im, y1f.re));
      y1f.im = fmaf(x.re, xc.im, fmaf(x.im, xc.re, y1f.im));
      zf t;
      t.re = h.re * xrf.re + h.im * xrf.im;               // conj(h) x~(row)
      t.im = h.re * xrf.im - h.im * xrf.re;
      hf[k].re = on ? t.re : 0.0f;
      hf[k].im = on ? t.im : 0.0f;
    }
    y1s[u][w][lane] = ldz(y1f);
    BUTTERFLY16F(hf, lane);
    if ((lane & 3) == 0) y2s[u][16 * w + ((lane >> 2) & 15)] = ldz(hf[0]);
  }
  __syncthreads();
  // thread tid: tile tid >> 7; rows (A x~) for (tid & 127) < 64, columns (A^H x~) above
  const int u = tid >> 7, i = tid & 63;
  if (u == 1 && !two) return;
  zc* yp = a.ypart;
  if ((tid & 64) == 0) {
    const zc sm = y1s[u][0][i] + y1s[u][1][i] + y1s[u][2][i] + y1s[u][3][i];
    yp[((size_t)bi[u] * a.nblk + bj[u]) * TB + i] = sm;
  } else {
    yp[((size_t)bj[u] * a.nblk + bi[u]) * TB + i] = y2s[u][i];
  }
}

// workgroups of the GEMV at column j (nt = row blocks of the trailing matrix): the tiles
// (COMPLEX*8: diagonal ones alone, off-diagonal ones in pairs), then nt workgroups for each
// pass of dot products after the first
template <class T>
__host__ __device__ inline int hemv_tile_blocks(int nt) {
  if (sizeof(T) == sizeof(zc)) return nt * (nt + 1) / 2;
  const int noff = nt * (nt - 1) / 2;
  return nt + (noff + 1) / 2;
}
template <class T>
__host__ __device__ inline int hemv_blocks(int nt, int jj, int nrhs) {
  return hemv_tile_blocks<T>(nt) + nt * (hemv_split_passes<T>(jj, nrhs) - 1);
}

template <class T>
__device__ __forceinline__ void tri_hemv(const TriArgs<T>& a, int bx, int j, int jj, int nprev) {
  const int nt = a.nblk - (j + 1) / TB;
  if constexpr (std::is_same<T, zf>::value) {
    const int ntile = hemv_tile_blocks<T>(nt);
    if (bx >= ntile) {
      const int q = bx - ntile;
      tri_hemv_quant(a, q % nt, 1 + q / nt, j, jj, nprev);
    } else if (bx < nt) {
      tri_hemv_tile(a, bx, bx, j, jj, nprev);
    } else {
      tri_hemv_pair(a, bx - nt, nt * (nt - 1) / 2, j, jj, nprev);
    }
  } else {
    int ti, tj;
    tri_tile(bx, ti, tj);
    tri_hemv_tile(a, ti, tj, j, jj, nprev);
  }
}
template <class T>
__global__ __launch_bounds__(256, 4) void k_tri_hemv(TriArgs<T> a, int j, int jj, int nprev) {
  tri_hemv(a, blockIdx.x, j, jj, nprev);
}

// ---- W(:, kp-1) of the panel's last reflector ----------------------------------------
template <class T>
__device__ __forceinline__ void tri_wlast(const TriArgs<T>& a, int bx, int jl, int kpm1, int nprev) {
  const int lane = threadIdx.x;
  const zc tp = a.tau[jl];
  const zc ap = alpha_from_puv(a.puv, nprev, tp);
  const int r = ((jl + 1) / TB + bx) * TB + lane;
  if (r > jl && r < a.N) a.W[(size_t)kpm1 * a.N + r] = tp * a.u[r] + ap * a.Vp[(size_t)kpm1 * a.N + r];
}
template <class T>
__global__ __launch_bounds__(64) void k_tri_wlast(TriArgs<T> a, int jl, int kpm1, int nprev) {
  tri_wlast(a, blockIdx.x, jl, kpm1, nprev);
}

// ---- trailing update A22 -= V W^H + W V^H on lower 64x64 tiles (MFMA f64) ---------------
template <class T>
__device__ __forceinline__ void tri_her2k(const TriArgs<T>& a, int bx, int s0, int kp) {
  const int N = a.N;
  const int b0 = s0 / TB;
  int t = bx;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  int tj = t - ti * (ti + 1) / 2;
  const int bi = b0 + ti, bj = b0 + tj;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = bi * TB + (wave >> 1) * 32;
  const int col0 = bj * TB + (wave & 1) * 32;
  // the strictly upper quarter of a diagonal tile is never stored
  const bool upper = bi == bj && (wave >> 1) < (wave & 1);
  const int lr = lane & 15, lk = lane >> 4;
  const int K = 2 * kp;
  if constexpr (sizeof(T) == sizeof(zc)) {
    if (upper) return;
    d4 cr[2][2], ci[2][2];
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) {
        cr[x][y] = d4{0, 0, 0, 0};
        ci[x][y] = d4{0, 0, 0, 0};
      }
    for (int k0 = 0; k0 < K; k0 += 4) {
      int k = k0 + lk;
      zc av[2], bv[2];
      for (int x = 0; x < 2; ++x) {
        int r = row0 + x * 16 + lr;
        int c = col0 + x * 16 + lr;
        zc va = mkz(0, 0), vb = mkz(0, 0);
        if (k < K) {
          if (r < N) va = (k < kp) ? a.Vp[(size_t)k * N + r] : a.W[(size_t)(k - kp) * N + r];
          if (c < N) vb = (k < kp) ? a.W[(size_t)k * N + c] : a.Vp[(size_t)(k - kp) * N + c];
        }
        av[x] = va;
        bv[x] = vb;   // B(k, c) = conj(vb)
      }
      for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
          // a * conj(b)
          cr[x][y] = mfma16x16x4(av[x].re, bv[y].re, cr[x][y]);
          cr[x][y] = mfma16x16x4(av[x].im, bv[y].im, cr[x][y]);
          ci[x][y] = mfma16x16x4(av[x].im, bv[y].re, ci[x][y]);
          ci[x][y] = mfma16x16x4(-av[x].re, bv[y].im, ci[x][y]);
        }
    }
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y)
        for (int q = 0; q < 4; ++q) {
          int r = row0 + x * 16 + lk + 4 * q;     // f64 MFMA: register q holds row lk + 4q
          int c = col0 + y * 16 + lr;
          if (r < N && c < N && r >= s0 && c >= s0 && r >= c) {
            size_t off = (size_t)r + (size_t)a.ld * c;
            zc h = ldz(a.H[off]);
            h.re -= cr[x][y][q];
            h.im -= ci[x][y][q];
            stz(a.H[off], h);
          }
        }
  } else {
    // precision form: FP32 operands, FP32 accumulation (v_mfma_f32_16x16x4_f32).  The
    // tile's operands, [Vp | W] for its rows and [W | Vp] for its columns, go through LDS
    // in FP32 once per workgroup, 32 of K at a time (rows padded against bank conflicts);
    // the products run in the order of k.  Every wave takes part in the staging.
    __shared__ zf As[32][TB + 1], Bs[32][TB + 1];
    const int rl0 = bi * TB, cl0 = bj * TB;
    const int rw = (wave >> 1) * 32, cw = (wave & 1) * 32;
    f4 cr[2][2], ci[2][2];
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) {
        cr[x][y] = f4{0, 0, 0, 0};
        ci[x][y] = f4{0, 0, 0, 0};
      }
    for (int kb = 0; kb < K; kb += 32) {
      if (kb > 0) __syncthreads();
#pragma unroll 4
      for (int u = 0; u < 32 * TB / 256; ++u) {
        const int e = threadIdx.x + 256 * u, kk = e >> 6, cc = e & 63, k = kb + kk;
        const int r = rl0 + cc, c = cl0 + cc;
        zc va = mkz(0, 0), vb = mkz(0, 0);
        if (k < K) {
          if (r < N) va = (k < kp) ? a.Vp[(size_t)k * N + r] : a.W[(size_t)(k - kp) * N + r];
          if (c < N) vb = (k < kp) ? a.W[(size_t)k * N + c] : a.Vp[(size_t)(k - kp) * N + c];
        }
        As[kk][cc] = zf{(float)va.re, (float)va.im};
        Bs[kk][cc] = zf{(float)vb.re, (float)vb.im};
      }
      __syncthreads();
      if (!upper) {
        for (int k0 = 0; k0 < 32 && kb + k0 < K; k0 += 4) {
          const int kk = k0 + lk;
          float ar[2], ai[2], br[2], bi[2];
          for (int x = 0; x < 2; ++x) {
            const zf va = As[kk][rw + x * 16 + lr], vb = Bs[kk][cw + x * 16 + lr];
            ar[x] = va.re;
            ai[x] = va.im;
            br[x] = vb.re;
            bi[x] = vb.im;
          }
          for (int x = 0; x < 2; ++x)
            for (int y = 0; y < 2; ++y) {
              cr[x][y] = mfma16x16x4f32(ar[x], br[y], cr[x][y]);
              cr[x][y] = mfma16x16x4f32(ai[x], bi[y], cr[x][y]);
              ci[x][y] = mfma16x16x4f32(ai[x], br[y], ci[x][y]);
              ci[x][y] = mfma16x16x4f32(-ar[x], bi[y], ci[x][y]);
            }
        }
      }
    }
    if (upper) return;
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y)
        for (int q = 0; q < 4; ++q) {
          int r = row0 + x * 16 + 4 * lk + q;     // f32 MFMA: register q holds row 4lk + q
          int c = col0 + y * 16 + lr;
          if (r < N && c < N && r >= s0 && c >= s0 && r >= c) {
            size_t off = (size_t)r + (size_t)a.ld * c;
            T h = a.H[off];
            h.re -= cr[x][y][q];
            h.im -= ci[x][y][q];
            a.H[off] = h;
          }
        }
  }
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_her2k(TriArgs<T> a, int s0, int kp) {
  tri_her2k(a, blockIdx.x, s0, kp);
}

// reflectors of the finished panel into the lower triangle: H(r, j) = v_j(r), r > j+1
template <class T>
__device__ __forceinline__ void tri_storev(const TriArgs<T>& a, int bx, int c, int p) {
  const int j = p + c;
  const int r = j + 2 + bx * 256 + threadIdx.x;
  if (r < a.N) stz(a.H[(size_t)r + (size_t)a.ld * j], a.Vp[(size_t)c * a.N + r]);
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_storev(TriArgs<T> a, int p) {
  tri_storev(a, blockIdx.x, blockIdx.y, p);
}

template <class T>
__device__ __forceinline__ void tri_last(const TriArgs<T>& a) {
  zc x = ldz(a.H[(size_t)(a.N - 1) + (size_t)a.ld * (a.N - 1)]);
  a.d[a.N - 1] = x.re;
  a.e[a.N - 1] = 0.0;
}
template <class T>
__global__ void k_tri_last(TriArgs<T> a) {
  tri_last(a);
}

// ---- batched launches --------------------------------------------------------------
// One launch serves up to kTriBatchMax matrices at the same column step: workgroup b
// belongs to the matrix whose range [pre[i], pre[i+1]) holds it, and runs the single
// matrix's kernel body on that matrix's arguments with block index b - pre[i], so every
// matrix goes through exactly the arithmetic of its own reduction.
template <class T>
struct TriBatch {
  const TriArgs<T>* args;     // device array, one per matrix of the group
  int n;
  int m[kTriBatchMax];        // matrix of slot i
  int pre[kTriBatchMax + 1];  // first workgroup of slot i
};

__device__ inline int tri_slot(const int* pre, int n, int b) {
  int i = 0;
  while (i + 1 < n && pre[i + 1] <= b) ++i;
  return i;
}

template <class T>
__global__ __launch_bounds__(256) void k_tri_scal_b(TriBatch<T> B, int j, int jj) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  tri_scal(a, j, jj);
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_rowstep_b(TriBatch<T> B, int j, int jj, int update_col) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  tri_rowstep<T, true>(a, (int)blockIdx.x - B.pre[i], j, jj, update_col);
}
template <class T>
__global__ __launch_bounds__(256, 4) void k_tri_hemv_b(TriBatch<T> B, int j, int jj) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  const int nprev = jj > 0 ? (a.N - j + RPW - 1) / RPW : 0;
  tri_hemv(a, (int)blockIdx.x - B.pre[i], j, jj, nprev);
}
template <class T>
__global__ __launch_bounds__(64) void k_tri_wlast_b(TriBatch<T> B, int jl, int kpm1) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  const int nF = (a.N - (jl + 1) + RPW - 1) / RPW;
  tri_wlast(a, (int)blockIdx.x - B.pre[i], jl, kpm1, nF);
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_her2k_b(TriBatch<T> B, int s0, int kp) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  tri_her2k(a, (int)blockIdx.x - B.pre[i], s0, kp);
}
template <class T>
__global__ __launch_bounds__(256) void k_tri_storev_b(TriBatch<T> B, int p) {
  const int i = tri_slot(B.pre, B.n, blockIdx.x);
  const TriArgs<T> a = B.args[B.m[i]];
  const int gx = (a.N - p - 2 + 255) / 256;
  const int b = (int)blockIdx.x - B.pre[i];
  tri_storev(a, b % gx, b / gx, p);
}
template <class T>
__global__ void k_tri_last_b(TriBatch<T> B) {
  const TriArgs<T> a = B.args[B.m[blockIdx.x]];
  tri_last(a);
}


// ---------------------------------------------------------------------------
size_t tridiag_work_bytes(int N, int nb, int nrhs) {
  size_t nblk = (N + TB - 1) / TB;
  size_t b = 0;
  b += sizeof(zc) * (size_t)N * nb * 2 + 512;   // W, Vp
  b += sizeof(zc) * (size_t)N * 2 + 512;        // u, tau
  b += sizeof(double) * (size_t)N * 2 + 512;    // d, e
  b += sizeof(double) * nblk + 256;             // pn
  b += sizeof(zc) * (N / RPW + 2) + 256;        // puv
  b += sizeof(zc) * nblk * (2 * nb + nrhs + 1) + 1024;
  b += sizeof(zc) * nblk * nblk * TB + 256;     // ypart
  b += sizeof(zc) * N * 2 + 512;                // xn, xp
  b += sizeof(zc) * 16 + 256;                   // gs
  b += sizeof(zc) * kScalLen + 256;             // scal
  return b + 4096;
}

int tridiag_bind(TridiagWork& tw, void* base, int N, int nb, int nrhs) {
  tw.N = N;
  tw.nb = nb;
  tw.nrhs = nrhs;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~size_t(255);
    return q;
  };
  size_t nblk = (N + TB - 1) / TB;
  tw.W = (zc*)take(sizeof(zc) * (size_t)N * nb);
  tw.Vp = (zc*)take(sizeof(zc) * (size_t)N * nb);
  tw.u = (zc*)take(sizeof(zc) * N);
  tw.tau = (zc*)take(sizeof(zc) * N);
  tw.d = (double*)take(sizeof(double) * N);
  tw.e = (double*)take(sizeof(double) * N);
  tw.red = (double*)take(sizeof(double) * nblk);
  tw.puv = (zc*)take(sizeof(zc) * (N / RPW + 2));
  tw.zred = (zc*)take(sizeof(zc) * nblk * (2 * nb + nrhs + 1));
  tw.ypart = (zc*)take(sizeof(zc) * nblk * nblk * TB);
  tw.xn = (zc*)take(sizeof(zc) * N);
  tw.xp = (zc*)take(sizeof(zc) * N);
  tw.gs = (zc*)take(sizeof(zc) * 16);
  tw.scal = (zc*)take(sizeof(zc) * kScalLen);
  return WTB_OK;
}

// the panel width a reduction runs with: the tuned one when it fits the binding
static int tri_panel(const TridiagWork& tw) {
  return (tw.panel >= 8 && tw.panel <= tw.nb) ? tw.panel : tw.nb;
}

template <class T>
static TriArgs<T> tri_args(const TridiagWork& tw, T* d_H, int ldh, zc* d_g, int nb) {
  const int N = tw.N;
  TriArgs<T> a;
  a.H = d_H;
  a.ld = ldh;
  a.N = N;
  a.nb = nb;
  a.W = tw.W;
  a.Vp = tw.Vp;
  a.u = tw.u;
  a.tau = tw.tau;
  a.d = tw.d;
  a.e = tw.e;
  a.nblk = (N + TB - 1) / TB;
  a.pn = tw.red;
  a.puv = tw.puv;
  a.pt1 = tw.zred;
  a.pt2 = a.pt1 + (size_t)a.nblk * nb;
  a.pvg = a.pt2 + (size_t)a.nblk * nb;
  a.ypart = tw.ypart;
  a.g = d_g;
  a.xn = tw.xn;
  a.xp = tw.xp;
  a.gs = tw.gs;
  a.scal = tw.scal;
  a.nrhs = tw.nrhs;
  return a;
}

template <class T>
static int run_tridiag_t(hipStream_t s, TridiagWork& tw, T* d_H, int ldh, zc* d_g) {
  const int N = tw.N;
  const int nb = tri_panel(tw);
  if (nb > 32 || tw.nrhs > 16) return WTB_ERR_LIMIT;
  const TriArgs<T> a = tri_args(tw, d_H, ldh, d_g, nb);
  tw.kev_used = 0;
  if (N == 1) {
    hipLaunchKernelGGL(k_tri_last<T>, dim3(1), dim3(1), 0, s, a);
    WTB_LAUNCH_CHECK();
    return WTB_OK;
  }
  for (int p = 0; p < N - 1; p += nb) {
    const int kp = (nb < N - 1 - p) ? nb : (N - 1 - p);
    int nprev = 0;                                  // workgroups of the last rowstep
    for (int jj = 0; jj < kp; ++jj) {
      const int j = p + jj;
      if (jj > 0) {
        nprev = (N - j + RPW - 1) / RPW;
        hipLaunchKernelGGL(k_tri_rowstep<T>, dim3(nprev), dim3(256), 0, s, a, j, jj, 1);
      }
      const int nt = a.nblk - (j + 1) / TB;
      const bool tev = tw.kev && tw.kev_used < tw.kev_cap;
      if (tev) hipEventRecord(tw.kev[2 * tw.kev_used], s);
      hipLaunchKernelGGL(k_tri_hemv<T>, dim3(hemv_blocks<T>(nt, jj, tw.nrhs)), dim3(256), 0, s, a, j, jj, nprev);
      if (tev) hipEventRecord(tw.kev[2 * tw.kev_used++ + 1], s);
    }
    // finish the panel's last reflector (no further column in this panel)
    const int jl = p + kp - 1;
    const int jn = jl + 1;
    const int nF = (N - jn + RPW - 1) / RPW;
    hipLaunchKernelGGL(k_tri_rowstep<T>, dim3(nF), dim3(256), 0, s, a, jn, kp, 0);
    hipLaunchKernelGGL(k_tri_wlast<T>, dim3(a.nblk - (jl + 1) / TB), dim3(64), 0, s, a, jl, kp - 1, nF);
    if (tw.store_v && N - p - 2 > 0)
      hipLaunchKernelGGL(k_tri_storev<T>, dim3((N - p - 2 + 255) / 256, kp), dim3(256), 0, s, a, p);
    const int s0 = p + kp;
    if (s0 < N) {
      const int nt2 = a.nblk - s0 / TB;
      hipLaunchKernelGGL(k_tri_her2k<T>, dim3(nt2 * (nt2 + 1) / 2), dim3(256), 0, s, a, s0, kp);
    }
  }
  hipLaunchKernelGGL(k_tri_last<T>, dim3(1), dim3(1), 0, s, a);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int run_tridiag(hipStream_t s, TridiagWork& tw, zc* d_H, int ldh, zc* d_g) {
  return run_tridiag_t<zc>(s, tw, d_H, ldh, d_g);
}
int run_tridiag_f32(hipStream_t s, TridiagWork& tw, zf* d_H, int ldh, zc* d_g) {
  return run_tridiag_t<zf>(s, tw, d_H, ldh, d_g);
}

// ---- a group of reductions in lockstep ------------------------------------------------
// One launch per kernel and column step for every matrix still reducing that column (in
// chunks of kTriBatchMax).  Each matrix sees the same launches, in the same order and
// with the same arguments, as run_tridiag_t gives it alone; only the matrices in between
// differ.
template <class T, class F, class... A>
static void launch_group(hipStream_t s, F kern, const TriArgs<T>* dargs, const std::vector<int>& ms,
                         const std::vector<int>& nbk, int threads, A... args) {
  for (size_t c0 = 0; c0 < ms.size(); c0 += kTriBatchMax) {
    TriBatch<T> B;
    B.args = dargs;
    B.n = 0;
    int tot = 0;
    for (size_t i = c0; i < ms.size() && i < c0 + kTriBatchMax; ++i) {
      if (nbk[i] <= 0) continue;
      B.m[B.n] = ms[i];
      B.pre[B.n] = tot;
      tot += nbk[i];
      ++B.n;
    }
    if (B.n == 0) continue;
    B.pre[B.n] = tot;
    hipLaunchKernelGGL(kern, dim3(tot), dim3(threads), 0, s, B, args...);
  }
}

template <class T>
static int run_tridiag_group_t(hipStream_t s, const TriGroup& G) {
  const TriArgs<T>* dargs = static_cast<const TriArgs<T>*>(G.d_args);
  const int M = (int)G.items.size(), nb = G.nb;
  std::vector<int> N(M), nblk(M);
  int maxN = 0;
  for (int i = 0; i < M; ++i) {
    N[i] = G.items[i].tw->N;
    nblk[i] = (N[i] + TB - 1) / TB;
    maxN = std::max(maxN, N[i]);
  }
  std::vector<int> ms, fs, sel, nbk;
  auto counts = [&](const std::vector<int>& list, auto f) {
    nbk.resize(list.size());
    for (size_t k = 0; k < list.size(); ++k) nbk[k] = f(list[k]);
  };
  for (int p = 0; p < maxN - 1; p += nb) {
    for (int jj = 0; jj < nb; ++jj) {
      const int j = p + jj;
      ms.clear();
      fs.clear();
      for (int i = 0; i < M; ++i) {
        if (N[i] - 1 <= p) continue;
        const int kp = std::min(nb, N[i] - 1 - p);
        if (jj < kp) ms.push_back(i);
        if (jj + 1 == kp) fs.push_back(i);
      }
      if (ms.empty()) break;
      if (jj > 0) {
        counts(ms, [&](int) { return 1; });
        launch_group<T>(s, k_tri_scal_b<T>, dargs, ms, nbk, 256, j, jj);
        counts(ms, [&](int i) { return (N[i] - j + RPW - 1) / RPW; });
        launch_group<T>(s, k_tri_rowstep_b<T>, dargs, ms, nbk, 256, j, jj, 1);
      }
      counts(ms, [&](int i) {
        const int nt = nblk[i] - (j + 1) / TB;
        return hemv_blocks<T>(nt, jj, G.items[i].tw->nrhs);
      });
      launch_group<T>(s, k_tri_hemv_b<T>, dargs, ms, nbk, 256, j, jj);
      if (fs.empty()) continue;
      // the panel ends at column j for these matrices
      const int kp = jj + 1, jl = j, jn = j + 1, s0 = p + kp;
      counts(fs, [&](int) { return 1; });
      launch_group<T>(s, k_tri_scal_b<T>, dargs, fs, nbk, 256, jn, kp);
      counts(fs, [&](int i) { return (N[i] - jn + RPW - 1) / RPW; });
      launch_group<T>(s, k_tri_rowstep_b<T>, dargs, fs, nbk, 256, jn, kp, 0);
      counts(fs, [&](int i) { return nblk[i] - (jl + 1) / TB; });
      launch_group<T>(s, k_tri_wlast_b<T>, dargs, fs, nbk, 64, jl, kp - 1);
      sel.clear();
      for (int i : fs)
        if (G.items[i].tw->store_v && N[i] - p - 2 > 0) sel.push_back(i);
      counts(sel, [&](int i) { return (N[i] - p - 2 + 255) / 256 * kp; });
      launch_group<T>(s, k_tri_storev_b<T>, dargs, sel, nbk, 256, p);
      sel.clear();
      for (int i : fs)
        if (s0 < N[i]) sel.push_back(i);
      counts(sel, [&](int i) {
        const int nt2 = nblk[i] - s0 / TB;
        return nt2 * (nt2 + 1) / 2;
      });
      launch_group<T>(s, k_tri_her2k_b<T>, dargs, sel, nbk, 256, s0, kp);
    }
  }
  ms.resize(M);
  for (int i = 0; i < M; ++i) ms[i] = i;
  counts(ms, [&](int) { return 1; });
  launch_group<T>(s, k_tri_last_b<T>, dargs, ms, nbk, 1);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int tri_group_prepare(TriGroup& G) {
  tri_group_free(G);
  const int M = (int)G.items.size();
  if (M == 0) return WTB_OK;
  for (const auto& it : G.items)
    if (tri_panel(*it.tw) != G.nb || G.nb > 32 || it.tw->nrhs > 16) return WTB_ERR_LIMIT;
  if (G.f32) {
    std::vector<TriArgs<zf>> h(M);
    for (int i = 0; i < M; ++i) h[i] = tri_args(*G.items[i].tw, (zf*)G.items[i].H, G.items[i].ldh, G.items[i].g, G.nb);
    WTB_HIP(hipMalloc(&G.d_args, sizeof(TriArgs<zf>) * M));
    WTB_HIP(hipMemcpy(G.d_args, h.data(), sizeof(TriArgs<zf>) * M, hipMemcpyHostToDevice));
  } else {
    std::vector<TriArgs<zc>> h(M);
    for (int i = 0; i < M; ++i) h[i] = tri_args(*G.items[i].tw, (zc*)G.items[i].H, G.items[i].ldh, G.items[i].g, G.nb);
    WTB_HIP(hipMalloc(&G.d_args, sizeof(TriArgs<zc>) * M));
    WTB_HIP(hipMemcpy(G.d_args, h.data(), sizeof(TriArgs<zc>) * M, hipMemcpyHostToDevice));
  }
  return WTB_OK;
}

void tri_group_free(TriGroup& G) {
  if (G.d_args) (void)hipFree(G.d_args);
  G.d_args = nullptr;
}

int run_tridiag_group(hipStream_t s, const TriGroup& G) {
  if (G.items.empty()) return WTB_OK;
  if (!G.d_args) return WTB_ERR_ARG;
  return G.f32 ? run_tridiag_group_t<zf>(s, G) : run_tridiag_group_t<zc>(s, G);
}

}  // namespace wtb
