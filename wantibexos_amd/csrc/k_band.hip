// Stage (d1) in two stages: a Hermitian matrix (lower triangle, column-major, ld) is
// reduced first to a lower band of width b (dense -> band, blocked Householder on panels
// of b columns with MFMA two-sided updates), then to a real symmetric tridiagonal (band ->
// tridiagonal by bulge chasing), applying the conjugate transpose of every orthogonal
// factor to a few tracked vectors, so that X^T h = Z^T Q2^T Q1^T h needs no eigenvectors.
// The arithmetic is that of LAPACK's ZHETRD_HE2HB (ZGEQRF panels with ZLARFG reflectors,
// ZLARFT compact WY factor, the ZHER2K-form two-sided update) and ZHETRD_HB2ST (the
// bulge-chasing sweeps); the organisation into kernels is this file's own.
//
// Stage 1, panel p (columns c0 = p b .. c0 + b - 1, trailing rows r0 = c0 + b .. N - 1):
//   k_pqr_col(c), c = 0..k     one launch per column: finish reflector c - 1 from the
//                              previous launch's partial sums (every workgroup, in a fixed
//                              order), apply it to the panel columns right of it, and form
//                              the partial sums of column c (|x|^2 and x^H P(:, j))
//   k_pfin                     explicit V (unit diagonal) and R back into the matrix
//   k_gram, k_gram_sum         G = V^H V, t1 = V^H g (FP64 MFMA on row chunks, partials summed in order)
//   k_pT                       T (ZLARFT forward/columnwise) and t2 = T^H t1
//   k_pY                       Y = V T, g -= V t2 (rows of the panel)
//   k_hmm_b, k_hmm_sum         X = A22 Y (MFMA, full Hermitian storage, K slices)
//   k_gram, k_gram_sum         M = V^H X
//   k_pW                       W = X - 1/2 V (T^H M)
//   k_her2k_full               A22 -= V W^H + W V^H (MFMA, the tiles on and below the diagonal)
// The tracked vectors are carried conjugated (gc = conj(g)): Q^H gc = conj(Q^T g).
//
// Stage 2 (bulge chasing, FP64, band storage BD[c][d] = A(c + d, c), d <= 2b): sweep s
// annihilates column s below its subdiagonal; its tasks k = 0, 1, ... each apply one
// reflector (ZLARFG) two-sidedly to a diagonal block and the previous one from the right
// to the block below it (creating the bulge whose first column the next reflector
// annihilates).  One persistent kernel whose workgroups claim sweeps in order from a
// counter; task k of sweep s waits until sweep s - 1 has finished its tasks 0..k+1 (its
// later tasks touch only rows below this task's), signalled through a per-sweep progress
// counter (device-scope accesses, no cache maintenance).  Every wait is bounded: past the limit the
// kernel sets an error flag that every waiting workgroup sees, and all of them return.
#include "wtb_kernels.h"

namespace wtb {

#define PR 256          // panel rows per workgroup (one per thread)
#define BMAX 32         // largest band width

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline zc zdiv(zc a, zc b) {
  // a / b with the scaling of Smith's algorithm, both cases formed and one selected
  const bool re_big = fabs(b.re) >= fabs(b.im);
  const double p = re_big ? b.re : b.im, q = re_big ? b.im : b.re;
  const double r = q / p, den = p + q * r;
  const double xr = re_big ? a.re + a.im * r : a.re * r + a.im;
  const double xi = re_big ? a.im - a.re * r : a.im * r - a.re;
  return mkz(xr / den, xi / den);
}

__device__ inline zc shfl_z(zc v, int src) { return mkz(__shfl(v.re, src), __shfl(v.im, src)); }
__device__ inline zc shfl_xor_z(zc v, int off) { return mkz(__shfl_xor(v.re, off), __shfl_xor(v.im, off)); }

// x of another lane by a DPP lane permutation (within rows of 16 lanes) or a swizzle
// (within 32 lanes); both halves of the double move
template <int CTRL>
__device__ inline double dpp_d(double x) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double swz16_d(double x) {   // lane ^ 16 within 32 lanes
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_ds_swizzle((int)v, 0x401F);
  const int hi = __builtin_amdgcn_ds_swizzle((int)(v >> 32), 0x401F);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline zc swz16_z(zc v) { return mkz(swz16_d(v.re), swz16_d(v.im)); }
template <int CTRL>
__device__ inline zc dpp_z(zc v) { return mkz(dpp_d<CTRL>(v.re), dpp_d<CTRL>(v.im)); }

// The value of the partner lane that differs from this one in lane bit OFF and agrees in
// the bits above it: ^1 and ^2 by quad permutations, i <-> 7 - i (row half mirror) for 4,
// a rotation by 8 in a row of 16 for 8, a swizzle for 16, a permute for 32.  Partners are
// mutual, so sums formed with them are the same in both lanes.
template <int OFF>
__device__ inline double xchg(double x) {
  if constexpr (OFF == 1) return dpp_d<0xB1>(x);
  else if constexpr (OFF == 2) return dpp_d<0x4E>(x);
  else if constexpr (OFF == 4) return dpp_d<0x141>(x);
  else if constexpr (OFF == 8) return dpp_d<0x128>(x);
  else if constexpr (OFF == 16) return swz16_d(x);
  else return __shfl_xor(x, 32);
}
template <int OFF>
__device__ inline zc xchg(zc v) { return mkz(xchg<OFF>(v.re), xchg<OFF>(v.im)); }

// Data handed between workgroups of one kernel (the panel's partial sums, the chase's band
// and tracked vectors) moves through device-scope (sc1) loads and stores, which the
// hardware keeps coherent across the XCDs' L2 caches; the writer's stores are complete
// (s_waitcnt) before its counter moves, so the hand-off needs no L2 write-back or
// invalidate.
__device__ inline zc ld_dev(const zc* p) {
  const double* q = reinterpret_cast<const double*>(p);
  return mkz(__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
             __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline void st_dev(zc* p, zc v) {
  double* q = reinterpret_cast<double*>(p);
  __hip_atomic_store(q, v.re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q + 1, v.im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// sum of p[b * stride] over b = b0, b0 + step, ... < b1 in that order, device-scope loads
// issued eight at a time
__device__ inline zc strided_sum_dev(const zc* p, size_t stride, int b0, int b1, int step) {
  zc acc = mkz(0, 0);
  for (int b = b0; b < b1; b += 8 * step) {
    zc v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int bu = b + u * step;
      v[u] = bu < b1 ? ld_dev(p + (size_t)bu * stride) : mkz(0, 0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (b + u * step < b1) acc += v[u];
  }
  return acc;
}

// ZLARFG on (alpha, x) with |x|^2 = xn2: beta (real, returned as the new alpha), tau and
// the scale 1/(alpha - beta) of x.  tau = 0 (H = I) when x = 0 and alpha is real.
__device__ inline void larfg(zc alpha, double xn2, zc& beta, zc& tau, zc& scale) {
  const bool ident = xn2 == 0.0 && alpha.im == 0.0;
  const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
  const double b = ident ? 1.0 : (alpha.re >= 0.0 ? -nrm : nrm);      // 1: any nonzero stand-in
  const zc t = mkz((b - alpha.re) / b, -alpha.im / b);
  const zc sc = zdiv(mkz(1, 0), mkz(alpha.re - b, alpha.im));
  beta = ident ? alpha : mkz(b, 0.0);
  tau = ident ? mkz(0, 0) : t;
  scale = ident ? mkz(0, 0) : sc;
}

// Sums over the 256 threads of a workgroup of one real value x0 and NB complex values
// v[0..NB) (NB a power of two <= 32).  Within a wave the NB sums are formed by a butterfly
// that halves the set each lane keeps (NB - 1 exchanges instead of 6 NB), after which lane
// l holds the wave's sum of value (l >> (6 - log2 NB)) & (NB - 1); the four waves are then
// added in index order.  Results: red[4 * (NB + 1) + q] (q = 0: x0, 1 + j: v[j]).
// one halving step of the butterfly: lanes with the OFF bit set keep values [n, 2n)
template <int OFF, int n, int NB>
__device__ inline void bf_step(zc (&v)[NB], int lane) {
  const bool up = (lane & OFF) != 0;
#pragma unroll
  for (int q = 0; q < n; ++q) {
    const zc recv = xchg<OFF>(up ? v[q] : v[q + n]);
    v[q] = (up ? v[q + n] : v[q]) + recv;
  }
}

template <int NB>
__device__ inline void block_sum_bf(double x0, zc (&v)[NB], zc* red /* [5][NB + 1] in LDS */) {
  static_assert(NB == 16 || NB == 32, "butterfly for 16 or 32 values");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  zc r;
  if constexpr (NB == 32) {
    bf_step<32, 16>(v, lane);
    bf_step<16, 8>(v, lane);
    bf_step<8, 4>(v, lane);
    bf_step<4, 2>(v, lane);
    bf_step<2, 1>(v, lane);
    r = v[0] + xchg<1>(v[0]);
  } else {
    bf_step<32, 8>(v, lane);
    bf_step<16, 4>(v, lane);
    bf_step<8, 2>(v, lane);
    bf_step<4, 1>(v, lane);
    r = v[0] + xchg<2>(v[0]);
    r += xchg<1>(r);
  }
  x0 += xchg<1>(x0);
  x0 += xchg<2>(x0);
  x0 += xchg<4>(x0);
  x0 += xchg<8>(x0);
  x0 += xchg<16>(x0);
  x0 += xchg<32>(x0);
  constexpr int SH = (NB == 32) ? 1 : 2;
  if ((lane & ((1 << SH) - 1)) == 0) red[wave * (NB + 1) + 1 + ((lane >> SH) & (NB - 1))] = r;
  if (lane == 0) red[wave * (NB + 1)] = mkz(x0, 0.0);
  __syncthreads();
  if ((int)threadIdx.x < NB + 1) {
    const int q = threadIdx.x;
    red[4 * (NB + 1) + q] = ((red[q] + red[(NB + 1) + q]) + red[2 * (NB + 1) + q]) + red[3 * (NB + 1) + q];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// stage 1: panel QR, one launch per column
// ---------------------------------------------------------------------------
template <class T>
struct PanelArgs {
  T* A;            // the matrix (full storage), ld
  int ld, N, b;
  int c0, r0, m, k;     // panel columns c0..c0+b-1, rows r0..N-1 (m of them), k reflectors
  zc* Pw;          // panel working copy [b][N] (row i of the panel at Pw[i + N*j])
  zc* part;        // [2][nwg][BMAX + 1] partial sums (column c: slot c & 1)
  zc* rowc;        // [2][BMAX] row c of the panel before reflector c
  zc* tau;         // [BMAX]
  int nwg;
};

template <class T, int B>
__global__ __launch_bounds__(PR) void k_pqr_col(PanelArgs<T> a, int c) {
  __shared__ zc red[5 * (B + 1)];
  __shared__ zc sh[B + 4];
  __shared__ zc shs[256 / (B + 1)][B + 1];
  const int i = blockIdx.x * PR + threadIdx.x;       // panel row
  const bool live = i < a.m;
  const size_t N = a.N;
  zc row[B];
#pragma unroll
  for (int j = 0; j < B; ++j) row[j] = mkz(0, 0);
  if (live) {
    if (c == 0) {
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j < a.b) row[j] = ldz(a.A[(size_t)(a.r0 + i) + (size_t)a.ld * (a.c0 + j)]);
    } else {
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j >= c - 1 && j < a.b) row[j] = a.Pw[i + N * j];
    }
  }
  if (c > 0) {
    // finish reflector c - 1 from the partial sums of the previous launch
    const int cp = c - 1;
    const zc* pp = a.part + (size_t)(cp & 1) * a.nwg * (BMAX + 1);
    // sh[0].re = |x|^2 below the diagonal, sh[1 + j] = x^H P(:, j): the workgroups' partial
    // sums, NS threads per value each adding every NS-th partial (eight loads in flight),
    // then the NS sub-sums in order
    constexpr int NS = 256 / (B + 1);
    if ((int)threadIdx.x < NS * (B + 1)) {
      const int q = threadIdx.x % (B + 1), sub = threadIdx.x / (B + 1);
      shs[sub][q] = strided_sum(pp + q, BMAX + 1, sub, a.nwg, NS, mkz(0, 0));
    }
    __syncthreads();
    if ((int)threadIdx.x < B + 1) {
      zc t = shs[0][threadIdx.x];
      for (int u = 1; u < NS; ++u) t += shs[u][threadIdx.x];
      sh[threadIdx.x] = t;
    }
    __syncthreads();
    const zc* rc = a.rowc + (cp & 1) * BMAX;
    const zc alpha = rc[cp];
    zc beta, tau, scale;
    larfg(alpha, sh[0].re, beta, tau, scale);
    const zc ctau = conj(tau);
    // v_i = 1 (i == cp), x_i scale (i > cp), 0 above
    zc xcp = mkz(0, 0);                // row[cp] without a runtime index into the array
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (j == cp) xcp = row[j];
    zc vi = mkz(0, 0);
    if (live && i == cp) vi = mkz(1, 0);
    if (live && i > cp) vi = xcp * scale;
    // w_j = v^H P(:, j) = P(cp, j) + conj(scale) (x^H P(:, j) - conj(alpha) P(cp, j))
    const zc cs = conj(scale);
#pragma unroll
    for (int j = 0; j < B; ++j) {
      if (j > cp && j < a.b) {
        const zc pj = rc[j];
        const zc wj = pj + cs * (sh[1 + j] - conj(alpha) * pj);
        if (live && i >= cp) row[j] -= (ctau * vi) * wj;
      }
    }
#pragma unroll
    for (int j = 0; j < B; ++j)
      if (j == cp && live && i >= cp) row[j] = (i == cp) ? beta : vi;
    if (blockIdx.x == 0 && threadIdx.x == 0) a.tau[cp] = tau;
    __syncthreads();
  }
  if (c < a.k) {
    // partial sums of column c over this workgroup's rows
    zc v[B];
    double x0 = 0.0;
#pragma unroll
    for (int q = 0; q < B; ++q) v[q] = mkz(0, 0);
    if (live && i >= c) {
      zc x = mkz(0, 0);
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j == c) x = row[j];
      if (i > c) x0 = norm2(x);
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j > c && j < a.b) v[j] = fma_conj(x, row[j], mkz(0, 0));
    }
    block_sum_bf<B>(x0, v, red);
    if ((int)threadIdx.x < B + 1) {
      zc* pp = a.part + (size_t)(c & 1) * a.nwg * (BMAX + 1) + (size_t)blockIdx.x * (BMAX + 1);
      pp[threadIdx.x] = red[4 * (B + 1) + threadIdx.x];
    }
    if (live && i == c) {
      zc* rc = a.rowc + (c & 1) * BMAX;
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j >= c && j < a.b) rc[j] = row[j];
    }
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < B; ++j)
      if ((c == 0 || j >= c - 1) && j < a.b) a.Pw[i + N * j] = row[j];
  }
}

// explicit V (unit lower trapezoidal, zero above) and R = upper triangle of the panel back
// into the matrix
template <class T, int B>
__global__ __launch_bounds__(PR) void k_pfin(PanelArgs<T> a, zc* __restrict__ V) {
  const int i = blockIdx.x * PR + threadIdx.x;
  if (i >= a.m) return;
  const size_t N = a.N;
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if (j >= a.b) continue;
    const zc p = a.Pw[i + N * j];
    if (i <= j) stz(a.A[(size_t)(a.r0 + i) + (size_t)a.ld * (a.c0 + j)], p);    // R
    zc x = mkz(0, 0);
    if (j < a.k) x = i == j ? mkz(1, 0) : (i > j ? p : mkz(0, 0));
    V[i + N * j] = x;
  }
}

// The panel in one launch: workgroup w keeps rows 256w .. 256w + 255 of the panel in
// registers through all of its k reflectors, and the workgroups meet once per column at a
// counter (bar[c]: the workgroups that have published their partial sums of column c),
// so the arithmetic is k_pqr_col's without the per-column launches and panel round trips.
// Every workgroup must be resident at once: the host launches it only with at most one
// workgroup per compute unit and no other kernel sharing the device, and every wait is
// bounded (err[0] set past the limit, after which every workgroup returns).  At the end
// the panel goes out as V (unit lower trapezoidal) and R (k_pfin's job).
template <class T, int B>
__global__ __launch_bounds__(PR) void k_panel(PanelArgs<T> a, zc* __restrict__ V, int* __restrict__ bar,
                                              int* __restrict__ err, long long spin_limit) {
  constexpr int NS = 256 / (B + 1);
  __shared__ zc red[5 * 17], red2[5 * 17];
  __shared__ zc sh[B + 4];
  __shared__ zc shs[NS][B + 1];
  __shared__ zc rcs[B];
  __shared__ int flag;
  const int i = blockIdx.x * PR + threadIdx.x;       // panel row
  const bool live = i < a.m;
  const size_t N = a.N;
  zc row[B];
#pragma unroll
  for (int j = 0; j < B; ++j)
    row[j] = (live && j < a.b) ? ldz(a.A[(size_t)(a.r0 + i) + (size_t)a.ld * (a.c0 + j)]) : mkz(0, 0);
  for (int c = 0; c <= a.k; ++c) {
    if (c > 0) {
      const int cp = c - 1;
      // every workgroup's partial sums of column cp are out
      if (threadIdx.x == 0) {
        long long n = 0;
        int bad = 0;
        while (__hip_atomic_load(bar + cp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < a.nwg) {
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++n > spin_limit) {
            bad = 1;
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        flag = bad;
      }
      __syncthreads();
      if (flag) return;
      // sh[0].re = |x|^2 below the diagonal, sh[1 + j] = x^H P(:, j): NS threads per value
      // each adding every NS-th workgroup's partial, then the NS sub-sums in order
      const zc* pp = a.part + (size_t)(cp & 1) * a.nwg * (BMAX + 1);
      if ((int)threadIdx.x < NS * (B + 1)) {
        const int q = threadIdx.x % (B + 1), sub = threadIdx.x / (B + 1);
        shs[sub][q] = strided_sum_dev(pp + q, BMAX + 1, sub, a.nwg, NS);
      }
      if ((int)threadIdx.x < B) rcs[threadIdx.x] = ld_dev(a.rowc + (cp & 1) * BMAX + threadIdx.x);
      __syncthreads();
      if ((int)threadIdx.x < B + 1) {
        zc t = shs[0][threadIdx.x];
        for (int u = 1; u < NS; ++u) t += shs[u][threadIdx.x];
        sh[threadIdx.x] = t;
      }
      __syncthreads();
      const zc alpha = rcs[cp];
      zc beta, tau, scale;
      larfg(alpha, sh[0].re, beta, tau, scale);
      const zc ctau = conj(tau);
      // v_i = 1 (i == cp), x_i scale (i > cp), 0 above
      zc xcp = mkz(0, 0);
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j == cp) xcp = row[j];
      zc vi = mkz(0, 0);
      if (live && i == cp) vi = mkz(1, 0);
      if (live && i > cp) vi = xcp * scale;
      // w_j = v^H P(:, j) = P(cp, j) + conj(scale) (x^H P(:, j) - conj(alpha) P(cp, j))
      const zc cs = conj(scale);
#pragma unroll
      for (int j = 0; j < B; ++j) {
        if (j > cp && j < a.b) {
          const zc pj = rcs[j];
          const zc wj = pj + cs * (sh[1 + j] - conj(alpha) * pj);
          if (live && i >= cp) row[j] -= (ctau * vi) * wj;
        }
      }
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j == cp && live && i >= cp) row[j] = (i == cp) ? beta : vi;
      if (blockIdx.x == 0 && threadIdx.x == 0) a.tau[cp] = tau;
    }
    if (c < a.k) {
      // partial sums of column c over this workgroup's rows, 16 columns at a time (fewer
      // live registers), out with row c
      const bool on = live && i >= c;
      zc x = mkz(0, 0);
#pragma unroll
      for (int j = 0; j < B; ++j)
        if (j == c) x = row[j];
      const double x0 = (on && i > c) ? norm2(x) : 0.0;
#pragma unroll
      for (int h = 0; h < B / 16; ++h) {
        zc v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int j = 16 * h + q;
          v[q] = (on && j > c && j < a.b) ? fma_conj(x, row[j], mkz(0, 0)) : mkz(0, 0);
        }
        block_sum_bf<16>(h == 0 ? x0 : 0.0, v, h == 0 ? red : red2);
      }
      if ((int)threadIdx.x < B + 1) {
        const int q = threadIdx.x;
        const zc val = (q <= 16) ? red[4 * 17 + q] : red2[4 * 17 + q - 16];
        st_dev(a.part + (size_t)(c & 1) * a.nwg * (BMAX + 1) + (size_t)blockIdx.x * (BMAX + 1) + q, val);
      }
      if (live && i == c) {
#pragma unroll
        for (int j = 0; j < B; ++j)
          if (j >= c && j < a.b) st_dev(a.rowc + (c & 1) * BMAX + j, row[j]);
      }
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (threadIdx.x == 0) __hip_atomic_fetch_add(bar + c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (!live) return;
#pragma unroll
  for (int j = 0; j < B; ++j) {
    if (j >= a.b) continue;
    if (i <= j) stz(a.A[(size_t)(a.r0 + i) + (size_t)a.ld * (a.c0 + j)], row[j]);    // R
    zc x = mkz(0, 0);
    if (j < a.k) x = i == j ? mkz(1, 0) : (i > j ? row[j] : mkz(0, 0));
    V[i + N * j] = x;
  }
}

// one workgroup: T (ZLARFT, forward, columnwise) from G = V^H V and tau, and t2 = T^H t1
// (t1 = V^H gc); G and t1 column-major with ld B, T stored [col][row].  T = (I + D U)^-1 D
// with D = diag(tau), U the strict upper triangle of G (the closed form of ZLARFT's
// recurrence T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c), T(c, c) = tau_c, which also gives
// T(:, c) = 0 when tau_c = 0), so every column is an independent back substitution:
// column c by the eight lanes t = 8c .. 8c + 7, lane j holding T(q, c) for q = j (mod 8).
template <int B>
__global__ __launch_bounds__(256) void k_pT(int k, const zc* __restrict__ G, const zc* __restrict__ t1,
                                            const zc* __restrict__ tau, int nrhs, zc* __restrict__ Tm,
                                            zc* __restrict__ t2) {
  __shared__ zc Gs[B * B];
  __shared__ zc Ts[B * B];
  __shared__ zc taus[B];
  for (int q = threadIdx.x; q < B * B; q += 256) Gs[q] = G[q];   // G(r, c) at [c][r]
  if ((int)threadIdx.x < B) taus[threadIdx.x] = ((int)threadIdx.x < k) ? tau[threadIdx.x] : mkz(0, 0);
  __syncthreads();
  const int c = threadIdx.x >> 3, j = threadIdx.x & 7;
  constexpr int NU = B / 8;
  if (c < B) {
    zc tq[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) tq[u] = (j + 8 * u == c) ? taus[c] : mkz(0, 0);
    for (int r = c - 1; r >= 0; --r) {
      zc acc = mkz(0, 0);
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int q = j + 8 * u;
        if (q > r && q <= c) acc = fma_z(Gs[q * B + r], tq[u], acc);
      }
#pragma unroll
      for (int off = 1; off < 8; off <<= 1) {
        acc.re += __shfl_xor(acc.re, off);
        acc.im += __shfl_xor(acc.im, off);
      }
      const zc val = mkz(0, 0) - taus[r] * acc;
#pragma unroll
      for (int u = 0; u < NU; ++u)
        if (j + 8 * u == r) tq[u] = val;
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) Ts[c * B + j + 8 * u] = tq[u];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < B * B; q += 256) Tm[q] = Ts[q];
  // t2(p, r) = sum_q conj(T(q, p)) t1(q, r), T(q, p) = Ts[p][q], q <= p
  for (int q = threadIdx.x; q < B * 6; q += 256) {
    const int p = q / 6, r = q % 6;
    zc sacc = mkz(0, 0);
    if (r < nrhs)
      for (int qq = 0; qq <= p; ++qq) sacc = fma_conj(Ts[p * B + qq], t1[(size_t)r * B + qq], sacc);
    t2[q] = sacc;
  }
}

// Y = V T (rows of the panel) and gc -= V t2
template <int B>
__global__ __launch_bounds__(PR) void k_pY(int m, int N, int r0, const zc* __restrict__ V, const zc* __restrict__ Tm,
                                           const zc* __restrict__ t2, int nrhs, zc* __restrict__ Y, zc* __restrict__ gc) {
  __shared__ zc Ts[B * B];
  __shared__ zc t2s[6 * B];
  for (int q = threadIdx.x; q < B * B; q += PR) Ts[q] = Tm[q];
  for (int q = threadIdx.x; q < 6 * B; q += PR) t2s[q] = t2[q];
  __syncthreads();
  const int i = blockIdx.x * PR + threadIdx.x;
  if (i >= m) return;
  zc v[B];
#pragma unroll
  for (int j = 0; j < B; ++j) v[j] = V[i + (size_t)N * j];
#pragma unroll
  for (int c = 0; c < B; ++c) {
    zc s = mkz(0, 0);
#pragma unroll
    for (int q = 0; q <= c; ++q) s = fma_z(v[q], Ts[c * B + q], s);
    Y[i + (size_t)N * c] = s;
  }
  for (int r = 0; r < nrhs; ++r) {
    zc s = mkz(0, 0);
#pragma unroll
    for (int q = 0; q < B; ++q) s = fma_z(v[q], t2s[q * 6 + r], s);
    gc[(size_t)r * N + r0 + i] -= s;
  }
}

// W = X - 1/2 V Z with Z = T^H M, M = V^H X (column-major, ld B)
template <int B>
__global__ __launch_bounds__(PR) void k_pW(int m, int N, const zc* __restrict__ V, const zc* __restrict__ Tm,
                                           const zc* __restrict__ M, zc* __restrict__ X) {
  __shared__ zc Zs[B * B];
  __shared__ zc Ts[B * B];
  __shared__ zc Ms[B * B];
  for (int q = threadIdx.x; q < B * B; q += PR) {
    Ts[q] = Tm[q];
    Ms[q] = M[q];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < B * B; q += PR) {
    const int p = q / B, c = q % B;      // Z(p, c) = sum_r conj(T(r, p)) M(r, c), T(r, p) = Ts[p][r], r <= p
    zc sacc = mkz(0, 0);
    for (int r = 0; r <= p; ++r) sacc = fma_conj(Ts[p * B + r], Ms[c * B + r], sacc);
    Zs[q] = sacc;
  }
  __syncthreads();
  const int i = blockIdx.x * PR + threadIdx.x;
  if (i >= m) return;
  zc v[B];
#pragma unroll
  for (int j = 0; j < B; ++j) v[j] = V[i + (size_t)N * j];
#pragma unroll
  for (int c = 0; c < B; ++c) {
    zc sacc = mkz(0, 0);
#pragma unroll
    for (int p = 0; p < B; ++p) sacc = fma_z(v[p], Zs[p * B + c], sacc);
    X[i + (size_t)N * c] -= 0.5 * sacc;
  }
}

// Tall-skinny products out0 = A^H Y0 (B x B) and out1 = A^H Y1 (B x n1, n1 <= 16) over m
// rows (A, Y1 with ld N, Y0 with ld ldy0): FP64 MFMA on 64-row chunks staged through LDS,
// chunks c = blockIdx.x, + gridDim.x, ..., one partial per workgroup, summed in workgroup
// order by k_gram_sum.  Wave w accumulates the 16 x 16 output tiles w and w + 4.
#define GRAM_ROWS 64
#define GRAM_PMAX 64
template <int B>
__global__ __launch_bounds__(256) void k_gram(int m, int N, const zc* __restrict__ A, const zc* __restrict__ Y0,
                                              int ldy0, const zc* __restrict__ Y1, int n1, zc* __restrict__ part) {
  constexpr int NY = B + 16;                 // columns of [Y0 | Y1]
  constexpr int TI = B / 16;                 // output tile rows
  __shared__ zc As[GRAM_ROWS][B];
  __shared__ zc Ys[GRAM_ROWS][NY];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int ntile = TI * (TI + (n1 > 0 ? 1 : 0));
  d4 acc[2][2];                              // [tile slot][re, im]
  for (int x = 0; x < 2; ++x) acc[x][0] = acc[x][1] = d4{0, 0, 0, 0};
  const int nch = (m + GRAM_ROWS - 1) / GRAM_ROWS;
  for (int ch = blockIdx.x; ch < nch; ch += gridDim.x) {
    const int row0 = ch * GRAM_ROWS;
    __syncthreads();
    for (int e = t; e < GRAM_ROWS * B; e += 256) {
      const int rr = e % GRAM_ROWS, cc = e / GRAM_ROWS, row = row0 + rr;
      As[rr][cc] = row < m ? A[row + (size_t)N * cc] : mkz(0, 0);
      Ys[rr][cc] = row < m ? Y0[row + (size_t)ldy0 * cc] : mkz(0, 0);
    }
    for (int e = t; e < GRAM_ROWS * 16; e += 256) {
      const int rr = e % GRAM_ROWS, cc = e / GRAM_ROWS, row = row0 + rr;
      Ys[rr][B + cc] = (row < m && cc < n1) ? Y1[row + (size_t)N * cc] : mkz(0, 0);
    }
    __syncthreads();
    for (int x = 0; x < 2; ++x) {
      const int tile = wave + 4 * x;
      if (tile >= ntile) continue;
      const int ti = tile % TI, tj = tile / TI;
      const int i = ti * 16 + (lane & 15), j = tj * 16 + (lane & 15);
#pragma unroll 4
      for (int k0 = 0; k0 < GRAM_ROWS; k0 += 4) {
        const int kk = k0 + (lane >> 4);
        const zc av = As[kk][i], yv = Ys[kk][j];
        // conj(a) y = (ar yr + ai yi) + i (ar yi - ai yr)
        acc[x][0] = mfma16x16x4(av.re, yv.re, acc[x][0]);
        acc[x][0] = mfma16x16x4(av.im, yv.im, acc[x][0]);
        acc[x][1] = mfma16x16x4(av.re, yv.im, acc[x][1]);
        acc[x][1] = mfma16x16x4(-av.im, yv.re, acc[x][1]);
      }
    }
  }
  // partial: [blockIdx][B][NY] (row i of the output, column j of [Y0 | Y1])
  zc* pp = part + (size_t)blockIdx.x * B * NY;
  for (int x = 0; x < 2; ++x) {
    const int tile = wave + 4 * x;
    if (tile >= ntile) continue;
    const int ti = tile % TI, tj = tile / TI;
    for (int q = 0; q < 4; ++q) {
      const int i = ti * 16 + (lane >> 4) + 4 * q, j = tj * 16 + (lane & 15);
      pp[(size_t)i * NY + j] = mkz(acc[x][0][q], acc[x][1][q]);
    }
  }
}

// out0 (B x B, ld B) and out1 (B x n1, ld B) from the P partials, in workgroup order: four
// threads per output each add a quarter of the partials, then the quarters in order
template <int B>
__global__ __launch_bounds__(256) void k_gram_sum(int P, const zc* __restrict__ part, int n1, zc* __restrict__ out0,
                                                  zc* __restrict__ out1) {
  constexpr int NY = B + 16;
  const int o = blockIdx.x * 64 + (threadIdx.x >> 2), sub = threadIdx.x & 3;
  const int i = o / NY, j = o % NY;
  const bool live = i < B && (j < B || j - B < n1);
  const int per = (P + 3) / 4, w0 = sub * per, w1 = min(P, w0 + per);
  zc x = mkz(0, 0);
  if (live && w0 < w1) x = strided_sum(part + (size_t)i * NY + j, (size_t)B * NY, w0, w1, 1, mkz(0, 0));
  const zc x1 = shfl_xor_z(x, 1);
  const zc x2 = shfl_xor_z(x, 2), x3 = shfl_xor_z(x, 3);
  if (!live || sub != 0) return;
  const zc sum = ((x + x1) + x2) + x3;       // quarters 0, 1, 2, 3
  if (j < B) out0[i + (size_t)B * j] = sum;
  else out1[i + (size_t)B * (j - B)] = sum;
}

// X = A Y for the Hermitian trailing matrix A (m x m, lower triangle, ld) and the panel's
// Y (m x BC, ld ldy) on MFMA: FP32 operands and accumulation for the COMPLEX*8 matrix,
// FP64 for the COMPLEX*16 one.  A workgroup owns 128 rows and a K slice (blockIdx.y) and
// walks it in steps of 16 staged through LDS, the next step's loads in flight while the
// current one is multiplied; wave w computes rows 32w..32w+31 for all BC columns.  A step
// below the diagonal reads A(i, k) down the columns; a step above it reads conj(A(k, i))
// down the columns of the transposed block; a step across it reads element by element.
// With more than one slice the partials go to part[z][col][row] and k_hmm_sum adds them in
// slice order.
#define HB_R 128
#define HB_SMAX 16      // most K slices
template <class T, int BC>
__global__ __launch_bounds__(256) void k_hmm_b(int m, const T* __restrict__ A, int ld, const zc* __restrict__ Y,
                                               int ldy, int kc, zc* __restrict__ part, zc* __restrict__ X, int ldx) {
  constexpr bool F64 = sizeof(T) == sizeof(zc);
  typedef typename std::conditional<F64, double2, float2>::type ct;
  typedef typename std::conditional<F64, d4, f4>::type acc_t;
  constexpr int KS = F64 ? 16 : 32;                 // K per step (32 KB of the matrix in LDS)
  constexpr int NA = KS * HB_R / 256;               // matrix elements per thread and step
  constexpr int CT = BC / 16;                       // 16-column tiles
  constexpr int YP = KS * BC / 256;                 // Y elements per thread and step
  // rows padded by one element: the four k rows a wave reads at once (and the 16 rows the
  // transposed loads store at once) fall in different LDS banks
  __shared__ ct As[KS][HB_R + 1];
  __shared__ ct Ys[KS][BC + 1];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int R0 = blockIdx.x * HB_R;
  const int kbeg = blockIdx.y * kc, kend = min(m, kbeg + kc);
  // loaders: A(row R0 + (t & 127), k + (t >> 7) + 2u) below or across the diagonal,
  // conj(A(k + (t & 15) + 16 (u / 8), row R0 + (t >> 4) + 16 (u % 8))) above it (rows past m
  // read row m - 1, never stored); Y(k + (t % KS), t / KS + (256 / KS) u)
  const int arow = min(R0 + (t & 127), m - 1), ak = t >> 7;
  const int tk = t & 15, ti = t >> 4;
  const int yk = t % KS, ycol = t / KS;
  int lmode = 0;                                    // 0 below, 1 above, 2 across the diagonal
  acc_t cr[2][CT], ci[2][CT];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b) cr[a][b] = ci[a][b] = acc_t{0, 0, 0, 0};
  T ra[NA];
  zc ry[YP];
  auto gload = [&](int k0) {
    lmode = (k0 + KS - 1 <= R0) ? 0 : (k0 > R0 + HB_R - 1) ? 1 : 2;
    if (lmode == 1) {
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        const int k = min(k0 + tk + 16 * (u / 8), m - 1);
        const int i = min(R0 + ti + 16 * (u % 8), m - 1);
        ra[u] = A[(size_t)k + (size_t)ld * i];
      }
    } else {
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        const int k = min(k0 + ak + 2 * u, m - 1);  // past kend: Y is zero there
        ra[u] = (lmode == 0 || arow >= k) ? A[(size_t)arow + (size_t)ld * k] : A[(size_t)k + (size_t)ld * arow];
        if (lmode == 2 && arow < k) ra[u].im = -ra[u].im;
      }
    }
#pragma unroll
    for (int u = 0; u < YP; ++u) {
      const int k = k0 + yk;
      ry[u] = (k < kend) ? Y[(size_t)k + (size_t)ldy * (ycol + (256 / KS) * u)] : mkz(0, 0);
    }
  };
  gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += KS) {
    __syncthreads();                                  // the previous step is consumed
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      ct v;
      v.x = ra[u].re;
      v.y = ra[u].im;
      if (lmode == 1) {
        v.y = -v.y;                                 // A(i, k) = conj(A(k, i))
        As[tk + 16 * (u / 8)][ti + 16 * (u % 8)] = v;
      } else {
        As[ak + 2 * u][t & 127] = v;
      }
    }
#pragma unroll
    for (int u = 0; u < YP; ++u) {
      ct v;
      v.x = ry[u].re;
      v.y = ry[u].im;
      Ys[yk][ycol + (256 / KS) * u] = v;
    }
    __syncthreads();
    if (k0 + KS < kend) gload(k0 + KS);
#pragma unroll
    for (int kk = 0; kk < KS / 4; ++kk) {
      const int k = 4 * kk + kq;
      ct av[2], yv[CT];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[k][32 * w + 16 * a + i16];
#pragma unroll
      for (int b = 0; b < CT; ++b) yv[b] = Ys[k][16 * b + i16];
      // (ar + i ai)(yr + i yi) = (ar yr - ai yi) + i (ar yi + ai yr).  Each accumulator
      // takes its two products in this order; the first products of all accumulators go
      // before the second ones, so that no MFMA waits on the one just issued
#pragma unroll
      for (int b = 0; b < CT; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          if constexpr (F64) {
            cr[a][b] = mfma16x16x4(av[a].x, yv[b].x, cr[a][b]);
            ci[a][b] = mfma16x16x4(av[a].x, yv[b].y, ci[a][b]);
          } else {
            cr[a][b] = mfma16x16x4f32(av[a].x, yv[b].x, cr[a][b]);
            ci[a][b] = mfma16x16x4f32(av[a].x, yv[b].y, ci[a][b]);
          }
        }
#pragma unroll
      for (int b = 0; b < CT; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          if constexpr (F64) {
            cr[a][b] = mfma16x16x4(-av[a].y, yv[b].y, cr[a][b]);
            ci[a][b] = mfma16x16x4(av[a].y, yv[b].x, ci[a][b]);
          } else {
            cr[a][b] = mfma16x16x4f32(-av[a].y, yv[b].y, cr[a][b]);
            ci[a][b] = mfma16x16x4f32(av[a].y, yv[b].x, ci[a][b]);
          }
        }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b) {
      const int col = 16 * b + i16;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // f64 MFMA: register q holds row kq + 4q of the tile; f32 MFMA: row 4kq + q
        const int row = R0 + 32 * w + 16 * a + (F64 ? kq + 4 * q : 4 * kq + q);
        if (row >= m) continue;
        const zc v = mkz((double)cr[a][b][q], (double)ci[a][b][q]);
        if (part) part[(size_t)blockIdx.y * m * BC + (size_t)col * m + row] = v;
        else X[(size_t)row + (size_t)ldx * col] = v;
      }
    }
}

// X(row, col) = sum over the ns slices of part[z][col][row], in slice order
template <int BC>
__global__ __launch_bounds__(256) void k_hmm_sum(int m, int ns, const zc* __restrict__ part, zc* __restrict__ X,
                                                 int ldx) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)m * BC) return;
  const int row = (int)(e % m), col = (int)(e / m);
  X[(size_t)row + (size_t)ldx * col] = strided_sum(part + e, (size_t)m * BC, 0, ns, 1, mkz(0, 0));
}

// A22 -= U, U = V W^H + W V^H, on the 64 x 64 tiles of the trailing matrix (rows and
// columns r0..N-1) on and below the diagonal, K = 2 b: FP64 MFMA on the COMPLEX*16 matrix, FP32 MFMA (FP32 operands
// and accumulation) on the COMPLEX*8 one.  The MFMA rows run over the tile's columns and
// its columns over the tile's rows, D(i, j) = U(c_i, r_j) = conj(U(r_j, c_i)) since U is
// Hermitian, so that a lane's result lands next to its neighbours' in the column-major
// matrix.
template <class T, int B>
__global__ __launch_bounds__(256, 4) void k_her2k_full(T* __restrict__ A, int ld, int m, int r0, int N,
                                                    const zc* __restrict__ V, const zc* __restrict__ W) {
  // one workgroup per tile on or below the diagonal: t -> tile row ti >= tile column tj
  const int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  const int tj = t - ti * (ti + 1) / 2;
  constexpr bool F64 = sizeof(T) == sizeof(zc);
  constexpr int K = 2 * B;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci0 = tj * 64 + (wave >> 1) * 32;           // tile columns (MFMA rows)
  const int rj0 = ti * 64 + (wave & 1) * 32;            // tile rows (MFMA columns)
  const int lr = lane & 15, lk = lane >> 4;
  typedef typename std::conditional<F64, d4, f4>::type acc_t;
  // COMPLEX*8: the tile's current values are loaded first, so that their latency overlaps
  // the products (at N = 36864 6.2 s against 6.5 s per matrix); the COMPLEX*16 tile would
  // cost the registers of two more waves per SIMD (8.5 % slower at N = 12000)
  constexpr int HV = F64 ? 1 : 16;
  T hv[HV];
  if constexpr (!F64) {
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int c = ci0 + x * 16 + 4 * lk + q, r = rj0 + y * 16 + lr;
          if (r < m && c < m) hv[(x * 2 + y) * 4 + q] = A[(size_t)(r0 + r) + (size_t)ld * (r0 + c)];
        }
  }
  acc_t cr[2][2], ci[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) cr[x][y] = ci[x][y] = acc_t{0, 0, 0, 0};
  if constexpr (!F64) {
    // COMPLEX*8: the tile's operands go through LDS once per workgroup, in FP32, half of K
    // at a time (a wave pair shares each operand); the products run in the order of k
    __shared__ zf As[B][65], Bs[B][65];     // padded rows: no bank conflicts between k rows
    const int cl0 = tj * 64, rl0 = ti * 64;
    const int ciw = (wave >> 1) * 32, rjw = (wave & 1) * 32;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      // half 0: A = V(c, k), B = W(r, k); half 1: A = W(c, k), B = V(r, k)
      const zc* Asrc = h == 0 ? V : W;
      const zc* Bsrc = h == 0 ? W : V;
      if (h > 0) __syncthreads();
#pragma unroll 4
      for (int u = 0; u < B * 64 / 256; ++u) {
        const int e = threadIdx.x + 256 * u, kk = e >> 6, cc = e & 63;
        const int c = cl0 + cc, r = rl0 + cc;
        const zc a = (c < m) ? Asrc[(size_t)kk * N + c] : mkz(0, 0);
        const zc b = (r < m) ? Bsrc[(size_t)kk * N + r] : mkz(0, 0);
        As[kk][cc] = zf{(float)a.re, (float)a.im};
        Bs[kk][cc] = zf{(float)b.re, (float)b.im};
      }
      __syncthreads();
#pragma unroll
      for (int k0 = 0; k0 < B; k0 += 4) {
        const int kk = k0 + lk;
        zf af[2], bf[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          af[x] = As[kk][ciw + x * 16 + lr];
          bf[x] = Bs[kk][rjw + x * 16 + lr];
        }
        // first products of all accumulators, then the second ones (the order within
        // each accumulator is kept)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            cr[x][y] = mfma16x16x4f32(af[x].re, bf[y].re, cr[x][y]);
            ci[x][y] = mfma16x16x4f32(af[x].im, bf[y].re, ci[x][y]);
          }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) {
            cr[x][y] = mfma16x16x4f32(af[x].im, bf[y].im, cr[x][y]);
            ci[x][y] = mfma16x16x4f32(-af[x].re, bf[y].im, ci[x][y]);
          }
      }
    }
  } else {
  // A operand (i = column c, k): [V | W](c, k); B operand (k, j = row r): conj([W | V](r, k));
  // the operands of step k0 + 4 are loaded while step k0 is multiplied
  zc av[2], bv[2];
  auto oload = [&](int k0, zc (&pa)[2], zc (&pb)[2]) {
    const int k = k0 + lk;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int c = ci0 + x * 16 + lr, r = rj0 + x * 16 + lr;
      pa[x] = (c < m) ? ((k < B) ? V[(size_t)k * N + c] : W[(size_t)(k - B) * N + c]) : mkz(0, 0);
      pb[x] = (r < m) ? ((k < B) ? W[(size_t)k * N + r] : V[(size_t)(k - B) * N + r]) : mkz(0, 0);
    }
  };
  oload(0, av, bv);
#pragma unroll
  for (int k0 = 0; k0 < K; k0 += 4) {
    zc an[2], bn[2];
    if (k0 + 4 < K) oload(k0 + 4, an, bn);
    // a * conj(b): the first products of all accumulators, then the second ones (the order
    // within each accumulator is kept)
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        cr[x][y] = mfma16x16x4(av[x].re, bv[y].re, cr[x][y]);
        ci[x][y] = mfma16x16x4(av[x].im, bv[y].re, ci[x][y]);
      }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        cr[x][y] = mfma16x16x4(av[x].im, bv[y].im, cr[x][y]);
        ci[x][y] = mfma16x16x4(-av[x].re, bv[y].im, ci[x][y]);
      }
    if (k0 + 4 < K) {
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        av[x] = an[x];
        bv[x] = bn[x];
      }
    }
  }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // f64 MFMA: register q holds D row lk + 4q; f32 MFMA: row 4 lk + q
        const int c = ci0 + x * 16 + (F64 ? lk + 4 * q : 4 * lk + q);
        const int r = rj0 + y * 16 + lr;
        if (r < m && c < m) {
          zc v;
          if constexpr (F64) v = ldz(A[(size_t)(r0 + r) + (size_t)ld * (r0 + c)]);
          else v = ldz(hv[(x * 2 + y) * 4 + q]);
          v.re -= cr[x][y][q];
          v.im += ci[x][y][q];       // - conj(D)
          stz(A[(size_t)(r0 + r) + (size_t)ld * (r0 + c)], v);
        }
      }
}

// band of width b of the lower triangle into BD (FP64), zero for b < d <= 2b; gc = conj(g)
template <class T>
__global__ void k_band_extract(const T* __restrict__ A, int ld, int N, int b, int LB, zc* __restrict__ BD) {
  const int c = blockIdx.x;
  for (int d = threadIdx.x; d < LB; d += blockDim.x) {
    zc v = mkz(0, 0);
    if (d <= b && c + d < N) v = ldz(A[(size_t)(c + d) + (size_t)ld * c]);
    if (d == 0) v.im = 0.0;
    BD[(size_t)c * LB + d] = v;
  }
}

__global__ void k_conj_vec(int n, const zc* __restrict__ x, zc* __restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = conj(x[i]);
}

// ---------------------------------------------------------------------------
// stage 2: bulge chasing
// ---------------------------------------------------------------------------
struct ChaseArgs {
  zc* BD;            // [N][LB]
  int N, b, LB;
  zc* gc;            // nrhs x N (conjugated tracked vectors)
  int nrhs;
  int* prog;         // [N] tasks completed per sweep
  int* ctl;          // [0] error flag, [32] next sweep to claim
  long long spin_limit;
};

__device__ inline int chase_ntasks(int N, int b, int s) { return 1 + (N - 2 - s) / b; }

// element (r, c) of the band, r >= c
__device__ inline zc& bd(const ChaseArgs& a, int r, int c) { return a.BD[(size_t)c * a.LB + (r - c)]; }


// sum over the B (16 or 32) lanes of a row group: lane pairs ^1, ^2 (quad permutations),
// then i <-> 7 - i, i <-> 15 - i within rows (half mirror, mirror), then ^16.  Every step
// adds two equal partial sums in either order, so every lane of the group ends with the
// same bits.
template <int B>
__device__ inline double group_sum(double x) {
  x += dpp_d<0xB1>(x);    // quad_perm [1, 0, 3, 2]
  x += dpp_d<0x4E>(x);    // quad_perm [2, 3, 0, 1]
  x += dpp_d<0x141>(x);   // row_half_mirror
  x += dpp_d<0x140>(x);   // row_mirror
  if constexpr (B == 32) x += swz16_d(x);
  return x;
}
template <int B>
__device__ inline zc group_sum(zc x) {
  return mkz(group_sum<B>(x.re), group_sum<B>(x.im));
}

// Column sums of a B x B block held as thread (r, h) = (t % B, t / B) owning row r and
// columns h NC .. h NC + NC - 1 (NC = B^2 / 256); the B rows of a column group lie in one
// wave.  B = 32: a halving butterfly on row bits 4 (swizzle) and 3 (DPP rotate by 8 in a
// row of 16), each lane keeping the half its bit selects, then a DPP sum over the 8-lane
// block: thread (r, h) ends with the sum of local column r >> 3.  B = 16 (one column per
// thread): the group sum.
template <int B, int NC>
__device__ inline zc col_sums(zc (&t)[NC], int lane) {
  if constexpr (NC == 4) {
    static_assert(B == 32, "four columns per thread for B = 32");
    bf_step<16, 2>(t, lane);
    bf_step<8, 1>(t, lane);
    zc x = t[0] + xchg<4>(t[0]);
    x += xchg<2>(x);
    x += xchg<1>(x);
    return x;
  } else {
    static_assert(B == 16 && NC == 1, "one column per thread for B = 16");
    return group_sum<16>(t[0]);
  }
}

// Row sums: a thread's own columns, the column groups of its wave, then the four waves
// through LDS in wave order (red: [4][B]); every thread gets the sum of its row.
template <int B>
__device__ inline zc row_sum(zc x, zc* red, int t) {
#pragma unroll
  for (int off = B; off < 64; off <<= 1) x += shfl_xor_z(x, off);
  const int wave = t >> 6, lane = t & 63;
  if (lane < B) red[wave * B + lane] = x;
  __syncthreads();
  const int r = t & (B - 1);
  return ((red[r] + red[B + r]) + red[2 * B + r]) + red[3 * B + r];
}

// One sweep per 256-thread workgroup, claimed in order from a counter; task k of sweep
// s waits until sweep s - 1 has finished its tasks 0..k+1 (the later tasks of sweep
// s - 1 touch only rows below this task's last row).  Thread (r, h) keeps row r, columns
// h NC .. of the task's blocks in registers: column sums stay inside a wave, row sums
// cross the waves through LDS (two barriers per task), and the reflector's scalars and
// v are formed redundantly by every column group from column 0, which every thread
// holds for its row.
template <int B>
__global__ __launch_bounds__(256) void k_chase(ChaseArgs a) {
  constexpr int LB2 = (B == 32) ? 5 : 4;
  constexpr int NC = B * B / 256;
  constexpr int LNC = (NC == 4) ? 2 : 0;
  constexpr int NG = 256 / B;                  // column groups
  __shared__ zc red1[4 * B], red2[4 * B];
  __shared__ int flag;
  const int t = threadIdx.x, lane = t & 63;
  const int r = t & (B - 1), h = t >> LB2;
  const int c0 = h * NC;
  const int gb = lane & ~(B - 1);              // lane of row 0 of this thread's group
  const int N = a.N, b = B;
  int* const err = a.ctl;
  for (;;) {
    if (t == 0) flag = atomicAdd(a.ctl + 32, 1);
    __syncthreads();
    const int s = flag;
    __syncthreads();
    if (s > N - 2) return;
    const int nt = chase_ntasks(N, b, s);
    const int ntp = s > 0 ? chase_ntasks(N, b, s - 1) : 0;
    int lenp = 0;
    zc taup = mkz(0, 0);
    zc vp[NC];                                 // previous reflector at this thread's columns
#pragma unroll
    for (int q = 0; q < NC; ++q) vp[q] = mkz(0, 0);
    for (int k = 0; k < nt; ++k) {
      if (s > 0) {
        if (t == 0) {
          const int need = min(k + 2, ntp);
          long long n = 0;
          int bad = 0;
          while (__hip_atomic_load(&a.prog[s - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++n > a.spin_limit) {
              bad = 1;
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          flag = bad;
        }
        __syncthreads();
        if (flag) return;
      }
      // rows st..ed of this task; for k >= 1 the previous reflector's rows sp..sp+lenp-1
      const int sp = s + 1 + (k - 1) * b;
      const int st = (k == 0) ? s + 1 : sp + lenp;
      const int ed = min(st + b - 1, N - 1), len = ed - st + 1;
      const bool rin = r < len;
      // loads: diagonal block (full Hermitian), the block left of it (k >= 1) and its
      // column 0 for every row, or column s (k = 0); the tracked vector of this group
      zc D[NC], O[NC];
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        const int c = c0 + q;
        zc v = mkz(0, 0);
        if (rin && c < len) v = (r >= c) ? ld_dev(&bd(a, st + r, st + c)) : conj(ld_dev(&bd(a, st + c, st + r)));
        D[q] = v;
        zc o = mkz(0, 0);
        if (k > 0 && rin && c < lenp) o = ld_dev(&bd(a, st + r, sp + c));
        O[q] = o;
      }
      zc x0 = mkz(0, 0);
      if (rin) x0 = (k == 0) ? ld_dev(&bd(a, st + r, s)) : ld_dev(&bd(a, st + r, sp));
      const bool gq = h < a.nrhs;
      zc g = (rin && gq) ? ld_dev(&a.gc[(size_t)h * N + st + r]) : mkz(0, 0);
      if (k > 0) {
        // (i) the block times the previous reflector from the right: O -= taup (O vp) vp^H
        zc w = mkz(0, 0);
#pragma unroll
        for (int q = 0; q < NC; ++q) w = fma_z(O[q], vp[q], w);
        w = row_sum<B>(w, red1, t);
        const zc tw = taup * w;
#pragma unroll
        for (int q = 0; q < NC; ++q) O[q] = O[q] - tw * conj(vp[q]);
        x0 = x0 - tw;                          // column 0: vp[0] = 1
      }
      // (ii) reflector of x = rows st..ed of the column (ZLARFG), in every column group
      const double xn2 = group_sum<B>((r >= 1 && rin) ? norm2(x0) : 0.0);
      const zc alpha = shfl_z(x0, gb);
      zc beta, tau, scale;
      larfg(alpha, xn2, beta, tau, scale);
      const zc vr = (r == 0) ? mkz(1, 0) : (rin ? x0 * scale : mkz(0, 0));
      zc vc[NC];                               // v at this thread's columns
#pragma unroll
      for (int q = 0; q < NC; ++q) vc[q] = shfl_z(vr, gb + c0 + q);
      if (k > 0) {
        // (iii) O(:, 1:) <- H^H O(:, 1:): z_c = v^H O(:, c), O -= conj(tau) v z
        zc tq[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) tq[q] = fma_conj(vr, O[q], mkz(0, 0));
        const zc z = col_sums<B, NC>(tq, lane);
        const zc ctv = conj(tau) * vr;
#pragma unroll
        for (int q = 0; q < NC; ++q) {
          const int c = c0 + q;
          const zc zq = shfl_z(z, gb + (q << (LB2 - LNC)));   // the lane holding column c0 + q
          if (c == 0) O[q] = (r == 0) ? beta : mkz(0, 0);
          else O[q] = O[q] - ctv * zq;
          if (rin && c < lenp) st_dev(&bd(a, st + r, sp + c), O[q]);
        }
      }
      // (iv) D <- H^H D H: y = D v, g = v^H y, w = tau y - |tau|^2 g / 2 v, D -= v w^H + w v^H
      zc y = mkz(0, 0);
#pragma unroll
      for (int q = 0; q < NC; ++q) y = fma_z(D[q], vc[q], y);
      y = row_sum<B>(y, red2, t);
      // column s (task 0): beta then zeros, stored only now that every wave has loaded it
      // (row_sum's barrier)
      if (k == 0 && h == 0 && rin) st_dev(&bd(a, st + r, s), (r == 0) ? beta : mkz(0, 0));
      const double gam = group_sum<B>(vr.re * y.re + vr.im * y.im);
      const zc wr = tau * y - (0.5 * norm2(tau) * gam) * vr;
#pragma unroll
      for (int q = 0; q < NC; ++q) {
        const int c = c0 + q;
        const zc wc = shfl_z(wr, gb + c);
        const zc d = D[q] - vr * conj(wc) - wr * conj(vc[q]);
        if (rin && c <= r) st_dev(&bd(a, st + r, st + c), (r == c) ? mkz(d.re, 0.0) : d);
      }
      // (v) tracked vector h, rows st..ed: gc -= conj(tau) v (v^H gc)
      const zc al = group_sum<B>(fma_conj(vr, g, mkz(0, 0)));
      if (rin && gq) st_dev(&a.gc[(size_t)h * N + st + r], g - vr * (conj(tau) * al));
      lenp = len;
      taup = tau;
#pragma unroll
      for (int q = 0; q < NC; ++q) vp[q] = vc[q];
      // publish: every wave's stores have completed before the counter moves
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
      __syncthreads();
      if (t == 0) __hip_atomic_store(&a.prog[s], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// d, e of the tridiagonal from the band; g = conj(gc)
__global__ void k_band_finish(const zc* __restrict__ BD, int N, int LB, double* __restrict__ d, double* __restrict__ e,
                              const zc* __restrict__ gc, zc* __restrict__ g, int nrhs, const int* __restrict__ err,
                              int* __restrict__ status) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0 && status && (err[0] || err[8])) status[0] |= 1 << WTB_ERR_EIG;
  if (j >= N) return;
  d[j] = BD[(size_t)j * LB].re;
  e[j] = (j + 1 < N) ? BD[(size_t)j * LB + 1].re : 0.0;
  for (int r = 0; r < nrhs; ++r) g[(size_t)r * N + j] = conj(gc[(size_t)r * N + j]);
}

// control block of the chase: its error flag and sweep counter 0 (the panel's flag, [8],
// is kept)
__global__ void k_chase_init(int* ctl) {
  const int i = threadIdx.x;
  if (i == 0 || i == 32) ctl[i] = 0;
}

__global__ void k_zero_ints(int* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
size_t band_work_bytes(int N, int b, int nrhs) {
  const size_t nwg = (N + PR - 1) / PR + 1;
  const size_t LB = 2 * (size_t)b + 1;
  size_t s = 0;
  auto add = [&](size_t bytes) { s += (bytes + 255) & ~size_t(255); };
  add(sizeof(zc) * (size_t)N * b);            // Pw
  add(sizeof(zc) * (size_t)N * b);            // V
  add(sizeof(zc) * (size_t)N * b);            // Y
  add(sizeof(zc) * (size_t)N * b);            // X / W
  add(sizeof(zc) * 2 * nwg * (BMAX + 1));     // part
  add(sizeof(zc) * 2 * BMAX);                 // rowc
  add(sizeof(zc) * BMAX);                     // tau
  add(sizeof(zc) * BMAX * BMAX);              // G
  add(sizeof(zc) * BMAX * 6);                 // t1
  add(sizeof(zc) * BMAX * BMAX);              // T
  add(sizeof(zc) * 6 * BMAX);                 // t2
  add(sizeof(zc) * BMAX * BMAX);              // M
  add(sizeof(zc) * nwg * BMAX * BMAX);        // K slices of the small products
  add(sizeof(zc) * GRAM_PMAX * BMAX * (BMAX + 16));   // partials of the tall-skinny products
  add(sizeof(zc) * HB_SMAX * (size_t)N * b);           // K-slice partials of X = A22 Y
  add(sizeof(zc) * (size_t)N * LB);           // band
  add(sizeof(zc) * (size_t)N * (nrhs > 0 ? nrhs : 1));   // gc
  add(sizeof(int) * ((size_t)N + 1));         // prog
  add(sizeof(int) * 64);                      // err
  add(sizeof(int) * ((size_t)N / b + 2) * (BMAX + 1));   // panel column counters
  return s + 4096;
}

int band_bind(BandWork& w, void* base, int N, int b, int nrhs) {
  if (b != 16 && b != 32) {
    set_error("band reduction: band width must be 16 or 32");
    return WTB_ERR_ARG;
  }
  w.N = N;
  w.b = b;
  w.nrhs = nrhs;
  w.LB = 2 * b + 1;
  w.nwg_max = (N + PR - 1) / PR + 1;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~size_t(255);
    return q;
  };
  const size_t nwg = w.nwg_max;
  w.Pw = (zc*)take(sizeof(zc) * (size_t)N * b);
  w.V = (zc*)take(sizeof(zc) * (size_t)N * b);
  w.Y = (zc*)take(sizeof(zc) * (size_t)N * b);
  w.X = (zc*)take(sizeof(zc) * (size_t)N * b);
  w.part = (zc*)take(sizeof(zc) * 2 * nwg * (BMAX + 1));
  w.rowc = (zc*)take(sizeof(zc) * 2 * BMAX);
  w.tau = (zc*)take(sizeof(zc) * BMAX);
  w.G = (zc*)take(sizeof(zc) * BMAX * BMAX);
  w.t1 = (zc*)take(sizeof(zc) * BMAX * 6);
  w.T = (zc*)take(sizeof(zc) * BMAX * BMAX);
  w.t2 = (zc*)take(sizeof(zc) * 6 * BMAX);
  w.M = (zc*)take(sizeof(zc) * BMAX * BMAX);
  w.gw = (zc*)take(sizeof(zc) * nwg * BMAX * BMAX);
  w.gw_elems = nwg * BMAX * BMAX;
  w.gp = (zc*)take(sizeof(zc) * GRAM_PMAX * BMAX * (BMAX + 16));
  w.hb = (zc*)take(sizeof(zc) * HB_SMAX * (size_t)N * b);
  w.BD = (zc*)take(sizeof(zc) * (size_t)N * w.LB);
  w.gc = (zc*)take(sizeof(zc) * (size_t)N * (nrhs > 0 ? nrhs : 1));
  w.prog = (int*)take(sizeof(int) * ((size_t)N + 1));
  w.err = (int*)take(sizeof(int) * 64);
  w.bar = (int*)take(sizeof(int) * ((size_t)N / b + 2) * (BMAX + 1));
  return WTB_OK;
}

// X = A22 Y (m x B) with enough workgroups to fill the chip: K slices until the grid
// holds ~1024 workgroups (at most HB_SMAX slices), partials summed in slice order
template <class T, int B>
static void hmm_band(hipStream_t s, BandWork& w, int m, const T* A, int ld, const zc* Y, zc* X) {
  const int rb = (m + HB_R - 1) / HB_R;
  int ns = std::max(1, std::min(HB_SMAX, (1024 + rb - 1) / rb));
  const int kc = ((m + ns - 1) / ns + 31) / 32 * 32;   // a multiple of either K step
  ns = (m + kc - 1) / kc;
  hipLaunchKernelGGL((k_hmm_b<T, B>), dim3(rb, ns), dim3(256), 0, s, m, A, ld, Y, w.N, kc, ns > 1 ? w.hb : nullptr,
                     X, w.N);
  if (ns > 1) {
    const size_t tot = (size_t)m * B;
    hipLaunchKernelGGL(k_hmm_sum<B>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, ns, w.hb, X, w.N);
  }
}

// out0 = A^H Y0 (B x B) and out1 = A^H Y1 (B x n1) over m rows, ld B
template <int B>
static void gram(hipStream_t s, BandWork& w, int m, const zc* A, const zc* Y0, int ldy0, const zc* Y1, int n1,
                 zc* out0, zc* out1) {
  const int nch = (m + GRAM_ROWS - 1) / GRAM_ROWS;
  const int P = min(nch, GRAM_PMAX);
  hipLaunchKernelGGL(k_gram<B>, dim3(P), dim3(256), 0, s, m, w.N, A, Y0, ldy0, Y1, n1, w.gp);
  const int nout = B * (B + 16);
  hipLaunchKernelGGL(k_gram_sum<B>, dim3((nout + 63) / 64), dim3(256), 0, s, P, w.gp, n1, out0, out1);
}

template <class T, int B>
static int stage1(hipStream_t s, BandWork& w, T* H, int ld, int nrhs) {
  const int N = w.N, b = B;
  for (int p = 0;; ++p) {
    const int c0 = p * b, r0 = c0 + b;
    if (r0 >= N) break;
    const int m = N - r0;
    const int k = min(b, m - 1);
    if (k <= 0) break;
    PanelArgs<T> a;
    a.A = H;
    a.ld = ld;
    a.N = N;
    a.b = b;
    a.c0 = c0;
    a.r0 = r0;
    a.m = m;
    a.k = k;
    a.Pw = w.Pw;
    a.part = w.part;
    a.rowc = w.rowc;
    a.tau = w.tau;
    a.nwg = (m + PR - 1) / PR;
    const dim3 g(a.nwg), blk(PR);
    if (w.panel_resident > 0 && a.nwg <= w.panel_resident) {
      hipLaunchKernelGGL((k_panel<T, B>), g, blk, 0, s, a, w.V, w.bar + (size_t)p * (BMAX + 1), w.err + 8,
                         (long long)1 << 20);
    } else {
      for (int c = 0; c <= k; ++c) hipLaunchKernelGGL((k_pqr_col<T, B>), g, blk, 0, s, a, c);
      hipLaunchKernelGGL((k_pfin<T, B>), g, blk, 0, s, a, w.V);
    }
    int rc;
    gram<B>(s, w, m, w.V, w.V, N, nrhs > 0 ? w.gc + r0 : nullptr, nrhs, w.G, w.t1);
    hipLaunchKernelGGL((k_pT<B>), dim3(1), dim3(256), 0, s, k, w.G, w.t1, w.tau, nrhs, w.T, w.t2);
    hipLaunchKernelGGL((k_pY<B>), g, blk, 0, s, m, N, r0, w.V, w.T, w.t2, nrhs, w.Y, w.gc);
    hmm_band<T, B>(s, w, m, H + r0 + (size_t)ld * r0, ld, w.Y, w.X);
    gram<B>(s, w, m, w.V, w.X, N, nullptr, 0, w.M, nullptr);
    hipLaunchKernelGGL((k_pW<B>), g, blk, 0, s, m, N, w.V, w.T, w.M, w.X);
    const int nt = (m + 63) / 64;
    const bool tev = w.kev && w.kev_used < w.kev_cap;
    if (tev) hipEventRecord(w.kev[2 * w.kev_used], s);
    hipLaunchKernelGGL((k_her2k_full<T, B>), dim3(nt * (nt + 1) / 2), dim3(256), 0, s, H, ld, m, r0, N, w.V,
                       w.X);
    if (tev) {
      hipEventRecord(w.kev[2 * w.kev_used++ + 1], s);
      w.kev_flops += 8.0 * 0.5 * (double)m * (double)(m + 1) * (2.0 * k);   // lower triangle, 2k products each
    }
  }
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

template <class T>
static int run_band_t(hipStream_t s, BandWork& w, T* H, int ld, zc* g, double* d, double* e, int* status) {
  const int N = w.N, b = w.b, nrhs = g ? w.nrhs : 0;
  if (N < 2) {
    set_error("band reduction: N < 2");
    return WTB_ERR_ARG;
  }
  if (nrhs > 6) return WTB_ERR_LIMIT;
  if (nrhs > 0)
    hipLaunchKernelGGL(k_conj_vec, dim3((unsigned)(((size_t)nrhs * N + 255) / 256)), dim3(256), 0, s, nrhs * N, g,
                       w.gc);
  w.kev_used = 0;
  w.kev_flops = 0.0;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, w.err, 64);
  {
    const int nb = (N / b + 2) * (BMAX + 1);
    hipLaunchKernelGGL(k_zero_ints, dim3((nb + 255) / 256), dim3(256), 0, s, w.bar, nb);
  }
  int rc = (b == 16) ? stage1<T, 16>(s, w, H, ld, nrhs) : stage1<T, 32>(s, w, H, ld, nrhs);
  if (rc) return rc;
  hipLaunchKernelGGL(k_band_extract<T>, dim3(N), dim3(64), 0, s, H, ld, N, b, w.LB, w.BD);
  hipLaunchKernelGGL(k_zero_ints, dim3((N + 64 + 255) / 256), dim3(256), 0, s, w.prog, N + 1);
  hipLaunchKernelGGL(k_chase_init, dim3(1), dim3(64), 0, s, w.err);
  ChaseArgs ca;
  ca.BD = w.BD;
  ca.N = N;
  ca.b = b;
  ca.LB = w.LB;
  ca.gc = w.gc;
  ca.nrhs = nrhs;
  ca.prog = w.prog;
  ca.ctl = w.err;
  ca.spin_limit = 1ll << 20;   // polls of about a microsecond each: about a second
  const int P = min(N - 1, 128);   // workgroups claiming sweeps
  if (b == 16) hipLaunchKernelGGL(k_chase<16>, dim3(P), dim3(256), 0, s, ca);
  else hipLaunchKernelGGL(k_chase<32>, dim3(P), dim3(256), 0, s, ca);
  hipLaunchKernelGGL(k_band_finish, dim3((N + 255) / 256), dim3(256), 0, s, w.BD, N, w.LB, d, e, w.gc, g, nrhs,
                     w.err, status);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int run_band_tridiag(hipStream_t s, BandWork& w, zc* H, int ld, zc* g, double* d, double* e, int* status) {
  return run_band_t<zc>(s, w, H, ld, g, d, e, status);
}
int run_band_tridiag_f32(hipStream_t s, BandWork& w, zf* H, int ld, zc* g, double* d, double* e, int* status) {
  return run_band_t<zf>(s, w, H, ld, g, d, e, status);
}

}  // namespace wtb
