// Stage (d1) in two stages: a Hermitian matrix (full storage, column-major, ld) is
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
//   zgemm                      G = V^H V, t1 = V^H g (MFMA, K split over workgroups)
//   k_pT                       T (ZLARFT forward/columnwise) and t2 = T^H t1
//   k_pY                       Y = V T, g -= V t2 (rows of the panel)
//   hmm                        X = A22 Y (MFMA, full Hermitian storage)
//   zgemm                      M = V^H X
//   k_pW                       W = X - 1/2 V (T^H M)
//   k_her2k_full               A22 -= V W^H + W V^H (MFMA, every tile of the trailing matrix)
// The tracked vectors are carried conjugated (gc = conj(g)): Q^H gc = conj(Q^T g).
//
// Stage 2 (bulge chasing, FP64, band storage BD[c][d] = A(c + d, c), d <= 2b): sweep s
// annihilates column s below its subdiagonal; its tasks k = 0, 1, ... each apply one
// reflector (ZLARFG) two-sidedly to a diagonal block and the previous one from the right
// to the block below it (creating the bulge whose first column the next reflector
// annihilates).  One persistent kernel: workgroup w runs the sweeps w, w + P, ...; task k
// of sweep s waits until sweep s - 1 has finished task k + 2 (their blocks then share no
// row), signalled through a per-sweep progress counter with agent-scope release/acquire.
// Every wait is bounded: past the limit the kernel sets an error flag that every waiting
// workgroup sees, and all of them return.
#include "wtb_kernels.h"

namespace wtb {

#define PR 256          // panel rows per workgroup (one per thread)
#define BMAX 32         // largest band width

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__device__ inline zc zdiv(zc a, zc b) {
  // a / b with the scaling of Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = b.re + b.im * r;
    return mkz((a.re + a.im * r) / den, (a.im - a.re * r) / den);
  }
  const double r = b.re / b.im, den = b.im + b.re * r;
  return mkz((a.re * r + a.im) / den, (a.im * r - a.re) / den);
}

// ZLARFG on (alpha, x) with |x|^2 = xn2: beta (real, returned as the new alpha), tau and
// the scale 1/(alpha - beta) of x.  tau = 0 (H = I) when x = 0 and alpha is real.
__device__ inline void larfg(zc alpha, double xn2, zc& beta, zc& tau, zc& scale) {
  if (xn2 == 0.0 && alpha.im == 0.0) {
    beta = alpha;
    tau = mkz(0, 0);
    scale = mkz(0, 0);
    return;
  }
  const double nrm = sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2);
  const double b = alpha.re >= 0.0 ? -nrm : nrm;
  tau = mkz((b - alpha.re) / b, -alpha.im / b);
  scale = zdiv(mkz(1, 0), mkz(alpha.re - b, alpha.im));
  beta = mkz(b, 0.0);
}

// sum of a zc[NV] over the 256 threads of a workgroup into out[NV] (thread 0..NV-1 hold
// the results in red[]); fixed order: waves by butterfly, then waves in index order
template <int NV>
__device__ inline void block_sum(zc (&v)[NV], zc* red /* [4][NV] in LDS */) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    zc x = v[q];
    for (int off = 32; off > 0; off >>= 1) {
      x.re += __shfl_xor(x.re, off);
      x.im += __shfl_xor(x.im, off);
    }
    if (lane == 0) red[wave * NV + q] = x;
  }
  __syncthreads();
  if ((int)threadIdx.x < NV) {
    const int q = threadIdx.x;
    red[4 * NV + q] = ((red[q] + red[NV + q]) + red[2 * NV + q]) + red[3 * NV + q];
  }
  __syncthreads();
}

// Sums over the 256 threads of a workgroup of one real value x0 and NB complex values
// v[0..NB) (NB a power of two <= 32).  Within a wave the NB sums are formed by a butterfly
// that halves the set each lane keeps (NB - 1 exchanges instead of 6 NB), after which lane
// l holds the wave's sum of value (l >> (6 - log2 NB)) & (NB - 1); the four waves are then
// added in index order.  Results: red[4 * (NB + 1) + q] (q = 0: x0, 1 + j: v[j]).
// one halving step of the butterfly: lanes with the `off` bit set keep values [n, 2n)
template <int n, int NB>
__device__ inline void bf_step(zc (&v)[NB], int lane, int off) {
  const bool up = (lane & off) != 0;
#pragma unroll
  for (int q = 0; q < n; ++q) {
    const zc send = up ? v[q] : v[q + n];
    zc recv;
    recv.re = __shfl_xor(send.re, off);
    recv.im = __shfl_xor(send.im, off);
    v[q] = (up ? v[q + n] : v[q]) + recv;
  }
}

template <int NB>
__device__ inline void block_sum_bf(double x0, zc (&v)[NB], zc* red /* [5][NB + 1] in LDS */) {
  static_assert(NB == 16 || NB == 32, "butterfly for 16 or 32 values");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if constexpr (NB == 32) {
    bf_step<16>(v, lane, 32);
    bf_step<8>(v, lane, 16);
    bf_step<4>(v, lane, 8);
    bf_step<2>(v, lane, 4);
    bf_step<1>(v, lane, 2);
  } else {
    bf_step<8>(v, lane, 32);
    bf_step<4>(v, lane, 16);
    bf_step<2>(v, lane, 8);
    bf_step<1>(v, lane, 4);
  }
  zc r = v[0];
#pragma unroll
  for (int off = 32 / NB; off > 0; off >>= 1) {
    r.re += __shfl_xor(r.re, off);
    r.im += __shfl_xor(r.im, off);
  }
  for (int off = 32; off > 0; off >>= 1) x0 += __shfl_xor(x0, off);
  constexpr int SH = (NB == 32) ? 1 : (NB == 16) ? 2 : (NB == 8) ? 3 : (NB == 4) ? 4 : (NB == 2) ? 5 : 6;
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
    if ((int)threadIdx.x < B + 1) {
      const int q = threadIdx.x;
      zc s = mkz(0, 0);
      for (int w = 0; w < a.nwg; ++w) s += pp[(size_t)w * (BMAX + 1) + q];
      sh[q] = s;                       // sh[0].re = |x|^2 below the diagonal, sh[1 + j] = x^H P(:, j)
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

// one workgroup: T (ZLARFT, forward, columnwise) from G = V^H V and tau, and t2 = T^H t1
// (t1 = V^H gc); G and t1 column-major with ld B, T stored [col][row]
template <int B>
__global__ __launch_bounds__(256) void k_pT(int k, const zc* __restrict__ G, const zc* __restrict__ t1,
                                            const zc* __restrict__ tau, int nrhs, zc* __restrict__ Tm,
                                            zc* __restrict__ t2) {
  __shared__ zc Gs[B * B];
  __shared__ zc Ts[B * B];
  for (int q = threadIdx.x; q < B * B; q += 256) {
    Gs[q] = G[q];                        // G(r, c) at [c][r]
    Ts[q] = mkz(0, 0);
  }
  __syncthreads();
  // T(0:c, c) = -tau_c T(0:c, 0:c) G(0:c, c), T(c, c) = tau_c
  for (int c = 0; c < k; ++c) {
    const zc tc = tau[c];
    zc val = mkz(0, 0);
    if ((int)threadIdx.x < c) {
      const int r = threadIdx.x;
      zc sacc = mkz(0, 0);
      for (int q = r; q < c; ++q) sacc += Ts[q * B + r] * Gs[c * B + q];
      val = mkz(0, 0) - tc * sacc;
    }
    __syncthreads();
    if ((int)threadIdx.x < c) Ts[c * B + threadIdx.x] = val;
    if (threadIdx.x == 0) Ts[c * B + c] = tc;
    __syncthreads();
  }
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
  for (int q = threadIdx.x; q < B * B; q += PR) {
    const int p = q / B, c = q % B;      // Z(p, c) = sum_r conj(T(r, p)) M(r, c), T(r, p) = Tm[p][r], r <= p
    zc sacc = mkz(0, 0);
    for (int r = 0; r <= p; ++r) sacc = fma_conj(Tm[p * B + r], M[(size_t)c * B + r], sacc);
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

// A22 -= U, U = V W^H + W V^H, on every 64 x 64 tile of the trailing matrix (rows and
// columns r0..N-1), K = 2 b: FP64 MFMA on the COMPLEX*16 matrix, FP32 MFMA (FP32 operands
// and accumulation) on the COMPLEX*8 one.  The MFMA rows run over the tile's columns and
// its columns over the tile's rows, D(i, j) = U(c_i, r_j) = conj(U(r_j, c_i)) since U is
// Hermitian, so that a lane's result lands next to its neighbours' in the column-major
// matrix.
template <class T>
__global__ __launch_bounds__(256) void k_her2k_full(T* __restrict__ A, int ld, int m, int r0, int N, int b,
                                                    const zc* __restrict__ V, const zc* __restrict__ W) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ci0 = blockIdx.y * 64 + (wave >> 1) * 32;   // tile columns (MFMA rows)
  const int rj0 = blockIdx.x * 64 + (wave & 1) * 32;    // tile rows (MFMA columns)
  const int lr = lane & 15, lk = lane >> 4;
  const int K = 2 * b;
  typedef typename std::conditional<sizeof(T) == sizeof(zc), d4, f4>::type acc_t;
  acc_t cr[2][2], ci[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) cr[x][y] = ci[x][y] = acc_t{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + lk;
    zc av[2], bv[2];
    for (int x = 0; x < 2; ++x) {
      const int c = ci0 + x * 16 + lr, r = rj0 + x * 16 + lr;
      zc va = mkz(0, 0), vb = mkz(0, 0);
      if (k < K) {
        // A operand (i = column c, k): [V | W](c, k); B operand (k, j = row r): conj([W | V](r, k))
        if (c < m) va = (k < b) ? V[(size_t)k * N + c] : W[(size_t)(k - b) * N + c];
        if (r < m) vb = (k < b) ? W[(size_t)k * N + r] : V[(size_t)(k - b) * N + r];
      }
      av[x] = va;
      bv[x] = vb;
    }
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) {
        // a * conj(b)
        if constexpr (sizeof(T) == sizeof(zc)) {
          cr[x][y] = mfma16x16x4(av[x].re, bv[y].re, cr[x][y]);
          cr[x][y] = mfma16x16x4(av[x].im, bv[y].im, cr[x][y]);
          ci[x][y] = mfma16x16x4(av[x].im, bv[y].re, ci[x][y]);
          ci[x][y] = mfma16x16x4(-av[x].re, bv[y].im, ci[x][y]);
        } else {
          const float ar = (float)av[x].re, ai = (float)av[x].im, br = (float)bv[y].re, bi = (float)bv[y].im;
          cr[x][y] = mfma16x16x4f32(ar, br, cr[x][y]);
          cr[x][y] = mfma16x16x4f32(ai, bi, cr[x][y]);
          ci[x][y] = mfma16x16x4f32(ai, br, ci[x][y]);
          ci[x][y] = mfma16x16x4f32(-ar, bi, ci[x][y]);
        }
      }
  }
  constexpr bool F64 = sizeof(T) == sizeof(zc);
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      for (int q = 0; q < 4; ++q) {
        // f64 MFMA: register q holds D row lk + 4q; f32 MFMA: row 4 lk + q
        const int c = ci0 + x * 16 + (F64 ? lk + 4 * q : 4 * lk + q);
        const int r = rj0 + y * 16 + lr;
        if (r < m && c < m) {
          T& h = A[(size_t)(r0 + r) + (size_t)ld * (r0 + c)];
          zc v = ldz(h);
          v.re -= cr[x][y][q];
          v.im += ci[x][y][q];       // - conj(D)
          stz(h, v);
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
  int* err;          // set when a wait passes its bound
  int P;             // workgroups
  long long spin_limit;
};

__device__ inline int chase_ntasks(int N, int b, int s) { return 1 + (N - 2 - s) / b; }

// element (r, c) of the band, r >= c
__device__ inline zc& bd(const ChaseArgs& a, int r, int c) { return a.BD[(size_t)c * a.LB + (r - c)]; }

template <int B>
__global__ __launch_bounds__(256) void k_chase(ChaseArgs a) {
  __shared__ zc O[B][B + 1];      // off-diagonal block (rows st.., columns of the previous block)
  __shared__ zc D[B][B + 1];      // diagonal block, full Hermitian
  __shared__ zc vprev[B], vcur[B], wv[B];
  __shared__ zc sc[4];            // tau_prev, tau, beta, misc
  __shared__ int stop;
  const int t = threadIdx.x;
  const int N = a.N, b = a.b;
  for (int s = blockIdx.x; s <= N - 2; s += a.P) {
    const int nt = chase_ntasks(N, b, s);
    int lenp = 0;                 // length of the previous reflector
    for (int k = 0; k < nt; ++k) {
      // wait for sweep s - 1 to have finished task k + 2 (or all of its tasks)
      if (s > 0) {
        if (t == 0) {
          const int need = min(k + 3, chase_ntasks(N, b, s - 1));
          long long n = 0;
          int bad = 0;
          while (__hip_atomic_load(&a.prog[s - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (__hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || ++n > a.spin_limit) {
              bad = 1;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          if (bad) __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // the data sweep s - 1 wrote
          stop = bad;
        }
        __syncthreads();
        if (stop) return;
      }
      if (k == 0) {
        // column s, rows s+1..ed: reflector H0; D = A(st..ed, st..ed)
        const int st = s + 1, ed = min(s + b, N - 1), len = ed - st + 1;
        if (t < B) vcur[t] = (t < len) ? bd(a, st + t, s) : mkz(0, 0);
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          zc v = mkz(0, 0);
          if (r < len && c < len) v = r >= c ? bd(a, st + r, st + c) : conj(bd(a, st + c, st + r));
          D[r][c] = v;
        }
        __syncthreads();
        if (t == 0) {
          double xn2 = 0.0;
          for (int r = 1; r < len; ++r) xn2 += norm2(vcur[r]);
          zc beta, tau, scale;
          larfg(vcur[0], xn2, beta, tau, scale);
          sc[1] = tau;
          sc[2] = beta;
          sc[3] = scale;
        }
        __syncthreads();
        const zc tau = sc[1], scale = sc[3];
        if (t < B) {
          zc x = mkz(0, 0);
          if (t == 0) x = mkz(1, 0);
          else if (t < len) x = vcur[t] * scale;
          vcur[t] = x;
        }
        __syncthreads();
        // column s: beta then zeros
        if (t < len) bd(a, st + t, s) = t == 0 ? sc[2] : mkz(0, 0);
        lenp = len;
        // two-sided update of D with vcur
        if (t < B) {
          zc y = mkz(0, 0);
          for (int c = 0; c < len; ++c) y = fma_z(D[t][c], vcur[c], y);
          wv[t] = y;                                       // y = D v
        }
        __syncthreads();
        if (t == 0) {
          double g = 0.0;                                  // v^H y (real)
          for (int r = 0; r < len; ++r) g += vcur[r].re * wv[r].re + vcur[r].im * wv[r].im;
          sc[0] = mkz(g, 0);
        }
        __syncthreads();
        if (t < B) {                                       // w = tau y - 1/2 |tau|^2 (v^H y) v
          const double h = 0.5 * norm2(tau) * sc[0].re;
          wv[t] = tau * wv[t] - h * vcur[t];
        }
        __syncthreads();
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          if (r < len && c <= r) {
            const zc d = D[r][c] - vcur[r] * conj(wv[c]) - wv[r] * conj(vcur[c]);
            bd(a, st + r, st + c) = (r == c) ? mkz(d.re, 0.0) : d;
          }
        }
        // tracked vectors rows st..ed: gc -= conj(tau) v (v^H gc)
        if (t < a.nrhs) {
          zc* g = a.gc + (size_t)t * N + st;
          zc al = mkz(0, 0);
          for (int r = 0; r < len; ++r) al = fma_conj(vcur[r], g[r], al);
          const zc f = conj(tau) * al;
          for (int r = 0; r < len; ++r) g[r] -= vcur[r] * f;
        }
        if (t < B) vprev[t] = vcur[t];
        if (t == 0) sc[0] = tau;
      } else {
        // block O = A(st..ed, sp..ep), sp..ep the previous reflector's rows
        const int sp = s + 1 + (k - 1) * b, ep = sp + lenp - 1;
        const int st = ep + 1, ed = min(st + b - 1, N - 1), len = ed - st + 1;
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          O[r][c] = (r < len && c < lenp) ? bd(a, st + r, sp + c) : mkz(0, 0);
          zc v = mkz(0, 0);
          if (r < len && c < len) v = r >= c ? bd(a, st + r, st + c) : conj(bd(a, st + c, st + r));
          D[r][c] = v;
        }
        __syncthreads();
        // (i) O <- O Hprev: w = O vprev, O -= tau_prev w vprev^H
        const zc taup = sc[0];
        if (t < B) {
          zc w = mkz(0, 0);
          for (int c = 0; c < lenp; ++c) w = fma_z(O[t][c], vprev[c], w);
          wv[t] = w;
        }
        __syncthreads();
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          if (r < len && c < lenp) O[r][c] -= (taup * wv[r]) * conj(vprev[c]);
        }
        __syncthreads();
        // (ii) reflector from O(:, 0)
        if (t == 0) {
          double xn2 = 0.0;
          for (int r = 1; r < len; ++r) xn2 += norm2(O[r][0]);
          zc beta, tau, scale;
          larfg(O[0][0], xn2, beta, tau, scale);
          sc[1] = tau;
          sc[2] = beta;
          sc[3] = scale;
        }
        __syncthreads();
        const zc tau = sc[1], scale = sc[3];
        if (t < B) {
          zc x = mkz(0, 0);
          if (t == 0) x = mkz(1, 0);
          else if (t < len) x = O[t][0] * scale;
          vcur[t] = x;
        }
        __syncthreads();
        if (t < B) O[t][0] = (t == 0) ? sc[2] : mkz(0, 0);
        // (iii) O(:, 1:) <- H^H O(:, 1:): w'_c = v^H O(:, c), O -= conj(tau) v w'
        if (t < B) {
          zc w = mkz(0, 0);
          if (t >= 1 && t < lenp)
            for (int r = 0; r < len; ++r) w = fma_conj(vcur[r], O[r][t], w);
          wv[t] = w;
        }
        __syncthreads();
        const zc ct = conj(tau);
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          if (r < len && c >= 1 && c < lenp) O[r][c] -= (ct * vcur[r]) * wv[c];
        }
        __syncthreads();
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          if (r < len && c < lenp) bd(a, st + r, sp + c) = O[r][c];
        }
        // (iv) D <- H^H D H
        if (t < B) {
          zc y = mkz(0, 0);
          for (int c = 0; c < len; ++c) y = fma_z(D[t][c], vcur[c], y);
          wv[t] = y;
        }
        __syncthreads();
        if (t == 0) {
          double g = 0.0;
          for (int r = 0; r < len; ++r) g += vcur[r].re * wv[r].re + vcur[r].im * wv[r].im;
          sc[0] = mkz(g, 0);
        }
        __syncthreads();
        if (t < B) {
          const double h = 0.5 * norm2(tau) * sc[0].re;
          wv[t] = tau * wv[t] - h * vcur[t];
        }
        __syncthreads();
        for (int e = t; e < B * B; e += 256) {
          const int r = e % B, c = e / B;
          if (r < len && c <= r) {
            const zc d = D[r][c] - vcur[r] * conj(wv[c]) - wv[r] * conj(vcur[c]);
            bd(a, st + r, st + c) = (r == c) ? mkz(d.re, 0.0) : d;
          }
        }
        if (t < a.nrhs) {
          zc* g = a.gc + (size_t)t * N + st;
          zc al = mkz(0, 0);
          for (int r = 0; r < len; ++r) al = fma_conj(vcur[r], g[r], al);
          const zc f = conj(tau) * al;
          for (int r = 0; r < len; ++r) g[r] -= vcur[r] * f;
        }
        lenp = len;
        if (t < B) vprev[t] = vcur[t];
        if (t == 0) sc[0] = tau;
      }
      // publish: every thread's stores are done before the counter moves
      __syncthreads();
      if (t == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(&a.prog[s], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
    }
  }
}

// d, e of the tridiagonal from the band; g = conj(gc)
__global__ void k_band_finish(const zc* __restrict__ BD, int N, int LB, double* __restrict__ d, double* __restrict__ e,
                              const zc* __restrict__ gc, zc* __restrict__ g, int nrhs, const int* __restrict__ err,
                              int* __restrict__ status) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j == 0 && status && err[0]) status[0] |= 1 << WTB_ERR_EIG;
  if (j >= N) return;
  d[j] = BD[(size_t)j * LB].re;
  e[j] = (j + 1 < N) ? BD[(size_t)j * LB + 1].re : 0.0;
  for (int r = 0; r < nrhs; ++r) g[(size_t)r * N + j] = conj(gc[(size_t)r * N + j]);
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
  add(hmm_work_bytes(N, b));                  // hmm K slices
  add(sizeof(zc) * (size_t)N * LB);           // band
  add(sizeof(zc) * (size_t)N * (nrhs > 0 ? nrhs : 1));   // gc
  add(sizeof(int) * ((size_t)N + 1));         // prog
  add(sizeof(int) * 64);                      // err
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
  w.hw = (zc*)take(hmm_work_bytes(N, b));
  w.BD = (zc*)take(sizeof(zc) * (size_t)N * w.LB);
  w.gc = (zc*)take(sizeof(zc) * (size_t)N * (nrhs > 0 ? nrhs : 1));
  w.prog = (int*)take(sizeof(int) * ((size_t)N + 1));
  w.err = (int*)take(sizeof(int) * 64);
  return WTB_OK;
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
    for (int c = 0; c <= k; ++c) hipLaunchKernelGGL((k_pqr_col<T, B>), g, blk, 0, s, a, c);
    hipLaunchKernelGGL((k_pfin<T, B>), g, blk, 0, s, a, w.V);
    const zc one = mkz(1, 0), zero = mkz(0, 0);
    int rc;
    if ((rc = zgemm(s, 'C', 'N', b, b, m, one, w.V, N, w.V, N, zero, w.G, b, w.gw, w.gw_elems))) return rc;
    if (nrhs > 0 &&
        (rc = zgemm(s, 'C', 'N', b, nrhs, m, one, w.V, N, w.gc + r0, N, zero, w.t1, b, w.gw, w.gw_elems)))
      return rc;
    hipLaunchKernelGGL((k_pT<B>), dim3(1), dim3(256), 0, s, k, w.G, w.t1, w.tau, nrhs, w.T, w.t2);
    hipLaunchKernelGGL((k_pY<B>), g, blk, 0, s, m, N, r0, w.V, w.T, w.t2, nrhs, w.Y, w.gc);
    if constexpr (sizeof(T) == sizeof(zc)) {
      rc = hmm_f64(s, m, b, reinterpret_cast<const zc*>(H) + r0 + (size_t)ld * r0, ld, w.Y, N, w.X, N, w.hw);
    } else {
      rc = hmm_f32(s, m, b, reinterpret_cast<const zf*>(H) + r0 + (size_t)ld * r0, ld, w.Y, N, w.X, N, w.hw);
    }
    if (rc) return rc;
    if ((rc = zgemm(s, 'C', 'N', b, b, m, one, w.V, N, w.X, N, zero, w.M, b, w.gw, w.gw_elems))) return rc;
    hipLaunchKernelGGL((k_pW<B>), g, blk, 0, s, m, N, w.V, w.T, w.M, w.X);
    const int nt = (m + 63) / 64;
    const bool tev = w.kev && w.kev_used < w.kev_cap;
    if (tev) hipEventRecord(w.kev[2 * w.kev_used], s);
    hipLaunchKernelGGL((k_her2k_full<T>), dim3(nt, nt), dim3(256), 0, s, H, ld, m, r0, N, b, w.V, w.X);
    if (tev) {
      hipEventRecord(w.kev[2 * w.kev_used++ + 1], s);
      w.kev_flops += 8.0 * (double)m * (double)m * (2.0 * k);   // m^2 entries, 2k complex products each
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
  int rc = (b == 16) ? stage1<T, 16>(s, w, H, ld, nrhs) : stage1<T, 32>(s, w, H, ld, nrhs);
  if (rc) return rc;
  hipLaunchKernelGGL(k_band_extract<T>, dim3(N), dim3(64), 0, s, H, ld, N, b, w.LB, w.BD);
  hipLaunchKernelGGL(k_zero_ints, dim3((N + 64 + 255) / 256), dim3(256), 0, s, w.prog, N + 1);
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, s, w.err, 64);
  ChaseArgs ca;
  ca.BD = w.BD;
  ca.N = N;
  ca.b = b;
  ca.LB = w.LB;
  ca.gc = w.gc;
  ca.nrhs = nrhs;
  ca.prog = w.prog;
  ca.err = w.err;
  ca.P = min(N - 1, 256);
  ca.spin_limit = 1ll << 20;   // polls of about a microsecond each: about a second
  if (b == 16) hipLaunchKernelGGL(k_chase<16>, dim3(ca.P), dim3(256), 0, s, ca);
  else hipLaunchKernelGGL(k_chase<32>, dim3(ca.P), dim3(256), 0, s, ca);
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
