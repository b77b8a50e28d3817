// Stage (d3): eigenvectors of the Hermitian matrix from those of its tridiagonal,
// X = Q Z, with Q = H(0) H(1) ... H(N-2) the product of the reduction's
// reflectors H(j) = I - tau_j v_j v_j^H (v_j(j+1) = 1, v_j(r) for r > j+1 kept in
// the lower triangle of the reduced matrix, column j; the layout of ZHETRD with
// UPLO = 'L').  Panels of nb reflectors are applied last to first in compact WY
// form, I - V T V^H, with T upper triangular as ZLARFT ('F', 'C') builds it:
//   T(i,i) = tau_i,  T(0:i-1, i) = -tau_i T(0:i-1, 0:i-1) V(:, 0:i-1)^H v_i.
// Per panel: Y = V^H X (row chunks on FP64 MFMA, partial sums), Y <- T Y, and
// X -= V Y (FP64 MFMA), so the work is two complex GEMMs of N' x nb x m.
#include "wtb_kernels.h"

namespace wtb {

// reflectors per block: consecutive reflectors in any grouping form a block I - V T V^H,
// so the blocks need not match the reduction's panels; 64 (two 32-column panels) halves
// the launches, whose fixed costs dominate each block's time
#define BT_NB 64
constexpr int kBtVR = 16;                                    // T factor: V rows per LDS slice
constexpr int kBtTyCols = 256 / BT_NB;                       // Y = T Y: columns per workgroup
constexpr int kBtPairs = (BT_NB * (BT_NB - 1) / 2 + 255) / 256;   // Gram pairs per thread
constexpr int kBtSteps = 4;   // V^H X: MFMA steps (4 rows each) per group of loads

template <class T>
__device__ inline zc vref(const T* H, int ld, int r, int j) {
  // v_j(r): 0 above row j+1, 1 at row j+1, the stored entry below
  if (r <= j) return mkz(0, 0);
  if (r == j + 1) return mkz(1, 0);
  return ldz(H[(size_t)r + (size_t)ld * j]);
}

// Gram matrix of block blockIdx.x (reflectors p .. p+kp-1), G(i', i) = v_{p+i'}^H v_{p+i}
// for i' < i, in kBtSplit row ranges (blockIdx.y), each summed by its own workgroup into
// Gp; pair t of the strictly upper triangle (row i holds pairs i(i-1)/2 .. i(i+1)/2 - 1)
// is thread t % 256, slot t / 256
constexpr int kBtSplit = 8;
constexpr int kBtPairSlots = 256 * kBtPairs;
__device__ inline void bt_pair(int t, int& pi, int& pj) {
  int i = 1;
  while (t >= i && i < BT_NB) {
    t -= i;
    ++i;
  }
  pi = t;
  pj = i;
}
template <class T>
__global__ __launch_bounds__(256) void k_bt_gram(const T* __restrict__ H, int ld, int N, zc* __restrict__ Gp) {
  __shared__ zc Vs[kBtVR][BT_NB + 1];
  const int p = blockIdx.x * BT_NB;
  const int kp = min(BT_NB, N - 1 - p);
  const int tid = threadIdx.x;
  const int npair = kp * (kp - 1) / 2;
  const int rtot = N - 1 - p;
  const int per = ((rtot + kBtSplit - 1) / kBtSplit + kBtVR - 1) / kBtVR * kBtVR;
  const int rlo = p + 1 + (int)blockIdx.y * per, rhi = min(N, rlo + per);
  zc acc[kBtPairs];
  int pi[kBtPairs], pj[kBtPairs];
  for (int s = 0; s < kBtPairs; ++s) {
    acc[s] = mkz(0, 0);
    bt_pair(tid + 256 * s, pi[s], pj[s]);
  }
  for (int r0 = rlo; r0 < rhi; r0 += kBtVR) {
    for (int e = tid; e < kBtVR * BT_NB; e += 256) {
      const int rl = e % kBtVR, c = e / kBtVR;
      const int r = r0 + rl;
      Vs[rl][c] = (r < rhi && c < kp) ? vref(H, ld, r, p + c) : mkz(0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kBtPairs; ++s) {
      if (tid + 256 * s >= npair) continue;
      for (int rl = 0; rl < kBtVR; ++rl) acc[s] = fma_conj(Vs[rl][pi[s]], Vs[rl][pj[s]], acc[s]);
    }
    __syncthreads();
  }
  zc* out = Gp + ((size_t)blockIdx.x * kBtSplit + blockIdx.y) * kBtPairSlots;
  for (int s = 0; s < kBtPairs; ++s)
    if (tid + 256 * s < npair) out[tid + 256 * s] = acc[s];
}

// T factor of block blockIdx.x from its Gram partials (summed in range order), one
// workgroup per block
__global__ __launch_bounds__(256) void k_bt_tfactor(int N, const zc* __restrict__ tau, const zc* __restrict__ Gp,
                                                    zc* __restrict__ Tf) {
  __shared__ zc G[BT_NB][BT_NB + 1];
  __shared__ zc Ts[BT_NB][BT_NB + 1];
  const int p = blockIdx.x * BT_NB;
  const int kp = min(BT_NB, N - 1 - p);
  const int tid = threadIdx.x;
  const int npair = kp * (kp - 1) / 2;
  const zc* gp = Gp + (size_t)blockIdx.x * kBtSplit * kBtPairSlots;
  for (int s = 0; s < kBtPairs; ++s) {
    const int t = tid + 256 * s;
    if (t >= npair) continue;
    zc g = mkz(0, 0);
    for (int q = 0; q < kBtSplit; ++q) g += gp[(size_t)q * kBtPairSlots + t];
    int pi, pj;
    bt_pair(t, pi, pj);
    G[pi][pj] = g;
  }
  __syncthreads();
  // T column by column; thread r < i computes T(r, i) = -tau_i sum_{k=r}^{i-1} T(r,k) G(k,i)
  for (int i = 0; i < kp; ++i) {
    const zc ti = tau[p + i];
    if (tid < i) {
      zc sacc = mkz(0, 0);
      for (int k = tid; k < i; ++k) sacc = fma_z(Ts[tid][k], G[k][i], sacc);
      Ts[tid][i] = mkz(0, 0) - ti * sacc;
    } else if (tid == i) {
      Ts[i][i] = ti;
    }
    __syncthreads();
  }
  zc* out = Tf + (size_t)blockIdx.x * BT_NB * BT_NB;
  for (int e = tid; e < BT_NB * BT_NB; e += 256) {
    const int r = e % BT_NB, c = e / BT_NB;
    out[e] = (r < kp && c < kp && r <= c) ? Ts[r][c] : mkz(0, 0);
  }
}

// Ypart[chunk][c][col] = sum over the chunk's rows of conj(V(r, c)) X(r, col).
// Workgroup: 64 reflectors x 64 columns; wave w: reflectors 16w .. 16w + 15, all 64 columns.
template <class T>
__global__ __launch_bounds__(256) void k_bt_vhx(const T* __restrict__ H, int ld, int N, int p, int kp,
                                                const zc* __restrict__ X, int ldx, int m, int chunk,
                                                zc* __restrict__ Ypart) {
  const int col0 = blockIdx.x * 64;
  const int rbeg = p + 1 + blockIdx.y * chunk;
  const int rend = min(N, rbeg + chunk);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  static_assert(BT_NB == 64, "one reflector block of 16 per wave");
  constexpr int YC = 4;                      // column blocks of 16 per wave
  const int cb = wave * 16;                  // reflector block
  const int nb0 = col0;                      // column block
  d4 yr[YC], yi[YC];
  for (int y = 0; y < YC; ++y) {
    yr[y] = d4{0, 0, 0, 0};
    yi[y] = d4{0, 0, 0, 0};
  }
  // rows in groups of kBtSteps MFMA steps: the group's loads are all issued before its
  // products (each accumulator still takes its steps in ascending row order)
  const int c = cb + lr;
  for (int g0 = rbeg; g0 < rend; g0 += 4 * kBtSteps) {
    zc v[kBtSteps], xv[kBtSteps][YC];
#pragma unroll
    for (int t = 0; t < kBtSteps; ++t) {
      const int r = g0 + 4 * t + lk;
      v[t] = (r < rend && c < kp) ? vref(H, ld, r, p + c) : mkz(0, 0);
#pragma unroll
      for (int y = 0; y < YC; ++y) {
        const int col = nb0 + y * 16 + lr;
        xv[t][y] = (r < rend && col < m) ? X[(size_t)r + (size_t)ldx * col] : mkz(0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < kBtSteps; ++t) {
      if (g0 + 4 * t >= rend) break;               // the same steps as one step per 4 rows
#pragma unroll
      for (int y = 0; y < YC; ++y) {
        // conj(v) x = (vr xr + vi xi) + i (vr xi - vi xr)
        yr[y] = mfma16x16x4(v[t].re, xv[t][y].re, yr[y]);
        yr[y] = mfma16x16x4(v[t].im, xv[t][y].im, yr[y]);
        yi[y] = mfma16x16x4(v[t].re, xv[t][y].im, yi[y]);
        yi[y] = mfma16x16x4(-v[t].im, xv[t][y].re, yi[y]);
      }
    }
  }
  zc* out = Ypart + (size_t)blockIdx.y * BT_NB * m;
  for (int y = 0; y < YC; ++y) {
    const int col = nb0 + y * 16 + lr;
    if (col >= m) continue;
    for (int q = 0; q < 4; ++q) {
      const int c = cb + lk + 4 * q;           // f64 MFMA: register q holds row lk + 4q
      out[(size_t)c * m + col] = mkz(yr[y][q], yi[y][q]);
    }
  }
}

// Y(:, col) = T sum_chunks Ypart(:, col) for kBtTyCols columns per workgroup: thread (c, col)
// adds its chunk sums in order into LDS, then thread (r, col) applies row r of T
__global__ __launch_bounds__(256) void k_bt_ty(const zc* __restrict__ Ypart, int nchunk, int m, int kp,
                                               const zc* __restrict__ Tf, zc* __restrict__ Y) {
  static_assert(BT_NB * kBtTyCols == 256, "one (row, column) pair per thread");
  __shared__ zc Ts[BT_NB * BT_NB];
  __shared__ zc Ss[BT_NB][kBtTyCols];
  const int tid = threadIdx.x;
  for (int e = tid; e < BT_NB * BT_NB; e += 256) Ts[e] = Tf[e];
  const int cl = tid % kBtTyCols, col = blockIdx.x * kBtTyCols + cl;
  const int c = tid / kBtTyCols;                       // row of the block
  zc acc = mkz(0, 0);
  if (c < kp && col < m) acc = strided_sum(Ypart + (size_t)c * m + col, (size_t)BT_NB * m, 0, nchunk, 1, acc);
  Ss[c][cl] = acc;
  __syncthreads();
  if (col >= m) return;
  const int r = c;
  zc y = mkz(0, 0);
  for (int k = r; k < kp; ++k) y = fma_z(Ts[r + BT_NB * k], Ss[k][cl], y);   // T column-major
  Y[(size_t)r * m + col] = y;
}

// rows per V^H X chunk: 256, or down to 64 when the columns alone give few workgroups
static int bt_chunk(int rows, int ncol) {
  int nch = (rows + 255) / 256;
  if (ncol * nch >= 512) return 256;
  nch = std::min((768 + ncol - 1) / ncol, (rows + 63) / 64);
  const int ch = (rows + nch - 1) / nch;
  return std::max(64, std::min(256, (ch + 3) & ~3));
}

// X(r, col) -= sum_c V(r, c) Y(c, col) for rows r > p; 64 x 64 tiles, waves of 32 x 32
template <class T>
__global__ __launch_bounds__(256) void k_bt_update(const T* __restrict__ H, int ld, int N, int p, int kp,
                                                   const zc* __restrict__ Y, zc* __restrict__ X, int ldx,
                                                   int m) {
  const int rt = p + 1 + blockIdx.x * 64, col0 = blockIdx.y * 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int rb = rt + (wave >> 1) * 32, cb = col0 + (wave & 1) * 32;
  d4 ar[2][2], ai[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      ar[x][y] = d4{0, 0, 0, 0};
      ai[x][y] = d4{0, 0, 0, 0};
    }
  // the panel's reflectors in two halves of four MFMA steps: a half's loads are all issued
  // before its products (each accumulator still takes its steps in ascending k)
#pragma unroll
  for (int h = 0; h < BT_NB / 16; ++h) {
    zc v[4][2], yv[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = 16 * h + 4 * t + lk;
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = rb + x * 16 + lr;
        v[t][x] = (k < kp && r < N) ? vref(H, ld, r, p + k) : mkz(0, 0);
      }
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int col = cb + y * 16 + lr;
        yv[t][y] = (k < kp && col < m) ? Y[(size_t)k * m + col] : mkz(0, 0);
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (16 * h + 4 * t >= kp) break;             // the same steps as one step per 4 of k
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          ar[x][y] = mfma16x16x4(v[t][x].re, yv[t][y].re, ar[x][y]);
          ar[x][y] = mfma16x16x4(-v[t][x].im, yv[t][y].im, ar[x][y]);
          ai[x][y] = mfma16x16x4(v[t][x].re, yv[t][y].im, ai[x][y]);
          ai[x][y] = mfma16x16x4(v[t][x].im, yv[t][y].re, ai[x][y]);
        }
    }
  }
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      const int col = cb + y * 16 + lr;
      if (col >= m) continue;
      for (int q = 0; q < 4; ++q) {
        const int r = rb + x * 16 + lk + 4 * q;
        if (r >= N) continue;
        zc& d = X[(size_t)r + (size_t)ldx * col];
        d.re -= ar[x][y][q];
        d.im -= ai[x][y][q];
      }
    }
}

// X(:, c) = Z(:, c0 + c) as complex
__global__ void k_bt_load(int N, int m, const double* __restrict__ Z, int c0, zc* __restrict__ X, int ldx) {
  const size_t tot = (size_t)N * m;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(t % N), c = (int)(t / N);
    X[(size_t)r + (size_t)ldx * c] = mkz(Z[(size_t)r + (size_t)N * (c0 + c)], 0.0);
  }
}

// phase convention: the last component of each vector real and >= 0 (the eigenvectors
// of ZHEEV with UPLO = 'U' have a real last component)
__global__ __launch_bounds__(256) void k_bt_phase(int N, int m, zc* __restrict__ X, int ldx) {
  __shared__ zc s_last;
  const int c = blockIdx.x;
  if (c >= m) return;
  zc* x = X + (size_t)ldx * c;
  if (threadIdx.x == 0) s_last = x[N - 1];
  __syncthreads();
  const zc last = s_last;
  const double a = sqrt(norm2(last));
  if (!(a > 0.0)) return;
  const zc ph = mkz(last.re / a, -last.im / a);
  for (int r = threadIdx.x; r < N - 1; r += 256) x[r] = x[r] * ph;
  if (threadIdx.x == 0) x[N - 1] = mkz(a, 0.0);
}

size_t backtr_work_bytes(int N, int m) {
  const size_t np = (N + BT_NB - 1) / BT_NB + 1;
  const size_t nch = (N + 63) / 64 + 1;          // chunks of at least 64 rows
  return sizeof(zc) * (np * BT_NB * BT_NB + nch * BT_NB * (size_t)m + BT_NB * (size_t)m +
                       np * kBtSplit * (size_t)kBtPairSlots) +
         4 * 256;
}

template <class T>
static int run_backtr_t(hipStream_t s, const T* H, int ld, int N, const zc* tau, const double* Z, int c0,
                        int m, zc* X, int ldx, void* work) {
  if (N < 1 || m < 1) return WTB_OK;
  char* w = static_cast<char*>(work);
  const int np = (N - 1 + BT_NB - 1) / BT_NB;   // panels of reflectors 0 .. N-2
  zc* Tf = (zc*)w;
  w += ((sizeof(zc) * (size_t)(np + 1) * BT_NB * BT_NB + 255) & ~size_t(255));
  zc* Ypart = (zc*)w;
  const size_t nchmax = (N + 63) / 64 + 1;
  w += ((sizeof(zc) * nchmax * BT_NB * (size_t)m + 255) & ~size_t(255));
  zc* Y = (zc*)w;
  w += ((sizeof(zc) * BT_NB * (size_t)m + 255) & ~size_t(255));
  zc* Gp = (zc*)w;                                // Gram partials, kBtSplit per block
  const int gl = (int)std::min<size_t>(((size_t)N * m + 255) / 256, 65536);
  hipLaunchKernelGGL(k_bt_load, dim3(gl), dim3(256), 0, s, N, m, Z, c0, X, ldx);
  if (np > 0) {
    hipLaunchKernelGGL(k_bt_gram<T>, dim3(np, kBtSplit), dim3(256), 0, s, H, ld, N, Gp);
    hipLaunchKernelGGL(k_bt_tfactor, dim3(np), dim3(256), 0, s, N, tau, Gp, Tf);
    const int ncol = (m + 63) / 64;
    for (int b = np - 1; b >= 0; --b) {
      const int p = b * BT_NB;
      const int kp = std::min(BT_NB, N - 1 - p);
      const int rows = N - p - 1;
      const int chunk = bt_chunk(rows, ncol);
      const int nch = (rows + chunk - 1) / chunk;
      hipLaunchKernelGGL(k_bt_vhx<T>, dim3(ncol, nch), dim3(256), 0, s, H, ld, N, p, kp, X, ldx, m, chunk,
                         Ypart);
      hipLaunchKernelGGL(k_bt_ty, dim3((m + kBtTyCols - 1) / kBtTyCols), dim3(256), 0, s, Ypart, nch, m, kp,
                         Tf + (size_t)b * BT_NB * BT_NB, Y);
      hipLaunchKernelGGL(k_bt_update<T>, dim3((rows + 63) / 64, ncol), dim3(256), 0, s, H, ld, N, p, kp,
                         Y, X, ldx, m);
    }
  }
  hipLaunchKernelGGL(k_bt_phase, dim3(m), dim3(256), 0, s, N, m, X, ldx);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int run_backtr(hipStream_t s, const zc* H, int ld, int N, const zc* tau, const double* Z, int c0, int m,
               zc* X, int ldx, void* work) {
  return run_backtr_t<zc>(s, H, ld, N, tau, Z, c0, m, X, ldx, work);
}
int run_backtr_f32(hipStream_t s, const zf* H, int ld, int N, const zc* tau, const double* Z, int c0,
                   int m, zc* X, int ldx, void* work) {
  return run_backtr_t<zf>(s, H, ld, N, tau, Z, c0, m, X, ldx, work);
}

}  // namespace wtb
