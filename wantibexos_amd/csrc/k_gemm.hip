// Complex GEMM on MFMA for the iterative (Davidson / Lanczos) solvers:
//   C(M x N) = alpha op(A) op(B) + beta C,   op = identity or conjugate transpose,
// column-major operands.  64 x 64 output tile per workgroup, K staged through
// LDS in slices of 16 with loads arranged along the contiguous index of each
// operand (so A^H and B^H read whole column segments), four waves of 32 x 32,
// complex products as four real MFMAs.  FP64 operands use
// v_mfma_f64_16x16x4_f64; the precision form (F32) rounds both operands to FP32
// and accumulates in FP32 on v_mfma_f32_16x16x4_f32.
#include "wtb_kernels.h"
#include <algorithm>
#include <cstdint>

namespace wtb {

#define GT 64   // output tile edge
#define GK 16   // K slice

// Split K (gridDim.z > 1): slice z covers k in [z kc, (z+1) kc) and writes its raw sums to
// part + z M N (ld M); k_zgemm_reduce then forms alpha sum_z part + beta C in slice order.
template <int OPA, int OPB, class TA, class TB, bool F32>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, zc alpha, const TA* __restrict__ A,
                                               int lda, const TB* __restrict__ B, int ldb, zc beta,
                                               zc* __restrict__ C, int ldc, int kc, zc* __restrict__ part) {
  __shared__ zc As[GK][GT + 1];   // (k, i)
  __shared__ zc Bs[GK][GT + 1];   // (k, j)
  const int i0 = blockIdx.x * GT, j0 = blockIdx.y * GT;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  typedef typename std::conditional<F32, f4, d4>::type acc_t;
  acc_t cr[2][2], ci[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      cr[x][y] = acc_t{0, 0, 0, 0};
      ci[x][y] = acc_t{0, 0, 0, 0};
    }
  const int kbeg = blockIdx.z * kc;
  const int kend = min(K, kbeg + kc);
  // this thread's four staging slots of op(A) and op(B) per K slice; the next slice is
  // loaded into registers while the current one is multiplied
  zc ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (OPA == 0) {
        const int gi = i0 + (tid & 63), gk = k0 + (tid >> 6) + 4 * u;
        ra[u] = (gi < M && gk < kend) ? ldz(A[(size_t)gi + (size_t)lda * gk]) : mkz(0, 0);
      } else {
        const int gi = i0 + (tid >> 4) + 16 * u, gk = k0 + (tid & 15);
        ra[u] = (gi < M && gk < kend) ? conj(ldz(A[(size_t)gk + (size_t)lda * gi])) : mkz(0, 0);
      }
      if (OPB == 0) {
        const int gj = j0 + (tid >> 4) + 16 * u, gk = k0 + (tid & 15);
        rb[u] = (gj < N && gk < kend) ? ldz(B[(size_t)gk + (size_t)ldb * gj]) : mkz(0, 0);
      } else {
        const int gj = j0 + (tid & 63), gk = k0 + (tid >> 6) + 4 * u;
        rb[u] = (gj < N && gk < kend) ? conj(ldz(B[(size_t)gj + (size_t)ldb * gk])) : mkz(0, 0);
      }
    }
  };
  if (kbeg < kend) gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += GK) {
    __syncthreads();                                  // the previous slice is consumed
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (OPA == 0) As[(tid >> 6) + 4 * u][tid & 63] = ra[u];
      else As[tid & 15][(tid >> 4) + 16 * u] = ra[u];
      if (OPB == 0) Bs[tid & 15][(tid >> 4) + 16 * u] = rb[u];
      else Bs[(tid >> 6) + 4 * u][tid & 63] = rb[u];
    }
    __syncthreads();
    if (k0 + GK < kend) gload(k0 + GK);
    for (int kq = 0; kq < GK; kq += 4) {
      zc a[2], b[2];
      for (int x = 0; x < 2; ++x) a[x] = As[kq + lk][wi + x * 16 + lr];
      for (int y = 0; y < 2; ++y) b[y] = Bs[kq + lk][wj + y * 16 + lr];
      for (int x = 0; x < 2; ++x)
        for (int y = 0; y < 2; ++y) {
          if constexpr (F32) {
            const float ar = (float)a[x].re, ai = (float)a[x].im, br = (float)b[y].re, bi = (float)b[y].im;
            cr[x][y] = mfma16x16x4f32(ar, br, cr[x][y]);
            cr[x][y] = mfma16x16x4f32(-ai, bi, cr[x][y]);
            ci[x][y] = mfma16x16x4f32(ar, bi, ci[x][y]);
            ci[x][y] = mfma16x16x4f32(ai, br, ci[x][y]);
          } else {
            cr[x][y] = mfma16x16x4(a[x].re, b[y].re, cr[x][y]);
            cr[x][y] = mfma16x16x4(-a[x].im, b[y].im, cr[x][y]);
            ci[x][y] = mfma16x16x4(a[x].re, b[y].im, ci[x][y]);
            ci[x][y] = mfma16x16x4(a[x].im, b[y].re, ci[x][y]);
          }
        }
    }
  }
  const bool use_beta = beta.re != 0.0 || beta.im != 0.0;
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      const int j = j0 + wj + y * 16 + lr;
      if (j >= N) continue;
      for (int q = 0; q < 4; ++q) {
        // f64 MFMA: register q holds row lk + 4q; f32 MFMA: row 4 lk + q
        const int i = i0 + wi + x * 16 + (F32 ? 4 * lk + q : lk + 4 * q);
        if (i >= M) continue;
        const zc v = mkz((double)cr[x][y][q], (double)ci[x][y][q]);
        if (part) {
          part[(size_t)blockIdx.z * M * N + (size_t)i + (size_t)M * j] = v;
          continue;
        }
        zc& c = C[(size_t)i + (size_t)ldc * j];
        zc out = alpha * v;
        if (use_beta) out += beta * c;
        c = out;
      }
    }
}

__global__ void k_zgemm_reduce(int M, int N, int nz, const zc* __restrict__ part, zc alpha, zc beta,
                               zc* __restrict__ C, int ldc) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (size_t)M * N) return;
  const int i = (int)(t % M), j = (int)(t / M);
  zc acc = mkz(0, 0);
  for (int z = 0; z < nz; ++z) acc += part[(size_t)z * M * N + t];
  zc& c = C[(size_t)i + (size_t)ldc * j];
  zc out = alpha * acc;
  if (beta.re != 0.0 || beta.im != 0.0) out += beta * c;
  c = out;
}

int zgemm(hipStream_t s, char opa, char opb, int M, int N, int K, zc alpha, const zc* A, int lda,
          const zc* B, int ldb, zc beta, zc* C, int ldc, zc* work, size_t work_elems) {
  if (M <= 0 || N <= 0) return WTB_OK;
  if (K <= 0) {
    set_error("zgemm: K = 0");
    return WTB_ERR_ARG;
  }
  const int gx = (M + GT - 1) / GT, gy = (N + GT - 1) / GT;
  // K slices: aim at ~1024 workgroups, slices of at least 256
  int nz = 1, kc = K;
  const size_t mn = (size_t)M * N;
  if (work && gx * gy < 512 && K >= 512) {
    nz = std::min((1024 + gx * gy - 1) / (gx * gy), K / 256);
    nz = (int)std::min<size_t>((size_t)nz, work_elems / mn);
    if (nz >= 2) {
      kc = (((K + nz - 1) / nz + GK - 1) / GK) * GK;
      nz = (K + kc - 1) / kc;
    } else {
      nz = 1;
    }
  }
  dim3 g(gx, gy, nz), b(256);
  const bool ca = opa == 'C', cb = opb == 'C';
  zc* pt = nz > 1 ? work : nullptr;
  if (!ca && !cb) hipLaunchKernelGGL((k_zgemm<0, 0, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kc, pt);
  else if (ca && !cb) hipLaunchKernelGGL((k_zgemm<1, 0, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kc, pt);
  else if (!ca && cb) hipLaunchKernelGGL((k_zgemm<0, 1, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kc, pt);
  else hipLaunchKernelGGL((k_zgemm<1, 1, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, kc, pt);
  if (nz > 1)
    hipLaunchKernelGGL(k_zgemm_reduce, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, M, N, nz, work, alpha,
                       beta, C, ldc);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// Thin product Y = H X (nc <= 16 columns) for the Lanczos chains.  A workgroup covers
// THIN_ROWS rows (wave w: 128 of them, two per lane, so every load of H moves 16 B per lane
// and 1 KB of one column per wave) and THIN_KC columns of K, staged THIN_SUB at a time
// through LDS as FP32 pairs.  FP32 accumulation for the COMPLEX*8 matrix; the partial sums
// of the K slices go to part[z] (ld N) and are added in slice order.
#define THIN_ROWS 512
#define THIN_KC 512
#define THIN_SUB 256
template <bool VEC>
__global__ __launch_bounds__(256) void k_thin_hmm(int N, int nc, const zf* __restrict__ H, int ldh,
                                                  const zc* __restrict__ X, int ldx, zc* __restrict__ part) {
  __shared__ float2 sm[THIN_SUB * 16];            // X sub-slice [k][16]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = blockIdx.x * THIN_ROWS + w * 128 + 2 * lane;   // rows r and r + 1
  const int k0 = blockIdx.y * THIN_KC;
  const int kend = min(N, k0 + THIN_KC);
  const bool rv = r < N;
  float ar0[16], ai0[16], ar1[16], ai1[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) ar0[q] = ai0[q] = ar1[q] = ai1[q] = 0.f;
  for (int ks = k0; ks < kend; ks += THIN_SUB) {
    const int kn = min(THIN_SUB, kend - ks);
    __syncthreads();
    for (int e = threadIdx.x; e < THIN_SUB * 16; e += 256) {
      const int kk = e >> 4, q = e & 15;
      float2 v = make_float2(0.f, 0.f);
      if (q < nc && kk < kn) {
        const zc x = X[(size_t)(ks + kk) + (size_t)ldx * q];
        v = make_float2((float)x.re, (float)x.im);
      }
      sm[e] = v;
    }
    __syncthreads();
    if (!rv) continue;
    for (int kb = 0; kb < kn; kb += 8) {
      float4 h[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kk = kb + u;
        const zf* src = H + (size_t)r + (size_t)ldh * (ks + kk);
        if (kk < kn) {
          if (VEC) {
            h[u] = *reinterpret_cast<const float4*>(src);
          } else {
            const zf a0 = src[0];
            const zf a1 = (r + 1 < N) ? src[1] : zf{0.f, 0.f};
            h[u] = make_float4(a0.re, a0.im, a1.re, a1.im);
          }
        } else {
          h[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          if (q < nc) {
            const float2 x = sm[(kb + u) * 16 + q];
            ar0[q] = fmaf(h[u].x, x.x, fmaf(-h[u].y, x.y, ar0[q]));
            ai0[q] = fmaf(h[u].x, x.y, fmaf(h[u].y, x.x, ai0[q]));
            ar1[q] = fmaf(h[u].z, x.x, fmaf(-h[u].w, x.y, ar1[q]));
            ai1[q] = fmaf(h[u].z, x.y, fmaf(h[u].w, x.x, ai1[q]));
          }
        }
      }
    }
  }
  if (!rv) return;
  zc* dst = part + (size_t)blockIdx.y * N * nc;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    if (q < nc) {
      dst[(size_t)r + (size_t)N * q] = mkz((double)ar0[q], (double)ai0[q]);
      if (r + 1 < N) dst[(size_t)r + 1 + (size_t)N * q] = mkz((double)ar1[q], (double)ai1[q]);
    }
  }
}

// The same thin product on FP32 MFMA (v_mfma_f32_16x16x4_f32, complex as four real
// products), for an even N and leading dimension: lane l loads rows (2i, 2i+1), i = l%16,
// of column k0 + l/16 as one 16-byte pair; the even and the odd rows form two 16-row
// blocks of the A operand, so a wave covers 64 rows (four blocks) per 4-column step and
// the B operand (X, columns padded to 16) comes from LDS.  Same K slices and partials as
// k_thin_hmm.
#define THIN_MROWS 256   // rows per workgroup of k_thin_mfma: 4 waves x 64
__global__ __launch_bounds__(256) void k_thin_mfma(int N, int nc, const zf* __restrict__ H, int ldh,
                                                   const zc* __restrict__ X, int ldx, zc* __restrict__ part) {
  __shared__ float2 sm[THIN_SUB * 16];            // X sub-slice [k][16]
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = lane & 15, kq = lane >> 4;
  const int R0 = blockIdx.x * THIN_MROWS + w * 64;
  const int k0 = blockIdx.y * THIN_KC;
  const int kend = min(N, k0 + THIN_KC);
  // rows of the two 32-row halves this lane loads (clamped in range; their results are dropped)
  const int ra = R0 + 2 * i, rb = R0 + 32 + 2 * i;
  const zf* pa = H + (ra < N ? ra : 0);
  const zf* pb = H + (rb < N ? rb : 0);
  f4 cr[4], ci[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) cr[b] = ci[b] = f4{0.f, 0.f, 0.f, 0.f};
  for (int ks = k0; ks < kend; ks += THIN_SUB) {
    const int kn = min(THIN_SUB, kend - ks);
    __syncthreads();
    for (int e = threadIdx.x; e < THIN_SUB * 16; e += 256) {
      const int kk = e >> 4, q = e & 15;
      float2 v = make_float2(0.f, 0.f);
      if (q < nc && kk < kn) {
        const zc x = X[(size_t)(ks + kk) + (size_t)ldx * q];
        v = make_float2((float)x.re, (float)x.im);
      }
      sm[e] = v;
    }
    __syncthreads();
    for (int kb = 0; kb < kn; kb += 16) {
      float4 av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        int col = ks + kb + 4 * u + kq;
        col = col < N ? col : N - 1;                // staged X is zero there
        av[u] = *reinterpret_cast<const float4*>(pa + (size_t)ldh * col);
        bv[u] = *reinterpret_cast<const float4*>(pb + (size_t)ldh * col);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float2 x = sm[(kb + 4 * u + kq) * 16 + i];
        const float xr = x.x, xi = x.y, nxi = -x.y;
        const float A[4][2] = {{av[u].x, av[u].y}, {av[u].z, av[u].w}, {bv[u].x, bv[u].y}, {bv[u].z, bv[u].w}};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          cr[b] = mfma16x16x4f32(A[b][0], xr, cr[b]);
          cr[b] = mfma16x16x4f32(A[b][1], nxi, cr[b]);
          ci[b] = mfma16x16x4f32(A[b][0], xi, ci[b]);
          ci[b] = mfma16x16x4f32(A[b][1], xr, ci[b]);
        }
      }
    }
  }
  if (i >= nc) return;
  zc* dst = part + (size_t)blockIdx.y * N * nc + (size_t)N * i;
  // block b = 2 h + parity; register q holds block row 4 kq + q = local pair index
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int row = R0 + 32 * (b >> 1) + 2 * (4 * kq + q) + (b & 1);
      if (row < N) dst[row] = mkz((double)cr[b][q], (double)ci[b][q]);
    }
}

size_t thin_work_bytes(int N, int nc) {
  return sizeof(zc) * (size_t)((N + THIN_KC - 1) / THIN_KC) * N * nc + 256;
}

int hmm_thin_f32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                 zc* work) {
  if (nc < 1 || nc > 16) {
    set_error("thin product: 1..16 columns");
    return WTB_ERR_ARG;
  }
  const int nz = (N + THIN_KC - 1) / THIN_KC;
  const dim3 g((N + THIN_ROWS - 1) / THIN_ROWS, nz);
  // 16-byte loads of a row pair need an even leading dimension and a 16-byte aligned base
  const bool pairs = (ldh % 2) == 0 && (reinterpret_cast<uintptr_t>(H) % 16) == 0;
  if (pairs && (N % 2) == 0)
    hipLaunchKernelGGL(k_thin_mfma, dim3((N + THIN_MROWS - 1) / THIN_MROWS, nz), dim3(256), 0, s, N, nc, H, ldh, X,
                       ldx, work);
  else if (pairs)
    hipLaunchKernelGGL(k_thin_hmm<true>, g, dim3(256), 0, s, N, nc, H, ldh, X, ldx, work);
  else
    hipLaunchKernelGGL(k_thin_hmm<false>, g, dim3(256), 0, s, N, nc, H, ldh, X, ldx, work);
  const size_t tot = (size_t)N * nc;
  hipLaunchKernelGGL(k_zgemm_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, nc, nz, work,
                     mkz(1, 0), mkz(0, 0), Y, ldy);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

// Y = H X for the COMPLEX*8 matrix (even N and leading dimension) on FP32 MFMA with FP32
// accumulation, or (F64) on FP64 MFMA with the matrix widened exactly to FP64: a
// workgroup computes 128 rows x 64 columns, wave w rows 32w..32w+31 (two 16-row blocks x
// four 16-column blocks, complex as four real products).  K advances in slices of 16:
// H as row pairs (one 16-byte load per lane, 1 KB per column per wave) and X converted to
// FP32 pairs once, both staged in LDS; the next slice is loaded into registers while the
// current one is multiplied.  K slices of the grid (z) write partial sums.
#define H32_R 128
#define H32_C 64
template <bool F64>
__global__ __launch_bounds__(256) void k_hmm32(int N, int nc, const zf* __restrict__ H, int ldh,
                                               const zc* __restrict__ X, int ldx, zc* __restrict__ Y, int ldy,
                                               int kc, zc* __restrict__ part) {
  typedef typename std::conditional<F64, double2, float2>::type bt;
  typedef typename std::conditional<F64, d4, f4>::type acc_t;
  // rows padded by one element: the four k rows a wave reads at once, and the 16 rows the
  // X loader stores at once, fall in different LDS banks
  __shared__ float2 As[16][H32_R + 1];
  __shared__ bt Bs[16][H32_C + 1];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i16 = lane & 15, kq = lane >> 4;
  const int R0 = blockIdx.x * H32_R, C0 = blockIdx.y * H32_C;
  const int kbeg = blockIdx.z * kc, kend = min(N, kbeg + kc);
  // loader: thread t stages row pair (t & 63) for k = (t >> 6) + 4u, and X(k = t & 15,
  // column (t >> 4) + 16u)
  const int lrp = tid & 63, lk = tid >> 6;
  const int arow = R0 + 2 * lrp;
  const zf* abase = H + (arow < N ? arow : 0);     // rows past N: any in-range pair, never written
  const int bk = tid & 15, bcol = tid >> 4;
  acc_t cr[2][4], ci[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) cr[a][b] = ci[a][b] = acc_t{0, 0, 0, 0};
  float4 ra[4];
  zc rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = min(k0 + lk + 4 * u, N - 1);   // past kend: X is zero there
      ra[u] = *reinterpret_cast<const float4*>(abase + (size_t)ldh * k);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = k0 + bk, col = C0 + bcol + 16 * u;
      rb[u] = (k < kend && col < nc) ? X[(size_t)k + (size_t)ldx * col] : mkz(0, 0);
    }
  };
  gload(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += 16) {
    __syncthreads();                                  // the previous slice is consumed
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      As[lk + 4 * u][2 * lrp] = make_float2(ra[u].x, ra[u].y);
      As[lk + 4 * u][2 * lrp + 1] = make_float2(ra[u].z, ra[u].w);
      if constexpr (F64) Bs[bk][bcol + 16 * u] = make_double2(rb[u].re, rb[u].im);
      else Bs[bk][bcol + 16 * u] = make_float2((float)rb[u].re, (float)rb[u].im);
    }
    __syncthreads();
    if (k0 + 16 < kend) gload(k0 + 16);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      const int k = 4 * kk + kq;
      float2 av[2];
      bt bv[4];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = As[k][32 * w + 16 * a + i16];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[k][16 * b + i16];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          if constexpr (F64) {
            const double ar = av[a].x, ai = av[a].y;
            cr[a][b] = mfma16x16x4(ar, bv[b].x, cr[a][b]);
            cr[a][b] = mfma16x16x4(ai, -bv[b].y, cr[a][b]);
            ci[a][b] = mfma16x16x4(ar, bv[b].y, ci[a][b]);
            ci[a][b] = mfma16x16x4(ai, bv[b].x, ci[a][b]);
          } else {
            cr[a][b] = mfma16x16x4f32(av[a].x, bv[b].x, cr[a][b]);
            cr[a][b] = mfma16x16x4f32(av[a].y, -bv[b].y, cr[a][b]);
            ci[a][b] = mfma16x16x4f32(av[a].x, bv[b].y, ci[a][b]);
            ci[a][b] = mfma16x16x4f32(av[a].y, bv[b].x, ci[a][b]);
          }
        }
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col = C0 + 16 * b + i16;
      if (col >= nc) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // f32 MFMA: register q holds row 4 kq + q; f64 MFMA: row kq + 4 q
        const int row = R0 + 32 * w + 16 * a + (F64 ? kq + 4 * q : 4 * kq + q);
        if (row >= N) continue;
        const zc v = mkz((double)cr[a][b][q], (double)ci[a][b][q]);
        if (part) part[(size_t)blockIdx.z * N * nc + (size_t)row + (size_t)N * col] = v;
        else Y[(size_t)row + (size_t)ldy * col] = v;
      }
    }
}

// Y = H X with enough workgroups to fill the chip: when the output tiles alone would
// give fewer than ~1024 workgroups, K is split into slices (partial sums in work, of
// hmm_work_bytes) reduced in a fixed order.
static int split_of(int N, int nc) {
  const int tiles = ((N + GT - 1) / GT) * ((nc + GT - 1) / GT);
  int ns = (1024 + tiles - 1) / tiles;
  const int maxs = N / 256 > 1 ? N / 256 : 1;
  return ns < 1 ? 1 : (ns > maxs ? maxs : (ns > 32 ? 32 : ns));
}

// enough for every column count up to nc
size_t hmm_work_bytes(int N, int nc) {
  size_t mx = 0;
  for (int c = 1; c <= nc; ++c) {
    const size_t b = (size_t)split_of(N, c) * N * c;
    mx = b > mx ? b : mx;
  }
  return sizeof(zc) * mx + 256;
}

template <class TA, bool F32>
static int hmm_t(hipStream_t s, int N, int nc, const TA* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                 zc* work) {
  if (N <= 0 || nc <= 0) return WTB_OK;
  const int ns = work ? split_of(N, nc) : 1;
  const int kc = (((N + ns - 1) / ns + GK - 1) / GK) * GK;
  dim3 g((N + GT - 1) / GT, (nc + GT - 1) / GT, ns), b(256);
  hipLaunchKernelGGL((k_zgemm<0, 0, TA, zc, F32>), g, b, 0, s, N, nc, N, mkz(1, 0), H, ldh, X, ldx, mkz(0, 0), Y,
                     ldy, kc, ns > 1 ? work : nullptr);
  if (ns > 1) {
    const size_t tot = (size_t)N * nc;
    hipLaunchKernelGGL(k_zgemm_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, nc, ns, work,
                       mkz(1, 0), mkz(0, 0), Y, ldy);
  }
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

template <bool F64>
static int hmm_h32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                   zc* work) {
  if (N <= 0 || nc <= 0) return WTB_OK;
  if ((N % 2) || (ldh % 2) || (reinterpret_cast<uintptr_t>(H) % 16))
    return hmm_t<zf, !F64>(s, N, nc, H, ldh, X, ldx, Y, ldy, work);
  // K slices so that the grid reaches ~1024 workgroups; the partials fit the workspace
  // (hmm_work_bytes), which is sized for the 64 x 64 tiling's slices (never fewer)
  const int tiles = ((N + H32_R - 1) / H32_R) * ((nc + H32_C - 1) / H32_C);
  int ns = work ? std::min((1024 + tiles - 1) / tiles, std::min(split_of(N, nc), std::max(1, N / 512))) : 1;
  if (ns < 1) ns = 1;
  const int kc = (((N + ns - 1) / ns + 15) / 16) * 16;
  ns = (N + kc - 1) / kc;
  dim3 g((N + H32_R - 1) / H32_R, (nc + H32_C - 1) / H32_C, ns), b(256);
  hipLaunchKernelGGL(k_hmm32<F64>, g, b, 0, s, N, nc, H, ldh, X, ldx, Y, ldy, kc, ns > 1 ? work : nullptr);
  if (ns > 1) {
    const size_t tot = (size_t)N * nc;
    hipLaunchKernelGGL(k_zgemm_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, N, nc, ns, work,
                       mkz(1, 0), mkz(0, 0), Y, ldy);
  }
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int hmm_f32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy, zc* work) {
  return hmm_h32<false>(s, N, nc, H, ldh, X, ldx, Y, ldy, work);
}
int hmm_f32_acc64(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                  zc* work) {
  return hmm_h32<true>(s, N, nc, H, ldh, X, ldx, Y, ldy, work);
}
int hmm_f64(hipStream_t s, int N, int nc, const zc* H, int ldh, const zc* X, int ldx, zc* Y, int ldy, zc* work) {
  return hmm_t<zc, false>(s, N, nc, H, ldh, X, ldx, Y, ldy, work);
}

}  // namespace wtb
