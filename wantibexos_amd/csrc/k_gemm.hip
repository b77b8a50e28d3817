// Complex GEMM on MFMA for the iterative (Davidson / Lanczos) solvers:
//   C(M x N) = alpha op(A) op(B) + beta C,   op = identity or conjugate transpose,
// column-major operands.  64 x 64 output tile per workgroup, K staged through
// LDS in slices of 16 with loads arranged along the contiguous index of each
// operand (so A^H and B^H read whole column segments), four waves of 32 x 32,
// complex products as four real MFMAs.  FP64 operands use
// v_mfma_f64_16x16x4_f64; the precision form (F32) rounds both operands to FP32
// and accumulates in FP32 on v_mfma_f32_16x16x4_f32.
#include "wtb_kernels.h"

namespace wtb {

#define GT 64   // output tile edge
#define GK 16   // K slice

template <int OPA, int OPB, class TA, class TB, bool F32>
__global__ __launch_bounds__(256) void k_zgemm(int M, int N, int K, zc alpha, const TA* __restrict__ A,
                                               int lda, const TB* __restrict__ B, int ldb, zc beta,
                                               zc* __restrict__ C, int ldc) {
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
  for (int k0 = 0; k0 < K; k0 += GK) {
    // stage op(A)(i0:i0+64, k0:k0+16)
    if (OPA == 0) {
      const int i = tid & 63;
      for (int kk = tid >> 6; kk < GK; kk += 4) {
        const int gi = i0 + i, gk = k0 + kk;
        As[kk][i] = (gi < M && gk < K) ? ldz(A[(size_t)gi + (size_t)lda * gk]) : mkz(0, 0);
      }
    } else {
      const int kk = tid & 15;
      for (int i = tid >> 4; i < GT; i += 16) {
        const int gi = i0 + i, gk = k0 + kk;
        As[kk][i] = (gi < M && gk < K) ? conj(ldz(A[(size_t)gk + (size_t)lda * gi])) : mkz(0, 0);
      }
    }
    // stage op(B)(k0:k0+16, j0:j0+64)
    if (OPB == 0) {
      const int kk = tid & 15;
      for (int j = tid >> 4; j < GT; j += 16) {
        const int gj = j0 + j, gk = k0 + kk;
        Bs[kk][j] = (gj < N && gk < K) ? ldz(B[(size_t)gk + (size_t)ldb * gj]) : mkz(0, 0);
      }
    } else {
      const int j = tid & 63;
      for (int kk = tid >> 6; kk < GK; kk += 4) {
        const int gj = j0 + j, gk = k0 + kk;
        Bs[kk][j] = (gj < N && gk < K) ? conj(ldz(B[(size_t)gj + (size_t)ldb * gk])) : mkz(0, 0);
      }
    }
    __syncthreads();
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
    __syncthreads();
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
        zc& c = C[(size_t)i + (size_t)ldc * j];
        zc out = alpha * v;
        if (use_beta) out += beta * c;
        c = out;
      }
    }
}

int zgemm(hipStream_t s, char opa, char opb, int M, int N, int K, zc alpha, const zc* A, int lda,
          const zc* B, int ldb, zc beta, zc* C, int ldc) {
  if (M <= 0 || N <= 0) return WTB_OK;
  if (K <= 0) {
    set_error("zgemm: K = 0");
    return WTB_ERR_ARG;
  }
  dim3 g((M + GT - 1) / GT, (N + GT - 1) / GT), b(256);
  const bool ca = opa == 'C', cb = opb == 'C';
  if (!ca && !cb) hipLaunchKernelGGL((k_zgemm<0, 0, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (ca && !cb) hipLaunchKernelGGL((k_zgemm<1, 0, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else if (!ca && cb) hipLaunchKernelGGL((k_zgemm<0, 1, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  else hipLaunchKernelGGL((k_zgemm<1, 1, zc, zc, false>), g, b, 0, s, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int hmm_f32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy) {
  if (N <= 0 || nc <= 0) return WTB_OK;
  dim3 g((N + GT - 1) / GT, (nc + GT - 1) / GT), b(256);
  hipLaunchKernelGGL((k_zgemm<0, 0, zf, zc, true>), g, b, 0, s, N, nc, N, mkz(1, 0), H, ldh, X, ldx,
                     mkz(0, 0), Y, ldy);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int hmm_f32_acc64(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy) {
  if (N <= 0 || nc <= 0) return WTB_OK;
  dim3 g((N + GT - 1) / GT, (nc + GT - 1) / GT), b(256);
  hipLaunchKernelGGL((k_zgemm<0, 0, zf, zc, false>), g, b, 0, s, N, nc, N, mkz(1, 0), H, ldh, X, ldx,
                     mkz(0, 0), Y, ldy);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int hmm_f64(hipStream_t s, int N, int nc, const zc* H, int ldh, const zc* X, int ldx, zc* Y, int ldy) {
  if (N <= 0 || nc <= 0) return WTB_OK;
  dim3 g((N + GT - 1) / GT, (nc + GT - 1) / GT), b(256);
  hipLaunchKernelGGL((k_zgemm<0, 0, zc, zc, false>), g, b, 0, s, N, nc, N, mkz(1, 0), H, ldh, X, ldx,
                     mkz(0, 0), Y, ldy);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
