// Block Davidson for the lowest excitons of the BSE matrix (the iterative form of
// the eigensolve that subprograms/bse_solver-tool-diel.f90:485-492 does with a
// full ZHEEV).  The matrix is the full Hermitian COMPLEX*8 copy; the products
// H V run on FP32 MFMA, everything else (basis, projected matrix, Ritz pairs,
// residuals) is FP64.
//
//   V (N x m) orthonormal, HV = H V, G = V^H H V; V starts as unit vectors at the nk
//   smallest D, or as the lowest eigenvectors of the preconditioner's coarse block
//   loop: G = S diag(theta) S^H (tridiagonal reduction + divide and conquer +
//         back-transform on the GPU), U = V S_k, HU = HV S_k, R = HU - U theta
//         stop when the first nev residual norms are below tol
//         T = P(R_unconverged)  (diagonal or learned preconditioner; the learned one
//         only in the FP32 phase)
//         T -= V (V^H T) twice, then twice T <- T L^-H with T^H T = L L^H (columns
//         of tiny pivots dropped) and T -= V (V^H T) between the two, V += T,
//         HV += H T, G grows by V^H (H T)
//         restart with the k = nev + guard Ritz vectors when m would pass max_sub
// The products run on FP32 MFMA until the residuals reach switch_tol, then the basis
// restarts on the Ritz vectors with H U recomputed in FP64 arithmetic (same COMPLEX*8
// matrix) and the remaining iterations use FP64 products, so that the residual bound
// refers to the stored matrix and not to the FP32 rounding of the products.
#include "wtb_kernels.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

namespace wtb {

// G(r, c) for r >= m0 > c from the computed block G(c, r) (c < m0 <= r): Hermitian mirror
__global__ void k_mirror_block(int m0, int m1, zc* G, int ld) {
  const int r = m0 + blockIdx.x;
  const int c = blockIdx.y * 256 + threadIdx.x;
  if (r >= m1 || c >= m0) return;
  G[(size_t)r + (size_t)ld * c] = conj(G[(size_t)c + (size_t)ld * r]);
}

__global__ void k_diag_set(int k, const double* theta, zc* G, int ld) {
  const int c = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
  if (r >= k || c >= k) return;
  G[(size_t)r + (size_t)ld * c] = mkz(r == c ? theta[r] : 0.0, 0.0);
}

__global__ void k_unit_cols(int N, int k, const int* rows, zc* V, int ld) {
  const int c = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
  if (r >= N || c >= k) return;
  V[(size_t)r + (size_t)ld * c] = mkz(r == rows[c] ? 1.0 : 0.0, 0.0);
}

// columns src(map q) -> dst q
__global__ void k_gather_cols(int N, int nq, const int* map, const zc* src, int lds, zc* dst, int ldd) {
  const int q = blockIdx.y, r = blockIdx.x * 256 + threadIdx.x;
  if (r >= N || q >= nq) return;
  dst[(size_t)r + (size_t)ldd * q] = src[(size_t)r + (size_t)lds * map[q]];
}

// coarse block of H: G(i, j) = H(rows[i], rows[j]) widened to FP64 (H full Hermitian)
__global__ void k_gather_block(int M, const int* rows, const zf* H, int ldh, zc* G) {
  const int j = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M || j >= M) return;
  G[(size_t)i + (size_t)M * j] = ldz(H[(size_t)rows[i] + (size_t)ldh * rows[j]]);
}
// RS(i, q) = R(rows[i], idx[q])
__global__ void k_gather_rows(int M, int nq, const int* rows, const int* idx, const zc* R, int ldr, zc* RS) {
  const int q = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M || q >= nq) return;
  RS[(size_t)i + (size_t)M * q] = R[(size_t)rows[i] + (size_t)ldr * idx[q]];
}
// Y(i, q) /= lam_i - theta_idx[q], the difference kept at least dmin away from zero
__global__ void k_coarse_scale(int M, int nq, const int* idx, const double* lam, const double* theta, double dmin,
                               zc* Y) {
  const int q = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M || q >= nq) return;
  double d = lam[i] - theta[idx[q]];
  if (fabs(d) < dmin) d = d < 0.0 ? -dmin : dmin;
  zc& y = Y[(size_t)i + (size_t)M * q];
  y = y * (1.0 / d);
}
// T(rows[i], q) = Z(i, q)
__global__ void k_scatter_rows(int M, int nq, const int* rows, const zc* Z, zc* T, int ldt) {
  const int q = blockIdx.y, i = blockIdx.x * 256 + threadIdx.x;
  if (i >= M || q >= nq) return;
  T[(size_t)rows[i] + (size_t)ldt * q] = Z[(size_t)i + (size_t)M * q];
}

namespace {

struct SmallEig {
  int mmax = 0, kmax = 0;
  zc* A = nullptr;       // mmax x mmax working copy
  void* tri = nullptr;
  void* dcw = nullptr;
  void* btw = nullptr;
  double* lam = nullptr;
  double* Z = nullptr;   // mmax x mmax
};

int small_alloc(SmallEig& e, int mmax, int kmax) {
  e.mmax = mmax;
  e.kmax = kmax;
  WTB_HIP(hipMalloc(&e.A, sizeof(zc) * (size_t)mmax * mmax));
  WTB_HIP(hipMalloc(&e.tri, tridiag_work_bytes(mmax, 32, 0)));
  WTB_HIP(hipMalloc(&e.dcw, dc_work_bytes(mmax, 0, 1)));
  WTB_HIP(hipMalloc(&e.btw, backtr_work_bytes(mmax, kmax)));
  WTB_HIP(hipMalloc(&e.lam, sizeof(double) * mmax));
  WTB_HIP(hipMalloc(&e.Z, sizeof(double) * (size_t)mmax * mmax));
  return WTB_OK;
}

void small_free(SmallEig& e) {
  hipFree(e.A);
  hipFree(e.tri);
  hipFree(e.dcw);
  hipFree(e.btw);
  hipFree(e.lam);
  hipFree(e.Z);
}

// eigenpairs of the m x m Hermitian G (lower triangle used): all eigenvalues into
// theta, the first k eigenvectors into S (m x k, ld lds)
int small_heev(hipStream_t s, SmallEig& e, int m, const zc* G, int ldg, double* theta, zc* S, int lds, int k) {
  WTB_HIP(hipMemcpy2DAsync(e.A, sizeof(zc) * m, G, sizeof(zc) * ldg, sizeof(zc) * m, m,
                           hipMemcpyDeviceToDevice, s));
  TridiagWork tw;
  tridiag_bind(tw, e.tri, m, 32, 0);
  tw.store_v = 1;
  DCTree tree;
  int rc = dc_prepare(e.dcw, m, 0, tree, 1);
  if (rc) return rc;
  if ((rc = run_tridiag(s, tw, e.A, m, nullptr))) return rc;
  if ((rc = run_dc(s, e.dcw, tree, tw.d, tw.e, nullptr, theta, e.Z, k))) return rc;
  return run_backtr(s, e.A, m, m, tw.tau, e.Z, 0, k, S, lds, e.btw);
}

}  // namespace

struct DavidsonWork {
  int N = 0, nk = 0, msub = 0, M = 0;   // sizes the buffers were allocated for
  zc *V = nullptr, *HV = nullptr, *G = nullptr, *S = nullptr, *U = nullptr, *HU = nullptr, *R = nullptr,
     *T = nullptr, *T2 = nullptr, *P = nullptr, *Gb = nullptr, *Linv = nullptr;
  double *theta = nullptr, *rn = nullptr;
  int *idx = nullptr, *keep = nullptr, *map = nullptr;
  zc* hw = nullptr;
  zc* gw = nullptr;
  int* crow = nullptr;
  zc *QS = nullptr, *CS = nullptr, *CY = nullptr, *CZ = nullptr;
  double* lamS = nullptr;
  SmallEig se, sc;
  bool have = false;

  void release() {
    for (void* p : {(void*)V, (void*)HV, (void*)G, (void*)S, (void*)U, (void*)HU, (void*)R, (void*)T,
                    (void*)T2, (void*)P, (void*)Gb, (void*)Linv, (void*)theta, (void*)rn, (void*)idx,
                    (void*)keep, (void*)map, (void*)hw, (void*)gw, (void*)crow, (void*)QS, (void*)CS,
                    (void*)CY, (void*)CZ, (void*)lamS})
      hipFree(p);
    small_free(se);
    small_free(sc);
    *this = DavidsonWork();
  }
  // buffers for an N x N matrix, nk Ritz pairs, a basis of msub and a coarse block of M;
  // kept when the held sizes cover the request, else reallocated for the larger of both
  int reserve(int n, int k, int ms, int m) {
    if (have && n == N && k <= nk && ms <= msub && m <= M) return WTB_OK;
    const int n2 = n, k2 = (have && n == N) ? std::max(k, nk) : k;
    const int ms2 = (have && n == N) ? std::max(ms, msub) : ms;
    const int m2 = (have && n == N) ? std::max(m, M) : m;
    release();
    N = n2;
    nk = k2;
    msub = ms2;
    M = m2;
    have = true;
    bool ok = true;
    auto take = [&](auto*& ptr, size_t bytes) {
      if (ok && hipMalloc((void**)&ptr, bytes) != hipSuccess) {
        ptr = nullptr;
        ok = false;
      }
    };
    take(V, sizeof(zc) * (size_t)N * msub);
    take(HV, sizeof(zc) * (size_t)N * msub);
    take(G, sizeof(zc) * (size_t)msub * msub);
    take(S, sizeof(zc) * (size_t)msub * nk);
    take(U, sizeof(zc) * (size_t)N * nk);
    take(HU, sizeof(zc) * (size_t)N * nk);
    take(R, sizeof(zc) * (size_t)N * nk);
    take(T, sizeof(zc) * (size_t)N * nk);
    take(T2, sizeof(zc) * (size_t)N * nk);
    take(P, sizeof(zc) * (size_t)msub * nk);
    take(Gb, sizeof(zc) * (size_t)nk * nk);
    take(Linv, sizeof(zc) * (size_t)nk * nk);
    take(theta, sizeof(double) * msub);
    take(rn, sizeof(double) * nk);
    take(idx, sizeof(int) * nk);
    take(keep, sizeof(int) * nk);
    take(map, sizeof(int) * nk);
    take(hw, hmm_work_bytes(N, nk));
    take(gw, sizeof(zc) * (size_t)8 * msub * nk);
    if (M > 0) {
      take(crow, sizeof(int) * M);
      take(QS, sizeof(zc) * (size_t)M * M);
      take(CS, sizeof(zc) * (size_t)M * std::max(M, nk));
      take(CY, sizeof(zc) * (size_t)M * nk);
      take(CZ, sizeof(zc) * (size_t)M * nk);
      take(lamS, sizeof(double) * M);
    }
    if (ok && small_alloc(se, msub, nk) != WTB_OK) ok = false;
    if (ok && M > 0 && small_alloc(sc, M, M) != WTB_OK) ok = false;
    if (!ok) {
      release();
      set_error("davidson: device allocation failed");
      return WTB_ERR_NOMEM;
    }
    return WTB_OK;
  }
};

DavidsonWork* davidson_work_new() { return new DavidsonWork(); }
void davidson_work_delete(DavidsonWork* w) {
  if (!w) return;
  w->release();
  delete w;
}

int run_davidson(hipStream_t s, int N, const zf* H, int ldh, const double* D, const float* stats,
                 const DavidsonCfg& cfg, const MlpDev* mlp, double* d_W, double* d_resid, zc* d_X,
                 DavidsonInfo& info, DavidsonWork& work) {
  const int nev = cfg.nev;
  const int nk = std::min(N, nev + cfg.guard);
  const int msub = std::min(N, std::max(cfg.max_sub, 2 * nk));
  if (nev < 1 || nev > N) {
    set_error("davidson: nev must lie in 1..N");
    return WTB_ERR_ARG;
  }
  info = DavidsonInfo();
  const int M = std::min(N, std::max(0, cfg.coarse));
  const size_t gws = (size_t)8 * msub * nk;
  int rc = work.reserve(N, nk, msub, M);
  if (rc) return rc;
  // the buffers, named as the algorithm above names them
  zc *V = work.V, *HV = work.HV, *G = work.G, *S = work.S, *U = work.U, *HU = work.HU, *R = work.R,
     *T = work.T, *T2 = work.T2, *P = work.P, *Gb = work.Gb, *Linv = work.Linv;
  double *theta = work.theta, *rn = work.rn;
  int *idx = work.idx, *keep = work.keep, *map = work.map;
  zc* hw = work.hw;
  zc* gw = work.gw;   // K-slice partials of the projections V^H X
  // coarse block of the preconditioner: rows, eigenvectors QS and eigenvalues lamS of
  // H(rows, rows), and three M x nk work blocks
  int* crow = work.crow;
  zc *QS = work.QS, *CS = work.CS, *CY = work.CY, *CZ = work.CZ;
  double* lamS = work.lamS;
  SmallEig& se = work.se;
  SmallEig& sc = work.sc;
  auto cleanup = [&]() {};
#define DAV_TRY(call)      \
  if ((rc = (call))) {     \
    cleanup();             \
    return rc;             \
  }
  const zc one = mkz(1, 0), zero = mkz(0, 0), mone = mkz(-1, 0);
  // start: unit vectors at the nk smallest diagonal entries
  double Dmin = 0.0, fp32_floor = 0.0;
  {
    std::vector<double> hD(N);
    DAV_TRY(hipStreamSynchronize(s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);   // D comes from stream s
    DAV_TRY(hipMemcpy(hD.data(), D, sizeof(double) * N, hipMemcpyDeviceToHost) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    std::vector<int> order(N);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hD[a] < hD[b]; });
    Dmin = hD[order[0]];
    double dmax = 0.0;
    for (double v : hD) dmax = std::max(dmax, std::fabs(v));
    // residual level the FP32 products reach: their rounding grows like sqrt(N) |H|
    fp32_floor = 4e-7 * std::sqrt((double)N) * dmax;
    if (M > 0) {
      // the coarse block: the M states of lowest diagonal energy, where the bound excitons
      // live; H(rows, rows) = QS diag(lamS) QS^H once, on the same small eigensolver
      DAV_TRY(hipMemcpy(crow, order.data(), sizeof(int) * M, hipMemcpyHostToDevice) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      hipLaunchKernelGGL(k_gather_block, dim3((M + 255) / 256, M), dim3(256), 0, s, M, crow, H, ldh, CS);
      DAV_TRY(small_heev(s, sc, M, CS, M, lamS, QS, M, M));
    }
    order.resize(nk);
    DAV_TRY(hipMemcpy(idx, order.data(), sizeof(int) * nk, hipMemcpyHostToDevice) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    hipLaunchKernelGGL(k_unit_cols, dim3((N + 255) / 256, nk), dim3(256), 0, s, N, nk, idx, V, N);
    // with a coarse block of at least nk states, start from its nk lowest eigenvectors
    // instead (the unit vectors' rows lie inside the block, so the scatter replaces them)
    if (M >= nk)
      hipLaunchKernelGGL(k_scatter_rows, dim3((M + 255) / 256, nk), dim3(256), 0, s, M, nk, crow, QS, V, N);
  }
  bool hi = false;   // FP64 products
  auto matvec = [&](int nc, const zc* X, zc* Y) {
    return hi ? hmm_f32_acc64(s, N, nc, H, ldh, X, N, Y, N, hw) : hmm_f32(s, N, nc, H, ldh, X, N, Y, N, hw);
  };
  int m = nk;
  DAV_TRY(matvec(m, V, HV));
  info.matvecs += m;
  DAV_TRY(zgemm(s, 'C', 'N', m, m, N, one, V, N, HV, N, zero, G, msub, gw, gws));
  std::vector<double> h_rn(nk);
  double best_mx = 1e300;
  int since_best = 0;
  const char* tr = getenv("WTB_DAVIDSON_TRACE");   // per-iteration residuals on stderr
  const bool trace = tr && tr[0] == '1';
  std::vector<int> h_keep(nk), h_idx(nk), h_map(nk);   // live across the asynchronous copies
  for (int it = 0; it < cfg.max_iter; ++it) {
    info.iters = it + 1;
    DAV_TRY(small_heev(s, se, m, G, msub, theta, S, msub, nk));
    DAV_TRY(zgemm(s, 'N', 'N', N, nk, m, one, V, N, S, msub, zero, U, N));
    DAV_TRY(zgemm(s, 'N', 'N', N, nk, m, one, HV, N, S, msub, zero, HU, N));
    DAV_TRY(launch_residual(s, N, nk, U, HU, N, theta, R, rn));
    DAV_TRY(hipMemcpyAsync(h_rn.data(), rn, sizeof(double) * nk, hipMemcpyDeviceToHost, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    DAV_TRY(hipStreamSynchronize(s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    double mx = 0.0;
    for (int c = 0; c < nev; ++c) mx = std::max(mx, h_rn[c]);
    info.max_resid = mx;
    if (trace) {
      int nun = 0;
      for (int c = 0; c < nk; ++c) nun += h_rn[c] > cfg.tol;
      fprintf(stderr, "davidson it %d %s m=%d max resid %.3e unconverged %d\n", it, hi ? "fp64" : "fp32", m, mx, nun);
    }
    if (hi) ++info.iters_fp64;
    if (mx <= cfg.tol && hi) {
      info.converged = 1;
      break;
    }
    if (it + 1 == cfg.max_iter) break;
    // FP32 products stall at a residual level set by their rounding (grows like sqrt(N));
    // switch below that level, or when three iterations in a row bring no new best
    // residual (by a factor 0.7) once the residuals are below 1e-2
    if (mx < 0.7 * best_mx) {
      best_mx = mx;
      since_best = 0;
    } else {
      ++since_best;
    }
    const bool stall = since_best >= 3;
    if (!hi && (mx <= std::max(cfg.switch_tol, fp32_floor) || mx <= cfg.tol || (stall && mx < 1e-2))) {
      // switch to FP64 products: restart on the Ritz vectors with H U recomputed
      hi = true;
      DAV_TRY(hipMemcpyAsync(V, U, sizeof(zc) * (size_t)N * nk, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      DAV_TRY(matvec(nk, V, HV));
      info.matvecs += nk;
      info.matvecs_fp64 += nk;
      m = nk;
      DAV_TRY(zgemm(s, 'C', 'N', m, m, N, one, V, N, HV, N, zero, G, msub, gw, gws));
      best_mx = 1e300;
      since_best = 0;
      continue;
    }
    int nq = 0;
    for (int c = 0; c < nk; ++c)
      if (h_rn[c] > cfg.tol) h_idx[nq++] = c;
    if (m + nq > msub) {
      // restart on the Ritz vectors
      DAV_TRY(hipMemcpyAsync(V, U, sizeof(zc) * (size_t)N * nk, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      DAV_TRY(hipMemcpyAsync(HV, HU, sizeof(zc) * (size_t)N * nk, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      hipLaunchKernelGGL(k_diag_set, dim3((nk + 255) / 256, nk), dim3(256), 0, s, nk, theta, G, msub);
      m = nk;
    }
    DAV_TRY(hipMemcpyAsync(idx, h_idx.data(), sizeof(int) * nq, hipMemcpyHostToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    if (mlp && mlp->w && !hi) {   // the FP64 polish uses the diagonal preconditioner
      DAV_TRY(launch_precond_mlp(s, N, nq, idx, R, N, theta, D, stats, Dmin, *mlp, T, N));
    } else {
      DAV_TRY(launch_precond_diag(s, N, nq, idx, R, N, theta, D, cfg.dmin, T, N));
    }
    if (M > 0) {
      // on the coarse rows: T = QS (lamS - theta)^-1 QS^H R, the block's exact inverse
      const dim3 gq((M + 255) / 256, nq);
      hipLaunchKernelGGL(k_gather_rows, gq, dim3(256), 0, s, M, nq, crow, idx, R, N, CS);
      DAV_TRY(zgemm(s, 'C', 'N', M, nq, M, one, QS, M, CS, M, zero, CY, M, gw, gws));
      hipLaunchKernelGGL(k_coarse_scale, gq, dim3(256), 0, s, M, nq, idx, lamS, theta, cfg.coarse_dmin, CY);
      DAV_TRY(zgemm(s, 'N', 'N', M, nq, M, one, QS, M, CY, M, zero, CZ, M, gw, gws));
      hipLaunchKernelGGL(k_scatter_rows, gq, dim3(256), 0, s, M, nq, crow, CZ, T, N);
    }
    // T -= V (V^H T), twice
    for (int pass = 0; pass < 2; ++pass) {
      DAV_TRY(zgemm(s, 'C', 'N', m, nq, N, one, V, N, T, N, zero, P, msub, gw, gws));
      DAV_TRY(zgemm(s, 'N', 'N', N, nq, m, mone, V, N, P, msub, one, T, N));
    }
    // Cholesky QR twice (T <- T L^-H with T^H T = L L^H, columns of tiny pivots dropped),
    // with a projection against V in between: one pass squares the condition number
    int nn = nq;
    for (int pass = 0; pass < 2 && nn > 0; ++pass) {
      if (pass == 1) {
        DAV_TRY(zgemm(s, 'C', 'N', m, nn, N, one, V, N, T, N, zero, P, msub, gw, gws));
        DAV_TRY(zgemm(s, 'N', 'N', N, nn, m, mone, V, N, P, msub, one, T, N));
      }
      DAV_TRY(zgemm(s, 'C', 'N', nn, nn, N, one, T, N, T, N, zero, Gb, nk, gw, gws));
      DAV_TRY(launch_chol_inv(s, nn, Gb, nk, Linv, keep, 1e-10));
      DAV_TRY(zgemm(s, 'N', 'C', N, nn, nn, one, T, N, Linv, nn, zero, T2, N));
      DAV_TRY(hipMemcpyAsync(h_keep.data(), keep, sizeof(int) * nn, hipMemcpyDeviceToHost, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      DAV_TRY(hipStreamSynchronize(s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      int nk2 = 0;
      for (int q = 0; q < nn; ++q)
        if (h_keep[q]) h_map[nk2++] = q;
      nn = nk2;
      if (nn == 0) break;
      DAV_TRY(hipMemcpyAsync(map, h_map.data(), sizeof(int) * nn, hipMemcpyHostToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
      hipLaunchKernelGGL(k_gather_cols, dim3((N + 255) / 256, nn), dim3(256), 0, s, N, nn, map, T2, N, T, N);
    }
    if (nn == 0) break;   // no new direction: stagnation
    zc* Vn = V + (size_t)N * m;
    DAV_TRY(hipMemcpyAsync(Vn, T, sizeof(zc) * (size_t)N * nn, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
    DAV_TRY(matvec(nn, Vn, HV + (size_t)N * m));
    info.matvecs += nn;
    if (hi) info.matvecs_fp64 += nn;
    // G(0:m+nn, m:m+nn) = V^H (H Vn); mirror into G(m:m+nn, 0:m)
    DAV_TRY(zgemm(s, 'C', 'N', m + nn, nn, N, one, V, N, HV + (size_t)N * m, N, zero, G + (size_t)msub * m, msub, gw, gws));
    hipLaunchKernelGGL(k_mirror_block, dim3(nn, (m + 255) / 256), dim3(256), 0, s, m, m + nn, G, msub);
    m += nn;
  }
  DAV_TRY(hipMemcpyAsync(d_W, theta, sizeof(double) * nev, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
  DAV_TRY(hipMemcpyAsync(d_resid, rn, sizeof(double) * nev, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
  if (d_X) DAV_TRY(hipMemcpyAsync(d_X, U, sizeof(zc) * (size_t)N * nev, hipMemcpyDeviceToDevice, s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
  DAV_TRY(hipStreamSynchronize(s) == hipSuccess ? WTB_OK : WTB_ERR_HIP);
#undef DAV_TRY
  cleanup();
  return WTB_OK;
}

}  // namespace wtb
