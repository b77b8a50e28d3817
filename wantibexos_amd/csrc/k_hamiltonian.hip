// Stages (a) and (b): k grid, H(k) phase GEMM on MFMA, batched per-k Jacobi
// eigensolver, band window, interband dipoles.
//
// Reference: subroutines/mhkpack_subs.f90:359-432 (k grid, recvec),
// subroutines/hamiltonian_tb.f90:2-137,371-388 (eigsys, fermilvl),
// subroutines/optics.f90:1-156 (hlm, optsp), general_subs.f90:69-112
// (sandwich, prodintsq), subprograms/bse_solver-tool-diel.f90:257-377.
#include "wtb_kernels.h"
#include <algorithm>

namespace wtb {

// pi = acos(-1.) evaluated in single precision (mhkpack_subs.f90:413)
__device__ __constant__ double kPi32 = 3.14159274101257324;

// recvec (mhkpack_subs.f90:409-432) with the reference's operation order.
__device__ inline void recvec_dev(const double* rl, double b[3][3]) {
#pragma clang fp contract(off)
  // rl[i + 3*c] = component c of a_i
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) a[i][c] = rl[i + 3 * c];
  auto cross = [](const double* v1, const double* v2, double* vx) {
    vx[0] = (v1[1] * v2[2]) - (v1[2] * v2[1]);
    vx[1] = (v1[2] * v2[0]) - (v1[0] * v2[2]);
    vx[2] = (v1[0] * v2[1]) - (v1[1] * v2[0]);
  };
  double v23[3], v31[3], v12[3];
  cross(a[1], a[2], v23);
  cross(a[2], a[0], v31);
  cross(a[0], a[1], v12);
  double vol = fabs((a[0][0] * v23[0]) + (a[0][1] * v23[1]) + (a[0][2] * v23[2]));
  double f = (2.0 * kPi32) / vol;
  for (int c = 0; c < 3; ++c) {
    b[0][c] = f * v23[c];
    b[1][c] = f * v31[c];
    b[2][c] = f * v12[c];
  }
}

// monhkhorst_pack: kpt(nk,3), index i*n2*n3 + j*n3 + l
// dq (optional): a q-path row (coordinate, qx, qy, qz); the grid is then k + q with q added
// first, as monhkhorst_packq does (mhkpack_subs.f90:305-356)
__global__ void k_kgrid(const double* __restrict__ rl, int n1, int n2, int n3, double s1, double s2,
                        double s3, const double* __restrict__ dq, double* __restrict__ kpt) {
#pragma clang fp contract(off)
  int nk = n1 * n2 * n3;
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nk) return;
  double b[3][3];
  recvec_dev(rl, b);
  int i = idx / (n2 * n3), j = (idx / n3) % n2, l = idx % n3;
  for (int c = 0; c < 3; ++c) {
    double ksh = b[0][c] * s1 + b[1][c] * s2 + b[2][c] * s3;
    double v;
    if (dq)
      v = dq[1 + c] + (b[0][c] / (double)n1) * (double)i + (b[1][c] / (double)n2) * (double)j +
          (b[2][c] / (double)n3) * (double)l + ksh;
    else
      v = (b[0][c] / (double)n1) * (double)i + (b[1][c] / (double)n2) * (double)j +
          (b[2][c] / (double)n3) * (double)l + ksh;
    kpt[idx + (size_t)nk * c] = v;
  }
}

// Phase operand of the H(k) GEMM, row-major A[4*nk][nvec]:
//   rows [0,nk):        e^{i fl32(k.R)}                       (hamiltonian_tb.f90:62-77)
//   rows [a*nk,(a+1)nk) (i fl32(R_a)) e^{i fl32(k.R)}, a=1..3  (optics.f90:39-80)
__global__ void k_phase(const double* __restrict__ kpt, int nk, const int* __restrict__ rvec,
                        int nvec, const double* __restrict__ rl, zc* __restrict__ A) {
#pragma clang fp contract(off)
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nk * nvec) return;
  int k = idx / nvec, r = idx % nvec;
  double R1 = rvec[r], R2 = rvec[r + nvec], R3 = rvec[r + 2 * nvec];
  double rc[3];
  for (int c = 0; c < 3; ++c)
    rc[c] = (R1 * rl[0 + 3 * c] + R2 * rl[1 + 3 * c]) + R3 * rl[2 + 3 * c];
  double kx = kpt[k], ky = kpt[k + nk], kz = kpt[k + 2 * nk];
  double x = kx * rc[0] + ky * rc[1] + kz * rc[2];
  double ph = (double)(float)x;          // cmplx(0.0, x) has default (single) kind
  double sn, cs;
  sincos(ph, &sn, &cs);
  A[(size_t)k * nvec + r] = mkz(cs, sn);
  for (int a = 0; a < 3; ++a) {
    double ra = (double)(float)rc[a];
    // (0, ra) * (cs, sn)
    A[((size_t)(a + 1) * nk + k) * nvec + r] = mkz(0.0 * cs - ra * sn, 0.0 * sn + ra * cs);
  }
}

// Stacked H(R)/ndeg, row-major B[nvec][nw*nw], column e = n*nw + m <-> H(R)_{mn}
__global__ void k_hr_stack(const double* __restrict__ hop, const double* __restrict__ ihop,
                           const int* __restrict__ ff, int nvec, int nw, zc* __restrict__ B) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  int nw2 = nw * nw;
  if (idx >= nvec * nw2) return;
  int r = idx / nw2, e = idx % nw2;
  int m = e % nw, n = e / nw;
  size_t src = r + (size_t)nvec * (m + (size_t)nw * n);
  double f = 1.0 / (double)ff[r];
  B[idx] = mkz(hop[src] * f, ihop[src] * f);
}

// C[M][N] = A[M][K] * B[K][N] (row-major, complex), MFMA f64 16x16x4.
// 256 threads = 4 waves, 64x64 output tile, each wave a 32x32 quarter.
__global__ __launch_bounds__(256) void k_zgemm_rr(const zc* __restrict__ A, const zc* __restrict__ B,
                                                  zc* __restrict__ C, int M, int N, int K) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = blockIdx.x * 64 + (wave >> 1) * 32;
  const int col0 = blockIdx.y * 64 + (wave & 1) * 32;
  const int lr = lane & 15, lk = lane >> 4;
  d4 cr[2][2], ci[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      cr[a][b] = d4{0, 0, 0, 0};
      ci[a][b] = d4{0, 0, 0, 0};
    }
  for (int k0 = 0; k0 < K; k0 += 4) {
    int k = k0 + lk;
    zc av[2], bv[2];
    for (int a = 0; a < 2; ++a) {
      int r = row0 + a * 16 + lr;
      av[a] = (r < M && k < K) ? A[(size_t)r * K + k] : mkz(0, 0);
      int c = col0 + a * 16 + lr;
      bv[a] = (c < N && k < K) ? B[(size_t)k * N + c] : mkz(0, 0);
    }
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b) {
        cr[a][b] = mfma16x16x4(av[a].re, bv[b].re, cr[a][b]);
        cr[a][b] = mfma16x16x4(-av[a].im, bv[b].im, cr[a][b]);
        ci[a][b] = mfma16x16x4(av[a].re, bv[b].im, ci[a][b]);
        ci[a][b] = mfma16x16x4(av[a].im, bv[b].re, ci[a][b]);
      }
  }
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int q = 0; q < 4; ++q) {
        int r = row0 + a * 16 + lk + 4 * q;
        int c = col0 + b * 16 + lr;
        if (r < M && c < N) C[(size_t)r * N + c] = mkz(cr[a][b][q], ci[a][b][q]);
      }
}

// ---------------------------------------------------------------------------
// Batched Hermitian eigensolver: cyclic two-sided Jacobi with round-robin
// ordering, one wavefront per k-point, matrix and vectors resident in LDS.
// Output follows eigsys: ascending eigenvalues, eigenvectors in ZHEEV('L')'s
// gauge (orbital-1 component real, chosen >= 0), nocp = #(E <= 0), scissor on
// E above the Fermi level.
// ---------------------------------------------------------------------------
__device__ inline int rr_player(int pos, int round, int ne) {
  return pos == 0 ? 0 : ((pos - 1 + round) % (ne - 1)) + 1;
}

__device__ __forceinline__ void jacobi_body(int k, const zc* __restrict__ Hk, int nk, int nw, double ef32,
                                            double scs, int nc, int nv, double* __restrict__ Eall,
                                            zc* __restrict__ Uout, int* __restrict__ nocp_out,
                                            double* __restrict__ eigv, zc* __restrict__ vec,
                                            int* __restrict__ status) {
  extern __shared__ unsigned char smem[];
  const int n = nw;
  const int ne = n + (n & 1);
  const int np = ne / 2;
  const int ld = n;
  zc* A = reinterpret_cast<zc*>(smem);
  zc* V = A + n * ld;
  zc* rs = V + n * ld;                        // s*phi per pair
  double* rcv = reinterpret_cast<double*>(rs + np);
  int* rp = reinterpret_cast<int*>(rcv + np);
  int* rq = rp + np;
  int* flag = rq + np;                        // [0]: rotations this sweep
  double* lam = reinterpret_cast<double*>(flag + 2);
  int* rank = reinterpret_cast<int*>(lam + n);

  const int lane = threadIdx.x;
  const zc* H = Hk + (size_t)k * n * n;
  double amax = 0.0;
  for (int e = lane; e < n * n; e += 64) {
    int m = e % n, c = e / n;
    zc h = H[e];
    if (m == c) h = mkz(h.re + ef32, h.im);   // htb + fermimt, fermimt(i,i) = cmplx(-efermi,0.0)
    A[m + c * ld] = h;
    V[m + c * ld] = mkz(m == c ? 1.0 : 0.0, 0.0);
    amax = fmax(amax, fabs(h.re) + fabs(h.im));
  }
  for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off));
  __syncthreads();

  const double tiny = 1e-300 + amax * 1e-30;
  int sweep = 0;
  for (; sweep < 40; ++sweep) {
    if (lane == 0) flag[0] = 0;
    __syncthreads();
    for (int round = 0; round < ne - 1; ++round) {
      if (lane < np) {
        int p = rr_player(lane, round, ne), q = rr_player(ne - 1 - lane, round, ne);
        if (p > q) { int t = p; p = q; q = t; }
        rp[lane] = -1;
        if (q < n) {
          double app = A[p + p * ld].re, aqq = A[q + q * ld].re;
          zc apq = A[p + q * ld];
          double g = sqrt(norm2(apq));
          if (g > tiny && g > 1.1102230246251565e-16 * 0.5 * sqrt(fabs(app) * fabs(aqq))) {
            double tau = (aqq - app) / (2.0 * g);
            double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            double c = 1.0 / sqrt(1.0 + t * t);
            double s = t * c;
            rp[lane] = p;
            rq[lane] = q;
            rcv[lane] = c;
            rs[lane] = mkz(s * apq.re / g, s * apq.im / g);
            flag[0] = 1;
          }
        }
      }
      __syncthreads();
      // columns: A <- A J, V <- V J
      for (int it = lane; it < np * n; it += 64) {
        int j = it / n, r = it % n;
        int p = rp[j];
        if (p < 0) continue;
        int q = rq[j];
        double c = rcv[j];
        zc sp = rs[j];
        zc ap = A[r + p * ld], aq = A[r + q * ld];
        A[r + p * ld] = c * ap - conj(sp) * aq;
        A[r + q * ld] = sp * ap + c * aq;
        zc vp = V[r + p * ld], vq = V[r + q * ld];
        V[r + p * ld] = c * vp - conj(sp) * vq;
        V[r + q * ld] = sp * vp + c * vq;
      }
      __syncthreads();
      // rows: A <- J^H A
      for (int it = lane; it < np * n; it += 64) {
        int j = it / n, r = it % n;
        int p = rp[j];
        if (p < 0) continue;
        int q = rq[j];
        double c = rcv[j];
        zc sp = rs[j];
        zc ap = A[p + r * ld], aq = A[q + r * ld];
        zc np_ = c * ap - sp * aq;
        zc nq_ = conj(sp) * ap + c * aq;
        if (r == q) np_ = mkz(0, 0);
        if (r == p) nq_ = mkz(0, 0);
        if (r == p) np_.im = 0.0;
        if (r == q) nq_.im = 0.0;
        A[p + r * ld] = np_;
        A[q + r * ld] = nq_;
      }
      __syncthreads();
    }
    if (flag[0] == 0) break;
    __syncthreads();
  }
  if (sweep >= 40 && lane == 0) atomicOr(status, 1 << WTB_ERR_EIG);

  // sort ascending (rank by counting, ties by index)
  for (int j = lane; j < n; j += 64) lam[j] = A[j + j * ld].re;
  __syncthreads();
  int nle = 0;
  for (int j = lane; j < n; j += 64) {
    double lj = lam[j];
    int rk = 0;
    for (int i = 0; i < n; ++i) {
      double li = lam[i];
      rk += (li < lj) || (li == lj && i < j);
    }
    rank[j] = rk;
    nle += (lj <= 0.0);
  }
  for (int off = 32; off > 0; off >>= 1) nle += __shfl_xor(nle, off);
  __syncthreads();
  const int nocp = nle;                        // first E > 0 is at index nocp (0-based)
  if (lane == 0) {
    nocp_out[k] = nocp;
    if (nocp >= n) atomicOr(status, 1 << WTB_ERR_NOCP);
    else if (nv > nocp) atomicOr(status, 1 << WTB_ERR_NVAL);
    else if (nc > n - nocp) atomicOr(status, 1 << WTB_ERR_NCOND);
  }
  const int nb = nc + nv;
  for (int j = lane; j < n; j += 64) {
    int rk = rank[j];
    double e = lam[j] + (rk >= nocp ? scs : 0.0);
    if (Eall) Eall[(size_t)k * n + rk] = e;
    zc x1 = V[0 + j * ld];
    double mag = sqrt(norm2(x1));
    zc ph = mag > 0 ? mkz(x1.re / mag, -x1.im / mag) : mkz(1.0, 0.0);
    int b = rk - (nocp - nv);
    bool sel = (b >= 0 && b < nb && eigv != nullptr);
    if (sel) eigv[(size_t)k * nb + b] = e;
    for (int m = 0; m < n; ++m) {
      zc u = V[m + j * ld] * ph;
      if (m == 0) u = mkz(mag, 0.0);
      // autovetores(i,m) = component m of eigenvector i, per k
      if (Uout) Uout[((size_t)k * n + m) * n + rk] = u;
      if (sel) vec[((size_t)k * nb + b) * n + m] = u;
    }
  }
}
__global__ __launch_bounds__(64) void k_jacobi(const zc* __restrict__ Hk, int nk, int nw,
                                               double ef32, double scs, int nc, int nv,
                                               double* __restrict__ Eall, zc* __restrict__ Uout,
                                               int* __restrict__ nocp_out, double* __restrict__ eigv,
                                               zc* __restrict__ vec, int* __restrict__ status) {
  jacobi_body(blockIdx.x, Hk, nk, nw, ef32, scs, nc, nv, Eall, Uout, nocp_out, eigv, vec, status);
}

// the k-points of up to 64 models in one launch: workgroup b belongs to the model whose
// range [pre[i], pre[i+1]) holds it and runs that model's k-point b - pre[i]
struct JacobiBatch {
  const JacobiItem* items;   // device array
  int n;
  int m[64];
  int pre[65];
};
__global__ __launch_bounds__(64) void k_jacobi_b(JacobiBatch B) {
  int i = 0;
  while (i + 1 < B.n && B.pre[i + 1] <= (int)blockIdx.x) ++i;
  const JacobiItem it = B.items[B.m[i]];
  jacobi_body((int)blockIdx.x - B.pre[i], it.hk, it.nk, it.nw, it.ef32, it.scs, it.nc, it.nv, it.eall, it.uout,
              it.nocp, it.eigv, it.vec, it.status);
}

// ---------------------------------------------------------------------------
// Dipoles (optsp, optics.f90:102-156): for every BSE state I = (k,c,v),
//   h_a(I) = Re(1/(Ec - Ev + i fl32(sme))) * conj(U_c) . (dH_a(k) U_v)
// and the direct gap min_k(E_{nv+1} - E_{nv}) (bse_solver-tool-diel.f90:305-311).
// One wavefront per k.  dHk = rows [nk, 4nk) of the phase GEMM output.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_dipoles(const zc* __restrict__ Hk4, int nk, int nw, int nc,
                                                int nv, double sme32, const double* __restrict__ eigv,
                                                const zc* __restrict__ vec, zc* __restrict__ dip,
                                                double* __restrict__ gapk) {
  extern __shared__ unsigned char smem[];
  const int n = nw, nb = nc + nv;
  zc* dH = reinterpret_cast<zc*>(smem);           // n*n
  zc* U = dH + n * n;                             // nb*n
  zc* y = U + nb * n;                             // nv*n
  const int k = blockIdx.x, lane = threadIdx.x;
  const size_t N = (size_t)nk * nc * nv;
  for (int e = lane; e < nb * n; e += 64) U[e] = vec[(size_t)k * nb * n + e];
  if (lane == 0) gapk[k] = eigv[(size_t)k * nb + nv] - eigv[(size_t)k * nb + nv - 1];
  for (int a = 0; a < 3; ++a) {
    __syncthreads();
    const zc* src = Hk4 + ((size_t)(a + 1) * nk + k) * n * n;
    for (int e = lane; e < n * n; e += 64) dH[e] = src[e];
    __syncthreads();
    // matvec (general_subs.f90:132-160): y_v(m) = sum_j dH(m,j) U_v(j), sequential in j
    for (int it = lane; it < nv * n; it += 64) {
      int v = it / n, m = it % n;
      zc acc = mkz(0, 0);
      for (int j = 0; j < n; ++j) acc = acc + dH[m + j * n] * U[v * n + j];
      y[v * n + m] = acc;
    }
    __syncthreads();
    for (int it = lane; it < nc * nv; it += 64) {
      int c = it / nv, v = it % nv;
      zc acc = mkz(0, 0);
      for (int m = 0; m < n; ++m) acc = acc + conj(U[(nv + c) * n + m]) * y[v * n + m];
      double ec = eigv[(size_t)k * nb + nv + c], ev = eigv[(size_t)k * nb + v];
      double ar = ec - ev, ai = sme32;
      double factor = ar / (ar * ar + ai * ai);     // REAL factor = 1./esp keeps Re(1/esp)
      size_t I = (size_t)k * nc * nv + (size_t)c * nv + v;
      dip[(size_t)a * N + I] = factor * acc;
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
int launch_kgrid(hipStream_t s, const double* d_rlat, const int ng[3], const double sh[3],
                 double* d_kpt, const double* d_q) {
  int nk = ng[0] * ng[1] * ng[2];
  hipLaunchKernelGGL(k_kgrid, dim3((nk + 127) / 128), dim3(128), 0, s, d_rlat, ng[0], ng[1], ng[2],
                     sh[0], sh[1], sh[2], d_q, d_kpt);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_hk(hipStream_t s, const DevModel& dm, const double* d_kpt, int nk, zc* d_phase,
              zc* d_hk4) {
  int nvec = dm.nvec, nw = dm.nw;
  hipLaunchKernelGGL(k_phase, dim3((nk * nvec + 255) / 256), dim3(256), 0, s, d_kpt, nk, dm.rvec,
                     nvec, dm.rlat, d_phase);
  WTB_LAUNCH_CHECK();
  int M = 4 * nk, N = nw * nw;
  hipLaunchKernelGGL(k_zgemm_rr, dim3((M + 63) / 64, (N + 63) / 64), dim3(256), 0, s, d_phase,
                     dm.hstack, d_hk4, M, N, nvec);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_hr_stack(hipStream_t s, const double* hop, const double* ihop, const int* ff, int nvec,
                    int nw, zc* B) {
  int tot = nvec * nw * nw;
  hipLaunchKernelGGL(k_hr_stack, dim3((tot + 255) / 256), dim3(256), 0, s, hop, ihop, ff, nvec, nw, B);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

size_t jacobi_lds_bytes(int nw) {
  int np = (nw + 1) / 2;
  return sizeof(zc) * (2 * nw * nw + np) + sizeof(double) * np + sizeof(int) * (2 * np + 2) +
         sizeof(double) * nw + sizeof(int) * nw + 64;
}

int launch_jacobi(hipStream_t s, const zc* d_hk, int nk, int nw, double ef32, double scs, int nc,
                  int nv, double* d_eall, zc* d_uout, int* d_nocp, double* d_eigv, zc* d_vec,
                  int* d_status) {
  size_t lds = jacobi_lds_bytes(nw);
  hipLaunchKernelGGL(k_jacobi, dim3(nk), dim3(64), lds, s, d_hk, nk, nw, ef32, scs, nc, nv, d_eall,
                     d_uout, d_nocp, d_eigv, d_vec, d_status);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_jacobi_batch(hipStream_t s, const JacobiItem* d_items, const std::vector<JacobiItem>& items) {
  int nwmax = 0;
  for (const auto& it : items) nwmax = std::max(nwmax, it.nw);
  const size_t lds = jacobi_lds_bytes(nwmax);
  for (size_t c0 = 0; c0 < items.size(); c0 += 64) {
    JacobiBatch B;
    B.items = d_items;
    B.n = 0;
    int tot = 0;
    for (size_t i = c0; i < items.size() && i < c0 + 64; ++i) {
      B.m[B.n] = (int)i;
      B.pre[B.n] = tot;
      tot += items[i].nk;
      ++B.n;
    }
    B.pre[B.n] = tot;
    if (tot > 0) hipLaunchKernelGGL(k_jacobi_b, dim3(tot), dim3(64), lds, s, B);
  }
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int launch_dipoles(hipStream_t s, const zc* d_hk4, int nk, int nw, int nc, int nv, double sme32,
                   const double* d_eigv, const zc* d_vec, zc* d_dip, double* d_gapk) {
  size_t lds = sizeof(zc) * (nw * nw + (nc + nv) * nw + nv * nw);
  hipLaunchKernelGGL(k_dipoles, dim3(nk), dim3(64), lds, s, d_hk4, nk, nw, nc, nv, sme32, d_eigv,
                     d_vec, d_dip, d_gapk);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
