// Stage (d1): blocked Householder reduction of a Hermitian matrix (lower
// triangle, column-major, ld) to a real symmetric tridiagonal (d, e), the
// algorithm of LAPACK ZHETRD/ZLATRD/ZHER2K for UPLO = 'L' (the reference
// calls ZHEEV, bse_solver-tool-diel.f90:485-488, whose first step this is).
//
// Per reduced column j (panel column jj = j - p) three kernels run:
//   A  column update with the panel's V, W; |x|^2 partials; d_j; and the
//      completion of W(:,jj-1) from the previous column's u.
//   B  reflector (ZLARFG) from the partials, then y = A22 v on lower tiles of
//      64x64: each tile is read once and serves A v (rows) and A^H v (cols);
//      diagonal tiles also produce W^H v, V^H v and v^T g partials.
//   C  u = y - V (W^H v) - W (V^H v); u^H v partials; g <- g - tau conj(v)(v^T g).
// After a panel: W(:,last) and the trailing update A22 -= V W^H + W V^H on
// MFMA (lower tiles only).  g accumulates Q^T g, Q = H(1)...H(N-1), so that
// the eigenvectors X = Q Z never have to be formed for X^T h.
// All reductions are written as partials and summed in a fixed order, so the
// result is reproducible run to run.
#include "wtb_kernels.h"

namespace wtb {

#define TB 64   // tile / row-block edge

struct TriArgs {
  zc* H;
  int ld, N, nb;
  zc* W;       // N x nb, ld N
  zc* Vp;      // N x nb, ld N
  zc* u;       // N
  zc* tau;     // N
  double* d;
  double* e;
  double* rnorm;   // [nblk]
  zc* ruv;         // [nblk]
  zc* rt1;         // [nblk][nb]
  zc* rt2;         // [nblk][nb]
  zc* rvg;         // [nblk][nrhs]
  zc* ypart;       // [nblk][nblk][TB]
  zc* g;           // N x nrhs
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
__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// sum of partials p[b*stride] for b in [b0, nb): every lane of the wave gets it
__device__ inline zc sum_partials(const zc* p, int b0, int nbk, int stride) {
  zc acc = mkz(0, 0);
  for (int b = b0 + (int)(threadIdx.x & 63); b < nbk; b += 64) acc += p[(size_t)b * stride];
  return wave_sum(acc);
}
__device__ inline double sum_partials(const double* p, int b0, int nbk) {
  double acc = 0;
  for (int b = b0 + (int)(threadIdx.x & 63); b < nbk; b += 64) acc += p[b];
  return wave_sum(acc);
}

// W(r, jj-1) = tau_{j-1} u(r) + alpha' v_{j-1}(r),  alpha' = -1/2 |tau|^2 (u^H v)
__device__ inline zc w_complete(const TriArgs& a, int r, int jjm1, zc taup, zc alphap) {
  return taup * a.u[r] + alphap * a.Vp[(size_t)jjm1 * a.N + r];
}

// ---- kernel A -----------------------------------------------------------------
__global__ __launch_bounds__(64) void k_tri_colupd(TriArgs a, int j, int jj, int bprev0) {
  __shared__ zc sWj[64], sVj[64];
  const int lane = threadIdx.x;
  const int blk = j / TB + blockIdx.x;
  const int r = blk * TB + lane;
  const int N = a.N;
  zc taup = mkz(0, 0), alphap = mkz(0, 0);
  if (jj > 0) {
    zc uv = sum_partials(a.ruv, bprev0, a.nblk, 1);
    taup = a.tau[j - 1];
    double t2 = norm2(taup);
    alphap = mkz(-0.5 * t2 * uv.re, -0.5 * t2 * uv.im);
  }
  // row j of the panel's W and V (W(j, jj-1) completed here as well)
  if (lane < jj) {
    zc wj = (lane == jj - 1) ? w_complete(a, j, jj - 1, taup, alphap) : a.W[(size_t)lane * N + j];
    sWj[lane] = wj;
    sVj[lane] = a.Vp[(size_t)lane * N + j];
  }
  __syncthreads();
  double nrm = 0.0;
  if (r >= j && r < N) {
    if (jj > 0) a.W[(size_t)(jj - 1) * N + r] = w_complete(a, r, jj - 1, taup, alphap);
    zc x = a.H[(size_t)r + (size_t)a.ld * j];
    // A(j:,j) -= V(j:,0:jj) conj(W(j,0:jj))^T ; then -= W(j:,0:jj) conj(V(j,0:jj))^T
    for (int c = 0; c < jj; ++c) {
      zc v = a.Vp[(size_t)c * N + r];
      x = x - v * conj(sWj[c]);
    }
    for (int c = 0; c < jj; ++c) {
      zc w = (c == jj - 1) ? w_complete(a, r, jj - 1, taup, alphap) : a.W[(size_t)c * N + r];
      x = x - w * conj(sVj[c]);
    }
    if (r == j) {
      x.im = 0.0;
      a.d[j] = x.re;
    }
    a.H[(size_t)r + (size_t)a.ld * j] = x;
    if (r >= j + 2) nrm = norm2(x);
  }
  nrm = wave_sum(nrm);
  if (lane == 0) a.rnorm[blk] = nrm;
}

// reflector of column j (ZLARFG): beta, tau, scale s with v = [1; s x]
struct Refl {
  zc tau;
  double beta;
  zc s;
  bool ident;
};
__device__ inline Refl make_refl(const TriArgs& a, int j) {
  double xn2 = sum_partials(a.rnorm, j / TB, a.nblk);
  zc alpha = a.H[(size_t)(j + 1) + (size_t)a.ld * j];
  Refl R;
  if (xn2 == 0.0 && alpha.im == 0.0) {
    R.tau = mkz(0, 0);
    R.beta = alpha.re;
    R.s = mkz(0, 0);
    R.ident = true;
    return R;
  }
  double beta = -copysign(sqrt(alpha.re * alpha.re + alpha.im * alpha.im + xn2), alpha.re);
  R.beta = beta;
  R.tau = mkz((beta - alpha.re) / beta, -alpha.im / beta);
  zc den = mkz(alpha.re - beta, alpha.im);
  double dd = den.re * den.re + den.im * den.im;
  R.s = mkz(den.re / dd, -den.im / dd);
  R.ident = false;
  return R;
}
__device__ inline zc refl_v(const TriArgs& a, const Refl& R, int j, int r) {
  if (r == j + 1) return mkz(1.0, 0.0);
  if (r <= j || r >= a.N || R.ident) return mkz(0, 0);
  return R.s * a.H[(size_t)r + (size_t)a.ld * j];
}

// ---- kernel B -----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tri_hemv(TriArgs a, int j, int jj) {
  __shared__ zc vR[TB], vC[TB];
  __shared__ zc yw[4][TB];
  __shared__ zc yc[TB];
  const int N = a.N;
  const int b0 = (j + 1) / TB;
  const int nt = a.nblk - b0;
  int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  int tj = t - ti * (ti + 1) / 2;
  (void)nt;
  const int bi = b0 + ti, bj = b0 + tj;
  const Refl R = make_refl(a, j);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int R0 = bi * TB, C0 = bj * TB;
  if (tid < TB) vR[tid] = refl_v(a, R, j, R0 + tid);
  else if (tid < 2 * TB) vC[tid - TB] = refl_v(a, R, j, C0 + tid - TB);
  __syncthreads();
  const int r = R0 + lane;
  const bool rvalid = r > j && r < N;
  zc acc = mkz(0, 0);
  const zc vr = vR[lane];
  for (int cc = wave; cc < TB; cc += 4) {
    const int c = C0 + cc;
    const bool cvalid = c > j && c < N;
    zc x = mkz(0, 0);
    bool lower = (bi != bj) || (lane >= cc);
    if (rvalid && cvalid && lower) {
      x = a.H[(size_t)r + (size_t)a.ld * c];
      if (bi == bj && lane == cc) x.im = 0.0;       // diagonal: real part (ZHEMV)
    }
    acc = fma_z(x, vC[cc], acc);
    // transposed contribution: strictly lower elements only
    zc tcon = mkz(0, 0);
    if (bi != bj || lane > cc) tcon = conj(x) * vr;
    tcon = wave_sum(tcon);
    if (lane == 0) yc[cc] = tcon;
  }
  yw[wave][lane] = acc;
  __syncthreads();
  zc* yp = a.ypart;
  if (tid < TB) {
    zc s = yw[0][tid] + yw[1][tid] + yw[2][tid] + yw[3][tid];
    if (bi == bj) s += yc[tid];
    yp[((size_t)bi * a.nblk + bj) * TB + tid] = s;
  } else if (tid < 2 * TB && bi != bj) {
    yp[((size_t)bj * a.nblk + bi) * TB + (tid - TB)] = yc[tid - TB];
  }
  if (bi == bj) {
    // reflector storage, tau/e, and the small partial products for this row block
    if (tid < TB) {
      int rr = R0 + tid;
      if (rr < N) a.Vp[(size_t)jj * N + rr] = vR[tid];
    }
    if (bi == b0 && tid == 0) {
      a.tau[j] = R.tau;
      a.e[j] = R.beta;
    }
    // t1[c] = sum_r conj(W(r,c)) v(r), t2[c] = sum_r conj(V(r,c)) v(r), vg[q] = sum_r v(r) g(r,q)
    const int nsm = 2 * jj + a.nrhs;
    for (int q = wave; q < nsm; q += 4) {
      zc s = mkz(0, 0);
      if (rvalid) {
        if (q < jj) s = conj(a.W[(size_t)q * N + r]) * vr;
        else if (q < 2 * jj) s = conj(a.Vp[(size_t)(q - jj) * N + r]) * vr;
        else s = vr * a.g[(size_t)(q - 2 * jj) * N + r];
      }
      s = wave_sum(s);
      if (lane == 0) {
        if (q < jj) a.rt1[(size_t)bi * a.nb + q] = s;
        else if (q < 2 * jj) a.rt2[(size_t)bi * a.nb + (q - jj)] = s;
        else a.rvg[(size_t)bi * a.nrhs + (q - 2 * jj)] = s;
      }
    }
  }
}

// ---- kernel C -----------------------------------------------------------------
__global__ __launch_bounds__(64) void k_tri_u(TriArgs a, int j, int jj) {
  __shared__ zc st1[64], st2[64], svg[16];
  const int lane = threadIdx.x;
  const int N = a.N;
  const int b0 = (j + 1) / TB;
  const int blk = b0 + blockIdx.x;
  const int r = blk * TB + lane;
  // reduce t1, t2, vg over row blocks (fixed order)
  for (int q = 0; q < 2 * jj + a.nrhs; ++q) {
    zc s;
    if (q < jj) s = sum_partials(a.rt1 + q, b0, a.nblk, a.nb);
    else if (q < 2 * jj) s = sum_partials(a.rt2 + (q - jj), b0, a.nblk, a.nb);
    else s = sum_partials(a.rvg + (q - 2 * jj), b0, a.nblk, a.nrhs);
    if (lane == 0) {
      if (q < jj) st1[q] = s;
      else if (q < 2 * jj) st2[q - jj] = s;
      else svg[q - 2 * jj] = s;
    }
  }
  __syncthreads();
  zc uv = mkz(0, 0);
  if (r > j && r < N) {
    zc y = mkz(0, 0);
    const zc* yp = a.ypart + ((size_t)blk * a.nblk) * TB + lane;
    for (int sl = b0; sl < a.nblk; ++sl) y += yp[(size_t)sl * TB];
    for (int c = 0; c < jj; ++c) y = y - a.Vp[(size_t)c * N + r] * st1[c];
    for (int c = 0; c < jj; ++c) y = y - a.W[(size_t)c * N + r] * st2[c];
    a.u[r] = y;
    zc v = a.Vp[(size_t)jj * N + r];
    uv = conj(y) * v;
    zc tj = a.tau[j];
    for (int q = 0; q < a.nrhs; ++q) {
      zc gq = a.g[(size_t)q * N + r];
      gq = gq - (tj * conj(v)) * svg[q];
      a.g[(size_t)q * N + r] = gq;
    }
  }
  uv = wave_sum(uv);
  if (lane == 0) a.ruv[blk] = uv;
}

// ---- W completion at the panel end ------------------------------------------------
__global__ __launch_bounds__(64) void k_tri_wlast(TriArgs a, int j, int jj) {
  const int lane = threadIdx.x;
  const int b0 = (j + 1) / TB;
  const int r = (b0 + blockIdx.x) * TB + lane;
  zc uv = sum_partials(a.ruv, b0, a.nblk, 1);
  zc tp = a.tau[j];
  double t2 = norm2(tp);
  zc ap = mkz(-0.5 * t2 * uv.re, -0.5 * t2 * uv.im);
  if (r > j && r < a.N) a.W[(size_t)jj * a.N + r] = tp * a.u[r] + ap * a.Vp[(size_t)jj * a.N + r];
}

// ---- trailing update A22 -= V W^H + W V^H on lower 64x64 tiles (MFMA f64) ---------
__global__ __launch_bounds__(256) void k_tri_her2k(TriArgs a, int s0, int kp) {
  const int N = a.N;
  const int b0 = s0 / TB;
  int t = blockIdx.x;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  while (ti * (ti + 1) / 2 > t) --ti;
  int tj = t - ti * (ti + 1) / 2;
  const int bi = b0 + ti, bj = b0 + tj;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row0 = bi * TB + (wave >> 1) * 32;
  const int col0 = bj * TB + (wave & 1) * 32;
  if (bi == bj && (wave >> 1) < (wave & 1)) return;   // strictly upper quarter of a diagonal tile
  const int lr = lane & 15, lk = lane >> 4;
  d4 cr[2][2], ci[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      cr[x][y] = d4{0, 0, 0, 0};
      ci[x][y] = d4{0, 0, 0, 0};
    }
  const int K = 2 * kp;
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
        int r = row0 + x * 16 + lk + 4 * q;
        int c = col0 + y * 16 + lr;
        if (r < N && c < N && r >= s0 && c >= s0 && r >= c) {
          size_t off = (size_t)r + (size_t)a.ld * c;
          zc h = a.H[off];
          h.re -= cr[x][y][q];
          h.im -= ci[x][y][q];
          a.H[off] = h;
        }
      }
}

__global__ void k_tri_last(TriArgs a) {
  zc x = a.H[(size_t)(a.N - 1) + (size_t)a.ld * (a.N - 1)];
  a.d[a.N - 1] = x.re;
  a.e[a.N - 1] = 0.0;
}

// ---------------------------------------------------------------------------
size_t tridiag_work_bytes(int N, int nb, int nrhs) {
  size_t nblk = (N + TB - 1) / TB;
  size_t b = 0;
  b += sizeof(zc) * (size_t)N * nb * 2;         // W, Vp
  b += sizeof(zc) * (size_t)N * 2;              // u, tau
  b += sizeof(double) * (size_t)N * 2;          // d, e
  b += sizeof(double) * (nblk + 1);             // rnorm (+1: keeps the complex partials aligned)
  b += sizeof(zc) * nblk * (1 + 2 * nb + nrhs); // ruv, rt1, rt2, rvg
  b += sizeof(zc) * nblk * nblk * TB;           // ypart
  return b + 1024;
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
  tw.red = (double*)take(sizeof(double) * (nblk + 1) + sizeof(zc) * nblk * (1 + 2 * nb + nrhs));
  tw.ypart = (zc*)take(sizeof(zc) * nblk * nblk * TB);
  return WTB_OK;
}

int run_tridiag(hipStream_t s, TridiagWork& tw, zc* d_H, int ldh, zc* d_g) {
  const int N = tw.N, nb = tw.nb;
  if (nb > 32 || tw.nrhs > 16) return WTB_ERR_LIMIT;
  TriArgs a;
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
  a.rnorm = tw.red;
  zc* zr = reinterpret_cast<zc*>(tw.red + ((a.nblk + 1) & ~1));   // 16-byte aligned
  a.ruv = zr;
  a.rt1 = zr + a.nblk;
  a.rt2 = a.rt1 + (size_t)a.nblk * nb;
  a.rvg = a.rt2 + (size_t)a.nblk * nb;
  a.ypart = tw.ypart;
  a.g = d_g;
  a.nrhs = tw.nrhs;
  if (N == 1) {
    hipLaunchKernelGGL(k_tri_last, dim3(1), dim3(1), 0, s, a);
    WTB_LAUNCH_CHECK();
    return WTB_OK;
  }
  for (int p = 0; p < N - 1; p += nb) {
    const int kp = (nb < N - 1 - p) ? nb : (N - 1 - p);
    int prev_b0 = 0;
    for (int jj = 0; jj < kp; ++jj) {
      const int j = p + jj;
      const int nA = a.nblk - j / TB;
      hipLaunchKernelGGL(k_tri_colupd, dim3(nA), dim3(64), 0, s, a, j, jj, prev_b0);
      const int nt = a.nblk - (j + 1) / TB;
      hipLaunchKernelGGL(k_tri_hemv, dim3(nt * (nt + 1) / 2), dim3(256), 0, s, a, j, jj);
      hipLaunchKernelGGL(k_tri_u, dim3(nt), dim3(64), 0, s, a, j, jj);
      prev_b0 = (j + 1) / TB;
    }
    const int jl = p + kp - 1;
    const int ntl = a.nblk - (jl + 1) / TB;
    hipLaunchKernelGGL(k_tri_wlast, dim3(ntl), dim3(64), 0, s, a, jl, kp - 1);
    const int s0 = p + kp;
    if (s0 < N) {
      const int nt2 = a.nblk - s0 / TB;
      hipLaunchKernelGGL(k_tri_her2k, dim3(nt2 * (nt2 + 1) / 2), dim3(256), 0, s, a, s0, kp);
    }
  }
  hipLaunchKernelGGL(k_tri_last, dim3(1), dim3(1), 0, s, a);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
