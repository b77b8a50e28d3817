// Stage (d2): eigenvalues of the real symmetric tridiagonal (d, e) by
// divide and conquer, carrying Z^T x for a few tracked vectors instead of the
// eigenvector matrix Z.  The algorithm is the one of LAPACK DSTEDC/DLAED0-4
// (rank-one tearing, deflation of small z components and of close poles by
// Givens rotations, secular equation, Gu-Eisenstat recomputation of z):
//
//   T = diag(T1, T2) + rho u u^T,  Ti = Zi Li Zi^T,  z = [last row Z1; +-first row Z2]
//   D + rho z z^T = U L U^T  =>  Z = diag(Z1, Z2) U
//
// Each node keeps, for its index range, its eigenvalues (ascending), the first
// and last rows of its Z (needed by its parent's z) and Z^T x for the tracked
// vectors x.  The parent applies U^T to [first1; 0], [0; last2], [x1; x2],
// which costs O(n^2) per merge, so the whole solve is O(N^2).
//
// With vectors requested, each node also keeps its block of the eigenvector
// matrix Z (N x N, column-major, the node's diagonal block): the parent gathers
// its children's columns in sorted order, applies the deflation rotations row by
// row, and forms diag(Z1, Z2) U for the non-deflated columns with an FP64 MFMA
// GEMM, writing every column straight to its sorted position.
//
// Leaves are single elements.  All nodes of one tree level are processed by the
// same kernel launches (grid.y = node).  The secular equation is solved by
// bisection over the bit patterns of the root's offset from the nearer pole,
// which reaches the last unit of that offset in at most 64 steps.
#include "wtb_kernels.h"
#include <algorithm>
#include <vector>

namespace wtb {

struct DCNode {
  int lo, mid, hi;
  int pad;
};

struct DCArgs {
  int N, nt;            // size, tracked vectors (<= 8)
  const DCNode* nodes;  // this level
  int nnodes;
  const double* e;      // off-diagonal
  double* lam;          // [N]
  double* FR;           // [N]
  double* LR;           // [N]
  double* T;            // [nt][N]
  // work
  double* sD;           // [N] sorted poles
  double* sz;           // [N]
  double* sx;           // [(2+nt)][N]: xF, xL, T...
  double* cD;           // [N] compacted: dlamda (first K), deflated after
  double* cw;           // [N]
  double* cx;           // [(2+nt)][N]
  int* slot;            // [N]
  int* org;             // [N]
  double* tau;          // [N]
  double* what;         // [N]
  double* nl;           // [N] new eigenvalues (unsorted)
  double* nx;           // [(2+nt)][N]
  int* K;               // [nnodes]
  double* rho;          // [nnodes]
  // eigenvector mode (Z != nullptr)
  double* Z;            // N x N: eigenvectors of the nodes (diagonal blocks)
  double* Z2;           // N x N: gathered, rotated columns
  double* U;            // N x N: secular eigenvector blocks
  int* perm;            // [N] sorted position -> source column
  int* cmap;            // [N] slot -> sorted position (node-relative)
  int* rank;            // [N] position -> sorted destination (node-relative)
  int* rotp;            // [N] deflation rotations: positions, cosine, sine
  int* rotn;
  double* rotc;
  double* rots;
  int* nrot;            // [nnodes]
};

__device__ __constant__ double kEps = 1.1102230246251565e-16;   // DLAMCH('E')

// d_i -= |e_{i-1}| + |e_i| (every coupling is a split point when leaves are size 1);
// leaves: lambda = d_i, first = last row = 1, T = x
__global__ void k_dc_leaves(int N, const double* __restrict__ d, const double* __restrict__ e,
                            double* lam, double* FR, double* LR) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double v = d[i];
  if (i > 0) v -= fabs(e[i - 1]);
  if (i < N - 1) v -= fabs(e[i]);
  lam[i] = v;
  FR[i] = 1.0;
  LR[i] = 1.0;
}

// M1: merge the children's sorted eigenvalues, build z and the tracked sources
__global__ void k_dc_gather(DCArgs a) {
  const DCNode nd = a.nodes[blockIdx.y];
  const int n = nd.hi - nd.lo, n1 = nd.mid - nd.lo;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = a.e[nd.mid - 1];
  const double sgn = beta < 0 ? -1.0 : 1.0;
  const double isq2 = 0.70710678118654752440;
  int src, rank;
  double val;
  if (i < n1) {
    src = nd.lo + i;
    val = a.lam[src];
    int l = nd.mid, h = nd.hi;       // count child-2 values < val
    while (l < h) {
      int m = (l + h) >> 1;
      if (a.lam[m] < val) l = m + 1; else h = m;
    }
    rank = i + (l - nd.mid);
  } else {
    src = nd.mid + (i - n1);
    val = a.lam[src];
    int l = nd.lo, h = nd.mid;       // count child-1 values <= val
    while (l < h) {
      int m = (l + h) >> 1;
      if (a.lam[m] <= val) l = m + 1; else h = m;
    }
    rank = (i - n1) + (l - nd.lo);
  }
  const int dst = nd.lo + rank;
  const int N = a.N;
  a.sD[dst] = val;
  a.sz[dst] = (i < n1 ? a.LR[src] : sgn * a.FR[src]) * isq2;
  a.sx[dst] = (i < n1) ? a.FR[src] : 0.0;              // first row of [Z1 0; 0 Z2]
  a.sx[(size_t)N + dst] = (i < n1) ? 0.0 : a.LR[src];  // last row
  for (int q = 0; q < a.nt; ++q) a.sx[(size_t)(2 + q) * N + dst] = a.T[(size_t)q * N + src];
  if (a.Z) a.perm[dst] = src;
}

// M2: deflation scan (DLAED2), one workgroup per node.  The poles and weights
// are staged in LDS (when they fit) for the sequential scan of thread 0, which
// decides each sorted position's slot (non-deflated first), applies the Givens
// rotations of close-pole deflation to the tracked sources in place, and writes
// the poles and weights of the secular equation.
__global__ __launch_bounds__(256) void k_dc_deflate(DCArgs a, int use_lds) {
  extern __shared__ double lds_dz[];
  __shared__ double red[2][256];
  const int nodeid = blockIdx.x;
  const DCNode nd = a.nodes[nodeid];
  const int lo = nd.lo, n = nd.hi - nd.lo;
  const int N = a.N, nx = 2 + a.nt;
  const double rho = fabs(2.0 * a.e[nd.mid - 1]);
  double zm = 0.0, dm = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    double zi = a.sz[lo + i], di = a.sD[lo + i];
    zm = fmax(zm, fabs(zi));
    dm = fmax(dm, fabs(di));
    if (use_lds) {
      lds_dz[i] = di;
      lds_dz[n + i] = zi;
    }
  }
  red[0][threadIdx.x] = zm;
  red[1][threadIdx.x] = dm;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + st]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + st]);
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double* D = use_lds ? lds_dz : a.sD + lo;
  double* Z = use_lds ? lds_dz + n : a.sz + lo;
  const double zmax = red[0][0], dmax = red[1][0];
  const double tol = 8.0 * kEps * fmax(dmax, zmax);
  a.rho[nodeid] = rho;
  int K = 0, nd_ = 0, nr = 0;
  auto deflate = [&](int i) {
    int s = n - 1 - nd_;
    ++nd_;
    a.slot[lo + i] = s;
    a.cD[lo + s] = D[i];
  };
  auto keep = [&](int i) {
    int s = K;
    ++K;
    a.slot[lo + i] = s;
    a.cD[lo + s] = D[i];
    a.cw[lo + s] = Z[i];
  };
  if (rho * zmax <= tol) {
    for (int i = 0; i < n; ++i) deflate(i);
    a.K[nodeid] = 0;
    if (a.Z) a.nrot[nodeid] = 0;
    return;
  }
  int j = 0, pj = -1;
  for (; j < n; ++j) {
    if (rho * fabs(Z[j]) <= tol) deflate(j);
    else {
      pj = j;
      ++j;
      break;
    }
  }
  if (pj >= 0) {
    for (; j < n; ++j) {
      const int nj = j;
      const double znj = Z[nj];
      if (rho * fabs(znj) <= tol) {
        deflate(nj);
        continue;
      }
      double s = Z[pj], c = znj;
      const double tau = hypot(c, s);
      const double t = D[nj] - D[pj];
      c = c / tau;
      s = -s / tau;
      if (fabs(t * c * s) <= tol) {
        Z[nj] = tau;
        Z[pj] = 0.0;
        for (int q = 0; q < nx; ++q) {
          double* x = a.sx + (size_t)q * N + lo;
          double xp = x[pj], xn = x[nj];
          x[pj] = c * xp + s * xn;
          x[nj] = c * xn - s * xp;
        }
        if (a.Z) {
          a.rotp[lo + nr] = pj;
          a.rotn[lo + nr] = nj;
          a.rotc[lo + nr] = c;
          a.rots[lo + nr] = s;
          ++nr;
        }
        const double dp = D[pj], dn = D[nj];
        const double t2 = dp * c * c + dn * s * s;
        D[nj] = dp * s * s + dn * c * c;
        D[pj] = t2;
        deflate(pj);
        pj = nj;
      } else {
        keep(pj);
        pj = nj;
      }
    }
    keep(pj);
  }
  a.K[nodeid] = K;
  if (a.Z) a.nrot[nodeid] = nr;
}

// M2b: move the tracked sources to their slots
__global__ void k_dc_compact(DCArgs a) {
  const DCNode nd = a.nodes[blockIdx.y];
  const int n = nd.hi - nd.lo;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = a.slot[nd.lo + i];
  const int N = a.N;
  for (int q = 0; q < 2 + a.nt; ++q) a.cx[(size_t)q * N + nd.lo + s] = a.sx[(size_t)q * N + nd.lo + i];
  if (a.Z) a.cmap[nd.lo + s] = i;
}

// secular function f(d_o + tau) = 1 + rho sum_i w_i^2 / ((d_i - d_o) - tau), summed by a
// 16-lane group
__device__ inline double secular16(const double* __restrict__ dl, const double* __restrict__ w, int K,
                                   int o, double tau, double rho, int sub) {
  const double dor = dl[o];
  double s = 0.0;
  for (int i = sub; i < K; i += 16) {
    double wi = w[i];
    s += wi * wi / ((dl[i] - dor) - tau);
  }
  for (int off = 8; off > 0; off >>= 1) s += __shfl_xor(s, off, 16);
  return 1.0 + rho * s;
}

// M3: roots of the secular equation, 16 lanes per root
__global__ __launch_bounds__(256) void k_dc_secular(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int K = a.K[node];
  const int jr = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int sub = threadIdx.x & 15;
  if (blockIdx.x * (blockDim.x >> 4) >= K) return;   // whole block idle (uniform)
  const bool active = jr < K;
  const int j = active ? jr : K - 1;
  const double* dl = a.cD + nd.lo;
  const double* w = a.cw + nd.lo;
  const double rho = a.rho[node];
  int o;
  bool neg;
  double hi;
  if (j < K - 1) {
    double gap = dl[j + 1] - dl[j];
    double mid = 0.5 * gap;
    double fm = secular16(dl, w, K, j, mid, rho, sub);
    if (fm >= 0.0) {
      o = j;
      neg = false;
      hi = mid;
    } else {
      o = j + 1;
      neg = true;
      hi = gap - mid;
    }
  } else {
    double ws = 0.0;
    for (int i = sub; i < K; i += 16) ws += w[i] * w[i];
    for (int off = 8; off > 0; off >>= 1) ws += __shfl_xor(ws, off, 16);
    o = K - 1;
    neg = false;
    hi = rho * ws;
  }
  // bisection on the magnitude m of tau (tau = +-m) over the bit patterns of m
  unsigned long long lob = 0ull, hib = __double_as_longlong(hi);
  for (int it = 0; it < 64; ++it) {
    if (hib - lob <= 1ull) break;
    unsigned long long mb = lob + ((hib - lob) >> 1);
    double m = __longlong_as_double(mb);
    double f = secular16(dl, w, K, o, neg ? -m : m, rho, sub);
    // f increases with tau: for tau = +m the root lies above m iff f < 0;
    // for tau = -m it lies at larger m iff f > 0
    bool up = neg ? (f > 0.0) : (f < 0.0);
    if (up) lob = mb; else hib = mb;
  }
  double m = __longlong_as_double(hib);
  if (active && sub == 0) {
    a.org[nd.lo + j] = o;
    a.tau[nd.lo + j] = neg ? -m : m;
  }
}

// d_i - lambda_j with lambda_j = d_{o_j} + tau_j
__device__ inline double dmin_lam(const double* dl, const int* org, const double* tau, int i, int j) {
  return (dl[i] - dl[org[j]]) - tau[j];
}

// M4: Gu-Eisenstat weights (DLAED3): w_i = sign(z_i) sqrt(-prod_j (d_i-l_j) / prod_{j!=i} (d_i-d_j))
__global__ __launch_bounds__(256) void k_dc_lowner(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int K = a.K[node];
  if (blockIdx.x * (blockDim.x >> 4) >= K) return;
  const int ir = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int sub = threadIdx.x & 15;
  const bool active = ir < K;
  const int i = active ? ir : K - 1;
  const double* dl = a.cD + nd.lo;
  const int* org = a.org + nd.lo;
  const double* tau = a.tau + nd.lo;
  double p = 1.0;
  for (int j = sub; j < K; j += 16) {
    double dlj = dmin_lam(dl, org, tau, i, j);
    p *= (j == i) ? dlj : dlj / (dl[i] - dl[j]);
  }
  for (int off = 8; off > 0; off >>= 1) p *= __shfl_xor(p, off, 16);
  if (active && sub == 0) a.what[nd.lo + i] = copysign(sqrt(fmax(-p, 0.0)), a.cw[nd.lo + i]);
}

// M5: new eigenvalues and U^T x for the tracked sources; deflated slots copy through.
// NXM >= 2 + a.nt: the accumulators (10 for up to 8 tracked sources, 14 for 12)
template <int NXM>
__global__ __launch_bounds__(256) void k_dc_apply(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo;
  const int K = a.K[node];
  const int jr = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
  const int sub = threadIdx.x & 15;
  const int N = a.N, nx = 2 + a.nt;
  if (blockIdx.x * (blockDim.x >> 4) >= n) return;
  if (jr >= n) return;                       // lanes of a 16-group share jr
  const int lo = nd.lo;
  if (jr >= K) {
    if (sub == 0) {
      a.nl[lo + jr] = a.cD[lo + jr];
      for (int q = 0; q < nx; ++q) a.nx[(size_t)q * N + lo + jr] = a.cx[(size_t)q * N + lo + jr];
    }
    return;
  }
  const int j = jr;
  const double* dl = a.cD + lo;
  const int* org = a.org + lo;
  const double* tau = a.tau + lo;
  const double* wh = a.what + lo;
  double acc[NXM];
  for (int q = 0; q < nx; ++q) acc[q] = 0.0;
  double nrm = 0.0;
  for (int i = sub; i < K; i += 16) {
    double s = wh[i] / dmin_lam(dl, org, tau, i, j);
    nrm += s * s;
    for (int q = 0; q < nx; ++q) acc[q] += s * a.cx[(size_t)q * N + lo + i];
  }
  for (int off = 8; off > 0; off >>= 1) {
    nrm += __shfl_xor(nrm, off, 16);
    for (int q = 0; q < nx; ++q) acc[q] += __shfl_xor(acc[q], off, 16);
  }
  const double inv = 1.0 / sqrt(nrm);
  if (sub == 0) {
    a.nl[lo + j] = dl[org[j]] + tau[j];
    for (int q = 0; q < nx; ++q) a.nx[(size_t)q * N + lo + j] = acc[q] * inv;
  }
  if (a.Z) {
    double* Uj = a.U + (size_t)(lo + j) * N + lo;
    for (int i = sub; i < K; i += 16) Uj[i] = wh[i] / dmin_lam(dl, org, tau, i, j) * inv;
  }
}

// M6: sort the node's eigenvalues ascending and write lam/FR/LR/T
__global__ void k_dc_sort(DCArgs a) {
  const DCNode nd = a.nodes[blockIdx.y];
  const int n = nd.hi - nd.lo, lo = nd.lo;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = a.nl[lo + i];
  int rank = 0;
  for (int k = 0; k < n; ++k) {
    double u = a.nl[lo + k];
    rank += (u < v) || (u == v && k < i);
  }
  const int N = a.N, dst = lo + rank;
  if (a.Z) a.rank[lo + i] = rank;
  a.lam[dst] = v;
  a.FR[dst] = a.nx[lo + i];
  a.LR[dst] = a.nx[(size_t)N + lo + i];
  for (int q = 0; q < a.nt; ++q) a.T[(size_t)q * N + dst] = a.nx[(size_t)(2 + q) * N + lo + i];
}


// ---- eigenvector mode ------------------------------------------------------------------
__global__ void k_dcv_init(int N, double* Z) {
  const size_t n2 = (size_t)N * N;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n2; t += (size_t)gridDim.x * blockDim.x)
    Z[t] = (t % (N + 1) == 0) ? 1.0 : 0.0;
}

// Z2(:, lo+i) = Z(:, perm(lo+i)) over the node's rows: its children's columns in sorted order
__global__ __launch_bounds__(256) void k_dcv_gather(DCArgs a) {
  const DCNode nd = a.nodes[blockIdx.y];
  const int n = nd.hi - nd.lo, i = blockIdx.x;
  if (i >= n) return;
  const size_t N = a.N;
  const double* src = a.Z + (size_t)a.perm[nd.lo + i] * N;
  double* dst = a.Z2 + (size_t)(nd.lo + i) * N;
  for (int r = nd.lo + threadIdx.x; r < nd.hi; r += 256) dst[r] = src[r];
}

// the close-pole deflation rotations, in scan order, one thread per row
__global__ __launch_bounds__(256) void k_dcv_rotate(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int r = nd.lo + blockIdx.x * 256 + threadIdx.x;
  if (r >= nd.hi) return;
  const int nr = a.nrot[node];
  const size_t N = a.N;
  for (int k = 0; k < nr; ++k) {
    const int q = nd.lo + k;
    double* xp = a.Z2 + (size_t)(nd.lo + a.rotp[q]) * N + r;
    double* xn = a.Z2 + (size_t)(nd.lo + a.rotn[q]) * N + r;
    const double c = a.rotc[q], sn = a.rots[q];
    const double p = *xp, v = *xn;
    *xp = c * p + sn * v;
    *xn = c * v - sn * p;
  }
}

// Z(:, lo + rank(j)) = Z2(:, lo + cmap(0..K-1)) U(0..K-1, j) for the non-deflated
// slots j < K; 64 x 64 output tiles, four waves of 32 x 32 on v_mfma_f64_16x16x4
__global__ __launch_bounds__(256) void k_dcv_gemm(DCArgs a) {
  const int node = blockIdx.z;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo, K = a.K[node], lo = nd.lo;
  const int r0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  if (r0 >= n || j0 >= K) return;
  const size_t N = a.N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int rb = r0 + (wave >> 1) * 32, cb = j0 + (wave & 1) * 32;
  d4 acc[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0, 0, 0, 0};
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + lk;
    double av[2] = {0.0, 0.0}, bv[2] = {0.0, 0.0};
    if (k < K) {
      const double* zc_ = a.Z2 + (size_t)(lo + a.cmap[lo + k]) * N + lo;
      for (int x = 0; x < 2; ++x) {
        const int r = rb + x * 16 + lr;
        if (r < n) av[x] = zc_[r];
      }
      for (int y = 0; y < 2; ++y) {
        const int c = cb + y * 16 + lr;
        if (c < K) bv[y] = a.U[(size_t)(lo + c) * N + lo + k];
      }
    }
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y) acc[x][y] = mfma16x16x4(av[x], bv[y], acc[x][y]);
  }
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) {
      const int c = cb + y * 16 + lr;
      if (c >= K) continue;
      double* dst = a.Z + (size_t)(lo + a.rank[lo + c]) * N + lo;
      for (int q = 0; q < 4; ++q) {
        const int r = rb + x * 16 + lk + 4 * q;     // f64 MFMA: register q holds row lk + 4q
        if (r < n) dst[r] = acc[x][y][q];
      }
    }
}

// deflated slots j >= K: the gathered, rotated column moves to its sorted position
__global__ __launch_bounds__(256) void k_dcv_copy(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo, j = blockIdx.x;
  if (j >= n || j < a.K[node]) return;
  const size_t N = a.N;
  const double* src = a.Z2 + (size_t)(nd.lo + a.cmap[nd.lo + j]) * N;
  double* dst = a.Z + (size_t)(nd.lo + a.rank[nd.lo + j]) * N;
  for (int r = nd.lo + threadIdx.x; r < nd.hi; r += 256) dst[r] = src[r];
}

// ---------------------------------------------------------------------------
// host side: tree and launches
// ---------------------------------------------------------------------------
static void build_levels(int N, std::vector<std::vector<DCNode>>& levels) {
  struct Item { int lo, hi, depth; };
  std::vector<Item> st;
  st.push_back({0, N, 0});
  std::vector<std::pair<int, DCNode>> merges;
  int maxd = 0;
  while (!st.empty()) {
    Item it = st.back();
    st.pop_back();
    if (it.hi - it.lo <= 1) continue;
    int mid = it.lo + (it.hi - it.lo) / 2;
    merges.push_back({it.depth, DCNode{it.lo, mid, it.hi, 0}});
    maxd = it.depth > maxd ? it.depth : maxd;
    st.push_back({it.lo, mid, it.depth + 1});
    st.push_back({mid, it.hi, it.depth + 1});
  }
  levels.assign(maxd + 1, {});
  for (auto& m : merges) levels[maxd - m.first].push_back(m.second);   // deepest first
}

size_t dc_work_bytes(int N, int ntrack, int vectors) {
  size_t nx = 2 + ntrack;
  size_t b = 0;
  if (vectors) {
    b += sizeof(double) * (size_t)N * N * 2;   // Z2 U
    b += sizeof(int) * (size_t)N * 5;          // perm cmap rank rotp rotn
    b += sizeof(double) * (size_t)N * 2;       // rotc rots
    b += sizeof(int) * (N / 2 + 2) + 8 * 256;  // nrot
  }
  b += sizeof(double) * N * 3;             // lam FR LR
  b += sizeof(double) * N * 2;             // sD sz
  b += sizeof(double) * N * nx * 3;        // sx cx nx
  b += sizeof(double) * N * 5;             // cD cw tau what nl
  b += sizeof(int) * N * 2;                // slot org
  b += (sizeof(int) + sizeof(double)) * (N / 2 + 2);   // K rho
  b += sizeof(DCNode) * (N + 2);           // nodes
  return b + 64 * 256;
}

struct DCLayout {
  DCArgs a;
  DCNode* dnodes;
};

static DCLayout dc_layout(void* work, int N, int ntrack, int vectors) {
  char* p = static_cast<char*>(work);
  auto take = [&](size_t bytes) {
    char* q = p;
    p += (bytes + 255) & ~size_t(255);
    return q;
  };
  const size_t nx = 2 + ntrack;
  DCLayout L;
  DCArgs& a = L.a;
  a.N = N;
  a.nt = ntrack;
  a.FR = (double*)take(sizeof(double) * N);
  a.LR = (double*)take(sizeof(double) * N);
  a.sD = (double*)take(sizeof(double) * N);
  a.sz = (double*)take(sizeof(double) * N);
  a.sx = (double*)take(sizeof(double) * N * nx);
  a.cD = (double*)take(sizeof(double) * N);
  a.cw = (double*)take(sizeof(double) * N);
  a.cx = (double*)take(sizeof(double) * N * nx);
  a.slot = (int*)take(sizeof(int) * N);
  a.org = (int*)take(sizeof(int) * N);
  a.tau = (double*)take(sizeof(double) * N);
  a.what = (double*)take(sizeof(double) * N);
  a.nl = (double*)take(sizeof(double) * N);
  a.nx = (double*)take(sizeof(double) * N * nx);
  a.K = (int*)take(sizeof(int) * (N / 2 + 2));
  a.rho = (double*)take(sizeof(double) * (N / 2 + 2));
  L.dnodes = (DCNode*)take(sizeof(DCNode) * (N + 2));
  a.Z = nullptr;
  a.Z2 = a.U = nullptr;
  a.perm = a.cmap = a.rank = a.rotp = a.rotn = a.nrot = nullptr;
  a.rotc = a.rots = nullptr;
  if (vectors) {
    a.Z2 = (double*)take(sizeof(double) * (size_t)N * N);
    a.U = (double*)take(sizeof(double) * (size_t)N * N);
    a.perm = (int*)take(sizeof(int) * N);
    a.cmap = (int*)take(sizeof(int) * N);
    a.rank = (int*)take(sizeof(int) * N);
    a.rotp = (int*)take(sizeof(int) * N);
    a.rotn = (int*)take(sizeof(int) * N);
    a.rotc = (double*)take(sizeof(double) * N);
    a.rots = (double*)take(sizeof(double) * N);
    a.nrot = (int*)take(sizeof(int) * (N / 2 + 2));
  }
  return L;
}

int dc_prepare(void* work, int N, int ntrack, DCTree& tree, int vectors) {
  std::vector<std::vector<DCNode>> levels;
  build_levels(N, levels);
  std::vector<DCNode> flat;
  tree.level_off.clear();
  tree.level_cnt.clear();
  tree.level_maxn.clear();
  for (auto& L : levels) {
    if (L.empty()) continue;
    int maxn = 0;
    for (auto& nd : L) maxn = std::max(maxn, nd.hi - nd.lo);
    tree.level_off.push_back((int)flat.size());
    tree.level_cnt.push_back((int)L.size());
    tree.level_maxn.push_back(maxn);
    flat.insert(flat.end(), L.begin(), L.end());
  }
  tree.N = N;
  tree.ntrack = ntrack;
  tree.vectors = vectors;
  DCLayout lay = dc_layout(work, N, ntrack, vectors);
  // the deflation scan stages up to 2 x 8192 doubles in LDS
  WTB_HIP(hipFuncSetAttribute((const void*)k_dc_deflate, hipFuncAttributeMaxDynamicSharedMemorySize,
                              2 * 8192 * (int)sizeof(double)));
  if (!flat.empty()) {
    WTB_HIP(hipMemcpy(lay.dnodes, flat.data(), sizeof(DCNode) * flat.size(), hipMemcpyHostToDevice));
  }
  return WTB_OK;
}

int run_dc(hipStream_t s, void* work, const DCTree& tree, const double* d_d, const double* d_e,
           double* d_track, double* d_lambda, double* d_Z) {
  const int N = tree.N;
  if (tree.ntrack > 12) {
    set_error("divide and conquer: at most 12 tracked vectors");
    return WTB_ERR_LIMIT;
  }
  if (d_Z && !tree.vectors) {
    set_error("divide and conquer: eigenvectors requested from a plan prepared without them");
    return WTB_ERR_ARG;
  }
  DCLayout lay = dc_layout(work, N, tree.ntrack, tree.vectors);
  DCArgs a = lay.a;
  a.e = d_e;
  a.lam = d_lambda;
  a.T = d_track;
  a.Z = d_Z;
  if (d_Z) hipLaunchKernelGGL(k_dcv_init, dim3(1024), dim3(256), 0, s, N, d_Z);
  hipLaunchKernelGGL(k_dc_leaves, dim3((N + 255) / 256), dim3(256), 0, s, N, d_d, d_e, d_lambda, a.FR,
                     a.LR);
  WTB_LAUNCH_CHECK();
  for (size_t l = 0; l < tree.level_off.size(); ++l) {
    const int cnt = tree.level_cnt[l], maxn = tree.level_maxn[l];
    a.nodes = lay.dnodes + tree.level_off[l];
    a.nnodes = cnt;
    dim3 g1((maxn + 255) / 256, cnt);
    dim3 g16((maxn * 16 + 255) / 256, cnt);
    hipLaunchKernelGGL(k_dc_gather, g1, dim3(256), 0, s, a);
    const int use_lds = maxn <= 8192 ? 1 : 0;
    hipLaunchKernelGGL(k_dc_deflate, dim3(cnt), dim3(256), use_lds ? 2 * sizeof(double) * maxn : 0, s,
                       a, use_lds);
    hipLaunchKernelGGL(k_dc_compact, g1, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_dc_secular, g16, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_dc_lowner, g16, dim3(256), 0, s, a);
    if (tree.ntrack <= 8) hipLaunchKernelGGL(k_dc_apply<10>, g16, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_dc_apply<14>, g16, dim3(256), 0, s, a);
    hipLaunchKernelGGL(k_dc_sort, g1, dim3(256), 0, s, a);
    if (d_Z) {
      hipLaunchKernelGGL(k_dcv_gather, dim3(maxn, cnt), dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_dcv_rotate, dim3((maxn + 255) / 256, cnt), dim3(256), 0, s, a);
      const int nt64 = (maxn + 63) / 64;
      hipLaunchKernelGGL(k_dcv_gemm, dim3(nt64, nt64, cnt), dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_dcv_copy, dim3(maxn, cnt), dim3(256), 0, s, a);
    }
    WTB_LAUNCH_CHECK();
  }
  return WTB_OK;
}

}  // namespace wtb
