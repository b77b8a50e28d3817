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
// same kernel launches (grid.y = node).  The secular equation is solved by the
// middle-way iteration (DLAED4's rational interpolation of the two nearest poles)
// inside a bracket over the bit patterns of the root's offset from the nearer pole,
// which ends, like the plain bisection over those patterns, at the last unit of the
// offset.
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
  int zcols;            // eigenvector columns formed at this level: sorted positions < zcols
};

__device__ __constant__ double kEps = 1.1102230246251565e-16;   // DLAMCH('E')
// merges above this size use a whole wave per root in the secular, weight and apply
// kernels (16 lanes per root below it, where a wave would be mostly idle)
static const int kDCWideMin = 256;

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

// exclusive prefix count of f over the workgroup's 256 threads (every thread calls it);
// *total receives the count over all of them
__device__ inline int block_prefix256(bool f, int* wsum, int* total) {
  const unsigned long long b = __ballot(f);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int inw = __popcll(b & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wv] = __popcll(b);
  __syncthreads();
  int off = 0, tot = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wv) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return off + inw;
}

// M2: deflation (DLAED2), one workgroup of 256 per node.  The decisions and values are
// those of the sequential scan over the sorted positions (small weights deflate; a
// close pair (pj, nj) of kept candidates is rotated so that pj deflates, and the rotated
// nj becomes the next pair's pj), found in three parts:
//   - the small-weight test of every position and the close-pole test of every pair of
//     consecutive candidates, on the original values, in parallel;
//   - one wave walks the pairs in order, 64 at a time by ballot while no rotation is due,
//     and applies the rotations one by one, testing the pair after a rotation with the
//     rotated values (the only values a test can see changed);
//   - the slots from prefix counts: kept positions first, in order; deflated ones from the
//     end, in the order of deflation (a small weight at its own position, a rotated pj at
//     its nj's).
// The Givens rotations then act on the tracked sources in their order.
__global__ __launch_bounds__(256) void k_dc_deflate(DCArgs a, int use_lds) {
  extern __shared__ double lds_dz[];
  __shared__ double red[2][256];
  __shared__ int wsum[4];
  __shared__ int s_nr;
  const int nodeid = blockIdx.x;
  const DCNode nd = a.nodes[nodeid];
  const int lo = nd.lo, n = nd.hi - nd.lo;
  const int N = a.N, nx = 2 + a.nt;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double rho = fabs(2.0 * a.e[nd.mid - 1]);
  double zm = 0.0, dm = 0.0;
  for (int i = tid; i < n; i += 256) {
    double zi = a.sz[lo + i], di = a.sD[lo + i];
    zm = fmax(zm, fabs(zi));
    dm = fmax(dm, fabs(di));
    if (use_lds) {
      lds_dz[i] = di;
      lds_dz[n + i] = zi;
    }
  }
  red[0][tid] = zm;
  red[1][tid] = dm;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[0][tid] = fmax(red[0][tid], red[0][tid + st]);
      red[1][tid] = fmax(red[1][tid], red[1][tid + st]);
    }
    __syncthreads();
  }
  double* D = use_lds ? lds_dz : a.sD + lo;
  double* Z = use_lds ? lds_dz + n : a.sz + lo;
  const double zmax = red[0][0], dmax = red[1][0];
  const double tol = 8.0 * kEps * fmax(dmax, zmax);
  if (tid == 0) a.rho[nodeid] = rho;
  if (rho * zmax <= tol) {
    // every position deflates, in order
    for (int i = tid; i < n; i += 256) {
      const int sl = n - 1 - i;
      a.slot[lo + i] = sl;
      a.cD[lo + sl] = D[i];
    }
    if (tid == 0) {
      a.K[nodeid] = 0;
      a.nrot[nodeid] = 0;
    }
    return;
  }
  // flags per position: bit 0 small weight, bit 1 the pair ending here passes the
  // close-pole test on original values, bit 2 nj of a rotation, bit 3 pj of a rotation,
  // bits 4.. the pj of the rotation at this nj.  P: the candidates (not small), in order.
  int* flag = a.org + lo;
  int* P = reinterpret_cast<int*>(a.tau + lo);
  int m = 0;
  for (int c0 = 0; c0 < n; c0 += 256) {
    const int i = c0 + tid;
    const bool small = i < n && rho * fabs(Z[i]) <= tol;
    if (i < n) flag[i] = small ? 1 : 0;
    int tot;
    const int pos = block_prefix256(i < n && !small, wsum, &tot);
    if (i < n && !small) P[m + pos] = i;
    m += tot;
  }
  __syncthreads();
  for (int k = 1 + tid; k < m; k += 256) {
    const int pj = P[k - 1], nj = P[k];
    double s = Z[pj], c = Z[nj];
    const double tau = hypot(c, s);
    const double t = D[nj] - D[pj];
    c = c / tau;
    s = -s / tau;
    if (fabs(t * c * s) <= tol) flag[nj] |= 2;
  }
  __syncthreads();
  if (wv == 0) {
    int k = 1, nr = 0;
    bool kmod = false;          // P[k-1] carries a rotation's values
    double dmod = 0.0, zmod = 0.0;
    while (k < m) {
      if (!kmod) {
        const int kk = k + lane;
        const unsigned long long b = __ballot(kk < m && (flag[P[kk < m ? kk : 0]] & 2));
        if (b == 0ull) {
          k += 64;
          continue;
        }
        k += __ffsll((long long)b) - 1;
      }
      const int pj = P[k - 1], nj = P[k];
      const double dp = kmod ? dmod : D[pj], zp = kmod ? zmod : Z[pj];
      const double dn = D[nj];
      double s = zp, c = Z[nj];
      const double tau = hypot(c, s);
      const double t = dn - dp;
      c = c / tau;
      s = -s / tau;
      if (fabs(t * c * s) <= tol) {
        const double t2 = dp * c * c + dn * s * s;
        const double dnew = dp * s * s + dn * c * c;
        if (lane == 0) {
          Z[nj] = tau;
          Z[pj] = 0.0;
          D[nj] = dnew;
          D[pj] = t2;
          flag[nj] |= 4 | (pj << 4);
          flag[pj] |= 8;
          a.rotp[lo + nr] = pj;
          a.rotn[lo + nr] = nj;
          a.rotc[lo + nr] = c;
          a.rots[lo + nr] = s;
        }
        ++nr;
        kmod = true;
        dmod = dnew;
        zmod = tau;
      } else {
        kmod = false;
      }
      ++k;
    }
    if (lane == 0) s_nr = nr;
  }
  __syncthreads();
  const int nr = s_nr;
  int K = 0, E = 0;
  for (int c0 = 0; c0 < n; c0 += 256) {
    const int i = c0 + tid;
    const int fl = i < n ? flag[i] : 0;
    const bool kept = i < n && (fl & 9) == 0;
    const bool ev = i < n && (fl & 5) != 0;
    int tk, te;
    const int pk = block_prefix256(kept, wsum, &tk);
    const int pe = block_prefix256(ev, wsum, &te);
    if (kept) {
      const int sl = K + pk;
      a.slot[lo + i] = sl;
      a.cD[lo + sl] = D[i];
      a.cw[lo + sl] = Z[i];
    }
    if (ev) {
      const int e = (fl & 1) ? i : (fl >> 4);
      const int sl = n - 1 - (E + pe);
      a.slot[lo + e] = sl;
      a.cD[lo + sl] = D[e];
    }
    K += tk;
    E += te;
  }
  if (tid == 0) {
    a.K[nodeid] = K;
    a.nrot[nodeid] = nr;
  }
  // the rotations on the tracked sources, in order (a source per thread)
  if (nr > 0 && tid < nx) {
    double* x = a.sx + (size_t)tid * N + lo;
    for (int r = 0; r < nr; ++r) {
      const int pj = a.rotp[lo + r], nj = a.rotn[lo + r];
      const double c = a.rotc[lo + r], s = a.rots[lo + r];
      const double xp = x[pj], xn = x[nj];
      x[pj] = c * xp + s * xn;
      x[nj] = c * xn - s * xp;
    }
  }
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

// x of another lane of the same 16-lane DPP row, both halves of the double
template <int CTRL>
__device__ inline double dpp16(double x) {
  const long long v = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (long long)(unsigned)lo);
}
// sum over the 16 lanes of a DPP row (lanes 16r .. 16r + 15): pairs ^1, ^2 (quad
// permutations), i <-> 7 - i, i <-> 15 - i; every step adds two equal partial sums in
// either order, so every lane ends with the same bits
__device__ inline double row16_sum(double x) {
  x += dpp16<0xB1>(x);
  x += dpp16<0x4E>(x);
  x += dpp16<0x141>(x);
  x += dpp16<0x140>(x);
  return x;
}

// the sums of the secular function at d_o + tau, split at pole js: psi = sum_{i <= js}
// w_i^2 / (delta_i - tau), phi over i > js, and their derivatives, each summed by a 16-lane
// group (every lane gets the same values)
// sum over a G-lane group (G = 16: one DPP row; G = 64: the wave, the rows then joined
// pairwise by xor 16 and 32, which again leaves the same bits in every lane)
template <int G>
__device__ inline double group_sum(double x) {
  x = row16_sum(x);
  if (G == 64) {
    x += __shfl_xor(x, 16);
    x += __shfl_xor(x, 32);
  }
  return x;
}

struct SecSums {
  double psi, dpsi, phi, dphi;
};
template <int G>
__device__ inline SecSums secular_sums(const double* __restrict__ dl, const double* __restrict__ w, int K, int o,
                                       int js, double tau, int sub) {
  const double dor = dl[o];
  SecSums r{0.0, 0.0, 0.0, 0.0};
  // eight poles per step, all loads issued before the first use; the terms are added in
  // the order of i either way, each to psi (i <= js) or phi
  auto term = [&](double di, double wi, bool lo) {
    const double inv = 1.0 / ((di - dor) - tau);
    const double t = wi * wi * inv;
    const double t2 = t * inv;
    r.psi += lo ? t : 0.0;
    r.dpsi += lo ? t2 : 0.0;
    r.phi += lo ? 0.0 : t;
    r.dphi += lo ? 0.0 : t2;
  };
  int i = sub;
  for (; i + 7 * G < K; i += 8 * G) {
    double dv[8], wv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      dv[u] = dl[i + u * G];
      wv[u] = w[i + u * G];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) term(dv[u], wv[u], i + u * G <= js);
  }
  for (; i < K; i += G) term(dl[i], w[i], i <= js);
  r.psi = group_sum<G>(r.psi);
  r.dpsi = group_sum<G>(r.dpsi);
  r.phi = group_sum<G>(r.phi);
  r.dphi = group_sum<G>(r.dphi);
  return r;
}

// M3: roots of the secular equation, G lanes per root (16, or a wave for merges above
// 256).  Root j lies between poles j and j + 1 (above the last pole for j = K - 1); its
// offset tau from the nearer pole o is bracketed by the bit patterns of |tau|, [lob, hib],
// with f on the two sides of zero, and every evaluation narrows the bracket.  The first
// point is the midpoint of the bracket in value.  The next point is the middle-way step
// (rational interpolation of the two nearest poles' parts of f, with value and derivative
// at the last point) when it falls strictly inside the bracket, else the bracket's bit
// midpoint.  The iterate has converged when, after a first middle-way step, a step moves
// by at most 8 bit patterns or lands just past the end the last point set; after 12
// middle-way steps it is taken as converged too.  From there the other end is found by
// galloping (1, 2, 4, ... patterns) until the sign flips, and the patterns left are
// bisected.  The iteration ends where the plain bisection over the bit patterns ends, at
// adjacent patterns.  A CPU emulation on the merges of an N = 5184 BSE tridiagonal gave
// the bisection's bits on every root, in 8.4 evaluations on average at the root merge
// (at most 30) instead of 62.
template <int G>
__global__ __launch_bounds__(256) void k_dc_secular(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int K = a.K[node];
  const int jr = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int sub = threadIdx.x & (G - 1);
  if (blockIdx.x * (blockDim.x / G) >= K) return;    // whole block idle (uniform)
  const bool active = jr < K;
  const int j = active ? jr : K - 1;
  const double* dl = a.cD + nd.lo;
  const double* w = a.cw + nd.lo;
  const double rho = a.rho[node];
  const bool last = j == K - 1;
  const int js = last ? K - 2 : j;                  // psi: poles <= js, phi: above
  const bool mw = K >= 2;                            // a two-pole model exists
  int o;
  bool neg;
  double hi;
  SecSums sm{0.0, 0.0, 0.0, 0.0};
  double tc = 0.0;                                   // tau of the last evaluation (origin o)
  bool have = false;
  if (!last) {
    const double gap = dl[j + 1] - dl[j];
    const double mid = 0.5 * gap;
    sm = secular_sums<G>(dl, w, K, j, js, mid, sub);
    const double fm = 1.0 + rho * (sm.psi + sm.phi);
    if (fm >= 0.0) {
      o = j;
      neg = false;
      hi = mid;
      tc = mid;
    } else {
      o = j + 1;
      neg = true;
      hi = gap - mid;
      tc = -hi;                                      // the same point, from pole j + 1
    }
    have = true;                                     // the sums do not depend on the origin
  } else {
    double ws = 0.0;
    for (int i = sub; i < K; i += G) ws += w[i] * w[i];
    ws = group_sum<G>(ws);
    o = K - 1;
    neg = false;
    hi = rho * ws;
    // a first point at the bracket's midpoint in value (the bit midpoint of [0, hi] is
    // a tiny number, from which the middle-way model says nothing)
    tc = 0.5 * hi;
    sm = secular_sums<G>(dl, w, K, o, js, tc, sub);
    have = true;
  }
  // the poles bounding the two-pole model, relative to the origin
  const double dlo = mw ? dl[js] - dl[o] : 0.0, dhi = mw ? dl[js + 1] - dl[o] : 0.0;
  unsigned long long lob = 0ull, hib = __double_as_longlong(hi);
  int side = 0;                                      // last evaluation moved lob (+1) or hib (-1)
  if (last) {
    if (1.0 + rho * (sm.psi + sm.phi) < 0.0) {
      lob = __double_as_longlong(tc);
      side = 1;
    } else {
      hib = __double_as_longlong(tc);
      side = -1;
    }
  }
  int nmw = 0, gdir = 0;
  bool conv = false;
  unsigned long long step = 1ull;
  unsigned long long lastb = have ? __double_as_longlong(fabs(tc)) : 0ull;
  // at most 12 middle-way steps and 64 gallop steps come before the bisection's 64, so
  // the bracket always closes
  for (int it = 0; it < 144; ++it) {
    if (hib - lob <= 1ull) break;
    unsigned long long mb = lob + ((hib - lob) >> 1);
    if (conv) {
      // converged: gallop by 1, 2, 4, ... bit patterns towards the end not yet next to
      // the root until the sign flips, then bisect the few patterns left
      if (gdir == 0) gdir = side;
      if (side == gdir) {
        const unsigned long long pb = gdir > 0 ? lob + step : hib - step;
        step <<= 1;
        if (pb > lob && pb < hib) mb = pb;
      }
    } else if (mw && have && nmw >= 12) {
      // twelve middle-way steps without convergence (the steps creep): gallop from the
      // last point
      conv = true;
      gdir = side;
      const unsigned long long pb = gdir > 0 ? lob + 1ull : hib - 1ull;
      step = 2ull;
      if (pb > lob && pb < hib) mb = pb;
    } else if (mw && have) {
      // middle way: f / rho = 1 / rho + psi + phi; c eta^2 - a eta + b = 0 with the
      // two poles' distances from the last point
      const double W = 1.0 / rho + sm.psi + sm.phi;
      const double d1 = dlo - tc, d2 = dhi - tc;
      double c = W - d1 * sm.dpsi - d2 * sm.dphi;
      const double aa = (d1 + d2) * W - d1 * d2 * (sm.dpsi + sm.dphi);
      const double bb = d1 * d2 * W;
      const double disc = sqrt(fabs(aa * aa - 4.0 * bb * c));
      double eta;
      if (!last) {
        eta = (aa <= 0.0) ? (aa - disc) / (2.0 * c) : 2.0 * bb / (aa + disc);
      } else {
        c = fabs(c);
        eta = (aa >= 0.0) ? (aa + disc) / (2.0 * c) : 2.0 * bb / (aa - disc);
      }
      const double tn = tc + eta;
      const double mn = neg ? -tn : tn;
      if (isfinite(mn) && mn > 0.0) {
        unsigned long long pb = __double_as_longlong(mn);
        const unsigned long long dist = pb > lastb ? pb - lastb : lastb - pb;
        // converged: a step (after a first middle-way step) of at most 8 bit patterns
        if (nmw >= 1 && dist <= 8ull) {
          conv = true;
          gdir = side;
          pb = gdir > 0 ? lob + 1ull : hib - 1ull;
          step = 2ull;
          if (pb > lob && pb < hib) mb = pb;
        } else if (pb > lob && pb < hib) {
          mb = pb;
          ++nmw;
        } else if (nmw >= 1 && dist <= (1ull << 20)) {
          // a proposal just past the end the last point set: the steps have converged
          // there, so gallop from it (a bisection would restart from the far end)
          conv = true;
          gdir = side;
          pb = gdir > 0 ? lob + 1ull : hib - 1ull;
          step = 2ull;
          if (pb > lob && pb < hib) mb = pb;
        }
      }
    }
    const double m = __longlong_as_double(mb);
    tc = neg ? -m : m;
    lastb = mb;
    sm = secular_sums<G>(dl, w, K, o, js, tc, sub);
    have = true;
    const double f = 1.0 + rho * (sm.psi + sm.phi);
    // f increases with tau: for tau = +m the root lies above m iff f < 0;
    // for tau = -m it lies at larger m iff f > 0
    const bool up = neg ? (f > 0.0) : (f < 0.0);
    if (up) {
      lob = mb;
      side = 1;
    } else {
      hib = mb;
      side = -1;
    }
  }
  const double m = __longlong_as_double(hib);
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
template <int G>
__global__ __launch_bounds__(256) void k_dc_lowner(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int K = a.K[node];
  if (blockIdx.x * (blockDim.x / G) >= K) return;
  const int ir = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int sub = threadIdx.x & (G - 1);
  const bool active = ir < K;
  const int i = active ? ir : K - 1;
  const double* dl = a.cD + nd.lo;
  const int* org = a.org + nd.lo;
  const double* tau = a.tau + nd.lo;
  double p = 1.0;
  const double di = dl[i];
  auto factor = [&](int j, int oj, double tj, double dj) {
    const double dlj = (di - dl[oj]) - tj;
    p *= (j == i) ? dlj : dlj / (di - dj);
  };
  int j = sub;
  for (; j + 3 * G < K; j += 4 * G) {
    int ov[4];
    double tv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ov[u] = org[j + u * G];
      tv[u] = tau[j + u * G];
      dv[u] = dl[j + u * G];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) factor(j + u * G, ov[u], tv[u], dv[u]);
  }
  for (; j < K; j += G) factor(j, org[j], tau[j], dl[j]);
  for (int off = G / 2; off > 0; off >>= 1) p *= __shfl_xor(p, off, G);
  if (active && sub == 0) a.what[nd.lo + i] = copysign(sqrt(fmax(-p, 0.0)), a.cw[nd.lo + i]);
}

// M5: new eigenvalues and U^T x for the tracked sources; deflated slots copy through.
// NXM >= 2 + a.nt: the accumulators (10 for up to 8 tracked sources, 14 for 12)
template <int NXM, int G>
__global__ __launch_bounds__(256) void k_dc_apply(DCArgs a) {
  const int node = blockIdx.y;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo;
  const int K = a.K[node];
  const int jr = (blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int sub = threadIdx.x & (G - 1);
  const int N = a.N, nx = 2 + a.nt;
  if (blockIdx.x * (blockDim.x / G) >= n) return;
  if (jr >= n) return;                       // lanes of a G-group share jr
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
#pragma unroll
  for (int q = 0; q < NXM; ++q) acc[q] = 0.0;
  double nrm = 0.0;
  const double lj = dl[org[j]], tj = tau[j];
  const double* cx = a.cx + lo;
  int i = sub;
  // four poles per step, loads first; the terms are added in the order of i
  for (; i + 3 * G < K; i += 4 * G) {
    double sv[4], xv[4][NXM];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      sv[u] = wh[i + u * G] / ((dl[i + u * G] - lj) - tj);
#pragma unroll
      for (int q = 0; q < NXM; ++q)
        if (q < nx) xv[u][q] = cx[(size_t)q * N + i + u * G];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      nrm += sv[u] * sv[u];
#pragma unroll
      for (int q = 0; q < NXM; ++q)
        if (q < nx) acc[q] += sv[u] * xv[u][q];
    }
  }
  for (; i < K; i += G) {
    double s = wh[i] / ((dl[i] - lj) - tj);
    nrm += s * s;
#pragma unroll
    for (int q = 0; q < NXM; ++q)
      if (q < nx) acc[q] += s * cx[(size_t)q * N + i];
  }
  for (int off = G / 2; off > 0; off >>= 1) {
    nrm += __shfl_xor(nrm, off, G);
#pragma unroll
    for (int q = 0; q < NXM; ++q)
      if (q < nx) acc[q] += __shfl_xor(acc[q], off, G);
  }
  const double inv = 1.0 / sqrt(nrm);
  if (sub == 0) {
    a.nl[lo + j] = dl[org[j]] + tau[j];
#pragma unroll
    for (int q = 0; q < NXM; ++q)
      if (q < nx) a.nx[(size_t)q * N + lo + j] = acc[q] * inv;
  }
  if (a.Z) {
    double* Uj = a.U + (size_t)(lo + j) * N + lo;
    for (int i = sub; i < K; i += G) Uj[i] = wh[i] / dmin_lam(dl, org, tau, i, j) * inv;
  }
}

// M6: sort the node's eigenvalues ascending (stable in the slot order) and write
// lam/FR/LR/T.  The rank of value i counts the values below it: 8 lanes share an element,
// each counting every 8th value of the node as it passes through LDS in chunks, and their
// counts are added.
constexpr int kSortLanes = 8;
__global__ __launch_bounds__(256) void k_dc_sort(DCArgs a) {
  constexpr int CH = 4096;
  __shared__ double buf[CH];
  const DCNode nd = a.nodes[blockIdx.y];
  const int n = nd.hi - nd.lo, lo = nd.lo;
  constexpr int EPB = 256 / kSortLanes;                 // elements per workgroup
  if ((int)(blockIdx.x * EPB) >= n) return;             // whole block idle (uniform)
  const int part = threadIdx.x & (kSortLanes - 1);
  const int i = blockIdx.x * EPB + threadIdx.x / kSortLanes;
  const double v = i < n ? a.nl[lo + i] : 0.0;
  int rank = 0;
  for (int k0 = 0; k0 < n; k0 += CH) {
    const int len = min(CH, n - k0);
    __syncthreads();
    for (int k = threadIdx.x; k < len; k += blockDim.x) buf[k] = a.nl[lo + k0 + k];
    __syncthreads();
    int k = part;
    for (; k + 3 * kSortLanes < len; k += 4 * kSortLanes) {
      double u[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) u[q] = buf[k + q * kSortLanes];
#pragma unroll
      for (int q = 0; q < 4; ++q) rank += (u[q] < v) || (u[q] == v && k0 + k + q * kSortLanes < i);
    }
    for (; k < len; k += kSortLanes) {
      const double u = buf[k];
      rank += (u < v) || (u == v && k0 + k < i);
    }
  }
  for (int off = kSortLanes / 2; off > 0; off >>= 1) rank += __shfl_xor(rank, off, kSortLanes);
  if (i >= n || part != 0) return;
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

constexpr int kDcvSteps = 8;

// Z(:, lo + rank(j)) = Z2(:, lo + cmap(0..K-1)) U(0..K-1, j) for the non-deflated
// slots j < K; 64 x 64 output tiles, four waves of 32 x 32 on v_mfma_f64_16x16x4
__global__ __launch_bounds__(256) void k_dcv_gemm(DCArgs a) {
  const int node = blockIdx.z;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo, K = a.K[node], lo = nd.lo;
  const int r0 = blockIdx.x * 64, j0 = blockIdx.y * 64;
  if (r0 >= n || j0 >= K) return;
  // the secular roots' sorted positions increase with the slot: a tile whose first
  // column lies past the wanted ones has no wanted column
  if (a.rank[lo + j0] >= a.zcols) return;
  const size_t N = a.N;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int rb = r0 + (wave >> 1) * 32, cb = j0 + (wave & 1) * 32;
  d4 acc[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) acc[x][y] = d4{0, 0, 0, 0};
  // K in groups of kDcvSteps MFMA steps: a group's operand loads are all issued before its
  // products, and the next group's column map is loaded while they run (each accumulator
  // still takes its steps in ascending k)
  int cm[kDcvSteps];
#pragma unroll
  for (int t = 0; t < kDcvSteps; ++t) {
    const int k = 4 * t + lk;
    cm[t] = (k < K) ? a.cmap[lo + k] : -1;
  }
  for (int k0 = 0; k0 < K; k0 += 4 * kDcvSteps) {
    double av[kDcvSteps][2], bv[kDcvSteps][2];
#pragma unroll
    for (int t = 0; t < kDcvSteps; ++t) {
      const int k = k0 + 4 * t + lk;
      const double* zc_ = a.Z2 + (size_t)(lo + (cm[t] < 0 ? 0 : cm[t])) * N + lo;
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        const int r = rb + x * 16 + lr;
        av[t][x] = (cm[t] >= 0 && r < n) ? zc_[r] : 0.0;
      }
#pragma unroll
      for (int y = 0; y < 2; ++y) {
        const int c = cb + y * 16 + lr;
        bv[t][y] = (k < K && c < K) ? a.U[(size_t)(lo + c) * N + lo + k] : 0.0;
      }
    }
#pragma unroll
    for (int t = 0; t < kDcvSteps; ++t) {
      const int k = k0 + 4 * kDcvSteps + 4 * t + lk;
      cm[t] = (k < K) ? a.cmap[lo + k] : -1;
    }
#pragma unroll
    for (int t = 0; t < kDcvSteps; ++t) {
      if (k0 + 4 * t >= K) break;                  // the same steps as one step per 4 of K
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = mfma16x16x4(av[t][x], bv[t][y], acc[x][y]);
    }
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

// The same product for the large merges: 128 x 128 output tiles (four waves of 64 x 64)
// with each 16-deep slice of K staged once per workgroup in LDS, so the workgroup reads
// its rows of Z2 and columns of U once instead of once per wave pair (half the traffic of
// the 64 x 64 tiles).  Every output block takes its K steps in the same order as there, so
// the results are bitwise those of k_dcv_gemm.
constexpr int kDcvT = 128;     // output tile edge
constexpr int kDcvK = 16;      // K per LDS slice (four MFMA steps)
__global__ __launch_bounds__(256) void k_dcv_gemm128(DCArgs a) {
  const int node = blockIdx.z;
  const DCNode nd = a.nodes[node];
  const int n = nd.hi - nd.lo, K = a.K[node], lo = nd.lo;
  const int r0 = blockIdx.x * kDcvT, j0 = blockIdx.y * kDcvT;
  if (r0 >= n || j0 >= K) return;
  if (a.rank[lo + j0] >= a.zcols) return;          // no wanted column in the tile
  const size_t N = a.N;
  // one row of padding: the four k rows a wave reads together fall in different banks
  __shared__ double As[kDcvK][kDcvT + 1];
  __shared__ double Bs[kDcvK][kDcvT + 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int rw = (wave >> 1) * 64, cw = (wave & 1) * 64;      // the wave's 64 x 64 block
  // loaders: A(r, k) for r = t % T and k = t / T + (256 / T) u; B(k, c) for k = t % KS and
  // c = t / KS + (256 / KS) u (T the tile edge, KS the slice depth)
  static_assert(256 % kDcvT == 0 && 256 % kDcvK == 0 && (kDcvT * kDcvK) % 256 == 0 && kDcvK % 4 == 0,
                "loader shapes");
  const int ar = t % kDcvT, ak = t / kDcvT;
  const int bk = t % kDcvK, bc = t / kDcvK;
  constexpr int NA = kDcvT * kDcvK / 256, NB = kDcvT * kDcvK / 256;
  double ra[NA], rb[NB];
  auto gload = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int k = k0 + ak + (256 / kDcvT) * u;
      const int r = r0 + ar;
      ra[u] = (k < K && r < n) ? a.Z2[(size_t)(lo + a.cmap[lo + k]) * N + lo + r] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int k = k0 + bk;
      const int c = j0 + bc + (256 / kDcvK) * u;
      rb[u] = (k < K && c < K) ? a.U[(size_t)(lo + c) * N + lo + k] : 0.0;
    }
  };
  d4 acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = d4{0, 0, 0, 0};
  gload(0);
  for (int k0 = 0; k0 < K; k0 += kDcvK) {
    __syncthreads();                                 // the previous slice is consumed
#pragma unroll
    for (int u = 0; u < NA; ++u) As[ak + (256 / kDcvT) * u][ar] = ra[u];
#pragma unroll
    for (int u = 0; u < NB; ++u) Bs[bk][bc + (256 / kDcvK) * u] = rb[u];
    __syncthreads();
    if (k0 + kDcvK < K) gload(k0 + kDcvK);
#pragma unroll
    for (int s4 = 0; s4 < kDcvK / 4; ++s4) {
      if (k0 + 4 * s4 >= K) break;                   // the same steps as one step per 4 of K
      const int kr = 4 * s4 + lk;
      double av[4], bv[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) av[x] = As[kr][rw + 16 * x + lr];
#pragma unroll
      for (int y = 0; y < 4; ++y) bv[y] = Bs[kr][cw + 16 * y + lr];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = mfma16x16x4(av[x], bv[y], acc[x][y]);
    }
  }
#pragma unroll
  for (int y = 0; y < 4; ++y) {
    const int c = j0 + cw + 16 * y + lr;
    if (c >= K) continue;
    double* dst = a.Z + (size_t)(lo + a.rank[lo + c]) * N + lo;
#pragma unroll
    for (int x = 0; x < 4; ++x)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = r0 + rw + 16 * x + lk + 4 * q;     // f64 MFMA: register q holds row lk + 4q
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
  if (a.rank[nd.lo + j] >= a.zcols) return;
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
    b += sizeof(int) * (size_t)N * 3;          // perm cmap rank
  }
  b += sizeof(int) * (size_t)N * 2;            // rotp rotn
  b += sizeof(double) * (size_t)N * 2;         // rotc rots
  b += sizeof(int) * (N / 2 + 2) + 8 * 256;    // nrot
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
  a.perm = a.cmap = a.rank = nullptr;
  // the deflation's rotations (sources of both modes, eigenvector columns in vector mode)
  a.rotp = (int*)take(sizeof(int) * N);
  a.rotn = (int*)take(sizeof(int) * N);
  a.rotc = (double*)take(sizeof(double) * N);
  a.rots = (double*)take(sizeof(double) * N);
  a.nrot = (int*)take(sizeof(int) * (N / 2 + 2));
  if (vectors) {
    a.Z2 = (double*)take(sizeof(double) * (size_t)N * N);
    a.U = (double*)take(sizeof(double) * (size_t)N * N);
    a.perm = (int*)take(sizeof(int) * N);
    a.cmap = (int*)take(sizeof(int) * N);
    a.rank = (int*)take(sizeof(int) * N);
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
           double* d_track, double* d_lambda, double* d_Z, int zcols) {
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
    // below the root every column feeds the parent
    a.zcols = (l + 1 == tree.level_off.size() && zcols >= 0) ? zcols : N;
    dim3 g1((maxn + 255) / 256, cnt);
    dim3 g16((maxn * 16 + 255) / 256, cnt);
    dim3 g64((maxn * 64 + 255) / 256, cnt);
    const bool wide = maxn > kDCWideMin;       // one root per wave for the large merges
    hipLaunchKernelGGL(k_dc_gather, g1, dim3(256), 0, s, a);
    const int use_lds = maxn <= 8192 ? 1 : 0;
    hipLaunchKernelGGL(k_dc_deflate, dim3(cnt), dim3(256), use_lds ? 2 * sizeof(double) * maxn : 0, s,
                       a, use_lds);
    hipLaunchKernelGGL(k_dc_compact, g1, dim3(256), 0, s, a);
    if (wide) {
      hipLaunchKernelGGL(k_dc_secular<64>, g64, dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_dc_lowner<64>, g64, dim3(256), 0, s, a);
      if (tree.ntrack <= 8) hipLaunchKernelGGL((k_dc_apply<10, 64>), g64, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((k_dc_apply<14, 64>), g64, dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL(k_dc_secular<16>, g16, dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_dc_lowner<16>, g16, dim3(256), 0, s, a);
      if (tree.ntrack <= 8) hipLaunchKernelGGL((k_dc_apply<10, 16>), g16, dim3(256), 0, s, a);
      else hipLaunchKernelGGL((k_dc_apply<14, 16>), g16, dim3(256), 0, s, a);
    }
    hipLaunchKernelGGL(k_dc_sort, dim3((maxn + 256 / kSortLanes - 1) / (256 / kSortLanes), cnt), dim3(256), 0, s,
                       a);
    if (d_Z) {
      hipLaunchKernelGGL(k_dcv_gather, dim3(maxn, cnt), dim3(256), 0, s, a);
      hipLaunchKernelGGL(k_dcv_rotate, dim3((maxn + 255) / 256, cnt), dim3(256), 0, s, a);
      if (maxn > 1024) {   // below that, 128-wide tiles leave too few workgroups
        const int nt = (maxn + kDcvT - 1) / kDcvT;
        hipLaunchKernelGGL(k_dcv_gemm128, dim3(nt, nt, cnt), dim3(256), 0, s, a);
      } else {
        const int nt64 = (maxn + 63) / 64;
        hipLaunchKernelGGL(k_dcv_gemm, dim3(nt64, nt64, cnt), dim3(256), 0, s, a);
      }
      hipLaunchKernelGGL(k_dcv_copy, dim3(maxn, cnt), dim3(256), 0, s, a);
    }
    WTB_LAUNCH_CHECK();
  }
  return WTB_OK;
}

}  // namespace wtb
