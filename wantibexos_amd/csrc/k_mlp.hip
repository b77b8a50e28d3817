// Learned preconditioner for the Davidson form: a small MLP (5 -> 128 -> 128 -> 1,
// tanh, FP32) that corrects the diagonal preconditioner entry of each basis state:
//   t_i = M_i r_i,  M_i = (D_i - theta) / ((D_i - theta)^2 + eta^2) + y_i / sigma,
// the first term being 1/(D_i - theta) regularised at resonance.
//
// Features of basis state i at Ritz value theta (D_i = Re H_ii; row statistics over
// j != i of the Hermitian matrix):
//   f0 = D_i - theta            f1 = sum_j |H_ij|
//   f2 = sqrt(N sum_j |H_ij|^2) f3 = N max_j |H_ij|       f4 = theta - min_i D_i
// (the N factors keep f2, f3 of order one: the kernel entries scale as 1/N_k).
// Inputs are standardised with the (mu, sd) stored in front of the weights.  The
// training target is y_i = sigma (Re[(H - theta - i eta)^-1]_ii - (D_i - theta) /
// ((D_i - theta)^2 + eta^2)) at theta = an exciton energy, eta = sigma = 0.05 eV: the
// difference between the regularised diagonal of the reduced resolvent and the
// regularised diagonal preconditioner.
//
// Parameter pack (floats): mu[5] sd[5] W1[H][5] b1[H] W2[H][H] b2[H] w3[H] b3.
// Training: one workgroup per 64 samples computes the forward pass into LDS and the
// backward pass into per-workgroup partial gradients, a second kernel sums the
// partials in a fixed order, a third applies Adam.
#include "wtb_kernels.h"

namespace wtb {

#define MLP_H 128
#define MLP_F 5
#define MLP_S 64          // samples per training workgroup
#define MLP_SIGMA 0.05f

constexpr int kOffMu = 0, kOffSd = 5, kOffW1 = 10;
constexpr int kOffB1 = kOffW1 + MLP_H * MLP_F;
constexpr int kOffW2 = kOffB1 + MLP_H;
constexpr int kOffB2 = kOffW2 + MLP_H * MLP_H;
constexpr int kOffW3 = kOffB2 + MLP_H;
constexpr int kOffB3 = kOffW3 + MLP_H;
constexpr int kNParams = kOffB3 + 1;

int mlp_nparams(int hidden) { return hidden == MLP_H ? kNParams : -1; }

__device__ inline void mlp_features(double Di, float s0, float s1, float s2, double theta, double Dmin, int N,
                                    float f[MLP_F]) {
  f[0] = (float)(Di - theta);
  f[1] = s0;
  f[2] = s1 * sqrtf((float)N);
  f[3] = s2 * (float)N;
  f[4] = (float)(theta - Dmin);
}

// y for one standardised input; P in LDS
__device__ inline float mlp_eval(const float* P, const float x[MLP_F]) {
  float h1[MLP_H];
  for (int c = 0; c < MLP_H; ++c) {
    float a = P[kOffB1 + c];
    for (int f = 0; f < MLP_F; ++f) a = fmaf(P[kOffW1 + c * MLP_F + f], x[f], a);
    h1[c] = tanhf(a);
  }
  float y = P[kOffB3];
  for (int r = 0; r < MLP_H; ++r) {
    float a = P[kOffB2 + r];
    const float* w = P + kOffW2 + r * MLP_H;
    for (int c = 0; c < MLP_H; ++c) a = fmaf(w[c], h1[c], a);
    y = fmaf(P[kOffW3 + r], tanhf(a), y);
  }
  return y;
}

__device__ inline void load_params(const float* __restrict__ Pg, float* P) {
  for (int i = threadIdx.x; i < kNParams; i += blockDim.x) P[i] = Pg[i];
  __syncthreads();
}

// T(:,q) = M(theta_idx q) R(:, idx q)
__global__ __launch_bounds__(256) void k_precond_mlp(int N, const int* __restrict__ idx, const zc* __restrict__ R,
                                                     int ld, const double* __restrict__ theta,
                                                     const double* __restrict__ D, const float* __restrict__ st,
                                                     double Dmin, const float* __restrict__ Pg, zc* __restrict__ T,
                                                     int ldt) {
  extern __shared__ float P[];
  load_params(Pg, P);
  const int q = blockIdx.y, c = idx[q];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  float f[MLP_F];
  mlp_features(D[i], st[i], st[(size_t)N + i], st[2 * (size_t)N + i], theta[c], Dmin, N, f);
  const double d = D[i] - theta[c];
  const double eta = (double)MLP_SIGMA;
  for (int k = 0; k < MLP_F; ++k) f[k] = (f[k] - P[kOffMu + k]) / P[kOffSd + k];
  const double M = d / (d * d + eta * eta) + (double)mlp_eval(P, f) / (double)MLP_SIGMA;
  T[(size_t)i + (size_t)ldt * q] = R[(size_t)i + (size_t)ld * c] * M;
}

// inference on raw features X[5][S] (feature-major)
__global__ __launch_bounds__(256) void k_mlp_forward(int S, const float* __restrict__ X, const float* __restrict__ Pg,
                                                     float* __restrict__ y) {
  extern __shared__ float P[];
  load_params(Pg, P);
  const int s = blockIdx.x * 256 + threadIdx.x;
  if (s >= S) return;
  float f[MLP_F];
  for (int k = 0; k < MLP_F; ++k) f[k] = (X[(size_t)k * S + s] - P[kOffMu + k]) / P[kOffSd + k];
  y[s] = mlp_eval(P, f);
}

// forward + backward of the mean squared error over a batch: samples perm[b0 .. b0+B),
// MLP_S per workgroup; partial gradients gp[blockIdx][kNParams], partial loss lp[blockIdx]
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(int S, const float* __restrict__ X, const float* __restrict__ t,
                                                     const int* __restrict__ perm, int b0, int B,
                                                     const float* __restrict__ Pg, float* __restrict__ gp,
                                                     float* __restrict__ lp) {
  extern __shared__ float sm[];
  float* P = sm;                               // kNParams
  float* H1 = P + ((kNParams + 3) & ~3);       // [MLP_S][MLP_H]  tanh(a1), then dA1
  float* H2 = H1 + MLP_S * MLP_H;              // [MLP_S][MLP_H]  tanh(a2), then dA2
  float* XS = H2 + MLP_S * MLP_H;              // [MLP_S][MLP_F]
  float* DY = XS + MLP_S * MLP_F;              // [MLP_S]
  load_params(Pg, P);
  const int tid = threadIdx.x;
  const int s0 = blockIdx.x * MLP_S;
  const int ns = min(MLP_S, B - s0);
  const float invB = 1.0f / (float)B;
  // forward: thread (sample s = tid / 4, quarter qd = tid % 4) computes 32 hidden units
  {
    const int sl = tid >> 2, qd = tid & 3;
    float xn[MLP_F];
    if (sl < ns) {
      const int sg = perm[b0 + s0 + sl];
      for (int k = 0; k < MLP_F; ++k) xn[k] = (X[(size_t)k * S + sg] - P[kOffMu + k]) / P[kOffSd + k];
      if (qd == 0)
        for (int k = 0; k < MLP_F; ++k) XS[sl * MLP_F + k] = xn[k];
      for (int c = qd * 32; c < qd * 32 + 32; ++c) {
        float a = P[kOffB1 + c];
        for (int f = 0; f < MLP_F; ++f) a = fmaf(P[kOffW1 + c * MLP_F + f], xn[f], a);
        H1[sl * MLP_H + c] = tanhf(a);
      }
    }
    __syncthreads();
    if (sl < ns) {
      for (int r = qd * 32; r < qd * 32 + 32; ++r) {
        float a = P[kOffB2 + r];
        const float* w = P + kOffW2 + r * MLP_H;
        const float* h = H1 + sl * MLP_H;
        for (int c = 0; c < MLP_H; ++c) a = fmaf(w[c], h[c], a);
        H2[sl * MLP_H + r] = tanhf(a);
      }
    }
    __syncthreads();
    if (tid < MLP_S) {
      float dy = 0.0f, l = 0.0f;
      if (tid < ns) {
        float y = P[kOffB3];
        for (int r = 0; r < MLP_H; ++r) y = fmaf(P[kOffW3 + r], H2[tid * MLP_H + r], y);
        const float e = y - t[perm[b0 + s0 + tid]];
        l = e * e * invB;
        dy = 2.0f * e * invB;
      }
      DY[tid] = dy;
      // loss partial: wave 0 sums its 64 lanes
      for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
      if (tid == 0) lp[blockIdx.x] = l;
    }
    __syncthreads();
  }
  float* g = gp + (size_t)blockIdx.x * kNParams;
  // dw3[r] = sum_s dy_s h2[s][r]; db3 = sum_s dy_s
  if (tid < MLP_H) {
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a = fmaf(DY[sl], H2[sl * MLP_H + tid], a);
    g[kOffW3 + tid] = a;
  }
  if (tid == MLP_H) {
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a += DY[sl];
    g[kOffB3] = a;
  }
  __syncthreads();
  // dA2[s][r] = dy_s w3[r] (1 - h2^2), in place
  for (int e = tid; e < MLP_S * MLP_H; e += 256) {
    const int sl = e / MLP_H, r = e % MLP_H;
    const float h = H2[e];
    H2[e] = sl < ns ? DY[sl] * P[kOffW3 + r] * (1.0f - h * h) : 0.0f;
  }
  __syncthreads();
  // dW2[r][c] = sum_s dA2[s][r] h1[s][c]; db2[r] = sum_s dA2[s][r]
  for (int e = tid; e < MLP_H * MLP_H; e += 256) {
    const int r = e / MLP_H, c = e % MLP_H;
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a = fmaf(H2[sl * MLP_H + r], H1[sl * MLP_H + c], a);
    g[kOffW2 + e] = a;
  }
  if (tid < MLP_H) {
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a += H2[sl * MLP_H + tid];
    g[kOffB2 + tid] = a;
  }
  // dA1[s][c] = (sum_r dA2[s][r] W2[r][c]) (1 - h1^2): computed for all (s, c), then written in place
  float dA1[MLP_S * MLP_H / 256];
  for (int k = 0; k < MLP_S * MLP_H / 256; ++k) {
    const int e = tid + 256 * k;
    const int sl = e / MLP_H, c = e % MLP_H;
    float a = 0.0f;
    for (int r = 0; r < MLP_H; ++r) a = fmaf(H2[sl * MLP_H + r], P[kOffW2 + r * MLP_H + c], a);
    const float h = H1[e];
    dA1[k] = sl < ns ? a * (1.0f - h * h) : 0.0f;
  }
  __syncthreads();
  for (int k = 0; k < MLP_S * MLP_H / 256; ++k) H1[tid + 256 * k] = dA1[k];
  __syncthreads();
  // dW1[c][f] = sum_s dA1[s][c] x[s][f]; db1[c] = sum_s dA1[s][c]
  for (int e = tid; e < MLP_H * MLP_F; e += 256) {
    const int c = e / MLP_F, f = e % MLP_F;
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a = fmaf(H1[sl * MLP_H + c], XS[sl * MLP_F + f], a);
    g[kOffW1 + e] = a;
  }
  if (tid < MLP_H) {
    float a = 0.0f;
    for (int sl = 0; sl < ns; ++sl) a += H1[sl * MLP_H + tid];
    g[kOffB1 + tid] = a;
  }
  if (tid < 10) g[tid] = 0.0f;   // mu, sd are not trained
}

// grad[p] = sum over workgroups, in workgroup order; loss = sum of partial losses
__global__ __launch_bounds__(256) void k_mlp_reduce(int nwg, const float* __restrict__ gp, const float* __restrict__ lp,
                                                    float* __restrict__ grad, float* __restrict__ loss) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < kNParams) {
    float a = 0.0f;
    for (int w = 0; w < nwg; ++w) a += gp[(size_t)w * kNParams + p];
    grad[p] = a;
  }
  if (p == 0) {
    float l = 0.0f;
    for (int w = 0; w < nwg; ++w) l += lp[w];
    loss[0] = l;
  }
}

// Adam on the trainable parameters (all but mu, sd)
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ P, const float* __restrict__ grad,
                                                  float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < 10 || p >= kNParams) return;
  const float gr = grad[p];
  const float mi = b1 * m[p] + (1.0f - b1) * gr;
  const float vi = b2 * v[p] + (1.0f - b2) * gr * gr;
  m[p] = mi;
  v[p] = vi;
  P[p] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static size_t fwd_bwd_lds() {
  return sizeof(float) * (((kNParams + 3) & ~3) + 2 * MLP_S * MLP_H + MLP_S * MLP_F + MLP_S);
}

int mlp_setup() {
  WTB_HIP(hipFuncSetAttribute((const void*)k_precond_mlp, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(sizeof(float) * kNParams)));
  WTB_HIP(hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(sizeof(float) * kNParams)));
  WTB_HIP(hipFuncSetAttribute((const void*)k_mlp_fwd_bwd, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)fwd_bwd_lds()));
  return WTB_OK;
}

int launch_precond_mlp(hipStream_t s, int N, int nq, const int* idx, const zc* R, int ld, const double* theta,
                       const double* D, const float* stats, double Dmin, const MlpDev& mlp, zc* T, int ldt) {
  if (nq <= 0) return WTB_OK;
  if (mlp.hidden != MLP_H) {
    set_error("learned preconditioner: hidden width must be 128");
    return WTB_ERR_ARG;
  }
  int rc = mlp_setup();
  if (rc) return rc;
  hipLaunchKernelGGL(k_precond_mlp, dim3((N + 255) / 256, nq), dim3(256), sizeof(float) * kNParams, s, N, idx, R,
                     ld, theta, D, stats, Dmin, mlp.w, T, ldt);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int mlp_forward(hipStream_t s, int S, const float* X, const float* P, float* y) {
  int rc = mlp_setup();
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_forward, dim3((S + 255) / 256), dim3(256), sizeof(float) * kNParams, s, S, X, P, y);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

size_t mlp_grad_work_floats(int B) { return (size_t)((B + MLP_S - 1) / MLP_S) * (kNParams + 1) + 64; }

int mlp_loss_grad(hipStream_t s, int S, const float* X, const float* t, const int* perm, int b0, int B,
                  const float* P, float* work, float* grad, float* loss) {
  int rc = mlp_setup();
  if (rc) return rc;
  const int nwg = (B + MLP_S - 1) / MLP_S;
  float* gp = work;
  float* lp = work + (size_t)nwg * kNParams;
  hipLaunchKernelGGL(k_mlp_fwd_bwd, dim3(nwg), dim3(256), fwd_bwd_lds(), s, S, X, t, perm, b0, B, P, gp, lp);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParams + 255) / 256), dim3(256), 0, s, nwg, gp, lp, grad, loss);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

int mlp_adam(hipStream_t s, float* P, const float* grad, float* m, float* v, float lr, int step) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNParams + 255) / 256), dim3(256), 0, s, P, grad, m, v, lr, b1, b2, eps, c1,
                     c2);
  WTB_LAUNCH_CHECK();
  return WTB_OK;
}

}  // namespace wtb
