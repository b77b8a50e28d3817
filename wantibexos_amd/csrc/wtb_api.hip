// C-ABI of the BSE pipeline (include/wtb_bse.h): context, plan, entry points.
#include "wtb_kernels.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace wtb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

int launch_xth(hipStream_t s, int n, const zc* X, const zc* h1, const zc* h2, double* act);

__global__ void k_zero_int(int* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_split_g(int N, int nrhs, const zc* __restrict__ g, double* __restrict__ track) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int q = 0; q < nrhs; ++q) {
    zc v = g[(size_t)q * N + i];
    track[(size_t)(2 * q) * N + i] = v.re;
    track[(size_t)(2 * q + 1) * N + i] = v.im;
  }
}
// lower triangle of the Hermitian matrix whose upper triangle is given (ZHEEV 'U')
__global__ void k_upper_to_lower(int n, const zc* __restrict__ a, int lda, zc* __restrict__ h, int ldh) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int j = blockIdx.y;
  if (i >= n || i < j) return;
  zc v = a[(size_t)j + (size_t)lda * i];      // a(j,i), j <= i
  v = conj(v);
  if (i == j) v.im = 0.0;
  h[(size_t)i + (size_t)ldh * j] = v;
}

static int status_to_code(int st) {
  if (st & (1 << WTB_ERR_NOCP)) return WTB_ERR_NOCP;
  if (st & (1 << WTB_ERR_NVAL)) return WTB_ERR_NVAL;
  if (st & (1 << WTB_ERR_NCOND)) return WTB_ERR_NCOND;
  if (st & (1 << WTB_ERR_EIG)) return WTB_ERR_EIG;
  return WTB_OK;
}
static const char* status_text(int code) {
  switch (code) {
    case WTB_ERR_NOCP: return "no band above the Fermi level at some k (nocp undefined)";
    case WTB_ERR_NVAL: return "Number of valence states higher then ocuppied states";
    case WTB_ERR_NCOND: return "Number of conduction states higher then unocuppied states";
    case WTB_ERR_EIG: return "The algorithm failed to compute eigenvalues.";
    default: return "";
  }
}

template <class T>
static int dmalloc(T** p, size_t count) {
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * count);
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(") + std::to_string(sizeof(T) * count) + " bytes): " +
              hipGetErrorString(e));
    *p = nullptr;
    return WTB_ERR_NOMEM;
  }
  return WTB_OK;
}

// Scratch memory of wtb_bse_spectrum_noalloc: a file-scope device array (every GPU the
// library runs on holds its own copy until the process ends).  Host code never uses the
// array's name as an address: wtb_create looks the copy of the current device up with
// hipGetSymbolAddress once per device and keeps that pointer in g_scratch_dev.
constexpr size_t kScratchDoubles = size_t(1) << 20;   // 8 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// the scratch array's address on `device` (the current device), looked up once
static int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("scratch: device index beyond the library's table");
    return WTB_ERR_LIMIT;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_dev[device]) return WTB_OK;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return WTB_ERR_HIP;
  }
  g_scratch_dev[device] = static_cast<double*>(ptr);
  return WTB_OK;
}

}  // namespace wtb

using namespace wtb;

struct wtb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};

struct wtb_bse_plan {
  wtb_ctx* ctx = nullptr;
  wtb_bse_params prm{};
  bool filled = false;           // the k grid, bands and dipoles have been computed (build/execute)
  bool sp_only = false;          // single-particle table only: no BSE matrix or solver buffers
  int nk = 0, nw = 0, nvec = 0, nb = 0, N = 0;
  double ef32 = 0, scs = 0, sme32 = 0, vc3d = 0, nkf32 = 0;
  // finite temperature: gap correction (added to scs), 1/(kB T) of the Fermi factors,
  // and the tracked vectors: h (3) at T = 0, h and f h (6) otherwise
  double tcor = 0, beta = 0;
  int nrhs = 3;
  CoulombArgs ca{};
  DevModel dm;
  double* kpt = nullptr;
  zc* phase = nullptr;
  zc* hk4 = nullptr;
  double* eall = nullptr;
  int* nocp = nullptr;
  double* eigv = nullptr;
  zc* vec = nullptr;
  zc* dip = nullptr;
  zc* g = nullptr;
  double* fdeh = nullptr;   // f_v - f_c per basis state (finite temperature)
  // finite-Q exciton bands (q mode): the q-path row (device), the k+q grid and its band
  // window, the exchange vector a_I = <c_I|v_I> and V(q)
  int qmode = 0;
  double* dq = nullptr;
  double* kq = nullptr;
  double* eigvq = nullptr;
  zc* vecq = nullptr;
  int* nocpq = nullptr;
  zc* xa = nullptr;
  double* dvq = nullptr;
  double* gapk = nullptr;
  double* wtab = nullptr;
  int wtab_n = 0;
  zc* H = nullptr;      // FP64 BSE matrix (plain form)
  zf* Hf = nullptr;     // FP32 BSE matrix (precision form)
  bool mixed = false;
  size_t ldh = 0;
  void* triwork = nullptr;
  TridiagWork tw;
  void* dcwork = nullptr;
  DCTree tree;
  double* track = nullptr;
  double* lambda = nullptr;
  double* act = nullptr;
  double* wq = nullptr;
  double* actq = nullptr;
  double* eps = nullptr;
  double* smeavg = nullptr;
  int* status = nullptr;
  hipStream_t stream = nullptr;             // launches go here (the context stream unless batched)
  hipStream_t lz_stream = nullptr;          // iterative form: the Lanczos chains, beside the Davidson
  // exciton eigenvectors [wf_first, wf_first + wf_count) (BSE_WF); zero count: none formed
  int wf_first = 0, wf_count = 0;
  double* Zv = nullptr;                     // N x N tridiagonal eigenvectors
  zc* Xv = nullptr;                         // N x wf_count exciton vectors
  // precision form: FP64 Rayleigh-quotient refinement of the nref lowest excitons (0: off).
  // H holds the FP64 matrix (full Hermitian) beside the reduced COMPLEX*8 copy; Xr the
  // back-transformed vectors, HXr = H Xr, rq_res the residual norms |H x - theta x| / |x|
  // stage (d1) form: two-stage (dense -> band -> tridiagonal) when the path needs no
  // eigenvectors, one-stage Householder otherwise (its reflectors feed the back-transform)
  int two_stage = -1;          // -1: two-stage from kTwoStageMinN on, 0: one-stage, 1: two-stage
  int band_b = 32;
  BandWork bw;
  void* bandwork = nullptr;
  int nref = 0;
  zc* Xr = nullptr;
  zc* HXr = nullptr;
  void* btw_r = nullptr;
  zc* hw_r = nullptr;
  double* rq_res = nullptr;
  double* rq_lam0 = nullptr;                // the unrefined energies of those excitons
  void* btw = nullptr;                      // back-transform workspace
  bool timing = false;
  bool kernel_timing = false;              // per-launch events on the GEMV kernel
  std::vector<hipEvent_t> kev;             // 2 per tridiagonal column
  double kernel_ms = 0.0;
  int kernel_launches = 0;
  double kernel_flops = 0.0;                // two-stage: flops of the timed trailing updates
  bool use_graph = true;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipEvent_t ev[WTB_NSTAGES + 1] = {};
  double stage_ms[WTB_NSTAGES] = {};
  // iterative form: device buffers kept between calls of wtb_bse_plan_lowest (allocated at
  // its first call, grown when a later call needs more)
  DavidsonWork* dav = nullptr;
  double* lw_D = nullptr;       // diagonal of H [N]
  float* lw_stats = nullptr;    // row statistics [3][N]
  double* lw_W = nullptr;       // Ritz values [nev]
  double* lw_R = nullptr;       // residual norms [nev]
  double* lw_sm = nullptr;      // width and E_max of the spectrum [4]
  zc* lw_X = nullptr;           // Ritz vectors [nev][N]
  void* lw_hay = nullptr;       // Lanczos chains
  size_t lw_nev = 0, lw_xcols = 0, lw_hay_bytes = 0;
};

extern "C" {

const char* wtb_last_error(void) { return g_err.c_str(); }
int wtb_version(void) { return 10000; }

int wtb_create(int device, wtb_ctx** out) {
  if (!out) {
    set_error("wtb_create: null output pointer");
    return WTB_ERR_ARG;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    set_error("wtb_create: no HIP device available (the library runs on MI355X/gfx950 only)");
    return WTB_ERR_HIP;
  }
  if (device < 0 || device >= ndev) {
    set_error("wtb_create: device index out of range");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  WTB_HIP(hipGetDeviceProperties(&prop, device));
  if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) {
    set_error(std::string("wtb_create: device is ") + prop.gcnArchName + ", this library is built for gfx950");
    return WTB_ERR_HIP;
  }
  // the no-allocation entry point's scratch array on this device (device is current)
  int rc = scratch_setup(device);
  if (rc) return rc;
  wtb_ctx* c = new wtb_ctx;
  c->device = device;
  WTB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = c;
  return WTB_OK;
}

int wtb_destroy(wtb_ctx* c) {
  if (!c) return WTB_OK;
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return WTB_OK;
}

void* wtb_stream(wtb_ctx* c) { return c ? (void*)c->stream : nullptr; }

int wtb_sync(wtb_ctx* c) {
  if (!c) return WTB_ERR_ARG;
  WTB_HIP(hipStreamSynchronize(c->stream));
  return WTB_OK;
}

// ---------------------------------------------------------------------------
int wtb_monkhorst_pack(wtb_ctx* c, const int ngrid[3], const double shift[3], const double* rlat,
                       double* kpt) {
  if (!c || !ngrid || !shift || !rlat || !kpt) {
    set_error("wtb_monkhorst_pack: null argument");
    return WTB_ERR_ARG;
  }
  if (ngrid[0] < 1 || ngrid[1] < 1 || ngrid[2] < 1) {
    set_error("wtb_monkhorst_pack: grid sizes must be >= 1");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  int nk = ngrid[0] * ngrid[1] * ngrid[2];
  double *d_rl = nullptr, *d_k = nullptr;
  int rc;
  if ((rc = dmalloc(&d_rl, 9)) || (rc = dmalloc(&d_k, (size_t)nk * 3))) {
    hipFree(d_rl);
    return rc;
  }
  hipMemcpyAsync(d_rl, rlat, 9 * sizeof(double), hipMemcpyHostToDevice, c->stream);
  rc = launch_kgrid(c->stream, d_rl, ngrid, shift, d_k);
  if (rc == WTB_OK) {
    hipMemcpyAsync(kpt, d_k, sizeof(double) * nk * 3, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = WTB_ERR_HIP;
  }
  hipFree(d_rl);
  hipFree(d_k);
  return rc;
}

static int check_model(const wtb_tb_model* m) {
  if (!m || !m->rlat || !m->rvec || !m->ffactor || !m->hop || !m->ihop) {
    set_error("tight-binding model: null array");
    return WTB_ERR_ARG;
  }
  if (m->nw < 1 || m->nvec < 1) {
    set_error("tight-binding model: nw and nvec must be >= 1");
    return WTB_ERR_ARG;
  }
  if (m->nw > 64) {
    set_error("tight-binding model: nw > 64 exceeds the LDS-resident per-k eigensolver");
    return WTB_ERR_LIMIT;
  }
  size_t cnt = (size_t)m->nvec * m->nw * m->nw;
  for (size_t i = 0; i < cnt; ++i)
    if (!std::isfinite(m->hop[i]) || !std::isfinite(m->ihop[i])) {
      set_error("tight-binding model: NaN or Inf in H(R)");
      return WTB_ERR_NONFINITE;
    }
  for (int i = 0; i < m->nvec; ++i)
    if (m->ffactor[i] == 0) {
      set_error("tight-binding model: zero degeneracy factor");
      return WTB_ERR_ARG;
    }
  if (!std::isfinite(m->efermi) || !std::isfinite(m->scs)) {
    set_error("tight-binding model: non-finite E_F or scissor");
    return WTB_ERR_NONFINITE;
  }
  return WTB_OK;
}

static int upload_model(hipStream_t s, const wtb_tb_model* m, DevModel& dm) {
  int rc;
  dm.nw = m->nw;
  dm.nvec = m->nvec;
  if ((rc = dmalloc(&dm.rlat, 9)) || (rc = dmalloc(&dm.rvec, (size_t)m->nvec * 3)) ||
      (rc = dmalloc(&dm.hstack, (size_t)m->nvec * m->nw * m->nw)))
    return rc;
  size_t cnt = (size_t)m->nvec * m->nw * m->nw;
  double *dh = nullptr, *dih = nullptr;
  int* dff = nullptr;
  if ((rc = dmalloc(&dh, cnt)) || (rc = dmalloc(&dih, cnt)) || (rc = dmalloc(&dff, m->nvec))) {
    hipFree(dh);
    hipFree(dih);
    return rc;
  }
  WTB_HIP(hipMemcpy(dm.rlat, m->rlat, 9 * sizeof(double), hipMemcpyHostToDevice));
  WTB_HIP(hipMemcpy(dm.rvec, m->rvec, sizeof(int) * 3 * m->nvec, hipMemcpyHostToDevice));
  WTB_HIP(hipMemcpy(dh, m->hop, sizeof(double) * cnt, hipMemcpyHostToDevice));
  WTB_HIP(hipMemcpy(dih, m->ihop, sizeof(double) * cnt, hipMemcpyHostToDevice));
  WTB_HIP(hipMemcpy(dff, m->ffactor, sizeof(int) * m->nvec, hipMemcpyHostToDevice));
  rc = launch_hr_stack(s, dh, dih, dff, m->nvec, m->nw, dm.hstack);
  hipStreamSynchronize(s);
  hipFree(dh);
  hipFree(dih);
  hipFree(dff);
  return rc;
}

static void free_model(DevModel& dm) {
  hipFree(dm.rlat);
  hipFree(dm.rvec);
  hipFree(dm.hstack);
  dm = DevModel();
}

int wtb_eigsys(wtb_ctx* c, const wtb_tb_model* m, int nk, const double* kpt, double* energias,
               double* autovetores, int* nocp) {
  if (!c || !kpt || !energias || nk < 1) {
    set_error("wtb_eigsys: bad argument");
    return WTB_ERR_ARG;
  }
  int rc = check_model(m);
  if (rc) return rc;
  WTB_HIP(hipSetDevice(c->device));
  DevModel dm;
  if ((rc = upload_model(c->stream, m, dm))) {
    free_model(dm);
    return rc;
  }
  const int nw = m->nw;
  double *dk = nullptr, *de = nullptr;
  zc *dph = nullptr, *dhk = nullptr, *du = nullptr;
  int *dn = nullptr, *dst = nullptr;
  if ((rc = dmalloc(&dk, (size_t)nk * 3)) || (rc = dmalloc(&dph, (size_t)4 * nk * m->nvec)) ||
      (rc = dmalloc(&dhk, (size_t)4 * nk * nw * nw)) || (rc = dmalloc(&de, (size_t)nk * nw)) ||
      (rc = dmalloc(&du, (size_t)nk * nw * nw)) || (rc = dmalloc(&dn, nk)) || (rc = dmalloc(&dst, 1))) {
  } else {
    hipMemcpyAsync(dk, kpt, sizeof(double) * nk * 3, hipMemcpyHostToDevice, c->stream);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c->stream, dst, 1);
    double ef32 = (double)(float)(-m->efermi);
    rc = launch_hk(c->stream, dm, dk, nk, dph, dhk);
    if (!rc)
      rc = launch_jacobi(c->stream, dhk, nk, nw, ef32, m->scs, 0, 0, de, du, dn, nullptr, nullptr, dst);
    int st = 0;
    if (!rc) {
      hipMemcpyAsync(energias, de, sizeof(double) * nk * nw, hipMemcpyDeviceToHost, c->stream);
      if (autovetores)
        hipMemcpyAsync(autovetores, du, sizeof(zc) * nk * nw * nw, hipMemcpyDeviceToHost, c->stream);
      if (nocp) hipMemcpyAsync(nocp, dn, sizeof(int) * nk, hipMemcpyDeviceToHost, c->stream);
      hipMemcpyAsync(&st, dst, sizeof(int), hipMemcpyDeviceToHost, c->stream);
      if (hipStreamSynchronize(c->stream) != hipSuccess) {
        set_error("wtb_eigsys: stream error");
        rc = WTB_ERR_HIP;
      }
    }
    if (!rc) {
      st &= (1 << WTB_ERR_NOCP) | (1 << WTB_ERR_EIG);
      rc = status_to_code(st);
      if (rc) set_error(status_text(rc));
    }
  }
  hipFree(dk);
  hipFree(dph);
  hipFree(dhk);
  hipFree(de);
  hipFree(du);
  hipFree(dn);
  hipFree(dst);
  free_model(dm);
  return rc;
}

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
static void plan_free(wtb_bse_plan* p) {
  free_model(p->dm);
  void* bufs[] = {p->kpt, p->phase, p->hk4, p->eall, p->nocp, p->eigv, p->vec, p->dip, p->g,
                  p->fdeh, p->gapk, p->wtab, p->H, p->Hf, p->triwork, p->dcwork, p->track, p->lambda, p->act,
                  p->wq, p->actq, p->eps, p->smeavg, p->status, p->Zv, p->Xv, p->btw,
                  p->dq, p->kq, p->eigvq, p->vecq, p->nocpq, p->xa, p->dvq,
                  p->lw_D, p->lw_stats, p->lw_W, p->lw_R, p->lw_sm, p->lw_X, p->lw_hay,
                  p->Xr, p->HXr, p->btw_r, p->hw_r, p->rq_res, p->rq_lam0, p->bandwork};
  for (void* b : bufs) hipFree(b);
  davidson_work_delete(p->dav);
  for (auto& e : p->ev)
    if (e) hipEventDestroy(e);
  for (auto& e : p->kev)
    if (e) hipEventDestroy(e);
  if (p->gexec) hipGraphExecDestroy(p->gexec);
  if (p->graph) hipGraphDestroy(p->graph);
}

static double vcell3d_host(const double* rl) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) a[i][c] = rl[i + 3 * c];
  double vx0 = (a[0][1] * a[1][2]) - (a[0][2] * a[1][1]);
  double vx1 = (a[0][2] * a[1][0]) - (a[0][0] * a[1][2]);
  double vx2 = (a[0][0] * a[1][1]) - (a[0][1] * a[1][0]);
  return std::fabs(vx0 * a[2][0] + vx1 * a[2][1] + vx2 * a[2][2]);
}

// k_B of general_subs.f90:575,594,613: a REAL literal stored in a double parameter
static const double kKbF32 = (double)8.617330350e-5f;

// tcor of bse_solver-tool-diel-temp.f90:262-283: "VE" gapcortemp (general_subs.f90:569-586),
// "BE" gapcortemp2 (:588-604), otherwise 0
static double gap_correction(int ta, double st, double phavg, double temp) {
  if (ta == WTB_TA_VE) {
    const double aux1 = phavg / (2.0 * kKbF32 * temp);
    const double aux2 = 1.0 / std::tanh(aux1);
    return -(st * phavg * (aux2 - 1.0));
  }
  if (ta == WTB_TA_BE) {
    const double aux1 = 2.0 * st;
    const double aux2 = std::exp(phavg / temp);
    return -(aux1 / (aux2 - 1.0));
  }
  return 0.0;
}

// beta = 1/(k_B T) of fermidisteh (general_subs.f90:606-635); T = 0 is taken as the REAL 1E-8
static double fermi_beta(double temp) {
  const double t = temp == 0.0 ? (double)1e-8f : temp;
  return 1.0 / (kKbF32 * t);
}

// sp_only: a plan for the single-particle table alone (bands, dipoles, Fermi factors):
// no BSE matrix, reduction or divide-and-conquer buffers, O(N) device memory as the
// reference's spoptics (sp_solver-tool-diel.f90:137-196)
static int plan_create_impl(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, wtb_bse_plan** out,
                            bool sp_only);

int wtb_bse_plan_create(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm,
                        wtb_bse_plan** out) {
  return plan_create_impl(c, m, prm, out, false);
}

static int plan_create_impl(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, wtb_bse_plan** out,
                            bool sp_only) {
  if (!c || !prm || !out) {
    set_error("wtb_bse_plan_create: null argument");
    return WTB_ERR_ARG;
  }
  *out = nullptr;
  int rc = check_model(m);
  if (rc) return rc;
  if (prm->ngrid[0] < 1 || prm->ngrid[1] < 1 || prm->ngrid[2] < 1 || prm->nc < 1 || prm->nv < 1 ||
      prm->numbse < 2) {
    set_error("wtb_bse_plan_create: grid, nc, nv must be >= 1 and NESPEC >= 2");
    return WTB_ERR_ARG;
  }
  if (prm->coultype < WTB_V2DK || prm->coultype > WTB_V0DT) {
    set_error("Wrong Coulomb Potential");
    return WTB_ERR_COULOMB;
  }
  if (prm->nc + prm->nv > m->nw) {
    set_error("nc + nv exceeds the number of bands");
    return WTB_ERR_ARG;
  }
  if (prm->precision != WTB_FP64 && prm->precision != WTB_MIXED) {
    set_error("wtb_bse_plan_create: unknown precision form");
    return WTB_ERR_ARG;
  }
  double tcor = 0.0;
  if (prm->bset) {
    if (prm->bset != 1 || prm->ta < WTB_TA_FA || prm->ta > WTB_TA_BE || !(prm->temp >= 0.0) ||
        !std::isfinite(prm->temp) || !std::isfinite(prm->st) || !std::isfinite(prm->phavg)) {
      set_error("wtb_bse_plan_create: finite temperature needs bset = 1, TA in {FA, VE, BE}, finite TEMP >= 0, "
                "ST, PHAVG");
      return WTB_ERR_ARG;
    }
    tcor = gap_correction(prm->ta, prm->st, prm->phavg, prm->temp);
    if (!std::isfinite(tcor)) {
      set_error("wtb_bse_plan_create: the gap correction is not finite (TEMP = 0 with PHAVG = 0?)");
      return WTB_ERR_NONFINITE;
    }
  }
  WTB_HIP(hipSetDevice(c->device));
  wtb_bse_plan* p = new wtb_bse_plan;
  p->ctx = c;
  p->stream = c->stream;
  p->prm = *prm;
  p->mixed = prm->precision == WTB_MIXED;
  p->sp_only = sp_only;
  p->nk = prm->ngrid[0] * prm->ngrid[1] * prm->ngrid[2];
  p->nw = m->nw;
  p->nvec = m->nvec;
  p->nb = prm->nc + prm->nv;
  p->N = p->nk * prm->nc * prm->nv;
  p->ef32 = (double)(float)(-m->efermi);
  // eigsys(scs + tcor, ...) (bse_solver-tool-diel-temp.f90:289)
  p->tcor = tcor;
  p->scs = m->scs + tcor;
  p->nrhs = prm->bset ? 6 : 3;
  p->beta = prm->bset ? fermi_beta(prm->temp) : 0.0;
  p->sme32 = (double)(float)prm->cshift;
  p->vc3d = vcell3d_host(m->rlat);
  p->nkf32 = (double)(float)p->nk;
  p->ca.type = prm->coultype;
  p->ca.n1 = prm->ngrid[0];
  p->ca.n2 = prm->ngrid[1];
  p->ca.n3 = prm->ngrid[2];
  for (int i = 0; i < 3; ++i) p->ca.ediel[i] = prm->ediel[i];
  p->ca.lc = prm->lc;
  p->ca.ez = prm->ez;
  p->ca.w = prm->w;
  p->ca.r0 = prm->r0;
  p->ca.ktol = prm->ktol;
  const int nk = p->nk, nw = p->nw, N = p->N;
  p->ldh = (size_t)N;
  p->wtab_n = (2 * prm->ngrid[0] - 1) * (2 * prm->ngrid[1] - 1) * (2 * prm->ngrid[2] - 1);
  hipStream_t s = c->stream;
  const int nb_panel = 32;
  if ((rc = upload_model(s, m, p->dm)) || (rc = dmalloc(&p->kpt, (size_t)nk * 3)) ||
      (rc = dmalloc(&p->phase, (size_t)4 * nk * p->nvec)) ||
      (rc = dmalloc(&p->hk4, (size_t)4 * nk * nw * nw)) || (rc = dmalloc(&p->eall, (size_t)nk * nw)) ||
      (rc = dmalloc(&p->nocp, nk)) || (rc = dmalloc(&p->eigv, (size_t)nk * p->nb)) ||
      (rc = dmalloc(&p->vec, (size_t)nk * p->nb * nw)) || (rc = dmalloc(&p->dip, (size_t)3 * N)) ||
      (rc = dmalloc(&p->g, (size_t)p->nrhs * N)) || (rc = dmalloc(&p->gapk, nk)) ||
      (rc = prm->bset ? dmalloc(&p->fdeh, N) : WTB_OK) || (rc = dmalloc(&p->status, 4))) {
    plan_free(p);
    delete p;
    return rc;
  }
  if (sp_only) {
    for (auto& e : p->ev) hipEventCreate(&e);
    *out = p;
    return WTB_OK;
  }
  if ((rc = dmalloc(&p->wtab, p->wtab_n)) ||
      (rc = (prm->precision == WTB_MIXED) ? dmalloc(&p->Hf, p->ldh * N) : dmalloc(&p->H, p->ldh * N)) ||
      (rc = dmalloc((char**)&p->triwork, tridiag_work_bytes(N, nb_panel, p->nrhs))) ||
      (rc = dmalloc((char**)&p->dcwork, dc_work_bytes(N, 2 * p->nrhs))) ||
      (rc = dmalloc(&p->track, (size_t)2 * p->nrhs * N)) || (rc = dmalloc(&p->lambda, N)) ||
      (rc = dmalloc(&p->act, (size_t)8 * N)) || (rc = dmalloc(&p->wq, N)) ||
      (rc = dmalloc(&p->actq, (size_t)6 * N)) || (rc = dmalloc(&p->eps, (size_t)12 * prm->numbse)) ||
      (rc = dmalloc(&p->smeavg, 4))) {
    plan_free(p);
    delete p;
    return rc;
  }
  tridiag_bind(p->tw, p->triwork, N, nb_panel, p->nrhs);
  if (N >= 2 && ((rc = dmalloc((char**)&p->bandwork, band_work_bytes(N, p->band_b, p->nrhs))) ||
                 (rc = band_bind(p->bw, p->bandwork, N, p->band_b, p->nrhs)))) {
    plan_free(p);
    delete p;
    return rc;
  }
  // one-launch panels need every workgroup resident: at most one per compute unit of a
  // device the plan does not share (a batch sets 0: per-column launches)
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) cus = 0;
    p->bw.panel_resident = cus;
  }
  if ((rc = dc_prepare(p->dcwork, N, 2 * p->nrhs, p->tree))) {
    plan_free(p);
    delete p;
    return rc;
  }
  for (auto& e : p->ev) hipEventCreate(&e);
  *out = p;
  return WTB_OK;
}

int wtb_bse_plan_dim(const wtb_bse_plan* p) { return p ? p->N : 0; }

int wtb_bse_plan_set_timing(wtb_bse_plan* p, int on) {
  if (!p || on < 0 || on > 2) return WTB_ERR_ARG;
  p->timing = on != 0;
  p->kernel_timing = on == 2;
  if (p->kernel_timing && p->kev.empty()) {
    p->kev.assign(2 * (size_t)p->N, nullptr);
    for (auto& e : p->kev) WTB_HIP(hipEventCreate(&e));
  }
  return WTB_OK;
}

int wtb_bse_plan_kernel_ms(const wtb_bse_plan* p, double* total_ms, int* launches) {
  if (!p || !total_ms || !launches) return WTB_ERR_ARG;
  *total_ms = p->kernel_ms;
  *launches = p->kernel_launches;
  return WTB_OK;
}

int wtb_bse_plan_kernel_flops(const wtb_bse_plan* p, double* flops) {
  if (!p || !flops) return WTB_ERR_ARG;
  *flops = p->kernel_flops;
  return WTB_OK;
}

int wtb_bse_plan_stage_ms(const wtb_bse_plan* p, double* ms) {
  if (!p || !ms) return WTB_ERR_ARG;
  for (int i = 0; i < WTB_NSTAGES; ++i) ms[i] = p->stage_ms[i];
  return WTB_OK;
}

static void plan_drop_graph(wtb_bse_plan* p);

// Stage (d1): the lower triangle of the plan's matrix (FP64 or COMPLEX*8) to the real
// tridiagonal tw.d, tw.e, with Q^T applied to g (may be NULL).  Two-stage when nothing
// downstream needs the Householder reflectors (exciton vectors, the refinement); its
// bulge chase reports a stalled wait through status.
// Matrices of at least this size take the two-stage reduction by default: its MFMA panel
// updates overtake the one-stage GEMV chain between N = 5184 (one-stage faster) and
// N = 36864 (two-stage 3.4x to 4.4x faster) on MI355X.
constexpr int kTwoStageMinN = 8192;

static int plan_reduce(wtb_bse_plan* p, zc* g, int* status) {
  hipStream_t s = p->stream;
  const bool vectors = p->wf_count > 0 || (p->mixed && p->nref > 0);
  const bool two = p->two_stage == 1 || (p->two_stage < 0 && p->N >= kTwoStageMinN);
  if (two && !vectors && p->bandwork) {
    p->bw.kev = p->kernel_timing ? p->kev.data() : nullptr;
    p->bw.kev_cap = p->kernel_timing ? p->N : 0;
    // the band reduction reads and updates the lower triangle only
    if (p->mixed) return run_band_tridiag_f32(s, p->bw, p->Hf, (int)p->ldh, g, p->tw.d, p->tw.e, status);
    return run_band_tridiag(s, p->bw, p->H, (int)p->ldh, g, p->tw.d, p->tw.e, status);
  }
  return p->mixed ? run_tridiag_f32(s, p->tw, p->Hf, (int)p->ldh, g) : run_tridiag(s, p->tw, p->H, (int)p->ldh, g);
}

// q mode (bsebnds, subprograms/bse_kpath-tool.f90:187-389): per q, the band windows at k
// (valence) and k+q (conduction), the matrix with the exchange term, its eigenvalues
// (ZHEEV('N','U'), :353-389).  With bset it is bsebndstemp (bse_kpath-tool-temp.f90): the
// scissor carries tcor on both grids and the kernel terms the row factors f_v - f_c.
// status[0] holds the k grid's checks, status[1] the k+q grid's.
static int plan_enqueue_q(wtb_bse_plan* p, bool build_only) {
  hipStream_t s = p->stream;
  const wtb_bse_params& P = p->prm;
  const int nk = p->nk, nw = p->nw;
  int rc;
  auto mark = [&](int i) {
    if (p->timing) hipEventRecord(p->ev[i], s);
  };
  mark(0);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, p->status, 2);
  if ((rc = launch_kgrid(s, p->dm.rlat, P.ngrid, P.mshift, p->kpt))) return rc;
  if ((rc = launch_hk(s, p->dm, p->kpt, nk, p->phase, p->hk4))) return rc;
  if ((rc = launch_jacobi(s, p->hk4, nk, nw, p->ef32, p->scs, P.nc, P.nv, nullptr, nullptr, p->nocp, p->eigv,
                          p->vec, p->status)))
    return rc;
  if ((rc = launch_kgrid(s, p->dm.rlat, P.ngrid, P.mshift, p->kq, p->dq))) return rc;
  if ((rc = launch_hk(s, p->dm, p->kq, nk, p->phase, p->hk4))) return rc;
  if ((rc = launch_jacobi(s, p->hk4, nk, nw, p->ef32, p->scs, P.nc, P.nv, nullptr, nullptr, p->nocpq, p->eigvq,
                          p->vecq, p->status + 1)))
    return rc;
  if ((rc = launch_mix_q(s, nk, P.nc, P.nv, nw, p->eigvq, p->vecq, p->eigv, p->vec))) return rc;
  // bsebndstemp: f_v(k) - f_c(k+q) of every basis state, from the mixed window
  if (P.bset && (rc = launch_fermi_factors(s, nk, P.nc, P.nv, p->eigv, p->beta, p->fdeh))) return rc;
  mark(1);
  if ((rc = launch_exchange_vec(s, nk, P.nc, P.nv, nw, p->vec, p->xa))) return rc;
  if ((rc = launch_vq(s, p->dm.rlat, p->ca, p->dq, p->dvq))) return rc;
  mark(2);
  if ((rc = launch_wtable(s, p->dm.rlat, p->ca, p->wtab))) return rc;
  if ((rc = launch_bse_build(s, nk, P.nc, P.nv, nw, P.ngrid, p->wtab, p->eigv, p->vec, p->H, (int)p->ldh, p->Hf,
                             P.bset ? p->fdeh : nullptr, p->xa, p->dvq)))
    return rc;
  mark(3);
  if (build_only) return WTB_OK;
  p->tw.kev = p->kernel_timing ? p->kev.data() : nullptr;
  p->tw.kev_cap = p->kernel_timing ? p->N : 0;
  if ((rc = plan_reduce(p, nullptr, p->status + 1))) return rc;
  mark(4);
  if ((rc = run_dc(s, p->dcwork, p->tree, p->tw.d, p->tw.e, nullptr, p->lambda,
                   p->wf_count > 0 ? p->Zv : nullptr, p->wf_first + p->wf_count)))
    return rc;
  // BSE_WF along the path (bse_kpath-tool.f90:359-366, 402-410): the exciton vectors
  if (p->wf_count > 0) {
    rc = p->mixed ? run_backtr_f32(s, p->Hf, (int)p->ldh, p->N, p->tw.tau, p->Zv, p->wf_first, p->wf_count, p->Xv,
                                   p->N, p->btw)
                  : run_backtr(s, p->H, (int)p->ldh, p->N, p->tw.tau, p->Zv, p->wf_first, p->wf_count, p->Xv, p->N,
                               p->btw);
    if (rc) return rc;
  }
  mark(5);
  mark(6);
  return WTB_OK;
}

// the status of a q-mode run: the reference checks the valence window on the k grid and
// the conduction window on the k+q grid (quantumnumbers2(.., nocpk, nocpq, ..),
// general_subs.f90:517-529)
static int qmode_status(const int st[2]) {
  const int k_bits = st[0] & ~(1 << WTB_ERR_NCOND);
  const int q_bits = st[1] & ~(1 << WTB_ERR_NVAL);
  return k_bits | q_bits;
}

// the plan's per-k Jacobi eigensolves (stage b) as a batch item
static JacobiItem plan_jacobi_item(const wtb_bse_plan* p) {
  return JacobiItem{p->hk4, p->nk, p->nw, p->ef32, p->scs, p->prm.nc, p->prm.nv, p->eall, nullptr,
                    p->nocp, p->eigv, p->vec, p->status};
}

// stage (a): the k grid and the H(k) phase GEMM
static int plan_enqueue_pre_a(wtb_bse_plan* p) {
  hipStream_t s = p->stream;
  const wtb_bse_params& P = p->prm;
  int rc;
  if (p->timing) hipEventRecord(p->ev[0], s);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, s, p->status, 1);
  if ((rc = launch_kgrid(s, p->dm.rlat, P.ngrid, P.mshift, p->kpt))) return rc;
  return launch_hk(s, p->dm, p->kpt, p->nk, p->phase, p->hk4);
}

// stages (a)-(c) and the tracked vectors' copy: everything before the reduction (stages
// (a) and (b) only when `after_b`: the batch runs them itself)
static int plan_enqueue_pre(wtb_bse_plan* p, bool build_only, bool after_b = false) {
  hipStream_t s = p->stream;
  const wtb_bse_params& P = p->prm;
  const int nk = p->nk, nw = p->nw, N = p->N;
  int rc;
  auto mark = [&](int i) {
    if (p->timing) hipEventRecord(p->ev[i], s);
  };
  if (!after_b) {
    // (a) k grid + H(k) phase GEMM, (b) per-k Jacobi
    if ((rc = plan_enqueue_pre_a(p))) return rc;
    const JacobiItem j = plan_jacobi_item(p);
    if ((rc = launch_jacobi(s, j.hk, j.nk, j.nw, j.ef32, j.scs, j.nc, j.nv, j.eall, j.uout, j.nocp, j.eigv,
                            j.vec, j.status)))
      return rc;
  }
  mark(1);
  if ((rc = launch_dipoles(s, p->hk4, nk, nw, P.nc, P.nv, p->sme32, p->eigv, p->vec, p->dip, p->gapk)))
    return rc;
  // finite temperature: f_v - f_c of every basis state
  if (P.bset && (rc = launch_fermi_factors(s, nk, P.nc, P.nv, p->eigv, p->beta, p->fdeh))) return rc;
  mark(2);
  if (p->sp_only) return WTB_OK;
  // (c) W table + assembly (lower triangle); the refinement of the precision form also
  // keeps the FP64 matrix, filled to full Hermitian storage for its products
  if ((rc = launch_wtable(s, p->dm.rlat, p->ca, p->wtab))) return rc;
  const bool refine = p->mixed && p->nref > 0;
  if ((rc = launch_bse_build(s, nk, P.nc, P.nv, nw, P.ngrid, p->wtab, p->eigv, p->vec,
                             (p->mixed && !refine) ? nullptr : p->H, (int)p->ldh, p->Hf, p->fdeh)))
    return rc;
  if (refine && (rc = sym_fill_f64(s, N, p->H, (int)p->ldh))) return rc;
  mark(3);
  if (build_only) return WTB_OK;
  // (d) Householder tridiagonalisation, g <- Q^T h (and Q^T f h at finite temperature)
  WTB_HIP(hipMemcpyAsync(p->g, p->dip, sizeof(zc) * 3 * N, hipMemcpyDeviceToDevice, s));
  if (P.bset && (rc = launch_scale_rows(s, 3, N, p->dip, p->fdeh, p->g + (size_t)3 * N))) return rc;
  p->tw.kev = p->kernel_timing ? p->kev.data() : nullptr;
  p->tw.kev_cap = p->kernel_timing ? p->N : 0;
  return WTB_OK;
}

// whether the plan's reduction is the one-stage form, which a batch runs in lockstep
static bool plan_one_stage(const wtb_bse_plan* p) {
  if (p->qmode || p->sp_only) return false;
  const bool vectors = p->wf_count > 0 || (p->mixed && p->nref > 0);
  const bool two = p->two_stage == 1 || (p->two_stage < 0 && p->N >= kTwoStageMinN);
  return !(two && !vectors && p->bandwork);
}

// stages (d2)-(e): everything after the reduction
static int plan_enqueue_post(wtb_bse_plan* p) {
  hipStream_t s = p->stream;
  const wtb_bse_params& P = p->prm;
  const int N = p->N;
  const bool refine = p->mixed && p->nref > 0;
  int rc;
  auto mark = [&](int i) {
    if (p->timing) hipEventRecord(p->ev[i], s);
  };
  mark(4);
  // tridiagonal divide and conquer: eigenvalues and Z^T (Q^T h)
  hipLaunchKernelGGL(k_split_g, dim3((N + 255) / 256), dim3(256), 0, s, N, p->nrhs, p->g, p->track);
  // only the eigenvector columns the refinement and BSE_WF read are formed
  const int zcols = std::max(refine ? p->nref : 0, p->wf_count > 0 ? p->wf_first + p->wf_count : 0);
  if ((rc = run_dc(s, p->dcwork, p->tree, p->tw.d, p->tw.e, p->track, p->lambda,
                   (p->wf_count > 0 || refine) ? p->Zv : nullptr, zcols)))
    return rc;
  if (refine) {
    // lowest nref exciton vectors x = Q z (FP32 reflectors, FP64 arithmetic), H x with the
    // FP64 matrix, then theta = x^H H x / x^H x replaces the energy and |H x - theta x| / |x|
    // bounds its distance to an eigenvalue of the FP64 matrix
    const int nr = p->nref;
    if ((rc = run_backtr_f32(s, p->Hf, (int)p->ldh, N, p->tw.tau, p->Zv, 0, nr, p->Xr, N, p->btw_r))) return rc;
    if ((rc = hmm_f64(s, N, nr, p->H, (int)p->ldh, p->Xr, N, p->HXr, N, p->hw_r))) return rc;
    if ((rc = launch_rq_refine(s, N, nr, p->Xr, p->HXr, N, p->lambda, p->rq_lam0, p->rq_res))) return rc;
  }
  if (p->wf_count > 0) {
    rc = p->mixed ? run_backtr_f32(s, p->Hf, (int)p->ldh, N, p->tw.tau, p->Zv, p->wf_first, p->wf_count, p->Xv, N,
                                   p->btw)
                  : run_backtr(s, p->H, (int)p->ldh, N, p->tw.tau, p->Zv, p->wf_first, p->wf_count, p->Xv, N,
                               p->btw);
    if (rc) return rc;
  }
  mark(5);
  // (e) activities, F15.4 round trip, eps(omega)
  rc = P.bset ? launch_activities_temp(s, N, p->track, P.dtfull, P.cpol, p->act)
              : launch_activities(s, N, p->track, P.dtfull, P.cpol, p->act);
  if (rc) return rc;
  const double* e_in = p->lambda;
  const double* a_in = p->act;
  if (P.f154) {
    if ((rc = launch_quantize(s, N, p->lambda, p->wq))) return rc;
    if ((rc = launch_quantize(s, 6 * N, p->act, p->actq))) return rc;
    e_in = p->wq;
    a_in = p->actq;
  }
  if ((rc = launch_spectrum(s, N, e_in, a_in, p->vc3d, p->nkf32, P.ebse0, P.ebsef, P.numbse, P.cshift,
                            P.renorm, p->smeavg, p->eps)))
    return rc;
  mark(6);
  return WTB_OK;
}

static int plan_enqueue(wtb_bse_plan* p, bool build_only = false) {
  if (p->qmode) return plan_enqueue_q(p, build_only);
  int rc = plan_enqueue_pre(p, build_only);
  if (rc || build_only || p->sp_only) return rc;
  if ((rc = plan_reduce(p, p->g, p->status))) return rc;
  return plan_enqueue_post(p);
}

// The whole pipeline is captured once into a HIP graph (a few thousand kernel
// nodes for the column-by-column reduction) and replayed on later executes.
static int plan_capture(wtb_bse_plan* p) {
  hipStream_t s = p->stream;
  WTB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = plan_enqueue(p);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    return WTB_ERR_HIP;
  }
  e = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(g);
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return WTB_ERR_HIP;
  }
  p->graph = g;
  return WTB_OK;
}

int wtb_bse_plan_execute(wtb_bse_plan* p) {
  if (!p) {
    set_error("wtb_bse_plan_execute: null plan");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  if (p->timing || !p->use_graph) {
    const int rc = plan_enqueue(p);
    if (!rc) p->filled = true;
    return rc;
  }
  if (!p->gexec) {
    int rc = plan_capture(p);
    if (rc) return rc;
  }
  WTB_HIP(hipGraphLaunch(p->gexec, p->stream));
  p->filled = true;
  return WTB_OK;
}

int wtb_bse_plan_set_graph(wtb_bse_plan* p, int on) {
  if (!p) return WTB_ERR_ARG;
  p->use_graph = on != 0;
  return WTB_OK;
}

int wtb_bse_plan_build(wtb_bse_plan* p) {
  if (!p) {
    set_error("wtb_bse_plan_build: null plan");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  int rc = plan_enqueue(p, true);
  if (rc) return rc;
  WTB_HIP(hipStreamSynchronize(p->ctx->stream));
  p->filled = true;
  return WTB_OK;
}

int wtb_bse_plan_results(wtb_bse_plan* p, double* w, double* act, double* eps, double* egap) {
  if (!p) {
    set_error("wtb_bse_plan_results: null plan");
    return WTB_ERR_ARG;
  }
  hipStream_t s = p->ctx->stream;
  const int N = p->N;
  if (p->qmode && (act || eps || egap)) {
    set_error("wtb_bse_plan_results: a finite-Q plan gives the exciton energies only (act, eps, egap NULL)");
    return WTB_ERR_ARG;
  }
  int st = 0, stq[2] = {0, 0};
  WTB_HIP(hipMemcpyAsync(p->qmode ? stq : &st, p->status, sizeof(int) * (p->qmode ? 2 : 1), hipMemcpyDeviceToHost,
                         s));
  if (w) WTB_HIP(hipMemcpyAsync(w, p->lambda, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  if (act) WTB_HIP(hipMemcpyAsync(act, p->act, sizeof(double) * 8 * N, hipMemcpyDeviceToHost, s));
  if (eps)
    WTB_HIP(hipMemcpyAsync(eps, p->eps, sizeof(double) * 12 * p->prm.numbse, hipMemcpyDeviceToHost, s));
  std::vector<double> gk;
  if (egap) {
    gk.resize(p->nk);
    WTB_HIP(hipMemcpyAsync(gk.data(), p->gapk, sizeof(double) * p->nk, hipMemcpyDeviceToHost, s));
  }
  WTB_HIP(hipStreamSynchronize(s));
  if (egap) {
    // egap = 50.0; keep the minimum (bse_solver-tool-diel.f90:255,305-311)
    double g = 50.0;
    for (double v : gk)
      if (v <= g) g = v;
    *egap = g;
  }
  if (p->timing) {
    // stage boundaries: [0]=start [1]=eig [2]=dip [3]=build [4]=tridiag [5]=dc [6]=end
    float t;
    const int idx[WTB_NSTAGES + 1] = {0, 1, 2, 3, 4, 5, 6};
    for (int i = 0; i < WTB_NSTAGES; ++i) {
      if (hipEventElapsedTime(&t, p->ev[idx[i]], p->ev[idx[i + 1]]) == hipSuccess) p->stage_ms[i] = t;
    }
  }
  if (p->kernel_timing) {
    // the one-stage GEMV's launches, or the two-stage trailing updates' (same event array)
    const int used = p->tw.kev_used > 0 ? p->tw.kev_used : p->bw.kev_used;
    p->kernel_ms = 0.0;
    p->kernel_launches = used;
    p->kernel_flops = p->tw.kev_used > 0 ? 0.0 : p->bw.kev_flops;
    for (int c = 0; c < used; ++c) {
      float t;
      if (hipEventElapsedTime(&t, p->kev[2 * c], p->kev[2 * c + 1]) == hipSuccess) p->kernel_ms += t;
    }
  }
  if (p->qmode) st = qmode_status(stq);
  int rc = status_to_code(st);
  if (rc) set_error(status_text(rc));
  return rc;
}

size_t wtb_bse_plan_buffer_bytes(const wtb_bse_plan* p, int what) {
  if (!p) return 0;
  switch (what) {
    case WTB_BUF_KPT: return sizeof(double) * p->nk * 3;
    case WTB_BUF_EALL: return sizeof(double) * p->nk * p->nw;
    case WTB_BUF_NOCP: return sizeof(int) * p->nk;
    case WTB_BUF_EIGV: return sizeof(double) * p->nk * p->nb;
    case WTB_BUF_VECTOR: return sizeof(zc) * p->nk * p->nb * p->nw;
    case WTB_BUF_DIPOLE: return sizeof(zc) * 3 * p->N;
    case WTB_BUF_HBSE: return (p->mixed ? sizeof(zf) : sizeof(zc)) * p->ldh * p->N;
    case WTB_BUF_WTAB: return sizeof(double) * p->wtab_n;
    case WTB_BUF_TRIDIAG: return sizeof(double) * 2 * p->N;
    case WTB_BUF_EXCWF: return sizeof(zc) * (size_t)p->N * p->wf_count;
    case WTB_BUF_FDEH: return p->fdeh ? sizeof(double) * p->N : 0;
    default: return 0;
  }
}

int wtb_bse_plan_fetch(wtb_bse_plan* p, int what, void* dst, size_t bytes) {
  if (!p || !dst || bytes != wtb_bse_plan_buffer_bytes(p, what) || bytes == 0) {
    set_error("wtb_bse_plan_fetch: bad buffer id or size");
    return WTB_ERR_ARG;
  }
  hipStream_t s = p->ctx->stream;
  const void* src = nullptr;
  switch (what) {
    case WTB_BUF_KPT: src = p->kpt; break;
    case WTB_BUF_EALL: src = p->eall; break;
    case WTB_BUF_NOCP: src = p->nocp; break;
    case WTB_BUF_EIGV: src = p->eigv; break;
    case WTB_BUF_VECTOR: src = p->vec; break;
    case WTB_BUF_DIPOLE: src = p->dip; break;
    case WTB_BUF_HBSE: src = p->mixed ? (const void*)p->Hf : (const void*)p->H; break;
    case WTB_BUF_WTAB: src = p->wtab; break;
    case WTB_BUF_EXCWF: src = p->Xv; break;
    case WTB_BUF_FDEH: src = p->fdeh; break;
    case WTB_BUF_TRIDIAG: {
      WTB_HIP(hipMemcpyAsync(dst, p->tw.d, sizeof(double) * p->N, hipMemcpyDeviceToHost, s));
      WTB_HIP(hipMemcpyAsync((char*)dst + sizeof(double) * p->N, p->tw.e, sizeof(double) * p->N,
                             hipMemcpyDeviceToHost, s));
      WTB_HIP(hipStreamSynchronize(s));
      return WTB_OK;
    }
  }
  WTB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  WTB_HIP(hipStreamSynchronize(s));
  return WTB_OK;
}

int wtb_bse_plan_destroy(wtb_bse_plan* p) {
  if (!p) return WTB_OK;
  hipSetDevice(p->ctx->device);
  hipStreamSynchronize(p->ctx->stream);
  if (p->lz_stream) {
    hipStreamSynchronize(p->lz_stream);
    hipStreamDestroy(p->lz_stream);
  }
  plan_free(p);
  delete p;
  return WTB_OK;
}

int wtb_bsesolver(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, double* w, double* act,
                  double* eps, double* egap) {
  wtb_bse_plan* p = nullptr;
  int rc = wtb_bse_plan_create(c, m, prm, &p);
  if (rc) return rc;
  rc = wtb_bse_plan_execute(p);
  if (!rc) rc = wtb_bse_plan_results(p, w, act, eps, egap);
  wtb_bse_plan_destroy(p);
  return rc;
}

// ---------------------------------------------------------------------------
int wtb_zheev(wtb_ctx* c, int want_vectors, int n, double* a, int lda, double* w) {
  if (!c || !a || !w || n < 1 || lda < n) {
    set_error("wtb_zheev: bad argument");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int vec = want_vectors ? 1 : 0;
  zc *da = nullptr, *dh = nullptr, *dx = nullptr;
  void *tri = nullptr, *dcw = nullptr, *btw = nullptr;
  double *dl = nullptr, *dz = nullptr;
  int rc;
  if ((rc = dmalloc(&da, (size_t)lda * n)) || (rc = dmalloc(&dh, (size_t)n * n)) ||
      (rc = dmalloc((char**)&tri, tridiag_work_bytes(n, 32, 0))) ||
      (rc = dmalloc((char**)&dcw, dc_work_bytes(n, 0, vec))) || (rc = dmalloc(&dl, n)) ||
      (vec && ((rc = dmalloc(&dz, (size_t)n * n)) || (rc = dmalloc(&dx, (size_t)n * n)) ||
               (rc = dmalloc((char**)&btw, backtr_work_bytes(n, n)))))) {
  } else {
    TridiagWork tw;
    DCTree tree;
    tridiag_bind(tw, tri, n, 32, 0);
    tw.store_v = vec;
    rc = dc_prepare(dcw, n, 0, tree, vec);
    if (!rc) {
      hipMemcpyAsync(da, a, sizeof(zc) * lda * n, hipMemcpyHostToDevice, s);
      hipLaunchKernelGGL(k_upper_to_lower, dim3((n + 255) / 256, n), dim3(256), 0, s, n, da, lda, dh, n);
      rc = run_tridiag(s, tw, dh, n, nullptr);
    }
    if (!rc) rc = run_dc(s, dcw, tree, tw.d, tw.e, nullptr, dl, dz);
    if (!rc && vec) rc = run_backtr(s, dh, n, n, tw.tau, dz, 0, n, dx, n, btw);
    if (!rc) {
      hipMemcpyAsync(w, dl, sizeof(double) * n, hipMemcpyDeviceToHost, s);
      if (vec)
        hipMemcpy2DAsync(a, sizeof(zc) * lda, dx, sizeof(zc) * n, sizeof(zc) * n, n, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) {
        set_error("wtb_zheev: stream error");
        rc = WTB_ERR_HIP;
      }
    }
  }
  hipFree(da);
  hipFree(dh);
  hipFree(dx);
  hipFree(tri);
  hipFree(dcw);
  hipFree(btw);
  hipFree(dl);
  hipFree(dz);
  return rc;
}

static int xth_call(wtb_ctx* c, int n, const double* x, const double* h1, const double* h2, double* act) {
  if (!c || !x || !h1 || !act || n < 1) {
    set_error("dielbse: bad argument");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  zc *dx = nullptr, *d1 = nullptr, *d2 = nullptr;
  double* da = nullptr;
  int rc;
  if ((rc = dmalloc(&dx, (size_t)n * n)) || (rc = dmalloc(&d1, n)) || (rc = dmalloc(&d2, n)) ||
      (rc = dmalloc(&da, n))) {
  } else {
    hipMemcpyAsync(dx, x, sizeof(zc) * n * n, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(d1, h1, sizeof(zc) * n, hipMemcpyHostToDevice, s);
    if (h2) hipMemcpyAsync(d2, h2, sizeof(zc) * n, hipMemcpyHostToDevice, s);
    rc = launch_xth(s, n, dx, d1, h2 ? d2 : nullptr, da);
    if (!rc) {
      hipMemcpyAsync(act, da, sizeof(double) * n, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
    }
  }
  hipFree(dx);
  hipFree(d1);
  hipFree(d2);
  hipFree(da);
  return rc;
}

int wtb_dielbsev(wtb_ctx* c, int n, const double* x, const double* h, double* act) {
  return xth_call(c, n, x, h, nullptr, act);
}
int wtb_dielbsep(wtb_ctx* c, int n, const double* x, const double* h1, const double* h2, double* act) {
  if (!h2) {
    set_error("wtb_dielbsep: null h2");
    return WTB_ERR_ARG;
  }
  return xth_call(c, n, x, h1, h2, act);
}

int wtb_bse_spectrum(wtb_ctx* c, int n, const double* e, const double* act6, const double* rlat,
                     const int ngrid[3], double ebse0, double ebsef, int numbse, double sme, int renorm,
                     double* eps) {
  if (!c || !e || !act6 || !rlat || !ngrid || !eps || n < 1 || numbse < 2) {
    set_error("wtb_bse_spectrum: bad argument");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double *de = nullptr, *da = nullptr, *dw = nullptr, *deps = nullptr;
  int rc;
  if ((rc = dmalloc(&de, n)) || (rc = dmalloc(&da, (size_t)6 * n)) || (rc = dmalloc(&dw, 4)) ||
      (rc = dmalloc(&deps, (size_t)12 * numbse))) {
  } else {
    hipMemcpyAsync(de, e, sizeof(double) * n, hipMemcpyHostToDevice, s);
    hipMemcpyAsync(da, act6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s);
    double nkf = (double)(float)(ngrid[0] * ngrid[1] * ngrid[2]);
    rc = launch_spectrum(s, n, de, da, vcell3d_host(rlat), nkf, ebse0, ebsef, numbse, sme, renorm, dw, deps);
    if (!rc) {
      hipMemcpyAsync(eps, deps, sizeof(double) * 12 * numbse, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
    }
  }
  hipFree(de);
  hipFree(da);
  hipFree(dw);
  hipFree(deps);
  return rc;
}

int wtb_bse_spectrum_noalloc(wtb_ctx* c, int n, const double* e, const double* act6, const double* rlat,
                             const int ngrid[3], double ebse0, double ebsef, int numbse, double sme, int renorm,
                             double* eps) {
  if (!c || !e || !act6 || !rlat || !ngrid || !eps || n < 1 || numbse < 2) {
    set_error("wtb_bse_spectrum_noalloc: bad argument");
    return WTB_ERR_ARG;
  }
  // layout in the scratch array: e [n], act6 [6 n], eps [12 numbse], avgsme [4]
  const size_t need = 7 * (size_t)n + 12 * (size_t)numbse + 4;
  if (need > kScratchDoubles) {
    set_error("wtb_bse_spectrum_noalloc: 7 n + 12 numbse + 4 = " + std::to_string(need) +
              " doubles exceeds the scratch array (" + std::to_string(kScratchDoubles) + ")");
    return WTB_ERR_LIMIT;
  }
  double* base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_dev[c->device] : nullptr;
  if (!base) {
    set_error("wtb_bse_spectrum_noalloc: the scratch array of this device was not set up");
    return WTB_ERR_HIP;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double* de = base;
  double* da = de + n;
  double* deps = da + (size_t)6 * n;
  double* dw = deps + (size_t)12 * numbse;
  int rc = WTB_OK;
  if (hipMemcpyAsync(de, e, sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(da, act6, sizeof(double) * 6 * n, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = WTB_ERR_HIP;
  const double nkf = (double)(float)(ngrid[0] * ngrid[1] * ngrid[2]);
  if (!rc) rc = launch_spectrum(s, n, de, da, vcell3d_host(rlat), nkf, ebse0, ebsef, numbse, sme, renorm, dw, deps);
  if (!rc && hipMemcpyAsync(eps, deps, sizeof(double) * 12 * numbse, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = WTB_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = WTB_ERR_HIP;
  if (rc == WTB_ERR_HIP) set_error("wtb_bse_spectrum_noalloc: HIP error");
  return rc;
}

// |a1 x a2| (vcell2D, general_subs.f90:195-206, with prodvec :182-193 and vecsize :286-294)
static double vcell2d_host(const double* rl) {
#pragma clang fp contract(off)
  const double a1[3] = {rl[0], rl[3], rl[6]}, a2[3] = {rl[1], rl[4], rl[7]};
  const double x = (a1[1] * a2[2]) - (a1[2] * a2[1]);
  const double y = (a1[2] * a2[0]) - (a1[0] * a2[2]);
  const double z = (a1[0] * a2[1]) - (a1[1] * a2[0]);
  return std::sqrt(((x * x) + (y * y)) + (z * z));
}

int wtb_optprop(wtb_ctx* c, int n, int ncomp, const double* energy, const double* eps_re, const double* eps_im,
                double* out) {
  if (!c || !energy || !eps_re || !eps_im || !out || n < 1 || ncomp < 1 || ncomp > 64) {
    set_error("wtb_optprop: null argument, n < 1 or ncomp outside 1..64");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  double* d = nullptr;
  const size_t nn = (size_t)n, m = nn * ncomp;
  int rc = dmalloc(&d, nn + 2 * m + 5 * m);
  if (rc) return rc;
  double *den = d, *der = d + nn, *dei = der + m, *dout = dei + m;
  if (hipMemcpyAsync(den, energy, sizeof(double) * nn, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(der, eps_re, sizeof(double) * m, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dei, eps_im, sizeof(double) * m, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = WTB_ERR_HIP;
  if (!rc) rc = launch_optprop(s, n, ncomp, den, der, dei, dout);
  if (!rc && hipMemcpyAsync(out, dout, sizeof(double) * 5 * m, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = WTB_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = WTB_ERR_HIP;
  hipFree(d);
  if (rc == WTB_ERR_HIP) set_error("wtb_optprop: HIP error");
  return rc;
}

int wtb_exciton_lifetime(wtb_ctx* c, int sysdim, const int ngrid[3], const double* rlat, int n, int ncomp, int nout,
                         const double* fosc, double* lft) {
  if (sysdim < 1 || sysdim > 3) {
    set_error("Wrong System dimension");
    return WTB_ERR_ARG;
  }
  if (!c || !ngrid || !rlat || !fosc || !lft || n < 1 || ncomp < 3 || nout < 1 || nout > ncomp) {
    set_error("wtb_exciton_lifetime: null argument, n < 1, ncomp < 3 or nout outside 1..ncomp");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const size_t nn = (size_t)n, ncol = (size_t)ncomp + 1;
  double* d = nullptr;
  int rc = dmalloc(&d, nn * ncol + nn * (nout + 1));
  if (rc) return rc;
  const double nkd = (double)(ngrid[0] * ngrid[1] * ngrid[2]);
  if (hipMemcpyAsync(d, fosc, sizeof(double) * nn * ncol, hipMemcpyHostToDevice, s) != hipSuccess) rc = WTB_ERR_HIP;
  if (!rc) rc = launch_lifetime(s, n, (int)ncol, nout, sysdim, nkd, vcell2d_host(rlat), rlat[0], d, d + nn * ncol);
  if (!rc && hipMemcpyAsync(lft, d + nn * ncol, sizeof(double) * nn * (nout + 1), hipMemcpyDeviceToHost, s) !=
                 hipSuccess)
    rc = WTB_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = WTB_ERR_HIP;
  hipFree(d);
  if (rc == WTB_ERR_HIP) set_error("wtb_exciton_lifetime: HIP error");
  return rc;
}

// ---------------------------------------------------------------------------
// learned preconditioner object
// ---------------------------------------------------------------------------
struct wtb_mlp {
  wtb_ctx* ctx = nullptr;
  int hidden = 0, np = 0;
  float* P = nullptr;     // parameters
  float* m = nullptr;     // Adam moments
  float* v = nullptr;
  int step = 0;
};

int wtb_mlp_create(wtb_ctx* c, int hidden, wtb_mlp** out) {
  if (!c || !out || mlp_nparams(hidden) <= 0) {
    set_error("wtb_mlp_create: hidden width must be 128");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  auto* q = new wtb_mlp();
  q->ctx = c;
  q->hidden = hidden;
  q->np = mlp_nparams(hidden);
  int rc;
  if ((rc = dmalloc(&q->P, q->np)) || (rc = dmalloc(&q->m, q->np)) || (rc = dmalloc(&q->v, q->np))) {
    wtb_mlp_destroy(q);
    return rc;
  }
  hipMemsetAsync(q->P, 0, sizeof(float) * q->np, c->stream);
  hipMemsetAsync(q->m, 0, sizeof(float) * q->np, c->stream);
  hipMemsetAsync(q->v, 0, sizeof(float) * q->np, c->stream);
  WTB_HIP(hipStreamSynchronize(c->stream));
  *out = q;
  return WTB_OK;
}

int wtb_mlp_nparams(const wtb_mlp* q) { return q ? q->np : 0; }

int wtb_mlp_set_params(wtb_mlp* q, const float* p, int n) {
  if (!q || !p || n != q->np) {
    set_error("wtb_mlp_set_params: bad argument");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipMemcpyAsync(q->P, p, sizeof(float) * n, hipMemcpyHostToDevice, q->ctx->stream));
  hipMemsetAsync(q->m, 0, sizeof(float) * q->np, q->ctx->stream);
  hipMemsetAsync(q->v, 0, sizeof(float) * q->np, q->ctx->stream);
  q->step = 0;
  WTB_HIP(hipStreamSynchronize(q->ctx->stream));
  return WTB_OK;
}

int wtb_mlp_get_params(const wtb_mlp* q, float* p, int n) {
  if (!q || !p || n != q->np) {
    set_error("wtb_mlp_get_params: bad argument");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipMemcpyAsync(p, q->P, sizeof(float) * n, hipMemcpyDeviceToHost, q->ctx->stream));
  WTB_HIP(hipStreamSynchronize(q->ctx->stream));
  return WTB_OK;
}

int wtb_mlp_forward(wtb_mlp* q, int S, const float* X, float* y) {
  if (!q || !X || !y || S < 1) {
    set_error("wtb_mlp_forward: bad argument");
    return WTB_ERR_ARG;
  }
  hipStream_t s = q->ctx->stream;
  float *dX = nullptr, *dy = nullptr;
  int rc;
  if (!(rc = dmalloc(&dX, (size_t)5 * S)) && !(rc = dmalloc(&dy, (size_t)S))) {
    hipMemcpyAsync(dX, X, sizeof(float) * 5 * S, hipMemcpyHostToDevice, s);
    rc = mlp_forward(s, S, dX, q->P, dy);
    if (!rc) {
      hipMemcpyAsync(y, dy, sizeof(float) * S, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
    }
  }
  hipFree(dX);
  hipFree(dy);
  return rc;
}

static int mlp_upload(hipStream_t s, int S, const float* X, const float* t, float** dX, float** dt, int** dperm) {
  int rc;
  if ((rc = dmalloc(dX, (size_t)5 * S)) || (rc = dmalloc(dt, (size_t)S)) || (rc = dmalloc(dperm, (size_t)S)))
    return rc;
  WTB_HIP(hipMemcpyAsync(*dX, X, sizeof(float) * 5 * S, hipMemcpyHostToDevice, s));
  WTB_HIP(hipMemcpyAsync(*dt, t, sizeof(float) * S, hipMemcpyHostToDevice, s));
  return WTB_OK;
}

int wtb_mlp_loss_grad(wtb_mlp* q, int S, const float* X, const float* t, float* loss, float* grad) {
  if (!q || !X || !t || !loss || !grad || S < 1) {
    set_error("wtb_mlp_loss_grad: bad argument");
    return WTB_ERR_ARG;
  }
  hipStream_t s = q->ctx->stream;
  float *dX = nullptr, *dt = nullptr, *work = nullptr, *dg = nullptr, *dl = nullptr;
  int* dperm = nullptr;
  int rc = mlp_upload(s, S, X, t, &dX, &dt, &dperm);
  if (!rc && !(rc = dmalloc(&work, mlp_grad_work_floats(S))) && !(rc = dmalloc(&dg, (size_t)q->np)) &&
      !(rc = dmalloc(&dl, 4))) {
    std::vector<int> perm(S);
    for (int i = 0; i < S; ++i) perm[i] = i;
    hipMemcpy(dperm, perm.data(), sizeof(int) * S, hipMemcpyHostToDevice);
    rc = mlp_loss_grad(s, S, dX, dt, dperm, 0, S, q->P, work, dg, dl);
    if (!rc) {
      hipMemcpyAsync(grad, dg, sizeof(float) * q->np, hipMemcpyDeviceToHost, s);
      hipMemcpyAsync(loss, dl, sizeof(float), hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
    }
  }
  for (void* b : {(void*)dX, (void*)dt, (void*)dperm, (void*)work, (void*)dg, (void*)dl}) hipFree(b);
  return rc;
}

int wtb_mlp_train(wtb_mlp* q, int S, const float* X, const float* t, int epochs, int batch, float lr,
                  unsigned seed, float* loss_hist) {
  if (!q || !X || !t || S < 1 || epochs < 1 || batch < 1) {
    set_error("wtb_mlp_train: bad argument");
    return WTB_ERR_ARG;
  }
  hipStream_t s = q->ctx->stream;
  float *dX = nullptr, *dt = nullptr, *work = nullptr, *dg = nullptr, *dl = nullptr;
  int* dperm = nullptr;
  const int B = std::min(batch, S);
  int rc = mlp_upload(s, S, X, t, &dX, &dt, &dperm);
  const int nb = S / B;   // whole batches per epoch
  if (!rc && !(rc = dmalloc(&work, mlp_grad_work_floats(B))) && !(rc = dmalloc(&dg, (size_t)q->np)) &&
      !(rc = dmalloc(&dl, (size_t)nb + 1))) {
    std::vector<int> perm(S);
    std::vector<float> ls(nb);
    unsigned long long st = 0x9E3779B97F4A7C15ull ^ seed;
    for (int ep = 0; ep < epochs && !rc; ++ep) {
      for (int i = 0; i < S; ++i) perm[i] = i;
      for (int i = S - 1; i > 0; --i) {   // Fisher-Yates with a 64-bit LCG
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        const int j = (int)((st >> 33) % (unsigned long long)(i + 1));
        std::swap(perm[i], perm[j]);
      }
      if (hipMemcpy(dperm, perm.data(), sizeof(int) * S, hipMemcpyHostToDevice) != hipSuccess) {
        rc = WTB_ERR_HIP;
        break;
      }
      for (int b = 0; b < nb && !rc; ++b) {
        rc = mlp_loss_grad(s, S, dX, dt, dperm, b * B, B, q->P, work, dg, dl + b);
        if (!rc) rc = mlp_adam(s, q->P, dg, q->m, q->v, lr, ++q->step);
      }
      if (!rc) {
        hipMemcpyAsync(ls.data(), dl, sizeof(float) * nb, hipMemcpyDeviceToHost, s);
        if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
        double a = 0.0;
        for (float v : ls) a += v;
        if (loss_hist) loss_hist[ep] = (float)(a / nb);
      }
    }
  }
  for (void* b : {(void*)dX, (void*)dt, (void*)dperm, (void*)work, (void*)dg, (void*)dl}) hipFree(b);
  return rc;
}

int wtb_mlp_destroy(wtb_mlp* q) {
  if (!q) return WTB_OK;
  hipFree(q->P);
  hipFree(q->m);
  hipFree(q->v);
  delete q;
  return WTB_OK;
}

// per-state diagonal and row statistics of the plan's BSE matrix (after build):
// D[N], stats[3][N] = sum |H_ij|, sqrt(sum |H_ij|^2), max |H_ij| over j != i
int wtb_bse_plan_row_stats(wtb_bse_plan* p, double* D, float* stats) {
  if (!p || !D || !stats) {
    set_error("wtb_bse_plan_row_stats: bad argument");
    return WTB_ERR_ARG;
  }
  hipStream_t s = p->ctx->stream;
  const int N = p->N;
  double* dD = nullptr;
  float* dst = nullptr;
  int rc;
  if (!p->mixed) {
    set_error("wtb_bse_plan_row_stats: needs a plan in the precision form (WTB_MIXED)");
    return WTB_ERR_ARG;
  }
  if (!(rc = dmalloc(&dD, N)) && !(rc = dmalloc(&dst, (size_t)3 * N))) {
    rc = plan_enqueue(p, true);
    if (!rc) rc = sym_fill_f32(s, N, p->Hf, (int)p->ldh);
    if (!rc) rc = row_stats_f32(s, N, p->Hf, (int)p->ldh, dD, dst);
    if (!rc) {
      hipMemcpyAsync(D, dD, sizeof(double) * N, hipMemcpyDeviceToHost, s);
      hipMemcpyAsync(stats, dst, sizeof(float) * 3 * N, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
    }
  }
  hipFree(dD);
  hipFree(dst);
  return rc;
}

// The iterative form's parameters, checked before the plan is looked at (so that the
// check runs without a device): every value that would size a buffer or a loop.
static int check_davidson_params(const wtb_davidson_params* dp) {
  if (!dp) {
    set_error("wtb_bse_plan_lowest: null parameters");
    return WTB_ERR_ARG;
  }
  const char* bad = nullptr;
  if (dp->nev < 1) bad = "nev must be >= 1";
  else if (dp->guard < 0) bad = "guard must be >= 0";
  else if (dp->max_iter < 1) bad = "max_iter must be >= 1";
  else if (dp->max_sub < 0) bad = "max_sub must be >= 0";
  else if (dp->coarse < 0) bad = "coarse must be >= 0";
  else if (!(dp->tol > 0.0) || !std::isfinite(dp->tol)) bad = "tol must be a positive finite number";
  else if (dp->lanczos_steps < 0) bad = "lanczos_steps must be >= 0";
  if (bad) {
    set_error(std::string("wtb_bse_plan_lowest: ") + bad);
    return WTB_ERR_ARG;
  }
  return WTB_OK;
}

// grow a device buffer to at least `bytes` (contents are not kept)
static int grow_buffer(void** ptr, size_t* cap, size_t bytes) {
  if (*ptr && *cap >= bytes) return WTB_OK;
  hipFree(*ptr);
  *ptr = nullptr;
  *cap = 0;
  int rc = dmalloc((char**)ptr, bytes);
  if (!rc) *cap = bytes;
  return rc;
}

int wtb_bse_plan_lowest(wtb_bse_plan* p, const wtb_davidson_params* dp, const wtb_mlp* mlp, double* W,
                        double* resid, double* eps, double* X, wtb_davidson_info* info) {
  int rc0 = check_davidson_params(dp);
  if (rc0) return rc0;
  if (!p || !W || !resid || dp->nev > p->N) {
    set_error("wtb_bse_plan_lowest: null plan or output, or nev > N");
    return WTB_ERR_ARG;
  }
  if (!p->mixed) {
    set_error("wtb_bse_plan_lowest: needs a plan in the precision form (WTB_MIXED)");
    return WTB_ERR_ARG;
  }
  if (p->qmode && eps && dp->lanczos_steps > 0) {
    // the spectrum is a Q = 0 quantity; a finite-Q plan has no dipoles
    set_error("wtb_bse_plan_lowest: eps(omega) needs a Q = 0 plan (pass eps = NULL or lanczos_steps = 0 with a "
              "finite-Q plan)");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  const int N = p->N, nev = dp->nev;
  const wtb_bse_params& P = p->prm;
  wtb_davidson_info inf{};
  hipEvent_t ev[4] = {};
  for (auto& e : ev) WTB_HIP(hipEventCreate(&e));
  int rc = WTB_OK;
  const int steps = dp->lanczos_steps;
  // the plan keeps these between calls: no allocation once their sizes are reached
  if (!p->dav && !(p->dav = davidson_work_new())) rc = WTB_ERR_NOMEM;
  if (!rc && !p->lw_D) {
    if ((rc = dmalloc(&p->lw_D, N)) || (rc = dmalloc(&p->lw_stats, (size_t)3 * N)) || (rc = dmalloc(&p->lw_sm, 4))) {
      hipFree(p->lw_D);
      hipFree(p->lw_stats);
      hipFree(p->lw_sm);
      p->lw_D = nullptr;
      p->lw_stats = nullptr;
      p->lw_sm = nullptr;
    }
  }
  if (!rc && (size_t)nev > p->lw_nev) {
    hipFree(p->lw_W);
    hipFree(p->lw_R);
    p->lw_W = p->lw_R = nullptr;
    p->lw_nev = 0;
    if (!(rc = dmalloc(&p->lw_W, nev)) && !(rc = dmalloc(&p->lw_R, nev))) p->lw_nev = nev;
  }
  if (!rc && X && (size_t)nev > p->lw_xcols) {
    hipFree(p->lw_X);
    p->lw_X = nullptr;
    p->lw_xcols = 0;
    if (!(rc = dmalloc(&p->lw_X, (size_t)N * nev))) p->lw_xcols = nev;
  }
  if (!rc && eps && steps > 0) rc = grow_buffer(&p->lw_hay, &p->lw_hay_bytes, haydock_work_bytes(N, steps));
  double* dD = p->lw_D;
  float* dst = p->lw_stats;
  double *dW = p->lw_W, *dR = p->lw_R, *dsm = p->lw_sm;
  zc* dX = X ? p->lw_X : nullptr;
  void* hw = (eps && steps > 0) ? p->lw_hay : nullptr;
  if (!rc) {
    hipEventRecord(ev[0], s);
    rc = plan_enqueue(p, true);
    if (!rc) rc = sym_fill_f32(s, N, p->Hf, (int)p->ldh);
    if (!rc) rc = row_stats_f32(s, N, p->Hf, (int)p->ldh, dD, dst);
    hipEventRecord(ev[1], s);
    // the Lanczos chains need only the matrix and the dipoles: they run on a stream of
    // their own beside the Davidson (whose Rayleigh-Ritz steps leave most of the chip
    // idle); eps is formed from their coefficients once E_1 is known
    const bool lz = eps && steps > 0;
    hipStream_t s2 = nullptr;
    hipEvent_t ev_lz = nullptr;
    if (!rc && lz) {
      if (!p->lz_stream && hipStreamCreateWithFlags(&p->lz_stream, hipStreamNonBlocking) != hipSuccess) {
        p->lz_stream = nullptr;
        rc = WTB_ERR_HIP;
      }
      s2 = p->lz_stream;
      if (!rc && hipEventCreateWithFlags(&ev_lz, hipEventDisableTiming) != hipSuccess) rc = WTB_ERR_HIP;
      if (!rc && hipStreamWaitEvent(s2, ev[1], 0) != hipSuccess) rc = WTB_ERR_HIP;
      if (!rc)
        rc = run_haydock_chains(s2, N, p->Hf, (int)p->ldh, p->dip, P.dtfull, steps, hw, P.bset ? p->fdeh : nullptr);
      if (!rc && hipEventRecord(ev_lz, s2) != hipSuccess) rc = WTB_ERR_HIP;
    }
    DavidsonCfg cfg;
    cfg.nev = nev;
    cfg.guard = dp->guard;
    cfg.max_sub = dp->max_sub;
    cfg.max_iter = dp->max_iter;
    cfg.tol = dp->tol;
    cfg.coarse = dp->coarse;
    MlpDev md;
    if (mlp) {
      md.w = mlp->P;
      md.hidden = mlp->hidden;
    }
    DavidsonInfo di;
    if (!rc) rc = run_davidson(s, N, p->Hf, (int)p->ldh, dD, dst, cfg, mlp ? &md : nullptr, dW, dR, dX, di, *p->dav);
    hipEventRecord(ev[2], s);
    inf.iters = di.iters;
    inf.matvecs = di.matvecs;
    inf.converged = di.converged;
    inf.iters_fp64 = di.iters_fp64;
    inf.matvecs_fp64 = di.matvecs_fp64;
    inf.max_resid = di.max_resid;
    if (!rc && lz && di.converged) {
      const double gama = (double)180.90f;
      const double caux = gama / (p->nkf32 * p->vc3d);
      if (hipStreamWaitEvent(s, ev_lz, 0) != hipSuccess) rc = WTB_ERR_HIP;
      if (!rc)
        rc = run_haydock_eps(s, N, P.dtfull, steps, dW, caux, P.ebse0, P.ebsef, P.numbse, P.cshift, P.renorm, hw,
                             p->eps, dsm, nullptr, P.bset);
    }
    // every path: the chains are done before their workspace is freed
    if (s2 && hipStreamSynchronize(s2) != hipSuccess && !rc) rc = WTB_ERR_HIP;
    if (ev_lz) hipEventDestroy(ev_lz);
    hipEventRecord(ev[3], s);
    if (!rc && di.converged) {
      hipMemcpyAsync(W, dW, sizeof(double) * nev, hipMemcpyDeviceToHost, s);
      hipMemcpyAsync(resid, dR, sizeof(double) * nev, hipMemcpyDeviceToHost, s);
      if (X) hipMemcpyAsync(X, dX, sizeof(zc) * N * nev, hipMemcpyDeviceToHost, s);
      if (eps && steps > 0)
        hipMemcpyAsync(eps, p->eps, sizeof(double) * 12 * P.numbse, hipMemcpyDeviceToHost, s);
      if (hipStreamSynchronize(s) != hipSuccess) rc = WTB_ERR_HIP;
      float t;
      if (hipEventElapsedTime(&t, ev[0], ev[1]) == hipSuccess) inf.ms_build = t;
      if (hipEventElapsedTime(&t, ev[1], ev[2]) == hipSuccess) inf.ms_solve = t;
      if (hipEventElapsedTime(&t, ev[2], ev[3]) == hipSuccess) inf.ms_eps = t;
    } else if (!rc && dp->fallback) {
      // not converged: the full FP64 solve decides (plain path on this plan's model)
      inf.used_fallback = 1;
      if (!p->H) rc = dmalloc(&p->H, p->ldh * N);
      if (!rc) {
        p->mixed = false;
        rc = plan_enqueue(p);
        p->mixed = true;
      }
      if (!rc) {
        std::vector<double> wall(N);
        rc = wtb_bse_plan_results(p, wall.data(), nullptr, eps, nullptr);
        if (!rc) {
          for (int i = 0; i < nev; ++i) {
            W[i] = wall[i];
            resid[i] = 0.0;
          }
          if (X) memset(X, 0, sizeof(zc) * N * nev);
        }
      }
    } else if (!rc) {
      set_error("wtb_bse_plan_lowest: Davidson did not converge (fallback off)");
      rc = WTB_ERR_EIG;
    }
  }
  for (auto& e : ev)
    if (e) hipEventDestroy(e);
  if (info) *info = inf;
  return rc;
}

// ---------------------------------------------------------------------------
// GEMV tile shape and its autotuner
// ---------------------------------------------------------------------------
static void plan_drop_graph(wtb_bse_plan* p) {
  if (p->gexec) hipGraphExecDestroy(p->gexec);
  if (p->graph) hipGraphDestroy(p->graph);
  p->gexec = nullptr;
  p->graph = nullptr;
}

int wtb_bse_plan_set_panel(wtb_bse_plan* p, int nb) {
  if (!p || nb < 8 || nb > 32) {
    set_error("wtb_bse_plan_set_panel: panel width must lie in 8..32");
    return WTB_ERR_ARG;
  }
  if (p->tw.panel != nb) plan_drop_graph(p);
  p->tw.panel = nb;
  return WTB_OK;
}

int wtb_bse_plan_tune(wtb_bse_plan* p, int* best, double* ms) {
  if (!p) {
    set_error("wtb_bse_plan_tune: null plan");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  const int cand[4] = {16, 20, 24, 32};
  const bool timing0 = p->timing;
  const int nb0 = p->tw.panel;
  double tbest = 1e300;
  int sbest = nb0;
  p->timing = true;
  int rc = WTB_OK;
  for (int k = 0; k < 4 && !rc; ++k) {
    p->tw.panel = cand[k];
    double t = 1e300;
    for (int rep = 0; rep < 2 && !rc; ++rep) {   // the first run warms up
      rc = plan_enqueue(p);
      if (!rc && hipStreamSynchronize(p->stream) != hipSuccess) rc = WTB_ERR_HIP;
      float f;
      if (!rc && hipEventElapsedTime(&f, p->ev[3], p->ev[4]) == hipSuccess) t = f;
    }
    if (ms) ms[k] = t;
    if (t < tbest) {
      tbest = t;
      sbest = cand[k];
    }
  }
  p->timing = timing0;
  p->tw.panel = sbest;
  if (sbest != nb0) plan_drop_graph(p);
  if (best) *best = sbest;
  return rc;
}

// ---------------------------------------------------------------------------
// batching form: many models in one captured graph.  Stages (a)-(c) of every matrix run
// on parallel streams (lanes); the one-stage reductions then run as groups in lockstep
// (one launch per kernel and column step for all matrices of a group, the groups being
// the matrices of one element type and panel width); stages (d2)-(e) run on the lanes
// again.  Matrices with another reduction (two-stage) run their whole pipeline on
// their lane before the groups.
// ---------------------------------------------------------------------------
struct wtb_bse_batch {
  wtb_ctx* ctx = nullptr;
  std::vector<wtb_bse_plan*> plans;
  std::vector<hipStream_t> lanes;
  std::vector<int> lane_of;             // plan -> lane
  std::vector<hipEvent_t> ev_done;      // one per lane
  hipEvent_t ev_start = nullptr;
  hipEvent_t ev_mid = nullptr;          // the groups' reductions are done
  hipEvent_t ev_jac = nullptr;          // the batched Jacobi eigensolves are done
  std::vector<TriGroup> groups;
  std::vector<char> grouped;            // plan -> reduced in a group
  std::vector<JacobiItem> jitems;       // the grouped plans' Jacobi eigensolves (stage b)
  JacobiItem* d_jitems = nullptr;
  bool stale = true;                    // groups and graph to be (re)built
  hipGraphExec_t gexec = nullptr;
};

static void batch_drop(wtb_bse_batch* b) {
  if (b->gexec) hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  for (auto& g : b->groups) tri_group_free(g);
  b->groups.clear();
  if (b->d_jitems) hipFree(b->d_jitems);
  b->d_jitems = nullptr;
  b->jitems.clear();
  b->stale = true;
}

// the groups: one per (element type, panel width) among the plans with a one-stage
// reduction
static int batch_group(wtb_bse_batch* b) {
  batch_drop(b);
  const int n = (int)b->plans.size();
  b->grouped.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    wtb_bse_plan* p = b->plans[i];
    if (!plan_one_stage(p)) continue;
    const int nb = (p->tw.panel >= 8 && p->tw.panel <= p->tw.nb) ? p->tw.panel : p->tw.nb;
    TriGroup* g = nullptr;
    for (auto& x : b->groups)
      if (x.f32 == p->mixed && x.nb == nb) g = &x;
    if (!g) {
      b->groups.emplace_back();
      g = &b->groups.back();
      g->f32 = p->mixed;
      g->nb = nb;
    }
    g->items.push_back(TriGroupItem{&p->tw, p->mixed ? (void*)p->Hf : (void*)p->H, (int)p->ldh, p->g});
    b->grouped[i] = 1;
  }
  for (auto& g : b->groups) {
    const int rc = tri_group_prepare(g);
    if (rc) return rc;
  }
  for (int i = 0; i < n; ++i)
    if (b->grouped[i]) b->jitems.push_back(plan_jacobi_item(b->plans[i]));
  if (!b->jitems.empty()) {
    WTB_HIP(hipMalloc(&b->d_jitems, sizeof(JacobiItem) * b->jitems.size()));
    WTB_HIP(hipMemcpy(b->d_jitems, b->jitems.data(), sizeof(JacobiItem) * b->jitems.size(),
                      hipMemcpyHostToDevice));
  }
  b->stale = false;
  return WTB_OK;
}

// the set's pipeline on the context stream, forking to the lanes and joining back
static int batch_enqueue(wtb_bse_batch* b) {
  hipStream_t s = b->ctx->stream;
  const int n = (int)b->plans.size();
  auto fork = [&](hipEvent_t e) -> int {
    WTB_HIP(hipEventRecord(e, s));
    for (hipStream_t l : b->lanes) WTB_HIP(hipStreamWaitEvent(l, e, 0));
    return WTB_OK;
  };
  auto join = [&]() -> int {
    for (size_t l = 0; l < b->lanes.size(); ++l) {
      WTB_HIP(hipEventRecord(b->ev_done[l], b->lanes[l]));
      WTB_HIP(hipStreamWaitEvent(s, b->ev_done[l], 0));
    }
    return WTB_OK;
  };
  int rc;
  // stage (a) of the grouped matrices on the lanes, every model's Jacobi eigensolves in one
  // batched launch, then stages (b')-(c) on the lanes again; the other plans run whole
  if ((rc = fork(b->ev_start))) return rc;
  for (int i = 0; i < n; ++i) {
    wtb_bse_plan* p = b->plans[i];
    rc = b->grouped[i] ? plan_enqueue_pre_a(p) : plan_enqueue(p);
    if (rc) return rc;
  }
  if ((rc = join())) return rc;
  if (!b->jitems.empty() && (rc = launch_jacobi_batch(s, b->d_jitems, b->jitems))) return rc;
  if ((rc = fork(b->ev_jac))) return rc;
  for (int i = 0; i < n; ++i)
    if (b->grouped[i] && (rc = plan_enqueue_pre(b->plans[i], false, true))) return rc;
  if ((rc = join())) return rc;
  for (const auto& g : b->groups)
    if ((rc = run_tridiag_group(s, g))) return rc;
  if ((rc = fork(b->ev_mid))) return rc;
  for (int i = 0; i < n; ++i)
    if (b->grouped[i] && (rc = plan_enqueue_post(b->plans[i]))) return rc;
  return join();
}

int wtb_bse_batch_create(wtb_ctx* c, int n, const wtb_tb_model* models, const wtb_bse_params* prms, int nlanes,
                         wtb_bse_batch** out) {
  if (!c || n < 1 || !models || !prms || !out || nlanes < 1 || nlanes > 8) {
    set_error("wtb_bse_batch_create: bad argument (1 <= nlanes <= 8)");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(c->device));
  auto* b = new wtb_bse_batch();
  b->ctx = c;
  int rc = WTB_OK;
  for (int i = 0; i < n && !rc; ++i) {
    wtb_bse_plan* p = nullptr;
    rc = wtb_bse_plan_create(c, &models[i], &prms[i], &p);
    if (!rc) b->plans.push_back(p);
  }
  for (int l = 0; l < nlanes && !rc; ++l) {
    hipStream_t st;
    hipEvent_t e;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      rc = WTB_ERR_HIP;
      break;
    }
    b->lanes.push_back(st);
    b->ev_done.push_back(e);
  }
  if (!rc && (hipEventCreateWithFlags(&b->ev_start, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&b->ev_mid, hipEventDisableTiming) != hipSuccess ||
              hipEventCreateWithFlags(&b->ev_jac, hipEventDisableTiming) != hipSuccess))
    rc = WTB_ERR_HIP;
  if (rc) {
    wtb_bse_batch_destroy(b);
    return rc;
  }
  // longest processing time first: the largest N^3 goes to the least loaded lane
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return b->plans[x]->N > b->plans[y]->N; });
  std::vector<double> load(nlanes, 0.0);
  b->lane_of.assign(n, 0);
  for (int i : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->lane_of[i] = l;
    const double N = b->plans[i]->N;
    load[l] += N * N * N;
    b->plans[i]->stream = b->lanes[l];
    if (nlanes > 1) b->plans[i]->bw.panel_resident = 0;   // lanes share the device
  }
  *out = b;
  return WTB_OK;
}

int wtb_bse_batch_size(const wtb_bse_batch* b) { return b ? (int)b->plans.size() : 0; }

int wtb_bse_batch_dim(const wtb_bse_batch* b, int i) {
  return (b && i >= 0 && i < (int)b->plans.size()) ? b->plans[i]->N : 0;
}

int wtb_bse_batch_tune(wtb_bse_batch* b, int* tiles) {
  if (!b) return WTB_ERR_ARG;
  WTB_HIP(hipStreamSynchronize(b->ctx->stream));
  batch_drop(b);                          // the panel widths, hence the groups, change
  for (size_t i = 0; i < b->plans.size(); ++i) {
    int best = 32;
    int rc = wtb_bse_plan_tune(b->plans[i], &best, nullptr);
    if (rc) return rc;
    if (tiles) tiles[i] = best;
  }
  return WTB_OK;
}

// the first call (and the first after a tune) forms the groups and captures the set's
// graph; every call replays it on the context stream
int wtb_bse_batch_execute(wtb_bse_batch* b) {
  if (!b) return WTB_ERR_ARG;
  WTB_HIP(hipSetDevice(b->ctx->device));
  int rc;
  if (b->stale && (rc = batch_group(b))) return rc;
  if (!b->gexec) {
    hipStream_t s = b->ctx->stream;
    WTB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("wtb_bse_batch_execute: hipStreamEndCapture: ") + hipGetErrorString(e));
      return WTB_ERR_HIP;
    }
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
      b->gexec = nullptr;
      set_error(std::string("wtb_bse_batch_execute: hipGraphInstantiate: ") + hipGetErrorString(e));
      return WTB_ERR_HIP;
    }
  }
  WTB_HIP(hipGraphLaunch(b->gexec, b->ctx->stream));
  for (auto* p : b->plans) p->filled = true;
  return WTB_OK;
}

int wtb_bse_batch_results(wtb_bse_batch* b, int i, double* w, double* act, double* eps, double* egap) {
  if (!b || i < 0 || i >= (int)b->plans.size()) {
    set_error("wtb_bse_batch_results: bad index");
    return WTB_ERR_ARG;
  }
  return wtb_bse_plan_results(b->plans[i], w, act, eps, egap);
}

int wtb_bse_batch_destroy(wtb_bse_batch* b) {
  if (!b) return WTB_OK;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(b->ctx->stream);
  batch_drop(b);
  if (b->ev_mid) hipEventDestroy(b->ev_mid);
  if (b->ev_jac) hipEventDestroy(b->ev_jac);
  for (auto* p : b->plans) wtb_bse_plan_destroy(p);
  for (auto l : b->lanes) {
    hipStreamSynchronize(l);
    hipStreamDestroy(l);
  }
  for (auto e : b->ev_done) hipEventDestroy(e);
  if (b->ev_start) hipEventDestroy(b->ev_start);
  delete b;
  return WTB_OK;
}

int wtb_bse_plan_set_vectors(wtb_bse_plan* p, int first, int count) {
  if (!p || first < 0 || count < 0 || first + count > p->N) {
    set_error("wtb_bse_plan_set_vectors: need 0 <= first, first + count <= N");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  WTB_HIP(hipStreamSynchronize(p->stream));
  const int N = p->N;
  int rc = WTB_OK;
  // Everything new is allocated and prepared into temporaries first; the plan is changed
  // only when all of it succeeded, so a failed call leaves a plan that still executes.
  const bool need_z = count > 0 && !p->Zv;
  const bool need_x = count != p->wf_count && count > 0;
  void* ndc = nullptr;
  double* nzv = nullptr;
  zc* nxv = nullptr;
  void* nbt = nullptr;
  DCTree ntree{};
  auto drop = [&]() {
    hipFree(ndc);
    hipFree(nzv);
    hipFree(nxv);
    hipFree(nbt);
  };
  if (need_z) {
    // the divide and conquer needs its eigenvector workspace, the reduction keeps its reflectors
    if ((rc = dmalloc((char**)&ndc, dc_work_bytes(N, 2 * p->nrhs, 1))) || (rc = dmalloc(&nzv, (size_t)N * N)) ||
        (rc = dc_prepare(ndc, N, 2 * p->nrhs, ntree, 1))) {
      drop();
      return rc;
    }
  }
  if (need_x &&
      ((rc = dmalloc(&nxv, (size_t)N * count)) || (rc = dmalloc((char**)&nbt, backtr_work_bytes(N, count))))) {
    drop();
    return rc;
  }
  if (need_z) {
    hipFree(p->dcwork);
    p->dcwork = ndc;
    p->Zv = nzv;
    p->tree = ntree;
    p->tw.store_v = 1;
  }
  if (count != p->wf_count) {
    hipFree(p->Xv);
    hipFree(p->btw);
    p->Xv = nxv;
    p->btw = nbt;
  }
  p->wf_first = first;
  p->wf_count = count;
  plan_drop_graph(p);
  return WTB_OK;
}

int wtb_bse_plan_set_reduction(wtb_bse_plan* p, int two_stage, int band) {
  if (!p || two_stage < -1 || two_stage > 1 || (band != 16 && band != 32)) {
    set_error("wtb_bse_plan_set_reduction: two_stage -1, 0 or 1, band 16 or 32");
    return WTB_ERR_ARG;
  }
  if (two_stage == 1 && p->N < 2) {
    set_error("wtb_bse_plan_set_reduction: the two-stage reduction needs N >= 2");
    return WTB_ERR_ARG;
  }
  if (band != p->band_b && p->N >= 2) {
    WTB_HIP(hipSetDevice(p->ctx->device));
    WTB_HIP(hipStreamSynchronize(p->stream));
    void* nw = nullptr;
    BandWork nb;
    int rc;
    if ((rc = dmalloc((char**)&nw, band_work_bytes(p->N, band, p->nrhs)))) return rc;
    if ((rc = band_bind(nb, nw, p->N, band, p->nrhs))) {
      hipFree(nw);
      return rc;
    }
    hipFree(p->bandwork);
    p->bandwork = nw;
    nb.panel_resident = p->bw.panel_resident;
    p->bw = nb;
    p->band_b = band;
  }
  if (two_stage != p->two_stage || band != p->band_b) plan_drop_graph(p);
  p->two_stage = two_stage;
  plan_drop_graph(p);
  return WTB_OK;
}

int wtb_bse_plan_set_refine(wtb_bse_plan* p, int nref) {
  if (!p) {
    set_error("wtb_bse_plan_set_refine: null plan");
    return WTB_ERR_ARG;
  }
  if (nref < 0 || nref > p->N) {
    set_error("wtb_bse_plan_set_refine: need 0 <= nref <= N");
    return WTB_ERR_ARG;
  }
  if (nref > 0 && (!p->mixed || p->qmode || p->sp_only)) {
    set_error("wtb_bse_plan_set_refine: the refinement belongs to the precision form (WTB_MIXED) of a Q = 0 plan");
    return WTB_ERR_ARG;
  }
  if (nref == p->nref) return WTB_OK;
  WTB_HIP(hipSetDevice(p->ctx->device));
  WTB_HIP(hipStreamSynchronize(p->stream));
  const int N = p->N;
  int rc = WTB_OK;
  // new buffers into temporaries first; the plan changes only when all of them exist
  const bool need_z = nref > 0 && !p->Zv;
  const bool need_h = nref > 0 && !p->H;
  void* ndc = nullptr;
  double* nzv = nullptr;
  zc* nh = nullptr;
  zc *nx = nullptr, *nhx = nullptr, *nhw = nullptr;
  void* nbt = nullptr;
  double *nres = nullptr, *nl0 = nullptr;
  DCTree ntree{};
  auto drop = [&]() {
    for (void* b : {(void*)ndc, (void*)nzv, (void*)nh, (void*)nx, (void*)nhx, (void*)nhw, nbt, (void*)nres,
                    (void*)nl0})
      hipFree(b);
  };
  if (nref > 0) {
    if ((need_z && ((rc = dmalloc((char**)&ndc, dc_work_bytes(N, 2 * p->nrhs, 1))) ||
                    (rc = dmalloc(&nzv, (size_t)N * N)) || (rc = dc_prepare(ndc, N, 2 * p->nrhs, ntree, 1)))) ||
        (need_h && (rc = dmalloc(&nh, p->ldh * N))) || (rc = dmalloc(&nx, (size_t)N * nref)) ||
        (rc = dmalloc(&nhx, (size_t)N * nref)) || (rc = dmalloc((char**)&nbt, backtr_work_bytes(N, nref))) ||
        (rc = dmalloc((char**)&nhw, hmm_work_bytes(N, nref))) || (rc = dmalloc(&nres, nref)) ||
        (rc = dmalloc(&nl0, nref))) {
      drop();
      return rc;
    }
  }
  if (need_z) {
    hipFree(p->dcwork);
    p->dcwork = ndc;
    p->Zv = nzv;
    p->tree = ntree;
    p->tw.store_v = 1;
  }
  if (need_h) p->H = nh;
  for (void* b : {(void*)p->Xr, (void*)p->HXr, p->btw_r, (void*)p->hw_r, (void*)p->rq_res, (void*)p->rq_lam0})
    hipFree(b);
  p->Xr = nx;
  p->HXr = nhx;
  p->btw_r = nbt;
  p->hw_r = nhw;
  p->rq_res = nres;
  p->rq_lam0 = nl0;
  p->nref = nref;
  plan_drop_graph(p);
  return WTB_OK;
}

int wtb_bse_plan_refine_results(wtb_bse_plan* p, int n, double* resid, double* w_unrefined) {
  if (!p || n < 0 || n > p->nref || (n > 0 && !resid)) {
    set_error("wtb_bse_plan_refine_results: need 0 <= n <= nref of wtb_bse_plan_set_refine");
    return WTB_ERR_ARG;
  }
  if (n == 0) return WTB_OK;
  if (!p->filled) {
    set_error("wtb_bse_plan_refine_results: execute the plan first");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  WTB_HIP(hipStreamSynchronize(p->stream));
  WTB_HIP(hipMemcpyAsync(resid, p->rq_res, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  if (w_unrefined) WTB_HIP(hipMemcpyAsync(w_unrefined, p->rq_lam0, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  WTB_HIP(hipStreamSynchronize(s));
  return WTB_OK;
}

int wtb_bsesolver_wf(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, double* w, double* act,
                     double* eps, double* egap, int wf_first, int wf_count, double* wf, double* kpt, int* nocp) {
  wtb_bse_plan* p = nullptr;
  int rc = wtb_bse_plan_create(c, m, prm, &p);
  if (rc) return rc;
  rc = wtb_bse_plan_set_vectors(p, wf_first, wf_count);
  if (!rc) rc = wtb_bse_plan_execute(p);
  if (!rc) rc = wtb_bse_plan_results(p, w, act, eps, egap);
  if (!rc && wf && wf_count > 0)
    rc = wtb_bse_plan_fetch(p, WTB_BUF_EXCWF, wf, sizeof(zc) * (size_t)p->N * wf_count);
  if (!rc && kpt) rc = wtb_bse_plan_fetch(p, WTB_BUF_KPT, kpt, sizeof(double) * 3 * p->nk);
  if (!rc && nocp) rc = wtb_bse_plan_fetch(p, WTB_BUF_NOCP, nocp, sizeof(int) * p->nk);
  wtb_bse_plan_destroy(p);
  return rc;
}


// ---------------------------------------------------------------------------
// finite-Q exciton bands (bsebnds)
// ---------------------------------------------------------------------------
static int plan_make_qmode(wtb_bse_plan* p) {
  const int nk = p->nk, nb = p->nb, nw = p->nw, N = p->N;
  int rc;
  if ((rc = dmalloc(&p->dq, 4)) || (rc = dmalloc(&p->kq, (size_t)3 * nk)) || (rc = dmalloc(&p->eigvq, (size_t)nk * nb)) ||
      (rc = dmalloc(&p->vecq, (size_t)nk * nb * nw)) || (rc = dmalloc(&p->nocpq, nk)) || (rc = dmalloc(&p->xa, N)) ||
      (rc = dmalloc(&p->dvq, 1)))
    return rc;
  const double q0[4] = {0.0, 0.0, 0.0, 0.0};
  WTB_HIP(hipMemcpy(p->dq, q0, sizeof(q0), hipMemcpyHostToDevice));
  // eigenvalues only (ZHEEV('N')): no tracked vectors
  const int panel = p->tw.panel;
  tridiag_bind(p->tw, p->triwork, N, 32, 0);
  p->tw.panel = panel;
  if ((rc = dc_prepare(p->dcwork, N, 0, p->tree))) return rc;
  p->nrhs = 0;
  p->qmode = 1;
  plan_drop_graph(p);
  return WTB_OK;
}

int wtb_bse_qplan_create(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, wtb_bse_plan** out) {
  int rc = wtb_bse_plan_create(c, m, prm, out);
  if (rc) return rc;
  if ((rc = plan_make_qmode(*out))) {
    wtb_bse_plan_destroy(*out);
    *out = nullptr;
  }
  return rc;
}

int wtb_bse_plan_set_q(wtb_bse_plan* p, const double* q) {
  if (!p || !q || !p->qmode) {
    set_error("wtb_bse_plan_set_q: needs a finite-Q plan (wtb_bse_qplan_create) and q[4]");
    return WTB_ERR_ARG;
  }
  for (int i = 0; i < 4; ++i)
    if (!std::isfinite(q[i])) {
      set_error("wtb_bse_plan_set_q: q is not finite");
      return WTB_ERR_NONFINITE;
    }
  WTB_HIP(hipSetDevice(p->ctx->device));
  WTB_HIP(hipMemcpyAsync(p->dq, q, sizeof(double) * 4, hipMemcpyHostToDevice, p->stream));
  WTB_HIP(hipStreamSynchronize(p->stream));
  return WTB_OK;
}

int wtb_kpathbse(const double* rlat, int nks, const double* ks, int npts, double* qauxv, double* labels) {
  if (!rlat || !ks || !qauxv || nks < 2 || npts < 2) {
    set_error("wtb_kpathbse: need rlat, ks, qauxv, nks >= 2 and npts >= 2");
    return WTB_ERR_ARG;
  }
  // recvec with the REAL pi (mhkpack_subs.f90:409-432)
  double a[3][3], b[3][3];
  for (int i = 0; i < 3; ++i)
    for (int c = 0; c < 3; ++c) a[i][c] = rlat[i + 3 * c];
  auto cross = [](const double* u, const double* v, double* w) {
    w[0] = (u[1] * v[2]) - (u[2] * v[1]);
    w[1] = (u[2] * v[0]) - (u[0] * v[2]);
    w[2] = (u[0] * v[1]) - (u[1] * v[0]);
  };
  double v23[3], v31[3], v12[3];
  cross(a[1], a[2], v23);
  cross(a[2], a[0], v31);
  cross(a[0], a[1], v12);
  const double vol = std::fabs((a[0][0] * v23[0]) + (a[0][1] * v23[1]) + (a[0][2] * v23[2]));
  const double f = (2.0 * (double)3.14159265358979f) / vol;
  for (int c = 0; c < 3; ++c) {
    b[0][c] = f * v23[c];
    b[1][c] = f * v31[c];
    b[2][c] = f * v12[c];
  }
  // kpathbse (mhkpack_subs.f90:132-208)
  const int nseg = nks / 2, nq = nseg * npts;
  std::vector<double> ksx((size_t)nks * 3), kdis(nseg);
  for (int i = 0; i < nks; ++i)
    for (int c = 0; c < 3; ++c)
      ksx[3 * i + c] = ks[i] * b[0][c] + ks[i + nks] * b[1][c] + ks[i + 2 * nks] * b[2][c];
  double tot = 0.0;
  for (int j = 0; j < nseg; ++j) {
    const double* e = &ksx[3 * (2 * j + 1)];
    const double* s0 = &ksx[3 * (2 * j)];
    const double d0 = e[0] - s0[0], d1 = e[1] - s0[1], d2 = e[2] - s0[2];
    kdis[j] = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    tot = tot + kdis[j];
  }
  for (int j = 0; j < nseg; ++j) kdis[j] = kdis[j] / tot;
  double acc = 0.0;
  if (labels) labels[0] = 0.0;
  for (int j = 0; j < nseg; ++j) {
    const double* s0 = &ksx[3 * (2 * j)];
    const double* e = &ksx[3 * (2 * j + 1)];
    for (int i = 0; i < npts; ++i) {
      const int r = j * npts + i;
      qauxv[r] = acc + (kdis[j] * (double)i) / (double)(npts - 1);
      for (int c = 0; c < 3; ++c) qauxv[r + (size_t)nq * (1 + c)] = s0[c] + ((e[c] - s0[c]) * (double)i) / (double)(npts - 1);
    }
    acc = acc + kdis[j];
    if (labels) labels[j + 1] = acc;
  }
  return WTB_OK;
}

struct wtb_bse_bands {
  wtb_ctx* ctx = nullptr;
  std::vector<wtb_bse_plan*> plans;     // one finite-Q plan per lane
  std::vector<hipStream_t> streams;
  int N = 0;
  double* hq = nullptr;                 // pinned staging: q rows, energies, status words
  double* hw = nullptr;
  int* hst = nullptr;
  int cap = 0;                          // q-points the staging holds
  int wf_count = 0;                     // exciton vectors per q (wtb_bse_bands_set_vectors)
  zc* hx = nullptr;                     // pinned staging of the vectors
  size_t hx_cap = 0;
};

int wtb_bse_bands_destroy(wtb_bse_bands* b) {
  if (!b) return WTB_OK;
  hipSetDevice(b->ctx->device);
  for (wtb_bse_plan* p : b->plans) {
    hipStreamSynchronize(p->stream);
    p->stream = b->ctx->stream;
    wtb_bse_plan_destroy(p);
  }
  for (hipStream_t st : b->streams) hipStreamDestroy(st);
  if (b->hq) hipHostFree(b->hq);
  if (b->hw) hipHostFree(b->hw);
  if (b->hst) hipHostFree(b->hst);
  if (b->hx) hipHostFree(b->hx);
  delete b;
  return WTB_OK;
}

int wtb_bse_bands_set_vectors(wtb_bse_bands* b, int first, int count) {
  if (!b) return WTB_ERR_ARG;
  for (wtb_bse_plan* p : b->plans) {
    hipStream_t st = p->stream;
    int rc = wtb_bse_plan_set_vectors(p, first, count);
    p->stream = st;
    if (rc) return rc;
  }
  b->wf_count = count;
  return WTB_OK;
}

int wtb_bse_bands_create(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, int lanes,
                         wtb_bse_bands** out) {
  if (!c || !m || !prm || !out || lanes < 1 || lanes > 8) {
    set_error("wtb_bse_bands_create: null argument or lanes outside 1..8");
    return WTB_ERR_ARG;
  }
  *out = nullptr;
  WTB_HIP(hipSetDevice(c->device));
  auto* b = new wtb_bse_bands();
  b->ctx = c;
  int rc = WTB_OK;
  for (int l = 0; l < lanes && !rc; ++l) {
    wtb_bse_plan* p = nullptr;
    if ((rc = wtb_bse_qplan_create(c, m, prm, &p))) break;
    b->plans.push_back(p);
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      set_error("wtb_bse_bands_create: hipStreamCreate failed");
      rc = WTB_ERR_HIP;
      break;
    }
    b->streams.push_back(st);
    p->stream = st;
    if (lanes > 1) p->bw.panel_resident = 0;   // lanes share the device
  }
  if (rc) {
    wtb_bse_bands_destroy(b);
    return rc;
  }
  b->N = b->plans[0]->N;
  *out = b;
  return WTB_OK;
}

int wtb_bse_bands_set_graph(wtb_bse_bands* b, int on) {
  if (!b) return WTB_ERR_ARG;
  for (wtb_bse_plan* p : b->plans) p->use_graph = on != 0;
  return WTB_OK;
}

int wtb_bse_bands_run(wtb_bse_bands* b, int nq, const double* qauxv, double* exk) {
  return wtb_bse_bands_run_wf(b, nq, qauxv, exk, nullptr);
}

int wtb_bse_bands_run_wf(wtb_bse_bands* b, int nq, const double* qauxv, double* exk, double* wf) {
  if (!b || nq < 1 || !qauxv || !exk || (wf && b->wf_count < 1)) {
    set_error("wtb_bse_bands_run: null argument, nq < 1, or vectors asked without wtb_bse_bands_set_vectors");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(b->ctx->device));
  const int N = b->N, nl = (int)b->plans.size();
  if (nq > b->cap) {
    if (b->hq) hipHostFree(b->hq);
    if (b->hw) hipHostFree(b->hw);
    if (b->hst) hipHostFree(b->hst);
    b->hq = b->hw = nullptr;
    b->hst = nullptr;
    b->cap = 0;
    if (hipHostMalloc((void**)&b->hq, sizeof(double) * 4 * nq) != hipSuccess ||
        hipHostMalloc((void**)&b->hw, sizeof(double) * (size_t)N * nq) != hipSuccess ||
        hipHostMalloc((void**)&b->hst, sizeof(int) * 2 * nq) != hipSuccess) {
      set_error("wtb_bse_bands_run: pinned host allocation failed");
      return WTB_ERR_NOMEM;
    }
    b->cap = nq;
  }
  const size_t per_q = (size_t)b->N * b->wf_count;      // complex entries of one q's vectors
  if (wf && per_q * nq > b->hx_cap) {
    if (b->hx) hipHostFree(b->hx);
    b->hx = nullptr;
    b->hx_cap = 0;
    if (hipHostMalloc((void**)&b->hx, sizeof(zc) * per_q * nq) != hipSuccess) {
      set_error("wtb_bse_bands_run: pinned host allocation failed");
      return WTB_ERR_NOMEM;
    }
    b->hx_cap = per_q * nq;
  }
  for (int iq = 0; iq < nq; ++iq)
    for (int k = 0; k < 4; ++k) b->hq[4 * iq + k] = qauxv[iq + (size_t)nq * k];
  // the q-points round-robin over the lanes; each lane's plan runs its points in order
  int rc = WTB_OK;
  for (int iq = 0; iq < nq && !rc; ++iq) {
    wtb_bse_plan* p = b->plans[iq % nl];
    hipStream_t st = p->stream;
    if (hipMemcpyAsync(p->dq, b->hq + 4 * iq, sizeof(double) * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error("wtb_bse_bands_run: q upload failed");
      rc = WTB_ERR_HIP;
      break;
    }
    if ((rc = wtb_bse_plan_execute(p))) break;
    if (hipMemcpyAsync(b->hw + (size_t)N * iq, p->lambda, sizeof(double) * N, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(b->hst + 2 * iq, p->status, sizeof(int) * 2, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (wf && hipMemcpyAsync(b->hx + per_q * iq, p->Xv, sizeof(zc) * per_q, hipMemcpyDeviceToHost, st) !=
                   hipSuccess)) {
      set_error("wtb_bse_bands_run: result copy failed");
      rc = WTB_ERR_HIP;
    }
  }
  for (hipStream_t st : b->streams)
    if (hipStreamSynchronize(st) != hipSuccess && !rc) {
      set_error("wtb_bse_bands_run: HIP error while running the q-points");
      rc = WTB_ERR_HIP;
    }
  for (int iq = 0; iq < nq && !rc; ++iq) {
    const int code = status_to_code(qmode_status(b->hst + 2 * iq));
    if (code) {
      set_error(status_text(code));
      rc = code;
    }
  }
  if (!rc) std::memcpy(exk, b->hw, sizeof(double) * (size_t)N * nq);
  if (!rc && wf) std::memcpy(wf, b->hx, sizeof(zc) * per_q * nq);
  return rc;
}

int wtb_bsebnds_wf(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, int nq, const double* qauxv,
                   int lanes, double* exk, int wf_first, int wf_count, double* wf, double* kpt, int* nocp) {
  if (!c || !m || !prm || nq < 1 || !qauxv || !exk || wf_count < 0 || (wf_count > 0 && !wf)) {
    set_error("wtb_bsebnds_wf: null argument, nq < 1 or bad vector range");
    return WTB_ERR_ARG;
  }
  wtb_bse_bands* b = nullptr;
  int rc = wtb_bse_bands_create(c, m, prm, std::max(1, std::min({lanes, nq, 8})), &b);
  if (!rc && wf_count > 0) rc = wtb_bse_bands_set_vectors(b, wf_first, wf_count);
  if (!rc) rc = wtb_bse_bands_run_wf(b, nq, qauxv, exk, wf_count > 0 ? wf : nullptr);
  // the k grid and its nocp (the same at every q), as excwfi2 prints them
  if (!rc && kpt) rc = wtb_bse_plan_fetch(b->plans[0], WTB_BUF_KPT, kpt, sizeof(double) * 3 * b->plans[0]->nk);
  if (!rc && nocp) rc = wtb_bse_plan_fetch(b->plans[0], WTB_BUF_NOCP, nocp, sizeof(int) * b->plans[0]->nk);
  wtb_bse_bands_destroy(b);
  return rc;
}

int wtb_bsebnds(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, int nq, const double* qauxv, int lanes,
                double* exk) {
  if (!c || !m || !prm || nq < 1 || !qauxv || !exk) {
    set_error("wtb_bsebnds: null argument or nq < 1");
    return WTB_ERR_ARG;
  }
  wtb_bse_bands* b = nullptr;
  int rc = wtb_bse_bands_create(c, m, prm, std::max(1, std::min({lanes, nq, 8})), &b);
  if (!rc) rc = wtb_bse_bands_run(b, nq, qauxv, exk);
  wtb_bse_bands_destroy(b);
  return rc;
}


// ---------------------------------------------------------------------------
// single-particle side output (sp_optics.dat, tm_coef.dat)
// ---------------------------------------------------------------------------
int wtb_bse_plan_sp_table(wtb_bse_plan* p, float* vecres) {
  if (!p || !vecres || p->qmode) {
    set_error("wtb_bse_plan_sp_table: needs a Q = 0 plan and vecres(N, 15)");
    return WTB_ERR_ARG;
  }
  if (!p->filled) {
    set_error("wtb_bse_plan_sp_table: build or execute the plan first (the table comes from its bands)");
    return WTB_ERR_ARG;
  }
  WTB_HIP(hipSetDevice(p->ctx->device));
  hipStream_t s = p->ctx->stream;
  WTB_HIP(hipStreamSynchronize(p->stream));
  const int N = p->N;
  float* d = nullptr;
  int rc = dmalloc(&d, (size_t)30 * N);
  if (rc) return rc;
  rc = launch_sp_table(s, N, p->nk, p->prm.nc, p->prm.nv, p->kpt, p->nocp, p->eigv, p->dip,
                       p->prm.bset ? p->fdeh : nullptr, d, d + (size_t)15 * N);
  if (!rc && hipMemcpyAsync(vecres, d + (size_t)15 * N, sizeof(float) * 15 * N, hipMemcpyDeviceToHost, s) != hipSuccess)
    rc = WTB_ERR_HIP;
  if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = WTB_ERR_HIP;
  hipFree(d);
  if (rc == WTB_ERR_HIP) set_error("wtb_bse_plan_sp_table: HIP error");
  return rc;
}

int wtb_sp_table(wtb_ctx* c, const wtb_tb_model* m, const wtb_bse_params* prm, float* vecres) {
  wtb_bse_plan* p = nullptr;
  int rc = plan_create_impl(c, m, prm, &p, true);
  if (rc) return rc;
  rc = wtb_bse_plan_build(p);
  if (!rc) {
    // the band-window checks of the stages the table comes from
    int st = 0;
    if (hipMemcpy(&st, p->status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = WTB_ERR_HIP;
    else if ((rc = status_to_code(st))) set_error(status_text(rc));
  }
  if (!rc) rc = wtb_bse_plan_sp_table(p, vecres);
  wtb_bse_plan_destroy(p);
  return rc;
}

}  // extern "C"
