// Internal declarations shared by the kernel translation units and the C-ABI.
#pragma once
#include "wtb_common.h"
#include "../../include/wtb_bse.h"
#include <vector>

namespace wtb {

// Device copy of the tight-binding model.
struct DevModel {
  int nw = 0, nvec = 0;
  double* rlat = nullptr;   // [9], Fortran rlat(3,3)
  int* rvec = nullptr;      // [nvec*3], Fortran rvec(nvec,3)
  zc* hstack = nullptr;     // [nvec][nw*nw] H(R)/ndeg, column e = n*nw + m
};

// Coulomb parameters in the form the table kernel needs.
struct CoulombArgs {
  int type;
  int n1, n2, n3;
  double ediel[3];
  double lc, ez, w, r0, ktol;
};

// stage (a)+(b)
// d_q (optional, device): a q-path row (coordinate, qx, qy, qz): the grid k + q
int launch_kgrid(hipStream_t s, const double* d_rlat, const int ng[3], const double sh[3],
                 double* d_kpt, const double* d_q = nullptr);
int launch_hr_stack(hipStream_t s, const double* hop, const double* ihop, const int* ff, int nvec,
                    int nw, zc* B);
int launch_hk(hipStream_t s, const DevModel& dm, const double* d_kpt, int nk, zc* d_phase,
              zc* d_hk4);
size_t jacobi_lds_bytes(int nw);
// one model's Jacobi eigensolves (the arguments of launch_jacobi) for a batched launch
struct JacobiItem {
  const zc* hk;
  int nk, nw;
  double ef32, scs;
  int nc, nv;
  double* eall;
  zc* uout;
  int* nocp;
  double* eigv;
  zc* vec;
  int* status;
};
// every model's k-points in launches of up to 64 models (d_items: the items on the device)
int launch_jacobi_batch(hipStream_t s, const JacobiItem* d_items, const std::vector<JacobiItem>& items);
int launch_jacobi(hipStream_t s, const zc* d_hk, int nk, int nw, double ef32, double scs, int nc,
                  int nv, double* d_eall, zc* d_uout, int* d_nocp, double* d_eigv, zc* d_vec,
                  int* d_status);
int launch_dipoles(hipStream_t s, const zc* d_hk4, int nk, int nw, int nc, int nv, double sme32,
                   const double* d_eigv, const zc* d_vec, zc* d_dip, double* d_gapk);

// stage (c)
int launch_wtable(hipStream_t s, const double* d_rlat, const CoulombArgs& ca, double* d_wtab);
// d_fdeh (optional, finite temperature): kernel terms of column J of the lower triangle
// (row J of the reference's upper one) times fdeh[J].  d_xa, d_vq (optional, finite Q):
// the exchange term -V(q) a_I conj(a_J)
int launch_bse_build(hipStream_t s, int nk, int nc, int nv, int nw, const int ng[3],
                     const double* d_wtab, const double* d_eigv, const zc* d_vec, zc* d_H, int ldh,
                     zf* d_Hf, const double* d_fdeh = nullptr, const zc* d_xa = nullptr,
                     const double* d_vq = nullptr);
// finite-Q exciton bands: conduction rows [nv, nb) of eigv/vec from the k+q window (eigvq/vecq);
// a[I] = <c_I|v_I> (c at k+q, v at k); vq = V(q) (0 when q = 0)
int launch_mix_q(hipStream_t s, int nk, int nc, int nv, int nw, const double* d_eigvq, const zc* d_vecq,
                 double* d_eigv, zc* d_vec);
int launch_exchange_vec(hipStream_t s, int nk, int nc, int nv, int nw, const zc* d_vec, zc* d_a);
int launch_vq(hipStream_t s, const double* d_rlat, const CoulombArgs& ca, const double* d_q, double* d_vq);
// fdeh[I] = f(E_v) - f(E_c) of basis state I, f(E) = 1/(exp(E beta) + 1)
int launch_fermi_factors(hipStream_t s, int nk, int nc, int nv, const double* d_eigv, double beta,
                         double* d_fdeh);
// y[q][i] = x[q][i] f[i], q < nvec (complex vectors of length N)
int launch_scale_rows(hipStream_t s, int nvec, int N, const zc* x, const double* f, zc* y);

// stage (d): Householder reduction of the lower triangle of H (ldh) to a real
// tridiagonal (d, e), applying Q^T to nrhs complex vectors g (g[N*nrhs]).
struct TridiagWork {
  int N = 0, nb = 32, nrhs = 0;
  zc* W = nullptr;        // N x nb
  zc* Vp = nullptr;       // N x nb (reflectors of the current panel)
  zc* u = nullptr;        // N
  zc* tau = nullptr;      // N
  double* d = nullptr;    // N
  double* e = nullptr;    // N
  double* red = nullptr;  // real partial sums
  zc* zred = nullptr;     // complex partial sums
  zc* ypart = nullptr;    // hemv partials [nblk][nblk][64]
  zc* xn = nullptr;       // reduced column below the diagonal
  zc* xp = nullptr;       // panel-updated column before the deferred term
  zc* puv = nullptr;      // per-workgroup u^H v partials
  zc* gs = nullptr;       // saved row of the tracked vectors
  zc* scal = nullptr;     // one reflector's scalars (batched form)
  // optional per-launch timing of the Hermitian GEMV kernel: when kev is set,
  // launch c is bracketed by events kev[2c], kev[2c+1] (c < kev_cap)
  hipEvent_t* kev = nullptr;
  int kev_cap = 0;
  int kev_used = 0;
  // keep the reflectors (v_j below row j+1 in column j of H) for the back-transform
  int store_v = 0;
  // panel width of the blocked reduction (columns per trailing update), <= nb of the
  // binding; the autotuner picks it per matrix
  int panel = 32;
};
size_t tridiag_work_bytes(int N, int nb, int nrhs);
int tridiag_bind(TridiagWork& tw, void* base, int N, int nb, int nrhs);
int run_tridiag(hipStream_t s, TridiagWork& tw, zc* d_H, int ldh, zc* d_g);
// the same reduction on a COMPLEX*8 matrix (precision form)
int run_tridiag_f32(hipStream_t s, TridiagWork& tw, zf* d_H, int ldh, zc* d_g);
// A group of such reductions of one element type and panel width, run in lockstep: one
// launch per kernel and column step serves every matrix of the group, and each matrix
// gets bitwise the results of run_tridiag / run_tridiag_f32 on it alone.
struct TriGroupItem {
  TridiagWork* tw;
  void* H;     // zc* or zf* (f32)
  int ldh;
  zc* g;
};
struct TriGroup {
  bool f32 = false;
  int nb = 32;                       // the panel width of every member
  std::vector<TriGroupItem> items;
  void* d_args = nullptr;            // the members' kernel arguments on the device
};
int tri_group_prepare(TriGroup& G);  // allocates d_args (the members must be bound)
void tri_group_free(TriGroup& G);
int run_tridiag_group(hipStream_t s, const TriGroup& G);

// stage (d), two-stage form: dense -> band of width b (16 or 32) -> real tridiagonal
// (d, e), applying Q^T to nrhs (<= 6) complex vectors g.  The matrix is in FULL Hermitian
// storage (both triangles) and is overwritten.  A wait of the bulge-chasing kernel that
// passes its bound sets WTB_ERR_EIG in status[0].
struct BandWork {
  int N = 0, b = 32, nrhs = 0, LB = 0, nwg_max = 0;
  zc *Pw = nullptr, *V = nullptr, *Y = nullptr, *X = nullptr;
  zc *part = nullptr, *rowc = nullptr, *tau = nullptr;
  zc *G = nullptr, *t1 = nullptr, *T = nullptr, *t2 = nullptr, *M = nullptr;
  zc* gw = nullptr;
  size_t gw_elems = 0;
  zc* gp = nullptr;             // partials of the panel's tall-skinny products
  zc* hb = nullptr;             // K-slice partials of the trailing product X = A22 Y
  zc* BD = nullptr;
  zc* gc = nullptr;
  int* prog = nullptr;
  int* err = nullptr;           // [0] the chase's error flag, [8] the panel's, [32] next sweep
  int* bar = nullptr;           // per panel and column: workgroups past the column
  int panel_resident = 0;       // one-launch panels up to this many workgroups (0: per column)
  // optional per-launch timing of the trailing-update kernel (k_her2k_full): launch c is
  // bracketed by kev[2c], kev[2c+1] (c < kev_cap); kev_flops[c] its flop count
  hipEvent_t* kev = nullptr;
  int kev_cap = 0;
  int kev_used = 0;
  double kev_flops = 0.0;       // summed over the timed launches
};
size_t band_work_bytes(int N, int b, int nrhs);
int band_bind(BandWork& w, void* base, int N, int b, int nrhs);
int run_band_tridiag(hipStream_t s, BandWork& w, zc* H, int ld, zc* g, double* d, double* e, int* status);
int run_band_tridiag_f32(hipStream_t s, BandWork& w, zf* H, int ld, zc* g, double* d, double* e, int* status);

// stage (d): tridiagonal divide and conquer (eigenvalues + Z^T x for tracked vectors)
struct DCTree {
  int N = 0, ntrack = 0, vectors = 0;
  std::vector<int> level_off, level_cnt, level_maxn;
};
size_t dc_work_bytes(int N, int ntrack, int vectors = 0);
int dc_prepare(void* work, int N, int ntrack, DCTree& tree, int vectors = 0);
// d_track: ntrack x N, in: the tridiagonal-basis vectors, out: Z^T (.)
// d_Z (optional, N x N column-major): the eigenvectors, column j for d_lambda[j];
// needs a tree prepared with vectors = 1.  zcols >= 0: only columns j < zcols are
// formed (the others of d_Z are left undefined)
int run_dc(hipStream_t s, void* work, const DCTree& tree, const double* d_d, const double* d_e,
           double* d_track, double* d_lambda, double* d_Z = nullptr, int zcols = -1);

// stage (d3): X(:, 0:m) = Q Z(:, c0:c0+m) from the stored reflectors (store_v) and tau;
// the last component of each column is made real and >= 0
size_t backtr_work_bytes(int N, int m);
int run_backtr(hipStream_t s, const zc* H, int ld, int N, const zc* tau, const double* Z, int c0, int m,
               zc* X, int ldx, void* work);
int run_backtr_f32(hipStream_t s, const zf* H, int ld, int N, const zc* tau, const double* Z, int c0,
                   int m, zc* X, int ldx, void* work);

// complex GEMM C = alpha op(A) op(B) + beta C, op 'N' or 'C' (FP64 MFMA).  With a
// workspace of work_elems complex entries, a long K over few output tiles is split into
// slices whose partial sums are added in slice order (deterministic).
int zgemm(hipStream_t s, char opa, char opb, int M, int N, int K, zc alpha, const zc* A, int lda,
          const zc* B, int ldb, zc beta, zc* C, int ldc, zc* work = nullptr, size_t work_elems = 0);
// Y = H X for a full Hermitian H: COMPLEX*8 operand on FP32 MFMA, or FP64; work (optional,
// hmm_work_bytes) lets a thin product split K over more workgroups
size_t hmm_work_bytes(int N, int nc);
int hmm_f32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
            zc* work = nullptr);
int hmm_f64(hipStream_t s, int N, int nc, const zc* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
            zc* work = nullptr);
// thin product (nc <= 16) for the Lanczos chains; work of thin_work_bytes
size_t thin_work_bytes(int N, int nc);
int hmm_thin_f32(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                 zc* work);
// COMPLEX*8 operand, FP64 MFMA arithmetic (the final iterations of the iterative solver)
int hmm_f32_acc64(hipStream_t s, int N, int nc, const zf* H, int ldh, const zc* X, int ldx, zc* Y, int ldy,
                  zc* work = nullptr);

// stage (e)
int launch_activities(hipStream_t s, int N, const double* d_G /* 6 x N */, int dtfull, int cpol,
                      double* d_act /* 8 x N */);
// finite temperature: d_G rows 0-5 from h, rows 6-11 from f h
int launch_activities_temp(hipStream_t s, int N, const double* d_G /* 12 x N */, int dtfull, int cpol,
                           double* d_act /* 8 x N */);
int launch_quantize(hipStream_t s, int n, const double* src, double* dst);
// the single-particle table vecres(N, 15) (REAL), sorted by transition energy; work: 15 N floats
int launch_sp_table(hipStream_t s, int N, int nk, int nc, int nv, const double* kpt, const int* nocp,
                    const double* eigv, const zc* dip, const double* fdeh, float* work, float* out);
int launch_spectrum(hipStream_t s, int N, const double* d_e, const double* d_act6, double vc,
                    double nkf32, double ebse0, double ebsef, int numbse, double sme, int renorm,
                    double* d_work, double* d_eps);

int launch_imagrenorm(hipStream_t s, const double* d_e, double ebse0, double ebsef, int numbse, double* d_eps);
// optical constants n, k, R, alpha, loss of nc components at n photon energies:
// er, ei [nc][n], out [5][nc][n]
int launch_optprop(hipStream_t s, int n, int nc, const double* en, const double* er, const double* ei, double* out);
// exciton radiative lifetimes: fosc [ncol][n] (energy first), lft [1 + nout][n]
int launch_lifetime(hipStream_t s, int n, int ncol, int nout, int sysdim, double nkd, double vc2d, double a11,
                    const double* fosc, double* lft);

// iterative form: Davidson for the lowest excitons, Haydock chains for eps(omega)
int sym_fill_f32(hipStream_t s, int N, zf* H, int ld);
int sym_fill_f64(hipStream_t s, int N, zc* H, int ld);
int row_stats_f32(hipStream_t s, int N, const zf* H, int ld, double* D, float* st);
// theta = x^H H x / x^H x for the nc columns of X (FP64 products HX): lam[c] -> lam0[c],
// theta -> lam[c], |H x - theta x| / |x| -> res[c]
int launch_rq_refine(hipStream_t s, int N, int nc, const zc* X, const zc* HX, int ld, double* lam, double* lam0,
                     double* res);
int launch_residual(hipStream_t s, int N, int nc, const zc* U, const zc* HU, int ld, const double* theta,
                    zc* R, double* rn);
int launch_precond_diag(hipStream_t s, int N, int nq, const int* idx, const zc* R, int ld,
                        const double* theta, const double* D, double dmin, zc* T, int ldt);
int launch_chol_inv(hipStream_t s, int b, zc* G, int ld, zc* Linv, int* keep, double tol);
size_t haydock_work_bytes(int N, int steps);
// eps[6][2][numbse] from Lanczos chains of the dipoles (dip: 3 x N); d_E1: lowest exciton
// energy (device); d_smeavg[0] <- 10 (E_max - E_1) / N, [1] <- E_max; h_coef (host, optional):
// a then b coefficients, nchains x (steps + 1) each
int run_haydock(hipStream_t s, int N, const zf* H, int ldh, const zc* dip, int dtfull, int steps,
                const double* d_E1, double caux, double ebse0, double ebsef, int numbse, double sme,
                int renorm, void* work, double* d_eps, double* d_smeavg, double* h_coef);
// the same in two parts: the chains (no dependence on E_1, so they can run beside the
// eigensolver on a stream of their own), then eps from their coefficients; one workspace
// fdeh (optional, finite temperature): chains of h_a +- f h_b for every component; run
// run_haydock_eps with temp = 1 on them
int run_haydock_chains(hipStream_t s, int N, const zf* H, int ldh, const zc* dip, int dtfull, int steps,
                       void* work, const double* fdeh = nullptr);
int run_haydock_eps(hipStream_t s, int N, int dtfull, int steps, const double* d_E1, double caux, double ebse0,
                    double ebsef, int numbse, double sme, int renorm, void* work, double* d_eps, double* d_smeavg,
                    double* h_coef, int temp = 0);

// learned preconditioner: MLP (5 -> 128 -> 128 -> 1, tanh, FP32) on per-state features
struct MlpDev {
  const float* w = nullptr;   // packed parameters, see mlp_layout
  int hidden = 0;
};
int mlp_nparams(int hidden);
int launch_precond_mlp(hipStream_t s, int N, int nq, const int* idx, const zc* R, int ld, const double* theta,
                       const double* D, const float* stats, double Dmin, const MlpDev& mlp, zc* T, int ldt);

int mlp_forward(hipStream_t s, int S, const float* X, const float* P, float* y);
size_t mlp_grad_work_floats(int B);
// loss and gradient of the mean squared error over samples perm[b0 .. b0+B) of X[5][S], t[S]
int mlp_loss_grad(hipStream_t s, int S, const float* X, const float* t, const int* perm, int b0, int B,
                  const float* P, float* work, float* grad, float* loss);
int mlp_adam(hipStream_t s, float* P, const float* grad, float* m, float* v, float lr, int step);

struct DavidsonCfg {
  int nev = 256, guard = 32, max_sub = 864, max_iter = 60;
  double tol = 1e-6;
  double dmin = 1e-3;
  // residual level at which the products switch from FP32 to FP64 arithmetic (the
  // FP32 products stall near 1e-6 eV for matrices of norm ~10 eV)
  double switch_tol = 2e-5;
  // coarse block of the preconditioner: the `coarse` basis states of lowest diagonal
  // energy, whose block of H is inverted exactly (0: diagonal preconditioner only)
  int coarse = 0;
  double coarse_dmin = 1e-2;
};
struct DavidsonInfo {
  int iters = 0, matvecs = 0, converged = 0;
  int iters_fp64 = 0, matvecs_fp64 = 0;
  double max_resid = 0.0;
};
// lowest cfg.nev eigenpairs of the full Hermitian COMPLEX*8 matrix H; d_W (nev, device),
// d_resid (nev, device), d_X (N x nev, device, optional)
// Device workspace of the block Davidson, kept between calls (a plan owns one): it is
// allocated on the first call and grown only when a call needs more.
struct DavidsonWork;
DavidsonWork* davidson_work_new();
void davidson_work_delete(DavidsonWork* w);
int run_davidson(hipStream_t s, int N, const zf* H, int ldh, const double* D, const float* stats,
                 const DavidsonCfg& cfg, const MlpDev* mlp, double* d_W, double* d_resid, zc* d_X,
                 DavidsonInfo& info, DavidsonWork& work);

}  // namespace wtb
