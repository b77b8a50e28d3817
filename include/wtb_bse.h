/*
 * wtb_bse.h -- C-ABI of the MI355X (gfx950) BSE pipeline.
 *
 * Drop-in boundary for the WanTiBEXOS Bethe-Salpeter path.  Every entry point
 * replaces one reference subroutine (or loop of calls) and is callable from the
 * reference's Fortran drivers through ISO_C_BINDING (see
 * wantibexos_amd/fortran/wtb_gpu.f90) or from C / ctypes.
 *
 * Conventions
 *  - Arrays are passed in the reference's Fortran (column-major) order and
 *    index meaning, unchanged.  Complex arrays are interleaved (re, im) doubles
 *    (COMPLEX*16).  Layout changes happen on the device.
 *  - Flags are plain int (never Fortran LOGICAL).
 *  - Every call returns a status (WTB_OK = 0); wtb_last_error() gives the text.
 *    Status codes map onto the reference's STOP messages.
 *  - A context owns one HIP stream on one device; calls on a context are not
 *    re-entrant.  Host arrays are borrowed for the duration of the call only.
 *  - All work runs on the GPU.  There is no CPU fallback: without a usable
 *    gfx950 device wtb_create fails.
 */
#ifndef WTB_BSE_H
#define WTB_BSE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  WTB_OK = 0,
  WTB_ERR_ARG = 1,       /* invalid argument (size, pointer, enum)                          */
  WTB_ERR_HIP = 2,       /* HIP runtime error                                               */
  WTB_ERR_NOMEM = 3,     /* device allocation failed                                        */
  WTB_ERR_COULOMB = 4,   /* "Wrong Coulomb Potential"            bse_subs.f90:88-91         */
  WTB_ERR_NVAL = 5,      /* "Number of valence states higher then ocuppied states"
                            general_subs.f90:521-524                                        */
  WTB_ERR_NCOND = 6,     /* "Number of conduction states higher then unocuppied states"
                            general_subs.f90:525-529                                        */
  WTB_ERR_NOCP = 7,      /* no band above E_F at some k: nocp undefined in the reference
                            (hamiltonian_tb.f90:112-127)                                    */
  WTB_ERR_EIG = 8,       /* "The algorithm failed to compute eigenvalues."
                            hamiltonian_tb.f90:98-101, bse_solver-tool-diel.f90:489-492     */
  WTB_ERR_LIMIT = 9,     /* a documented size limit is exceeded (nw > 64, ...)              */
  WTB_ERR_NONFINITE = 10 /* NaN or Inf in the input                                         */
};

/* ---- COULOMB_POT selector (bse_subs.f90:45-91) ------------------------------ */
enum {
  WTB_V2DK = 1, WTB_V3D = 2, WTB_V3DL = 3, WTB_V2DT = 4, WTB_V2DT2 = 5,
  WTB_V2DOH = 6, WTB_V2DRK = 7, WTB_V1DT = 8, WTB_V0DT = 9
};

/* ---- arithmetic forms ------------------------------------------------------- */
enum {
  WTB_FP64 = 0,   /* plain: everything in FP64 (drop-in for the reference)            */
  WTB_MIXED = 1   /* precision form: FP32 assembly / tridiagonal reduction, FP64
                     tridiagonal solve and FP64 refinement of the lowest excitons    */
};

typedef struct wtb_ctx wtb_ctx;
typedef struct wtb_bse_plan wtb_bse_plan;

/* Tight-binding model as hamiltonian_input_read leaves it in the module
 * hamiltonian_input_variables (subroutines/module_input_read.f90:1-30,32-176). */
typedef struct {
  int nw;                /* w90basis                                                  */
  int nvec;              /* number of R vectors                                       */
  const double* rlat;    /* rlat(3,3): rlat[i + 3*c] = component c of lattice vector i */
  const int* rvec;       /* rvec(nvec,3): rvec[i + nvec*c]                              */
  const int* ffactor;    /* ndegen(nvec)                                              */
  const double* hop;     /* hopmatrices(nvec,nw,nw): hop[i + nvec*(m + nw*n)] = Re H(R_i)_mn */
  const double* ihop;    /* ihopmatrices(nvec,nw,nw), imaginary parts                 */
  double efermi;         /* E_F (the reference subtracts its float32 rounding)          */
  double scs;            /* scissor added to the bands above E_F                      */
} wtb_tb_model;

/* The bsesolver / bsedielraw arguments that change the arithmetic
 * (bse_solver-tool-diel.f90:2-5, bse_diel-tool.f90:5, input keys of
 * module_input_read.f90:364-658). */
typedef struct {
  int ngrid[3];          /* NGX NGY NGZ                                               */
  double mshift[3];      /* SHIFT_1..3                                                */
  int nc, nv;            /* NBANDSC NBANDSV                                           */
  int coultype;          /* COULOMB_POT (WTB_V2DRK, ...)                              */
  double ediel[3];       /* EDIEL_T EDIEL EDIEL_B                                     */
  double lc, ez, w, r0;  /* LC EDIEL_Z W_COUL R_0                                     */
  double ktol;           /* KTOL                                                      */
  double cshift;         /* CSHIFT (passed in bsesolver's sme slot and to bsedielraw)  */
  double ebse0, ebsef;   /* ENSPECI ENSPECF                                           */
  int numbse;            /* NESPEC                                                    */
  int dtfull;            /* DTDIAG: xy,xz,yz activities (else zero)                   */
  int cpol;              /* CPOL: sp,sm activities (else zero)                        */
  int renorm;            /* RNMD: imagrenorm on eps2                                  */
  int f154;              /* 1: round energies/activities as the bse_opt_diel.dat
                            "(7F15.4)" write/read does before eps(omega)              */
  int precision;         /* WTB_FP64 or WTB_MIXED                                     */
  /* finite temperature (bsesolvertemp, subprograms/bse_solver-tool-diel-temp.f90),
   * BSET= T; all zero: bsesolver's T = 0 path */
  int bset;              /* 1: kernel terms times f_v - f_c of the row state, activities
                            Re[(X^T h_a) conj(X^T (f h_b))] (bse_subs_temp.f90:2-126,340-417) */
  int ta;                /* TA: WTB_TA_FA, WTB_TA_VE or WTB_TA_BE (gap correction)       */
  double temp;           /* TEMP (K); 0 is taken as 1E-8 (general_subs.f90:620-624)      */
  double st, phavg;      /* ST, PHAVG: parameters of the gap correction                 */
} wtb_bse_params;

/* ---- TA selector: the gap correction tcor added to the scissor
 *      (bse_solver-tool-diel-temp.f90:262-283) -------------------------------- */
enum {
  WTB_TA_FA = 0,   /* "FA" and any other value: tcor = 0                              */
  WTB_TA_VE = 1,   /* "VE": gapcortemp  (general_subs.f90:569-586)                    */
  WTB_TA_BE = 2    /* "BE": gapcortemp2 (general_subs.f90:588-604)                    */
};

/* ---- context ---------------------------------------------------------------- */
int wtb_create(int device, wtb_ctx** ctx);
int wtb_destroy(wtb_ctx* ctx);
const char* wtb_last_error(void);
int wtb_version(void);
void* wtb_stream(wtb_ctx* ctx);             /* the context's hipStream_t             */
int wtb_sync(wtb_ctx* ctx);                 /* wait for all work on the context      */

/* ---- single stages ---------------------------------------------------------- */

/* monhkhorst_pack (subroutines/mhkpack_subs.f90:359-407, recvec :409-432):
 * kpt(nk,3), nk = ngrid[0]*ngrid[1]*ngrid[2]; bit-identical to the reference. */
int wtb_monkhorst_pack(wtb_ctx* ctx, const int ngrid[3], const double shift[3],
                       const double* rlat, double* kpt);

/* eigsys (subroutines/hamiltonian_tb.f90:2-137) for nk k-points in one launch.
 * energias(nw,nk): ascending band energies, scissor applied above E_F.
 * autovetores(nw,nw,nk) complex: autovetores(i,:,k) = eigenvector i (the
 * reference's transposed layout), gauge of ZHEEV('V','L'): component 1 real.
 * nocp(nk): number of bands with E <= 0.  Limit: nw <= 64. */
int wtb_eigsys(wtb_ctx* ctx, const wtb_tb_model* model, int nk, const double* kpt,
               double* energias, double* autovetores, int* nocp);

/* ZHEEV('N' or 'V', 'U', n, a, lda, w) for the BSE matrix
 * (subprograms/bse_solver-tool-diel.f90:485-492).  The upper triangle of
 * a(lda,n) is read.  want_vectors = 0: eigenvalues only.  want_vectors = 1:
 * a is overwritten with the orthonormal eigenvectors (column j <-> w[j]),
 * each scaled so that its last component is real and >= 0 (ZHEEV's vectors
 * have a real last component of either sign). */
int wtb_zheev(wtb_ctx* ctx, int want_vectors, int n, double* a, int lda, double* w);

/* dielbsev (subroutines/bse_subs.f90:336-378): act(i) = |sum_j x(j,i) h(j)|^2 */
int wtb_dielbsev(wtb_ctx* ctx, int n, const double* x, const double* h, double* act);
/* dielbsep (bse_subs.f90:278-332): act(i) = Re[(x^T h1)_i conj((x^T h2)_i)]     */
int wtb_dielbsep(wtb_ctx* ctx, int n, const double* x, const double* h1, const double* h2,
                 double* act);

/* The bsedielraw loop (subprograms/bse_diel-tool.f90:111,135-201): optdiel
 * (optics.f90:160-244) with avgsme (:302-324) for xx,yy,zz (delta = 1) and
 * xy,xz,yz (delta = 0) on numbse photon energies, then imagrenorm
 * (optics.f90:327-390) if renorm.  e(n), act6(n,6) in order xx,yy,zz,xy,xz,yz.
 * eps(numbse,2,6): eps[j + numbse*(part + 2*comp)], part 0 = Re, 1 = Im,
 * without the spin factor and float32 conversion of the file write. */
int wtb_bse_spectrum(wtb_ctx* ctx, int n, const double* e, const double* act6,
                     const double* rlat, const int ngrid[3], double ebse0, double ebsef,
                     int numbse, double sme, int renorm, double* eps);

/* No counterpart in the reference.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * wtb_bse_spectrum with its staging and partial results in a library-owned static
 * device array (2^20 doubles, 8 MiB, held by every device the library runs on until the
 * process ends) instead of allocations: it calls no hipMalloc or hipFree.  Same
 * arguments and results as wtb_bse_spectrum.  Limit: 7 n + 12 numbse + 4 <= 2^20
 * (else WTB_ERR_LIMIT).  Every call on a device shares that device's array, so calls
 * that overlap on one device race. */
int wtb_bse_spectrum_noalloc(wtb_ctx* ctx, int n, const double* e, const double* act6,
                             const double* rlat, const int ngrid[3], double ebse0, double ebsef,
                             int numbse, double sme, int renorm, double* eps);

/* bseoptprop and spoptprop (subprograms/diel-pp-bse.f90:3-173, diel-pp.f90:3-172) and
 * their circular-polarisation forms bseoptproppol, spoptproppol (diel-pp-bse-pol.f90,
 * diel-pp-pol.f90): the optical constants of subroutines/diel-pp-subs.f90:1-73 for ncomp
 * spectra as the bse_diel*.dat / sp_diel*.dat files hold them.
 *   energy(n), eps_re(n, ncomp), eps_im(n, ncomp)
 *   out(n, ncomp, 5): refractive index, extinction coefficient, reflectivity,
 *                     absorption coefficient, energy-loss function               */
int wtb_optprop(wtb_ctx* ctx, int n, int ncomp, const double* energy, const double* eps_re,
                const double* eps_im, double* out);

/* lifetime / lifetimepol (subprograms/exciton_lifetime.f90:2-92, 95-187) with exclft
 * (subroutines/bse_subs.f90:472-530): radiative lifetimes (s).
 *   fosc(n, 1 + ncomp): energy, then the activities as bse_opt_diel.dat (ncomp = 6:
 *                       xx yy zz xy xz yz) or bse_opt_diel-pol.dat (ncomp = 5:
 *                       x y z sp sm) holds them
 *   lft(n, 1 + nout):   the lifetime for the summed strength of components 1..3, then
 *                       one per component 1..nout (3 for lifetime, 5 for lifetimepol)
 *   sysdim 1, 2, 3 for SYSDIM= "1D", "2D", "3D"; another value is WTB_ERR_ARG with the
 *   reference's "Wrong System dimension".                                          */
int wtb_exciton_lifetime(wtb_ctx* ctx, int sysdim, const int ngrid[3], const double* rlat, int n,
                         int ncomp, int nout, const double* fosc, double* lft);

/* ---- the whole pipeline -------------------------------------------------------
 * bsesolver (subprograms/bse_solver-tool-diel.f90:2-660) followed by the
 * spectrum of bsedielraw (subprograms/bse_diel-tool.f90:5-245), on the GPU:
 * k grid, H(k) phase GEMM, per-k Jacobi eigensolves, dipoles, W(k-k') table,
 * BSE assembly, Hermitian eigensolve, activities, eps(omega).
 *   w(N)           exciton energies, ascending (N = nk*nc*nv)
 *   act(N,8)       xx,yy,zz,xy,xz,yz,sp,sm activities (zero where the reference
 *                  zeroes them: dtfull/cpol off)
 *   eps(numbse,2,6) as wtb_bse_spectrum (may be NULL)
 *   egap           direct gap logged by the reference (:305-317) (may be NULL)
 * With prm->bset = 1 it is bsesolvertemp
 * (subprograms/bse_solver-tool-diel-temp.f90:2-660) instead: bands with the
 * scissor scs + tcor, the matrix of matrizelbsetemp, act from dielbseptemp.  */
int wtb_bsesolver(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm,
                  double* w, double* act, double* eps, double* egap);

/* Plan form of wtb_bsesolver for repeated runs: create uploads the model and
 * allocates every device buffer; execute enqueues the whole pipeline on the
 * context stream (inputs resident in HBM, no host synchronisation, no
 * allocation); results synchronises and copies out. */
int wtb_bse_plan_create(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm,
                        wtb_bse_plan** plan);
int wtb_bse_plan_dim(const wtb_bse_plan* plan);
int wtb_bse_plan_execute(wtb_bse_plan* plan);
/* Stages up to the BSE matrix only (k grid, H(k), per-k eig, dipoles, W table,
 * assembly): afterwards WTB_BUF_HBSE holds the matrix the reference hands to
 * ZHEEV (bse_solver-tool-diel.f90:458), lower triangle.  execute() later
 * overwrites it during the reduction. */
int wtb_bse_plan_build(wtb_bse_plan* plan);
int wtb_bse_plan_results(wtb_bse_plan* plan, double* w, double* act, double* eps,
                         double* egap);
/* Copy an intermediate device buffer to the host (tests, diagnostics).
 * what: see WTB_BUF_*; bytes must equal the buffer size (wtb_bse_plan_buffer_bytes). */
enum {
  WTB_BUF_KPT = 0,       /* kpt(nk,3) double                                          */
  WTB_BUF_EALL = 1,      /* energias per k: [k][nw] double                            */
  WTB_BUF_NOCP = 2,      /* nocp [k] int                                              */
  WTB_BUF_EIGV = 3,      /* selected bands [k][nc+nv] double                          */
  WTB_BUF_VECTOR = 4,    /* selected vectors [k][nc+nv][nw] complex                   */
  WTB_BUF_DIPOLE = 5,    /* hrx,hry,hrz [3][N] complex                                */
  WTB_BUF_HBSE = 6,      /* BSE matrix, column-major N x N complex (before the solve) */
  WTB_BUF_WTAB = 7,      /* W(Delta) table, [D1][D2][D3] double                       */
  WTB_BUF_TRIDIAG = 8,   /* d[N] then e[N] double (after the reduction)                */
  WTB_BUF_EXCWF = 9,     /* exciton vectors [count][N] complex (wtb_bse_plan_set_vectors) */
  WTB_BUF_FDEH = 10,     /* f_v - f_c per basis state [N] double (bset = 1)              */
  WTB_BUF_NBUF = 11
};
size_t wtb_bse_plan_buffer_bytes(const wtb_bse_plan* plan, int what);
int wtb_bse_plan_fetch(wtb_bse_plan* plan, int what, void* dst, size_t bytes);
/* Per-stage device time of the last execute (ms), measured with HIP events on
 * the context stream when timing is on: [0] k grid + H(k) + per-k eig,
 * [1] dipoles, [2] W table + assembly, [3] tridiagonal reduction,
 * [4] tridiagonal eigensolve, [5] activities + eps(omega). */
enum { WTB_NSTAGES = 6 };
/* on = 0 off, 1 stage events, 2 stage events plus an event pair around every
 * launch of the tridiagonal reduction's Hermitian GEMV kernel (the dominant
 * kernel), summed by wtb_bse_plan_kernel_ms. */
int wtb_bse_plan_set_timing(wtb_bse_plan* plan, int on);
int wtb_bse_plan_kernel_ms(const wtb_bse_plan* plan, double* total_ms, int* launches);
/* With the two-stage reduction the timed kernel is instead the trailing update of the
 * dense -> band stage (one launch per panel); flops = their algorithmic flop count
 * (8 m^2 2k per launch: m trailing rows, k reflectors); 0 for the one-stage GEMV. */
int wtb_bse_plan_kernel_flops(const wtb_bse_plan* plan, double* flops);
/* execute() captures the pipeline into a HIP graph on its first call and
 * replays it afterwards (default on); off launches the kernels directly.
 * Timing mode always launches directly. */
int wtb_bse_plan_set_graph(wtb_bse_plan* plan, int on);
int wtb_bse_plan_stage_ms(const wtb_bse_plan* plan, double* ms);
int wtb_bse_plan_destroy(wtb_bse_plan* plan);

/* ------------------------------------------------------------------------
 * Iterative form (no counterpart in the reference, which always runs the
 * full ZHEEV of bse_solver-tool-diel.f90:485-492): the lowest nev excitons of
 * the BSE matrix by block Davidson on the COMPLEX*8 matrix (FP32 MFMA
 * products, FP64 Rayleigh-Ritz and residuals), and eps(omega) from Lanczos
 * continued fractions of the dipole vectors instead of all N oscillator
 * strengths (so without the F15.4 round trip).  The plan must be created
 * with precision = WTB_MIXED.  The preconditioner is 1/(D_i - theta) or,
 * with an MLP, the network's estimate of the regularised resolvent
 * diagonal; with coarse > 0 it is replaced on the coarse states by the exact
 * inverse of their block of H - theta (eigendecomposed once).  When Davidson does not converge and fallback = 1, the full
 * FP64 solve of the plain path answers instead (used_fallback = 1).
 * ------------------------------------------------------------------------ */
typedef struct wtb_mlp wtb_mlp;
typedef struct {
  int nev;            /* lowest excitons wanted                              */
  int guard;          /* extra Ritz pairs carried along (e.g. 32)            */
  int max_sub;        /* basis size that triggers a restart (e.g. 5 (nev+guard)) */
  int max_iter;
  double tol;         /* residual 2-norm bound, eV                           */
  int lanczos_steps;  /* chain length for eps(omega); 0: no eps              */
  int fallback;       /* 1: full FP64 solve when Davidson does not converge   */
  int coarse;         /* states of lowest diagonal energy whose block of H the
                         preconditioner inverts exactly (e.g. 10 (nev+guard),
                         at most N/4);
                         0: diagonal (or learned) preconditioner only          */
} wtb_davidson_params;
typedef struct {
  int iters, matvecs, converged, used_fallback;
  int iters_fp64, matvecs_fp64;     /* of those, with FP64 products (final polish) */
  double max_resid;                 /* eV */
  double ms_build, ms_solve, ms_eps; /* device time of the three phases */
} wtb_davidson_info;
/* W[nev] ascending, resid[nev] residual norms, eps[6][2][numbse] (optional),
 * X[nev][N] complex eigenvectors (optional) */
int wtb_bse_plan_lowest(wtb_bse_plan* plan, const wtb_davidson_params* dp, const wtb_mlp* mlp, double* W,
                        double* resid, double* eps, double* X, wtb_davidson_info* info);
/* diagonal D[N] and row statistics stats[3][N] (sum |H_ij|, sqrt(sum |H_ij|^2),
 * max |H_ij|, j != i) of the plan's matrix: the preconditioner's inputs */
int wtb_bse_plan_row_stats(wtb_bse_plan* plan, double* D, float* stats);

/* The preconditioner network: 5 -> 128 -> 128 -> 1, tanh, FP32, trained and
 * evaluated by the library's own kernels.  Features X[5][S] as described in
 * wantibexos_amd/csrc/k_mlp.hip; parameters packed as mu[5] sd[5] W1 b1 W2 b2 w3 b3. */
int wtb_mlp_create(wtb_ctx* ctx, int hidden, wtb_mlp** out);
int wtb_mlp_nparams(const wtb_mlp* mlp);
int wtb_mlp_set_params(wtb_mlp* mlp, const float* p, int n);
int wtb_mlp_get_params(const wtb_mlp* mlp, float* p, int n);
int wtb_mlp_forward(wtb_mlp* mlp, int S, const float* X, float* y);
/* mean squared error over all S samples and its gradient (nparams floats) */
int wtb_mlp_loss_grad(wtb_mlp* mlp, int S, const float* X, const float* t, float* loss, float* grad);
/* Adam, shuffled batches, loss_hist[epochs] = mean batch loss per epoch */
int wtb_mlp_train(wtb_mlp* mlp, int S, const float* X, const float* t, int epochs, int batch, float lr,
                  unsigned seed, float* loss_hist);
int wtb_mlp_destroy(wtb_mlp* mlp);

/* ------------------------------------------------------------------------
 * Panel width of the blocked Householder reduction (8..32 columns per
 * trailing MFMA update; default 32) and its autotuner, which times the
 * reduction for widths 16, 20, 24, 32 on the plan's own matrix and keeps the
 * fastest (ms[4]: their times).  No counterpart in the reference.
 * ------------------------------------------------------------------------ */
int wtb_bse_plan_set_panel(wtb_bse_plan* plan, int nb);
int wtb_bse_plan_tune(wtb_bse_plan* plan, int* best, double* ms);

/* ------------------------------------------------------------------------
 * Batching form (no counterpart in the reference, which solves one model
 * per run): many independent models of uneven sizes, one plan each, assigned
 * to nlanes streams largest-first; execute() replays every matrix's captured
 * graph on its stream and joins on the context stream; results() as for a plan.
 * ------------------------------------------------------------------------ */
typedef struct wtb_bse_batch wtb_bse_batch;
int wtb_bse_batch_create(wtb_ctx* ctx, int n, const wtb_tb_model* models, const wtb_bse_params* prms,
                         int nlanes, wtb_bse_batch** out);
int wtb_bse_batch_size(const wtb_bse_batch* batch);
int wtb_bse_batch_dim(const wtb_bse_batch* batch, int i);
int wtb_bse_batch_tune(wtb_bse_batch* batch, int* tiles);
int wtb_bse_batch_execute(wtb_bse_batch* batch);
int wtb_bse_batch_results(wtb_bse_batch* batch, int i, double* W, double* act, double* eps, double* egap);
int wtb_bse_batch_destroy(wtb_bse_batch* batch);

/* ---- finite-Q exciton bands (BSE_BND= T) ----------------------------------------
 * kpathbse (subroutines/mhkpack_subs.f90:132-208): ks(nks,3) fractional segment ends
 * (pairs), npts points per segment -> qauxv((nks/2)*npts, 4): path coordinate, qx, qy,
 * qz; labels (nks/2 + 1, may be NULL): the coordinates of the segment ends
 * (KLABELS-BSE.dat).  Host only.                                                  */
int wtb_kpathbse(const double* rlat, int nks, const double* ks, int npts, double* qauxv, double* labels);
/* bsebnds (subprograms/bse_kpath-tool.f90:3-479) for nq q-points: exk(N, nq) the
 * ascending exciton energies of the finite-Q BSE matrix (matrizelbsekq,
 * subroutines/bse_subs_kpath.f90:1-214: conduction bands at k+q, valence bands at
 * k, direct term V(k-k')<c|c'><v|v'>, exchange -V(q)<c|v><v'|c'> for q != 0) at each
 * q = qauxv(iq, 2:4).  The q-points run on `lanes` HIP streams (1..8), one plan
 * each.  With prm->bset = 1 it is bsebndstemp (subprograms/bse_kpath-tool-temp.f90,
 * matrizelbsekqtemp in subroutines/bse_subs_temp.f90:128-338): the gap correction
 * on both grids, the kernel (direct - exchange) times the row's f_v - f_c.         */
int wtb_bsebnds(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm, int nq, const double* qauxv,
                int lanes, double* exk);
/* The same with the plans kept: create (one finite-Q plan and stream per lane, each
 * captured into a HIP graph at its first q), run any number of paths, destroy.  */
typedef struct wtb_bse_bands wtb_bse_bands;
int wtb_bse_bands_create(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm, int lanes,
                         wtb_bse_bands** bands);
int wtb_bse_bands_run(wtb_bse_bands* bands, int nq, const double* qauxv, double* exk);
/* on = 0: launch the kernels directly instead of replaying the captured graphs
 * (for kernel tracers)                                                            */
int wtb_bse_bands_set_graph(wtb_bse_bands* bands, int on);
/* BSE_WF along the path (bse_kpath-tool.f90:359-366, excwfi2 in bse_subs.f90:578-622):
 * runs also form the vectors of excitons first .. first+count-1 (0-based, ascending)
 * at every q; run_wf returns them as wf[nq][count][N] complex, each with its last
 * component real and >= 0.                                                        */
int wtb_bse_bands_set_vectors(wtb_bse_bands* bands, int first, int count);
int wtb_bse_bands_run_wf(wtb_bse_bands* bands, int nq, const double* qauxv, double* exk, double* wf);
/* wtb_bsebnds plus those vectors and, for excwfi2, the k grid kpt(nk,3) and its
 * nocp(nk) (kpt, nocp may be NULL)                                                */
int wtb_bsebnds_wf(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm, int nq,
                   const double* qauxv, int lanes, double* exk, int wf_first, int wf_count, double* wf,
                   double* kpt, int* nocp);
int wtb_bse_bands_destroy(wtb_bse_bands* bands);
/* Plan form of one q-point: create, then wtb_bse_plan_set_q(plan, q[4]) (a qauxv
 * row) before build/execute; results gives w only (act, eps, egap NULL);
 * WTB_BUF_HBSE holds the finite-Q matrix after build.                            */
int wtb_bse_qplan_create(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm,
                         wtb_bse_plan** plan);
int wtb_bse_plan_set_q(wtb_bse_plan* plan, const double* q);

/* The single-particle table bsesolver writes as sp_optics.dat, sp_optics-pol.dat and
 * (TM_COEF) tm_coef*.dat (bse_solver-tool-diel.f90:336-420, with the finite-temperature
 * factor of bse_solver-tool-diel-temp.f90:405-424 when prm->bset): vecres(N, 15),
 * REAL, rows sorted by transition energy (column 7) as Bubblem sorts them
 * (general_subs.f90:405-440).  The plan form needs a built or executed plan.       */
int wtb_bse_plan_sp_table(wtb_bse_plan* plan, float* vecres);
int wtb_sp_table(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm, float* vecres);

/* BSE_WF (bse_solver-tool-diel.f90:513-519, excwf2 in bse_subs.f90:534-576):
 * executes also form the eigenvectors of excitons first .. first+count-1
 * (0-based, ascending energy), fetched as WTB_BUF_EXCWF, each with its last
 * component real and >= 0 (ZHEEV('V','U') leaves it real with either sign). */
int wtb_bse_plan_set_vectors(wtb_bse_plan* plan, int first, int count);
/* Form of the tridiagonal reduction of the BSE matrix (the first step of the ZHEEV of
 * bse_solver-tool-diel.f90:485-492; no counterpart in the reference's arithmetic):
 * two_stage = 0: one-stage blocked Householder; two_stage = 1: dense -> band of width
 * `band` (16 or 32, MFMA panel updates, lower triangle only) -> tridiagonal by bulge
 * chasing; two_stage = -1 (default): the two-stage form for N >= 8192, else one-stage.
 * Executes that form exciton vectors (BSE_WF, the refinement) always take the one-stage
 * form, whose reflectors they need.                                                    */
int wtb_bse_plan_set_reduction(wtb_bse_plan* plan, int two_stage, int band);

/* Precision form (WTB_MIXED, Q = 0): FP64 Rayleigh-quotient refinement of the nref lowest
 * excitons, the refinement of bse_solver-tool-diel.f90:485-492's eigenvalues that the
 * precision form adds (no counterpart in the reference).  Executes then also assemble the
 * FP64 matrix, form the exciton vectors x of the lowest nref (divide-and-conquer vectors
 * and the back-transform through the FP32 reflectors), and replace each of those energies
 * by theta = x^H H x / x^H x with the FP64 H; |H x - theta x| / |x| bounds the distance of
 * theta to an eigenvalue of the FP64 matrix.  nref = 0 turns it off (default).        */
int wtb_bse_plan_set_refine(wtb_bse_plan* plan, int nref);
/* after an execute: resid[n] those residual norms (eV) and, if w_unrefined is not NULL,
 * the energies before the refinement (n <= nref)                                       */
int wtb_bse_plan_refine_results(wtb_bse_plan* plan, int n, double* resid, double* w_unrefined);

/* wtb_bsesolver plus those vectors (wf: [count][N] complex) and the k grid
 * (kpt(nk,3) Fortran order) and nocpk(nk) that excwf2 prints beside them */
int wtb_bsesolver_wf(wtb_ctx* ctx, const wtb_tb_model* model, const wtb_bse_params* prm, double* W,
                     double* act, double* eps, double* egap, int wf_first, int wf_count, double* wf,
                     double* kpt, int* nocp);

#ifdef __cplusplus
}
#endif
#endif /* WTB_BSE_H */
