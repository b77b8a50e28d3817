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
int launch_kgrid(hipStream_t s, const double* d_rlat, const int ng[3], const double sh[3],
                 double* d_kpt);
int launch_hr_stack(hipStream_t s, const double* hop, const double* ihop, const int* ff, int nvec,
                    int nw, zc* B);
int launch_hk(hipStream_t s, const DevModel& dm, const double* d_kpt, int nk, zc* d_phase,
              zc* d_hk4);
size_t jacobi_lds_bytes(int nw);
int launch_jacobi(hipStream_t s, const zc* d_hk, int nk, int nw, double ef32, double scs, int nc,
                  int nv, double* d_eall, zc* d_uout, int* d_nocp, double* d_eigv, zc* d_vec,
                  int* d_status);
int launch_dipoles(hipStream_t s, const zc* d_hk4, int nk, int nw, int nc, int nv, double sme32,
                   const double* d_eigv, const zc* d_vec, zc* d_dip, double* d_gapk);

// stage (c)
int launch_wtable(hipStream_t s, const double* d_rlat, const CoulombArgs& ca, double* d_wtab);
int launch_bse_build(hipStream_t s, int nk, int nc, int nv, int nw, const int ng[3],
                     const double* d_wtab, const double* d_eigv, const zc* d_vec, zc* d_H, int ldh,
                     zf* d_Hf);

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
  void* pub = nullptr;    // scalars published by the last workgroup of a step
  unsigned* cnt = nullptr;  // arrival counters
};
size_t tridiag_work_bytes(int N, int nb, int nrhs);
int tridiag_bind(TridiagWork& tw, void* base, int N, int nb, int nrhs);
int run_tridiag(hipStream_t s, TridiagWork& tw, zc* d_H, int ldh, zc* d_g);
// the same reduction on a COMPLEX*8 matrix (precision form)
int run_tridiag_f32(hipStream_t s, TridiagWork& tw, zf* d_H, int ldh, zc* d_g);

// stage (d): tridiagonal divide and conquer (eigenvalues + Z^T x for tracked vectors)
struct DCTree {
  int N = 0, ntrack = 0;
  std::vector<int> level_off, level_cnt, level_maxn;
};
size_t dc_work_bytes(int N, int ntrack);
int dc_prepare(void* work, int N, int ntrack, DCTree& tree);
// d_track: ntrack x N, in: the tridiagonal-basis vectors, out: Z^T (.)
int run_dc(hipStream_t s, void* work, const DCTree& tree, const double* d_d, const double* d_e,
           double* d_track, double* d_lambda);

// stage (e)
int launch_activities(hipStream_t s, int N, const double* d_G /* 6 x N */, int dtfull, int cpol,
                      double* d_act /* 8 x N */);
int launch_quantize(hipStream_t s, int n, const double* src, double* dst);
int launch_spectrum(hipStream_t s, int N, const double* d_e, const double* d_act6, double vc,
                    double nkf32, double ebse0, double ebsef, int numbse, double sme, int renorm,
                    double* d_work, double* d_eps);

}  // namespace wtb
