"""Python host interface over the C-ABI of libwtb_bse.so (include/wtb_bse.h).

The functions mirror the reference subroutines they replace (same names,
argument meaning, and STOP conditions raised as exceptions):

    monkhorst_pack   subroutines/mhkpack_subs.f90:359-407
    eigsys           subroutines/hamiltonian_tb.f90:2-137 (all k at once)
    zheev            ZHEEV('N','U') of subprograms/bse_solver-tool-diel.f90:485-492
    dielbsev/p       subroutines/bse_subs.f90:278-378
    bse_spectrum     the bsedielraw loop, subprograms/bse_diel-tool.f90:135-201
    bsesolver        subprograms/bse_solver-tool-diel.f90:2-660 + spectrum

Everything runs on the GPU; the library refuses to start without a gfx950
device, and importing this module without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from pathlib import Path

import numpy as np

from .tbmodel import TBModel

_LIB_PATH = Path(__file__).resolve().parent / "libwtb_bse.so"

WTB_OK = 0
ERRORS = {
    1: "invalid argument", 2: "HIP runtime error", 3: "device allocation failed",
    4: "Wrong Coulomb Potential",
    5: "Number of valence states higher then ocuppied states",
    6: "Number of conduction states higher then unocuppied states",
    7: "no band above the Fermi level at some k",
    8: "The algorithm failed to compute eigenvalues.",
    9: "size limit exceeded", 10: "NaN or Inf in the input",
}
COULOMB = {"V2DK": 1, "V3D": 2, "V3DL": 3, "V2DT": 4, "V2DT2": 5, "V2DOH": 6, "V2DRK": 7,
           "V1DT": 8, "V0DT": 9}
FP64, MIXED = 0, 1
# TA (bse_solver-tool-diel-temp.f90:262-283): any value other than VE and BE means no gap correction
TA = {"FA": 0, "VE": 1, "BE": 2}
BUF = dict(kpt=0, eall=1, nocp=2, eigv=3, vector=4, dipole=5, hbse=6, wtab=7, tridiag=8, excwf=9, fdeh=10)
NSTAGES = 6
STAGES = ["eig", "dipoles", "build", "tridiag", "dc", "spectrum"]


class WTBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[wtb {code}] {msg}")
        self.code = code


class _Model(C.Structure):
    _fields_ = [("nw", C.c_int), ("nvec", C.c_int), ("rlat", C.c_void_p), ("rvec", C.c_void_p),
                ("ffactor", C.c_void_p), ("hop", C.c_void_p), ("ihop", C.c_void_p),
                ("efermi", C.c_double), ("scs", C.c_double)]


class _Params(C.Structure):
    _fields_ = [("ngrid", C.c_int * 3), ("mshift", C.c_double * 3), ("nc", C.c_int), ("nv", C.c_int),
                ("coultype", C.c_int), ("ediel", C.c_double * 3), ("lc", C.c_double),
                ("ez", C.c_double), ("w", C.c_double), ("r0", C.c_double), ("ktol", C.c_double),
                ("cshift", C.c_double), ("ebse0", C.c_double), ("ebsef", C.c_double),
                ("numbse", C.c_int), ("dtfull", C.c_int), ("cpol", C.c_int), ("renorm", C.c_int),
                ("f154", C.c_int), ("precision", C.c_int), ("bset", C.c_int), ("ta", C.c_int),
                ("temp", C.c_double), ("st", C.c_double), ("phavg", C.c_double)]


class _DavidsonParams(C.Structure):
    _fields_ = [("nev", C.c_int), ("guard", C.c_int), ("max_sub", C.c_int), ("max_iter", C.c_int),
                ("tol", C.c_double), ("lanczos_steps", C.c_int), ("fallback", C.c_int), ("coarse", C.c_int)]


class _DavidsonInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("matvecs", C.c_int), ("converged", C.c_int),
                ("used_fallback", C.c_int), ("iters_fp64", C.c_int), ("matvecs_fp64", C.c_int),
                ("max_resid", C.c_double), ("ms_build", C.c_double),
                ("ms_solve", C.c_double), ("ms_eps", C.c_double)]


def _load():
    if not _LIB_PATH.exists():
        raise ImportError(f"{_LIB_PATH} is not built: run python -m wantibexos_amd.build "
                          "(there is no CPU fallback)")
    lib = C.CDLL(str(_LIB_PATH))
    vp, ip, dp = C.c_void_p, C.c_int, C.POINTER(C.c_double)
    sig = {
        "wtb_create": (ip, [ip, C.POINTER(vp)]),
        "wtb_destroy": (ip, [vp]),
        "wtb_last_error": (C.c_char_p, []),
        "wtb_version": (ip, []),
        "wtb_stream": (vp, [vp]),
        "wtb_sync": (ip, [vp]),
        "wtb_monkhorst_pack": (ip, [vp, vp, vp, vp, vp]),
        "wtb_eigsys": (ip, [vp, vp, ip, vp, vp, vp, vp]),
        "wtb_zheev": (ip, [vp, ip, ip, vp, ip, vp]),
        "wtb_dielbsev": (ip, [vp, ip, vp, vp, vp]),
        "wtb_dielbsep": (ip, [vp, ip, vp, vp, vp, vp]),
        "wtb_bse_spectrum": (ip, [vp, ip, vp, vp, vp, vp, C.c_double, C.c_double, ip, C.c_double,
                                  ip, vp]),
        "wtb_bsesolver": (ip, [vp, vp, vp, vp, vp, vp, vp]),
        "wtb_bse_plan_create": (ip, [vp, vp, vp, C.POINTER(vp)]),
        "wtb_bse_plan_dim": (ip, [vp]),
        "wtb_bse_plan_execute": (ip, [vp]),
        "wtb_bse_plan_build": (ip, [vp]),
        "wtb_bse_plan_results": (ip, [vp, vp, vp, vp, vp]),
        "wtb_bse_plan_buffer_bytes": (C.c_size_t, [vp, ip]),
        "wtb_bse_plan_fetch": (ip, [vp, ip, vp, C.c_size_t]),
        "wtb_bse_plan_set_timing": (ip, [vp, ip]),
        "wtb_bse_plan_set_graph": (ip, [vp, ip]),
        "wtb_bse_plan_stage_ms": (ip, [vp, vp]),
        "wtb_bse_plan_kernel_ms": (ip, [vp, vp, vp]),
        "wtb_bse_plan_destroy": (ip, [vp]),
        "wtb_bse_plan_lowest": (ip, [vp, vp, vp, vp, vp, vp, vp, vp]),
        "wtb_bse_plan_row_stats": (ip, [vp, vp, vp]),
        "wtb_mlp_create": (ip, [vp, ip, C.POINTER(vp)]),
        "wtb_mlp_nparams": (ip, [vp]),
        "wtb_mlp_set_params": (ip, [vp, vp, ip]),
        "wtb_mlp_get_params": (ip, [vp, vp, ip]),
        "wtb_mlp_forward": (ip, [vp, ip, vp, vp]),
        "wtb_mlp_loss_grad": (ip, [vp, ip, vp, vp, vp, vp]),
        "wtb_mlp_train": (ip, [vp, ip, vp, vp, ip, ip, C.c_float, C.c_uint, vp]),
        "wtb_mlp_destroy": (ip, [vp]),
        "wtb_bse_plan_set_panel": (ip, [vp, ip]),
        "wtb_bse_plan_set_vectors": (ip, [vp, ip, ip]),
        "wtb_bsesolver_wf": (ip, [vp, vp, vp, vp, vp, vp, vp, ip, ip, vp, vp, vp]),
        "wtb_bse_plan_tune": (ip, [vp, vp, vp]),
        "wtb_bse_batch_create": (ip, [vp, ip, vp, vp, ip, C.POINTER(vp)]),
        "wtb_bse_batch_size": (ip, [vp]),
        "wtb_bse_batch_dim": (ip, [vp, ip]),
        "wtb_bse_batch_tune": (ip, [vp, vp]),
        "wtb_bse_batch_execute": (ip, [vp]),
        "wtb_bse_batch_results": (ip, [vp, ip, vp, vp, vp, vp]),
        "wtb_bse_batch_destroy": (ip, [vp]),
        "wtb_kpathbse": (ip, [vp, ip, vp, ip, vp, vp]),
        "wtb_bsebnds": (ip, [vp, vp, vp, ip, vp, ip, vp]),
        "wtb_bse_qplan_create": (ip, [vp, vp, vp, C.POINTER(vp)]),
        "wtb_bse_plan_set_q": (ip, [vp, vp]),
        "wtb_bse_bands_create": (ip, [vp, vp, vp, ip, C.POINTER(vp)]),
        "wtb_bse_bands_run": (ip, [vp, ip, vp, vp]),
        "wtb_bse_bands_destroy": (ip, [vp]),
        "wtb_bse_bands_set_graph": (ip, [vp, ip]),
        "wtb_bse_plan_sp_table": (ip, [vp, vp]),
        "wtb_sp_table": (ip, [vp, vp, vp, vp]),
        "wtb_bse_bands_set_vectors": (ip, [vp, ip, ip]),
        "wtb_bse_bands_run_wf": (ip, [vp, ip, vp, vp, vp]),
        "wtb_bsebnds_wf": (ip, [vp, vp, vp, ip, vp, ip, vp, ip, ip, vp, vp, vp]),
        "wtb_optprop": (ip, [vp, ip, ip, vp, vp, vp, vp]),
        "wtb_bse_plan_set_refine": (ip, [vp, ip]),
        "wtb_bse_plan_set_reduction": (ip, [vp, ip, ip]),
        "wtb_bse_plan_kernel_flops": (ip, [vp, vp]),
        "wtb_bse_plan_refine_results": (ip, [vp, ip, vp, vp]),
        "wtb_bse_spectrum_noalloc": (ip, [vp, ip, vp, vp, vp, vp, C.c_double, C.c_double, ip, C.c_double,
                                          ip, vp]),
        "wtb_exciton_lifetime": (ip, [vp, ip, vp, vp, ip, ip, ip, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _check(rc):
    if rc != WTB_OK:
        msg = lib().wtb_last_error().decode(errors="replace") or ERRORS.get(rc, "error")
        raise WTBError(rc, msg)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


@dataclasses.dataclass
class BSEParams:
    """bsesolver/bsedielraw arguments (input keys of module_input_read.f90:364-658)."""
    ngrid: tuple = (1, 1, 1)           # NGX NGY NGZ
    mshift: tuple = (0.0, 0.0, 0.0)    # SHIFT_1..3
    nc: int = 1                        # NBANDSC
    nv: int = 1                        # NBANDSV
    coultype: str = "V3D"              # COULOMB_POT
    ediel: tuple = (1.0, 1.0, 1.0)     # EDIEL_T EDIEL EDIEL_B
    lc: float = 1.0                    # LC
    ez: float = 1.0                    # EDIEL_Z
    w: float = 0.0                     # W_COUL
    r0: float = 1.0                    # R_0
    ktol: float = float(np.float32(0.001))   # KTOL default is the REAL literal 0.001
    cshift: float = float(np.float32(0.01))  # CSHIFT default 0.01 (REAL literal)
    ebse0: float = 0.0                 # ENSPECI
    ebsef: float = 3.0                 # ENSPECF
    numbse: int = 6001                 # NESPEC
    dtfull: bool = False               # DTDIAG
    cpol: bool = False                 # CPOL
    renorm: bool = True                # RNMD
    f154: bool = True                  # bse_opt_diel.dat F15.4 round trip
    precision: int = FP64
    # finite temperature (BSET= T: bsesolvertemp)
    bset: bool = False                 # BSET
    temp: float = 0.0                  # TEMP (K)
    ta: str = "FA"                     # TA: FA, VE or BE
    st: float = 0.0                    # ST
    phavg: float = 0.0                 # PHAVG

    @property
    def nk(self):
        return int(np.prod(self.ngrid))

    @property
    def dim(self):
        return self.nk * self.nc * self.nv

    def to_c(self) -> _Params:
        ct = self.coultype.strip().upper()
        p = _Params()
        p.ngrid[:] = [int(v) for v in self.ngrid]
        p.mshift[:] = [float(v) for v in self.mshift]
        p.nc, p.nv = int(self.nc), int(self.nv)
        p.coultype = COULOMB.get(ct, -1)
        p.ediel[:] = [float(v) for v in self.ediel]
        p.lc, p.ez, p.w, p.r0, p.ktol = map(float, (self.lc, self.ez, self.w, self.r0, self.ktol))
        p.cshift, p.ebse0, p.ebsef = map(float, (self.cshift, self.ebse0, self.ebsef))
        p.numbse = int(self.numbse)
        p.dtfull, p.cpol, p.renorm, p.f154 = (int(bool(v)) for v in
                                              (self.dtfull, self.cpol, self.renorm, self.f154))
        p.precision = int(self.precision)
        p.bset = int(bool(self.bset))
        p.ta = TA.get(str(self.ta).strip().upper(), 0)
        p.temp, p.st, p.phavg = map(float, (self.temp, self.st, self.phavg))
        return p


class _CModel:
    """Keeps Fortran-ordered copies of a TBModel alive for one call."""

    def __init__(self, m: TBModel):
        self.rlat = np.asfortranarray(np.asarray(m.rlat, dtype=np.float64))
        self.rvec = np.asfortranarray(np.asarray(m.rvec, dtype=np.int32))
        self.ff = np.ascontiguousarray(np.asarray(m.ndeg, dtype=np.int32))
        self.hop = np.asfortranarray(np.asarray(m.hop, dtype=np.float64))
        self.ihop = np.asfortranarray(np.asarray(m.ihop, dtype=np.float64))
        self.s = _Model(m.nw, m.nvec, _p(self.rlat), _p(self.rvec), _p(self.ff), _p(self.hop),
                        _p(self.ihop), float(m.efermi), float(m.scs))


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _check(lib().wtb_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().wtb_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return lib().wtb_stream(self.h)

    def sync(self):
        _check(lib().wtb_sync(self.h))


def monkhorst_pack(ctx: Context, ngrid, shift, rlat):
    ng = np.array(ngrid, dtype=np.int32)
    sh = np.array(shift, dtype=np.float64)
    rl = np.asfortranarray(np.asarray(rlat, dtype=np.float64))
    kpt = np.zeros((int(np.prod(ng)), 3), dtype=np.float64, order="F")
    _check(lib().wtb_monkhorst_pack(ctx.h, _p(ng), _p(sh), _p(rl), _p(kpt)))
    return np.ascontiguousarray(kpt)


def eigsys(ctx: Context, model: TBModel, kpt):
    """Returns (E (nk,nw) ascending incl. scissor, U (nk,nw,nw) with U[k,i,:] =
    eigenvector i (the reference's autovetores(i,:)), nocp (nk,))."""
    cm = _CModel(model)
    k = np.asfortranarray(np.asarray(kpt, dtype=np.float64))
    nk, nw = k.shape[0], model.nw
    E = np.zeros((nk, nw))
    U = np.zeros((nk, nw, nw), dtype=np.complex128)     # [k][m][i] in C order = autovetores(i,m,k)
    nocp = np.zeros(nk, dtype=np.int32)
    _check(lib().wtb_eigsys(ctx.h, C.byref(cm.s), nk, _p(k), _p(E), _p(U), _p(nocp)))
    return E, np.ascontiguousarray(U.transpose(0, 2, 1)), nocp


def zheev(ctx: Context, a, vectors: bool = False):
    """ZHEEV('N' or 'V', 'U') of the Hermitian matrix whose upper triangle is `a`:
    the eigenvalues, or (eigenvalues, eigenvectors) with each vector's last
    component real and >= 0."""
    a = np.array(a, dtype=np.complex128, order="F", copy=True)
    n = a.shape[0]
    w = np.zeros(n)
    _check(lib().wtb_zheev(ctx.h, int(vectors), n, _p(a), n, _p(w)))
    return (w, a) if vectors else w


def dielbsev(ctx: Context, x, h):
    x = np.asfortranarray(np.asarray(x, dtype=np.complex128))
    h = np.ascontiguousarray(h, dtype=np.complex128)
    act = np.zeros(x.shape[0])
    _check(lib().wtb_dielbsev(ctx.h, x.shape[0], _p(x), _p(h), _p(act)))
    return act


def dielbsep(ctx: Context, x, h1, h2):
    x = np.asfortranarray(np.asarray(x, dtype=np.complex128))
    h1 = np.ascontiguousarray(h1, dtype=np.complex128)
    h2 = np.ascontiguousarray(h2, dtype=np.complex128)
    act = np.zeros(x.shape[0])
    _check(lib().wtb_dielbsep(ctx.h, x.shape[0], _p(x), _p(h1), _p(h2), _p(act)))
    return act


def bse_spectrum(ctx: Context, e, act6, rlat, ngrid, ebse0, ebsef, numbse, sme, renorm=True):
    """eps (6 components xx,yy,zz,xy,xz,yz; 2 parts re,im; numbse)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    a6 = np.asfortranarray(np.asarray(act6, dtype=np.float64).T)   # act6(n,6)
    rl = np.asfortranarray(np.asarray(rlat, dtype=np.float64))
    ng = np.array(ngrid, dtype=np.int32)
    eps = np.zeros((6, 2, int(numbse)))
    _check(lib().wtb_bse_spectrum(ctx.h, len(e), _p(e), _p(a6), _p(rl), _p(ng), float(ebse0),
                                  float(ebsef), int(numbse), float(sme), int(bool(renorm)), _p(eps)))
    return eps


def optprop(ctx: Context, energy, eps_re, eps_im):
    """bseoptprop / spoptprop and their -pol forms (diel-pp-bse.f90:3-173): eps_re, eps_im
    (n, ncomp) -> (n, k, R, alpha, loss), each (n, ncomp)."""
    en = np.ascontiguousarray(energy, dtype=np.float64)
    er = np.asfortranarray(eps_re, dtype=np.float64)
    ei = np.asfortranarray(eps_im, dtype=np.float64)
    n, nc = er.shape
    out = np.zeros((n, nc, 5), order="F")
    _check(lib().wtb_optprop(ctx.h, n, nc, _p(en), _p(er), _p(ei), _p(out)))
    return tuple(out[:, :, q] for q in range(5))


SYSDIM = {"1D": 1, "2D": 2, "3D": 3}


def exciton_lifetime(ctx: Context, sysdim: str, ngrid, rlat, table, nout: int):
    """lifetime (nout = 3) / lifetimepol (nout = 5), exciton_lifetime.f90:2-187: table
    (N, 1 + ncomp) as bse_opt_diel(-pol).dat holds it -> (N, 1 + nout) lifetimes (s)."""
    t = np.asfortranarray(table, dtype=np.float64)
    n, ncol = t.shape
    rl = np.asfortranarray(np.asarray(rlat, dtype=np.float64))
    ng = np.array(ngrid, dtype=np.int32)
    out = np.zeros((n, nout + 1), order="F")
    _check(lib().wtb_exciton_lifetime(ctx.h, SYSDIM.get(sysdim, 0), _p(ng), _p(rl), n, ncol - 1, nout, _p(t),
                                      _p(out)))
    return out


def bse_spectrum_noalloc(ctx: Context, e, act6, rlat, ngrid, ebse0, ebsef, numbse, sme, renorm=True):
    """bse_spectrum without device allocations (the library's static scratch array;
    7 n + 12 numbse + 4 <= 2^20)."""
    e = np.ascontiguousarray(e, dtype=np.float64)
    a6 = np.asfortranarray(np.asarray(act6, dtype=np.float64).T)
    rl = np.asfortranarray(np.asarray(rlat, dtype=np.float64))
    ng = np.array(ngrid, dtype=np.int32)
    eps = np.zeros((6, 2, int(numbse)))
    _check(lib().wtb_bse_spectrum_noalloc(ctx.h, len(e), _p(e), _p(a6), _p(rl), _p(ng), float(ebse0),
                                          float(ebsef), int(numbse), float(sme), int(bool(renorm)), _p(eps)))
    return eps


class BSEPlan:
    """Device-resident pipeline: create once, execute many times.  finite_q=True: the
    finite-Q plan of the exciton bands (set_q, then build/execute; energies() only)."""

    def __init__(self, ctx: Context, model: TBModel, params: BSEParams, finite_q: bool = False):
        self.ctx = ctx
        self.params = params
        self.finite_q = finite_q
        self._cm = _CModel(model)
        self._cp = params.to_c()
        h = C.c_void_p()
        create = lib().wtb_bse_qplan_create if finite_q else lib().wtb_bse_plan_create
        _check(create(ctx.h, C.byref(self._cm.s), C.byref(self._cp), C.byref(h)))
        self.h = h
        self.N = lib().wtb_bse_plan_dim(h)

    def sp_table(self):
        """The single-particle table vecres (N, 15), float32, sorted by transition energy
        (the sp_optics.dat / tm_coef.dat content); needs build() or execute() first."""
        t = np.zeros((15, self.N), dtype=np.float32)
        _check(lib().wtb_bse_plan_sp_table(self.h, _p(t)))
        return t.T.copy()

    def set_q(self, q):
        """A row of the q path (coordinate, qx, qy, qz) for the next build/execute."""
        qa = np.ascontiguousarray(q, dtype=np.float64)
        assert qa.shape == (4,)
        _check(lib().wtb_bse_plan_set_q(self.h, _p(qa)))

    def energies(self):
        """The exciton energies of the last execute (the only result of a finite-Q plan)."""
        W = np.zeros(self.N)
        _check(lib().wtb_bse_plan_results(self.h, _p(W), None, None, None))
        return W

    def execute(self):
        _check(lib().wtb_bse_plan_execute(self.h))

    def build(self):
        """Stages up to the BSE matrix only; fetch('hbse') then holds it."""
        _check(lib().wtb_bse_plan_build(self.h))

    def set_timing(self, on=True):
        """0/False off, 1/True per-stage events, 2 also per-launch events on the
        tridiagonal reduction's GEMV kernel (kernel_ms)."""
        _check(lib().wtb_bse_plan_set_timing(self.h, int(on)))

    def kernel_ms(self):
        """(summed ms, launches) of the GEMV kernel in the last timing-2 execute."""
        t = np.zeros(1)
        n = np.zeros(1, dtype=np.int32)
        _check(lib().wtb_bse_plan_kernel_ms(self.h, _p(t), _p(n)))
        return float(t[0]), int(n[0])

    def kernel_flops(self):
        """Algorithmic flops of the timed two-stage trailing updates (0: one-stage GEMV)."""
        f = np.zeros(1)
        _check(lib().wtb_bse_plan_kernel_flops(self.h, _p(f)))
        return float(f[0])

    def set_graph(self, on=True):
        _check(lib().wtb_bse_plan_set_graph(self.h, int(on)))

    def results(self):
        N, nb = self.N, int(self.params.numbse)
        W = np.zeros(N)
        act = np.zeros((8, N))
        eps = np.zeros((6, 2, nb))
        egap = np.zeros(1)
        _check(lib().wtb_bse_plan_results(self.h, _p(W), _p(act), _p(eps), _p(egap)))
        return dict(W=W, act=act, eps=eps, egap=float(egap[0]))

    def stage_ms(self):
        ms = np.zeros(NSTAGES)
        _check(lib().wtb_bse_plan_stage_ms(self.h, _p(ms)))
        return dict(zip(STAGES, ms))

    def fetch(self, what: str):
        code = BUF[what]
        nbytes = lib().wtb_bse_plan_buffer_bytes(self.h, code)
        raw = np.zeros(nbytes, dtype=np.uint8)
        _check(lib().wtb_bse_plan_fetch(self.h, code, _p(raw), nbytes))
        p, nk, N = self.params, self.params.nk, self.N
        nb = p.nc + p.nv
        if what == "kpt":
            return raw.view(np.float64).reshape(3, nk).T.copy()
        if what == "eall":
            return raw.view(np.float64).reshape(nk, -1)
        if what == "nocp":
            return raw.view(np.int32).copy()
        if what == "eigv":
            return raw.view(np.float64).reshape(nk, nb)
        if what == "vector":
            return raw.view(np.complex128).reshape(nk, nb, -1)
        if what == "dipole":
            return raw.view(np.complex128).reshape(3, N)
        if what == "hbse":                                       # column-major N x N (lower valid)
            dt = np.complex64 if p.precision == MIXED else np.complex128
            return raw.view(dt).reshape(N, N).T
        if what == "wtab":
            return raw.view(np.float64).reshape(2 * p.ngrid[0] - 1, 2 * p.ngrid[1] - 1, 2 * p.ngrid[2] - 1)
        if what == "tridiag":
            return raw.view(np.float64).reshape(2, N)
        if what == "excwf":
            return raw.view(np.complex128).reshape(-1, N).T            # (N, count)
        if what == "fdeh":
            return raw.view(np.float64).copy()
        return raw

    def lowest(self, nev=256, guard=32, max_sub=None, max_iter=60, tol=1e-6, lanczos_steps=250,
               fallback=True, mlp=None, vectors=False, coarse=None):
        """Iterative form (precision-form plan): the lowest `nev` excitons by block
        Davidson and eps(omega) by Lanczos chains.  `coarse`: states of lowest diagonal
        energy whose block the preconditioner inverts exactly (default 10 (nev + guard),
        at most N/4; 0 for the diagonal preconditioner alone).  Returns dict W, resid,
        eps (or None), X (N, nev) (or None), info."""
        N = self.N
        if coarse is None:
            coarse = min(10 * (nev + guard), N // 4)
        dp = _DavidsonParams(nev, guard, max_sub or 5 * (nev + guard), max_iter, tol, lanczos_steps,
                             int(fallback), int(coarse))
        W = np.zeros(nev)
        res = np.zeros(nev)
        eps = np.zeros((6, 2, int(self.params.numbse))) if lanczos_steps > 0 else None
        X = np.zeros((nev, N), dtype=np.complex128) if vectors else None
        info = _DavidsonInfo()
        _check(lib().wtb_bse_plan_lowest(self.h, C.byref(dp), mlp.h if mlp is not None else None, _p(W),
                                         _p(res), _p(eps) if eps is not None else None,
                                         _p(X) if X is not None else None, C.byref(info)))
        inf = {k: getattr(info, k) for k, _ in _DavidsonInfo._fields_}
        return dict(W=W, resid=res, eps=eps, X=X.T if X is not None else None, info=inf)

    def set_vectors(self, first, count):
        """Also form the eigenvectors of excitons first .. first+count-1 (fetch('excwf'))."""
        _check(lib().wtb_bse_plan_set_vectors(self.h, int(first), int(count)))

    def set_reduction(self, two_stage=None, band=32):
        """Tridiagonal reduction: two-stage (True: dense -> band -> tridiagonal), one-stage
        Householder (False), or by size (None, the default: two-stage from N = 8192 on);
        band width 16 or 32."""
        mode = -1 if two_stage is None else int(bool(two_stage))
        _check(lib().wtb_bse_plan_set_reduction(self.h, mode, int(band)))

    def set_refine(self, nref=256):
        """Precision form: FP64 Rayleigh-quotient refinement of the nref lowest excitons
        in every execute (0: off)."""
        _check(lib().wtb_bse_plan_set_refine(self.h, int(nref)))
        self.nref = int(nref)

    def refine_results(self):
        """(residual norms, energies before the refinement) of the refined excitons."""
        n = getattr(self, "nref", 0)
        res = np.zeros(n)
        w0 = np.zeros(n)
        _check(lib().wtb_bse_plan_refine_results(self.h, n, _p(res), _p(w0)))
        return res, w0

    def set_panel(self, nb):
        """Panel width of the blocked reduction (8..32 columns per trailing update)."""
        _check(lib().wtb_bse_plan_set_panel(self.h, int(nb)))

    def tune(self):
        """Time the reduction for panel widths 16, 20, 24, 32 on this matrix and keep the
        fastest.  Returns (best, {width: ms})."""
        best = np.zeros(1, dtype=np.int32)
        ms = np.zeros(4)
        _check(lib().wtb_bse_plan_tune(self.h, _p(best), _p(ms)))
        return int(best[0]), dict(zip((16, 20, 24, 32), ms))

    def row_stats(self):
        """Diagonal D (N,) and row statistics (3, N) of the matrix (precision-form plan)."""
        D = np.zeros(self.N)
        st = np.zeros((3, self.N), dtype=np.float32)
        _check(lib().wtb_bse_plan_row_stats(self.h, _p(D), _p(st)))
        return D, st

    def close(self):
        if self.h:
            lib().wtb_bse_plan_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class BSEBatch:
    """Batching form: many independent models, one plan each, on `lanes` streams."""

    def __init__(self, ctx: Context, models, params, lanes: int = 4):
        self.ctx = ctx
        self.params = list(params)
        self._cms = [_CModel(m) for m in models]
        cm_arr = (_Model * len(models))(*[cm.s for cm in self._cms])
        cp_arr = (_Params * len(models))(*[p.to_c() for p in self.params])
        self._arrs = (cm_arr, cp_arr)
        h = C.c_void_p()
        _check(lib().wtb_bse_batch_create(ctx.h, len(models), C.cast(cm_arr, C.c_void_p),
                                          C.cast(cp_arr, C.c_void_p), lanes, C.byref(h)))
        self.h = h
        self.dims = [lib().wtb_bse_batch_dim(h, i) for i in range(len(models))]

    def tune(self):
        t = np.zeros(len(self.dims), dtype=np.int32)
        _check(lib().wtb_bse_batch_tune(self.h, _p(t)))
        return t

    def execute(self):
        _check(lib().wtb_bse_batch_execute(self.h))

    def results(self, i):
        N, nb = self.dims[i], int(self.params[i].numbse)
        W = np.zeros(N)
        act = np.zeros((8, N))
        eps = np.zeros((6, 2, nb))
        egap = np.zeros(1)
        _check(lib().wtb_bse_batch_results(self.h, i, _p(W), _p(act), _p(eps), _p(egap)))
        return dict(W=W, act=act, eps=eps, egap=float(egap[0]))

    def close(self):
        if self.h:
            lib().wtb_bse_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class MLP:
    """The learned preconditioner network on the GPU (5 -> 128 -> 128 -> 1, tanh, FP32)."""

    def __init__(self, ctx: Context, hidden: int = 128):
        h = C.c_void_p()
        _check(lib().wtb_mlp_create(ctx.h, hidden, C.byref(h)))
        self.h = h
        self.nparams = lib().wtb_mlp_nparams(h)

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().wtb_mlp_set_params(self.h, _p(p), p.size))

    def get_params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().wtb_mlp_get_params(self.h, _p(p), p.size))
        return p

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)     # (5, S)
        y = np.zeros(X.shape[1], dtype=np.float32)
        _check(lib().wtb_mlp_forward(self.h, X.shape[1], _p(X), _p(y)))
        return y

    def loss_grad(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(lib().wtb_mlp_loss_grad(self.h, X.shape[1], _p(X), _p(t), _p(loss), _p(g)))
        return float(loss[0]), g

    def train(self, X, t, epochs=20, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        _check(lib().wtb_mlp_train(self.h, X.shape[1], _p(X), _p(t), epochs, batch, C.c_float(lr),
                                   C.c_uint(seed), _p(hist)))
        return hist

    def close(self):
        if self.h:
            lib().wtb_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def kpathbse(rlat, ks, npts):
    """kpathbse (mhkpack_subs.f90:132-208): (qauxv (nq, 4), labels (nks/2 + 1,))."""
    rl = np.asfortranarray(np.asarray(rlat, dtype=np.float64))
    ksf = np.asfortranarray(np.asarray(ks, dtype=np.float64))
    nks = ksf.shape[0]
    nq = (nks // 2) * int(npts)
    q = np.zeros((nq, 4), order="F")
    lab = np.zeros(nks // 2 + 1)
    _check(lib().wtb_kpathbse(rl.ctypes.data, nks, ksf.ctypes.data, int(npts), q.ctypes.data, _p(lab)))
    return np.ascontiguousarray(q), lab


def bsebnds(ctx: Context, model: TBModel, params: BSEParams, qpath, lanes: int = 4):
    """Exciton bands (bsebnds, bse_kpath-tool.f90): exk (N, nq) at the q-path rows
    qpath (nq, 4) = (coordinate, qx, qy, qz)."""
    cm = _CModel(model)
    cp = params.to_c()
    qf = np.asfortranarray(np.asarray(qpath, dtype=np.float64))
    nq = qf.shape[0]
    exk = np.zeros((params.dim, nq), order="F")
    _check(lib().wtb_bsebnds(ctx.h, C.byref(cm.s), C.byref(cp), nq, qf.ctypes.data, int(lanes), exk.ctypes.data))
    return np.ascontiguousarray(exk)


class BSEBands:
    """Exciton bands with the plans kept (wtb_bse_bands_*): run(qpath) -> exk (N, nq)."""

    def __init__(self, ctx: Context, model: TBModel, params: BSEParams, lanes: int = 4):
        self.ctx = ctx
        self.params = params
        self._cm = _CModel(model)
        self._cp = params.to_c()
        h = C.c_void_p()
        _check(lib().wtb_bse_bands_create(ctx.h, C.byref(self._cm.s), C.byref(self._cp), int(lanes), C.byref(h)))
        self.h = h

    def kernel_flops(self):
        """Algorithmic flops of the timed two-stage trailing updates (0: one-stage GEMV)."""
        f = np.zeros(1)
        _check(lib().wtb_bse_plan_kernel_flops(self.h, _p(f)))
        return float(f[0])

    def set_graph(self, on=True):
        """False: launch the kernels directly instead of replaying the captured graphs."""
        _check(lib().wtb_bse_bands_set_graph(self.h, int(on)))

    def set_vectors(self, first, count):
        """Also form the vectors of excitons first .. first+count-1 at every q (run_wf)."""
        _check(lib().wtb_bse_bands_set_vectors(self.h, int(first), int(count)))
        self.wf_count = int(count)

    def run(self, qpath):
        qf = np.asfortranarray(np.asarray(qpath, dtype=np.float64))
        exk = np.zeros((self.params.dim, qf.shape[0]), order="F")
        _check(lib().wtb_bse_bands_run(self.h, qf.shape[0], qf.ctypes.data, exk.ctypes.data))
        return np.ascontiguousarray(exk)

    def run_wf(self, qpath):
        """(exk (N, nq), wf (nq, N, count)): the energies and the exciton vectors."""
        qf = np.asfortranarray(np.asarray(qpath, dtype=np.float64))
        nq, N = qf.shape[0], self.params.dim
        exk = np.zeros((N, nq), order="F")
        wf = np.zeros((nq, self.wf_count, N), dtype=np.complex128)
        _check(lib().wtb_bse_bands_run_wf(self.h, nq, qf.ctypes.data, exk.ctypes.data, wf.ctypes.data))
        return np.ascontiguousarray(exk), np.ascontiguousarray(wf.transpose(0, 2, 1))

    def close(self):
        if self.h:
            lib().wtb_bse_bands_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def bsesolver(ctx: Context, model: TBModel, params: BSEParams):
    """The whole BSE pipeline (bsesolver + the bsedielraw spectrum) on the GPU.
    Returns W (N,), act (8,N) [xx,yy,zz,xy,xz,yz,sp,sm], eps (6,2,numbse), egap."""
    with BSEPlan(ctx, model, params) as plan:
        plan.execute()
        return plan.results()
