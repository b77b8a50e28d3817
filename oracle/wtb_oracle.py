"""CPU oracle for the WanTiBEXOS BSE path — TEST INFRASTRUCTURE ONLY.

This module restates, in numpy, the reference algorithm of the BSE pipeline
(bsesolver + bsedielraw) function by function, each citing the reference
file:line it follows.  It is the checker for the HIP product path: only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and never as the thing measured or shipped.

Float32 quirks of the reference are reproduced on purpose:
  * phases k.R and Cartesian R components pass through COMPLEX(4)
    (``cmplx(0.0, x)`` has default kind), hamiltonian_tb.f90:64-66,
    optics.f90:39-47;
  * pi = acos(-1.) in single precision, mhkpack_subs.f90:413;
  * Coulomb constant -(0.0904756)*10**3 evaluated in single precision,
    coulomb_pot.f90:6 (and every potential);
  * cmplx(-efermi,0.0), cmplx(0.0,sme), 1.0/sqrt(2.), gama = 180.90,
    factor = 0.15 are single precision (hamiltonian_tb.f90:383,
    optics.f90:139,188,335, bse_solver-tool-diel.f90:376-377).
  * optsp stores 1/(Ec-Ev+i*sme) into a REAL variable, so only its real part
    multiplies the dipole (optics.f90:130,149).

Pinning (how each function is checked; see tests/test_oracle_*.py):
  k grid, stt, Delta tables     bit-exact vs the compiled reference dump
  bands / vectors (gauge)       vs reference dump (MKL ZHEEV), <=1e-12 eV
  dipoles                       vs reference dump, <=1e-10 relative
  BSE matrix                    vs reference dump (full upper triangle, small N)
  Coulomb potentials            vs reference dump (matrix rows) + closed forms
  BSE eigenvalues/activities    vs reference dump; sum rule; Hermiticity
  eps(omega), imagrenorm        vs reference dump (F15.4 path and full precision)
  finite T (tcor, f_v - f_c)    vs reference dump; closed forms; activity sum rule
  rkmesh / rkmesh2D             vs the reference's routines on 162 cases; closed forms
  optprop, exciton lifetimes    vs the reference program's files; closed forms
  finite-Q bands (kpathbse,     vs the reference program's bands_bse.dat / KLABELS-BSE.dat;
    k+q grid, exchange)         q = 0, q = G and rank-one interlacing invariants
The per-k and BSE eigensolves use LAPACK through numpy (zheevd) in place of
MKL ZHEEV; eigenvalues are basis independent and the per-k eigenvectors are
put into ZHEEV('L')'s gauge (component on orbital 1 real), which is all the
reference's outputs depend on.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

F32 = np.float32


def fl32(x):
    """Round a double to float32 and back (what a default-kind REAL/COMPLEX does)."""
    return np.asarray(x, dtype=np.float64).astype(np.float32).astype(np.float64)


PI32 = float(F32(math.pi))                                  # acos(-1.)  mhkpack_subs.f90:413
CIC = float(F32(-F32(0.0904756) * F32(1000.0)))             # coulomb_pot.f90:6
GAMA = float(F32(180.90))                                   # optics.f90:188
RENORM_FACTOR = float(F32(0.15))                            # optics.f90:335
INV_SQRT2_F32 = float(F32(1.0) / np.sqrt(F32(2.0)))         # bse_solver-tool-diel.f90:376


# ---------------------------------------------------------------------------
# geometry and k grid
# ---------------------------------------------------------------------------

def prodvec(v1, v2):
    """general_subs.f90:183-193"""
    return np.array([(v1[1] * v2[2]) - (v1[2] * v2[1]),
                     (v1[2] * v2[0]) - (v1[0] * v2[2]),
                     (v1[0] * v2[1]) - (v1[1] * v2[0])])


def vecsize(v):
    """general_subs.f90:286-294"""
    return math.sqrt((v[0] ** 2) + (v[1] ** 2) + (v[2] ** 2))


def vcell2D(rlat):
    """general_subs.f90:195-206: |a1 x a2|"""
    return vecsize(prodvec(rlat[0], rlat[1]))


def vcell3D(rlat):
    """general_subs.f90:208-221: |(a1 x a2).a3|"""
    vx = prodvec(rlat[0], rlat[1])
    return abs(vx[0] * rlat[2][0] + vx[1] * rlat[2][1] + vx[2] * rlat[2][2])


def alat2D(rlat):
    """general_subs.f90:248-260"""
    return 0.5 * (vecsize(rlat[0]) + vecsize(rlat[1]))


def recvec(rlat):
    """mhkpack_subs.f90:409-432 (single-precision pi)"""
    a1, a2, a3 = (np.asarray(r, dtype=np.float64) for r in rlat)
    v23, v31, v12 = prodvec(a2, a3), prodvec(a3, a1), prodvec(a1, a2)
    vol = abs((a1[0] * v23[0]) + (a1[1] * v23[1]) + (a1[2] * v23[2]))
    f = (2.0 * PI32) / vol
    return np.array([f * v23, f * v31, f * v12])


def monkhorst_pack(ngrid, shift, rlat):
    """mhkpack_subs.f90:359-407 -> kpt (nk,3); index = i*n2*n3 + j*n3 + l (l fastest)."""
    n1, n2, n3 = (int(v) for v in ngrid)
    b = recvec(rlat)
    sh = [float(v) for v in shift]
    kshift = [b[0][c] * sh[0] + b[1][c] * sh[1] + b[2][c] * sh[2] for c in range(3)]
    s1 = [b[0][c] / float(n1) for c in range(3)]
    s2 = [b[1][c] / float(n2) for c in range(3)]
    s3 = [b[2][c] / float(n3) for c in range(3)]
    out = np.empty((n1 * n2 * n3, 3))
    cnt = 0
    for i in range(n1):
        for j in range(n2):
            for l in range(n3):
                for c in range(3):
                    out[cnt, c] = s1[c] * float(i) + s2[c] * float(j) + s3[c] * float(l) + kshift[c]
                cnt += 1
    return out


def rkmesh(rk, rlat):
    """mhkpack_subs.f90:1-26 (MESH_TYPE= RK3D): n_i = int(max(1, rk |b_i| / (2 pi) + 1/2))
    with the reference's single-precision pi in recvec and in the division."""
    b = recvec(rlat)
    pi = PI32  # acos(-1.0) single precision, used in double arithmetic
    vs = [vecsize(b[i]) / (2 * pi) for i in range(3)]
    return [int(max(1.0, rk * vs[i] + 0.5)) for i in range(3)]


def rkmesh2D(rk, rlat):
    """mhkpack_subs.f90:28-52 (MESH_TYPE= RK2D): rkmesh's first two axes, n_3 = 1."""
    g = rkmesh(rk, rlat)
    return [g[0], g[1], 1]


# ---------------------------------------------------------------------------
# stages (a)+(b): H(k), per-k eigensolve
# ---------------------------------------------------------------------------

def phases(kpt, rvec, rlat):
    """fl32(k.R) of hamiltonian_tb.f90:62-68, shape (nk, nvec)."""
    r = np.asarray(rvec, dtype=np.float64)
    L = np.asarray(rlat, dtype=np.float64)
    rx = (r[:, 0] * L[0, 0] + r[:, 1] * L[1, 0]) + r[:, 2] * L[2, 0]
    ry = (r[:, 0] * L[0, 1] + r[:, 1] * L[1, 1]) + r[:, 2] * L[2, 1]
    rz = (r[:, 0] * L[0, 2] + r[:, 1] * L[1, 2]) + r[:, 2] * L[2, 2]
    k = np.asarray(kpt, dtype=np.float64)
    x = (k[:, 0:1] * rx[None, :] + k[:, 1:2] * ry[None, :]) + k[:, 2:3] * rz[None, :]
    return fl32(x)


def cart_r(rvec, rlat):
    """fl32 Cartesian R components of optics.f90:41-45, shape (nvec, 3)."""
    r = np.asarray(rvec, dtype=np.float64)
    L = np.asarray(rlat, dtype=np.float64)
    out = np.empty((len(r), 3))
    for a in range(3):
        out[:, a] = (r[:, 0] * L[0, a] + r[:, 1] * L[1, a]) + r[:, 2] * L[2, a]
    return fl32(out)


def hk_stack(kpt, model):
    """H(k) = sum_R e^{i fl32(k.R)} H(R)/ndeg  (hamiltonian_tb.f90:71-77), (nk,nw,nw)."""
    ph = np.exp(1j * phases(kpt, model.rvec, model.rlat))
    hr = (model.hop + 1j * model.ihop) * (1.0 / model.ndeg.astype(np.float64))[:, None, None]
    return np.einsum("kr,rmn->kmn", ph, hr)


def dhk_stack(kpt, model):
    """dH/dk_alpha = sum_R i fl32(R_alpha) e^{i fl32(k.R)} H(R)/ndeg (optics.f90:1-97),
    shape (3, nk, nw, nw).  The R=0 diagonal is zeroed as in optics.f90:61-66
    (its weight fl32(R_alpha) is 0 anyway)."""
    ph = np.exp(1j * phases(kpt, model.rvec, model.rlat))
    rc = cart_r(model.rvec, model.rlat)
    hr = (model.hop + 1j * model.ihop) * (1.0 / model.ndeg.astype(np.float64))[:, None, None]
    zero = np.where((model.rvec == 0).all(axis=1))[0]
    for i0 in zero:
        hr[i0] = hr[i0] - np.diag(np.diag(hr[i0]))
    return np.stack([np.einsum("kr,rmn->kmn", ph * (1j * rc[None, :, a]), hr) for a in range(3)])


def gauge_fix(U):
    """Put eigenvector columns in ZHEEV('L')'s gauge: component 1 real and >= 0."""
    x1 = U[..., 0:1, :]
    mag = np.abs(x1)
    ph = np.where(mag > 0, np.conj(x1) / np.where(mag > 0, mag, 1.0), 1.0)
    return U * ph


def eigsys(kpt, model):
    """hamiltonian_tb.f90:2-137 for all k: returns (E (nk,nw), U (nk,nw,nw) columns =
    eigenvectors in ZHEEV('L') gauge, nocp (nk,))."""
    H = hk_stack(kpt, model)
    ef = float(fl32(-model.efermi))                      # cmplx(-efermi,0.0), :383
    H = H + ef * np.eye(model.nw)[None]
    E, U = np.linalg.eigh(H, UPLO="L")
    U = gauge_fix(U)
    pos = E > 0.0
    if not pos.any(axis=1).all():
        raise ValueError("eigsys: no band above E_F at some k (nocp undefined in the reference)")
    nocp = np.argmax(pos, axis=1)                        # first E>0 (0-based) = nocp
    E = E.copy()
    for k in range(len(E)):
        E[k, nocp[k]:] += model.scs                      # :129-133
    return E, U, nocp.astype(np.int32)


def select_bands(E, U, nocp, nc, nv):
    """bse_solver-tool-diel.f90:262-303: eigv (nk, nv+nc), vector (nk, nv+nc, nw)."""
    nk = len(E)
    idx = nocp[:, None] - nv + np.arange(nc + nv)[None, :]
    if (idx < 0).any() or (idx >= E.shape[1]).any():
        raise ValueError("quantumnumbers2: band window outside 1..nw")
    eigv = np.take_along_axis(E, idx, axis=1)
    vector = np.stack([U[k][:, idx[k]].T for k in range(nk)])
    return eigv, vector


def quantumnumbers2(nk, nc, nv, nocp, nw):
    """general_subs.f90:501-567: stt (N,4) 1-based (I, v, c, k); k outer, c, v inner."""
    if (nv > nocp).any():
        raise ValueError("Number of valence states higher then ocuppied states")
    if (nc > (nw - nocp)).any():
        raise ValueError("Number of conduction states higher then unocuppied states")
    k, c, v = np.meshgrid(np.arange(1, nk + 1), np.arange(nv + 1, nv + nc + 1),
                          np.arange(1, nv + 1), indexing="ij")
    n = nk * nc * nv
    stt = np.empty((n, 4), dtype=np.int32)
    stt[:, 0] = np.arange(1, n + 1)
    stt[:, 1] = v.ravel()
    stt[:, 2] = c.ravel()
    stt[:, 3] = k.ravel()
    return stt


def dipoles(kpt, model, eigv, vector, stt, sme):
    """optsp/hlm/sandwich per transition (optics.f90:102-156, general_subs.f90:69-86),
    driver bse_solver-tool-diel.f90:365-377.  Returns (5, N): hrx, hry, hrz, hrsp, hrsm."""
    dH = dhk_stack(kpt, model)                           # (3, nk, nw, nw)
    ki = stt[:, 3] - 1
    vi = stt[:, 1] - 1
    ci = stt[:, 2] - 1
    vv = vector[ki, vi]                                  # (N, nw)
    vc = vector[ki, ci]
    ev = eigv[ki, vi]
    ec = eigv[ki, ci]
    esp = (ec - ev) + 1j * float(fl32(sme))
    factor = (1.0 / esp).real                            # REAL factor = 1./esp, optics.f90:130,149
    out = np.empty((5, len(stt)), dtype=np.complex128)
    for a in range(3):
        aux = np.einsum("nmj,nj->nm", dH[a][ki], vv)     # matvec
        out[a] = factor * np.einsum("nm,nm->n", np.conj(vc), aux)
    out[3] = (out[0] + 1j * out[1]) * INV_SQRT2_F32
    out[4] = (out[0] - 1j * out[1]) * INV_SQRT2_F32
    return out


# ---------------------------------------------------------------------------
# Coulomb potentials (coulomb_pot.f90), V(k1,k2) for arrays of k-differences
# ---------------------------------------------------------------------------

def _modk(dk):
    return np.sqrt((dk[..., 0] ** 2 + dk[..., 1] ** 2) + dk[..., 2] ** 2)


def coulomb(coultype, dk, ngrid, rlat, ediel, lc, ez, w, r0, tolr):
    """V for k1-k2 = dk (..., 3), selecting the potential as matrizelbse does
    (bse_subs.f90:45-91)."""
    dk = np.asarray(dk, dtype=np.float64)
    nkt = int(ngrid[0]) * int(ngrid[1]) * int(ngrid[2])
    modk = _modk(dk)
    small = modk < tolr
    safe = np.where(small, 1.0, modk)
    ct = coultype.strip().upper()
    if ct == "V2DRK":                                    # coulomb_pot.f90:276-340
        vc = vcell2D(rlat)
        vbz = 1.0 / (nkt * vc)
        epar, et, eb = float(ediel[1]), float(ediel[0]), float(ediel[2])
        eta = math.sqrt(epar / ez)
        kappa = math.sqrt(epar * ez)
        pb = (eb - kappa) / (eb + kappa)
        pt = (et - kappa) / (et + kappa)
        aux1 = (1.0 - (pb * pt * np.exp(-2.0 * safe * eta * lc))) * kappa
        aux2 = (1.0 - (pt * np.exp(-eta * safe * lc))) * (1.0 - (pb * np.exp(-eta * safe * lc)))
        aux3 = r0 * safe * np.exp(-safe * w)
        ew = (aux1 / aux2) + aux3
        v = (vbz * CIC) * np.exp(-safe * w) * (1.0 / ew) * (1.0 / safe)
        return np.where(small, 0.0, v)
    if ct in ("V3D", "V3DL"):                            # coulomb_pot.f90:55-122
        vc = vcell3D(rlat)
        ed = 1.0 if ct == "V3D" else float(ediel[1])
        vbz = 1.0 / (nkt * vc)
        v = vbz * (CIC / ed) * (1.0 / (safe * safe))
        return np.where(small, 0.0, v)
    if ct == "V2DK":                                     # coulomb_pot.f90:2-51
        a0 = alat2D(rlat)
        vc = vcell2D(rlat)
        rr0 = ((ediel[1] - 1.0) * lc) / (ediel[0] + ediel[2])
        vbz = 1.0 / (nkt * vc)
        g1 = float(int(ngrid[0]) * int(ngrid[1]))
        auxi = (2.0 * PI32 * rr0) / (a0 * math.sqrt(g1))
        ed = (ediel[0] + ediel[2]) / 2.0
        alpha1, alpha2, alpha3 = float(F32(1.76)), 1.0, 0.0
        v0 = vbz * (CIC / ed) * (a0 * math.sqrt(g1) / (2.0 * PI32)) * (alpha1 + auxi * alpha2 + alpha3 * auxi ** 2)
        v = vbz * (CIC / ed) * (1.0 / (safe * (1 + (rr0 * safe))))
        return np.where(small, v0, v)
    if ct == "V2DT":                                     # coulomb_pot.f90:126-183
        vc = vcell3D(rlat)
        vbz = 1.0 / (nkt * vc)
        gz = np.abs(dk[..., 2])
        gpar = np.sqrt((dk[..., 0] * dk[..., 0]) + (dk[..., 1] * dk[..., 1]))
        rc = 0.5 * rlat[2][2]
        c0 = (vbz * CIC) * (-0.5 * rc * rc)
        c1 = (vbz * CIC) * (1.0 / (safe * safe)) * (1.0 - np.cos(gz * rc) - (gz * rc * np.sin(gz * rc)))
        gps = np.where(gpar < tolr, 1.0, gpar)
        aux1 = gz / gps
        aux2 = gps * rc
        aux3 = gz * rc
        c2 = (vbz * CIC) * (1.0 / (safe * safe)) * (1.0 + (np.exp(-aux2) * (aux1 * np.sin(aux3) - np.cos(aux3))))
        return np.where((gpar < tolr) & (gz < tolr), c0, np.where(gpar < tolr, c1, c2))
    if ct == "V2DT2":                                    # coulomb_pot.f90:233-272
        vc = vcell3D(rlat)
        vbz = 1.0 / (nkt * vc)
        qxy = np.sqrt((dk[..., 0] * dk[..., 0]) + (dk[..., 1] * dk[..., 1]))
        v = (vbz * CIC) * (1.0 / (safe * safe)) * (1.0 - np.exp(-0.5 * qxy * lc) * np.cos(0.5 * lc * dk[..., 2]))
        return np.where(small, 0.0, v)
    if ct == "V0DT":                                     # coulomb_pot.f90:187-228
        vc = vcell3D(rlat)
        vbz = 1.0 / (nkt * vc)
        cr = 0.5 * min(vecsize(rlat[0]), vecsize(rlat[1]), vecsize(rlat[2]))
        v0 = (vbz * CIC) * 0.5 * cr * cr
        v = (vbz * CIC) * (1.0 / (safe * safe)) * (1.0 - np.cos(cr * safe))
        return np.where(small, v0, v)
    if ct == "V2DOH":                                    # coulomb_pot.f90:344-377
        vc = vcell2D(rlat)
        vbz = 1.0 / (nkt * vc)
        v = (vbz * CIC) * np.exp(-w * safe) * (1.0 / (ez * safe))
        return np.where(small, 0.0, v)
    if ct == "V1DT":                                     # coulomb_pot.f90:381-432
        from scipy.special import j0, j1, k0, k1
        vc = vcell3D(rlat)
        vbz = 1.0 / (nkt * vc)
        ax = np.abs(dk[..., 0])
        qyz = np.sqrt((dk[..., 1] * dk[..., 1]) + (dk[..., 2] * dk[..., 2]))
        out = np.empty(np.shape(modk))
        a_small = ax < tolr
        q_small = qyz < tolr
        out = np.where(a_small & q_small, (vbz * CIC) * ((lc * lc) / 16.0) * (2.0 * math.log(lc) - 1.0), 0.0)
        for idx in zip(*np.nonzero(a_small & ~q_small)):
            out[idx] = (vbz * CIC) * integral_j0pot(lc, qyz[idx])
        gen = ~a_small
        axs = np.where(gen, ax, 1.0)
        c1 = qyz * lc * j1(qyz * lc) * k0(axs * lc)
        c2 = axs * lc * j0(qyz * lc) * k1(axs * lc)
        out = np.where(gen, (vbz * CIC) * (1.0 + c1 + c2), out)
        return out
    raise ValueError("Wrong Coulomb Potential")


def integral_j0pot(rmax, gperp):
    """coulomb_pot.f90:437-469: trapezoid of r J0(g r) log r on 100 points."""
    from scipy.special import j0
    n = 100
    eta = float(F32(1e-4))
    r = ((np.arange(1, n + 1) - 1.0) / (n - 1.0)) * rmax + eta
    f = r * j0(r * gperp) * np.log(r)
    h = r[1] - r[0]
    res = 0.0
    for i in range(n - 1):
        res = res + (h / 2.0) * (f[i] + f[i + 1])
    return res


# ---------------------------------------------------------------------------
# stage (c): BSE matrix
# ---------------------------------------------------------------------------

def bse_matrix(kpt, eigv, vector, stt, params, ngrid, rlat, fdeh=None):
    """Full Hermitian BSE matrix whose upper triangle is matrizelbse's
    (bse_subs.f90:2-124, loop bse_solver-tool-diel.f90:431-458):
      H_II = (Ec-Ev) + V(0);  H_IJ = V(k_I-k_J) <c_I|c_J> <v_I|v_J>  (J>I).
    With fdeh (finite temperature, matrizelbsetemp, bse_subs_temp.f90:2-126 in the loop
    j >= i of bse_solver-tool-diel-temp.f90:470-480): the kernel terms times the ROW
    state's f_v - f_c, H_II = (Ec-Ev) + V(0) f_I; ZHEEV('U') reads the upper triangle only,
    so the lower one is its conjugate."""
    ki = stt[:, 3] - 1
    vi = stt[:, 1] - 1
    ci = stt[:, 2] - 1
    vc = vector[ki, ci]                                  # (N, nw)
    vv = vector[ki, vi]
    oc = np.conj(vc) @ vc.T                              # <c_I|c_J>
    ov = np.conj(vv) @ vv.T
    kI = kpt[ki]
    dk = kI[:, None, :] - kI[None, :, :]
    V = coulomb(params["coultype"], dk, ngrid, rlat, params["ediel"], params["lc"],
                params["ez"], params["w"], params["r0"], params["ktol"])
    H = V * oc * ov
    if fdeh is not None:
        H = H * fdeh[:, None]
    H = np.triu(H, 1)
    H = H + np.conj(H.T)
    vd = np.diag(V) if fdeh is None else np.diag(V) * fdeh
    diag = (eigv[ki, ci] - eigv[ki, vi]) + vd
    H[np.arange(len(H)), np.arange(len(H))] = diag
    return H


def delta_index(ngrid):
    """Integer k-k' table (Delta1, Delta2, Delta3) for the BSE basis order (k outer):
    row/col = k index, entries = (i-i', j-j', l-l'); the reference never folds
    k-k' (general_subs.f90:297-308)."""
    n1, n2, n3 = ngrid
    i, j, l = np.meshgrid(np.arange(n1), np.arange(n2), np.arange(n3), indexing="ij")
    g = np.stack([i.ravel(), j.ravel(), l.ravel()], axis=1)
    return g[:, None, :] - g[None, :, :]


# ---------------------------------------------------------------------------
# stage (d)+(e1): eigensolve and oscillator strengths
# ---------------------------------------------------------------------------

def bse_eig(H):
    """ZHEEV('V','U') (bse_solver-tool-diel.f90:485-488): ascending eigenpairs."""
    return np.linalg.eigh(H, UPLO="U")


def activities(X, hvec):
    """dielbsev/dielbsep (bse_subs.f90:278-378): f_n = |sum_j X_jn h_j|^2 (no conjugate
    on X) for xx,yy,zz,sp,sm and Re[(X^T h_a)(X^T h_b)^*] for xy,xz,yz.
    hvec = (5,N) [x,y,z,sp,sm]; returns (8,N) in the order xx,yy,zz,xy,xz,yz,sp,sm."""
    t = X.T @ hvec.T                                     # (N, 5)
    tx, ty, tz, tp, tm = (t[:, i] for i in range(5))
    sq = lambda a: (a * np.conj(a)).real
    cr = lambda a, b: (a * np.conj(b)).real
    return np.stack([sq(tx), sq(ty), sq(tz), cr(tx, ty), cr(tx, tz), cr(ty, tz), sq(tp), sq(tm)])


# ---------------------------------------------------------------------------
# finite temperature (bsesolvertemp, subprograms/bse_solver-tool-diel-temp.f90)
# ---------------------------------------------------------------------------

KB_F32 = float(fl32(8.617330350e-5))     # REAL literal into a double parameter, general_subs.f90:575,594,613
T0_F32 = float(fl32(1e-8))               # temp = 1E-8 when temp == 0, general_subs.f90:620-624


def gap_correction(ta, st, phavg, temp):
    """tcor added to the scissor (bse_solver-tool-diel-temp.f90:262-283): FA 0,
    VE gapcortemp (general_subs.f90:569-586), BE gapcortemp2 (:588-604)."""
    with np.errstate(divide="ignore", over="ignore"):
        if ta == "VE":
            aux1 = np.float64(phavg) / (2.0 * KB_F32 * np.float64(temp))
            aux2 = 1.0 / np.tanh(aux1)
            return float(-st * phavg * (aux2 - 1.0))
        if ta == "BE":
            aux2 = np.exp(np.float64(phavg) / np.float64(temp))
            return float(-(2.0 * st / (aux2 - 1.0)))
    return 0.0


def fermidisteh(ec, ev, temp):
    """f_v - f_c with E_F = 0 (general_subs.f90:606-635); temp 0 is taken as 1E-8."""
    t = temp if temp != 0.0 else T0_F32
    beta = 1.0 / (KB_F32 * t)
    with np.errstate(over="ignore"):
        fv = 1.0 / (np.exp(np.asarray(ev) * beta) + 1.0)
        fc = 1.0 / (np.exp(np.asarray(ec) * beta) + 1.0)
    return fv - fc


def activities_temp(X, hvec, fdeh):
    """dielbseptemp (bse_subs_temp.f90:340-417) for all eight components
    (bse_solver-tool-diel-temp.f90:570-618): Re[(X^T h_a) (X^T (h_b f))^*]."""
    t = X.T @ hvec.T                                     # (N, 5)
    tf = X.T @ (hvec * fdeh[None, :]).T
    cr = lambda a, b: (t[:, a] * np.conj(tf[:, b])).real
    return np.stack([cr(0, 0), cr(1, 1), cr(2, 2), cr(0, 1), cr(0, 2), cr(1, 2), cr(3, 3), cr(4, 4)])


def sp_table(kpt, nocp, eigv, h, stt, nv, fdeh=None):
    """The single-particle table vecres(N, 15) of bse_solver-tool-diel.f90:365-402 (REAL):
    k, nocp, band numbers, Ec - Ev, |h_a|^2 and Re(h_a conj h_b), |h_sp|^2, |h_sm|^2 (at
    finite temperature times f_v - f_c, bse_solver-tool-diel-temp.f90:405-424), rows in
    the stable ascending order of Bubblem on column 7 (general_subs.f90:405-440)."""
    ki = stt[:, 3] - 1
    f32 = np.float32
    cols = [kpt[ki, 0], kpt[ki, 1], kpt[ki, 2], nocp[ki], nocp[ki] - nv + stt[:, 2], nocp[ki] - nv + stt[:, 1],
            eigv[ki, stt[:, 2] - 1] - eigv[ki, stt[:, 1] - 1]]
    t = [c.astype(np.float64).astype(f32) for c in cols]
    hx, hy, hz, hp, hm = h
    q = [(hx * np.conj(hx)).real, (hy * np.conj(hy)).real, (hz * np.conj(hz)).real, (hx * np.conj(hy)).real,
         (hx * np.conj(hz)).real, (hy * np.conj(hz)).real, (hp * np.conj(hp)).real, (hm * np.conj(hm)).real]
    for v in q:
        r = v.astype(f32)
        t.append(r if fdeh is None else (r.astype(np.float64) * fdeh).astype(f32))
    tab = np.stack(t, axis=1)
    return tab[np.argsort(tab[:, 6], kind="stable")]


def f15_4(x):
    """Value after a write "(F15.4)" and list-directed read back
    (bse_solver-tool-diel.f90:610, bse_diel-tool.f90:117-123)."""
    x = np.asarray(x, dtype=np.float64)
    flat = [float(f"{v:.4f}") for v in x.ravel()]
    return np.array(flat, dtype=np.float64).reshape(x.shape)


# ---------------------------------------------------------------------------
# stage (e2): eps(omega)
# ---------------------------------------------------------------------------

def avgsme(E):
    """optics.f90:302-324: 10 * (sum of successive differences) / N."""
    s = 0.0
    for i in range(1, len(E)):
        s = s + E[i] - E[i - 1]
    return 10.0 * (s / float(len(E)))


def optdiel(rtype, vc, ngrid, elux, E, f, sme, smeavg=None):
    """optics.f90:160-244 for an array of photon energies elux (vectorised over elux)."""
    baux = float(F32(float(int(ngrid[0]) * int(ngrid[1]) * int(ngrid[2])))) * vc
    caux = GAMA / (baux * 1.0)
    delta = 1.0 if rtype == 1 else 0.0
    if smeavg is None:
        smeavg = avgsme(E)
    smef = smeavg + sme
    el = np.asarray(elux, dtype=np.float64)[:, None]
    eaux = ((el - E[None, :]) ** 2) + (smef ** 2)
    raux = f[None, :] * ((E[None, :] - el) / eaux)
    iaux = f[None, :] * (smef / eaux)
    rpart = raux.sum(axis=1)
    ipart = iaux.sum(axis=1)
    return delta + (rpart * caux), ipart * caux


def imagrenorm(E, omega, eps_im):
    """optics.f90:327-390 (in place on a copy; returns the new imaginary part)."""
    out = np.array(eps_im, dtype=np.float64, copy=True)
    aux = E[0] - RENORM_FACTOR
    if aux <= 0.0:
        return out
    ibegin = 0
    for i in range(len(omega)):
        if omega[i] <= aux:
            ibegin = i
        else:
            break
    if ibegin + 1 >= len(omega):
        # the reference reads one element past the grid here; the oracle zeroes all
        out[:] = 0.0
        return out
    aux3 = out[ibegin + 1]
    out[:ibegin + 1] = 0.0
    rest = out[ibegin + 1:] - aux3
    out[ibegin + 1:] = np.where(rest < 0.0, 0.0, rest)
    return out


def omega_grid(ebse0, ebsef, numbse):
    """bse_diel-tool.f90:137"""
    n = int(numbse)
    return np.array([ebse0 + ((ebsef - ebse0) * (j - 1)) / (numbse - 1.0) for j in range(1, n + 1)])


def spectrum(E, act6, rlat, ngrid, ebse0, ebsef, numbse, sme, renorm=True):
    """bsedielraw loop (bse_diel-tool.f90:111,135-201) without the final spinf/REAL
    output conversion.  act6 = (6,N) in order xx,yy,zz,xy,xz,yz.  Returns
    (omega, re (6,nw), im (6,nw)) with components in the same order."""
    vc = vcell3D(rlat)
    om = omega_grid(ebse0, ebsef, numbse)
    smeavg = avgsme(E)
    re = np.empty((6, len(om)))
    im = np.empty((6, len(om)))
    for c in range(6):
        r, i = optdiel(1 if c < 3 else 0, vc, ngrid, om, E, act6[c], sme, smeavg)
        re[c] = r
        im[c] = imagrenorm(E, om, i) if renorm else i
    return om, re, im


# ---------------------------------------------------------------------------
# the whole pipeline
# ---------------------------------------------------------------------------

DEFAULT_PARAMS = dict(nc=2, nv=2, coultype="V2DRK", ediel=(1.0, 1.0, 1.0), lc=1.0, ez=1.0,
                      w=0.0, r0=40.0, ktol=0.001, cshift=0.05, shift=(0.0, 0.0, 0.0),
                      ebse0=0.0, ebsef=6.0, numbse=1001, renorm=True, dtfull=True, f154=True,
                      bset=False, temp=0.0, ta="FA", st=0.0, phavg=0.0)


def run_pipeline(model, ngrid, **kw):
    """bsesolver (bse_solver-tool-diel.f90:223-620) + bsedielraw (bse_diel-tool.f90)."""
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    kpt = monkhorst_pack(ngrid, p["shift"], model.rlat)
    tcor = gap_correction(p["ta"], p["st"], p["phavg"], p["temp"]) if p["bset"] else 0.0
    if tcor != 0.0:                                       # eigsys(scs+tcor), ...-temp.f90:289
        model = dataclasses.replace(model, scs=model.scs + tcor)
    E, U, nocp = eigsys(kpt, model)
    eigv, vector = select_bands(E, U, nocp, p["nc"], p["nv"])
    stt = quantumnumbers2(len(kpt), p["nc"], p["nv"], nocp, model.nw)
    h = dipoles(kpt, model, eigv, vector, stt, p["cshift"])
    fdeh = None
    if p["bset"]:                                         # ...-temp.f90:394-399
        ki = stt[:, 3] - 1
        fdeh = fermidisteh(eigv[ki, stt[:, 2] - 1], eigv[ki, stt[:, 1] - 1], p["temp"])
    H = bse_matrix(kpt, eigv, vector, stt, p, ngrid, model.rlat, fdeh)
    W, X = bse_eig(H)
    act = activities(X, h) if fdeh is None else activities_temp(X, h, fdeh)
    if not p["dtfull"]:
        act[3:6] = 0.0
    Wq, actq = (f15_4(W), f15_4(act[:6])) if p["f154"] else (W, act[:6])
    om, re, im = spectrum(Wq, actq, model.rlat, ngrid, p["ebse0"], p["ebsef"], p["numbse"],
                          p["cshift"], p["renorm"])
    spinf = 2.0 if model.systype.strip() == "NP" else 1.0
    return dict(kpt=kpt, E=E, U=U, nocp=nocp, eigv=eigv, vector=vector, stt=stt, h=h, H=H, tcor=tcor, fdeh=fdeh,
                W=W, X=X, act=act, Wq=Wq, actq=actq, omega=om, eps_re=re, eps_im=im, spinf=spinf)


# ---------------------------------------------------------------------------
# finite-Q exciton bands (bsebnds, subprograms/bse_kpath-tool.f90)
# ---------------------------------------------------------------------------

def kpathbse(rlat, ks, npts):
    """kpathbse (mhkpack_subs.f90:132-208): the pairs (ks[2j], ks[2j+1]) (fractional) are
    segments; npts points per segment, ends included.  Returns qauxv ((nks/2) npts, 4):
    path coordinate (cumulative segment length over the total), then q (Cartesian), and
    the segment-end coordinates (KLABELS-BSE.dat)."""
    b = recvec(rlat)
    ks = np.asarray(ks, dtype=np.float64)
    nseg = len(ks) // 2
    ksaux = np.array([[ks[i, 0] * b[0][c] + ks[i, 1] * b[1][c] + ks[i, 2] * b[2][c] for c in range(3)]
                      for i in range(len(ks))])
    kdis = np.empty(nseg)
    for j in range(nseg):
        d = ksaux[2 * j + 1] - ksaux[2 * j]                  # distvec(ks(2j), ks(2j-1)), :292-302
        kdis[j] = math.sqrt((d[0] ** 2 + d[1] ** 2) + d[2] ** 2)
    tot = 0.0
    for j in range(nseg):
        tot = tot + kdis[j]
    kdis = kdis / tot
    out = np.empty((nseg * npts, 4))
    labels = [0.0]
    acc = 0.0
    for j in range(nseg):
        a, e = ksaux[2 * j], ksaux[2 * j + 1]
        for i in range(npts):
            t = float(i)                                      # (i-1.) with 1-based i: exact
            out[j * npts + i, 0] = acc + (kdis[j] * t) / float(npts - 1)
            for c in range(3):
                out[j * npts + i, 1 + c] = a[c] + ((e[c] - a[c]) * t) / float(npts - 1)
        acc = acc + kdis[j]
        labels.append(acc)
    return out, np.array(labels)


def monkhorst_pack_q(q, ngrid, shift, rlat):
    """monhkhorst_packq (mhkpack_subs.f90:305-356): the grid k + q, with q added first
    ((((q + b1 i/n1) + b2 j/n2) + b3 l/n3) + kshift), q = qauxv row (coordinate, qx, qy, qz)."""
    n1, n2, n3 = (int(v) for v in ngrid)
    b = recvec(rlat)
    sh = [float(v) for v in shift]
    kshift = [b[0][c] * sh[0] + b[1][c] * sh[1] + b[2][c] * sh[2] for c in range(3)]
    s1 = [b[0][c] / float(n1) for c in range(3)]
    s2 = [b[1][c] / float(n2) for c in range(3)]
    s3 = [b[2][c] / float(n3) for c in range(3)]
    out = np.empty((n1 * n2 * n3, 3))
    cnt = 0
    for i in range(n1):
        for j in range(n2):
            for l in range(n3):
                for c in range(3):
                    out[cnt, c] = q[1 + c] + s1[c] * float(i) + s2[c] * float(j) + s3[c] * float(l) + kshift[c]
                cnt += 1
    return out


def bse_matrix_q(kpt, eigv_k, vec_k, eigv_q, vec_q, stt, q, params, ngrid, rlat, fdeh=None):
    """matrizelbsekq (bse_subs_kpath.f90:1-214) in the loop of bse_kpath-tool.f90:325-345,
    j >= i: conduction bands at k+q, valence bands at k, V(k-k') on the k grid; for q != 0
    (modvecq, general_subs.f90:310-321) the exchange term
    - V(q) <c_i|v_i> <v_j|c_j> (c at k+q, v at k).  Returns the Hermitian completion of
    the upper triangle that ZHEEV('U') reads.  With fdeh (bsebndstemp, matrizelbsekqtemp,
    bse_subs_temp.f90:128-338): the kernel (direct - exchange) times the row state's
    f_v - f_c, diagonal (Ec - Ev) + (V(0) - exchange) f."""
    ki = stt[:, 3] - 1
    vi = stt[:, 1] - 1
    ci = stt[:, 2] - 1
    vc = vec_q[ki, ci]                                   # conduction at k+q (N, nw)
    vv = vec_k[ki, vi]                                   # valence at k
    oc = np.conj(vc) @ vc.T
    ov = np.conj(vv) @ vv.T
    kI = kpt[ki]
    dk = kI[:, None, :] - kI[None, :, :]
    p = params
    V = coulomb(p["coultype"], dk, ngrid, rlat, p["ediel"], p["lc"], p["ez"], p["w"], p["r0"], p["ktol"])
    H = V * oc * ov
    de = eigv_q[ki, ci] - eigv_k[ki, vi]
    kd = np.diag(V).astype(np.complex128)
    modq = math.sqrt((q[1] ** 2 + q[2] ** 2) + q[3] ** 2)
    if modq != 0.0:
        vq = float(coulomb(p["coultype"], np.asarray(q[1:4], dtype=np.float64).reshape(1, 3), ngrid, rlat,
                           p["ediel"], p["lc"], p["ez"], p["w"], p["r0"], p["ktol"])[0])
        a = np.einsum("nm,nm->n", np.conj(vc), vv)      # <c_i|v_i>
        X = (vq * a)[:, None] * np.conj(a)[None, :]
        H = H - X
        kd = kd - np.diag(X)
    if fdeh is not None:
        H = H * fdeh[:, None]
        diag = de + kd * fdeh
    else:
        diag = de + kd
    H = np.triu(H, 1)
    H = H + np.conj(H.T)
    H[np.arange(len(H)), np.arange(len(H))] = diag
    return H


def run_bands(model, ngrid, qpath, **kw):
    """bsebnds (bse_kpath-tool.f90:187-445): exk (N, nq), the ascending exciton energies
    of the matrix at every q of the path (ZHEEV('N','U')); with bset, bsebndstemp
    (bse_kpath-tool-temp.f90: gap correction on both grids, row factors f_v - f_c)."""
    p = dict(DEFAULT_PARAMS)
    p.update(kw)
    kpt = monkhorst_pack(ngrid, p["shift"], model.rlat)
    if p["bset"]:                                         # bsebndstemp: eigsys(scs + tcor) on both grids
        tcor = gap_correction(p["ta"], p["st"], p["phavg"], p["temp"])
        if tcor != 0.0:
            model = dataclasses.replace(model, scs=model.scs + tcor)
    E, U, nocp = eigsys(kpt, model)
    eigv_k, vec_k = select_bands(E, U, nocp, p["nc"], p["nv"])
    nq = len(qpath)
    exk = None
    mats = []
    for iq in range(nq):
        q = qpath[iq]
        kq = monkhorst_pack_q(q, ngrid, p["shift"], model.rlat)
        Eq, Uq, nocpq = eigsys(kq, model)
        eigv_q, vec_q = select_bands(Eq, Uq, nocpq, p["nc"], p["nv"])
        if (p["nv"] > nocp).any():                          # quantumnumbers2(..., nocpk, nocpq)
            raise ValueError("Number of valence states higher then ocuppied states")
        if (p["nc"] > model.nw - nocpq).any():
            raise ValueError("Number of conduction states higher then unocuppied states")
        stt = quantumnumbers2(len(kpt), p["nc"], p["nv"], np.full(len(kpt), p["nv"]), model.nw)
        fdeh = None
        if p["bset"]:                                     # f_v(k) - f_c(k+q) of the row state
            ki = stt[:, 3] - 1
            fdeh = fermidisteh(eigv_q[ki, stt[:, 2] - 1], eigv_k[ki, stt[:, 1] - 1], p["temp"])
        H = bse_matrix_q(kpt, eigv_k, vec_k, eigv_q, vec_q, stt, q, p, ngrid, model.rlat, fdeh)
        W = np.linalg.eigvalsh(H, UPLO="U")
        if exk is None:
            exk = np.empty((len(W), nq))
        exk[:, iq] = W
        mats.append(H)
    return dict(kpt=kpt, exk=exk, H=mats)


# ---------------------------------------------------------------------------
# post-processing of the spectra and exciton lifetimes
# ---------------------------------------------------------------------------
HC_F32 = float(F32(19.746e-6))        # diel-pp-subs.f90:57 (REAL literal)
SQRT2_F32 = float(np.sqrt(F32(2.0)))  # sqrt(2.0), diel-pp-subs.f90:60
CFE_F32 = float(F32(7.297e-3))        # bse_subs.f90:475
LSPD_F32 = float(F32(299.792e16))     # bse_subs.f90:476
HBAR_F32 = float(F32(6.582e-16))      # bse_subs.f90:477


def optprop(energy, e1, e2):
    """refracao, extincao, reflectibilidade, abscoef, enloss (diel-pp-subs.f90:1-73) as
    bseoptprop applies them to every photon energy and component (diel-pp-bse.f90:89-140).
    Returns (n, k, R, alpha, L) with the shapes of e1."""
    e1 = np.asarray(e1, dtype=np.float64)
    e2 = np.asarray(e2, dtype=np.float64)
    en = np.asarray(energy, dtype=np.float64).reshape((-1,) + (1,) * (e1.ndim - 1))
    aux = np.sqrt((e1 * e1) + (e2 * e2))
    refr = np.sqrt((aux + e1) / 2.0)
    ext = np.sqrt((aux - e1) / 2.0)
    refl = (((refr - 1.0) * (refr - 1.0)) + ext * ext) / (((refr + 1.0) * (refr + 1.0)) + ext * ext)
    absc = ((SQRT2_F32 * en) * (1.0 / HC_F32)) * np.sqrt(aux - e1)
    with np.errstate(divide="ignore", invalid="ignore"):
        eloss = e2 / ((e1 * e1) + (e2 * e2))
    return refr, ext, refl, absc, eloss


def exclft(sysdim, ngrid, rlat, fosc, enexc):
    """exclft (bse_subs.f90:472-530): radiative lifetime for SYSDIM "3D", "2D" or "1D"."""
    nkd = float(int(ngrid[0]) * int(ngrid[1]) * int(ngrid[2]))
    f = np.asarray(fosc, dtype=np.float64)
    E = np.asarray(enexc, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        if sysdim == "3D":
            aux1 = (4.0 * CFE_F32) * ((E * E) * E)
            aux2 = f / nkd
            aux3 = (3.0 * (LSPD_F32 * LSPD_F32)) * ((HBAR_F32 * HBAR_F32) * HBAR_F32)
            return 1.0 / ((aux1 * aux2) / aux3)
        if sysdim == "2D":
            vc = vcell2D(np.asarray(rlat, dtype=np.float64))
            return ((nkd * vc) * HBAR_F32) / ((((8.0 * PI32) * CFE_F32) * E) * f)
        if sysdim == "1D":
            aux2 = ((((2.0 * PI32) * CFE_F32) * E) * E) * f
            aux1 = (((nkd * float(rlat[0][0])) * HBAR_F32) * HBAR_F32) * LSPD_F32
            return aux1 / aux2
    raise ValueError("Wrong System dimension")


def lifetimes(sysdim, ngrid, rlat, table, nout):
    """lifetime / lifetimepol (exciton_lifetime.f90:2-187): table (N, 1 + ncomp) as read from
    bse_opt_diel(-pol).dat; returns (N, 1 + nout): the summed x+y+z strength's lifetime,
    then one per component."""
    t = np.asarray(table, dtype=np.float64)
    E = t[:, 0]
    fsum = (t[:, 1] + t[:, 2]) + t[:, 3]
    cols = [exclft(sysdim, ngrid, rlat, fsum, E)] + [exclft(sysdim, ngrid, rlat, t[:, c], E)
                                                     for c in range(1, nout + 1)]
    return np.stack(cols, axis=1)
