"""Pins the oracle's finite-Q exciton bands (bsebnds) to the reference.

The fixtures under tests/golden/bands/ hold what the unmodified reference program
wrote for BSE_BND= T runs (tests/golden/make_golden_bands.py): bands_bse.dat
(every exciton energy at every q of the path, printed to round-trip precision)
and KLABELS-BSE.dat.  Invariants the mathematics fixes cover what the fixtures
leave open: q = 0 is the Q = 0 BSE, q = a reciprocal lattice vector G gives the
q = 0 energies again (H(k+G) = H(k), so <c_k+G|v_k> = 0), and the exchange term is
a rank-one update -V(q) a a^H (V < 0 here), so the energies interlace with the
direct term's.  The host path generator of the C-ABI (wtb_kpathbse) is checked
against the same files.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import wtb_oracle as O
from wantibexos_amd.tbmodel import TBModel

BANDS = Path(__file__).resolve().parent / "golden" / "bands"
CASES = sorted(p.stem for p in BANDS.glob("*.npz"))
LIB = Path(__file__).resolve().parents[1] / "wantibexos_amd" / "libwtb_bse.so"


def load_bands(name):
    z = np.load(BANDS / f"{name}.npz")
    p = json.loads(str(z["params"]))
    p["ngrid"] = tuple(p["ngrid"])
    p["ediel"] = tuple(p["ediel"])
    p["shift"] = tuple(p["shift"])
    m = TBModel(str(z["model_systype"]), float(z["model_scs"]), float(z["model_efermi"]), z["model_rlat"],
                z["model_rvec"], z["model_ndeg"], z["model_hop"], z["model_ihop"])
    return m, p, z


def okw(p):
    return {k: p[k] for k in O.DEFAULT_PARAMS if k in p}


@pytest.mark.parametrize("name", CASES)
def test_exciton_bands_against_reference(name):
    m, p, z = load_bands(name)
    qp, lab = O.kpathbse(m.rlat, z["kpath"], int(z["npts"]))
    # REAL(qauxv(:,1)) and KLABELS as printed: identical in float32
    assert np.array_equal(qp[:, 0].astype(np.float32), z["qcoord"].astype(np.float32))
    assert np.array_equal(lab.astype(np.float32), z["klabels"].astype(np.float32))
    r = O.run_bands(m, p["ngrid"], qp, **okw(p))
    assert r["exk"].shape == z["exk"].shape
    assert np.abs(r["exk"] - z["exk"]).max() < 1e-12          # observed ~1e-14


def test_q_zero_is_the_q0_bse():
    m, p, z = load_bands("hex6_g4_v2drk")
    r0 = O.run_pipeline(m, p["ngrid"], **okw(p))
    rb = O.run_bands(m, p["ngrid"], np.zeros((1, 4)), **okw(p))
    assert np.abs(rb["exk"][:, 0] - r0["W"]).max() < 1e-12


@pytest.mark.parametrize("name", ["hex6_g3_v2dk_c1v2", "cub6_g3_v3d"])
def test_reciprocal_lattice_vector_gives_q_zero(name):
    m, p, z = load_bands(name)
    b = O.recvec(m.rlat)
    G = b[0] - 2.0 * b[1]
    rb = O.run_bands(m, p["ngrid"], np.array([[0.0, 0.0, 0.0, 0.0], [1.0, *G]]), **okw(p))
    # the phases k.R pass through COMPLEX(4) (hamiltonian_tb.f90:64-66), so H(k+G) equals
    # H(k) to float32 rounding of the larger phase only: 5e-9 .. 1.2e-7 eV here (a wrong
    # band window, sign or conjugate moves them by 1e-2 eV and more)
    assert np.abs(rb["exk"][:, 1] - rb["exk"][:, 0]).max() < 1e-6


def test_exchange_is_a_rank_one_update():
    """Energies with the exchange interlace with those of the direct term alone
    (-V(q) a a^H is positive semidefinite for V(q) < 0): lam_i(D) <= lam_i(H) <= lam_i+1(D)."""
    m, p, z = load_bands("hex6_g3_v2dk_c1v2")
    kw = okw(p)
    q = np.array([0.1, 0.05, -0.02, 0.0])
    r = O.run_bands(m, p["ngrid"], q[None, :], **kw)
    H = r["H"][0]
    vq = float(O.coulomb(p["coultype"], q[None, 1:], p["ngrid"], m.rlat, p["ediel"], p["lc"], p["ez"], p["w"],
                         p["r0"], p["ktol"])[0])
    assert vq < 0
    lam = np.linalg.eigvalsh(H)
    kpt = O.monkhorst_pack(p["ngrid"], p["shift"], m.rlat)
    E, U, nocp = O.eigsys(kpt, m)
    ek, vk = O.select_bands(E, U, nocp, p["nc"], p["nv"])
    Eq, Uq, nq = O.eigsys(O.monkhorst_pack_q(q, p["ngrid"], p["shift"], m.rlat), m)
    eq, vqv = O.select_bands(Eq, Uq, nq, p["nc"], p["nv"])
    stt = O.quantumnumbers2(len(kpt), p["nc"], p["nv"], np.full(len(kpt), p["nv"]), m.nw)
    # the direct term D of H: the q = 0 branch (no exchange) on the same k and k+q bands
    D = O.bse_matrix_q(kpt, ek, vk, eq, vqv, stt, np.zeros(4), dict(p, **kw), p["ngrid"], m.rlat)
    s = np.linalg.svd(H - D, compute_uv=False)
    assert s[1] < 1e-10 * s[0]
    ld = np.linalg.eigvalsh(D)
    assert (lam >= ld - 1e-10).all() and (lam[:-1] <= ld[1:] + 1e-10).all()


def test_kpathbse_shape_and_ends():
    rlat = np.array([[3.2, 0, 0], [-1.6, 2.7713, 0], [0, 0, 20.0]])
    ks = np.array([[0, 0, 0], [0.5, 0, 0], [0.5, 0, 0], [1 / 3, 1 / 3, 0]])
    qp, lab = O.kpathbse(rlat, ks, 7)
    assert qp.shape == (14, 4)
    assert qp[0, 0] == 0.0 and abs(lab[-1] - 1.0) < 1e-15
    assert np.all(np.diff(qp[:, 0]) >= 0)
    b = O.recvec(rlat)
    assert np.allclose(qp[6, 1:], 0.5 * b[0], atol=1e-15)       # segment end included


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
@pytest.mark.parametrize("name", CASES)
def test_c_abi_kpathbse_matches_reference(name):
    """wtb_kpathbse (host code of the C-ABI, no GPU) against the reference's coordinates and
    the oracle's q vectors."""
    from wantibexos_amd import bse
    m, p, z = load_bands(name)
    qc, labc = bse.kpathbse(m.rlat, z["kpath"], int(z["npts"]))
    qo, labo = O.kpathbse(m.rlat, z["kpath"], int(z["npts"]))
    assert np.array_equal(qc, qo)
    assert np.array_equal(labc, labo)
    assert np.array_equal(qc[:, 0].astype(np.float32), z["qcoord"].astype(np.float32))


@pytest.mark.skipif(not LIB.exists(), reason="library not built")
def test_c_abi_kpathbse_rejects_bad_arguments():
    """Host-side checks of wtb_kpathbse (no GPU): fewer than two points per segment or
    fewer than two path ends is an argument error, as the division by npts - 1 and the
    segment pairs require."""
    from wantibexos_amd import bse
    rlat = np.array([[3.2, 0, 0], [-1.6, 2.7713, 0], [0, 0, 20.0]])
    ks = np.array([[0, 0, 0], [0.5, 0, 0]])
    with pytest.raises(bse.WTBError):
        bse.kpathbse(rlat, ks, 1)
    with pytest.raises(bse.WTBError):
        bse.kpathbse(rlat, ks[:1], 4)
