"""Pins the CPU oracle (oracle/wtb_oracle.py) to the reference.

Expected values come from the golden fixtures that tests/golden/make_golden.py
wrote by running the reference's own subroutines (compiled from its sources
with flang + MKL) on seeded synthetic models.  Closed-form checks cover what
a plausible slip (wrong constant, sign, index or missing term) would break.
"""
import math
from pathlib import Path

import numpy as np
import pytest

import wtb_oracle as O
from conftest import golden_cases
from golden_io import load, oracle_kwargs, ref_signs, unpack_upper

CASES = golden_cases()


@pytest.fixture(scope="module", params=CASES)
def case(request):
    m, p, z = load(request.param)
    r = O.run_pipeline(m, p["ngrid"], **oracle_kwargs(p))
    return request.param, m, p, z, r


def test_tables_bit_exact(case):
    name, m, p, z, r = case
    assert np.array_equal(r["kpt"], z["kpt"]), "k grid must be bit-exact"
    assert np.array_equal(r["stt"], z["stt"]), "BSE basis table must be bit-exact"
    assert np.array_equal(r["nocp"], z["nocpk"])


def test_bands_and_gauge(case):
    name, m, p, z, r = case
    assert np.abs(r["E"] - z["eall"]).max() < 1e-12
    # ZHEEV('L') leaves the orbital-1 component real
    assert np.abs(z["vector"][:, :, 0].imag).max() == 0.0
    s, _ = ref_signs(z)
    assert np.abs(z["vector"] * s[:, :, None] - r["vector"]).max() < 1e-12
    assert abs(z["egap"][0] - (r["eigv"][:, p["nv"]] - r["eigv"][:, p["nv"] - 1]).min()) < 1e-12


def test_dipoles(case):
    name, m, p, z, r = case
    _, sg = ref_signs(z)
    for i, nm in enumerate(["hrx", "hry", "hrz", "hrsp", "hrsm"]):
        ref = z[nm] * sg
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(ref - r["h"][i]).max() / scale < 1e-12, nm


def test_bse_matrix(case):
    name, m, p, z, r = case
    _, sg = ref_signs(z)
    N = len(z["stt"])
    H = r["H"]
    assert np.abs(H - H.conj().T).max() == 0.0
    rows = {"hrow1": 0, "hrow2": min(1, N - 1), "hrowm": max(0, N // 2 - 1)}
    for nm, i in rows.items():
        ref = z[nm] * sg[i] * sg
        mine = np.where(np.arange(N) >= i, H[i], 0.0)      # the reference fills j >= i only
        assert np.abs(ref - mine).max() < 1e-12, nm
    if "hbse_up" in z:
        Hr = unpack_upper(z["hbse_up"], N) * np.outer(sg, sg)
        assert np.abs(Hr - H).max() < 1e-12


def test_exciton_energies_and_activities(case):
    name, m, p, z, r = case
    assert np.abs(r["W"] - z["wbse"]).max() < 1e-10
    a = z["act"]
    comps = [0, 1, 2, 6, 7] + ([3, 4, 5] if p["dtfull"] else [])
    scale = np.abs(a[comps]).max()
    assert np.abs(r["act"][comps] - a[comps]).max() / scale < 1e-10
    if not p["dtfull"]:
        assert (r["act"][3:6] == 0).all()


def test_f154_round_trip(case):
    name, m, p, z, r = case
    assert np.array_equal(r["Wq"], z["wq"])
    assert np.sum(r["actq"] != z["actq"]) <= 1


def test_spectrum(case):
    name, m, p, z, r = case
    dq = z["dielq"]
    for c in range(6):
        for part, mine in ((0, r["eps_re"][c]), (1, r["eps_im"][c])):
            ref = dq[c, part]
            scale = max(np.abs(ref).max(), 1e-30)
            assert np.abs(mine - ref).max() / scale < 1e-10, (c, part)
    # the full-precision (no F15.4) spectrum, from the unrounded activities
    act6 = r["act"][:6]
    om, re, im = O.spectrum(r["W"], act6, m.rlat, p["ngrid"], p["ebse0"], p["ebsef"], p["numbse"],
                            p["cshift"], p["renorm"])
    df = z["dielf"]
    comps = range(6) if p["dtfull"] else range(3)
    for c in comps:
        assert np.abs(re[c] - df[c, 0]).max() / max(np.abs(df[c, 0]).max(), 1e-30) < 1e-9
        assert np.abs(im[c] - df[c, 1]).max() / max(np.abs(df[c, 1]).max(), 1e-30) < 1e-9


# ---------------------------------------------------------------------------
# closed forms and invariants
# ---------------------------------------------------------------------------

def test_float32_constants():
    assert O.PI32 == 3.14159274101257324
    assert O.CIC == -90.4755935668945312
    assert O.GAMA == 180.899993896484375
    assert O.RENORM_FACTOR == float(np.float32(0.15)) != 0.15
    assert O.INV_SQRT2_F32 == float(np.float32(0.70710677))


def test_recvec_duality():
    rlat = np.array([[3.2, 0, 0], [-1.6, 2.7713, 0], [0, 0, 20.0]])
    b = O.recvec(rlat)
    # a_i . b_j = 2 pi32 delta_ij (the reference's single-precision pi)
    assert np.allclose(rlat @ b.T, 2 * O.PI32 * np.eye(3), atol=1e-12)


def test_rytova_keldysh_reduces_to_closed_form():
    rlat = np.array([[3.2, 0, 0], [-1.6, 2.7713, 0], [0, 0, 20.0]])
    q = np.array([[0.05, 0.0, 0.0], [0.0, 0.3, 0.0], [0.2, -0.1, 0.0]])
    ng = (12, 12, 1)
    A = abs(3.2 * 2.7713)
    v = O.coulomb("V2DRK", q, ng, rlat, (1.0, 1.0, 1.0), 5.0, 1.0, 0.0, 40.0, 1e-3)
    qq = np.linalg.norm(q, axis=1)
    expect = O.CIC / (144 * A) / (qq * (1 + 40.0 * qq))
    assert np.allclose(v, expect, rtol=1e-13)
    assert O.coulomb("V2DRK", np.zeros((1, 3)), ng, rlat, (1, 1, 1), 5.0, 1.0, 0.0, 40.0, 1e-3)[0] == 0.0


def test_v3dl_closed_form():
    rlat = np.eye(3) * 5.4
    q = np.array([[0.1, 0.2, 0.3]])
    v = O.coulomb("V3DL", q, (4, 4, 4), rlat, (1.0, 6.0, 1.0), 1.0, 1.0, 0.0, 1.0, 1e-3)
    assert np.allclose(v, O.CIC / (64 * 5.4 ** 3) / (6.0 * (q ** 2).sum()), rtol=1e-14)


def test_activity_sum_rule():
    rng = np.random.default_rng(0)
    n = 40
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    H = A + A.conj().T
    W, X = O.bse_eig(H)
    h = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    act = O.activities(X, h)
    # X is unitary, so sum_n |X_n^T h|^2 = ||h||^2 for every component
    for c, comp in ((0, 0), (1, 1), (2, 2), (6, 3), (7, 4)):
        assert math.isclose(act[c].sum(), np.vdot(h[comp], h[comp]).real, rel_tol=1e-12)


def test_finite_temperature_activity_sum_rule():
    """dielbseptemp summed over excitons: sum_n Re[(X^T h_a)_n conj(X^T (f h_b))_n]
    = Re[sum_j h_a,j conj(h_b,j) f_j] (X unitary); f on the wrong vector, a missing
    conjugate or swapped components break it."""
    rng = np.random.default_rng(1)
    n = 40
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    W, X = O.bse_eig(A + A.conj().T)
    h = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    f = rng.uniform(0.2, 1.0, n)
    act = O.activities_temp(X, h, f)
    for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (3, 3), (4, 4)]):
        assert math.isclose(act[c].sum(), (h[a] * np.conj(h[b]) * f).sum().real, rel_tol=1e-11, abs_tol=1e-11)
    # f = 1: the T = 0 activities
    assert np.allclose(O.activities_temp(X, h, np.ones(n)), O.activities(X, h), rtol=1e-13, atol=1e-13)
    # the polarisation the Lanczos form's finite-T chains use: Re[t conj(u)] =
    # (|X^T (h_a + f h_b)|^2 - |X^T (h_a - f h_b)|^2) / 4
    quad = lambda v: np.abs(X.T @ v) ** 2
    for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]):
        pol = 0.25 * (quad(h[a] + f * h[b]) - quad(h[a] - f * h[b]))
        assert np.abs(pol - act[c]).max() < 1e-12


def test_f154_rounding():
    x = np.array([1.23455, 1.23445, 0.00004999, -0.00006, 2.5])
    assert list(O.f15_4(x)) == [float(f"{v:.4f}") for v in x]
    assert O.f15_4(np.array([0.00004]))[0] == 0.0


def test_imagrenorm_cases():
    om = np.linspace(0, 6, 11)
    im = np.linspace(1, 2, 11)
    # first exciton below 0.15 eV: untouched
    assert np.array_equal(O.imagrenorm(np.array([0.1]), om, im), im)
    out = O.imagrenorm(np.array([2.0]), om, im)
    i0 = int(np.where(om <= 2.0 - O.RENORM_FACTOR)[0][-1])
    assert (out[:i0 + 1] == 0).all()
    assert np.allclose(out[i0 + 1:], im[i0 + 1:] - im[i0 + 1])


def test_lorentzian_width_is_telescoped_spacing():
    E = np.sort(np.random.default_rng(1).uniform(1, 3, 50))
    assert math.isclose(O.avgsme(E), 10 * (E[-1] - E[0]) / 50, rel_tol=1e-12)


def test_delta_table_is_unfolded():
    d = O.delta_index((3, 2, 1))
    assert d.shape == (6, 6, 3)
    assert d[:, :, 0].min() == -2 and d[:, :, 0].max() == 2
    assert (d + d.transpose(1, 0, 2) == 0).all()


def test_finite_temperature_factors(case):
    """bsesolvertemp: the gap correction added to the scissor and f_v - f_c per state."""
    name, m, p, z, r = case
    if not p.get("bset"):
        pytest.skip("T = 0 case")
    # tanh / exp of two libraries: equal to the last unit or two
    assert abs(r["tcor"] - z["tcor"][0]) <= 1e-15 * max(abs(z["tcor"][0]), 1.0)
    assert np.abs(r["fdeh"] - z["fdeh"]).max() < 1e-14


def test_fermi_factor_closed_forms():
    """f_v - f_c is 1 for a gapped state at T = 0 (taken as 1E-8 K), tends to 0 when both
    energies sit far on the same side of E_F, and equals tanh-form (1/2)(tanh(b ec/2) -
    tanh(b ev/2)) at any T; the VE and BE gap corrections vanish at T = 0."""
    assert O.fermidisteh(1.0, -1.0, 0.0) == 1.0
    b = 1.0 / (O.KB_F32 * 3000.0)
    for ec, ev in [(0.8, -1.1), (2.0, -0.3), (0.1, -0.05)]:
        ref = 0.5 * (math.tanh(b * ec / 2) - math.tanh(b * ev / 2))
        assert abs(O.fermidisteh(ec, ev, 3000.0) - ref) < 1e-14
    assert O.gap_correction("VE", 0.5, 0.03, 0.0) == 0.0
    assert O.gap_correction("BE", 0.5, 200.0, 0.0) == 0.0
    assert O.KB_F32 == float(np.float32(8.617330350e-5))


def test_rkmesh_pinned_to_reference():
    """rkmesh / rkmesh2D (MESH_TYPE= RK3D / RK2D, mhkpack_subs.f90:1-52) against the
    reference's own routines on 162 cases (tests/golden/rkmesh.json), including
    densities within 1e-6 of a rounding boundary."""
    import json
    data = json.loads((Path(__file__).resolve().parent / "golden" / "rkmesh.json").read_text())["cases"]
    for c in data:
        L = np.array(c["rlat"]).reshape(3, 3)
        assert O.rkmesh(c["rk"], L) == c["rk3d"], c
        assert O.rkmesh2D(c["rk"], L) == c["rk2d"], c


def test_rkmesh_closed_form_and_planted_mistakes():
    """Closed form on orthogonal cells (|b_i| / 2 pi = 1 / a_i).  The single-precision
    pi enters recvec and the division alike and cancels; the fixture catches a dropped
    +1/2 and the direct-lattice length 1 / |a_i| in place of |b_i| / 2 pi on oblique cells."""
    import json
    import math
    for a, rk in [(5.4, 40.0), (3.0, 30.0), (4.5, 7.0), (2.0, 0.5)]:
        L = np.diag([a, a * 1.5, a * 3.0])
        want = [int(max(1.0, rk / (a * f) + 0.5)) for f in (1.0, 1.5, 3.0)]
        assert O.rkmesh(rk, L) == want
        assert O.rkmesh2D(rk, L) == want[:2] + [1]
    data = json.loads((Path(__file__).resolve().parent / "golden" / "rkmesh.json").read_text())["cases"]

    def direct_length(rk, L):
        return [int(max(1.0, rk / np.linalg.norm(L[i]) + 0.5)) for i in range(3)]

    def no_half(rk, L):
        b = O.recvec(L)
        return [int(max(1.0, rk * O.vecsize(b[i]) / (2 * O.PI32))) for i in range(3)]

    for wrong in (direct_length, no_half):
        assert any(wrong(c["rk"], np.array(c["rlat"]).reshape(3, 3)) != c["rk3d"] for c in data), wrong.__name__
