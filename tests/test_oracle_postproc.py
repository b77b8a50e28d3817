"""Pins the oracle's post-processing of the spectra (optprop) and its exciton lifetimes
to the reference program's own files (tests/golden/postproc, written by
tests/golden/make_golden_postproc.py from oracle/_ref/wtb.x), plus closed forms.

The fixtures hold what the files hold: the spectra as printed (REAL, list-directed),
the optical constants in F15.4 and the lifetimes in E15.4, so the oracle, fed the
printed inputs, must land within half a unit of the last printed digit."""
from pathlib import Path

import numpy as np
import pytest

import wtb_oracle as O

GOLD = Path(__file__).resolve().parent / "golden" / "postproc"
CASES = sorted(p.stem for p in GOLD.glob("*.npz"))
QUANT = ["refractive_index", "extinction_coef", "reflectibility", "absorption_coef", "en_loss_func"]
DIEL_READ = ["xx", "xy", "xz", "yy", "yz", "zz"]    # bseoptprop's reading order
DIEL_COLS = ["xx", "yy", "zz", "xy", "xz", "yz"]    # its output columns
POL = ["x", "y", "z", "sp", "sm"]


def spectra(z, pol):
    names = POL if pol else DIEL_COLS
    key = "dielpol_" if pol else "diel_"
    en = z[key + names[0]][:, 0]
    e1 = np.stack([z[key + c][:, 1] for c in names], axis=1)
    e2 = np.stack([z[key + c][:, 2] for c in names], axis=1)
    return en, e1, e2


def e_format_halfulp(x):
    """half a unit of the last digit of E15.4 output (0.dddd E+xx)"""
    x = np.abs(x)
    e = np.floor(np.log10(np.where(x > 0, x, 1.0))) + 1
    return 0.5e-4 * 10.0 ** e


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("pol", [False, True], ids=["bseoptprop", "bseoptproppol"])
def test_optprop_matches_reference_files(name, pol):
    z = np.load(GOLD / f"{name}.npz")
    en, e1, e2 = spectra(z, pol)
    got = O.optprop(en, e1, e2)
    for q, g in zip(QUANT, got):
        ref = z[f"bse_{q}" + ("_pol" if pol else "")]
        assert np.array_equal(ref[:, 0], np.round(en, 4))
        r = ref[:, 1:]
        # 0/0 (a component that is zero at every energy) is NaN in both
        assert np.array_equal(np.isnan(g), np.isnan(r)), (name, q)
        ok = ~np.isnan(r)
        # F15.4: within half a unit of the 4th decimal (plus the printed value's own
        # rounding of a REAL input)
        assert np.abs(g[ok] - r[ok]).max() <= 0.5e-4 + 1e-9 * np.abs(r[ok]).max(), (name, q)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("pol", [False, True], ids=["lifetime", "lifetimepol"])
def test_lifetimes_match_reference_files(name, pol):
    import json
    z = np.load(GOLD / f"{name}.npz")
    p = json.loads(str(z["params"]))
    tab = z["opt_diel_pol" if pol else "opt_diel"]
    ref = z["lifetime_pol" if pol else "lifetime"]
    got = O.lifetimes(p["sysdim"], p["ngrid"], z["rlat"], tab, 5 if pol else 3)
    assert np.array_equal(ref[:, 0], tab[:, 0])
    fin = np.isfinite(got)
    tol = e_format_halfulp(ref[:, 1:]) * 1.0001
    assert np.all(np.abs(got[fin] - ref[:, 1:][fin]) <= tol[fin]), name
    # zero strengths give an infinite lifetime in both
    assert np.array_equal(~np.isfinite(got), ~np.isfinite(ref[:, 1:]))


def test_optprop_closed_forms():
    """n + ik = sqrt(eps) (principal root), R from n and k, L = Im(-1/eps) and
    alpha = 2 E k / (hbar c) with the reference's single-precision sqrt(2) and hc."""
    rng = np.random.default_rng(5)
    e1 = rng.uniform(-5, 10, 200)
    e2 = rng.uniform(0, 10, 200)
    en = rng.uniform(0, 6, 200)
    n, k, R, a, L = O.optprop(en, e1, e2)
    z = np.sqrt(e1 + 1j * e2)
    assert np.allclose(n, z.real, rtol=1e-13) and np.allclose(k, z.imag, rtol=1e-12, atol=1e-13)
    assert np.allclose(R, np.abs((z - 1) / (z + 1)) ** 2, rtol=1e-12)
    assert np.allclose(L, (-1 / (e1 + 1j * e2)).imag, rtol=1e-13)
    assert np.allclose(a, 2 * en * k / O.HC_F32 * (O.SQRT2_F32 / np.sqrt(2.0)), rtol=1e-12)
    assert O.HC_F32 != 19.746e-6 and O.SQRT2_F32 != np.sqrt(2.0)


def test_exclft_closed_forms():
    """tau_3D = 3 c^2 hbar^3 N_k / (4 alpha E^3 f), tau_2D = N_k A hbar / (8 pi alpha E f),
    tau_1D = N_k a hbar^2 c / (2 pi alpha E^2 f) with the REAL constants."""
    rl = np.array([[3.2, 0, 0], [-1.6, 2.77128, 0], [0, 0, 20.0]])
    g = (4, 5, 1)
    nk = 20.0
    E, f = np.array([1.7, 2.3]), np.array([0.5, 3.0])
    c, hb, al = O.LSPD_F32, O.HBAR_F32, O.CFE_F32
    assert np.allclose(O.exclft("3D", g, rl, f, E), 3 * c ** 2 * hb ** 3 * nk / (4 * al * E ** 3 * f), rtol=1e-13)
    A = abs(np.cross(rl[0], rl[1])[2])
    assert np.allclose(O.exclft("2D", g, rl, f, E), nk * A * hb / (8 * O.PI32 * al * E * f), rtol=1e-13)
    assert np.allclose(O.exclft("1D", g, rl, f, E), nk * 3.2 * hb ** 2 * c / (2 * O.PI32 * al * E ** 2 * f),
                       rtol=1e-13)
    with pytest.raises(ValueError):
        O.exclft("0D", g, rl, f, E)
