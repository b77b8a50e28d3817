"""Pins the oracle's single-particle table (sp_table: the vecres(N, 15) table of
bse_solver-tool-diel.f90:365-420 and its finite-temperature form,
bse_solver-tool-diel-temp.f90:405-424) to what the unmodified reference program wrote
with TMCOEF= T (tests/golden/make_golden_sp.py): tm_coef.dat, tm_coef-pol.dat and
sp_optics.dat, in the reference's Bubblem order.  The integer columns (nocp and the band
numbers) and the k coordinates fix the row order, so the stable sort on the transition
energy is checked row for row; the real columns agree to the printed 1e-4."""
import json
from pathlib import Path

import numpy as np
import pytest

import wtb_oracle as O
from wantibexos_amd.tbmodel import TBModel

SP = Path(__file__).resolve().parent / "golden" / "sp"
CASES = sorted(p.stem for p in SP.glob("*.npz"))
HALF_ULP_F10_4 = 1.00001e-4          # a value on a rounding boundary of the printed 4 decimals


def load_sp(name):
    z = np.load(SP / f"{name}.npz")
    p = json.loads(str(z["params"]))
    for k in ("ngrid", "ediel", "shift"):
        p[k] = tuple(p[k])
    m = TBModel(str(z["model_systype"]), float(z["model_scs"]), float(z["model_efermi"]), z["model_rlat"],
                z["model_rvec"], z["model_ndeg"], z["model_hop"], z["model_ihop"])
    return m, p, z


def oracle_table(m, p):
    kw = {k: p[k] for k in O.DEFAULT_PARAMS if k in p}
    o = O.run_pipeline(m, p["ngrid"], **kw)
    return O.sp_table(o["kpt"], o["nocp"], o["eigv"], o["h"], o["stt"], p["nv"], o["fdeh"]).astype(np.float64)


def test_cases_present():
    assert len(CASES) == 3


@pytest.mark.parametrize("name", CASES)
def test_sp_table_against_reference(name):
    m, p, z = load_sp(name)
    t = oracle_table(m, p)
    tm, tmp, spo = z["tm_coef"], z["tm_coef_pol"], z["sp_optics"]
    assert t.shape == (len(tm), 15)
    # row order: the same (k, nocp, bands) in every row
    assert np.array_equal(t[:, 3:6], tm[:, 3:6])
    assert np.abs(t[:, 0:3] - tm[:, 0:3]).max() <= HALF_ULP_F10_4
    # energies and the six |h|^2 columns ("F10.4"), sigma+/- ("F10.4"), sp_optics ("F15.4")
    assert np.abs(t[:, 6:13] - tm[:, 6:13]).max() <= HALF_ULP_F10_4
    assert np.array_equal(tmp[:, :10], tm[:, :10])
    assert np.abs(t[:, 13:15] - tmp[:, 10:12]).max() <= HALF_ULP_F10_4
    assert np.abs(t[:, 6:13] - spo).max() <= HALF_ULP_F10_4
    # not vacuous: the table has structure beyond zeros, and ascending energies
    assert np.all(np.diff(tm[:, 6]) >= 0) and np.abs(tm[:, 7:10]).max() > 5e-3


def test_finite_temperature_columns_carry_the_fermi_factor():
    """At T = 4000 K the |h|^2 columns are scaled by f_v - f_c < 1: the T = 0 table of the
    same model differs from the fixture, the scaled one matches it (a dropped factor
    fails here)."""
    m, p, z = load_sp("hex6_g4_t4000_ve_sp")
    kw = {k: p[k] for k in O.DEFAULT_PARAMS if k in p}
    o = O.run_pipeline(m, p["ngrid"], **kw)
    f = o["fdeh"]
    assert f is not None and f.min() < 0.999
    t0 = O.sp_table(o["kpt"], o["nocp"], o["eigv"], o["h"], o["stt"], p["nv"], None).astype(np.float64)
    assert np.abs(t0[:, 7:10] - z["tm_coef"][:, 7:10]).max() > 1e-3


def bubble_order(e):
    """Bubblem's exchange rule (general_subs.f90:405-440) on a list of indices: swap
    neighbours while the left one is strictly greater, so equal keys keep their order."""
    idx = list(range(len(e)))
    for i in range(len(idx)):
        for j in range(len(idx) - 1 - i):
            if e[idx[j]] > e[idx[j + 1]]:
                idx[j], idx[j + 1] = idx[j + 1], idx[j]
    return idx


def test_sort_keeps_the_order_of_equal_transition_energies():
    """Brute force on a table with many exactly equal energies (the random models above
    have none): the oracle's order equals the bubble sort's."""
    rng = np.random.default_rng(7)
    nk, nc, nv, nw = 6, 2, 2, 5
    eigv = np.round(rng.uniform(-2, 2, (nk, nc + nv)))                # integers: many ties
    kpt = rng.uniform(-1, 1, (nk, 3))
    nocp = np.full(nk, 3)
    stt = O.quantumnumbers2(nk, nc, nv, nocp, nw)
    N = len(stt)
    h = [rng.normal(size=N) + 1j * rng.normal(size=N) for _ in range(5)]
    t = O.sp_table(kpt, nocp, eigv, h, stt, nv)
    ki = stt[:, 3] - 1
    e = (eigv[ki, stt[:, 2] - 1] - eigv[ki, stt[:, 1] - 1]).astype(np.float32)
    assert len(set(e.tolist())) < N // 2
    want = np.array(bubble_order(e))
    assert np.array_equal(t[:, 4], (nocp[ki] - nv + stt[:, 2]).astype(np.float32)[want])
    assert np.array_equal(t[:, 7], (h[0] * np.conj(h[0])).real.astype(np.float32)[want])
