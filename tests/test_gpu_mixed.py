"""GPU: the precision form (COMPLEX*8 BSE matrix, FP32 MFMA overlaps and trailing
updates, FP64 reflector norms, tridiagonal solve and spectrum) against the
reference's FP64 results.  Bars: exciton energies within 1 meV, eps(omega)
(without the F15.4 round trip) within 1e-3 relative (normwise per component), the FP32 matrix within FP32
rounding of the FP64 one."""
import numpy as np
import pytest

from conftest import golden_cases
from golden_io import load
from test_gpu_parity import ctx, params_of, relerr  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

E_TOL_MIXED = 1e-3      # eV
EPS_TOL_MIXED = 1e-3    # relative, per eps component


def mixed(prm):
    """The precision form without the F15.4 round trip: a 1e-7 change of an
    activity can flip its fourth decimal, so eps is compared unquantized."""
    from wantibexos_amd import bse
    prm.precision = bse.MIXED
    prm.f154 = False
    return prm


@pytest.mark.parametrize("name", golden_cases())
def test_mixed_against_reference(ctx, name):
    from wantibexos_amd import bse
    m, p, z = load(name)
    with bse.BSEPlan(ctx, m, params_of(p)) as plan:
        plan.build()
        H64 = np.tril(plan.fetch("hbse"))
    with bse.BSEPlan(ctx, m, mixed(params_of(p))) as plan:
        plan.build()
        H32 = np.tril(plan.fetch("hbse"))
        plan.execute()
        r = plan.results()
    assert H32.dtype == np.complex64
    scale = np.abs(H64).max()
    assert np.abs(H32.astype(np.complex128) - H64).max() <= 4e-7 * scale
    assert np.abs(r["W"] - z["wbse"]).max() < E_TOL_MIXED
    for c in range(6 if p["dtfull"] else 3):
        for part in range(2):
            ref = z["dielf"][c, part]
            if np.abs(ref).max() > 1e-8:
                assert relerr(r["eps"][c, part], ref) < EPS_TOL_MIXED, (c, part)


def test_mixed_larger_model(ctx):
    """N = 576 (12x12 k, 2+2 bands, 22 orbitals): mixed against FP64 on the GPU."""
    from wantibexos_amd import bse, tbmodel
    m = tbmodel.hex2d(seed=2, nw=22)
    base = dict(ngrid=(12, 12, 1), nc=2, nv=2, coultype="V2DRK", r0=40.0, cshift=0.05,
                ebse0=0.0, ebsef=6.0, numbse=501, dtfull=True, cpol=True, f154=False)
    r64 = bse.bsesolver(ctx, m, bse.BSEParams(**base))
    r32 = bse.bsesolver(ctx, m, bse.BSEParams(**base, precision=bse.MIXED))
    assert np.abs(r32["W"] - r64["W"]).max() < E_TOL_MIXED
    # oscillator strengths summed over the spectrum are smooth in the matrix
    for c in range(6):
        for part in range(2):
            ref = r64["eps"][c, part]
            if np.abs(ref).max() > 1e-8:
                assert relerr(r32["eps"][c, part], ref) < EPS_TOL_MIXED, (c, part)


@pytest.mark.parametrize("name", ["hex8_g12_v2drk_c1", "hex8_g6_v2drk", "cub6_g3_v3dl", "hex6_g4_t2500_ve"])
def test_refinement_of_lowest_excitons(ctx, name):
    """FP64 Rayleigh-quotient refinement of the lowest excitons in the precision form: each
    refined energy lies within its residual norm of the reference's FP64 exciton energy
    (the FP64 matrix has an eigenvalue within |H x - theta x| / |x| of theta), is at least
    as close as the unrefined one up to that bound, and the excitons above the refined set,
    the activities and eps(omega) are those of the unrefined precision form."""
    from wantibexos_amd import bse
    m, p, z = load(name)
    N = int(np.prod(p["ngrid"])) * p["nc"] * p["nv"]
    nref = min(64, N // 2)
    with bse.BSEPlan(ctx, m, mixed(params_of(p))) as plan:
        plan.set_reduction(False)               # the refinement's own (one-stage) reduction
        plan.execute()
        r0 = plan.results()
        plan.set_refine(nref)
        for _ in range(2):                       # capture, then replay
            plan.execute()
            r = plan.results()
            res, w0 = plan.refine_results()
            assert np.array_equal(w0, r0["W"][:nref])
            ref = z["wbse"][:nref]
            d = np.abs(r["W"][:nref] - ref)
            assert np.all(d <= res + 1e-12), (d - res).max()
            assert d.max() < 1e-5 and res.max() < 1e-3
            assert np.array_equal(r["W"][nref:], r0["W"][nref:])
            assert np.array_equal(r["act"], r0["act"])
        plan.set_refine(0)
        plan.execute()
        assert np.array_equal(plan.results()["W"], r0["W"])
