"""GPU: the finite-temperature BSE (bsesolvertemp, subprograms/bse_solver-tool-diel-temp.f90)
at a size beyond the golden fixtures, against the oracle and through properties that
hold at any size.  The fixtures themselves run through test_gpu_parity /
test_gpu_mixed (their bset cases).
- N = 576 (12 x 12 k, 2 + 2 bands): exciton energies within 1e-9 eV of the oracle's
  FP64 eigh, eps(omega) without the F15.4 round trip within 1e-8 relative;
- the activity sum rule sum_n act_n = Re sum_j h_a,j conj(h_b,j) f_j (X unitary)
  on the GPU's own dipoles and Fermi factors, every component (1e-9 relative);
- the precision form on the same matrix: 1 meV, eps 1e-3;
- the iterative form on a finite-temperature plan: the full solve's lowest excitons,
  and eps(omega) from polarised Lanczos chains against the full solve's."""

import numpy as np
import pytest

import wtb_oracle as O
from test_gpu_parity import ctx, relerr  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

NG = (12, 12, 1)
KW = dict(nc=2, nv=2, coultype="V2DRK", r0=40.0, cshift=0.05, ebse0=0.0, ebsef=6.0, numbse=1001,
          dtfull=True, f154=False, bset=True, temp=3000.0, ta="VE", st=0.3, phavg=0.05)


def params(**kw):
    from wantibexos_amd import bse
    d = dict(KW, **kw)
    return bse.BSEParams(ngrid=NG, cpol=True, **d)


@pytest.fixture(scope="module")
def model():
    from wantibexos_amd import tbmodel
    return tbmodel.hex2d(seed=5, nw=10, ngrid=NG)


def test_finite_temperature_against_oracle(ctx, model):
    from wantibexos_amd import bse
    print("case: N=576 finite temperature VE 3000 K, FP64", flush=True)
    with bse.BSEPlan(ctx, model, params()) as plan:
        assert plan.N == 576
        plan.execute()
        r = plan.results()
        dip = plan.fetch("dipole")
        f = plan.fetch("fdeh")
    o = O.run_pipeline(model, NG, **{k: v for k, v in KW.items() if k in O.DEFAULT_PARAMS})
    assert o["tcor"] < 0 and 0.5 < f.min() < f.max() < 1.0        # a visible temperature effect
    assert np.abs(f - o["fdeh"]).max() < 1e-14
    assert np.abs(r["W"] - o["W"]).max() < 1e-9
    for c in range(6):
        assert relerr(r["eps"][c, 0], o["eps_re"][c]) < 1e-8, c
        assert relerr(r["eps"][c, 1], o["eps_im"][c]) < 1e-8, c
    # sum rule on the GPU's own h and f
    hp = (dip[0] + 1j * dip[1]) * float(np.float32(1 / np.sqrt(np.float32(2))))
    hm = (dip[0] - 1j * dip[1]) * float(np.float32(1 / np.sqrt(np.float32(2))))
    h = [dip[0], dip[1], dip[2], hp, hm]
    for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2), (3, 3), (4, 4)]):
        expect = (h[a] * np.conj(h[b]) * f).sum().real
        scale = max(abs(expect), np.abs(h[a]).max() * np.abs(h[b]).max())
        assert abs(r["act"][c].sum() - expect) <= 1e-9 * scale, c


def test_finite_temperature_precision_form(ctx, model):
    from wantibexos_amd import bse
    print("case: N=576 finite temperature VE 3000 K, mixed", flush=True)
    r64 = bse.bsesolver(ctx, model, params())
    r32 = bse.bsesolver(ctx, model, params(precision=bse.MIXED))
    assert np.abs(r32["W"] - r64["W"]).max() < 1e-3
    for c in range(6):
        for part in range(2):
            ref = r64["eps"][c, part]
            if np.abs(ref).max() > 1e-8:
                assert relerr(r32["eps"][c, part], ref) < 1e-3, (c, part)


def test_finite_temperature_iterative_form(ctx, model):
    """Davidson on the finite-temperature matrix, and eps(omega) from Lanczos chains of
    h_a + f h_b and h_a - f h_b (Re[t conj(u)] = (|t + u|^2 - |t - u|^2) / 4): the full
    FP64 solve's excitons (1 meV bar) and spectrum (1e-3 relative, normwise per component)."""
    from wantibexos_amd import bse
    print("case: N=576 finite temperature VE 3000 K, Davidson + Lanczos", flush=True)
    full = bse.bsesolver(ctx, model, params())
    with bse.BSEPlan(ctx, model, params(precision=bse.MIXED)) as plan:
        r = plan.lowest(nev=32, guard=8, tol=1e-6, lanczos_steps=300)
    assert r["info"]["converged"] == 1
    assert np.abs(r["W"] - full["W"][:32]).max() < 1e-4
    errs = {}
    for c in range(6):
        for part in range(2):
            ref = full["eps"][c, part]
            if np.abs(ref).max() > 1e-8:
                errs[(c, part)] = relerr(r["eps"][c, part], ref)
    print("eps relative errors:", errs)
    assert len(errs) >= 6 and max(errs.values()) < 1e-3


def test_temperature_zero_limit(ctx, model):
    """BSET at T = 0 (taken as 1E-8 K) with TA = FA: every f_v - f_c is 1, so the matrix and
    the energies are bsesolver's bit for bit; the activities are formed by another kernel
    (Re[a conj(b)] with b = a), equal to the last unit."""
    from wantibexos_amd import bse
    print("case: N=576 finite-temperature path at T = 0", flush=True)
    r0 = bse.bsesolver(ctx, model, params(bset=False))
    r1 = bse.bsesolver(ctx, model, params(temp=0.0, ta="FA"))
    assert np.array_equal(r0["W"], r1["W"])
    assert np.abs(r1["act"] - r0["act"]).max() <= 1e-14 * np.abs(r0["act"]).max()
    assert np.abs(r1["eps"] - r0["eps"]).max() <= 1e-13 * np.abs(r0["eps"]).max()


def test_finite_temperature_argument_errors(ctx, model):
    """The host rejects what the kernels cannot take, before anything is launched."""
    from wantibexos_amd import bse
    for bad in (dict(temp=-1.0), dict(temp=float("nan")), dict(st=float("inf"))):
        with pytest.raises(bse.WTBError):
            bse.BSEPlan(ctx, model, params(**bad))
    # TEMP = 0 with PHAVG = 0 makes the VE gap correction 0/0: refused, not propagated
    with pytest.raises(bse.WTBError):
        bse.BSEPlan(ctx, model, params(temp=0.0, ta="VE", phavg=0.0))
