"""GPU: the two-stage tridiagonal reduction (dense -> band of width 16 or 32 with MFMA
panel updates, band -> tridiagonal by bulge chasing) against the one-stage Householder
reduction and the reference's fixtures.  Both reduce the same Hermitian matrix by unitary
similarity, so the exciton energies agree to rounding, and the activities |X^T h|^2 (with
X^T h = Z^T Q^T h carried through either factorisation) agree to rounding wherever the
excitons are non-degenerate; eps(omega) sums them and agrees to rounding everywhere."""
import numpy as np
import pytest

from conftest import golden_cases
from golden_io import load
from test_gpu_parity import ctx, params_of, relerr  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu


def run(ctx, m, prm, two_stage, band=32):
    from wantibexos_amd import bse
    with bse.BSEPlan(ctx, m, prm) as plan:
        plan.set_reduction(two_stage, band)
        plan.execute()
        return plan.results()


@pytest.mark.parametrize("band", [16, 32])
@pytest.mark.parametrize("name", golden_cases())
def test_two_stage_matches_one_stage_and_reference(ctx, name, band):
    m, p, z = load(name)
    print(f"two-stage {name} band {band}", flush=True)
    prm = params_of(p)
    prm.f154 = False
    r1 = run(ctx, m, prm, False)
    r2 = run(ctx, m, prm, True, band)
    scale = np.abs(r1["W"]).max()
    assert np.abs(r2["W"] - r1["W"]).max() < 1e-12 * scale + 1e-13
    assert np.abs(r2["W"] - z["wbse"]).max() < 1e-9
    for c in range(6):
        for q in range(2):
            ref = r1["eps"][c, q]
            if np.abs(ref).max() > 1e-8:
                assert relerr(r2["eps"][c, q], ref) < 1e-9, (c, q)
    # activity sums per component (sum rule) and activities of isolated excitons
    for c in range(8):
        s1, s2 = r1["act"][c].sum(), r2["act"][c].sum()
        assert abs(s1 - s2) <= 1e-10 * max(abs(s1), np.abs(r1["act"][c]).max(), 1e-30), c
    W = r1["W"]
    gap = np.minimum(np.abs(np.diff(W, prepend=-np.inf)), np.abs(np.diff(W, append=np.inf)))
    iso = gap > 1e-6 * max(scale, 1.0)
    amax = np.abs(r1["act"][:3]).max()
    assert np.abs(r2["act"][:, iso] - r1["act"][:, iso]).max() < 1e-8 * amax


@pytest.mark.parametrize("band", [16, 32])
def test_two_stage_precision_form(ctx, band):
    """COMPLEX*8 matrix, FP32 MFMA panel updates, FP64 reflectors and bulge chasing:
    within 1 meV and 1e-3 (eps) of the plain FP64 path on N = 576 (nw = 22)."""
    from wantibexos_amd import bse, tbmodel
    m = tbmodel.hex2d(seed=2, nw=22)
    base = dict(ngrid=(12, 12, 1), nc=2, nv=2, coultype="V2DRK", r0=40.0, cshift=0.05,
                ebse0=0.0, ebsef=6.0, numbse=501, dtfull=True, cpol=True, f154=False)
    r64 = run(ctx, m, bse.BSEParams(**base), False)
    r32 = run(ctx, m, bse.BSEParams(**base, precision=bse.MIXED), True, band)
    assert np.abs(r32["W"] - r64["W"]).max() < 1e-3
    for c in range(6):
        for q in range(2):
            ref = r64["eps"][c, q]
            if np.abs(ref).max() > 1e-8:
                assert relerr(r32["eps"][c, q], ref) < 1e-3, (c, q)


def test_two_stage_small_and_ragged_sizes(ctx):
    """N from 2 to 144 with band widths 16 and 32 (panels shorter than the band, a last
    panel of one row, sweeps of one task) against the one-stage reduction."""
    from wantibexos_amd import bse, tbmodel
    m = tbmodel.hex2d(seed=9, nw=6, ngrid=(12, 12, 1))      # gap placed on a grid every test grid divides
    for g in [(1, 1, 1), (1, 2, 1), (3, 1, 1), (2, 3, 1), (4, 4, 1), (6, 6, 1)]:
        for nc, nv in [(1, 2), (2, 2)]:
            prm = bse.BSEParams(ngrid=g, nc=nc, nv=nv, coultype="V2DRK", r0=40.0, cshift=0.05, ebse0=0.0,
                                ebsef=6.0, numbse=101, dtfull=True, cpol=True, f154=False)
            N = g[0] * g[1] * nc * nv
            if N < 2:
                continue
            r1 = run(ctx, m, prm, False)
            for band in (16, 32):
                r2 = run(ctx, m, prm, True, band)
                assert np.abs(r2["W"] - r1["W"]).max() < 1e-12, (g, nc, nv, band)
                assert np.abs(r2["eps"] - r1["eps"]).max() <= 1e-9 * np.abs(r1["eps"]).max(), (g, band)


@pytest.mark.parametrize("band", [16, 32])
def test_two_stage_repeatable_at_full_size(ctx, band):
    """The full-size matrix (hex2d nw=22, 36x36x1, N = 5184: 81 bulge-chasing workgroups
    at once, 21-workgroup one-launch panels) executed six times in a row: every execute
    bitwise equal to the first (the kernels' sums run in a fixed order, so any other
    outcome is a data race), the trace preserved, and the energies those of the one-stage
    reduction to rounding."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    import gen_models
    from wantibexos_amd import bse
    m = gen_models.generate([("hex2d", 2, 22, (36, 36, 1))])[0]
    prm = bse.BSEParams(ngrid=(36, 36, 1), nc=2, nv=2, coultype="V2DRK", r0=40.0, cshift=0.05, ebse0=0.0,
                        ebsef=6.0, numbse=1001, dtfull=True, cpol=True, f154=False)
    r1 = run(ctx, m, prm, False)
    with bse.BSEPlan(ctx, m, prm) as plan:
        plan.set_reduction(True, band)
        first = None
        for rep in range(6):
            print(f"two-stage N=5184 band {band} execute {rep}", flush=True)
            plan.execute()
            r = plan.results()
            if first is None:
                first = r
                scale = np.abs(r1["W"]).max()
                assert np.abs(r["W"] - r1["W"]).max() < 1e-11 * scale, band
                assert abs(r["W"].sum() - r1["W"].sum()) < 1e-10 * abs(r1["W"].sum())
            else:
                assert np.array_equal(r["W"], first["W"]), (band, rep)
                assert np.array_equal(r["eps"], first["eps"]), (band, rep)
