"""GPU: the batching form (many uneven models, one captured graph per matrix,
replayed on parallel streams) and the GEMV tile autotuner.
- each matrix of a batch gives bitwise the results of its own plan run alone;
- every panel width gives the same excitons and spectra (to rounding, the
  updates are grouped differently) and the tuner picks one of them;
- a batch with lanes = 1 and lanes = 4 agree bitwise."""
import numpy as np
import pytest

from golden_io import load
from test_gpu_parity import ctx, params_of  # noqa: F401  (fixture)

pytestmark = pytest.mark.gpu

NAMES = ["hex8_g6_v2drk", "cub6_g3_v3dl", "hex8_g12_v2drk_c1", "hex8_g3_v2drk", "hex6_g3_v1dt"]


def singles(ctx, cases):
    from wantibexos_amd import bse
    out = []
    for m, p in cases:
        with bse.BSEPlan(ctx, m, params_of(p)) as plan:
            plan.execute()
            out.append(plan.results())
    return out


@pytest.mark.parametrize("lanes", [1, 4])
def test_batch_matches_single_plans(ctx, lanes):
    from wantibexos_amd import bse
    cases = []
    for nm in NAMES:
        try:
            m, p, _ = load(nm)
        except FileNotFoundError:
            continue
        cases.append((m, p))
    assert len(cases) >= 3
    ref = singles(ctx, cases)
    with bse.BSEBatch(ctx, [m for m, _ in cases], [params_of(p) for _, p in cases], lanes=lanes) as b:
        for _ in range(2):                    # first execute captures, second replays
            b.execute()
            for i, r in enumerate(ref):
                got = b.results(i)
                assert np.array_equal(got["W"], r["W"]), i
                assert np.array_equal(got["act"], r["act"]), i
                assert np.array_equal(got["eps"], r["eps"]), i


def test_panel_widths_and_tuner(ctx):
    from wantibexos_amd import bse, tbmodel
    m = tbmodel.hex2d(seed=4, nw=8, ngrid=(16, 16, 1))     # N = 1024
    prm = bse.BSEParams(ngrid=(16, 16, 1), nc=2, nv=2, coultype="V2DRK", r0=40.0, cshift=0.05, ebse0=0.0,
                        ebsef=6.0, numbse=501, dtfull=True, cpol=True, f154=False)
    res = {}
    with bse.BSEPlan(ctx, m, prm) as plan:
        for nb in (16, 20, 24, 32):
            plan.set_panel(nb)
            plan.execute()
            res[nb] = plan.results()
        best, ms = plan.tune()
        plan.execute()
        tuned = plan.results()
    print("panel times (ms):", ms, "best:", best)
    assert best in (16, 20, 24, 32)
    for nb in (16, 20, 24):
        assert np.abs(res[nb]["W"] - res[32]["W"]).max() < 1e-11
        assert np.abs(res[nb]["eps"] - res[32]["eps"]).max() < 1e-9 * np.abs(res[32]["eps"]).max()
    assert np.array_equal(tuned["W"], res[best]["W"])
