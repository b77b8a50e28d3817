"""GPU: finite-Q exciton bands (bsebnds, subprograms/bse_kpath-tool.f90, and bsebndstemp)
through the C-ABI.
- every bands fixture (two at finite temperature): wtb_bsebnds against the reference
  program's bands_bse.dat,
  1e-9 eV (the q = 0 path is bit-compatible with bsesolver's; ZHEEV against the
  GPU's Householder + divide and conquer);
- the plan form of one q: the finite-Q matrix (lower triangle) against the oracle's,
  1e-12;
- N = 576 along a Gamma-K-K'-Gamma path: against the oracle (1e-9 eV), lanes 1 and
  4 bit-identical, the precision form within 1 meV, and the iterative form's lowest
  excitons at one q against the full solve."""
import numpy as np
import pytest

import wtb_oracle as O
from test_gpu_parity import ctx  # noqa: F401  (fixture)
from test_oracle_bands import CASES, load_bands, okw

pytestmark = pytest.mark.gpu


def params_of(p, **kw):
    from wantibexos_amd import bse
    d = dict(ngrid=p["ngrid"], mshift=p["shift"], nc=p["nc"], nv=p["nv"], coultype=p["coultype"],
             ediel=p["ediel"], lc=p["lc"], ez=p["ez"], w=p["w"], r0=p["r0"], ktol=p["ktol"],
             cshift=p["cshift"], ebse0=p["ebse0"], ebsef=p["ebsef"], numbse=p["numbse"],
             bset=p.get("bset", False), temp=p.get("temp", 0.0), ta=p.get("ta", "FA"), st=p.get("st", 0.0),
             phavg=p.get("phavg", 0.0))
    d.update(kw)
    return bse.BSEParams(**d)


@pytest.mark.parametrize("name", CASES)
def test_bands_against_reference(ctx, name):
    from wantibexos_amd import bse
    m, p, z = load_bands(name)
    print("case:", name, flush=True)
    qp, _ = bse.kpathbse(m.rlat, z["kpath"], int(z["npts"]))
    exk = bse.bsebnds(ctx, m, params_of(p), qp, lanes=3)
    assert exk.shape == z["exk"].shape
    assert np.abs(exk - z["exk"]).max() < 1e-9


def test_finite_q_matrix_against_oracle(ctx):
    from wantibexos_amd import bse
    m, p, z = load_bands("hex6_g3_v2dk_c1v2")
    qp, _ = bse.kpathbse(m.rlat, z["kpath"], int(z["npts"]))
    q = qp[3]
    print("case: finite-Q matrix, q =", q, flush=True)
    with bse.BSEPlan(ctx, m, params_of(p), finite_q=True) as plan:
        plan.set_q(q)
        plan.build()
        H = np.tril(plan.fetch("hbse"))
        plan.execute()
        W = plan.energies()
    o = O.run_bands(m, p["ngrid"], q[None, :], **okw(p))
    assert np.abs(H - np.tril(o["H"][0])).max() < 1e-12
    assert np.abs(W - o["exk"][:, 0]).max() < 1e-10


@pytest.fixture(scope="module")
def big():
    from wantibexos_amd import tbmodel
    ng = (12, 12, 1)
    m = tbmodel.hex2d(seed=8, nw=10, ngrid=ng)
    p = dict(O.DEFAULT_PARAMS, ngrid=ng, coultype="V2DK", ediel=(1.5, 4.0, 2.0), lc=6.0)
    ks = [(0.0, 0.0, 0.0), (0.6667, -0.3333, 0.0), (0.6667, -0.3333, 0.0), (0.3333, 0.3333, 0.0)]
    qp, _ = O.kpathbse(m.rlat, np.array(ks), 4)
    return m, p, qp


def test_bands_n576(ctx, big):
    from wantibexos_amd import bse
    m, p, qp = big
    print("case: bands N=576, 8 q-points", flush=True)
    e1 = bse.bsebnds(ctx, m, params_of(p), qp, lanes=1)
    e4 = bse.bsebnds(ctx, m, params_of(p), qp, lanes=4)
    assert np.array_equal(e1, e4)
    o = O.run_bands(m, p["ngrid"], qp, **okw(p))
    assert np.abs(e4 - o["exk"]).max() < 1e-9
    em = bse.bsebnds(ctx, m, params_of(p, precision=bse.MIXED), qp, lanes=4)
    assert np.abs(em - e4).max() < 1e-3
    # bsebndstemp on the same path
    ft = dict(bset=True, temp=4000.0, ta="VE", st=0.3, phavg=0.05)
    et = bse.bsebnds(ctx, m, params_of(p, **ft), qp, lanes=4)
    ot = O.run_bands(m, p["ngrid"], qp, **okw(dict(p, **ft)))
    assert np.abs(et - ot["exk"]).max() < 1e-9
    assert np.abs(et - e4).max() > 1e-3                     # the temperature shows
    # the kept plans: a second path (the same points reversed) replays the captured graphs
    with bse.BSEBands(ctx, m, params_of(p), lanes=3) as b:
        r1 = b.run(qp)
        r2 = b.run(qp[::-1])
    assert np.array_equal(r1, e4)
    assert np.array_equal(r2, e4[:, ::-1])


def test_bands_iterative_form(ctx, big):
    from wantibexos_amd import bse
    m, p, qp = big
    print("case: bands N=576, Davidson at one q", flush=True)
    full = bse.bsebnds(ctx, m, params_of(p), qp[2:3], lanes=1)[:, 0]
    with bse.BSEPlan(ctx, m, params_of(p, precision=bse.MIXED), finite_q=True) as plan:
        plan.set_q(qp[2])
        r = plan.lowest(nev=32, guard=8, tol=1e-6, lanczos_steps=0)
    assert r["info"]["converged"] == 1
    assert np.abs(r["W"] - full[:32]).max() < 1e-4


def test_finite_q_plan_errors(ctx, big):
    """A finite-Q plan gives energies only; non-finite q and eps from Lanczos chains are
    refused on the host."""
    from wantibexos_amd import bse
    m, p, qp = big
    with bse.BSEPlan(ctx, m, params_of(p), finite_q=True) as plan:
        with pytest.raises(bse.WTBError):
            plan.set_q(np.array([0.0, np.nan, 0.0, 0.0]))
        plan.set_q(qp[1])
        plan.execute()
        with pytest.raises(bse.WTBError):
            plan.results()                     # act / eps / egap requested
        assert plan.energies().shape == (plan.N,)
    with bse.BSEPlan(ctx, m, params_of(p, precision=bse.MIXED), finite_q=True) as plan:
        plan.set_q(qp[1])
        with pytest.raises(bse.WTBError):
            plan.lowest(nev=8, guard=4, tol=1e-6, lanczos_steps=20)     # eps at finite Q
    with bse.BSEPlan(ctx, m, params_of(p)) as plan:
        with pytest.raises(bse.WTBError):
            plan.set_q(qp[0])                                            # not a finite-Q plan


def test_bands_exciton_vectors(ctx, big):
    """BSE_WF along the path: the vectors run_wf returns are orthonormal eigenvectors of
    each q's finite-Q matrix (the oracle's, residual 1e-9), with the last component real
    and >= 0, and equal in modulus to the oracle's eigenvectors where the exciton is
    non-degenerate."""
    from wantibexos_amd import bse
    m, p, qp = big
    sel = qp[[0, 2, 5]]
    with bse.BSEBands(ctx, m, params_of(p), lanes=2) as b:
        b.set_vectors(0, 4)
        exk, wf = b.run_wf(sel)
    o = O.run_bands(m, p["ngrid"], sel, **okw(p))
    assert np.abs(exk - o["exk"]).max() < 1e-9
    for iq in range(len(sel)):
        H = o["H"][iq]
        X = wf[iq]                                       # (N, 4)
        assert np.abs(X.conj().T @ X - np.eye(4)).max() < 1e-10
        assert np.abs(H @ X - X * exk[:4, iq]).max() < 1e-9
        assert (X[-1].real >= 0).all() and np.abs(X[-1].imag).max() == 0.0
        lam, V = np.linalg.eigh(H)
        for n in range(4):
            gap = min(abs(lam[n] - lam[k]) for k in range(len(lam)) if k != n)
            if gap > 1e-6:
                assert np.abs(np.abs(X[:, n]) - np.abs(V[:, n])).max() < 1e-8, (iq, n)
