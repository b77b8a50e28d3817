"""The post-processing of the spectra (wtb_optprop) and the exciton lifetimes
(wtb_exciton_lifetime) on the GPU against the oracle, on the inputs of the reference
program's own files (tests/golden/postproc): the same operation order and
single-precision constants, so the results agree to the last bit."""
import json
from pathlib import Path

import numpy as np
import pytest

import wtb_oracle as O
from wantibexos_amd import bse

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "postproc"
CASES = sorted(p.stem for p in GOLD.glob("*.npz"))
COLS = ["xx", "yy", "zz", "xy", "xz", "yz"]
POL = ["x", "y", "z", "sp", "sm"]


@pytest.fixture(scope="module")
def ctx():
    with bse.Context(0) as c:
        yield c


def same(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("pol", [False, True], ids=["bse", "pol"])
def test_optprop_bitwise_oracle(ctx, name, pol):
    z = np.load(GOLD / f"{name}.npz")
    names, key = (POL, "dielpol_") if pol else (COLS, "diel_")
    en = z[key + names[0]][:, 0]
    e1 = np.stack([z[key + c][:, 1] for c in names], axis=1)
    e2 = np.stack([z[key + c][:, 2] for c in names], axis=1)
    print(f"optprop {name} pol={pol} n={len(en)}", flush=True)
    got = bse.optprop(ctx, en, e1, e2)
    want = O.optprop(en, e1, e2)
    for g, w in zip(got, want):
        assert same(g, w)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("pol", [False, True], ids=["lifetime", "lifetimepol"])
def test_lifetime_bitwise_oracle(ctx, name, pol):
    z = np.load(GOLD / f"{name}.npz")
    p = json.loads(str(z["params"]))
    tab = z["opt_diel_pol" if pol else "opt_diel"]
    nout = 5 if pol else 3
    print(f"lifetime {name} pol={pol} N={len(tab)}", flush=True)
    got = bse.exciton_lifetime(ctx, p["sysdim"], p["ngrid"], z["rlat"], tab, nout)
    want = O.lifetimes(p["sysdim"], p["ngrid"], z["rlat"], tab, nout)
    assert same(got, want)


def test_lifetime_rejects_bad_sysdim(ctx):
    z = np.load(GOLD / f"{CASES[0]}.npz")
    with pytest.raises(bse.WTBError, match="Wrong System dimension"):
        bse.exciton_lifetime(ctx, "4D", (3, 3, 1), z["rlat"], z["opt_diel"], 3)
