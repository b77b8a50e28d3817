"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import json
from pathlib import Path

import numpy as np

from wantibexos_amd.tbmodel import TBModel

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name):
    z = np.load(GOLDEN / f"{name}.npz")
    p = json.loads(str(z["params"]))
    p["ngrid"] = tuple(p["ngrid"])
    p["ediel"] = tuple(p["ediel"])
    p["shift"] = tuple(p["shift"])
    m = TBModel(str(z["model_systype"]), float(z["model_scs"]), float(z["model_efermi"]),
                z["model_rlat"], z["model_rvec"], z["model_ndeg"], z["model_hop"], z["model_ihop"])
    return m, p, z


def oracle_kwargs(p):
    import wtb_oracle as O
    return {k: p[k] for k in O.DEFAULT_PARAMS if k in p}


def ref_signs(z):
    """Sign of the (real) orbital-1 component of every reference band vector, and the
    product sign_c*sign_v per BSE basis state: the reference's ZHEEV gauge is
    'component 1 real' with either sign; the oracle and the GPU use >= 0."""
    x1 = z["vector"][:, :, 0]
    s = np.where(x1.real < 0, -1.0, 1.0)
    stt = z["stt"]
    k = stt[:, 3] - 1
    return s, s[k, stt[:, 1] - 1] * s[k, stt[:, 2] - 1]


def unpack_upper(packed, n):
    """Packed-by-columns upper triangle (i<=j) -> full Hermitian matrix."""
    H = np.zeros((n, n), dtype=np.complex128)
    jj, ii = [], []
    for j in range(n):
        ii.extend(range(j + 1))
        jj.extend([j] * (j + 1))
    ii = np.array(ii)
    jj = np.array(jj)
    H[ii, jj] = packed
    H = H + np.triu(H, 1).conj().T
    return H
