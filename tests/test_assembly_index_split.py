"""CPU: the assembly's W-table index (k_bse_build in wantibexos_amd/csrc/k_bse_build.hip).

For basis rows on k-point k and columns on k-point k2 the table holds W at the grid
difference: index = ((d1 + n1 - 1) D2 + (d2 + n2 - 1)) D3 + (d3 + n3 - 1), D2 = 2 n2 - 1,
D3 = 2 n3 - 1, d = coords(k) - coords(k2), coords(k) = (k / (n2 n3), (k / n3) % n2, k % n3)
(the k ordering of mhkpack_subs.f90:359-407).  The kernel tabulates a row part RW(k) and a
column part CW(k2) once per tile and uses RW - CW.  Checked here against the per-element
formula on every k pair of random grids, including grids of one point along an axis.
"""
import numpy as np
import pytest


def coords(k, n1, n2, n3):
    return k // (n2 * n3), (k // n3) % n2, k % n3


def index_direct(k, k2, n):
    n1, n2, n3 = n
    D2, D3 = 2 * n2 - 1, 2 * n3 - 1
    a1, a2, a3 = coords(k, *n)
    b1, b2, b3 = coords(k2, *n)
    return ((a1 - b1 + n1 - 1) * D2 + (a2 - b2 + n2 - 1)) * D3 + (a3 - b3 + n3 - 1)


def index_split(k, k2, n):
    n1, n2, n3 = n
    D2, D3 = 2 * n2 - 1, 2 * n3 - 1
    a1, a2, a3 = coords(k, *n)
    b1, b2, b3 = coords(k2, *n)
    rw = (a1 * D2 + a2) * D3 + a3 + ((n1 - 1) * D2 + (n2 - 1)) * D3 + (n3 - 1)
    cw = (b1 * D2 + b2) * D3 + b3
    return rw - cw


GRIDS = [(3, 3, 1), (4, 4, 1), (36, 36, 1), (16, 16, 16), (1, 5, 7), (6, 1, 3), (2, 3, 4)]


@pytest.mark.parametrize("n", GRIDS)
def test_split_index_equals_direct_index(n):
    nk = n[0] * n[1] * n[2]
    rng = np.random.default_rng(sum(n))
    ks = np.arange(nk) if nk <= 400 else rng.choice(nk, 400, replace=False)
    k, k2 = np.meshgrid(ks, ks, indexing="ij")
    a = index_direct(k, k2, n)
    b = index_split(k, k2, n)
    assert np.array_equal(a, b)
    size = (2 * n[0] - 1) * (2 * n[1] - 1) * (2 * n[2] - 1)
    assert a.min() >= 0 and a.max() < size
    # k = k2 lands on the table's centre, W(0)
    assert (a[k == k2] == size // 2).all()
