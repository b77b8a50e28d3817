"""wtb_bse_spectrum_noalloc (the library's static scratch array instead of device
allocations) against wtb_bse_spectrum on the same inputs: bitwise equal results on two
calls in a row, the second with a smaller input (so stale scratch contents from the first
call would show), and a clear error beyond the scratch size."""
import numpy as np
import pytest

from wantibexos_amd import bse

pytestmark = pytest.mark.gpu


def inputs(rng, n):
    e = np.sort(rng.uniform(1.0, 5.0, n))
    act6 = np.abs(rng.standard_normal((6, n)))
    act6[3:] = rng.standard_normal((3, n)) * 0.1
    rlat = np.array([[3.2, 0.0, 0.0], [-1.6, 2.7713, 0.0], [0.0, 0.0, 20.0]])
    return e, act6, rlat


def test_noalloc_matches_allocating_entry_point():
    rng = np.random.default_rng(7)
    with bse.Context(0) as ctx:
        for n, numbse, ngrid in ((5000, 1001, (36, 36, 1)), (300, 201, (10, 10, 1))):
            e, act6, rlat = inputs(rng, n)
            print(f"spectrum noalloc n={n} numbse={numbse}", flush=True)
            want = bse.bse_spectrum(ctx, e, act6, rlat, ngrid, 0.0, 6.0, numbse, 0.05, renorm=True)
            got = bse.bse_spectrum_noalloc(ctx, e, act6, rlat, ngrid, 0.0, 6.0, numbse, 0.05, renorm=True)
            assert np.array_equal(got, want), n
        # beyond the scratch array: refused before any device work
        e, act6, rlat = inputs(rng, 150000)
        with pytest.raises(bse.WTBError, match="exceeds the scratch array"):
            bse.bse_spectrum_noalloc(ctx, e, act6, rlat, (36, 36, 1), 0.0, 6.0, 1001, 0.05)
