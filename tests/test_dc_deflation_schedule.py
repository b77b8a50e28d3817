"""CPU: the schedule of the divide and conquer's deflation (k_dc_deflate in
wantibexos_amd/csrc/k_dc.hip) against the sequential scan it replaces.

The sequential scan (DLAED2's rule) visits the sorted positions in order: a small weight
deflates at once; a close pair of kept candidates (pj, nj) is rotated so that pj deflates
and the rotated nj becomes the next pair's pj.  The kernel finds the same decisions in
three parts: the small-weight and pair tests on the original values in parallel, a walk
over the pairs that skips 64 at a time while no test passes and re-tests the pair after a
rotation with the rotated values, and slots from prefix counts (kept positions first, in
order; deflated ones from the end, in the order of deflation).  Both are emulated here
with the same floating-point operations and compared exactly, on inputs built to have
long rotation chains, small weights between chain members and chains that a test on the
original values alone would miss.
"""
import numpy as np
import pytest

EPS = 1.1102230246251565e-16


def pair_test(dp, zp, dn, zn, tol):
    s, c = zp, zn
    tau = np.hypot(c, s)
    t = dn - dp
    c = c / tau
    s = -s / tau
    return abs(t * c * s) <= tol, c, s, tau


def sequential(D, Z, rho):
    D, Z = D.copy(), Z.copy()
    n = len(D)
    tol = 8 * EPS * max(np.abs(D).max(), np.abs(Z).max())
    slot = np.full(n, -1)
    cD = np.zeros(n)
    cw = np.zeros(n)
    rots = []
    K = nd = 0

    def deflate(i):
        nonlocal nd
        s = n - 1 - nd
        nd += 1
        slot[i] = s
        cD[s] = D[i]

    def keep(i):
        nonlocal K
        slot[i] = K
        cD[K] = D[i]
        cw[K] = Z[i]
        K += 1

    if rho * np.abs(Z).max() <= tol:
        for i in range(n):
            deflate(i)
        return slot, cD, cw, rots, 0
    pj = -1
    for j in range(n):
        if rho * abs(Z[j]) <= tol:
            deflate(j)
            continue
        if pj < 0:
            pj = j
            continue
        ok, c, s, tau = pair_test(D[pj], Z[pj], D[j], Z[j], tol)
        if ok:
            Z[j] = tau
            Z[pj] = 0.0
            dp, dn = D[pj], D[j]
            t2 = dp * c * c + dn * s * s
            D[j] = dp * s * s + dn * c * c
            D[pj] = t2
            rots.append((pj, j, c, s))
            deflate(pj)
        else:
            keep(pj)
        pj = j
    keep(pj)
    return slot, cD, cw, rots, K


def parallel(D, Z, rho, width=64):
    D, Z = D.copy(), Z.copy()
    n = len(D)
    tol = 8 * EPS * max(np.abs(D).max(), np.abs(Z).max())
    slot = np.full(n, -1)
    cD = np.zeros(n)
    cw = np.zeros(n)
    if rho * np.abs(Z).max() <= tol:
        for i in range(n):
            slot[i] = n - 1 - i
            cD[n - 1 - i] = D[i]
        return slot, cD, cw, [], 0
    small = rho * np.abs(Z) <= tol
    P = np.flatnonzero(~small)
    m = len(P)
    pass0 = np.zeros(m, bool)                       # the pair ending at P[k], original values
    for k in range(1, m):
        pass0[k] = pair_test(D[P[k - 1]], Z[P[k - 1]], D[P[k]], Z[P[k]], tol)[0]
    rot_at = np.zeros(n, bool)
    rot_pj = np.full(n, -1)
    pj_defl = np.zeros(n, bool)
    rots = []
    k, kmod, dmod, zmod = 1, False, 0.0, 0.0
    while k < m:
        if not kmod:
            window = pass0[k:k + width]
            if not window.any():
                k += width
                continue
            k += int(np.argmax(window))
        pj, nj = P[k - 1], P[k]
        dp = dmod if kmod else D[pj]
        zp = zmod if kmod else Z[pj]
        dn = D[nj]
        ok, c, s, tau = pair_test(dp, zp, dn, Z[nj], tol)
        if ok:
            t2 = dp * c * c + dn * s * s
            dnew = dp * s * s + dn * c * c
            Z[nj], Z[pj], D[nj], D[pj] = tau, 0.0, dnew, t2
            rot_at[nj], rot_pj[nj], pj_defl[pj] = True, pj, True
            rots.append((pj, nj, c, s))
            kmod, dmod, zmod = True, dnew, tau
        else:
            kmod = False
        k += 1
    kept = ~small & ~pj_defl
    ev = small | rot_at
    kpre = np.cumsum(kept) - kept
    epre = np.cumsum(ev) - ev
    for i in range(n):
        if kept[i]:
            s = kpre[i]
            slot[i], cD[s], cw[s] = s, D[i], Z[i]
        if ev[i]:
            e = i if small[i] else rot_pj[i]
            s = n - 1 - epre[i]
            slot[e], cD[s] = s, D[e]
    return slot, cD, cw, rots, int(kept.sum())


def chains_case(seed, n=400):
    """sorted poles with clusters of near-equal values (rotation chains), weights of
    mixed sizes and a few negligible ones"""
    rng = np.random.default_rng(seed)
    D = np.sort(rng.uniform(-3.0, 3.0, n))
    for start in rng.choice(n - 12, 12, replace=False):
        length = int(rng.integers(2, 10))
        D[start:start + length] = D[start] + np.arange(length) * 1e-17 * rng.uniform(0.1, 3.0)
    D = np.sort(D)
    Z = rng.standard_normal(n) * 10.0 ** rng.uniform(-8, 0, n)
    Z[rng.choice(n, 20, replace=False)] = 1e-19
    return D, Z / np.linalg.norm(Z)


def hidden_chain_case(seed, n=300):
    """poles i, i+1 tied with weights a, then pole i+2 at offset delta with weight b << a:
    the pair (i+1, i+2) fails the test on the original values (delta * a b / (a^2 + b^2)
    > tol) and passes once the tie's rotation has merged the weights into a sqrt(2)
    (delta * a sqrt(2) b / (2 a^2 + b^2) <= tol)"""
    rng = np.random.default_rng(100 + seed)
    D = np.sort(rng.uniform(-3.0, 3.0, n))
    Z = rng.uniform(0.5, 1.0, n) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    Z /= np.linalg.norm(Z)
    D[0], D[-1] = -3.0, 3.0
    a, b = 1e-3, 1e-8
    tol = 8 * EPS * max(np.abs(D).max(), np.abs(Z).max(), a)
    lo, hi = tol * (a * a + b * b) / (a * b), tol * (2 * a * a + b * b) / (np.sqrt(2) * a * b)
    for i in rng.choice(np.arange(5, n - 5, 6), 12, replace=False):
        D[i + 1] = D[i]
        D[i + 2] = D[i] + 0.5 * (lo + hi)
        Z[i], Z[i + 1], Z[i + 2] = a, a, b
    assert np.all(np.diff(D) >= 0)
    return D, Z


CASES = [("chains", s) for s in range(8)] + [("hidden", s) for s in range(4)]


def make(kind, seed):
    return chains_case(seed) if kind == "chains" else hidden_chain_case(seed)


@pytest.mark.parametrize("kind,seed", CASES)
@pytest.mark.parametrize("width", [64, 3])
def test_parallel_walk_matches_the_sequential_scan(kind, seed, width):
    D, Z = make(kind, seed)
    rho = 0.7
    a = sequential(D, Z, rho)
    b = parallel(D, Z, rho, width)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] and a[4] == b[4]
    assert sorted(a[0]) == list(range(len(D)))


def test_cases_exercise_chains_beyond_the_first_test():
    """some rotation in the cases is due only after a preceding rotation (its pair fails the
    test on the original values), so the re-test with rotated values is what finds it"""
    found = 0
    for kind, seed in CASES:
        D, Z = make(kind, seed)
        _, _, _, rots, _ = sequential(D, Z, 0.7)
        tol = 8 * EPS * max(np.abs(D).max(), np.abs(Z).max())
        for pj, nj, _, _ in rots:
            found += not pair_test(D[pj], Z[pj], D[nj], Z[nj], tol)[0]
    assert found > 0


def test_all_deflated_and_no_rotation_cases():
    D = np.linspace(-1, 1, 50)
    Z = np.full(50, 1e-30)
    for f in (sequential, parallel):
        slot, cD, _, rots, K = f(D, Z, 1.0)
        assert K == 0 and not rots and np.array_equal(slot, 49 - np.arange(50))
    Z = np.ones(50) / np.sqrt(50)
    a, b = sequential(D, Z, 1.0), parallel(D, Z, 1.0)
    assert a[4] == b[4] == 50 and np.array_equal(a[0], b[0])
