"""CPU: the ceilings of what a learned preconditioner could contribute.

The network learns the regularised resolvent diagonal Re[(H - theta - i eta)^-1]_ii
(eta = 0.05 eV) as a diagonal preconditioner for block Davidson.  Here that quantity is
used exactly (from a full eigendecomposition of an oracle-built BSE matrix) inside a
plain numpy block Davidson, against the diagonal preconditioner 1/(D_i - theta): if the
exact target brings no fewer iterations, no network trained on it can, which is what the
GPU runs of the trained network show (more iterations, from its approximation error).
The same holds for choosing the coarse block's states: the exact exciton weights pick
no better set than the lowest diagonal energies.
"""
import numpy as np

import wtb_oracle as O
from wantibexos_amd import tbmodel


def block_davidson(H, nev, guard, tol, precond, max_it=60):
    """Thick-restart block Davidson (basis up to 5 (nev + guard)); returns the iteration count."""
    N = H.shape[0]
    nk = nev + guard
    D = np.real(np.diag(H))
    V = np.zeros((N, nk), complex)
    V[np.argsort(D)[:nk], np.arange(nk)] = 1.0
    HV = H @ V
    for it in range(max_it):
        G = V.conj().T @ HV
        th, S = np.linalg.eigh((G + G.conj().T) / 2)
        th, S = th[:nk], S[:, :nk]
        U, HU = V @ S, HV @ S
        R = HU - U * th
        rn = np.linalg.norm(R, axis=0)
        if rn[:nev].max() <= tol:
            return it + 1
        T = np.stack([precond(R[:, c], th[c], D) for c in np.where(rn > tol)[0]], axis=1)
        if V.shape[1] + T.shape[1] > 5 * nk:
            V, HV = U, HU
        for _ in range(2):
            T -= V @ (V.conj().T @ T)
        Q, Rr = np.linalg.qr(T)
        Q = Q[:, np.abs(np.diag(Rr)) > 1e-10 * np.abs(np.diag(Rr)).max()]
        Q -= V @ (V.conj().T @ Q)
        Q, _ = np.linalg.qr(Q)
        V, HV = np.hstack([V, Q]), np.hstack([HV, H @ Q])
    return None


def test_exact_resolvent_diagonal_is_no_better_than_the_diagonal():
    m = tbmodel.hex2d(seed=1, nw=8, ngrid=(12, 12, 1))
    r = O.run_pipeline(m, (12, 12, 1), r0=40.0, f154=False)
    H = np.tril(r["H"]) + np.tril(r["H"], -1).conj().T          # N = 576
    lam, X = np.linalg.eigh(H)
    W2 = np.abs(X) ** 2
    eta = 0.05

    def plain(res, th, D):
        d = D - th
        d = np.where(np.abs(d) < 1e-3, np.where(d < 0, -1e-3, 1e-3), d)
        return res / d

    def exact_target(res, th, D):
        dl = lam - th
        return res * (W2 @ (dl / (dl * dl + eta * eta)))

    it_plain = block_davidson(H, 32, 8, 1e-6, plain)
    it_exact = block_davidson(H, 32, 8, 1e-6, exact_target)
    print("iterations: plain diagonal", it_plain, "exact resolvent diagonal", it_exact)
    assert it_plain is not None and it_exact is not None
    assert it_exact >= it_plain - 1


def test_exciton_weight_coarse_set_is_no_better_than_lowest_diagonal():
    """The coarse block's states: the lowest diagonal energies against the states that
    carry the most weight in the lowest excitons (sum_n<nev |X_in|^2, exact): the best a
    learned selector could pick.  Same iteration count, so the simple rule stays."""
    m = tbmodel.hex2d(seed=7, nw=10, ngrid=(14, 14, 1))
    r = O.run_pipeline(m, (14, 14, 1), r0=15.0, f154=False)
    H = np.tril(r["H"]) + np.tril(r["H"], -1).conj().T          # N = 784
    D = np.real(np.diag(H))
    lam, X = np.linalg.eigh(H)
    nev, guard, M = 32, 8, 160
    w = (np.abs(X[:, :nev]) ** 2).sum(1)

    def coarse(S):
        lamS, QS = np.linalg.eigh(H[np.ix_(S, S)])

        def P(res, th, Dd):
            d = Dd - th
            t = res / np.where(np.abs(d) < 1e-3, np.where(d < 0, -1e-3, 1e-3), d)
            dl = lamS - th
            dl = np.where(np.abs(dl) < 1e-2, np.where(dl < 0, -1e-2, 1e-2), dl)
            t[S] = QS @ ((QS.conj().T @ res[S]) / dl)
            return t
        return P

    it_d = block_davidson(H, nev, guard, 1e-6, coarse(np.argsort(D)[:M]))
    it_w = block_davidson(H, nev, guard, 1e-6, coarse(np.argsort(-w)[:M]))
    print("iterations: lowest-D coarse set", it_d, "exciton-weight coarse set", it_w)
    assert it_d is not None and it_w is not None
    assert it_w >= it_d - 1
