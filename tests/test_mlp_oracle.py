"""CPU: pins the oracle of the learned preconditioner (oracle/mlp_oracle.py) by
mathematics, and checks the package's training-data functions against it.
- gradient vs central finite differences of the loss (FP64);
- target vs the diagonal of an explicitly inverted (H - theta - i eta);
- row statistics and features vs their definitions on a small matrix."""
import numpy as np

import mlp_oracle as M
from wantibexos_amd import learned


def small_net(rng):
    mu = rng.normal(size=5)
    sd = rng.uniform(0.5, 2.0, 5)
    P = M.init_params(rng, mu, sd)
    P[M.OFF_B1:M.OFF_W2] = rng.normal(scale=0.1, size=M.H)
    P[M.OFF_B2:M.OFF_W3] = rng.normal(scale=0.1, size=M.H)
    P[M.OFF_B3] = 0.3
    return P


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = small_net(rng)
    X = rng.normal(size=(5, 40))
    t = rng.normal(size=40)
    loss, g = M.loss_grad(P, X, t)
    assert np.allclose(loss, np.mean((M.forward(P, X) - t) ** 2))
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 20, replace=False),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 10, replace=False),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 40, replace=False),
                          rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 10, replace=False),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 10, replace=False), [M.OFF_B3]])
    h = 1e-6
    for k in idx:
        Pp = P.copy()
        Pm = P.copy()
        Pp[k] += h
        Pm[k] -= h
        fd = (M.loss_grad(Pp, X, t)[0] - M.loss_grad(Pm, X, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), k
    assert np.all(g[:10] == 0.0)


def test_target_is_the_regularised_resolvent_diagonal():
    rng = np.random.default_rng(5)
    n = 30
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = (A + A.conj().T) / 2
    lam, X = np.linalg.eigh(A)
    theta = np.array([lam[0] - 0.1, lam[3] + 0.01, 0.5 * (lam[5] + lam[6])])
    t = M.target(lam, X, theta).reshape(theta.size, n)
    for k, th in enumerate(theta):
        R = np.linalg.inv(A - (th + 1j * M.SIGMA) * np.eye(n))
        assert np.allclose(t[k], M.SIGMA * np.real(np.diag(R)), atol=1e-12)
    D = np.real(np.diag(A))
    tc = M.target_correction(lam, X, theta, D).reshape(theta.size, n)
    for k, th in enumerate(theta):
        d = D - th
        assert np.allclose(tc[k], t[k] - M.SIGMA * d / (d * d + M.SIGMA ** 2), atol=1e-12)
    assert np.allclose(learned.targets(lam, X, theta, D), tc.reshape(-1), rtol=1e-5, atol=1e-6)


def test_features_and_row_stats():
    rng = np.random.default_rng(6)
    n = 12
    A = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    A = (A + A.conj().T) / 2
    D, st = M.row_stats(A)
    i = 4
    off = np.abs(np.delete(A[i], i))
    assert np.isclose(st[0, i], off.sum()) and np.isclose(st[1, i], np.sqrt((off ** 2).sum()))
    assert np.isclose(st[2, i], off.max()) and np.isclose(D[i], A[i, i].real)
    theta = np.array([-1.0, 0.5])
    f = M.features(D, st, theta, D.min(), n).reshape(5, 2, n)
    assert np.allclose(f[0, 1], D - 0.5) and np.allclose(f[2, 0], st[1] * np.sqrt(n))
    assert np.allclose(f[3, 1], st[2] * n) and np.allclose(f[4, 0], -1.0 - D.min())
    assert np.allclose(learned.features(D, st, theta, D.min(), n), f.reshape(5, -1), rtol=1e-6)
