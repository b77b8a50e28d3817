"""CPU restatement (numpy, FP64) of the learned Davidson preconditioner: the
network (5 -> 128 -> 128 -> 1, tanh), its mean-squared-error gradient, the
per-state features and the training target.

Test infrastructure only: tests/ compare the GPU kernels (forward, backward,
features) against these functions.  The reference has no counterpart (its
eigensolve is a full ZHEEV, subprograms/bse_solver-tool-diel.f90:485-492), so
this module is pinned by mathematics instead: the gradient against central
finite differences of the loss, the target against a direct resolvent
solve, and the features against their definitions on small matrices.

Parameter pack: mu[5] sd[5] W1[128][5] b1[128] W2[128][128] b2[128] w3[128] b3.
"""
from __future__ import annotations

import numpy as np

H = 128
F = 5
SIGMA = 0.05     # eV, output scale and resolvent broadening
OFF_MU, OFF_SD, OFF_W1 = 0, 5, 10
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAMS = OFF_B3 + 1


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return dict(mu=P[OFF_MU:OFF_MU + F], sd=P[OFF_SD:OFF_SD + F],
                W1=P[OFF_W1:OFF_B1].reshape(H, F), b1=P[OFF_B1:OFF_W2],
                W2=P[OFF_W2:OFF_B2].reshape(H, H), b2=P[OFF_B2:OFF_W3],
                w3=P[OFF_W3:OFF_B3], b3=P[OFF_B3])


def init_params(rng, mu, sd):
    """Glorot-uniform weights, zero biases, the given input standardisation."""
    P = np.zeros(NPARAMS)
    P[OFF_MU:OFF_MU + F] = mu
    P[OFF_SD:OFF_SD + F] = sd
    a1 = np.sqrt(6.0 / (F + H))
    a2 = np.sqrt(6.0 / (H + H))
    a3 = np.sqrt(6.0 / (H + 1))
    P[OFF_W1:OFF_B1] = rng.uniform(-a1, a1, H * F)
    P[OFF_W2:OFF_B2] = rng.uniform(-a2, a2, H * H)
    P[OFF_W3:OFF_B3] = rng.uniform(-a3, a3, H)
    return P


def forward(P, X):
    """y for raw features X (5, S)."""
    q = unpack(P)
    xn = (np.asarray(X, dtype=np.float64).T - q["mu"]) / q["sd"]
    h1 = np.tanh(xn @ q["W1"].T + q["b1"])
    h2 = np.tanh(h1 @ q["W2"].T + q["b2"])
    return h2 @ q["w3"] + q["b3"]


def loss_grad(P, X, t):
    """mean((y - t)^2) and its gradient in the parameter pack (zero for mu, sd)."""
    q = unpack(P)
    t = np.asarray(t, dtype=np.float64)
    S = t.size
    xn = (np.asarray(X, dtype=np.float64).T - q["mu"]) / q["sd"]
    h1 = np.tanh(xn @ q["W1"].T + q["b1"])
    h2 = np.tanh(h1 @ q["W2"].T + q["b2"])
    y = h2 @ q["w3"] + q["b3"]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / S
    g = np.zeros(NPARAMS)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    da2 = dy[:, None] * q["w3"][None, :] * (1.0 - h2 * h2)
    g[OFF_W2:OFF_B2] = (da2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = da2.sum(0)
    da1 = (da2 @ q["W2"]) * (1.0 - h1 * h1)
    g[OFF_W1:OFF_B1] = (da1.T @ xn).ravel()
    g[OFF_B1:OFF_W2] = da1.sum(0)
    return loss, g


def features(D, stats, theta, Dmin, N):
    """(5, S) features for every basis state i and every theta (i fastest)."""
    D = np.asarray(D, dtype=np.float64)
    st = np.asarray(stats, dtype=np.float64)
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    out = np.empty((5, theta.size, D.size))
    out[0] = D[None, :] - theta[:, None]
    out[1] = st[0][None, :]
    out[2] = (st[1] * np.sqrt(N))[None, :]
    out[3] = (st[2] * N)[None, :]
    out[4] = (theta - Dmin)[:, None]
    return out.reshape(5, -1)


def row_stats(Hfull):
    """D_i = Re H_ii and over j != i: sum |H_ij|, sqrt(sum |H_ij|^2), max |H_ij|."""
    A = np.abs(np.asarray(Hfull))
    np.fill_diagonal(A, 0.0)
    return np.real(np.diag(Hfull)).copy(), np.stack([A.sum(1), np.sqrt((A * A).sum(1)), A.max(1)])


def target_correction(lam, X, theta, D, eta=SIGMA):
    """The network's training target: target() minus sigma g(D_i - theta), g(d) = d/(d^2+eta^2)."""
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    dd = np.asarray(D, dtype=np.float64)[None, :] - theta[:, None]
    return target(lam, X, theta, eta) - (SIGMA * dd / (dd * dd + eta * eta)).reshape(-1)


def target(lam, X, theta, eta=SIGMA):
    """sigma Re[(H - theta - i eta)^-1]_ii from the eigenpairs (lam, X), (S,) with i fastest."""
    lam = np.asarray(lam, dtype=np.float64)
    W2 = np.abs(np.asarray(X)) ** 2                       # (N, N): |X_in|^2
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    d = lam[:, None] - theta[None, :]                     # (n, theta)
    g = d / (d * d + eta * eta)
    return (SIGMA * (W2 @ g)).T.reshape(-1)
