"""Training of the learned Davidson preconditioner (the network is evaluated and
trained by the library's own GPU kernels, k_mlp.hip).

Training data are BSE matrices the pipeline itself generates: seeded 2D
hexagonal models with r0 drawn uniformly from 10-80 A.  For each matrix the
full FP64 eigendecomposition (GPU, ZHEEV path with vectors) gives the target

    t_i(theta) = sigma (Re[(H - theta - i eta)^-1]_ii - g(D_i - theta)),   g(d) = d / (d^2 + eta^2),
    Re[(H - theta - i eta)^-1]_ii = sum_n |X_in|^2 g(l_n - theta),

at theta equal to one of the lowest exciton energies l_k (eta = sigma = 0.05 eV), where
the k-th term vanishes (its Lorentzian real part is zero at the pole) so that t is the
regularised diagonal of the reduced resolvent sum_{n != k} x_n x_n^H / (l_n - l_k), the
operator of the exact (Jacobi-Davidson) correction, minus the regularised diagonal
preconditioner g(D_i - theta) that the network corrects; and the features of every
basis state come from the matrix's diagonal and row
statistics (computed on the GPU).  The feature definition is the one of the
preconditioner kernel:

    f0 = D_i - theta, f1 = sum_j |H_ij|, f2 = sqrt(N sum_j |H_ij|^2), f3 = N max_j |H_ij|,
    f4 = theta - min_i D_i.
"""
from __future__ import annotations

import dataclasses
from pathlib import Path

import numpy as np

from . import bse, tbmodel

SIGMA = 0.05
HIDDEN = 128
NPARAMS = 10 + 5 * HIDDEN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1
WEIGHTS = Path(__file__).resolve().parent / "precond_mlp.npz"


def features(D, stats, theta, Dmin, N):
    """(5, S) features, basis state fastest, for each theta in turn."""
    D = np.asarray(D, dtype=np.float64)
    st = np.asarray(stats, dtype=np.float64)
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    f = np.empty((5, theta.size, D.size))
    f[0] = D[None, :] - theta[:, None]
    f[1] = st[0][None, :]
    f[2] = (st[1] * np.sqrt(N))[None, :]
    f[3] = (st[2] * N)[None, :]
    f[4] = (theta - Dmin)[:, None]
    return f.reshape(5, -1).astype(np.float32)


def targets(lam, X, theta, D, eta=SIGMA):
    """sigma (Re[(H - theta - i eta)^-1]_ii - g(D_i - theta)), basis state fastest, for each theta."""
    w2 = np.abs(X) ** 2
    theta = np.atleast_1d(np.asarray(theta, dtype=np.float64))
    d = np.asarray(lam)[:, None] - theta[None, :]
    g = d / (d * d + eta * eta)
    dd = np.asarray(D)[None, :] - theta[:, None]
    reg = dd / (dd * dd + eta * eta)
    return (SIGMA * ((w2 @ g).T - reg)).reshape(-1).astype(np.float32)


def init_params(rng, mu, sd):
    P = np.zeros(NPARAMS, dtype=np.float32)
    P[0:5] = mu
    P[5:10] = sd
    o = 10
    a1, a2, a3 = np.sqrt(6.0 / (5 + HIDDEN)), np.sqrt(6.0 / (2 * HIDDEN)), np.sqrt(6.0 / (HIDDEN + 1))
    P[o:o + 5 * HIDDEN] = rng.uniform(-a1, a1, 5 * HIDDEN)
    o += 5 * HIDDEN + HIDDEN
    P[o:o + HIDDEN * HIDDEN] = rng.uniform(-a2, a2, HIDDEN * HIDDEN)
    o += HIDDEN * HIDDEN + HIDDEN
    P[o:o + HIDDEN] = rng.uniform(-a3, a3, HIDDEN)
    return P


@dataclasses.dataclass
class TrainingSet:
    X: np.ndarray          # (5, S) features
    t: np.ndarray          # (S,) targets
    n_models: int
    dims: list


def model_params(r0, ngrid=(12, 12, 1)):
    return bse.BSEParams(ngrid=ngrid, nc=2, nv=2, coultype="V2DRK", r0=float(r0), cshift=0.05, ebse0=0.0,
                         ebsef=6.0, numbse=501, dtfull=True, cpol=True, f154=False)


def sample_matrix(ctx, model, prm, rng, n_theta=8, n_low=64):
    """Features and targets of one generated BSE matrix."""
    p64 = dataclasses.replace(prm, precision=bse.FP64)
    p32 = dataclasses.replace(prm, precision=bse.MIXED)
    with bse.BSEPlan(ctx, model, p32) as plan:
        D, st = plan.row_stats()
    with bse.BSEPlan(ctx, model, p64) as plan:
        plan.build()
        H = plan.fetch("hbse")
    N = H.shape[0]
    Hf = np.tril(H) + np.tril(H, -1).conj().T
    lam, X = bse.zheev(ctx, Hf, vectors=True)
    # theta at an exciton energy: g(0) = 0 removes that exciton's own pole, so the target is
    # the regularised diagonal of the reduced resolvent sum_{n != k} x_n x_n^H / (l_n - l_k),
    # the operator of the exact (Jacobi-Davidson) correction
    ks = rng.choice(min(n_low, N), size=min(n_theta, N), replace=False)
    theta = lam[ks]
    return features(D, st, theta, D.min(), N), targets(lam, X, theta, D), N


def make_training_set(ctx, n_models=200, seed=1234, nw_range=(6, 16), ngrid=(12, 12, 1), n_theta=8):
    rng = np.random.default_rng(seed)
    Xs, ts, dims = [], [], []
    for i in range(n_models):
        nw = int(rng.integers(nw_range[0], nw_range[1] + 1))
        nw += nw % 2
        model = tbmodel.hex2d(seed=10_000 + i, nw=nw, ngrid=ngrid)
        prm = model_params(rng.uniform(10.0, 80.0), ngrid)
        X, t, N = sample_matrix(ctx, model, prm, rng, n_theta)
        Xs.append(X)
        ts.append(t)
        dims.append(N)
    return TrainingSet(np.concatenate(Xs, axis=1), np.concatenate(ts), n_models, dims)


def train(ctx, data: TrainingSet, epochs=20, batch=4096, lr=2e-3, seed=7):
    """Train on the GPU; returns (params, loss per epoch)."""
    rng = np.random.default_rng(seed)
    mu = data.X.mean(axis=1)
    sd = data.X.std(axis=1) + 1e-6
    P0 = init_params(rng, mu, sd)
    with bse.MLP(ctx) as net:
        net.set_params(P0)
        hist = net.train(data.X, data.t, epochs=epochs, batch=batch, lr=lr, seed=seed)
        return net.get_params(), hist


def save(P, path=WEIGHTS, **meta):
    np.savez(path, params=np.asarray(P, dtype=np.float32), **meta)


def load(path=WEIGHTS):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32)


def main(argv=None):
    """python -m wantibexos_amd.learned [--models 200] [--epochs 20] [--out path]"""
    import argparse
    import sys
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", default=str(WEIGHTS))
    a = ap.parse_args(argv)
    with bse.Context(0) as ctx:
        t0 = time.perf_counter()
        data = make_training_set(ctx, n_models=a.models, seed=a.seed)
        t1 = time.perf_counter()
        print(f"training set: {data.X.shape[1]} samples from {data.n_models} matrices "
              f"(N = {min(data.dims)}..{max(data.dims)}) in {t1 - t0:.1f} s", file=sys.stderr, flush=True)
        P, hist = train(ctx, data, epochs=a.epochs)
        t2 = time.perf_counter()
        print(f"trained {a.epochs} epochs in {t2 - t1:.1f} s; loss {hist[0]:.5f} -> {hist[-1]:.5f}; "
              f"target variance {float(np.var(data.t)):.5f}", file=sys.stderr, flush=True)
        save(P, a.out, loss=hist, n_models=data.n_models, seed=a.seed, target_var=float(np.var(data.t)))


if __name__ == "__main__":
    main()
