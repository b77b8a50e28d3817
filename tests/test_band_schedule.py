"""CPU: the schedule of the bulge chasing (k_chase in wantibexos_amd/csrc/k_band.hip).

Task (s, k) of sweep s waits only until sweep s - 1 has finished its tasks 0..k+1; the
later tasks of earlier sweeps may run at the same time.  Two checks pin that rule:

* the index sets: from the kernel's own formulas for a task's rows (st..ed, and the
  previous reflector's rows sp..sp+lenp-1 for k >= 1), the band elements and tracked-vector
  rows of (s, k) are disjoint from those of every task (s - j, k') with k' >= k + j + 1,
  all that can still be running;
* the arithmetic: a numpy emulation of the sweeps (the same ZLARFG and two-sided updates on
  a dense Hermitian band matrix) gives bitwise the sequential result under random
  interleavings that respect the rule, while a rule one task looser (waiting only for
  task k) changes the result.
"""
import numpy as np
import pytest


def ntasks(N, b, s):
    return 1 + (N - 2 - s) // b


def task_rows(N, b):
    """(s, k) -> (st, ed, sp, lenp) as the kernel computes them"""
    out = {}
    for s in range(N - 1):
        lenp = 0
        for k in range(ntasks(N, b, s)):
            sp = s + 1 + (k - 1) * b
            st = s + 1 if k == 0 else sp + lenp
            ed = min(st + b - 1, N - 1)
            out[(s, k)] = (st, ed, sp, lenp)
            lenp = ed - st + 1
    return out


def touched(N, b):
    """(s, k) -> (band elements (row >= col), tracked rows)"""
    acc = {}
    for (s, k), (st, ed, sp, lenp) in task_rows(N, b).items():
        ln = ed - st + 1
        els = set()
        for r in range(ln):
            for c in range(ln):
                els.add((st + max(r, c), st + min(r, c)))
            if k == 0:
                els.add((st + r, s))
            else:
                for c in range(lenp):
                    els.add((st + r, sp + c))
        for (R, C) in els:
            assert 0 <= R - C <= 2 * b and R < N      # inside the band storage
        acc[(s, k)] = (els, set(range(st, ed + 1)))
    return acc


@pytest.mark.parametrize("N,b", [(70, 16), (100, 32), (97, 16), (130, 32), (33, 16), (2, 16)])
def test_concurrent_tasks_touch_disjoint_data(N, b):
    acc = touched(N, b)
    for (s, k), (e1, r1) in acc.items():
        for j in range(1, 5):
            if s - j < 0:
                break
            for kp in range(k + j + 1, ntasks(N, b, s - j)):
                e2, r2 = acc[(s - j, kp)]
                assert not (e1 & e2) and not (r1 & r2), (s, k, s - j, kp)
    # and the rule is tight: the task it waits for does overlap
    hits = sum(1 for (s, k), (e1, r1) in acc.items() if s > 0 and k + 1 < ntasks(N, b, s - 1)
               and (e1 & acc[(s - 1, k + 1)][0]))
    assert hits > 0 or N < 2 * b + 2


def larfg(alpha, xn2):
    if xn2 == 0.0 and alpha.imag == 0.0:
        return alpha, 0j, 0j
    nrm = np.sqrt(abs(alpha) ** 2 + xn2)
    beta = -nrm if alpha.real >= 0 else nrm
    return beta, complex((beta - alpha.real) / beta, -alpha.imag / beta), 1.0 / (alpha - beta)


def run_task(A, gc, state, s, k, rows):
    st, ed, sp, lenp = rows[(s, k)]
    ln = ed - st + 1
    if k == 0:
        x = A[st:ed + 1, s].copy()
    else:
        _, vprev, taup = state[s]
        O = A[st:ed + 1, sp:sp + lenp].copy()
        O -= taup * np.outer(O @ vprev, vprev.conj())
        x = O[:, 0].copy()
    beta, tau, scale = larfg(x[0], float(np.sum(np.abs(x[1:]) ** 2)))
    v = np.zeros(ln, complex)
    v[0] = 1.0
    v[1:] = x[1:] * scale
    if k == 0:
        A[st:ed + 1, s] = 0.0
        A[st, s] = beta
        A[s, st:ed + 1] = np.conj(A[st:ed + 1, s])
    else:
        O[:, 0] = 0.0
        O[0, 0] = beta
        O[:, 1:] -= np.conj(tau) * np.outer(v, v.conj() @ O[:, 1:])
        A[st:ed + 1, sp:sp + lenp] = O
        A[sp:sp + lenp, st:ed + 1] = O.conj().T
    D = A[st:ed + 1, st:ed + 1]
    y = D @ v
    g = float(np.real(v.conj() @ y))
    w = tau * y - 0.5 * abs(tau) ** 2 * g * v
    A[st:ed + 1, st:ed + 1] = D - np.outer(v, w.conj()) - np.outer(w, v.conj())
    al = v.conj() @ gc[st:ed + 1]
    gc[st:ed + 1] -= np.outer(v, np.conj(tau) * al)
    state[s] = (ln, v, tau)


def chase(A0, g0, b, rng=None, ahead=2):
    """sequential sweeps (rng None) or a random order of ready tasks, where (s, k) is ready
    once sweep s - 1 has finished min(k + ahead, all) tasks"""
    N = A0.shape[0]
    A, gc, state, rows = A0.copy(), g0.copy(), {}, task_rows(N, b)
    done = [0] * (N - 1)
    if rng is None:
        for s in range(N - 1):
            for k in range(ntasks(N, b, s)):
                run_task(A, gc, state, s, k, rows)
        return A, gc
    left = sum(ntasks(N, b, s) for s in range(N - 1))
    while left:
        ready = [s for s in range(N - 1) if done[s] < ntasks(N, b, s) and
                 (s == 0 or done[s - 1] >= min(done[s] + ahead, ntasks(N, b, s - 1)))]
        s = ready[rng.integers(len(ready))]
        run_task(A, gc, state, s, done[s], rows)
        done[s] += 1
        left -= 1
    return A, gc


def band_matrix(N, b, seed):
    rng = np.random.default_rng(seed)
    H = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    H = np.tril(H + H.conj().T)
    H[np.tril_indices(N, -b - 1)] = 0.0
    H = H + np.tril(H, -1).conj().T
    return H, rng.standard_normal((N, 2)) + 1j * rng.standard_normal((N, 2))


@pytest.mark.parametrize("N,b", [(60, 8), (70, 16), (90, 32)])
def test_interleavings_allowed_by_the_rule_give_the_sequential_result(N, b):
    H, g = band_matrix(N, b, N + b)
    T0, g0 = chase(H, g, b)
    # a real symmetric tridiagonal with H's eigenvalues
    assert np.abs(np.tril(T0, -2)).max() == 0.0 and np.abs(np.triu(T0, 2)).max() == 0.0
    assert np.allclose(np.linalg.eigvalsh(T0), np.linalg.eigvalsh(H), atol=1e-10)
    for trial in range(3):
        T1, g1 = chase(H, g, b, np.random.default_rng(trial))
        assert np.array_equal(T1, T0) and np.array_equal(g1, g0), trial


def test_a_looser_rule_changes_the_result():
    N, b = 70, 16
    H, g = band_matrix(N, b, 5)
    T0, _ = chase(H, g, b)
    T1, _ = chase(H, g, b, np.random.default_rng(0), ahead=1)
    assert not np.array_equal(T1, T0)
