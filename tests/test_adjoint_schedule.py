"""Binomial checkpointing schedule of the rollout adjoint (host logic, no GPU).

The dynamic-programming table is checked against brute-force recursion; the reverse sweep is
run on a scalar toy map x -> f(x), whose exact derivative of x_N with respect to x_0 is the
product of f'(x_n), and checked for the gradient, the number of forward steps it takes
(T[n, s] from the table) and the number of states alive at once (<= s + 1).
"""
import functools
import math

import pytest

from jaxfluids_amd.adjoint import binomial_table, first_level_snapshots, reverse_rollout


@functools.lru_cache(maxsize=None)
def brute(t, s):
    if t <= 1:
        return 0
    if s == 0:
        return t * (t - 1) // 2
    return min(m + brute(t - m, s - 1) + brute(m, s) for m in range(1, t))


def test_table_matches_brute_force():
    T, M = binomial_table(40, 6)
    for t in range(1, 41):
        for s in range(0, 7):
            assert T[t, s] == brute(t, s), (t, s)
            if t > 1 and s > 0:
                m = M[t, s]
                assert m + T[t - m, s - 1] + T[m, s] == T[t, s]


def test_table_is_binomial():
    """Griewank & Walther's closed form: with S checkpoints (here the start state plus s free
    slots, S = s + 1) the fewest forward steps to reverse t steps are r t - C(S + r, S + 1) for the
    smallest r with C(S + r, S) >= t."""
    T, _ = binomial_table(120, 5)
    for s in range(1, 6):
        S = s + 1
        for t in range(1, 121):
            r = 0
            while math.comb(S + r, S) < t:
                r += 1
            assert T[t, s] == r * t - math.comb(S + r, S + 1), (t, s)


f = lambda x: math.sin(x) + 0.5 * x
df = lambda x: math.cos(x) + 0.5


@pytest.mark.parametrize("n,s,prestore", [(1, 0, False), (7, 0, False), (20, 1, False), (37, 3, False),
                                          (64, 4, True), (50, 50, True), (23, 2, True)])
def test_reverse_sweep_on_a_scalar_map(n, s, prestore):
    xs = [0.3]
    for _ in range(n):
        xs.append(f(xs[-1]))
    exact = 1.0
    for x in xs[:-1]:
        exact *= df(x)
    snaps = {0: xs[0]}
    if prestore:
        for p in first_level_snapshots(n, s):
            snaps[p] = xs[p]
    pre = len(snaps) - 1
    steps = [0]
    alive = [len(snaps)]

    def advance(x, k):
        steps[0] += k
        for _ in range(k):
            x = f(x)
        return x

    def store(x):
        alive[0] = max(alive[0], len(snaps) + 1)
        return x

    grad = reverse_rollout(snaps, n, s, 1.0, advance, lambda x, lam: lam * df(x), store)
    assert abs(grad - exact) <= 1e-12 * abs(exact)
    assert alive[0] <= s + 1
    T, _ = binomial_table(n, s)
    if not prestore:
        assert steps[0] == T[n, s]
    else:
        assert steps[0] <= T[n, s] and pre == len(first_level_snapshots(n, s))


def test_enstrophy_of_taylor_green_closed_form():
    """Taylor-Green u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on [0, 2pi)^3: central
    differences scale each derivative of a single mode by s = sin(h)/h, so the discrete
    vorticity is s (-cos x sin y sin z, -sin x cos y sin z, 2 sin x sin y cos z) and its
    enstrophy 0.5 * 0.75 s^2 (2pi)^3 (grid means of sin^2, cos^2 are 1/2)."""
    import numpy as np
    import torch

    from jaxfluids_amd.adjoint import enstrophy

    n = 12
    h = 2 * np.pi / n
    x = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = np.stack([np.ones_like(X), np.sin(X) * np.cos(Y) * np.cos(Z), -np.cos(X) * np.sin(Y) * np.cos(Z),
                  np.zeros_like(X), np.ones_like(X)])
    got = float(enstrophy(torch.as_tensor(W), (h, h, h)))
    s = np.sin(h) / h
    assert abs(got - 0.5 * 0.75 * s * s * (2 * np.pi) ** 3) <= 1e-12 * got
    # 2-D shear u(y) = sin y: omega_z = -s cos y, enstrophy 0.5 * s^2 / 2 * area
    W2 = np.zeros((5, n, n, 1))
    W2[1] = np.sin(Y[:, :, :1])
    got2 = float(enstrophy(torch.as_tensor(W2), (h, h, 1.0)))
    assert abs(got2 - 0.5 * 0.5 * s * s * (2 * np.pi) ** 2) <= 1e-12 * got2


@pytest.mark.parametrize("n,s", [(1, 0), (7, 2), (30, 4), (50, 9)])
def test_reverse_sweep_with_a_state_pool(n, s):
    """The reverse sweep on pooled snapshot slots (StatePool on the CPU): snapshots + 1 slots
    suffice (snaps[0] and the first-level snapshots included), every stored state is released,
    and the gradient is the chain rule's."""
    import torch

    from jaxfluids_amd.adjoint import StatePool

    pool = StatePool(s + 1, (1,), "cpu", 0.0)
    xs = [0.3]
    for _ in range(n):
        xs.append(f(xs[-1]))
    exact = 1.0
    for x in xs[:-1]:
        exact *= df(x)
    t = lambda x: (torch.tensor([x], dtype=torch.float64), torch.zeros(1, dtype=torch.float64))  # noqa: E731
    snaps = {0: pool.store(t(xs[0]))}
    for p in first_level_snapshots(n, s):
        snaps[p] = pool.store(t(xs[p]))

    def advance(state, k):
        x = float(state[0][0])
        for _ in range(k):
            x = f(x)
        return t(x)

    grad = reverse_rollout(snaps, n, s, 1.0, advance, lambda st, lam: lam * df(float(st[0][0])), pool.store,
                           pool.release)
    assert abs(grad - exact) <= 1e-12 * abs(exact)
    assert set(snaps) <= {0}
    pool.release(snaps.pop(0))
    assert len(pool._free) == pool.nslots
