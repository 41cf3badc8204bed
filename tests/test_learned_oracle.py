"""Pins the oracle of the learned warm start (oracle/mlp.py) on the CPU."""
import itertools

import numpy as np

import oracle
from oracle import mlp


def small_params(rng, din, dh, dout):
    return np.concatenate([rng.standard_normal(int(np.prod(s))) * 0.5
                           for s in mlp.param_shapes(din, dh, dout)])


def test_backward_matches_central_differences():
    rng = np.random.default_rng(0)
    din, dh, dout, N = 5, 7, 3, 11
    p = small_params(rng, din, dh, dout)
    X = rng.standard_normal((N, din))
    T = rng.standard_normal((N, dout))
    L, g = mlp.loss_grad(p, X, T, din, dh, dout)
    h = 1e-6
    num = np.zeros_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = h
        num[i] = (mlp.loss_grad(p + e, X, T, din, dh, dout)[0] -
                  mlp.loss_grad(p - e, X, T, din, dh, dout)[0]) / (2 * h)
    assert np.allclose(g, num, rtol=1e-5, atol=1e-7)
    O, _ = mlp.forward(p, X, din, dh, dout)
    assert np.isclose(L, ((O - T) ** 2).sum() / N)


def test_adam_first_step_is_signed_lr():
    g = np.array([3.0, -0.5, 1e-3])
    p, m, v = mlp.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01, eps=0.0)
    assert np.allclose(p, -0.01 * np.sign(g))
    # a quadratic is minimised
    x = np.array([5.0, -3.0])
    m = np.zeros(2)
    v = np.zeros(2)
    for t in range(1, 3001):
        x, m, v = mlp.adam_step(x, 2 * x, m, v, t, lr=0.05)
    assert np.abs(x).max() < 1e-2


def test_procrustes_recovers_rotation():
    rng = np.random.default_rng(1)
    U = np.linalg.qr(rng.standard_normal((30, 4)))[0]
    Q = np.linalg.qr(rng.standard_normal((4, 4)))[0]
    assert np.allclose(mlp.procrustes_align(U @ Q, U), U, atol=1e-12)


def test_node_gauss_is_standard_normal_and_deterministic():
    g = mlp.node_gauss(4000, 62, 7)
    assert np.array_equal(g, mlp.node_gauss(4000, 62, 7))
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert not np.array_equal(g, mlp.node_gauss(4000, 62, 8))


def test_features_match_loop_restatement():
    rng = np.random.default_rng(2)
    N, n = 4, 9
    t = np.unique(np.sort(rng.integers(1, n + 1, size=(25, N)), axis=1), axis=0)
    v = rng.random(len(t))
    F = mlp.node_features(N, n, t, v, rounds=2, seed=5)
    # degree: owners with rem = node (every position of every nonzero)
    deg = np.zeros(n)
    for row in t:
        for a in row:
            deg[a - 1] += 1
    assert np.allclose(F[:, 0], np.log1p(deg))
    # partners: distinct other nodes sharing a nonzero
    partners = [set() for _ in range(n)]
    for row in t:
        for a, b in itertools.combinations(row, 2):
            if a != b:
                partners[a - 1].add(b)
                partners[b - 1].add(a)
    assert np.allclose(F[:, 1], np.log1p([len(s) for s in partners]))
    # pair multiplicity: occurrences of {a, b} over all position pairs
    mult = {}
    for row in t:
        for a, b in itertools.combinations(row, 2):
            if a != b:
                mult[(a, b)] = mult.get((a, b), 0) + 1
    maxmult, rep = np.zeros(n), np.zeros(n)
    for (a, b), m in mult.items():
        for x in (a, b):
            maxmult[x - 1] = max(maxmult[x - 1], m)
            rep[x - 1] += m >= 2
    assert np.allclose(F[:, 2], np.log1p(maxmult))
    assert np.allclose(F[:, 3], np.log1p(rep))
    assert rep.max() >= 1   # the case exercises repeated pairs
    # propagation: s'[r] += v * sum of the other positions' vectors, per position
    s = mlp.node_gauss(n, mlp.SKETCH, 5)
    for _ in range(2):
        new = np.zeros_like(s)
        for row, val in zip(t - 1, v):
            for j in range(N):
                others = [row[a] for a in range(N) if a != j]
                new[row[j]] += val * s[others].sum(axis=0)
        nrm = np.linalg.norm(new, axis=1, keepdims=True)
        s = np.where(nrm > 0, new / np.where(nrm > 0, nrm, 1), 0)
    assert np.allclose(F[:, 4:], s, atol=1e-12)
