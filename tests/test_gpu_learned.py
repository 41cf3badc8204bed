"""GPU kernels of the learned warm start against the oracle (oracle/mlp.py)."""
import sys

import numpy as np
import pytest

from oracle import mlp
import symprop_amd as sp
from symprop_amd import learned

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_features_match_oracle():
    inds, vals = sp.generate(sp.GEN_PLANTED, 5, 400, 3000, 0.8, 8, seed=3)
    say("features")
    t = sp.SpSymtensor(400, 5, inds, vals)
    F = learned.features(t, rounds=3, seed=11)
    Fo = mlp.node_features(5, 400, inds, vals, rounds=3, seed=11)
    assert F.shape == (400, 64)
    assert np.allclose(F, Fo, rtol=1e-5, atol=1e-6)


def test_features_hub_rows_split_over_pieces():
    # Zipf(1.2): node 1 owns tens of thousands of owners, so its sketch row
    # is summed over many pieces
    inds, vals = sp.generate(sp.GEN_ZIPF, 5, 300, 30000, 1.2, 1, seed=4)
    say("features zipf hubs")
    t = sp.SpSymtensor(300, 5, inds, vals)
    F = learned.features(t, rounds=3, seed=2)
    Fo = mlp.node_features(5, 300, inds, vals, rounds=3, seed=2)
    deg = np.expm1(Fo[:, 0])
    assert deg.max() > 8 * 2048
    assert np.allclose(F, Fo, rtol=1e-5, atol=1e-6)
    assert np.array_equal(learned.features(t, rounds=3, seed=2), F)


def test_loss_and_gradient_match_oracle():
    rng = np.random.default_rng(0)
    din, dh, dout, N = 64, 256, 8, 300
    m = learned.MLP(din, dh, dout, seed=1)
    p = m.params().astype(np.float64)
    X = rng.standard_normal((N, din)).astype(np.float32)
    T = rng.standard_normal((N, dout)).astype(np.float32)
    say("mlp loss/grad")
    L, g = m.loss_grad(X, T)
    Lo, go = mlp.loss_grad(p, X.astype(np.float64), T.astype(np.float64), din, dh, dout)
    assert abs(L - Lo) <= 1e-5 * abs(Lo)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go)
    # per parameter block
    off = 0
    for s in mlp.param_shapes(din, dh, dout):
        k = int(np.prod(s))
        gb, gob = g[off:off + k], go[off:off + k]
        assert np.linalg.norm(gb - gob) <= 1e-4 * np.linalg.norm(gob) + 1e-9
        off += k


def test_adam_trajectory_matches_oracle():
    rng = np.random.default_rng(1)
    din, dh, dout, N = 64, 32, 4, 128
    m = learned.MLP(din, dh, dout, seed=2)
    p = m.params().astype(np.float64)
    X = rng.standard_normal((N, din)).astype(np.float32)
    T = rng.standard_normal((N, dout)).astype(np.float32)
    say("adam trajectory")
    losses = m.train(X, T, steps=5, batch=N, lr=1e-3, seed=0)   # full batch
    mo = np.zeros_like(p)
    vo = np.zeros_like(p)
    ref = []
    for s in range(1, 6):
        L, g = mlp.loss_grad(p, X.astype(np.float64), T.astype(np.float64), din, dh, dout)
        ref.append(L)
        p, mo, vo = mlp.adam_step(p, g, mo, vo, s, 1e-3)
    assert np.allclose(losses, ref, rtol=1e-4)
    assert np.allclose(m.params(), p, rtol=1e-3, atol=1e-5)


def planted_basis(n, K):
    E = np.zeros((n, K))
    b = n // K
    for c in range(K):
        E[c * b:(c + 1) * b, c] = 1.0
    return np.linalg.qr(E)[0]


def test_learned_start_gain_over_random_and_hoevd():
    """Order-5 planted 8-community tensors (n=1000, 500k binary tuples,
    p_in=0.9), rank 8, HOOI with delta=0.5.  The network is trained on the
    factors plain HOOI converges to on 4 training seeds (started from the
    hypergraph sketch), then compared on 2 held-out seeds with random and
    HOEVD starts: iterations until the subspace moves by < 1e-6, converged
    fit, and angle to the planted communities."""
    N, n, nnz, K, R, delta, iters = 5, 1000, 500_000, 8, 8, 0.5, 30
    E = planted_basis(n, K)

    def tensor(seed):
        inds, vals = sp.generate(sp.GEN_PLANTED, N, n, nnz, 0.9, K, seed=seed)
        return sp.SpSymtensor(n, N, inds, vals)

    feats, facs = [], []
    for s in range(1, 5):
        say("teacher seed", s)
        t = tensor(s)
        F = learned.features(t)
        Us, _ = learned.hooi_trace(t, learned.spectral_start(F, R), iters, delta,
                                   stop_tol=1e-6)
        assert learned.converged_at(Us) is not None
        feats.append(F)
        facs.append(Us[-1])
    model = learned.WarmStartModel(R, seed=0)
    losses = model.fit(feats, facs, steps=600, batch=4096, lr=1e-3)
    assert losses[-10:].mean() < 0.05 * losses[:5].mean()
    for s in (101, 102):
        say("held-out seed", s)
        t = tensor(s)
        F = learned.features(t)
        Ul, el = learned.hooi_trace(t, model.predict(F), iters, delta, stop_tol=1e-6)
        Ur, er = learned.hooi_trace(t, np.random.default_rng(s).standard_normal((n, R)),
                                    iters, delta, stop_tol=1e-6)
        Uh0, _ = t.hoevd_init(R)
        Uh, eh = learned.hooi_trace(t, Uh0, iters, delta, stop_tol=1e-6)
        kl = learned.converged_at(Ul)
        kr = learned.converged_at(Ur) or iters + 1
        kh = learned.converged_at(Uh) or iters + 1
        say("iterations learned/random/hoevd", kl, kr, kh,
            "errors", el[-1], er[-1], eh[-1])
        assert kl is not None and kl <= 10
        assert kl < kr and kl < kh
        # converged fit at least as good as where the fallbacks end
        assert el[-1] <= min(er[-1], eh[-1]) + 1e-12
        assert learned.sin_angle(Ul[-1], E) < 0.1


def test_training_reduces_loss():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2000, 64)).astype(np.float32)
    W = rng.standard_normal((64, 8)).astype(np.float32) / 8
    T = np.tanh(X @ W).astype(np.float32)
    m = learned.MLP(64, 256, 8, seed=4)
    say("training")
    losses = m.train(X, T, steps=300, batch=512, lr=1e-3, seed=1)
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()
