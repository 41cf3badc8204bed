"""HOEVD's top-R solve on the implicit operator A = X_(1) X_(1)^T
(symprop_amd/csrc/hoevd.hip, block Krylov), against numpy's eigh of the
oracle's dense A (oracle.hoevd_factor) and against the library's own dense
path, which the reference's matrix::syevd call (src/sparse/spsymtensor.cpp:
376-382) is pinned to in test_gpu_parity.py.  SP_HOEVD=krylov / dense, read
when a tensor is created, selects the solver; by default dim > 4096 takes
the Krylov solve."""
import sys

import numpy as np
import pytest

import oracle
import symprop_amd as sp

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def check_pairs(U, w, Uo, wo, vec_tol=1e-9):
    R = len(wo)
    scale = wo[-1]
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wo).max() <= 1e-11 * scale
    assert np.allclose(U.T @ U, np.eye(R), atol=1e-12)
    # columns up to sign where the eigenvalue is separated from its neighbours
    full = np.r_[wo, np.inf]
    for k in range(R):
        gap = min(full[k + 1] - wo[k], wo[k] - (wo[k - 1] if k > 0 else -np.inf))
        if gap > 1e-4 * scale:
            assert abs(abs(U[:, k] @ Uo[:, k]) - 1.0) <= vec_tol, k


CASES = [
    # order, dim, tuples, generator, exponent, rank
    (4, 600, 20000, sp.GEN_ZIPF, 1.1, 6),     # hub leaves cut into pieces
    (3, 1500, 30000, sp.GEN_ZIPF, 1.3, 4),
    (5, 500, 12000, sp.GEN_UNIFORM, 0.0, 8),  # values 1.0: repeated eigenvalues
    (4, 800, 15000, sp.GEN_ZIPF, 1.1, 20),    # two 16-column chunks per apply
]


@pytest.mark.parametrize("N,n,cnt,gen,ex,R", CASES)
def test_krylov_matches_oracle(N, n, cnt, gen, ex, R, monkeypatch):
    monkeypatch.setenv("SP_HOEVD", "krylov")
    inds, vals = sp.generate(gen, N, n, cnt, ex, 1, seed=5)
    say("krylov hoevd vs oracle", N, n, len(vals), R)
    t = sp.SpSymtensor(n, N, inds, vals)
    U, w = t.hoevd_init(R)
    Uo, wo = oracle.hoevd_factor(oracle.hoevd_gram(N, n, inds, vals), R)
    check_pairs(U, w, Uo, wo)
    # the top-R subspace, where a gap separates it from the rest
    A = oracle.hoevd_gram(N, n, inds, vals)
    lam = np.linalg.eigvalsh(A)
    if lam[-R] - lam[-R - 1] > 1e-6 * lam[-1]:
        assert np.linalg.norm(U - Uo @ (Uo.T @ U)) <= 1e-9
    # residuals of the returned pairs
    assert np.linalg.norm(A @ U - U * w, axis=0).max() <= 1e-11 * lam[-1]
    # deterministic
    U2, w2 = t.hoevd_init(R)
    assert np.array_equal(U, U2) and np.array_equal(w, w2)


def test_krylov_is_the_default_above_4096_and_matches_dense(monkeypatch):
    N, n, R = 4, 6000, 8
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, 200000, 1.1, 1, seed=6)
    say("hoevd n = 6000: default (krylov) vs dense", len(vals))
    monkeypatch.delenv("SP_HOEVD", raising=False)
    t = sp.SpSymtensor(n, N, inds, vals)
    U, w = t.hoevd_init(R)
    monkeypatch.setenv("SP_HOEVD", "dense")
    td = sp.SpSymtensor(n, N, inds, vals)
    Ud, wd = td.hoevd_init(R)
    check_pairs(U, w, Ud, wd)
    # the decomposition loop started from either agrees
    a = t.tucker(alg=0, rank=R, max_iter=4, rel_tol=0.0, init="hoevd")
    b = td.tucker(alg=0, rank=R, max_iter=4, rel_tol=0.0, init="hoevd")
    assert np.allclose(a["errors"], b["errors"], rtol=0, atol=1e-10)


def test_krylov_rejects_rank_beyond_its_basis(monkeypatch):
    monkeypatch.setenv("SP_HOEVD", "krylov")
    inds, vals = sp.generate(sp.GEN_UNIFORM, 3, 40, 300, 0.0, 1, seed=1)
    t = sp.SpSymtensor(40, 3, inds, vals)
    say("krylov hoevd, dim 40 rank 4: basis of four blocks does not fit")
    with pytest.raises(sp.SpError):
        t.hoevd_init(4)
