"""leftsvd's warm-started Rayleigh-Ritz eigen-step against the direct solver.

In the Tucker loops the eigen-step first tries a Rayleigh-Ritz step on the
block Krylov space of X0 = Z^T U (U the factor the S3TTMc ran on, Z = Y
diag(sqrt W)) and keeps it only when every Ritz pair's residual is below
1e-13 of the largest eigenvalue; otherwise it takes the tridiagonal path.
The HOOI trajectories (src/sparse/spsymtensor.cpp:201-252) must not depend on
which of the two produced an iteration's factor.
"""
import sys

import numpy as np
import pytest

import oracle
from oracle import tucker
import symprop_amd as sp

from test_oracle import exact_rank_tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def planted(order, n, nnz, seed):
    inds, vals = sp.generate(sp.GEN_PLANTED, order, n, nnz, 0.8, 8, seed=seed)
    return inds, vals


@pytest.mark.parametrize("alg", [0, 2])
def test_warm_step_taken_and_trajectory_unchanged(alg, monkeypatch):
    # planted order-6 tensor: after a few iterations the Gram's top-8
    # eigenvectors lie in span(X0) and the warm step is taken
    N, n, R = 6, 3000, 8
    inds, vals = planted(N, n, 300_000, 3)
    kw = dict(alg=alg, rank=R, max_iter=25, rel_tol=0.0, delta=0.5, seed=0)
    say("warm eigen-step, planted order 6, alg", alg, "direct")
    monkeypatch.setenv("SP_EIG_WARM", "0")
    td = sp.SpSymtensor(n, N, inds, vals)
    td.timing(True)
    rd = td.tucker(**kw)
    assert td.timing_get("eig_warm")[1] == 0
    say("warm eigen-step, planted order 6, alg", alg, "warm")
    monkeypatch.delenv("SP_EIG_WARM")
    tw = sp.SpSymtensor(n, N, inds, vals)
    tw.timing(True)
    rw = tw.tucker(**kw)
    ok = tw.timing_get("eig_warm_ok")[1]
    tried = tw.timing_get("eig_warm")[1]
    say("warm steps taken", ok, "of", tried, "tried")
    assert ok >= 10
    assert np.allclose(rw["errors"], rd["errors"], rtol=0, atol=1e-12)
    assert np.allclose(rw["core_norms"], rd["core_norms"], rtol=1e-12, atol=0)
    assert tucker.subspace_cos(rw["U"], rd["U"]) > 1 - 1e-12
    assert np.allclose(rw["U"].T @ rw["U"], np.eye(R), atol=1e-12)


@pytest.mark.parametrize("alg,ofn", [(0, tucker.hooi), (2, tucker.althooi)])
def test_warm_trajectory_matches_oracle(alg, ofn):
    # the same comparison against the oracle's loops (numpy SVD of the
    # unpacked Y / eigh of Y W Y^T) on a tensor small enough for them: a
    # multilinear-rank-3 order-5 tensor plus noise, whose Gram has a clear
    # gap below the top 3, so the warm step is taken once U nears the
    # dominant subspace.  delta = 10 leaves the rows unclipped (with 12 rows
    # the reference's 0.5 clips some of them, and the regularised factor's
    # fixed point is then not an exact singular subspace of Y)
    N, n, R = 5, 12, 3
    tt, v, _, _ = exact_rank_tensor(n=n, R=R, N=N, seed=4)
    v = v + 1e-4 * np.abs(v).max() * np.random.default_rng(5).standard_normal(len(v))
    t = sp.SpSymtensor(n, N, tt, v)
    t.timing(True)
    say("warm eigen-step vs oracle, noisy rank-3 order 5, alg", alg)
    r = t.tucker(alg=alg, rank=R, max_iter=15, rel_tol=0.0, delta=10.0, seed=0)
    Uo, eo, _ = ofn(oracle.Tree(N, n, tt, v), R, 15, rel_tol=0.0, delta=10.0, seed=0)
    say("warm steps taken", t.timing_get("eig_warm_ok")[1], "errors", r["errors"])
    assert t.timing_get("eig_warm_ok")[1] >= 5
    assert np.allclose(r["errors"], eo, rtol=0, atol=1e-10)
    assert tucker.subspace_cos(r["U"], Uo) > 1 - 1e-9


def test_warm_leftsvd_matches_numpy_svd():
    # one leftsvd at a converged factor: the warm result against numpy's SVD
    # of the unpacked Y (singular values and left subspace)
    N, n, R = 5, 12, 3
    tt, v, _, _ = exact_rank_tensor(n=n, R=R, N=N, seed=6)
    v = v + 1e-4 * np.abs(v).max() * np.random.default_rng(7).standard_normal(len(v))
    t = sp.SpSymtensor(n, N, tt, v)
    r = t.tucker(alg=0, rank=R, max_iter=20, rel_tol=0.0, delta=10.0, seed=0)
    t.timing(True)
    Y = t.s3ttmc(r["U"])
    sig = t.leftsvd_y()
    Un = t.factor()
    say("converged leftsvd, warm steps", t.timing_get("eig_warm_ok")[1])
    assert t.timing_get("eig_warm_ok")[1] == 1
    Yu = np.stack([oracle.unpack(Y[i], R, N - 1) for i in range(n)])
    Us, s, _ = np.linalg.svd(Yu, full_matrices=False)
    assert np.allclose(sig, s[:R], rtol=1e-11)
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)
    c = np.linalg.svd(Un.T @ Us[:, :R], compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-10)


def test_krylov_step_on_zipf_order4(monkeypatch):
    # P = 220 (order 4, rank 10): leftsvd's block Krylov form.  On this kind of
    # tensor the loop's fixed point is not an exact singular subspace of Y
    # (row clipping), so the one-block step never applies; the Krylov space
    # of X0 reaches the direct solver's accuracy after a few iterations, and
    # the trajectory must not depend on which of the two gave an iteration
    N, n, R = 4, 2000, 10
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, 100_000, 1.1, 1, seed=2)
    kw = dict(alg=0, rank=R, max_iter=30, rel_tol=0.0, delta=0.5, seed=0)
    say("krylov eigen-step, zipf order 4, direct")
    monkeypatch.setenv("SP_EIG_WARM", "0")
    td = sp.SpSymtensor(n, N, inds, vals)
    rd = td.tucker(**kw)
    say("krylov eigen-step, zipf order 4, krylov")
    monkeypatch.delenv("SP_EIG_WARM")
    tw = sp.SpSymtensor(n, N, inds, vals)
    tw.timing(True)
    rw = tw.tucker(**kw)
    ok = tw.timing_get("eig_warm_ok")[1]
    say("krylov steps taken", ok, "of", tw.timing_get("eig_warm")[1], "tried")
    assert ok >= 15
    assert np.allclose(rw["errors"], rd["errors"], rtol=0, atol=1e-12)
    assert np.allclose(rw["core_norms"], rd["core_norms"], rtol=1e-12, atol=0)
    assert tucker.subspace_cos(rw["U"], rd["U"]) > 1 - 1e-12
