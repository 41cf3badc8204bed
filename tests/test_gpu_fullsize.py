"""S3TTMc at the benchmark sizes: the order-4 Zipf tensor (n = 10,000, 1M
nnz, R = 10) against the oracle itself, and the order-6 planted tensor
(n = 50,000, 5M nnz, R = 8) and the order-7 Zipf(1.2) tensor (n = 100,000,
10M nnz, R = 6) through size-independent properties (the oracle's opt5 level
arrays would need hundreds of GB there); the order-7 tensor's first 100,000
nonzeros also against the oracle.

Properties used:
- homogeneity: Y(2U) = 2^(N-1) Y(U) bit for bit (scaling by a power of two
  commutes with every rounding);
- linearity in the values: Y(2X) = 2 Y(X) bit for bit;
- full symmetry of the core U^T Y unpacked to an order-N tensor;
- additivity over a split of the nonzeros: Y(X1 + X2) = Y(X1) + Y(X2);
- the weighted Gram identity of the packed Y;
- determinism: repeated calls are identical;
and the eigen-step of the HOOI update on that Gram against numpy's eigh.
"""
import itertools
import sys

import numpy as np
import pytest

import oracle
import symprop_amd as sp

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def werr(Y, Yref, R, order):
    W = oracle.symtens_coeff(R, order - 1)
    return np.sqrt(((Y - Yref) ** 2 * W).sum()) / np.sqrt((Yref ** 2 * W).sum())


@pytest.fixture(scope="module")
def zipf4():
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 10000, 1_000_000, 1.1, 1, seed=0)
    return inds, vals


@pytest.fixture(scope="module")
def planted6():
    inds, vals = sp.generate(sp.GEN_PLANTED, 6, 50000, 5_000_000, 0.8, 8, seed=0)
    return inds, vals


@pytest.fixture(scope="module")
def zipf7():
    inds, vals = sp.generate(sp.GEN_ZIPF, 7, 100000, 10_000_000, 1.2, 1, seed=0)
    return inds, vals


SHAPES = {"zipf4": (10000, 4, 10), "planted6": (50000, 6, 8), "zipf7": (100000, 7, 6)}


def test_zipf7_prefix_matches_oracle(zipf7):
    inds, vals = zipf7
    n, N, R = SHAPES["zipf7"]
    k = 100_000
    a, v = np.ascontiguousarray(inds[:k]), np.ascontiguousarray(vals[:k])
    U = oracle.random_factor(0, n, R)
    say("zipf7 prefix vs oracle", k)
    t = sp.SpSymtensor(n, N, a, v)
    Y = t.s3ttmc(U)
    T = oracle.Tree(N, n, a, v)
    Yo = T.s3ttmc(U, R, nthreads=16)
    assert werr(Y, Yo, R, N) <= 1e-12
    for l in range(N - 1):
        assert t.level_size(l) == T.level_size(l)


def test_zipf4_full_size_matches_oracle(zipf4):
    inds, vals = zipf4
    n, N, R = 10000, 4, 10
    U = oracle.random_factor(0, n, R)
    say("zipf4 full size vs oracle")
    t = sp.SpSymtensor(n, N, inds, vals)
    Y = t.s3ttmc(U)
    T = oracle.Tree(N, n, inds, vals)
    Yo = T.s3ttmc(U, R, nthreads=16)
    assert werr(Y, Yo, R, N) <= 1e-12
    # construction: level sizes agree with the oracle's
    for l in range(N - 1):
        assert t.level_size(l) == T.level_size(l)
    del T
    # the reduced-precision forms against FP64 at full size
    Y32 = t.s3ttmc(U, precision=sp.PREC_FP32)
    Ybf = t.s3ttmc(U, precision=sp.PREC_BF16)
    assert werr(Y32, Y, R, N) < 2e-5
    assert werr(Ybf, Y, R, N) < 1e-2


def unpack_core(core, R, N):
    full = np.stack([oracle.unpack(core[r], R, N - 1) for r in range(R)])
    return full.reshape((R,) * N)


@pytest.mark.parametrize("which", ["zipf4", "planted6", "zipf7"])
def test_full_size_properties(which, request):
    inds, vals = request.getfixturevalue(which)
    n, N, R = SHAPES[which]
    rng = np.random.default_rng(7)
    U = np.linalg.qr(rng.standard_normal((n, R)))[0]
    say(which, "properties: build")
    t = sp.SpSymtensor(n, N, inds, vals)
    Y = t.s3ttmc(U)
    assert np.all(np.isfinite(Y))
    say(which, "determinism and homogeneity")
    assert np.array_equal(t.s3ttmc(U), Y)
    assert np.array_equal(t.s3ttmc(2.0 * U), 2.0 ** (N - 1) * Y)
    say(which, "linearity in the values")
    t2 = sp.SpSymtensor(n, N, inds, 2.0 * vals)
    assert np.array_equal(t2.s3ttmc(U), 2.0 * Y)
    del t2
    say(which, "additivity over a split of the nonzeros")
    half = len(vals) // 2
    ta = sp.SpSymtensor(n, N, inds[:half], vals[:half])
    Ya = ta.s3ttmc(U)
    del ta
    tb = sp.SpSymtensor(n, N, inds[half:], vals[half:])
    Yb = tb.s3ttmc(U)
    del tb
    assert werr(Ya + Yb, Y, R, N) <= 1e-12
    say(which, "symmetric core and Gram identity")
    core = U.T @ Y
    full = unpack_core(core, R, N)
    scale = np.abs(full).max()
    for perm in [(1, 0) + tuple(range(2, N)), tuple(range(1, N)) + (0,)]:
        assert np.allclose(full, np.transpose(full, perm), atol=1e-11 * scale)
    W = oracle.symtens_coeff(R, N - 1)
    assert np.array_equal(t.s3ttmc(U), Y)   # the device Y is Y(U) again
    G = t.gram()
    Gref = (Y * np.sqrt(W)).T @ (Y * np.sqrt(W))
    assert np.allclose(G, Gref, rtol=1e-10, atol=1e-12 * np.abs(Gref).max())
    say(which, "eigen-step on the full Gram")
    # the HOOI factor update: top-R eigenpairs of this Gram (the library's
    # reduction, host bisection/inverse iteration, back-transform) against
    # numpy's LAPACK eigh of the same matrix
    lam, V = np.linalg.eigh(Gref)
    top = lam[::-1][:R]
    sig = t.leftsvd_y()
    assert np.allclose(sig ** 2, top, rtol=0, atol=1e-11 * lam.max())
    Un = t.factor()
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-10)
    if top[-1] - lam[::-1][R] > 1e-6 * lam.max():
        Z = (Y * np.sqrt(W)) @ V[:, ::-1][:, :R] / np.sqrt(top)
        c = np.linalg.svd(Un.T @ Z, compute_uv=False)
        assert np.allclose(c, 1.0, atol=1e-8)
