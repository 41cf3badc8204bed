"""The eigen-step's host stage (symprop_amd/csrc/tridiag.cpp: bisection for
the R largest eigenvalues of a symmetric tridiagonal matrix, inverse iteration
for their vectors) against numpy's LAPACK eigh of the same matrix.  The
library's eigen-step follows matrix::syevd (src/dense/matrixops.cpp:77-100);
this pins its tridiagonal part on the CPU, including the cases inverse
iteration finds hard: clusters, Wilkinson pairs, exact multiplicities from a
split matrix, a rank-deficient Gram, zero and graded matrices."""
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "symprop_amd", "csrc")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    exe = str(tmp_path_factory.mktemp("tridiag") / "tridiag_harness")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC,
                    os.path.join(ROOT, "tests", "tridiag_harness.cpp"),
                    os.path.join(CSRC, "tridiag.cpp"), "-o", exe], check=True)
    return exe


def run(exe, d, e, R):
    P = len(d)
    text = "%d %d\n%s\n%s\n" % (P, R, " ".join(repr(float(x)) for x in d),
                                " ".join(repr(float(x)) for x in e))
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
    v = np.array(out.split(), dtype=float)
    return v[:R], v[R:].reshape(R, P)


def gram_tridiagonal(rng, n, P, decay, rank=None):
    Y = rng.standard_normal((n, P)) * (np.arange(P) + 1.0) ** -decay
    if rank is not None:
        Y = rng.standard_normal((n, rank)) @ rng.standard_normal((rank, P))
    T = sl.hessenberg(Y.T @ Y)
    return np.diag(T).copy(), np.diag(T, 1).copy()


def wilkinson(m):
    return np.abs(np.arange(-m, m + 1)).astype(float), np.ones(2 * m)


def cases():
    rng = np.random.default_rng(0)
    for P, R in [(1, 1), (2, 1), (2, 2), (3, 2), (10, 10), (50, 8), (220, 10)]:
        yield "random%d" % P, rng.standard_normal(P), rng.standard_normal(P - 1), R
    yield ("gram220",) + gram_tridiagonal(rng, 600, 220, 1.5) + (10,)
    yield ("rank3gram",) + gram_tridiagonal(rng, 300, 120, 0.0, rank=3) + (8,)
    d, e = wilkinson(10)
    yield "wilkinson21", d, e, 6
    yield ("glued_wilkinson", np.concatenate([d] * 5),
           np.concatenate([np.r_[e, 1e-10]] * 5)[:-1], 10)
    yield "split_repeated", np.tile([2.0, 1.0, 3.0], 4), np.tile([0.5, 0.5, 0.0], 4)[:-1], 6
    yield "zero", np.zeros(20), np.zeros(19), 4
    yield "identity", np.ones(20), np.zeros(19), 4
    yield "graded", 10.0 ** -np.arange(30), 10.0 ** -np.arange(1, 30), 5
    # far below / above the normal range of the pivot floor: solved scaled
    yield "tiny", 1e-250 * rng.standard_normal(40), 1e-250 * rng.standard_normal(39), 5
    yield "huge", 1e250 * rng.standard_normal(40), 1e250 * rng.standard_normal(39), 5
    # more than 16 wanted eigenvalues: the multisection works on groups of
    # 16 shifts over the open intervals
    yield "random200_R40", rng.standard_normal(200), rng.standard_normal(199), 40
    yield ("gram300_R24",) + gram_tridiagonal(rng, 800, 300, 1.0) + (24,)
    yield ("glued_wilkinson_R20", np.concatenate([d] * 6),
           np.concatenate([np.r_[e, 1e-10]] * 6)[:-1], 20)


@pytest.mark.parametrize("name,d,e,R", list(cases()), ids=[c[0] for c in cases()])
def test_tridiag_top_matches_eigh(harness, name, d, e, R):
    P = len(d)
    w, Z = run(harness, d, e, R)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    lam, Q = np.linalg.eigh(T)
    scale = max(np.abs(lam).max(), 1e-300)
    # eigenvalues: the R largest, descending, to a few ulps of ||T||
    assert np.abs(w - lam[::-1][:R]).max() <= 1e-14 * scale
    # eigenpairs: residual and orthonormality
    assert np.linalg.norm((T @ Z.T - Z.T * w) / scale, axis=0).max() <= 1e-13
    assert np.abs(Z @ Z.T - np.eye(R)).max() <= 1e-12
    # the wanted subspace, where a gap separates it from the rest
    if R < P and (lam[::-1][R - 1] - lam[::-1][R]) > 1e-8 * scale:
        Qr = Q[:, ::-1][:, :R]
        assert np.linalg.norm(Z.T - Qr @ (Qr.T @ Z.T)) <= 1e-10


def test_tridiag_top_rejects_bad_sizes(harness):
    p = subprocess.run([harness], input="3 4\n1 2 3\n0.5 0.5\n", capture_output=True, text=True)
    assert p.returncode != 0


def test_hoevd_gram_300_through_the_host_stage(harness):
    """The 300 x 300 HOEVD Gram of the order-4 Zipf tensor (n = 300, 6,000
    tuples, seed 8) whose top-4 eigenpairs once hung a rocSOLVER index-range
    solver: its tridiagonal form (scipy's Householder reduction) through the
    library's host stage, against numpy's eigh of the Gram itself."""
    import oracle
    import symprop_amd as sp
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 300, 6000, 1.1, 1, seed=8)
    A = oracle.hoevd_gram(4, 300, inds, vals)
    H, Q = sl.hessenberg(A, calc_q=True)
    d, e = np.diag(H).copy(), np.diag(H, 1).copy()
    R = 4
    w, Z = run(harness, d, e, R)
    lam, V = np.linalg.eigh(A)
    scale = lam.max()
    assert np.abs(w - lam[::-1][:R]).max() <= 1e-13 * scale
    X = Q @ Z.T   # eigenvectors of A
    assert np.linalg.norm(A @ X - X * w, axis=0).max() <= 1e-12 * scale
    top = V[:, ::-1][:, :R]
    assert np.allclose(np.abs(np.sum(X * top, axis=0)), 1.0, atol=1e-9)
