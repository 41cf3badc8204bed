"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each oracle function is checked against the reference's own known-answer
tests, the golden vectors written by the reference build (tests/golden), and
closed forms that a plausible slip (dropped term, wrong index or sign,
transposed operand) would break.
"""
import hashlib
import itertools
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle import ref

from conftest import GOLDEN


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        return json.load(f)


# ---- combinatorics / packed layout ------------------------------------------

def test_symtens_size_kat():
    # tests/test_symtensor.cpp:9-24
    assert oracle.symtens_size(3, 2) == 6
    assert oracle.symtens_size(4, 2) == 10
    assert oracle.symtens_size(3, 3) == 10
    assert oracle.symtens_size(5, 5) == 126


def test_symtens_size_closed_form():
    for R in range(1, 9):
        for k in range(1, 7):
            count = sum(1 for _ in itertools.combinations_with_replacement(range(R), k))
            assert oracle.symtens_size(R, k) == count == math.comb(R + k - 1, k)


def test_outer_product_kat():
    # tests/test_symtensor.cpp:26-58: packed lex order, Sym2 and Sym3 entries
    v1, v2, v3 = [1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]
    s2 = oracle.onesided(np.zeros(6), v1, v2, 1)
    s3 = oracle.onesided(np.zeros(10), s2, v3, 2)
    cnt = 0
    for i in range(3):
        for j in range(i, 3):
            assert s2[oracle.packed_index([i, j], 3)] == v1[i] * v2[j]
    for i in range(3):
        for j in range(i, 3):
            for k in range(j, 3):
                assert s3[cnt] == v1[i] * v2[j] * v3[k]
                assert oracle.packed_index([i, j, k], 3) == cnt
                cnt += 1


def test_packed_index_is_lex_rank():
    for R, k in [(3, 2), (4, 3), (5, 4), (2, 6)]:
        for rank, t in enumerate(itertools.combinations_with_replacement(range(R), k)):
            assert oracle.packed_index(list(t), R) == rank


def test_symtens_coeff_and_unpack():
    for R, k in [(3, 2), (4, 3), (3, 4)]:
        W = oracle.symtens_coeff(R, k)
        for p, t in enumerate(itertools.combinations_with_replacement(range(R), k)):
            perms = len(set(itertools.permutations(t)))
            assert W[p] == perms
        packed = np.arange(1, len(W) + 1, dtype=float)
        full = oracle.unpack(packed, R, k).reshape((R,) * k)
        for t in itertools.product(range(R), repeat=k):
            assert full[t] == packed[oracle.packed_index(sorted(t), R)]
        # ||full||^2 = sum W p^2 (the weighted norm identity used by symtens_norm)
        assert np.isclose((full ** 2).sum(), (W * packed ** 2).sum())


def test_gemmtn_kat():
    # tests/test_matrixops.cpp:8-17 (row-major A^T B with A 3x2, B 3x4)
    A = np.array([1, 2, 2, 3, 3, 4], float).reshape(3, 2)
    B = np.array([1, 2, 3, 4, 2, 3, 4, 5, 3, 4, 5, 6], float).reshape(3, 4)
    assert (A.T @ B).ravel().tolist() == [14, 20, 26, 32, 20, 29, 38, 47]


def test_multinomial_and_norm_closed_form(contact5):
    assert oracle.multinomial_sorted([1, 2, 2, 3, 4]) == math.factorial(5) // 2
    assert oracle.multinomial_sorted([7, 7, 7]) == 1
    o, d, inds, vals = contact5
    # contact-5 has no repeated index inside a tuple: ||X||^2 = nnz * 5!
    assert np.isclose(oracle.norm(o, inds, vals) ** 2, len(vals) * 120.0)
    assert oracle.norm(o, inds, vals) == golden()["contact5"]["norm"]
    t = np.array([[1, 1, 2], [2, 3, 3], [1, 2, 3]], np.uint32)
    v = np.array([2.0, -1.0, 0.5])
    assert np.isclose(oracle.norm(3, t, v) ** 2, 4 * 3 + 1 * 3 + 0.25 * 6)


# ---- loader ------------------------------------------------------------------

def test_loader_contact5(contact5):
    o, d, inds, vals = contact5
    assert (o, d, len(vals)) == (5, 245, 12704)
    assert np.all(np.diff(inds.astype(np.int64), axis=1) >= 0)
    assert np.all(vals == 1.0)
    assert inds[0].tolist() == [132, 208, 224, 244, 245]


def test_loader_sorts_and_stops_at_bad_line(tmp_path):
    p = tmp_path / "t.tns"
    p.write_text("3 9\n3 1 2 0.5\n9 9 1 2\nbad line\n4 4 4 1\n")
    o, d, inds, vals = oracle.load_tns(str(p))
    assert (o, d) == (3, 9)
    assert inds.tolist() == [[1, 2, 3], [1, 9, 9]]
    assert vals.tolist() == [0.5, 2.0]


# ---- tree construction -------------------------------------------------------

def test_tree_contact5_matches_golden(contact5):
    o, d, inds, vals = contact5
    g = golden()["contact5"]
    T = oracle.Tree(o, d, inds, vals)
    for l in range(o - 1):
        lv = T.level(l)
        ent = g["levels"][l]
        assert T.level_size(l) == ent["size"]
        assert sha(lv["level_idx"]) == ent["level_idx"]
        assert sha(lv["level_ptr"]) == ent["level_ptr"]
        if l >= 1:
            assert sha(lv["dep_list"]) == ent["dep_list"]
    rem, _ = T.owners()
    assert len(rem) == g["num_owners"] == 5 * 12704
    assert sha(rem) == g["owner_rem"]


def test_tree_small_repeats_match_golden():
    z = np.load(os.path.join(GOLDEN, "small_trees.npz"))
    ci = 0
    while "c%d_inds" % ci in z:
        t, v = z["c%d_inds" % ci], z["c%d_vals" % ci]
        N = t.shape[1]
        n = int(t.max())
        T = oracle.Tree(N, n, t, v)
        for l in range(N - 1):
            lv = T.level(l)
            assert np.array_equal(lv["level_idx"], z["c%d_l%d_idx" % (ci, l)])
            assert np.array_equal(lv["level_ptr"], z["c%d_l%d_ptr" % (ci, l)])
            if l >= 1:
                assert np.array_equal(lv["dep_list"], z["c%d_l%d_dep" % (ci, l)])
        rem, val = T.owners()
        assert np.array_equal(rem, z["c%d_rem" % ci])
        # CompressedDTree::print writes values with 6 significant digits
        assert np.allclose(val, z["c%d_val" % ci], rtol=1e-5, atol=0)
        ci += 1
    assert ci == 5


def brute_levels(N, inds):
    """Definition: distinct sorted sub-multisets of size s, lexicographic."""
    out = []
    for s in range(1, N):
        subs = set()
        for t in inds:
            for c in itertools.combinations(range(N), s):
                subs.add(tuple(int(t[i]) for i in c))
        out.append(sorted(subs))
    return out


def test_tree_definition_and_duplicate_owners():
    # (1,2,2,3,4): leaf (1,2,3,4) owned twice by rem 2 (dtree.cpp:62-80)
    t = np.array([[1, 2, 2, 3, 4]], np.uint32)
    T = oracle.Tree(5, 4, t, [1.0])
    leaf = T.level(3)
    keys = [tuple(k) for k in leaf["keys"]]
    rem, _ = T.owners()
    j = keys.index((1, 2, 3, 4))
    ptr = leaf["level_ptr"]
    assert rem[ptr[j]:ptr[j + 1]].tolist() == [2, 2]
    rng = np.random.default_rng(3)
    for _ in range(4):
        N, n = int(rng.integers(3, 6)), int(rng.integers(3, 7))
        tt = np.unique(np.sort(rng.integers(1, n + 1, size=(15, N)), axis=1), axis=0)
        T = oracle.Tree(N, n, tt, np.ones(len(tt)))
        for l, expect in enumerate(brute_levels(N, tt)):
            lv = T.level(l)
            assert [tuple(k) for k in lv["keys"]] == expect
            if l >= 1:
                prev = brute_levels(N, tt)[l - 1]
                for j, key in enumerate(expect):
                    for k in range(l + 1):
                        sub = key[:k] + key[k + 1:]
                        assert lv["dep_list"][j, k] == prev.index(sub)


@pytest.mark.skipif(not ref.available(), reason="oracle/_ref not built")
def test_tree_against_reference_build():
    rng = np.random.default_rng(11)
    for _ in range(6):
        N, n = int(rng.integers(3, 7)), int(rng.integers(3, 9))
        tt = np.unique(np.sort(rng.integers(1, n + 1, size=(30, N)), axis=1), axis=0)
        rng.shuffle(tt)
        v = rng.random(len(tt))
        T, RT = oracle.Tree(N, n, tt, v), ref.RefTree(N, n, tt, v)
        for l in range(N - 1):
            lv = T.level(l)
            assert np.array_equal(lv["level_idx"], RT.level_idx[l])
            assert np.array_equal(lv["level_ptr"], RT.level_ptr[l])
            if l >= 1:
                assert np.array_equal(lv["dep_list"], RT.dep_list(l))
        rem, _ = T.owners()
        assert np.array_equal(rem, RT.own_rem)


# ---- S3TTMc ------------------------------------------------------------------

def test_s3ttmc_contact5_golden(contact5):
    o, d, inds, vals = contact5
    T = oracle.Tree(o, d, inds, vals)
    Y = T.s3ttmc(oracle.uniform_factor(1, d, 4), 4, nthreads=1)
    G = np.load(os.path.join(GOLDEN, "contact5_y_r4.npy"))
    assert np.array_equal(Y, G)  # same schedule, one thread: bit-exact
    Y8 = T.s3ttmc(oracle.uniform_factor(1, d, 8), 8, nthreads=1)
    g8 = golden()["contact5_r8"]
    assert Y8.sum() == g8["sum"]
    assert np.allclose(Y8[[0, 100, 244]].sum(axis=1), g8["rows"], rtol=1e-15)


def test_s3ttmc_small_golden():
    z = np.load(os.path.join(GOLDEN, "small_trees.npz"))
    for ci in range(5):
        t, v = z["c%d_inds" % ci], z["c%d_vals" % ci]
        N, n = t.shape[1], int(t.max())
        T = oracle.Tree(N, n, t, v)
        Y = T.s3ttmc(z["c%d_U" % ci], 3, nthreads=1)
        assert np.array_equal(Y, z["c%d_Y" % ci])


def test_s3ttmc_dense_brute_force_with_repeats():
    rng = np.random.default_rng(5)
    for _ in range(8):
        N, n, R = int(rng.integers(3, 6)), int(rng.integers(2, 6)), int(rng.integers(1, 4))
        tt = np.unique(np.sort(rng.integers(1, n + 1, size=(12, N)), axis=1), axis=0)
        v = rng.standard_normal(len(tt))
        U = rng.standard_normal((n, R))
        Yt = oracle.Tree(N, n, tt, v).s3ttmc(U, R)
        Yd = oracle.s3ttmc_dense(N, n, tt, v, U, R)
        assert np.allclose(Yt, Yd, rtol=1e-12, atol=1e-12 * np.abs(Yd).max())


def permanent(M):
    n = M.shape[0]
    return sum(np.prod([M[i, p[i]] for i in range(n)])
               for p in itertools.permutations(range(n)))


def test_s3ttmc_single_nonzero_permanent():
    # tests/test_compressed_dtree.cpp:37-65 setting: {1,2,3,4,5}, v=0.5,
    # U=[[1,2],[2,3],[3,4],[4,5],[5,6]], R=2.  Y[rem, J] = v/prod f! * perm.
    t = np.array([[1, 2, 3, 4, 5]], np.uint32)
    U = np.array([[1, 2], [2, 3], [3, 4], [4, 5], [5, 6]], float)
    Y = oracle.Tree(5, 5, t, [0.5]).s3ttmc(U, 2)
    for rem in range(1, 6):
        leaf = [a for a in range(1, 6) if a != rem]
        for p, J in enumerate(itertools.combinations_with_replacement(range(2), 4)):
            M = np.array([[U[a - 1, j] for j in J] for a in leaf])
            assert np.isclose(Y[rem - 1, p], 0.5 * permanent(M), rtol=1e-13)


def test_s3ttmc_invariants():
    rng = np.random.default_rng(9)
    N, n, R = 4, 7, 3
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(40, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    T = oracle.Tree(N, n, tt, v)
    U = rng.standard_normal((n, R))
    Y = T.s3ttmc(U, R)
    # homogeneous of degree N-1 in U, linear in the values
    assert np.allclose(T.s3ttmc(2.0 * U, R), 2.0 ** (N - 1) * Y, rtol=1e-13)
    T2 = oracle.Tree(N, n, tt, 3.0 * v)
    assert np.allclose(T2.s3ttmc(U, R), 3.0 * Y, rtol=1e-13)
    # U^T Y unpacked is a fully symmetric order-N tensor
    core = U.T @ Y
    full = np.stack([oracle.unpack(core[r], R, N - 1) for r in range(R)])
    full = full.reshape((R,) * N)
    for perm in itertools.permutations(range(N)):
        assert np.allclose(full, np.transpose(full, perm), atol=1e-10)
    # Gram identity: unpacked Y Y^T = packed Y diag(W) Y^T
    W = oracle.symtens_coeff(R, N - 1)
    Yu = np.stack([oracle.unpack(Y[i], R, N - 1) for i in range(n)])
    assert np.allclose(Yu @ Yu.T, (Y * W) @ Y.T, rtol=1e-12)


# ---- HOEVD initialisation ----------------------------------------------------

def dense_tensor(N, n, tt, v):
    X = np.zeros((n,) * N)
    for row, val in zip(np.asarray(tt) - 1, v):
        for p in set(itertools.permutations(row)):
            X[p] = val
    return X


def hoevd_gram_leaf_form(N, n, tt, v):
    """A = sum over leaves S of mult(S) s_S s_S^T with the leaf's distinct
    owners: the form the device kernel evaluates."""
    T = oracle.Tree(N, n, tt, v)
    lv = T.level(T.depth - 1)
    rem, val = T.owners()
    A = np.zeros((n, n))
    for s, key in enumerate(lv["keys"]):
        mult = math.factorial(N - 1) / np.prod(
            [math.factorial(c) for c in np.unique(key, return_counts=True)[1]])
        lo, hi = int(lv["level_ptr"][s]), int(lv["level_ptr"][s + 1])
        own = {}
        for o in range(lo, hi):
            own[int(rem[o]) - 1] = val[o]   # a repeated owner is the same entry
        idx = np.array(sorted(own))
        sv = np.array([own[i] for i in idx])
        A[np.ix_(idx, idx)] += mult * np.outer(sv, sv)
    return A


def test_hoevd_gram_is_the_unfolding_gram():
    # spsymtensor.cpp:385-439 against X_(1) X_(1)^T of the dense tensor
    rng = np.random.default_rng(11)
    for _ in range(8):
        N, n = int(rng.integers(2, 6)), int(rng.integers(2, 6))
        tt = np.unique(np.sort(rng.integers(1, n + 1, size=(10, N)), axis=1), axis=0)
        v = rng.standard_normal(len(tt))
        A = oracle.hoevd_gram(N, n, tt, v)
        X1 = dense_tensor(N, n, tt, v).reshape(n, -1)
        assert np.allclose(A, X1 @ X1.T, rtol=1e-13, atol=1e-13)
        assert np.array_equal(A, A.T)
        # trace = sum of squares of all entries = ||X||^2
        assert np.isclose(np.trace(A), oracle.norm(N, tt, v) ** 2, rtol=1e-13)


def test_hoevd_gram_repeated_index_counts_once():
    # (1,2,2,3): X_(1) row 2 has the columns (1,2,3) in 3! orderings, each once
    tt = np.array([[1, 2, 2, 3]], np.uint32)
    A = oracle.hoevd_gram(4, 3, tt, [2.0])
    # row 1: columns = orderings of (2,2,3): 3 of them; row 2: of (1,2,3): 6;
    # row 3: of (1,2,2): 3
    assert np.allclose(np.diag(A), [3 * 4.0, 6 * 4.0, 3 * 4.0])
    assert np.allclose(A - np.diag(np.diag(A)), 0.0)


def test_hoevd_leaf_form_matches_coo_form():
    rng = np.random.default_rng(12)
    for _ in range(8):
        N, n = int(rng.integers(3, 7)), int(rng.integers(3, 8))
        tt = np.unique(np.sort(rng.integers(1, n + 1, size=(30, N)), axis=1), axis=0)
        v = rng.random(len(tt))
        assert np.allclose(hoevd_gram_leaf_form(N, n, tt, v),
                           oracle.hoevd_gram(N, n, tt, v), rtol=1e-13, atol=1e-13)


def test_hoevd_factor_is_top_eigenspace():
    rng = np.random.default_rng(13)
    N, n, R = 4, 9, 3
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(60, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    A = oracle.hoevd_gram(N, n, tt, v)
    U, w = oracle.hoevd_factor(A, R)
    assert np.all(np.diff(w) >= 0)                      # ascending, as syevd's tail
    assert np.allclose(A @ U, U * w, atol=1e-10 * w.max())
    # the largest R of the whole spectrum
    assert np.allclose(np.sort(np.linalg.eigvalsh(A))[-R:], w)


def test_s3ttmc_threads_agree(contact5):
    o, d, inds, vals = contact5
    T = oracle.Tree(o, d, inds, vals)
    U = oracle.uniform_factor(1, d, 5)
    Y1 = T.s3ttmc(U, 5, nthreads=1)
    Y4 = T.s3ttmc(U, 5, nthreads=4)
    assert np.allclose(Y1, Y4, rtol=1e-12)


def test_random_factor_is_libstdcxx_normal():
    a = oracle.random_factor(0, 3, 2)
    b = oracle.random_factor(0, 3, 2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, oracle.random_factor(1, 3, 2))
    big = oracle.random_factor(0, 20000, 5)
    assert abs(big.mean()) < 0.02 and abs(big.std() - 1.0) < 0.02


# ---- Tucker loops ----------------------------------------------------------

def exact_rank_tensor(n=7, R=2, N=3, seed=0):
    """Dense symmetric tensor with multilinear rank R, as unique sorted tuples."""
    rng = np.random.default_rng(seed)
    U0 = np.linalg.qr(rng.standard_normal((n, R)))[0]
    C = rng.standard_normal((R,) * N)
    C = sum(np.transpose(C, p) for p in itertools.permutations(range(N)))
    X = C
    for _ in range(N):
        X = np.tensordot(X, U0, axes=([0], [1]))
    tuples = np.array(list(itertools.combinations_with_replacement(range(1, n + 1), N)),
                      np.uint32)
    vals = np.array([X[tuple(t - 1)] for t in tuples])
    return tuples, vals, X, U0


def dense_fit_error(X, U):
    Xc = X
    for _ in range(X.ndim):
        Xc = np.tensordot(Xc, U, axes=([0], [0]))
    return np.sqrt(max((X ** 2).sum() - (Xc ** 2).sum(), 0.0)) / np.sqrt((X ** 2).sum())


def test_tucker_oracle_recovers_exact_rank():
    from oracle import tucker
    t, v, X, U0 = exact_rank_tensor()
    T = oracle.Tree(3, 7, t, v)
    assert np.isclose(oracle.norm(3, t, v), np.sqrt((X ** 2).sum()), rtol=1e-13)
    for alg, iters, tol in [(tucker.hooi, 10, 1e-6), (tucker.althooi, 10, 1e-6),
                            (tucker.hoqri, 200, 1e-4)]:
        U, errs, _ = alg(T, 2, iters, rel_tol=1e-12, delta=100.0)
        assert errs[-1] < tol, (alg.__name__, errs[-1])
        assert tucker.subspace_cos(U, U0) > 1 - 1e-6
        assert np.isclose(dense_fit_error(X, U), errs[-1], atol=1e-6)


def test_tucker_oracle_error_formula_matches_dense():
    from oracle import tucker
    rng = np.random.default_rng(4)
    n, N, R = 6, 4, 2
    t = np.unique(np.sort(rng.integers(1, n + 1, size=(60, N)), axis=1), axis=0)
    v = rng.random(len(t))
    T = oracle.Tree(N, n, t, v)
    X = np.zeros((n,) * N)
    for tt, vv in zip(t, v):
        for p in set(itertools.permutations(tt - 1)):
            X[p] = vv
    # one iteration from a given factor: the reported error contracts mode 1
    # with the new factor and modes 2..N with the regularised input factor
    U_in = rng.standard_normal((n, R))
    U, errs, _ = tucker.hooi(T, R, 1, rel_tol=0, delta=0.5, U0=U_in)
    U_reg = tucker.regularize(U_in, 0.5)
    Xc = np.tensordot(X, U, axes=([0], [0]))
    for _ in range(N - 1):
        Xc = np.tensordot(Xc, U_reg, axes=([0], [0]))
    expect = np.sqrt((X ** 2).sum() - (Xc ** 2).sum()) / np.sqrt((X ** 2).sum())
    assert np.isclose(expect, errs[-1], rtol=1e-10)
