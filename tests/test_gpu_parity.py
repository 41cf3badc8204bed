"""GPU parity: the HIP path (through the C-ABI) against the oracle.

Integer work (tree construction) must be bit-exact; FP64 S3TTMc must agree
with the oracle's opt5 restatement to a W-weighted relative Frobenius error
<= 1e-12 (different summation order); FP32 to <= 2e-5.
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
import symprop_amd as sp

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def werr(Y, Yref, R, order):
    W = oracle.symtens_coeff(R, order - 1)
    num = np.sqrt(((Y - Yref) ** 2 * W).sum())
    den = np.sqrt((Yref ** 2 * W).sum())
    return num / max(den, 1e-300)


def check_tree(t, o, d, inds, vals):
    T = oracle.Tree(o, d, inds, vals)
    assert t.depth() == o - 1
    for l in range(o - 1):
        a, b = t.level(l), T.level(l)
        assert t.level_size(l) == T.level_size(l), l
        assert np.array_equal(a["keys"], b["keys"]), l
        assert np.array_equal(a["level_idx"], b["level_idx"]), l
        assert np.array_equal(a["level_ptr"], b["level_ptr"]), l
        if l >= 1:
            assert np.array_equal(a["dep_list"], b["dep_list"]), l
    r1, v1 = t.owners()
    r2, v2 = T.owners()
    assert np.array_equal(r1, r2) and np.array_equal(v1, v2)
    return T


def test_construction_contact5(contact5):
    o, d, inds, vals = contact5
    say("construction contact-5")
    t = sp.SpSymtensor(d, o, inds, vals)
    check_tree(t, o, d, inds, vals)
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        g = json.load(f)["contact5"]
    assert [t.level_size(l) for l in range(4)] == [e["size"] for e in g["levels"]]
    assert t.norm() == g["norm"]


def test_ctx_create_rejects_bad_tuples():
    import ctypes as C
    L = sp.lib()
    h = C.c_void_p()
    v = np.ones(3)
    say("ctx create: unsorted and out-of-range tuples")
    t = np.array([[1, 2, 3], [2, 1, 3], [1, 1, 1]], np.uint32)
    rc = L.sp_ctx_create(3, 5, 3, t.ctypes.data_as(sp.u32p), v.ctypes.data_as(sp.f64p),
                         C.byref(h))
    assert rc == sp.SP_ERR_ARG and b"sorted" in L.sp_last_error() and not h.value
    for bad in ([1, 2, 9], [0, 1, 2]):
        t = np.array([[1, 2, 3], bad, [1, 1, 1]], np.uint32)
        rc = L.sp_ctx_create(3, 5, 3, t.ctypes.data_as(sp.u32p), v.ctypes.data_as(sp.f64p),
                             C.byref(h))
        assert rc == sp.SP_ERR_ARG and b"outside" in L.sp_last_error() and not h.value
    # the norm: a fixed-order sum of v^2 N!/prod f! (here 6 + 3 + 1 = 10 for values 1)
    t = sp.SpSymtensor(5, 3, np.array([[1, 2, 3], [1, 1, 2], [4, 4, 4]], np.uint32), v)
    assert t.norm() == np.sqrt(10.0)


def test_construction_small_golden():
    z = np.load(os.path.join(GOLDEN, "small_trees.npz"))
    for ci in range(5):
        tt, v = z["c%d_inds" % ci], z["c%d_vals" % ci]
        N, n = tt.shape[1], int(tt.max())
        say("construction golden case", ci, N, n)
        t = sp.SpSymtensor(n, N, tt, v)
        for l in range(N - 1):
            lv = t.level(l)
            assert np.array_equal(lv["level_idx"], z["c%d_l%d_idx" % (ci, l)])
            assert np.array_equal(lv["level_ptr"], z["c%d_l%d_ptr" % (ci, l)])
            if l >= 1:
                assert np.array_equal(lv["dep_list"], z["c%d_l%d_dep" % (ci, l)])
        rem, _ = t.owners()
        assert np.array_equal(rem, z["c%d_rem" % ci])


@pytest.mark.parametrize("N,n,cnt,seed", [
    (3, 9, 60, 1), (4, 30, 400, 2), (5, 12, 300, 3), (6, 8, 200, 4),
    (7, 6, 150, 5),
    (5, 70000, 3000, 6),   # 17-bit indices: 128-bit leaf keys
    (7, 100000, 2000, 7),  # 102-bit leaf keys
    (3, 1, 1, 8),          # single index value
])
def test_construction_random(N, n, cnt, seed):
    rng = np.random.default_rng(seed)
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(cnt, N)), axis=1), axis=0)
    rng.shuffle(tt)
    v = rng.random(len(tt))
    say("construction random", N, n, len(tt))
    t = sp.SpSymtensor(n, N, tt, v)
    check_tree(t, N, n, tt, v)


def test_s3ttmc_contact5_golden(contact5):
    o, d, inds, vals = contact5
    t = sp.SpSymtensor(d, o, inds, vals)
    say("s3ttmc contact-5 R=4")
    Y = t.s3ttmc(oracle.uniform_factor(1, d, 4))
    G = np.load(os.path.join(GOLDEN, "contact5_y_r4.npy"))
    assert werr(Y, G, 4, o) <= 1e-12
    say("s3ttmc contact-5 R=8")
    U8 = oracle.uniform_factor(1, d, 8)
    Y8 = t.s3ttmc(U8)
    Yo = oracle.Tree(o, d, inds, vals).s3ttmc(U8, 8)
    assert werr(Y8, Yo, 8, o) <= 1e-12


def test_s3ttmc_small_golden():
    z = np.load(os.path.join(GOLDEN, "small_trees.npz"))
    for ci in range(5):
        tt, v = z["c%d_inds" % ci], z["c%d_vals" % ci]
        N, n = tt.shape[1], int(tt.max())
        say("s3ttmc golden case", ci)
        t = sp.SpSymtensor(n, N, tt, v)
        Y = t.s3ttmc(z["c%d_U" % ci])
        assert werr(Y, z["c%d_Y" % ci], 3, N) <= 1e-12


@pytest.mark.parametrize("N,n,cnt,R,seed", [
    (3, 40, 500, 1, 1), (3, 40, 500, 7, 2), (4, 25, 800, 2, 3), (4, 60, 3000, 10, 4),
    (5, 20, 900, 3, 5), (5, 30, 2000, 8, 6), (6, 12, 600, 4, 7), (6, 15, 1500, 8, 8),
    (7, 9, 400, 3, 9), (7, 10, 800, 6, 10), (8, 6, 200, 3, 11), (4, 7, 300, 16, 12),
])
def test_s3ttmc_random_vs_oracle(N, n, cnt, R, seed):
    rng = np.random.default_rng(seed)
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(cnt, N)), axis=1), axis=0)
    v = rng.standard_normal(len(tt))
    U = rng.standard_normal((n, R))
    say("s3ttmc random", N, n, len(tt), R)
    t = sp.SpSymtensor(n, N, tt, v)
    Y = t.s3ttmc(U)
    Yo = oracle.Tree(N, n, tt, v).s3ttmc(U, R, nthreads=4)
    assert Y.shape == Yo.shape
    assert werr(Y, Yo, R, N) <= 1e-12
    if n <= 12 and N <= 6:
        Yd = oracle.s3ttmc_dense(N, n, tt, v, U, R)
        assert werr(Y, Yd, R, N) <= 1e-11


FAST_SHAPES = ([(3, R) for R in range(2, 17)] + [(4, R) for R in range(2, 17)] +
               [(5, R) for R in range(2, 13)] + [(6, R) for R in range(2, 9)] +
               [(7, R) for R in range(2, 7)])


@pytest.mark.parametrize("N,R", FAST_SHAPES)
def test_s3ttmc_fast_kernel_every_instantiated_shape(N, R):
    # every (order, rank) with a compiled MFMA kernel: FP64 against the
    # oracle, the FP32 and BF16 forms against FP64, and the MFMA kernels ran
    n = {3: 200, 4: 120, 5: 60, 6: 40, 7: 25}[N]
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, 3000, 1.1, 1, seed=N * 100 + R)
    U = np.linalg.qr(np.random.default_rng(R).standard_normal((n, R)))[0]
    say("s3ttmc fast shape", N, R, len(vals))
    t = sp.SpSymtensor(n, N, inds, vals)
    t.timing(True)
    Y = t.s3ttmc(U)
    assert t.timing_get("s3ttmc_mfma64")[1] == 1
    Yo = oracle.Tree(N, n, inds, vals).s3ttmc(U, R, nthreads=4)
    assert werr(Y, Yo, R, N) <= 1e-12
    assert werr(t.s3ttmc(U, precision=sp.PREC_FP32), Y, R, N) < 2e-5
    assert werr(t.s3ttmc(U, precision=sp.PREC_BF16), Y, R, N) < 1e-2
    assert t.timing_get("s3ttmc_mfma_f32")[1] == 1
    assert t.timing_get("s3ttmc_mfma_bf16")[1] == 1


def test_s3ttmc_hub_rows_split():
    # Zipf tensor: hub rows carry far more owners than one work chunk
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 2000, 60000, 1.1, 1, seed=3)
    R = 10
    U = oracle.random_factor(0, 2000, R)
    say("s3ttmc zipf hubs", len(vals))
    t = sp.SpSymtensor(2000, 4, inds, vals)
    Y = t.s3ttmc(U)
    Yo = oracle.Tree(4, 2000, inds, vals).s3ttmc(U, R, nthreads=8)
    assert werr(Y, Yo, R, 4) <= 1e-12
    # bit-identical repeat (no atomics)
    t.s3ttmc_dev()
    assert np.array_equal(t.y(), Y)


def test_s3ttmc_hub_rows_split_order6():
    inds, vals = sp.generate(sp.GEN_ZIPF, 6, 600, 20000, 1.2, 1, seed=5)
    R = 8
    U = oracle.random_factor(1, 600, R)
    say("s3ttmc zipf order 6 hubs", len(vals))
    t = sp.SpSymtensor(600, 6, inds, vals)
    Y = t.s3ttmc(U)
    Yo = oracle.Tree(6, 600, inds, vals).s3ttmc(U, R, nthreads=8)
    assert werr(Y, Yo, R, 6) <= 1e-12
    t.s3ttmc_dev()
    assert np.array_equal(t.y(), Y)


def test_s3ttmc_fp32(contact5):
    o, d, inds, vals = contact5
    t = sp.SpSymtensor(d, o, inds, vals)
    U = oracle.random_factor(0, d, 8)
    say("s3ttmc fp32 contact-5")
    Y32 = t.s3ttmc(U, precision=sp.PREC_FP32)
    Yo = oracle.Tree(o, d, inds, vals).s3ttmc(U, 8)
    assert werr(Y32, Yo, 8, o) <= 2e-5


@pytest.mark.parametrize("N,n,nnz,R,seed", [(4, 2000, 40000, 10, 1), (5, 300, 6000, 8, 2),
                                            (6, 400, 5000, 8, 3), (7, 60, 1500, 6, 4),
                                            (3, 500, 8000, 4, 5)])
def test_s3ttmc_fp32_mfma(N, n, nnz, R, seed):
    # FP32 operands and accumulation on v_mfma_f32_16x16x4f32: FP32-level
    # error, and the MFMA kernel (not the generic one) ran
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, nnz, 1.1, 1, seed=seed)
    U = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, R)))[0]
    say("s3ttmc fp32 mfma", N, n, len(vals), R)
    t = sp.SpSymtensor(n, N, inds, vals)
    Y32 = t.s3ttmc(U, precision=sp.PREC_FP32)
    Y64 = t.s3ttmc(U)
    e = werr(Y32, Y64, R, N)
    assert 1e-12 < e < 2e-5, e
    t.timing(True)
    t.s3ttmc_dev(sp.PREC_FP32)
    assert t.timing_get("s3ttmc_mfma_f32")[1] == 1
    # deterministic
    assert np.array_equal(t.s3ttmc(U, precision=sp.PREC_FP32), Y32)


@pytest.mark.parametrize("N,n,nnz,R,seed", [(4, 2000, 40000, 10, 1), (5, 300, 6000, 8, 2),
                                            (6, 400, 5000, 8, 3), (7, 60, 1500, 6, 4)])
def test_s3ttmc_bf16_mfma(N, n, nnz, R, seed):
    # BF16 operands, FP32 accumulation: error well below 1e-2 and well above
    # FP64 rounding (shows that the reduced-precision kernel ran)
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, nnz, 1.1, 1, seed=seed)
    U = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, R)))[0]
    say("s3ttmc bf16", N, n, len(vals), R)
    t = sp.SpSymtensor(n, N, inds, vals)
    Ybf = t.s3ttmc(U, precision=sp.PREC_BF16)
    Y64 = t.s3ttmc(U)
    e = werr(Ybf, Y64, R, N)
    assert 1e-7 < e < 1e-2, e
    t.timing(True)
    t.s3ttmc_dev(sp.PREC_BF16)
    assert t.timing_get("s3ttmc_mfma_bf16")[1] == 1


def test_dense_ops_vs_numpy():
    rng = np.random.default_rng(21)
    N, n, R = 4, 50, 5
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(2500, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    t = sp.SpSymtensor(n, N, tt, v)
    U = np.linalg.qr(rng.standard_normal((n, R)))[0]
    say("dense ops")
    Y = t.s3ttmc(U)
    W = oracle.symtens_coeff(R, N - 1)
    core = t.gemmtn()
    assert np.allclose(core, U.T @ Y, rtol=1e-12, atol=1e-12 * np.abs(core).max())
    cn = t.core_norm()
    assert np.isclose(cn, np.sqrt((core ** 2 * W).sum()), rtol=1e-13)
    G = t.gram()
    sw = np.sqrt(W)
    Gref = (Y * sw).T @ (Y * sw)
    assert np.allclose(G, Gref, rtol=1e-11, atol=1e-12 * np.abs(Gref).max())
    t.symgemm()
    F = t.factor()
    assert np.allclose(F, (Y * W) @ core.T, rtol=1e-11, atol=1e-12 * np.abs(F).max())
    t.qr()
    Q = t.factor()
    assert np.allclose(Q.T @ Q, np.eye(R), atol=1e-12)
    Qref = np.linalg.qr((Y * W) @ core.T)[0]
    assert np.allclose(np.abs(Q.T @ Qref), np.eye(R), atol=1e-8)


def test_rank_change_makes_y_and_core_stale():
    # A factor of another rank changes the packed width P; the device Y and
    # core of the previous rank must then be refused, not read with the new
    # width (they are P_old wide).
    rng = np.random.default_rng(22)
    N, n = 4, 40
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(1500, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    t = sp.SpSymtensor(n, N, tt, v)
    say("rank change: S3TTMc at R=4, factor at R=8")
    t.s3ttmc(np.linalg.qr(rng.standard_normal((n, 4)))[0])
    t.gemmtn()
    U8 = np.linalg.qr(rng.standard_normal((n, 8)))[0]
    t.set_factor(U8)
    for call in (t.y, t.gram, t.gemmtn, t.symgemm, t.leftsvd_y, t.core_norm):
        with pytest.raises(sp.SpError) as e:
            call()
        assert e.value.code == sp.SP_ERR_STATE, call
    # after an S3TTMc at the new rank everything works again
    Y = t.s3ttmc(U8)
    assert Y.shape == (n, oracle.symtens_coeff(8, N - 1).size)
    core = t.gemmtn()
    assert np.allclose(core, U8.T @ Y, rtol=1e-12, atol=1e-12 * np.abs(core).max())
    t.core_norm()
    # and back to the smaller rank: stale again until S3TTMc runs
    t.set_factor(U8[:, :4].copy())
    with pytest.raises(sp.SpError):
        t.gram()
    with pytest.raises(sp.SpError):
        t.core_norm()


@pytest.mark.parametrize("cond", [1.0, 1e5, 1e10])
def test_qr_well_and_ill_conditioned(cond):
    # CholeskyQR2 for well-conditioned factors, Householder (rocSOLVER) past the
    # Cholesky pivot guard; both must give Q = Householder Q up to column signs
    rng = np.random.default_rng(31)
    n, R = 3000, 8
    t = sp.SpSymtensor(n, 3, np.array([[1, 2, 3]], np.uint32), np.ones(1))
    A = np.linalg.qr(rng.standard_normal((n, R)))[0]
    B = np.linalg.qr(rng.standard_normal((R, R)))[0]
    U = (A * np.logspace(0, -np.log10(cond), R)) @ B
    t.set_factor(U)
    say("qr cond", cond)
    t.qr()
    Q = t.factor()
    assert np.allclose(Q.T @ Q, np.eye(R), atol=1e-12)
    Qref = np.linalg.qr(U)[0]
    assert np.allclose(np.abs(Q.T @ Qref), np.eye(R), atol=1e-5 if cond > 1e6 else 1e-10)


def spectrum_matrix(P, lam, seed):
    Q = np.linalg.qr(np.random.default_rng(seed).standard_normal((P, P)))[0]
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("eig", ["syevd", "syevj", "tridiag", "tridiag-wg", "tridiag-reg",
                                 "tridiag-rocsolver"])
@pytest.mark.parametrize("P,R,kind", [(1, 1, "random"), (2, 1, "random"), (3, 2, "random"),
                                      (32, 4, "random"), (33, 5, "gram"), (50, 8, "random"),
                                      (97, 8, "random"), (130, 6, "clusters"),
                                      (220, 10, "gram"), (240, 8, "rank3"), (256, 12, "gram"),
                                      (462, 6, "clusters"), (300, 8, "rank3"),
                                      (792, 8, "gram"), (1716, 8, "gram"),
                                      (2002, 8, "gram")])  # beyond the LDS-resident reduction
def test_syevd_top_matches_numpy(P, R, kind, eig, monkeypatch):
    # matrix::syevd: top-R eigenvectors, ascending, against numpy's LAPACK
    if eig == "syevj" and P > 500:
        pytest.skip("Jacobi sweeps at this order take seconds; the order is covered by the others")
    if eig in ("tridiag-wg", "tridiag-reg"):
        # the multi-workgroup reductions (rows in LDS; rows in registers) at
        # every order, also where the default takes another form
        if eig == "tridiag-reg" and P > 1024:
            pytest.skip("the same reduction as 'tridiag' at this order")
        monkeypatch.setenv("SP_SYTRD", eig.split("-")[1])
        eig = "tridiag"
    monkeypatch.setenv("SP_EIG", eig)
    rng = np.random.default_rng(P + R)
    if kind == "random":
        A = rng.standard_normal((P, P))
        A = A + A.T
    elif kind == "gram":
        # Gram-like spectrum: fast decay, then a flat tail
        lam = np.sort(np.concatenate([1e5 * 0.5 ** np.arange(R + 4), rng.random(P - R - 4)]))
        A = spectrum_matrix(P, lam, P)
    elif kind == "clusters":
        # tight clusters inside the top R, a gap of 2x below them
        top = np.repeat([10.0, 7.0, 7.0 + 1e-9], [2, 2, 2])
        lam = np.sort(np.concatenate([top, 3.0 * rng.random(P - R)]))
        A = spectrum_matrix(P, lam, P)
    else:  # rank 3: zero eigenvalues inside the wanted set
        B = rng.standard_normal((P, 3))
        A = B @ B.T
    say("syevd_top", P, R, kind, eig)
    t = sp.SpSymtensor(4, 3, np.array([[1, 2, 3]], dtype=np.uint32), np.ones(1))
    U, w = t.syevd_top(A, R)
    lam, V = np.linalg.eigh(A)
    scale = np.abs(lam).max()
    assert np.allclose(w, lam[P - R:], rtol=0, atol=1e-12 * scale)
    assert np.allclose(U.T @ U, np.eye(R), atol=1e-11)
    assert np.abs(A @ U - U * w).max() <= 1e-11 * scale
    if kind != "rank3":
        # the wanted subspace (separated from the rest by a gap)
        c = np.linalg.svd(U.T @ V[:, P - R:], compute_uv=False)
        assert c.min() > 1 - 1e-10


def test_syevd_top_reduction_timeout_falls_back(monkeypatch):
    """A tiny spin limit on the reduction's own bounded waits (set from the
    host) makes its workgroups give up; the library restores the matrix,
    counts the time-out and reduces it with rocSOLVER instead, with the same
    results."""
    monkeypatch.setenv("SP_TRD_SPIN", "1")
    P, R = 792, 8
    rng = np.random.default_rng(5)
    lam = np.sort(np.concatenate([1e5 * 0.5 ** np.arange(R + 4), rng.random(P - R - 4)]))
    A = spectrum_matrix(P, lam, 7)
    say("syevd_top timeout fallback", P, R)
    t = sp.SpSymtensor(4, 3, np.array([[1, 2, 3]], dtype=np.uint32), np.ones(1))
    t.timing(True)
    fallbacks = 0
    for _ in range(3):
        U, w = t.syevd_top(A, R)
        lam_ref, V = np.linalg.eigh(A)
        assert np.allclose(w, lam_ref[P - R:], rtol=0, atol=1e-12 * lam.max())
        assert np.abs(A @ U - U * w).max() <= 1e-11 * lam.max()
        c = np.linalg.svd(U.T @ V[:, P - R:], compute_uv=False)
        assert c.min() > 1 - 1e-10
    _, fallbacks = t.timing_get("sytrd_timeout")
    t.timing(False)
    assert fallbacks >= 1


@pytest.mark.parametrize("eig", ["syevd", "syevj", "tridiag", "tridiag-rocsolver"])
def test_leftsvd_subspace(eig, monkeypatch):
    monkeypatch.setenv("SP_EIG", eig)
    rng = np.random.default_rng(22)
    N, n, R = 4, 80, 4
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(4000, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    t = sp.SpSymtensor(n, N, tt, v)
    U = np.linalg.qr(rng.standard_normal((n, R)))[0]
    say("leftsvd")
    Y = t.s3ttmc(U)
    sig = t.leftsvd_y()
    Un = t.factor()
    Yu = np.stack([oracle.unpack(Y[i], R, N - 1) for i in range(n)])
    Us, s, _ = np.linalg.svd(Yu, full_matrices=False)
    assert np.allclose(sig, s[:R], rtol=1e-10)
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-10)
    # same subspace: singular values of Un^T Us all 1
    c = np.linalg.svd(Un.T @ Us[:, :R], compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-9)


@pytest.mark.parametrize("eig", ["tridiag", "tridiag-rocsolver", "syevd"])
@pytest.mark.parametrize("factor", [1e-200, 1e200])
def test_syevd_top_far_from_unit_scale(factor, eig, monkeypatch):
    # eigenpairs of c*A are c times A's with the same vectors; with c far from
    # 1 the eigen-step scales the matrix by a power of two first (as LAPACK's
    # dsyevd does), so its pivots and sums of squares stay in range
    monkeypatch.setenv("SP_EIG", eig)
    P, R = 220, 10
    rng = np.random.default_rng(31)
    lam = np.sort(np.concatenate([0.5 ** np.arange(R + 4), 1e-3 * rng.random(P - R - 4)]))
    A = spectrum_matrix(P, lam, 3) * factor
    say("syevd_top scaled", factor, eig)
    t = sp.SpSymtensor(4, 3, np.array([[1, 2, 3]], dtype=np.uint32), np.ones(1))
    U, w = t.syevd_top(A, R)
    assert np.all(np.isfinite(U)) and np.all(np.isfinite(w))
    assert np.allclose(w / factor, lam[P - R:], rtol=0, atol=1e-12)
    assert np.allclose(U.T @ U, np.eye(R), atol=1e-11)
    assert np.abs((A @ U - U * w) / factor).max() <= 1e-11


def test_leftsvd_scale_invariant():
    # U and 1e-40*U span the same space: for order 4 the second Y is 1e-120
    # times the first and its Gram about 1e-240 times; leftsvd must return
    # the same subspace, with singular values scaled by 1e-120
    rng = np.random.default_rng(23)
    N, n, R = 4, 80, 4
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(4000, N)), axis=1), axis=0)
    t = sp.SpSymtensor(n, N, tt, rng.random(len(tt)))
    U = np.linalg.qr(rng.standard_normal((n, R)))[0]
    say("leftsvd scale invariance")
    t.s3ttmc(U)
    sig1 = t.leftsvd_y()
    U1 = t.factor()
    t.s3ttmc(1e-40 * U)
    sig2 = t.leftsvd_y()
    U2 = t.factor()
    assert np.all(np.isfinite(U2))
    assert np.allclose(sig2 / 1e-120, sig1, rtol=1e-10)
    c = np.linalg.svd(U1.T @ U2, compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-9)


def test_leftsvd_rank_deficient_completes_basis():
    # one nonzero (1,2,3): Y has 3 nonzero rows, so rank(Y) <= 3 < R = 4;
    # LAPACK's SVD returns an orthonormal completion for the missing direction
    n, R = 12, 4
    t = sp.SpSymtensor(n, 3, np.array([[1, 2, 3]], np.uint32), np.ones(1))
    U = np.random.default_rng(3).standard_normal((n, R))
    say("leftsvd rank deficient")
    Y = t.s3ttmc(U)
    t.leftsvd_y()
    Un = t.factor()
    assert np.all(np.isfinite(Un))
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)
    W = oracle.symtens_coeff(R, 2)
    Z = Y * np.sqrt(W)
    Us, s, _ = np.linalg.svd(Z, full_matrices=False)
    k = int((s > s[0] * 1e-12).sum())
    assert k < R
    # the defined directions are kept
    c = np.linalg.svd(Un[:, :k].T @ Us[:, :k], compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-10)
    # a zero Y (no nonzero touches the factor's support) still gives a basis
    U0 = np.zeros((n, R))
    U0[6:10, :] = np.eye(R)
    t.s3ttmc(U0)
    t.leftsvd_y()
    Un = t.factor()
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)


def test_regularize_rank_deficient_input():
    n, R = 50, 4
    t = sp.SpSymtensor(n, 3, np.array([[1, 2, 3]], np.uint32), np.ones(1))
    rng = np.random.default_rng(4)
    U = rng.standard_normal((n, R))
    U[:, 3] = U[:, 0]               # two equal columns: a 3-dim span
    t.set_factor(U)
    say("regularize rank deficient")
    t.regularize(10.0)
    Un = t.factor()
    assert np.all(np.isfinite(Un))
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)
    # span(U) is inside span(Un)
    Q = np.linalg.qr(U[:, :3])[0]
    c = np.linalg.svd(Q.T @ Un, compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-10)


def test_regularize_matches_reference_math():
    rng = np.random.default_rng(23)
    n, R, delta = 60, 4, 0.1
    tt = np.array([[1, 2, 3], [2, 3, 4]], np.uint32)
    t = sp.SpSymtensor(n, 3, tt, np.ones(2))
    U = rng.standard_normal((n, R))
    t.set_factor(U)
    say("regularize")
    t.regularize(delta)
    Un = t.factor()
    # spsymtensor.cpp:168-199: clip rows to norm delta, then left singular vectors
    nr = np.linalg.norm(U, axis=1, keepdims=True)
    Uc = np.where(nr > delta, U * (delta / nr), U)
    Us = np.linalg.svd(Uc, full_matrices=False)[0]
    assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)
    c = np.linalg.svd(Un.T @ Us, compute_uv=False)
    assert np.allclose(c, 1.0, atol=1e-10)
    # columns come out as the singular vectors themselves (up to sign)
    assert np.allclose(np.abs(np.sum(Un * Us, axis=0)), 1.0, atol=1e-8)


@pytest.mark.parametrize("R", [3, 10, 16, 17])
def test_regularize_of_orthonormal_factor(R):
    # Repeated calls, as in the Tucker loops: for R <= 16 the symmetric second
    # pass runs its R x R step on the device (one-wave Jacobi), above that on
    # the host; both must give the clipped factor's left singular vectors
    rng = np.random.default_rng(40 + R)
    n, delta = 300, 0.2
    t = sp.SpSymtensor(n, 3, np.array([[1, 2, 3], [2, 3, 4]], np.uint32), np.ones(2))
    t.set_factor(rng.standard_normal((n, R)))
    say("regularize twice", R)
    t.regularize(delta)
    U1 = t.factor()
    for _ in range(2):
        t.regularize(delta)
        Un = t.factor()
        nr = np.linalg.norm(U1, axis=1, keepdims=True)
        Uc = np.where(nr > delta, U1 * (delta / nr), U1)
        Us, sv, _ = np.linalg.svd(Uc, full_matrices=False)
        assert np.allclose(Un.T @ Un, np.eye(R), atol=1e-12)
        assert np.allclose(np.linalg.svd(Un.T @ Us, compute_uv=False), 1.0, atol=1e-10)
        # singular vectors up to sign where the singular values are apart
        gap = np.minimum(np.r_[np.inf, sv[:-1] - sv[1:]], np.r_[sv[:-1] - sv[1:], np.inf])
        sep = gap > 1e-6 * sv[0]
        assert np.allclose(np.abs(np.sum(Un * Us, axis=0))[sep], 1.0, atol=1e-8)
        U1 = Un


# ---- HOEVD initialisation (spsymtensor.cpp:363-439) --------------------------

@pytest.mark.parametrize("N,n,cnt,seed", [(3, 7, 40, 1), (4, 9, 120, 2), (5, 6, 80, 3),
                                          (6, 5, 60, 4), (3, 40, 900, 5)])
def test_hoevd_gram_small(N, n, cnt, seed):
    rng = np.random.default_rng(seed)
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(cnt, N)), axis=1), axis=0)
    v = rng.standard_normal(len(tt))
    say("hoevd gram", N, n, len(tt))
    t = sp.SpSymtensor(n, N, tt, v)
    A = t.hoevd_gram()
    Ao = oracle.hoevd_gram(N, n, tt, v)
    assert np.array_equal(A, A.T)
    assert np.allclose(A, Ao, rtol=1e-13, atol=1e-13 * np.abs(Ao).max())


def test_hoevd_gram_hubs_and_column_tiles():
    # Zipf hubs: leaves with far more owners than one wavefront stages, and
    # dim > 4096 so rows span two column tiles
    inds, vals = sp.generate(sp.GEN_ZIPF, 3, 4500, 30000, 1.1, 1, seed=7)
    say("hoevd gram zipf", len(vals))
    t = sp.SpSymtensor(4500, 3, inds, vals)
    A = t.hoevd_gram()
    Ao = oracle.hoevd_gram(3, 4500, inds, vals)
    assert np.array_equal(A, A.T)
    assert np.allclose(A, Ao, rtol=1e-13, atol=1e-13 * np.abs(Ao).max())
    assert np.isclose(np.trace(A), t.norm() ** 2, rtol=1e-12)
    # deterministic
    assert np.array_equal(t.hoevd_gram(), A)


@pytest.mark.parametrize("eig", ["tridiag", "syevd"])
@pytest.mark.parametrize("n,cnt", [(280, 6000), (2300, 30000)])
def test_hoevd_factor_matches_oracle_eigenspace(n, cnt, eig, monkeypatch):
    # n = 2300 is past the LDS-resident reduction: rocSOLVER dsytrd, then the
    # host top-R stage and dormtr; "syevd" is the full rocSOLVER dsyevd path
    monkeypatch.setenv("SP_EIG", eig)
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, n, cnt, 1.1, 1, seed=9)
    R = 6
    say("hoevd factor", n, len(vals), eig)
    t = sp.SpSymtensor(n, 4, inds, vals)
    U, w = t.hoevd_init(R)
    Uo, wo = oracle.hoevd_factor(oracle.hoevd_gram(4, n, inds, vals), R)
    assert np.allclose(w, wo, rtol=1e-12)
    assert np.all(np.diff(w) >= 0)
    assert np.allclose(U.T @ U, np.eye(R), atol=1e-12)
    # column by column up to sign (distinct eigenvalues here)
    assert np.min(np.diff(wo)) > 1e-6 * wo[-1]
    assert np.allclose(np.abs(np.sum(U * Uo, axis=0)), 1.0, atol=1e-9)
    # it became the current factor
    assert np.array_equal(t.factor(), U)


def test_hoevd_init_in_tucker_run():
    # This 300 x 300 HOEVD matrix once hung rocSOLVER's index-range solver
    # dsyevdx (il = 297, iu = 300), which no path calls any more: the top-4
    # pairs come from the library's reduction, host stage and back-transform
    # (the host stage on this matrix is also checked on the CPU,
    # tests/test_tridiag_host.py::test_hoevd_gram_300_through_the_host_stage).
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 300, 6000, 1.1, 1, seed=8)
    R = 4
    say("tucker with hoevd init, n = 300")
    t = sp.SpSymtensor(300, 4, inds, vals)
    U0, w = t.hoevd_init(R)
    Uo, wo = oracle.hoevd_factor(oracle.hoevd_gram(4, 300, inds, vals), R)
    assert np.allclose(w, wo, rtol=1e-12)
    assert np.allclose(np.abs(np.sum(U0 * Uo, axis=0)), 1.0, atol=1e-9)
    a = t.tucker(alg=1, rank=R, max_iter=3, rel_tol=0.0, init="hoevd")
    b = t.tucker(alg=1, rank=R, max_iter=3, rel_tol=0.0, init="given", U0=U0)
    assert np.array_equal(a["errors"], b["errors"])


def test_hoevd_init_drives_tucker_run():
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 260, 6000, 1.1, 1, seed=10)
    R = 4
    say("tucker with hoevd init (dsyevd)")
    t = sp.SpSymtensor(260, 4, inds, vals)
    U0, w = t.hoevd_init(R)
    Uo, wo = oracle.hoevd_factor(oracle.hoevd_gram(4, 260, inds, vals), R)
    assert np.allclose(w, wo, rtol=1e-12)
    a = t.tucker(alg=1, rank=R, max_iter=3, rel_tol=0.0, init="hoevd")
    b = t.tucker(alg=1, rank=R, max_iter=3, rel_tol=0.0, init="given", U0=U0)
    assert np.array_equal(a["errors"], b["errors"])
