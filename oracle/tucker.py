"""Oracle Tucker loops (numpy LAPACK) -- TEST INFRASTRUCTURE ONLY.

Restates TuckerHOOI / TuckerHOQRI / TuckerAltHOOI of
src/sparse/spsymtensor.cpp:201-337 with regularize (168-199), the error
formula of src/tucker.cpp:40-45 and symtens_norm / symtens_rel_res of
src/dense/symtensor.cpp:78-106.  LAPACK calls map as in
src/dense/matrixops.cpp: leftsvd -> numpy.linalg.svd (left vectors, 'O'),
qr -> numpy.linalg.qr (thin Q), syevd -> numpy.linalg.eigh (ascending,
last R columns).  S3TTMc comes from the C oracle (liboracle.so).
Sizes: n x R^(N-1) dense unpacked Y for HOOI, so small tensors only.
"""
import numpy as np

import oracle


def fit_error(norm, core_norm):
    """sqrt(||X||^2 - ||C||^2) / ||X|| (tucker.cpp:40-45, spsymtensor.cpp:230-238)
    with the negative rounding residue of an exact fit clamped to 0 (the
    reference takes its sqrt, NaN)."""
    return np.sqrt(max(norm ** 2 - core_norm ** 2, 0.0)) / norm


def regularize(U, delta):
    U = U.copy()
    nr = np.sqrt((U * U).sum(axis=1))
    big = delta < nr
    U[big] *= (delta / nr[big])[:, None]
    return np.linalg.svd(U, full_matrices=False)[0]


_UNPACK_MAPS = {}


def unpack_rows(Y, R, k):
    """Unpacked rows: column (j1..jk) of the full R^k row takes the packed
    entry of the sorted multi-index (symtensor.hpp:96-116)."""
    key = (R, k)
    if key not in _UNPACK_MAPS:
        P = oracle.symtens_size(R, k)
        _UNPACK_MAPS[key] = oracle.unpack(np.arange(P, dtype=np.float64), R, k).astype(np.int64)
    return Y[:, _UNPACK_MAPS[key]]


def rel_residual(x, y):
    return ((x - y) ** 2).sum() / (x ** 2).sum()


def hooi(T, R, max_iter, rel_tol=1e-6, delta=0.5, seed=0, U0=None, nthreads=4):
    N, n = T.order, T.dim
    norm = oracle.norm(N, T.inds, T.vals)
    U = oracle.random_factor(seed, n, R) if U0 is None else np.array(U0, float)
    core_prev = np.zeros((R, R ** (N - 1)))
    errors = []
    prev = 0.0
    for _ in range(max_iter):
        U = regularize(U, delta)
        Y = unpack_rows(T.s3ttmc(U, R, nthreads=nthreads), R, N - 1)
        U = np.linalg.svd(Y, full_matrices=False)[0][:, :R]
        core = U.T @ Y
        err = fit_error(norm, np.sqrt((core ** 2).sum()))
        errors.append(err)
        if abs(prev - err) < rel_tol:
            break
        prev = err
        core_prev = core
    return U, np.array(errors), core


def hoqri(T, R, max_iter, rel_tol=1e-6, delta=0.5, seed=0, U0=None, nthreads=4):
    N, n = T.order, T.dim
    norm = oracle.norm(N, T.inds, T.vals)
    W = oracle.symtens_coeff(R, N - 1)
    U = oracle.random_factor(seed, n, R) if U0 is None else np.array(U0, float)
    errors = []
    prev = 0.0
    core = None
    for _ in range(max_iter):
        U = regularize(U, delta)
        Y = T.s3ttmc(U, R, nthreads=nthreads)
        core = U.T @ Y
        U = np.linalg.qr((Y * W) @ core.T)[0]
        cn = np.sqrt((core ** 2 * W).sum())
        err = fit_error(norm, cn)
        errors.append(err)
        if abs(prev - err) < rel_tol:
            break
        prev = err
    return U, np.array(errors), core


def althooi(T, R, max_iter, rel_tol=1e-6, delta=0.5, seed=0, U0=None, nthreads=4):
    N, n = T.order, T.dim
    norm = oracle.norm(N, T.inds, T.vals)
    W = oracle.symtens_coeff(R, N - 1)
    U = oracle.random_factor(seed, n, R) if U0 is None else np.array(U0, float)
    errors = []
    prev = 0.0
    core = None
    for _ in range(max_iter):
        U = regularize(U, delta)
        Y = T.s3ttmc(U, R, nthreads=nthreads)
        A = (Y * W) @ Y.T
        U = np.linalg.eigh(A)[1][:, n - R:]
        core = U.T @ Y
        cn = np.sqrt((core ** 2 * W).sum())
        err = fit_error(norm, cn)
        errors.append(err)
        if abs(prev - err) < rel_tol:
            break
        prev = err
    return U, np.array(errors), core


def subspace_cos(U1, U2):
    """Smallest cosine of the principal angles between span(U1), span(U2)."""
    Q1 = np.linalg.qr(U1)[0]
    Q2 = np.linalg.qr(U2)[0]
    return float(np.linalg.svd(Q1.T @ Q2, compute_uv=False).min())
