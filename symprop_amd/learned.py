"""Learned warm start for symmetric Tucker (no counterpart in the reference).

A node-feature MLP predicts an initial factor U that replaces the random /
HOEVD start at the initialisation hook (hosvd_init,
src/sparse/spsymtensor.cpp:339-383).  Features, forward/backward passes and
Adam run on the GPU in the library's HIP kernels (sp_node_features,
sp_mlp_*); this module is the host-side glue:

  features(t)            64 features per node (sp_node_features)
  WarmStartModel         feature standardisation + MLP 64 -> 256 -> 256 -> R
  aligned_targets(...)   converged factors rotated (orthogonal Procrustes)
                         onto one reference so that the regression target is
                         well defined (U is only defined up to rotation)
  hooi_trace(...)        TuckerHOOI's iteration (regularize, S3TTMc, leftsvd,
                         gemmtn) from a given start, keeping every factor, for
                         convergence measured by subspace angles
"""
import ctypes

import numpy as np

from . import SpSymtensor, _check, _p, f32p, f64p, lib, u64p, vp

NFEAT = 64


def features(t: SpSymtensor, rounds=3, seed=12345):
    out = np.zeros((t.dim, NFEAT), dtype=np.float32)
    _check(lib().sp_node_features(t.h, rounds, seed, _p(out, f32p)))
    return out


class MLP:
    """FP32 MLP din -> dh -> dh -> dout on the GPU (sp_mlp_*)."""

    def __init__(self, din=64, dh=256, dout=8, seed=0):
        self.din, self.dh, self.dout = din, dh, dout
        h = vp()
        _check(lib().sp_mlp_create(din, dh, dout, seed, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().sp_mlp_destroy(self.h)
            self.h = None

    def num_params(self):
        n = ctypes.c_uint64()
        _check(lib().sp_mlp_num_params(self.h, ctypes.byref(n)))
        return n.value

    def params(self):
        p = np.zeros(self.num_params(), dtype=np.float32)
        _check(lib().sp_mlp_get_params(self.h, _p(p, f32p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params()
        _check(lib().sp_mlp_set_params(self.h, _p(p, f32p)))

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.num_params(), dtype=np.float32)
        L = ctypes.c_double()
        _check(lib().sp_mlp_loss_grad(self.h, _p(X, f32p), _p(T, f32p), X.shape[0],
                                      ctypes.byref(L), _p(g, f32p)))
        return L.value, g

    def train(self, X, T, steps, batch=4096, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        losses = np.zeros(steps, dtype=np.float32)
        L = ctypes.c_double()
        _check(lib().sp_mlp_train(self.h, _p(X, f32p), _p(T, f32p), X.shape[0], steps,
                                  batch, lr, seed, _p(losses, f32p), ctypes.byref(L)))
        return losses

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        out = np.zeros((X.shape[0], self.dout), dtype=np.float32)
        _check(lib().sp_mlp_predict(self.h, _p(X, f32p), X.shape[0], _p(out, f32p)))
        return out


def procrustes(U, Uref):
    A, _, Bt = np.linalg.svd(U.T @ Uref)
    return U @ (A @ Bt)


def aligned_targets(factors, ref=None):
    """Rotate each orthonormal factor onto `ref` (default: the first one)."""
    ref = factors[0] if ref is None else ref
    return [procrustes(U, ref) for U in factors], ref


class WarmStartModel:
    """Feature standardisation + MLP; targets are scaled by sqrt(n) (O(1) rows)."""

    def __init__(self, rank, hidden=256, seed=0):
        self.rank = rank
        self.mlp = MLP(NFEAT, hidden, rank, seed)
        self.mu = None
        self.sd = None
        self.ref = None

    def fit(self, feats, factors, steps=600, batch=4096, lr=1e-3, seed=0):
        """feats: list of (n x 64); factors: list of converged orthonormal U."""
        targets, self.ref = aligned_targets(factors)
        X = np.concatenate(feats).astype(np.float64)
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-6
        Xs = ((X - self.mu) / self.sd).astype(np.float32)
        T = np.concatenate([np.sqrt(U.shape[0]) * U for U in targets]).astype(np.float32)
        return self.mlp.train(Xs, T, steps, batch, lr, seed)

    def predict(self, feats):
        Xs = ((feats - self.mu) / self.sd).astype(np.float32)
        return self.mlp.predict(Xs).astype(np.float64) / np.sqrt(feats.shape[0])


def spectral_start(feats, rank):
    """Non-learned start from the same features: the top-`rank` principal
    directions (left singular vectors) of the centred hypergraph sketch."""
    X = np.asarray(feats, dtype=np.float64)[:, 2:]
    X = X - X.mean(axis=0)
    return np.linalg.svd(X, full_matrices=False)[0][:, :rank]


def hooi_trace(t: SpSymtensor, U0, iters, delta=0.5, alg="hooi", stop_tol=None):
    """TuckerHOOI's iteration (src/sparse/spsymtensor.cpp:221-241) or
    TuckerHOQRI's (src/sparse/spsymtensor.cpp:267-281) from U0, returning the
    factor after every iteration and the reported errors.  With `stop_tol`,
    stops after the first iteration whose subspace moved by less than it
    (sin of the largest principal angle to the previous factor)."""
    if alg not in ("hooi", "hoqri"):
        raise ValueError("alg must be 'hooi' or 'hoqri'")
    t.set_factor(U0)
    norm = t.norm()
    Us, errs = [], []
    for _ in range(iters):
        t.regularize(delta)
        t.s3ttmc_dev()
        if alg == "hooi":
            t.leftsvd_y()
            t.gemmtn()
        else:
            t.gemmtn()
            t.symgemm()
            t.qr()
        cn = t.core_norm()
        errs.append(np.sqrt(max(norm * norm - cn * cn, 0.0)) / norm)
        Us.append(t.factor())
        if stop_tol is not None and len(Us) > 1 and sin_angle(Us[-1], Us[-2]) < stop_tol:
            break
    return Us, np.array(errs)


def converged_at(Us, tol=1e-6):
    """First iteration (1-based) whose factor moved by less than `tol` from
    the previous one; None if none did."""
    for k in range(1, len(Us)):
        if sin_angle(Us[k], Us[k - 1]) < tol:
            return k + 1
    return None


def sin_angle(U1, U2):
    """sin of the largest principal angle between two orthonormal bases."""
    c = np.linalg.svd(U1.T @ U2, compute_uv=False)
    return float(np.sqrt(max(0.0, 1.0 - min(1.0, c.min()) ** 2)))


def iterations_to(Us, target, tol=1e-6):
    """First iteration (1-based) whose factor is within `tol` (sin angle) of
    `target`; len(Us)+1 if never."""
    for k, U in enumerate(Us):
        if sin_angle(U, target) < tol:
            return k + 1
    return len(Us) + 1
