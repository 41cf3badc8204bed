"""Learned warm start for symmetric Tucker (no counterpart in the reference).

A node-feature MLP predicts an initial factor U that replaces the random /
HOEVD start at the initialisation hook (hosvd_init,
src/sparse/spsymtensor.cpp:339-383).  Features, forward/backward passes and
Adam run on the GPU in the library's HIP kernels (sp_node_features,
sp_mlp_*); this module is the host-side glue:

  features(t)            64 features per node (sp_node_features)
  WarmStartModel         feature standardisation + MLP 64 -> 256 -> 256 -> 1
                         scoring nodes by their row norm in the converged
                         factor (rotation invariant); the start is the
                         indicator columns of the R best-scored nodes
  feature_start(...)     the non-learned start from the same features
  aligned_targets(...)   converged factors rotated (orthogonal Procrustes)
                         onto one reference (for comparing factors)
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
    """Feature standardisation + MLP 64 -> 256 -> 256 -> 1 that scores nodes.

    Target: the node's row norm in the factor the plain method converged to,
    scaled so that the mean square over nodes is 1 (sqrt(n/R) * ||U*_i||).
    The row norm is invariant to the rotation and column order of U*, so
    factors of different tensors train one model without any alignment.
    Start: the indicator columns of the R highest-scoring nodes (orthonormal),
    the shape of the fixed points HOOI reaches on sparse planted tensors,
    where the converged factor sits on a handful of nodes."""

    def __init__(self, rank, hidden=256, seed=0):
        self.rank = rank
        self.mlp = MLP(NFEAT, hidden, 1, seed)
        self.mu = None
        self.sd = None

    def fit(self, feats, factors, steps=600, batch=4096, lr=1e-3, seed=0):
        """feats: list of (n x 64); factors: list of converged orthonormal U."""
        X = np.concatenate(feats).astype(np.float64)
        self.mu = X.mean(axis=0)
        self.sd = X.std(axis=0) + 1e-6
        Xs = ((X - self.mu) / self.sd).astype(np.float32)
        T = np.concatenate([np.sqrt(U.shape[0] / U.shape[1]) *
                            np.sqrt((U * U).sum(axis=1, keepdims=True)) for U in factors])
        return self.mlp.train(Xs, T.astype(np.float32), steps, batch, lr, seed)

    def scores(self, feats):
        Xs = ((feats - self.mu) / self.sd).astype(np.float32)
        return self.mlp.predict(Xs)[:, 0].astype(np.float64)

    def predict(self, feats, inds=None):
        """Start for a tensor: with its index tuples, the hyperedge
        neighbourhood of the best-scored node (pair_neighbourhood_start),
        else the indicator columns of the top-scored nodes."""
        sc = self.scores(feats)
        if inds is not None:
            return pair_neighbourhood_start(sc, inds, self.rank)
        return top_nodes_start(sc, self.rank)


def pair_neighbourhood_start(score, inds, rank):
    """Indicator columns on the nodes around the best-scored node i: its
    partner j sharing the most nonzeros with it, the nodes of the nonzeros
    holding both (most shared first), then the best-scored remaining nodes.
    A start on whole nonzeros keeps the first S3TTMc from vanishing (an
    owner contributes only when all its leaf nodes carry weight)."""
    score = np.asarray(score, dtype=np.float64)
    n = len(score)
    T = np.asarray(inds, dtype=np.int64) - 1
    rank_of = np.empty(n, dtype=np.int64)
    rank_of[np.lexsort((np.arange(n), -score))] = np.arange(n)
    i = int(np.argmin(rank_of))
    E = T[(T == i).any(axis=1)]
    chosen = [i]
    if len(E):
        cnt = np.bincount(E.ravel(), minlength=n).astype(np.float64)
        cnt[i] = -1.0
        j = int(np.lexsort((rank_of, -cnt))[0])
        if cnt[j] > 0:
            B = E[(E == j).any(axis=1)]
            cb = np.bincount(B.ravel(), minlength=n).astype(np.float64)
            cb[[i, j]] = -1.0
            nodes = np.nonzero(cb > 0)[0]
            nodes = nodes[np.lexsort((rank_of[nodes], -cb[nodes]))]
            chosen += [j] + [int(x) for x in nodes]
    seen = set(chosen)
    for x in np.argsort(rank_of):
        if len(chosen) >= rank:
            break
        if int(x) not in seen:
            chosen.append(int(x))
            seen.add(int(x))
    U = np.zeros((n, rank))
    U[chosen[:rank], np.arange(rank)] = 1.0
    return U


def top_nodes_start(score, rank):
    """Indicator columns of the `rank` highest-scoring nodes (ties: lower
    node index first)."""
    order = np.lexsort((np.arange(len(score)), -np.asarray(score)))
    U = np.zeros((len(score), rank))
    U[order[:rank], np.arange(rank)] = 1.0
    return U


def feature_score(feats):
    """Non-learned node score from the same features: largest pair
    multiplicity, then most repeated partners, then degree."""
    F = np.asarray(feats, dtype=np.float64)
    return F[:, 2] * 1e6 + F[:, 3] * 1e3 + F[:, 0]


def feature_start(feats, rank, inds=None):
    """Non-learned start: the same construction as the learned start, on
    feature_score instead of the network's score."""
    if inds is not None:
        return pair_neighbourhood_start(feature_score(feats), inds, rank)
    return top_nodes_start(feature_score(feats), rank)


def spectral_start(feats, rank):
    """Non-learned start from the same features: the top-`rank` principal
    directions (left singular vectors) of the centred hypergraph sketch."""
    X = np.asarray(feats, dtype=np.float64)[:, 4:]
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
