"""Oracle for the learned warm start -- TEST INFRASTRUCTURE ONLY.

The learned warm start has no counterpart in the reference (it plugs into the
initialisation hook hosvd_init, src/sparse/spsymtensor.cpp:339-383).  This
module restates, in FP64 numpy, exactly what the library's HIP kernels
compute, so that tests can check them:

  node features   64 per node: log(1+degree), log(1+distinct partners),
                  log(1+max pair multiplicity), log(1+repeated partners)
                  (multiplicity of {i, j}: its occurrences over all position
                  pairs of all nonzeros), and a 60-dim sketch: fixed hash-seeded Gaussian node
                  vectors propagated `rounds` times through the owner lists
                  (s'[r] = sum over owners (leaf L, rem r, value v) of
                  v * sum_{a in L} s[a]) with row L2 normalisation after
                  each round.
  MLP             64 -> 256 -> 256 -> R, ReLU, row-major W[in][out].
  loss            mean over rows of the squared row error.
  Adam            Kingma & Ba (2015), bias-corrected.

The backward pass is pinned by central finite differences in tests.
"""
import numpy as np

import oracle

SKETCH = 60


def splitmix(x):
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = x
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def node_gauss(n, dim, seed):
    """g[j, k]: standard normal from splitmix64 of (seed, node j, k), Box-Muller."""
    with np.errstate(over="ignore"):
        j = np.arange(n, dtype=np.uint64)[:, None]
        k = np.arange(dim, dtype=np.uint64)[None, :]
        base = (np.uint64(seed) * np.uint64(0x100000001B3) + j * np.uint64(2 * dim) +
                np.uint64(2) * k) & np.uint64(0xFFFFFFFFFFFFFFFF)
        a = splitmix(base)
        b = splitmix(base + np.uint64(1))
    u1 = ((a >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)
    u2 = (b >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def owner_lists(order, dim, inds, vals):
    """(rem, leaf tuple 0-based, value) for every owner (dtree.cpp:62-80)."""
    N = order
    inds = np.asarray(inds, dtype=np.int64).reshape(-1, N) - 1
    rems, leaves, vs = [], [], []
    for j in range(N - 1, -1, -1):
        rems.append(inds[:, j])
        leaves.append(np.delete(inds, j, axis=1))
        vs.append(np.asarray(vals, dtype=np.float64))
    return np.concatenate(rems), np.concatenate(leaves), np.concatenate(vs)


def node_features(order, dim, inds, vals, rounds=3, seed=12345):
    rem, leaf, v = owner_lists(order, dim, inds, vals)
    deg = np.bincount(rem, minlength=dim).astype(np.float64)
    t = np.asarray(inds, dtype=np.int64).reshape(-1, order) - 1
    pairs = set()
    for a in range(order):
        for b in range(a + 1, order):
            p = t[:, [a, b]]
            p = p[p[:, 0] != p[:, 1]]
            pairs.update(map(tuple, p))
    partners = np.zeros(dim)
    for a, b in pairs:
        partners[a] += 1
        partners[b] += 1
    mult = {}
    for a in range(order):
        for b in range(a + 1, order):
            for x, y in t[:, [a, b]]:
                if x != y:
                    mult[(x, y)] = mult.get((x, y), 0) + 1
    maxmult = np.zeros(dim)
    rep = np.zeros(dim)
    for (x, y), m in mult.items():
        maxmult[x] = max(maxmult[x], m)
        maxmult[y] = max(maxmult[y], m)
        if m >= 2:
            rep[x] += 1
            rep[y] += 1
    s = node_gauss(dim, SKETCH, seed)
    for _ in range(rounds):
        agg = s[leaf].sum(axis=1) * v[:, None]
        new = np.zeros_like(s)
        np.add.at(new, rem, agg)
        nrm = np.sqrt((new ** 2).sum(axis=1, keepdims=True))
        s = np.where(nrm > 0, new / np.where(nrm > 0, nrm, 1.0), 0.0)
    return np.concatenate([np.log1p(deg)[:, None], np.log1p(partners)[:, None],
                           np.log1p(maxmult)[:, None], np.log1p(rep)[:, None], s], axis=1)


def param_shapes(din=64, dh=256, dout=8):
    return [(din, dh), (dh,), (dh, dh), (dh,), (dh, dout), (dout,)]


def unflatten(p, din=64, dh=256, dout=8):
    out, off = [], 0
    for s in param_shapes(din, dh, dout):
        k = int(np.prod(s))
        out.append(p[off:off + k].reshape(s))
        off += k
    return out


def forward(p, X, din=64, dh=256, dout=8):
    W1, b1, W2, b2, W3, b3 = unflatten(p, din, dh, dout)
    Z1 = X @ W1 + b1
    H1 = np.maximum(Z1, 0)
    Z2 = H1 @ W2 + b2
    H2 = np.maximum(Z2, 0)
    return H2 @ W3 + b3, (Z1, H1, Z2, H2)


def loss_grad(p, X, T, din=64, dh=256, dout=8):
    W1, b1, W2, b2, W3, b3 = unflatten(p, din, dh, dout)
    O, (Z1, H1, Z2, H2) = forward(p, X, din, dh, dout)
    N = X.shape[0]
    D = O - T
    L = (D ** 2).sum() / N
    dO = 2.0 * D / N
    gW3 = H2.T @ dO
    gb3 = dO.sum(axis=0)
    dZ2 = (dO @ W3.T) * (Z2 > 0)
    gW2 = H1.T @ dZ2
    gb2 = dZ2.sum(axis=0)
    dZ1 = (dZ2 @ W2.T) * (Z1 > 0)
    gW1 = X.T @ dZ1
    gb1 = dZ1.sum(axis=0)
    g = np.concatenate([a.ravel() for a in (gW1, gb1, gW2, gb2, gW3, gb3)])
    return L, g


def adam_step(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def procrustes_align(U, Uref):
    """U Q with Q = argmin ||U Q - Uref||_F over orthogonal Q."""
    A, _, Bt = np.linalg.svd(U.T @ Uref)
    return U @ (A @ Bt)
