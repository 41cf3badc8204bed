"""Writes the golden fixtures of tests/golden from the reference's own code.

Run in a checkout where oracle/_ref/libsymprop_ref.so was built from the
reference sources (make -C oracle ref).  The outputs are data only: integer
tree arrays (or their SHA-256 when large) and FP64 S3TTMc results of the
reference's CompressedDTree / kronecker_product_opt5 at one OpenMP thread.

  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import ref  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def small_cases():
    """Small tensors with repeated indices inside tuples (seeded)."""
    rng = np.random.default_rng(20240601)
    cases = []
    for (N, n, cnt) in [(3, 5, 12), (4, 6, 25), (5, 5, 30), (5, 7, 40), (6, 4, 20)]:
        t = np.sort(rng.integers(1, n + 1, size=(cnt, N)), axis=1)
        t = np.unique(t, axis=0)
        rng.shuffle(t)          # construction must follow input order
        v = rng.random(len(t))
        cases.append((N, n, t.astype(np.uint32), v))
    return cases


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref` first")
    out = {}
    # contact-5: hashes of every tree array + Y at R=4 in full
    o, d, inds, vals = oracle.load_tns(os.path.join(HERE, "contact-5.tns"))
    rt = ref.RefTree(o, d, inds, vals)
    c5 = {"order": o, "dim": d, "nnz": int(len(vals)), "levels": []}
    for l in range(o - 1):
        ent = {"size": int(len(rt.level_idx[l])),
               "level_idx": sha(rt.level_idx[l].astype(np.uint32)),
               "level_ptr": sha(rt.level_ptr[l].astype(np.uint64))}
        if l >= 1:
            ent["dep_list"] = sha(rt.dep_list(l).astype(np.uint64))
        c5["levels"].append(ent)
    c5["owner_rem"] = sha(rt.own_rem.astype(np.uint32))
    c5["num_owners"] = int(len(rt.own_rem))
    c5["norm"] = oracle.norm(o, inds, vals)
    out["contact5"] = c5
    U = oracle.uniform_factor(1, d, 4)
    Y = rt.s3ttmc(U, 4, nthreads=1)
    np.save(os.path.join(HERE, "contact5_y_r4.npy"), Y)
    U8 = oracle.uniform_factor(1, d, 8)
    Y8 = rt.s3ttmc(U8, 8, nthreads=1)
    out["contact5_r8"] = {"sum": float(Y8.sum()), "sumsq": float((Y8 ** 2).sum()),
                          "rows": [float(x) for x in Y8[[0, 100, 244]].sum(axis=1)]}

    # small tensors with repeated indices: full arrays
    arrays = {}
    for ci, (N, n, t, v) in enumerate(small_cases()):
        rt = ref.RefTree(N, n, t, v)
        arrays["c%d_inds" % ci] = t
        arrays["c%d_vals" % ci] = v
        for l in range(N - 1):
            arrays["c%d_l%d_idx" % (ci, l)] = rt.level_idx[l].astype(np.uint32)
            arrays["c%d_l%d_ptr" % (ci, l)] = rt.level_ptr[l].astype(np.uint64)
            if l >= 1:
                arrays["c%d_l%d_dep" % (ci, l)] = rt.dep_list(l).astype(np.uint64)
        arrays["c%d_rem" % ci] = rt.own_rem.astype(np.uint32)
        arrays["c%d_val" % ci] = rt.own_val
        R = 3
        U = np.arange(1, n * R + 1, dtype=np.float64).reshape(n, R) / (n * R)
        arrays["c%d_U" % ci] = U
        arrays["c%d_Y" % ci] = rt.s3ttmc(U, R, nthreads=1)
    np.savez(os.path.join(HERE, "small_trees.npz"), **arrays)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote golden.json, contact5_y_r4.npy, small_trees.npz")


if __name__ == "__main__":
    main()
