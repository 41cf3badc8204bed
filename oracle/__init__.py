"""CPU oracle for the S3TTMc hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker.  The product (symprop_amd/) never
imports, links or runs anything under oracle/.

liboracle.so (oracle.cpp) restates the reference's construction and S3TTMc
(src/compressed_dtree.cpp:13-131, 421-518; src/dtree.cpp:38-80), its
combinatorics (include/detail/dense/symtensor_base.hpp:12-76), packed layout
(include/detail/dense/symtensor.hpp:20-142), norm
(src/sparse/spsymtensor.cpp:114-138) and loader (src/utils/loader.cpp:8-36).
oracle/tucker.py restates the HOOI/HOQRI/AltHOOI loops
(src/sparse/spsymtensor.cpp:168-337) over numpy's LAPACK.
oracle/ref.py binds the reference's own CompressedDTree compiled into
oracle/_ref/ (only where the reference checkout was present at build time).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all", "ref"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_binom.restype = ctypes.c_uint64
        L.orc_binom.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_symtens_size.restype = ctypes.c_uint64
        L.orc_symtens_size.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.orc_multinomial_sorted.restype = ctypes.c_uint64
        L.orc_multinomial_sorted.argtypes = [u32p, ctypes.c_int]
        L.orc_symtens_coeff.argtypes = [ctypes.c_uint32, ctypes.c_int, f64p]
        L.orc_packed_index.restype = ctypes.c_uint64
        L.orc_packed_index.argtypes = [u32p, ctypes.c_int, ctypes.c_uint32]
        L.orc_unpack.argtypes = [f64p, ctypes.c_uint32, ctypes.c_int, f64p]
        L.orc_onesided.argtypes = [f64p, f64p, f64p, ctypes.c_uint32, ctypes.c_int]
        L.orc_norm.restype = ctypes.c_double
        L.orc_norm.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, f64p]
        L.orc_tree_build.restype = ctypes.c_void_p
        L.orc_tree_build.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint64, u32p, f64p]
        L.orc_tree_free.argtypes = [ctypes.c_void_p]
        L.orc_tree_depth.restype = ctypes.c_int
        L.orc_tree_depth.argtypes = [ctypes.c_void_p]
        L.orc_tree_level_size.restype = ctypes.c_uint64
        L.orc_tree_level_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_tree_num_owners.restype = ctypes.c_uint64
        L.orc_tree_num_owners.argtypes = [ctypes.c_void_p]
        L.orc_tree_export.argtypes = [ctypes.c_void_p, ctypes.c_int, u32p, u32p,
                                      u64p, u64p]
        L.orc_tree_owners.argtypes = [ctypes.c_void_p, u32p, f64p]
        L.orc_s3ttmc.restype = ctypes.c_int
        L.orc_s3ttmc.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint32, f64p,
                                 ctypes.c_uint64, ctypes.c_int]
        L.orc_s3ttmc_dense.argtypes = [ctypes.c_uint32, ctypes.c_uint32,
                                       ctypes.c_uint64, u32p, f64p, f64p,
                                       ctypes.c_uint32, f64p]
        L.orc_hoevd_gram.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                     u32p, f64p, f64p]
        L.orc_random_factor.argtypes = [ctypes.c_uint64, ctypes.c_uint64, f64p]
        L.orc_uniform_factor.argtypes = [ctypes.c_uint64, ctypes.c_uint64, f64p]
        L.orc_load_tns.restype = ctypes.c_longlong
        L.orc_load_tns.argtypes = [ctypes.c_char_p, u32p, u32p, u32p, f64p]
        _LIB = L
    return _LIB


def _p(a, t):
    return a.ctypes.data_as(t)


def binom(n, k):
    return int(lib().orc_binom(n, k))


def symtens_size(R, k):
    return int(lib().orc_symtens_size(R, k))


def multinomial_sorted(t):
    t = np.ascontiguousarray(t, dtype=np.uint32)
    return int(lib().orc_multinomial_sorted(_p(t, u32p), len(t)))


def symtens_coeff(R, k):
    out = np.zeros(symtens_size(R, k))
    lib().orc_symtens_coeff(R, k, _p(out, f64p))
    return out


def packed_index(t, R):
    t = np.ascontiguousarray(t, dtype=np.uint32)
    return int(lib().orc_packed_index(_p(t, u32p), len(t), R))


def unpack(packed, R, k):
    packed = np.ascontiguousarray(packed, dtype=np.float64)
    out = np.zeros(R ** k)
    lib().orc_unpack(_p(packed, f64p), R, k, _p(out, f64p))
    return out


def onesided(dst, src, row, src_order):
    dst = np.ascontiguousarray(dst, dtype=np.float64).copy()
    src = np.ascontiguousarray(src, dtype=np.float64)
    row = np.ascontiguousarray(row, dtype=np.float64)
    lib().orc_onesided(_p(dst, f64p), _p(src, f64p), _p(row, f64p), len(row),
                       src_order)
    return dst


def norm(order, inds, vals):
    inds = np.ascontiguousarray(inds, dtype=np.uint32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    return float(lib().orc_norm(order, len(vals), _p(inds, u32p), _p(vals, f64p)))


class Tree:
    """Oracle compressed dimension tree (sort-based restatement)."""

    def __init__(self, order, dim, inds, vals):
        self.order, self.dim = int(order), int(dim)
        self.inds = np.ascontiguousarray(inds, dtype=np.uint32).reshape(-1, order)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.h = lib().orc_tree_build(order, dim, len(self.vals),
                                      _p(self.inds, u32p), _p(self.vals, f64p))
        self.depth = lib().orc_tree_depth(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    def level_size(self, l):
        return int(lib().orc_tree_level_size(self.h, l))

    def level(self, l):
        L = self.level_size(l)
        keys = np.zeros(L * (l + 1), dtype=np.uint32)
        idx = np.zeros(L, dtype=np.uint32)
        nptr = L + 1
        ptr = np.zeros(nptr, dtype=np.uint64)
        dep = np.zeros(L * (l + 1), dtype=np.uint64)
        lib().orc_tree_export(self.h, l, _p(keys, u32p), _p(idx, u32p),
                              _p(ptr, u64p), _p(dep, u64p) if l >= 1 else None)
        return dict(keys=keys.reshape(L, l + 1), level_idx=idx, level_ptr=ptr,
                    dep_list=dep.reshape(L, l + 1) if l >= 1 else None)

    def owners(self):
        no = int(lib().orc_tree_num_owners(self.h))
        rem = np.zeros(no, dtype=np.uint32)
        val = np.zeros(no)
        lib().orc_tree_owners(self.h, _p(rem, u32p), _p(val, f64p))
        return rem, val

    def s3ttmc(self, U, R, nthreads=1, ldy=None):
        """Packed Y (dim x P) by the opt5 schedule."""
        P = symtens_size(R, self.order - 1)
        ldy = P if ldy is None else ldy
        U = np.ascontiguousarray(U, dtype=np.float64)
        Y = np.zeros((self.dim, ldy))
        rc = lib().orc_s3ttmc(self.h, _p(U, f64p), R, _p(Y, f64p), ldy, nthreads)
        if rc != 0:
            raise ValueError("order < 3 has no S3TTMc levels")
        return Y[:, :P].copy()


def s3ttmc_dense(order, dim, inds, vals, U, R):
    inds = np.ascontiguousarray(inds, dtype=np.uint32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    U = np.ascontiguousarray(U, dtype=np.float64)
    P = symtens_size(R, order - 1)
    Y = np.zeros((dim, P))
    lib().orc_s3ttmc_dense(order, dim, len(vals), _p(inds, u32p), _p(vals, f64p),
                           _p(U, f64p), R, _p(Y, f64p))
    return Y


def hoevd_gram(order, dim, inds, vals):
    """X_(1) X_(1)^T as the reference's HOEVD builds it (dim x dim)."""
    inds = np.ascontiguousarray(inds, dtype=np.uint32)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    A = np.zeros((dim, dim))
    lib().orc_hoevd_gram(order, dim, len(vals), _p(inds, u32p), _p(vals, f64p), _p(A, f64p))
    return A


def hoevd_factor(A, R):
    """Eigenvectors of the R largest eigenvalues of A, ascending
    (src/dense/matrixops.cpp:77-100), and those eigenvalues."""
    w, V = np.linalg.eigh(A)
    return V[:, -R:], w[-R:]


def random_factor(seed, n, R):
    out = np.zeros(n * R)
    lib().orc_random_factor(seed, n * R, _p(out, f64p))
    return out.reshape(n, R)


def uniform_factor(seed, n, R):
    out = np.zeros(n * R)
    lib().orc_uniform_factor(seed, n * R, _p(out, f64p))
    return out.reshape(n, R)


def load_tns(path):
    o = ctypes.c_uint32()
    d = ctypes.c_uint32()
    cnt = lib().orc_load_tns(path.encode(), ctypes.byref(o), ctypes.byref(d),
                             None, None)
    if cnt < 0:
        raise OSError("cannot open " + path)
    inds = np.zeros(cnt * o.value, dtype=np.uint32)
    vals = np.zeros(cnt)
    lib().orc_load_tns(path.encode(), ctypes.byref(o), ctypes.byref(d),
                       _p(inds, u32p), _p(vals, f64p))
    return o.value, d.value, inds.reshape(cnt, o.value), vals
