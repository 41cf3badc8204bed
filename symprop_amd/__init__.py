"""MI355X-native S3TTMc hot path for symmetric sparse Tucker (SymProp).

Python access to the C-ABI of libsymprop_amd.so (include/symprop_amd.h).
The library is the product: construction, S3TTMc and the dense Y-side
operations run in its HIP kernels on the GPU.  There is no CPU fallback:
if the library is missing or no GPU is present, calls fail with an error.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsymprop_amd.so")
CLI_PATH = os.path.join(_HERE, "symtucker")
# the S3TTMc micro-drivers (examples/symtens_ttmc.cpp, examples/s3ttmctc.cpp)
TTMC_PATH = os.path.join(_HERE, "symtens_ttmc")
TTMCTC_PATH = os.path.join(_HERE, "s3ttmctc")

SP_OK = 0
SP_ERR_ARG = 1
SP_ERR_UNSUPPORTED = 2
SP_ERR_DEVICE = 3
SP_ERR_OOM = 4
SP_ERR_STATE = 5
SP_ERR_NUMERIC = 6
SP_ERR_IO = 7

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16 = 2

GEN_UNIFORM = 0
GEN_ZIPF = 1
GEN_PLANTED = 2

u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
f64p = ctypes.POINTER(ctypes.c_double)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p

# name -> (restype, argtypes) of every exported C-ABI function
SIGNATURES = {
    "sp_last_error": (ctypes.c_char_p, []),
    "sp_version": (ctypes.c_char_p, []),
    "sp_ctx_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                     u32p, f64p, ctypes.POINTER(vp)]),
    "sp_ctx_destroy": (ctypes.c_int, [vp]),
    "sp_norm": (ctypes.c_int, [vp, f64p]),
    "sp_cdtree_build": (ctypes.c_int, [vp]),
    "sp_cdtree_depth": (ctypes.c_int, [vp, u32p]),
    "sp_cdtree_level_size": (ctypes.c_int, [vp, ctypes.c_uint32, u64p]),
    "sp_cdtree_num_owners": (ctypes.c_int, [vp, u64p]),
    "sp_cdtree_export": (ctypes.c_int, [vp, ctypes.c_uint32, u32p, u32p, u64p, u64p]),
    "sp_cdtree_export_owners": (ctypes.c_int, [vp, u32p, f64p]),
    "sp_factor_upload": (ctypes.c_int, [vp, f64p, ctypes.c_uint32]),
    "sp_factor_download": (ctypes.c_int, [vp, f64p]),
    "sp_s3ttmc_dev": (ctypes.c_int, [vp, ctypes.c_int]),
    "sp_s3ttmc_noalloc": (ctypes.c_int, [vp, vp, ctypes.c_uint32, vp, ctypes.c_int]),
    "sp_s3ttmc": (ctypes.c_int, [vp, f64p, ctypes.c_uint32, f64p, ctypes.c_uint64,
                                 ctypes.c_int]),
    "sp_y_download": (ctypes.c_int, [vp, f64p, ctypes.c_uint64]),
    "sp_packed_width": (ctypes.c_int, [vp, u64p]),
    "sp_gemmtn": (ctypes.c_int, [vp, f64p]),
    "sp_symgemm": (ctypes.c_int, [vp]),
    "sp_gram": (ctypes.c_int, [vp, f64p]),
    "sp_leftsvd_y": (ctypes.c_int, [vp, f64p]),
    "sp_qr": (ctypes.c_int, [vp]),
    "sp_regularize": (ctypes.c_int, [vp, ctypes.c_double]),
    "sp_core_norm": (ctypes.c_int, [vp, f64p]),
    "sp_ctx_info": (ctypes.c_int, [vp, u32p, u32p, u64p]),
    "sp_hoevd_gram": (ctypes.c_int, [vp, f64p]),
    "sp_hoevd_init": (ctypes.c_int, [vp, ctypes.c_uint32, f64p, f64p]),
    "sp_syevd_top": (ctypes.c_int, [vp, f64p, ctypes.c_uint64, ctypes.c_uint32, f64p, f64p]),
    "sp_node_features": (ctypes.c_int, [vp, ctypes.c_uint32, ctypes.c_uint64, f32p]),
    "sp_mlp_create": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.c_uint64, ctypes.POINTER(vp)]),
    "sp_mlp_destroy": (ctypes.c_int, [vp]),
    "sp_mlp_num_params": (ctypes.c_int, [vp, u64p]),
    "sp_mlp_get_params": (ctypes.c_int, [vp, f32p]),
    "sp_mlp_set_params": (ctypes.c_int, [vp, f32p]),
    "sp_mlp_loss_grad": (ctypes.c_int, [vp, f32p, f32p, ctypes.c_uint64, f64p, f32p]),
    "sp_mlp_train": (ctypes.c_int, [vp, f32p, f32p, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_uint64, ctypes.c_float, ctypes.c_uint64, f32p,
                                    f64p]),
    "sp_mlp_predict": (ctypes.c_int, [vp, f32p, ctypes.c_uint64, f32p]),
    "sp_tucker_opts_default": (None, [ctypes.c_void_p]),
    "sp_tucker_run": (ctypes.c_int, [vp, vp, f64p, f64p, f64p, f64p, f64p, u64p, u64p,
                                     f64p]),
    "sp_factor_snapshot": (ctypes.c_int, [vp]),
    "sp_subspace_sin": (ctypes.c_int, [vp, f64p]),
    "sp_timing_enable": (ctypes.c_int, [vp, ctypes.c_int]),
    "sp_timing_get": (ctypes.c_int, [vp, ctypes.c_char_p, f64p, u64p]),
    "sp_timing_reset": (ctypes.c_int, [vp]),
    "sp_sync": (ctypes.c_int, [vp]),
    "sp_load_tns": (ctypes.c_int, [ctypes.c_char_p, u32p, u32p, u64p, u32p, f64p]),
    "sp_load_tns_cached": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, u32p, u32p, u64p,
                                          u32p, f64p]),
    "sp_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint64, ctypes.c_double, ctypes.c_uint32,
                                   ctypes.c_uint64, u32p, f64p, u64p]),
}

_LIB = None


class TuckerOpts(ctypes.Structure):
    """sp_tucker_opts (include/symprop_amd.h)."""
    _fields_ = [("alg", ctypes.c_int), ("rank", ctypes.c_uint32),
                ("max_iter", ctypes.c_uint64), ("rel_tol", ctypes.c_double),
                ("delta", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("init", ctypes.c_char_p), ("precision", ctypes.c_int),
                ("switch_tol", ctypes.c_double), ("subspace_tol", ctypes.c_double),
                ("verbose", ctypes.c_int)]


class SpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("symprop_amd error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads libsymprop_amd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsymprop_amd.so is not built: run "
                               "`python -m symprop_amd.build` (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != SP_OK:
        raise SpError(rc, lib().sp_last_error().decode())


def _p(a, t):
    return a.ctypes.data_as(t)


def load_tns(path, cache=None):
    """(order, dim, inds n x order, vals) of a .tns file; with `cache`, through
    the binary tuple cache (sp_load_tns_cached)."""
    o, d, n = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()

    def call(i, v):
        if cache is None:
            return lib().sp_load_tns(path.encode(), ctypes.byref(o), ctypes.byref(d),
                                     ctypes.byref(n), i, v)
        return lib().sp_load_tns_cached(path.encode(), cache.encode(), ctypes.byref(o),
                                        ctypes.byref(d), ctypes.byref(n), i, v)
    _check(call(None, None))
    inds = np.zeros(n.value * o.value, dtype=np.uint32)
    vals = np.zeros(n.value)
    _check(call(_p(inds, u32p), _p(vals, f64p)))
    return o.value, d.value, inds.reshape(n.value, o.value), vals


def generate(kind, order, dim, nnz, param=0.0, communities=1, seed=0):
    inds = np.zeros(nnz * order, dtype=np.uint32)
    vals = np.zeros(nnz)
    out = ctypes.c_uint64()
    _check(lib().sp_generate(kind, order, dim, nnz, param, communities, seed,
                             _p(inds, u32p), _p(vals, f64p), ctypes.byref(out)))
    n = out.value
    return inds[: n * order].reshape(n, order).copy(), vals[:n].copy()


class SpSymtensor:
    """Device-resident symmetric sparse tensor with its compressed tree.

    Mirrors SpSymtensor (include/detail/sparse/spsymtensor.hpp:26-65):
    construction builds the tree, S3TTMc returns the packed (or unpacked) Y.
    """

    def __init__(self, dim, order, inds, vals, build=True):
        self.order, self.dim = int(order), int(dim)
        self.inds = np.ascontiguousarray(inds, dtype=np.uint32).reshape(-1, self.order)
        self.vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.nnz = len(self.vals)
        h = vp()
        _check(lib().sp_ctx_create(self.order, self.dim, self.nnz,
                                   _p(self.inds, u32p), _p(self.vals, f64p),
                                   ctypes.byref(h)))
        self.h = h
        self.R = None
        if build:
            self.build()

    def close(self):
        if getattr(self, "h", None):
            lib().sp_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def build(self):
        _check(lib().sp_cdtree_build(self.h))

    def norm(self):
        v = ctypes.c_double()
        _check(lib().sp_norm(self.h, ctypes.byref(v)))
        return v.value

    # tree export -----------------------------------------------------
    def depth(self):
        d = ctypes.c_uint32()
        _check(lib().sp_cdtree_depth(self.h, ctypes.byref(d)))
        return d.value

    def level_size(self, l):
        s = ctypes.c_uint64()
        _check(lib().sp_cdtree_level_size(self.h, l, ctypes.byref(s)))
        return s.value

    def level(self, l):
        L = self.level_size(l)
        keys = np.zeros(L * (l + 1), dtype=np.uint32)
        idx = np.zeros(L, dtype=np.uint32)
        ptr = np.zeros(L + 1, dtype=np.uint64)
        dep = np.zeros(max(1, L * (l + 1)), dtype=np.uint64)
        _check(lib().sp_cdtree_export(self.h, l, _p(keys, u32p), _p(idx, u32p),
                                      _p(ptr, u64p), _p(dep, u64p) if l >= 1 else None))
        return dict(keys=keys.reshape(L, l + 1), level_idx=idx, level_ptr=ptr,
                    dep_list=dep[: L * (l + 1)].reshape(L, l + 1) if l >= 1 else None)

    def owners(self):
        n = ctypes.c_uint64()
        _check(lib().sp_cdtree_num_owners(self.h, ctypes.byref(n)))
        rem = np.zeros(n.value, dtype=np.uint32)
        val = np.zeros(n.value)
        _check(lib().sp_cdtree_export_owners(self.h, _p(rem, u32p), _p(val, f64p)))
        return rem, val

    # factor / S3TTMc ---------------------------------------------------
    def set_factor(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape[0] == self.dim
        self.R = U.shape[1]
        _check(lib().sp_factor_upload(self.h, _p(U, f64p), self.R))

    def factor(self):
        U = np.zeros((self.dim, self.R))
        _check(lib().sp_factor_download(self.h, _p(U, f64p)))
        return U

    def packed_width(self):
        P = ctypes.c_uint64()
        _check(lib().sp_packed_width(self.h, ctypes.byref(P)))
        return P.value

    def s3ttmc_dev(self, precision=PREC_FP64):
        _check(lib().sp_s3ttmc_dev(self.h, precision))

    def y(self):
        P = self.packed_width()
        Y = np.zeros((self.dim, P))
        _check(lib().sp_y_download(self.h, _p(Y, f64p), P))
        return Y

    def s3ttmc(self, U, precision=PREC_FP64):
        """Packed Y = S3TTMc(U) (n x P)."""
        self.set_factor(U)
        self.s3ttmc_dev(precision)
        return self.y()

    def gemmtn(self):
        core = np.zeros((self.R, self.packed_width()))
        _check(lib().sp_gemmtn(self.h, _p(core, f64p)))
        return core

    def symgemm(self):
        _check(lib().sp_symgemm(self.h))

    def gram(self):
        P = self.packed_width()
        G = np.zeros((P, P))
        _check(lib().sp_gram(self.h, _p(G, f64p)))
        return G

    def leftsvd_y(self):
        s = np.zeros(self.R)
        _check(lib().sp_leftsvd_y(self.h, _p(s, f64p)))
        return s

    def qr(self):
        _check(lib().sp_qr(self.h))

    def regularize(self, delta):
        _check(lib().sp_regularize(self.h, delta))

    def core_norm(self):
        v = ctypes.c_double()
        _check(lib().sp_core_norm(self.h, ctypes.byref(v)))
        return v.value

    def hoevd_gram(self):
        """A = X_(1) X_(1)^T (dim x dim)."""
        A = np.zeros((self.dim, self.dim))
        _check(lib().sp_hoevd_gram(self.h, _p(A, f64p)))
        return A

    def hoevd_init(self, rank):
        """HOEVD factor (top-`rank` eigenvectors of X_(1) X_(1)^T, ascending);
        it also becomes the current factor.  Returns (U, eigenvalues)."""
        U = np.zeros((self.dim, rank))
        w = np.zeros(rank)
        _check(lib().sp_hoevd_init(self.h, rank, _p(U, f64p), _p(w, f64p)))
        self.R = rank
        return U, w

    def syevd_top(self, A, rank):
        """Eigenvectors of the `rank` largest eigenvalues of the symmetric
        matrix A (matrix::syevd): (U, w), columns and values ascending."""
        A = np.ascontiguousarray(A, dtype=np.float64)
        n = A.shape[0]
        assert A.shape == (n, n)
        U = np.zeros((n, rank))
        w = np.zeros(rank)
        _check(lib().sp_syevd_top(self.h, _p(A, f64p), n, rank, _p(U, f64p), _p(w, f64p)))
        return U, w

    # decomposition -------------------------------------------------------
    def tucker(self, alg=0, rank=2, max_iter=10, rel_tol=1e-6, delta=0.5, seed=0,
               init="rnd", precision=PREC_FP64, switch_tol=-1.0, U0=None, verbose=False,
               subspace_tol=0.0):
        """One TuckerHOOI (alg 0) / TuckerHOQRI (1) / TuckerAltHOOI (2) run.

        Returns dict(U, errors, core_norms, times, iters, fp64_sweeps, seconds);
        errors, core_norms and times (seconds since the run started) are per
        iteration.
        """
        o = TuckerOpts()
        lib().sp_tucker_opts_default(ctypes.byref(o))
        init_b = init.encode()
        o.alg, o.rank, o.max_iter, o.rel_tol, o.delta, o.seed = (
            alg, rank, max_iter, rel_tol, delta, seed)
        o.init, o.precision, o.switch_tol, o.subspace_tol = (
            init_b, precision, switch_tol, subspace_tol)
        o.verbose = 1 if verbose else 0
        U = np.zeros((self.dim, rank))
        m = max(1, max_iter)
        errs, cns, ts = np.zeros(m), np.zeros(m), np.zeros(m)
        it, sw, sec = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_double()
        u0 = None
        if U0 is not None:
            u0 = np.ascontiguousarray(U0, dtype=np.float64)
            assert u0.shape == (self.dim, rank)
        _check(lib().sp_tucker_run(self.h, ctypes.byref(o),
                                   _p(u0, f64p) if u0 is not None else None,
                                   _p(U, f64p), _p(errs, f64p), _p(cns, f64p), _p(ts, f64p),
                                   ctypes.byref(it), ctypes.byref(sw), ctypes.byref(sec)))
        self.R = rank
        k = it.value
        return dict(U=U, errors=errs[:k].copy(), core_norms=cns[:k].copy(),
                    times=ts[:k].copy(), iters=k, fp64_sweeps=sw.value, seconds=sec.value)

    def factor_snapshot(self):
        _check(lib().sp_factor_snapshot(self.h))

    def subspace_sin(self):
        v = ctypes.c_double()
        _check(lib().sp_subspace_sin(self.h, ctypes.byref(v)))
        return v.value

    # timing ------------------------------------------------------------
    def timing(self, on=True):
        _check(lib().sp_timing_enable(self.h, 1 if on else 0))

    def timing_get(self, name):
        ms, cnt = ctypes.c_double(), ctypes.c_uint64()
        _check(lib().sp_timing_get(self.h, name.encode(), ctypes.byref(ms),
                                   ctypes.byref(cnt)))
        return ms.value, cnt.value

    def timing_reset(self):
        _check(lib().sp_timing_reset(self.h))

    def sync(self):
        _check(lib().sp_sync(self.h))
