"""Allocation-free S3TTMc (sp_s3ttmc_noalloc) against sp_s3ttmc_dev.

Both run the same kernels on the same work plan, so the results must be
bit-identical; the allocation-free form keeps its scratch (split-row partial
sums, the FP32 factor copy) in the library's static device array.  Two
tensors in a row, the second smaller, then the first again: the shared
scratch keeps the last call's contents, which must not leak into a result.
Device buffers come from the HIP runtime the library itself is linked to.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
import symprop_amd as sp

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def hip_runtime():
    sp.lib()
    path = None
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64.so" in line:
                path = line.split()[-1]
                break
    assert path, "libsymprop_amd.so did not map libamdhip64"
    H = ctypes.CDLL(path)
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipFree.argtypes = [ctypes.c_void_p]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    return H


class DeviceArray:
    def __init__(self, H, shape):
        self.H, self.shape = H, tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.p = ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.p), self.nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        assert self.H.hipMemcpy(self.p, a.ctypes.data, self.nbytes, H2D) == 0

    def fill_nan(self):
        self.upload(np.full(self.shape, np.nan))

    def download(self):
        a = np.empty(self.shape)
        assert self.H.hipMemcpy(a.ctypes.data, self.p, self.nbytes, D2H) == 0
        return a

    def free(self):
        if self.p:
            self.H.hipFree(self.p)
            self.p = None


def noalloc(t, dU, R, dY, prec):
    rc = sp.lib().sp_s3ttmc_noalloc(t.h, dU.p, R, dY.p, prec)
    if rc:
        raise sp.SpError(rc, sp.lib().sp_last_error().decode())
    t.sync()


def check_tensor(H, t, n, R, seed, expect_split):
    U = oracle.random_factor(seed, n, R)
    U2 = oracle.random_factor(seed + 100, n, R)
    dU = DeviceArray(H, (n, R))
    try:
        for prec in (sp.PREC_FP64, sp.PREC_FP32, sp.PREC_BF16):
            t.timing(True)
            t.timing_reset()
            Yref = t.s3ttmc(U, precision=prec)  # prepares the rank, allocating path
            assert (t.timing_get("reduce_split_rows")[1] == 1) == expect_split
            t.timing(False)
            P = Yref.shape[1]
            dY = DeviceArray(H, (n, P))
            try:
                dY.fill_nan()
                dU.upload(U)
                say("noalloc", n, R, "precision", prec)
                noalloc(t, dU, R, dY, prec)
                assert np.array_equal(dY.download(), Yref), prec
                # another factor: only the caller's Y changes
                dU.upload(U2)
                noalloc(t, dU, R, dY, prec)
                Y2 = dY.download()
                assert np.array_equal(t.y(), Yref), "context Y was touched"
                assert np.array_equal(t.s3ttmc(U2, precision=prec), Y2), prec
            finally:
                dY.free()
    finally:
        dU.free()


def test_noalloc_matches_s3ttmc_dev_two_tensors_in_a_row():
    H = hip_runtime()
    # a Zipf tensor whose hub rows are split over several work chunks (the
    # partial sums go through the scratch array), then a smaller order-5 one
    inds, vals = sp.generate(sp.GEN_ZIPF, 4, 2000, 60000, 1.1, 1, seed=3)
    big = sp.SpSymtensor(2000, 4, inds, vals)
    inds5, vals5 = sp.generate(sp.GEN_UNIFORM, 5, 300, 6000, 0.0, 1, seed=4)
    small = sp.SpSymtensor(300, 5, inds5, vals5)
    say("noalloc: order-4 zipf, 60k nnz, R=10")
    check_tensor(H, big, 2000, 10, 1, True)
    say("noalloc: order-5 uniform, 6k nnz, R=8 (smaller)")
    check_tensor(H, small, 300, 8, 2, False)
    say("noalloc: order-4 zipf again after the smaller call")
    check_tensor(H, big, 2000, 10, 3, True)


def test_noalloc_rejects_unprepared_rank_and_null():
    H = hip_runtime()
    inds, vals = sp.generate(sp.GEN_UNIFORM, 4, 100, 2000, 0.0, 1, seed=6)
    t = sp.SpSymtensor(100, 4, inds, vals)
    dU = DeviceArray(H, (100, 6))
    dY = DeviceArray(H, (100, oracle.symtens_coeff(6, 3).size))
    try:
        # no S3TTMc at rank 6 yet: the plan for it does not exist
        rc = sp.lib().sp_s3ttmc_noalloc(t.h, dU.p, 6, dY.p, sp.PREC_FP64)
        assert rc == sp.SP_ERR_STATE
        t.s3ttmc(oracle.random_factor(0, 100, 5))
        rc = sp.lib().sp_s3ttmc_noalloc(t.h, dU.p, 6, dY.p, sp.PREC_FP64)
        assert rc == sp.SP_ERR_STATE
        rc = sp.lib().sp_s3ttmc_noalloc(t.h, None, 5, dY.p, sp.PREC_FP64)
        assert rc == sp.SP_ERR_ARG
        rc = sp.lib().sp_s3ttmc_noalloc(t.h, dU.p, 5, dY.p, 7)
        assert rc == sp.SP_ERR_ARG
    finally:
        dU.free()
        dY.free()
