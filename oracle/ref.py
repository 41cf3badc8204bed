"""Bindings to the reference's own CompressedDTree (oracle/_ref) -- TEST ONLY.

oracle/_ref/libsymprop_ref.so is compiled by oracle/Makefile from the
reference sources (src/dtree.cpp, src/compressed_dtree.cpp,
src/dense/symtensor*.cpp, src/utils/function_timer.cpp) plus
oracle/ref_driver.cpp.  It exists only where the reference checkout was
present when the oracle was built; callers must handle available() == False.
"""
import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_HERE, "_ref", "libsymprop_ref.so")
_LIB = None

u32p = ctypes.POINTER(ctypes.c_uint32)
f64p = ctypes.POINTER(ctypes.c_double)


def available():
    return os.path.exists(_PATH)


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(_PATH)
        L.ref_build.restype = ctypes.c_void_p
        L.ref_build.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                u32p, f64p]
        L.ref_build_quiet.restype = ctypes.c_void_p
        L.ref_build_quiet.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                      u32p, f64p]
        L.ref_free.argtypes = [ctypes.c_void_p]
        L.ref_dump.restype = ctypes.c_char_p
        L.ref_dump.argtypes = [ctypes.c_void_p]
        L.ref_exclude_search.restype = ctypes.c_uint64
        L.ref_exclude_search.argtypes = [ctypes.c_void_p, u32p, ctypes.c_uint32,
                                         ctypes.c_uint32]
        L.ref_opt5.argtypes = [ctypes.c_void_p, f64p, ctypes.c_uint32, f64p,
                               ctypes.c_uint64, ctypes.c_int]
        _LIB = L
    return _LIB


class RefTree:
    """Reference CompressedDTree, read back through print() + exclude_search."""

    def __init__(self, order, dim, inds, vals):
        self.order, self.dim = int(order), int(dim)
        inds = np.ascontiguousarray(inds, dtype=np.uint32).reshape(-1, order)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.h = lib().ref_build(order, dim, len(vals),
                                 inds.ctypes.data_as(u32p),
                                 vals.ctypes.data_as(f64p))
        self._parse(lib().ref_dump(self.h).decode())

    def __del__(self):
        if getattr(self, "h", None):
            lib().ref_free(self.h)
            self.h = None

    def _parse(self, text):
        # compressed_dtree.cpp:530-556
        depth = self.order - 1
        lines = text.splitlines()
        self.level_idx, self.level_ptr = [], []
        pos = 1
        for l in range(depth - 1):
            size = int(re.match(r"level \d+ size (\d+)", lines[pos]).group(1))
            body = lines[pos + 1]
            nodes = re.findall(r"node (\d+) ptr (\d+) \|", body)
            tail = int(re.search(r" ptr (\d+)\s*$", body).group(1))
            assert len(nodes) == size
            self.level_idx.append(np.array([int(a) for a, _ in nodes], np.uint32))
            self.level_ptr.append(np.array([int(b) for _, b in nodes] + [tail],
                                           np.uint64))
            pos += 2
        idx, ptr, rem, val = [], [], [], []
        for line in lines[pos:]:
            m = re.match(r"node (\d+) ptr (\d+) \| ?(.*)$", line)
            if not m:
                continue
            idx.append(int(m.group(1)))
            ptr.append(int(m.group(2)))
            toks = m.group(3).split()
            for a, b in zip(toks[0::2], toks[1::2]):
                rem.append(int(a))
                val.append(float(b))
        ptr.append(len(rem))
        self.level_idx.append(np.array(idx, np.uint32))
        self.level_ptr.append(np.array(ptr, np.uint64))
        self.own_rem = np.array(rem, np.uint32)
        self.own_val = np.array(val)

    def keys(self, l):
        """Full tuples of level l, rebuilt by walking level_ptr."""
        prev = self.level_idx[0].reshape(-1, 1)
        for lv in range(1, l + 1):
            ptr = self.level_ptr[lv - 1]
            counts = np.diff(ptr.astype(np.int64))
            parent = np.repeat(np.arange(len(prev)), counts)
            prev = np.concatenate([prev[parent], self.level_idx[lv].reshape(-1, 1)],
                                  axis=1)
        return prev.astype(np.uint32)

    def dep_list(self, l):
        keys = self.keys(l)
        out = np.zeros(keys.shape, dtype=np.uint64)
        for j in range(keys.shape[0]):
            t = np.ascontiguousarray(keys[j])
            for k in range(l + 1):
                out[j, k] = lib().ref_exclude_search(self.h, t.ctypes.data_as(u32p),
                                                     l + 1, k)
        return out

    def s3ttmc(self, U, R, nthreads=1):
        from oracle import symtens_size
        P = symtens_size(R, self.order - 1)
        U = np.ascontiguousarray(U, dtype=np.float64)
        Y = np.zeros((self.dim, P))
        lib().ref_opt5(self.h, U.ctypes.data_as(f64p), R, Y.ctypes.data_as(f64p),
                       P, nthreads)
        return Y
