"""The host C++ drop-in (symprop_amd/host/symtucker.hpp) called as the
reference's callers call it, through tests/dropin_harness.cpp:
SpSymtensor::S3TTMc with the packed result and with unpacked rows
(partial_sym_res = false, the branch TuckerHOOI takes at
src/sparse/spsymtensor.cpp:223; unpacking as spsymtensor.cpp:85-107), and
TuckerHOOI from a given factor, against the oracle."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import oracle
from oracle import tucker
import symprop_amd as sp

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "symprop_amd")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    if not os.path.exists(os.path.join(PKG, "libsymprop_amd.so")):
        pytest.fail("libsymprop_amd.so is not built")
    exe = str(tmp_path_factory.mktemp("dropin") / "dropin_harness")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + os.path.join(PKG, "host"),
                    "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "dropin_harness.cpp"), "-L" + PKG,
                    "-lsymprop_amd", "-Wl,-rpath," + PKG, "-o", exe], check=True)
    return exe


def run(exe, tmp_path, N, n, inds, vals, U, R, iters):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as f:
        np.array([n, N, R, iters], dtype=np.uint32).tofile(f)
        np.array([len(vals)], dtype=np.uint64).tofile(f)
        np.ascontiguousarray(inds, dtype=np.uint32).tofile(f)
        np.ascontiguousarray(vals, dtype=np.float64).tofile(f)
        np.ascontiguousarray(U, dtype=np.float64).tofile(f)
    p = subprocess.run([exe, str(src), str(dst)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Finished creating compressed dtree" in p.stdout
    out = np.fromfile(dst, dtype=np.float64)
    P, F = oracle.symtens_size(R, N - 1), R ** (N - 1)
    sizes = [n * P, n * F, iters, n * R]
    assert len(out) == sum(sizes)
    parts = np.split(out, np.cumsum(sizes)[:-1])
    return (parts[0].reshape(n, P), parts[1].reshape(n, F), parts[2], parts[3].reshape(n, R))


@pytest.mark.parametrize("N,n,cnt,R", [(3, 50, 600, 4), (4, 40, 900, 3), (5, 30, 800, 3),
                                       (6, 25, 700, 2)])
def test_dropin_s3ttmc_packed_unpacked_and_hooi(harness, tmp_path, N, n, cnt, R):
    rng = np.random.default_rng(N * 100 + n)
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(cnt, N)), axis=1), axis=0)
    v = rng.standard_normal(len(tt))
    U = rng.standard_normal((n, R))
    say("drop-in S3TTMc / HOOI", N, n, len(tt), R)
    packed, full, errors, factor = run(harness, tmp_path, N, n, tt, v, U, R, 3)
    T = oracle.Tree(N, n, tt, v)
    Yo = T.s3ttmc(U, R)
    scale = np.abs(Yo).max()
    assert np.allclose(packed, Yo, rtol=1e-12, atol=1e-12 * scale)
    # unpacked rows: the packed row expanded over all R^(N-1) orderings
    for i in range(n):
        assert np.array_equal(full[i], oracle.unpack(packed[i], R, N - 1)), i
    # ... and the dense definition (every ordering of every nonzero)
    Yd = oracle.s3ttmc_dense(N, n, tt, v, U, R)
    assert np.allclose(full, np.stack([oracle.unpack(Yd[i], R, N - 1) for i in range(n)]),
                       rtol=1e-12, atol=1e-12 * scale)
    # HOOI from the given factor
    Uo, eo, _ = tucker.hooi(T, R, 3, rel_tol=0.0, delta=0.5, U0=U)
    assert np.allclose(errors, eo, rtol=0, atol=1e-10)
    assert tucker.subspace_cos(factor, Uo) > 1 - 1e-9
