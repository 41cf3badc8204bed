"""CPU-side checks of the product library: exports, loader, generators.

No compute call here touches a GPU.
"""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle
import symprop_amd as sp

from conftest import GOLDEN, ROOT


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(sp_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(sp.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 25
    for n in sorted(names):
        assert hasattr(L, n), n
    # the Python binding covers the same set
    assert names <= set(sp.SIGNATURES)


def test_version_and_error_string():
    assert b"gfx950" in sp.lib().sp_version()
    assert sp.lib().sp_last_error() is not None


def test_loader_matches_oracle_loader(tmp_path):
    path = os.path.join(GOLDEN, "contact-5.tns")
    a = sp.load_tns(path)
    b = oracle.load_tns(path)
    assert a[0] == b[0] and a[1] == b[1]
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    p = tmp_path / "odd.tns"
    p.write_text("4 10\n4 3 2 1 0.25\n10 1 1 5 -2\n7 7 7\n")
    a = sp.load_tns(str(p))
    b = oracle.load_tns(str(p))
    assert a[2].tolist() == b[2].tolist() == [[1, 2, 3, 4], [1, 1, 5, 10]]
    assert a[3].tolist() == b[3].tolist() == [0.25, -2.0]


def _write_tns(path, order, dim, tuples, vals, sep=" ", eol="\n", trailing=True):
    lines = ["%d %d" % (order, dim)]
    for t, v in zip(tuples, vals):
        lines.append(sep.join(str(int(x)) for x in t) + sep + repr(float(v)))
    text = eol.join(lines) + (eol if trailing else "")
    path.write_text(text)


@pytest.mark.parametrize("sep,eol,trailing", [(" ", "\n", True), ("\t", "\r\n", True),
                                              ("  ", "\n", False)])
def test_parallel_loader_matches_stream_reader(tmp_path, sep, eol, trailing):
    # several MiB, so the text is split over several parser threads; tuples
    # unsorted in the file, values with exponents and signs
    rng = np.random.default_rng(5)
    n = 120_000
    t = rng.integers(1, 5000, size=(n, 5))
    v = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, size=n)
    p = tmp_path / "big.tns"
    _write_tns(p, 5, 5000, t, v, sep, eol, trailing)
    a = sp.load_tns(str(p))
    b = oracle.load_tns(str(p))
    assert a[0] == b[0] == 5 and a[1] == b[1] == 5000
    # without a trailing newline the stream drops the last tuple
    assert len(a[3]) == n - (0 if trailing else 1)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])


def test_loader_falls_back_to_stream_semantics(tmp_path):
    # tokens the fast path does not take: the serial stream reader decides
    for body in ("1 2 3 0.5\n3 2 1 1.5abc\n4 4 4 2\n", "1 2 3 +0.5\n2 2 2 1\n",
                 "1 2 3 nan\n", "1 2 99999999999 1\n", "1 2 3 1\n1 2 x 1\n"):
        p = tmp_path / "odd.tns"
        p.write_text("3 9\n" + body)
        a = sp.load_tns(str(p))
        b = oracle.load_tns(str(p))
        assert a[2].tolist() == b[2].tolist(), body
        assert a[3].tolist() == b[3].tolist(), body


def test_loader_binary_cache(tmp_path):
    rng = np.random.default_rng(6)
    t = rng.integers(1, 300, size=(5000, 4))
    v = rng.random(5000)
    p = tmp_path / "c.tns"
    c = tmp_path / "c.spbin"
    _write_tns(p, 4, 300, t, v)
    a = sp.load_tns(str(p))
    b = sp.load_tns(str(p), cache=str(c))       # parses and writes the cache
    assert c.exists()
    c_mtime = os.stat(c).st_mtime_ns
    d = sp.load_tns(str(p), cache=str(c))       # reads the cache
    assert os.stat(c).st_mtime_ns == c_mtime
    for x in (b, d):
        assert x[0] == a[0] and x[1] == a[1]
        assert np.array_equal(x[2], a[2]) and np.array_equal(x[3], a[3])
    # a changed source invalidates the cache
    _write_tns(p, 4, 300, t[:100], v[:100])
    os.utime(p, ns=(os.stat(p).st_atime_ns, os.stat(p).st_mtime_ns + 10**9))
    e = sp.load_tns(str(p), cache=str(c))
    assert len(e[3]) == 100
    assert np.array_equal(e[2], oracle.load_tns(str(p))[2])


def test_loader_missing_file():
    with pytest.raises(sp.SpError) as e:
        sp.load_tns("/nonexistent/x.tns")
    assert e.value.code == sp.SP_ERR_IO


@pytest.mark.parametrize("kind,order,dim,nnz,param,k", [
    (sp.GEN_UNIFORM, 5, 327, 8000, 0.0, 1),
    (sp.GEN_ZIPF, 4, 1000, 20000, 1.1, 1),
    (sp.GEN_PLANTED, 6, 4000, 20000, 0.8, 8),
])
def test_generators(kind, order, dim, nnz, param, k):
    inds, vals = sp.generate(kind, order, dim, nnz, param, k, seed=0)
    inds2, vals2 = sp.generate(kind, order, dim, nnz, param, k, seed=0)
    assert np.array_equal(inds, inds2) and np.array_equal(vals, vals2)
    i3, _ = sp.generate(kind, order, dim, nnz, param, k, seed=1)
    assert not np.array_equal(inds[: len(i3)], i3[: len(inds)])
    assert inds.min() >= 1 and inds.max() <= dim
    assert np.all(np.diff(inds.astype(np.int64), axis=1) >= 0)
    # unique and lexicographically sorted
    assert len(np.unique(inds, axis=0)) == len(inds)
    assert np.all(np.lexsort(inds.T[::-1]) == np.arange(len(inds)))
    if kind == sp.GEN_UNIFORM:
        assert 0.98 * nnz <= len(inds) <= nnz
        assert np.all(vals == 1.0)
    else:
        assert len(inds) == nnz
    if kind == sp.GEN_ZIPF:
        assert np.all((vals >= 0) & (vals < 1))
        counts = np.bincount(inds.ravel(), minlength=dim + 1)
        assert counts[1] > counts[10] > counts[500]
    if kind == sp.GEN_PLANTED:
        block = dim // k
        comm = (inds - 1) // block
        mode = np.array([np.bincount(r, minlength=k).max() for r in comm])
        assert mode.mean() > 0.7 * order


def test_generator_bad_args():
    with pytest.raises(sp.SpError):
        sp.generate(7, 3, 10, 10)


def test_ctx_create_rejects_bad_arguments_before_touching_a_gpu():
    # per-tuple checks (range, order) run on the device right after the
    # upload (tests/test_gpu_parity.py); argument checks come first
    import ctypes as C
    L = sp.lib()
    h = C.c_void_p()
    v = np.ones(1)
    t = np.array([[1, 2]], np.uint32)
    rc = L.sp_ctx_create(2, 5, 1, t.ctypes.data_as(sp.u32p), v.ctypes.data_as(sp.f64p),
                         C.byref(h))
    assert rc == sp.SP_ERR_UNSUPPORTED
    rc = L.sp_ctx_create(3, 5, 1, None, None, C.byref(h))
    assert rc == sp.SP_ERR_ARG


@pytest.mark.parametrize("exe", ["TTMC_PATH", "TTMCTC_PATH"])
def test_micro_drivers_usage_and_missing_file(tmp_path, exe):
    # examples/symtens_ttmc.cpp:12-15 and :27-29 (s3ttmctc.cpp the same): the
    # usage line and the open error, both with exit status 1, before any
    # device work
    import subprocess
    path = getattr(sp, exe)
    if not os.path.exists(path):
        pytest.skip("drivers not built")
    p = subprocess.run([path], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert p.stderr == "Usage: %s <input_file> <R>\n" % path
    p = subprocess.run([path, str(tmp_path / "missing.tns"), "4"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 1
    assert p.stderr == "Error opening file: %s\n" % (tmp_path / "missing.tns")
