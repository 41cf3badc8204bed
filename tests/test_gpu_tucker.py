"""GPU Tucker loops (HOOI / HOQRI / AltHOOI) against the oracle's loops.

Both start from the same libstdc++ mt19937 + normal_distribution factor, so
the reported error trajectories must agree (they depend on subspaces only,
not on the sign/rotation choices of the eigen/QR/SVD routines).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from oracle import tucker
import symprop_amd as sp

from conftest import GOLDEN
from test_oracle import exact_rank_tensor

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


ALGS = [(0, tucker.hooi), (1, tucker.hoqri), (2, tucker.althooi)]


@pytest.mark.parametrize("alg,ofn", ALGS)
def test_trajectory_matches_oracle(alg, ofn):
    rng = np.random.default_rng(7)
    N, n, R = 5, 40, 3
    tt = np.unique(np.sort(rng.integers(1, n + 1, size=(700, N)), axis=1), axis=0)
    v = rng.random(len(tt))
    say("tucker alg", alg, N, n, len(tt), R)
    t = sp.SpSymtensor(n, N, tt, v)
    r = t.tucker(alg=alg, rank=R, max_iter=8, rel_tol=1e-14, delta=0.5, seed=0)
    Uo, eo, _ = ofn(oracle.Tree(N, n, tt, v), R, 8, rel_tol=1e-14, delta=0.5, seed=0)
    assert r["iters"] == len(eo)
    assert np.allclose(r["errors"], eo, rtol=0, atol=1e-10)
    assert tucker.subspace_cos(r["U"], Uo) > 1 - 1e-9
    assert np.allclose(r["U"].T @ r["U"], np.eye(R), atol=1e-10)


def test_hoqri_order7_zipf_matches_oracle():
    # the order-7 rank-6 shape of the HOQRI benchmark, at a size the oracle
    # runs in seconds
    N, n, R = 7, 400, 6
    inds, vals = sp.generate(sp.GEN_ZIPF, N, n, 20000, 1.2, 1, seed=3)
    say("hoqri order 7 zipf", n, len(vals), R)
    t = sp.SpSymtensor(n, N, inds, vals)
    r = t.tucker(alg=1, rank=R, max_iter=6, rel_tol=1e-14, delta=0.5, seed=0)
    Uo, eo, _ = tucker.hoqri(oracle.Tree(N, n, inds, vals), R, 6, rel_tol=1e-14, delta=0.5,
                             seed=0)
    assert r["iters"] == len(eo)
    assert np.allclose(r["errors"], eo, rtol=0, atol=1e-10)
    assert tucker.subspace_cos(r["U"], Uo) > 1 - 1e-9
    # the mixed forms reach the FP64 fit under the reference stopping rule
    e64 = t.tucker(alg=1, rank=R, max_iter=200, rel_tol=1e-6, delta=0.5, seed=0)["errors"][-1]
    for prec in (sp.PREC_FP32, sp.PREC_BF16):
        e = t.tucker(alg=1, rank=R, max_iter=200, rel_tol=1e-6, delta=0.5, seed=0,
                     precision=prec)["errors"][-1]
        assert abs(e - e64) / e64 <= 1e-6


@pytest.mark.parametrize("alg,ofn", ALGS)
def test_uniform_order5_reference_configuration(alg, ofn):
    # the reference's CPU-runnable case: order 5, n = 327, 8,000 uniform
    # tuples (deduplicated after sorting), values 1.0, rank 8, 20 iterations,
    # delta 0.1, random start seed 0
    inds, vals = sp.generate(sp.GEN_UNIFORM, 5, 327, 8000, 0.0, 1, seed=0)
    assert np.all(vals == 1.0)
    say("uniform order 5 config, alg", alg, len(vals))
    t = sp.SpSymtensor(327, 5, inds, vals)
    T = oracle.Tree(5, 327, inds, vals)
    # the reference stopping rule (it stops after two iterations here)
    r = t.tucker(alg=alg, rank=8, max_iter=20, rel_tol=1e-6, delta=0.1, seed=0)
    Uo, eo, _ = ofn(T, 8, 20, rel_tol=1e-6, delta=0.1, seed=0)
    assert r["iters"] == len(eo)
    assert np.allclose(r["errors"], eo, rtol=0, atol=1e-10)
    assert tucker.subspace_cos(r["U"], Uo) > 1 - 1e-8
    # all 20 iterations: the fit error of this structureless tensor is
    # 1 - 3e-8, so rounding differences between the GPU and numpy LAPACK
    # eigen-solvers grow from iteration 12 on (to ~3e-10 by iteration 19);
    # the first ten iterations agree to 1e-12 and every error to 1e-6
    r = t.tucker(alg=alg, rank=8, max_iter=20, rel_tol=0.0, delta=0.1, seed=0)
    _, eo, _ = ofn(T, 8, 20, rel_tol=0.0, delta=0.1, seed=0)
    d = np.abs(r["errors"] - eo)
    print("[case] uniform order 5, 20 iterations: max |error diff| per iteration", d)
    assert len(d) == 20
    assert d[:10].max() <= 1e-12
    assert d.max() <= 1e-6 * eo.max()


def test_exact_rank_recovered():
    tt, v, X, U0 = exact_rank_tensor(n=9, R=2, N=4, seed=1)
    say("tucker exact rank")
    t = sp.SpSymtensor(9, 4, tt, v)
    r = t.tucker(alg=0, rank=2, max_iter=10, rel_tol=1e-12, delta=100.0, seed=0)
    assert r["errors"][-1] < 1e-6
    assert tucker.subspace_cos(r["U"], U0) > 1 - 1e-8


def test_contact5_readme_run(contact5):
    # README.md:100: symtucker -f contact-5.tns -r 8 -i 20 -a 0 -d 0.1
    o, d, inds, vals = contact5
    say("tucker contact-5 HOOI r8 i20 d0.1")
    t = sp.SpSymtensor(d, o, inds, vals)
    r = t.tucker(alg=0, rank=8, max_iter=20, rel_tol=1e-6, delta=0.1, seed=0)
    Uo, eo, _ = tucker.hooi(oracle.Tree(o, d, inds, vals), 8, 20, rel_tol=1e-6,
                            delta=0.1, seed=0)
    assert r["iters"] == len(eo)
    assert np.allclose(r["errors"], eo, rtol=0, atol=1e-9)
    assert tucker.subspace_cos(r["U"], Uo) > 1 - 1e-7


@pytest.mark.parametrize("prec", [sp.PREC_FP32, sp.PREC_BF16])
def test_mixed_precision_reaches_fp64_fit(prec, contact5):
    o, d, inds, vals = contact5
    t = sp.SpSymtensor(d, o, inds, vals)
    say("tucker mixed", prec)
    # Config's default rel_tol (tucker.hpp:17); the contact-5 trajectory is slow and
    # not monotone, so much tighter tolerances stop at noise-dependent iterations
    r64 = t.tucker(alg=0, rank=8, max_iter=60, rel_tol=1e-6, delta=0.1, seed=0)
    rmx = t.tucker(alg=0, rank=8, max_iter=60, rel_tol=1e-6, delta=0.1, seed=0,
                   precision=prec)
    assert rmx["fp64_sweeps"] >= 1
    e64, emx = r64["errors"][-1], rmx["errors"][-1]
    assert abs(emx - e64) / e64 <= 1e-6


def test_cli_runs_and_writes_reference_format(tmp_path):
    import subprocess
    out = tmp_path / "out.txt"
    say("cli")
    p = subprocess.run([sp.CLI_PATH, "-f", os.path.join(GOLDEN, "contact-5.tns"), "-r", "4",
                        "-i", "3", "-a", "0", "-d", "0.1", "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Finished creating compressed dtree" in p.stdout
    assert "Insertion time:" in p.stdout and "Running HOOI algorithm" in p.stdout
    assert "Iteration: 0 relative residual:" in p.stdout
    assert "Accumulated function times:" in p.stdout
    lines = out.read_text().splitlines()
    assert lines[0] == "Mode-1 matricized C:"
    assert lines[5] == "Factor U:"
    assert len(lines) == 1 + 4 + 1 + 245
    assert all(l.endswith(",") for l in lines[1:5] + lines[6:])
    assert len(lines[1].split(",")) == 4 ** 4 + 1


def _parse_cli_output(path, R, n):
    lines = path.read_text().splitlines()
    assert lines[0] == "Mode-1 matricized C:"
    assert lines[R + 1] == "Factor U:"
    assert len(lines) == R + 2 + n
    assert all(l.endswith(",") for l in lines[1:R + 1] + lines[R + 2:])
    core = [np.array([float(x) for x in l.split(",")[:-1]]) for l in lines[1:R + 1]]
    U = np.array([[float(x) for x in l.split(",")[:-1]] for l in lines[R + 2:]])
    return core, U


@pytest.mark.parametrize("alg", [0, 2])
def test_cli_output_values_match_oracle(tmp_path, contact5, alg):
    """The written decomposition (tucker.cpp:47-61) value by value against the
    oracle's run of the same algorithm: HOOI (alg 0) writes the unpacked core
    R x R^(N-1), AltHOOI (alg 2) the packed R x P core into the same R^N
    array (quirk A2); both print core row i as the R^(N-1) entries starting at
    flat index i*R (quirk A1), so consecutive rows overlap.  U is defined up
    to the sign of each column (distinct singular values here), so factor
    columns are compared after fixing their signs, and core entries -- which
    carry products of those signs -- by absolute value.  Values are printed
    with 6 significant digits."""
    import subprocess
    o, d, inds, vals = contact5
    R, iters = 4, 3
    out = tmp_path / "out.txt"
    say("cli values, alg", alg)
    p = subprocess.run([sp.CLI_PATH, "-f", os.path.join(GOLDEN, "contact-5.tns"), "-r", str(R),
                        "-i", str(iters), "-a", str(alg), "-d", "0.1", "-o", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    core_rows, U = _parse_cli_output(out, R, d)
    ofn = tucker.hooi if alg == 0 else tucker.althooi
    Uo, eo, core_o = ofn(oracle.Tree(o, d, inds, vals), R, iters, rel_tol=1e-6, delta=0.1,
                         seed=0)
    # per-iteration errors printed with 12 (HOOI) / 10 (AltHOOI) digits
    errs = [float(l.split(" error ")[1]) for l in p.stdout.splitlines()
            if l.startswith("Iteration:")]
    assert len(errs) == len(eo)
    assert np.allclose(errs, eo, rtol=0, atol=1e-9)
    # the flat core array as the reference fills it
    flat = np.zeros(R ** o)
    flat[:core_o.size] = core_o.ravel()
    width = R ** (o - 1)
    scale = np.abs(flat).max()
    for i, row in enumerate(core_rows):
        assert len(row) == width
        ref = flat[i * R:i * R + width]
        assert np.allclose(np.abs(row), np.abs(ref), rtol=2e-5, atol=2e-6 * scale), i
        if i > 0:   # overlapping windows print the same entries
            assert np.array_equal(row[:width - R], core_rows[i - 1][R:])
    # factor columns up to sign
    sgn = np.sign((U * Uo).sum(axis=0))
    assert np.all(sgn != 0)
    assert np.allclose(U * sgn, Uo, rtol=0, atol=2e-6)
