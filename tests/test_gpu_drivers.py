"""The S3TTMc micro-drivers against the reference's (examples/symtens_ttmc.cpp,
examples/s3ttmctc.cpp): same command line and messages, and symtens_ttmc's
ttmc.out value by value against the golden packed Y of contact-5 at R = 4
(written by the reference's own opt5 from the same mt19937(1) U[1, 2)
factor, tests/golden/make_golden.py) and against the oracle."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle
import symprop_amd as sp

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

TNS = os.path.join(GOLDEN, "contact-5.tns")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_symtens_ttmc_writes_the_reference_output(tmp_path):
    say("symtens_ttmc contact-5 R=4")
    p = subprocess.run([sp.TTMC_PATH, TNS, "4"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 0, p.stderr
    # the constructor's message (spsymtensor.cpp:26), then the driver's
    lines = p.stdout.splitlines()
    assert lines[0] == "Finished creating compressed dtree"
    assert re.fullmatch(r"Insertion time: \d+ ms", lines[1])
    assert re.fullmatch(r"S3TTMc Time: \d+ ms", lines[2])
    assert len(lines) == 3
    text = (tmp_path / "ttmc.out").read_text()
    rows = text.split("\n")
    assert rows[-1] == ""
    rows = rows[:-1]
    G = np.load(os.path.join(GOLDEN, "contact5_y_r4.npy"))
    assert len(rows) == G.shape[0]
    for i, row in enumerate(rows):
        assert row.endswith(", ")
        cells = row[:-2].split(", ")
        assert len(cells) == G.shape[1]
        assert all(re.fullmatch(r"-?\d+\.\d{5}", c) for c in cells), i
        vals = np.array([float(c) for c in cells])
        assert np.allclose(vals, G[i], rtol=1e-12, atol=5.01e-6), i
    # the same Y from the oracle's opt5 restatement
    o, d, inds, vals = oracle.load_tns(TNS)
    Yo = oracle.Tree(o, d, inds, vals).s3ttmc(oracle.uniform_factor(1, d, 4), 4)
    assert np.allclose(Yo, G, rtol=1e-12, atol=1e-9)


def test_s3ttmctc_rounds_and_function_times(tmp_path):
    say("s3ttmctc contact-5 R=4")
    p = subprocess.run([sp.TTMCTC_PATH, TNS, "4"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    assert lines[0] == "Finished creating compressed dtree"
    assert re.fullmatch(r"Insertion time: \d+ ms", lines[1])
    rounds = [l for l in lines if l.startswith("S3TTMcTC Time: ")]
    assert len(rounds) == 10
    assert all(re.fullmatch(r"S3TTMcTC Time: \d+ ms", l) for l in rounds)
    assert "Accumulated function times:" in p.stdout
    for name in ("S3TTMc", "gemmtn", "symgemm"):
        assert re.search(r"\| %s$" % name, p.stdout, re.M), name
