"""Builds libsymprop_amd.so (HIP kernels + C-ABI + host C++) for gfx950.

Usage: python -m symprop_amd.build   (or __graft_entry__.build()).
Sources are compiled in parallel with hipcc into symprop_amd/_build/ and
linked in-tree as symprop_amd/libsymprop_amd.so; the command-line drivers
(symtucker, and the S3TTMc micro-drivers symtens_ttmc and s3ttmctc) are linked
next to it.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsymprop_amd.so")
CLI = os.path.join(HERE, "symtucker")
# driver executables: (main source under host/, executable in the package)
DRIVERS = [("symtucker_main.cpp", CLI),
           ("symtens_ttmc_main.cpp", os.path.join(HERE, "symtens_ttmc")),
           ("s3ttmctc_main.cpp", os.path.join(HERE, "s3ttmctc"))]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++20", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(HERE, "..", "include"), "-I" + CSRC, "-I" + HOST]
# Packed FP32 VALU forms are disabled for all device code.  MFMA operands and
# accumulators are kept in VGPRs: with the default heuristic the accumulators
# of the S3TTMc kernels were copied between VGPRs and AGPRs around every MFMA.
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form"]


def _sources():
    srcs = []
    for d in (CSRC, HOST):
        if not os.path.isdir(d):
            continue
        for f in sorted(os.listdir(d)):
            if f.endswith((".hip", ".cpp")) and f not in {m for m, _ in DRIVERS}:
                srcs.append(os.path.join(d, f))
    return srcs


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), err))
    return obj, err


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources()
    jobs = min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs +
                   ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
                    "-Wl,-rpath," + os.path.join(ROCM, "lib")], check=True)
    os.replace(tmp, LIB)
    for main, exe in DRIVERS:
        subprocess.run([HIPCC] + COMMON + [os.path.join(HOST, main), "-o", exe,
                        "-L" + HERE, "-lsymprop_amd", "-Wl,-rpath,$ORIGIN",
                        "-Wl,-rpath," + os.path.join(ROCM, "lib")], check=True)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print(LIB)
