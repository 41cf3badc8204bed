#!/usr/bin/env python3
"""Benchmark of the S3TTMc hot path on one MI355X.

Workload (default): the order-6 symmetric tensor the time-to-fit target is
stated on -- n = 50,000, 5,000,000 unique sorted 6-tuples from a planted
8-community model (p_in = 0.8), values 1.0, rank 8 (packed Y width
C(12,5) = 792), seed 0.  --workload zipf4 runs the order-4 Zipf(1.1) tensor
(n = 10,000, 1M nnz, U(0,1) values, rank 10) and --workload zipf7 the
order-7 Zipf(1.2) tensor (n = 100,000, 10M nnz, rank 6, HOQRI).

A step is one S3TTMc pass (sp_s3ttmc_dev) over the tensor with the factor
and tree resident in HBM.  W untimed warm-up steps, then K timed steps
between two stream synchronisations.  The JSON line also reports:
  - roofline of the dominant kernel, timed live with HIP events on the
    library's stream (algorithmic FP64 flops of the reference's tree
    schedule / kernel time, against the FP64 peak of 78.6 TFLOP/s);
  - HOOI time-to-fit in the FP64, FP32-mixed and BF16-mixed forms, each with
    its fit error relative to the FP64 run;
  - cpu_baseline: the reference's own construction + kronecker_product_opt5
    (oracle/_ref, compiled from the reference sources) on the host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import symprop_amd as sp  # noqa: E402

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix) peak
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def binom(n, k):
    from math import comb
    return comb(n, k)


def f_alg(t, N, R, nnz):
    """Tree-shared FP64 flops of one S3TTMc under the reference schedule:
    sum_{l=1}^{N-2} L_l (l+1) 2 P_{l+1}  +  N nnz 2 P_{N-1}."""
    P = lambda k: binom(R + k - 1, k)  # noqa: E731
    f = 0
    for l in range(1, N - 1):
        f += t.level_size(l) * (l + 1) * 2 * P(l + 1)
    f += N * nnz * 2 * P(N - 1)
    return float(f)


def gemm_work(t, N, R):
    """FP64 arithmetic of the MFMA S3TTMc kernel itself (its GEMM formulation:
    per owner of a row the P_H x P_T outer product of head and tail
    coefficients, H the head degree of the compiled shape): useful = owners x
    2 P_H P_T; issued = MFMA flops of the 16 x 16 tiles over the owner slots
    of the batches (rows cut into 64-owner batches, each padded to whole
    K-steps of 4 owners) plus the VALU FMAs of the tail strip (P_T = 33..36:
    columns 33.. on the vector pipe).  None for shapes without the kernel."""
    M = N - 1
    H = {2: 1, 3: 2, 4: 2, 5: 3, 6: 3}.get(M)
    if H is None:
        return None
    PH, PT = binom(R + H - 1, H), binom(R + M - H - 1, M - H)
    TH = (PH + 15) // 16
    strip = 32 < PT <= 36
    TT = 2 if strip else (PT + 15) // 16
    nsc = PT - 32 if strip else 0
    rem, _ = t.owners()
    c = np.bincount(rem.astype(np.int64))
    slots = float((64 * (c // 64) + 4 * ((c % 64 + 3) // 4)).sum())
    owners = float(len(rem))
    return {"head_degree": H, "P_H": PH, "P_T": PT, "mfma_tiles": TH * TT,
            "tail_strip_columns": nsc, "owners": owners, "owner_slots": slots,
            "useful_flops": owners * 2.0 * PH * PT,
            "issued_flops": slots * (TH * TT * 512.0 + 2.0 * 16 * TH * nsc)}


def workload(name):
    if name == "zipf4":
        return dict(kind=sp.GEN_ZIPF, order=4, dim=10000, nnz=1_000_000, param=1.1,
                    communities=1, rank=10, max_iter=10, delta=0.5, long_iters=4000,
                    desc="order-4 n=10000 1M unique Zipf(1.1) tuples U(0,1) values, rank 10")
    if name == "planted6":
        return dict(kind=sp.GEN_PLANTED, order=6, dim=50000, nnz=5_000_000, param=0.8,
                    communities=8, rank=8, max_iter=20, delta=0.5, long_iters=100,
                    desc="order-6 n=50000 5M unique planted-8-community tuples, rank 8")
    if name == "zipf7":
        return dict(kind=sp.GEN_ZIPF, order=7, dim=100000, nnz=10_000_000, param=1.2,
                    communities=1, rank=6, max_iter=20, delta=0.5, alg=1, long_iters=300,
                    desc="order-7 n=100000 10M unique Zipf(1.2) tuples U(0,1) values, rank 6, "
                         "HOQRI")
    if name == "uniform5":
        return dict(kind=sp.GEN_UNIFORM, order=5, dim=327, nnz=8000, param=0.0,
                    communities=1, rank=8, max_iter=20, delta=0.1, long_iters=300,
                    desc="order-5 n=327 8000 uniform tuples, rank 8")
    raise SystemExit("unknown workload " + name)


def measured_traffic(kernel, workload_name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary of this
    kernel and workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "%s_%s.md" % (kernel, workload_name))
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def uniform_sample(inds, vals, k, seed=0):
    """Seeded uniform random subset of k nonzeros, kept in the tensor's order."""
    if k >= len(vals):
        return inds, vals
    idx = np.sort(np.random.default_rng(seed).choice(len(vals), size=k, replace=False))
    return np.ascontiguousarray(inds[idx]), np.ascontiguousarray(vals[idx])


def host_threads():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    return max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))


def opt5_peak_bytes(t, N, R):
    """Peak FP64 bytes of the level arrays kronecker_product_opt5 holds at
    once (two consecutive levels, src/compressed_dtree.cpp:451-512)."""
    sizes = [t.level_size(l) * binom(R + l, l + 1) * 8 for l in range(N - 1)]
    return max(sizes[l - 1] + sizes[l] for l in range(1, N - 1)) if N > 2 else sizes[0]


def cpu_baseline(inds, vals, order, dim, R, budget_s=20.0):
    """The reference's own construction + kronecker_product_opt5 (oracle/_ref,
    compiled from the reference sources) on the host cores, on a seeded
    uniform random sample of the workload's nonzeros.  A 20,000-nonzero trial
    sizes the sample so that construction + one S3TTMc take about budget_s
    and opt5's level arrays stay within a quarter of host memory (at most
    64 GB).  The GPU then runs S3TTMc on the same sample with the same factor:
    its time on the sample and the W-weighted error of its FP64, FP32 and
    BF16 outputs against the reference's output are reported."""
    try:
        sys.path.insert(0, ROOT)
        from oracle import ref
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    if not ref.available():
        return {"error": "oracle/_ref not built"}
    threads = host_threads()
    U = np.linalg.qr(np.random.default_rng(0).standard_normal((dim, R)))[0]
    P = binom(R + order - 2, order - 1)
    W = np.array(oracle_coeff(R, order - 1))

    def run_ref(a, v):
        t0 = time.perf_counter()
        h = ref.lib().ref_build_quiet(order, dim, len(v), a.ctypes.data_as(ref.u32p),
                                      v.ctypes.data_as(ref.f64p))
        tb = time.perf_counter() - t0
        Y = np.zeros((dim, P))
        t0 = time.perf_counter()
        ref.lib().ref_opt5(h, U.ctypes.data_as(ref.f64p), R, Y.ctypes.data_as(ref.f64p),
                           P, threads)
        ts = time.perf_counter() - t0
        ref.lib().ref_free(h)
        return tb, ts, Y

    k0 = min(len(vals), 20_000)
    a0, v0 = uniform_sample(inds, vals, k0)
    tb0, ts0, _ = run_ref(a0, v0)
    t0g = sp.SpSymtensor(dim, order, a0, v0)
    peak0 = opt5_peak_bytes(t0g, order, R)
    del t0g
    mem_cap = min(64e9, 0.25 * host_memory_bytes()) if host_memory_bytes() else 16e9
    k = int(min(len(vals), k0 * budget_s / max(tb0 + ts0, 1e-3),
                k0 * mem_cap / max(peak0, 1.0)))
    k = max(k, k0)
    a, v = uniform_sample(inds, vals, k)
    tb, ts, Yref = run_ref(a, v)
    # the GPU on the same sample
    tg = sp.SpSymtensor(dim, order, a, v)
    F = f_alg(tg, order, R, len(v))
    peak = opt5_peak_bytes(tg, order, R)
    tg.set_factor(U)
    tg.s3ttmc_dev()
    tg.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        tg.s3ttmc_dev()
    tg.sync()
    g_ms = (time.perf_counter() - t0) * 1e3 / 10
    den = np.sqrt((Yref ** 2 * W).sum())
    errs = {}
    for name, prec in (("fp64", sp.PREC_FP64), ("fp32", sp.PREC_FP32), ("bf16", sp.PREC_BF16)):
        Yg = tg.s3ttmc(U, precision=prec)
        errs[name] = float(np.sqrt(((Yg - Yref) ** 2 * W).sum()) / den)
    del tg
    return {"value": len(v) / ts, "unit": "nnz/s", "cores": threads, "kind": "reference",
            "sample": "seeded uniform random sample of %d of the %d nonzeros (seed 0); "
                      "reference DTree+CompressedDTree build %.2f s, one "
                      "kronecker_product_opt5 %.2f s with %d OpenMP threads"
                      % (len(v), len(vals), tb, ts, threads),
            "sample_nnz": len(v), "insertion_s": tb, "s3ttmc_s": ts,
            "alg_flops_sample": F, "cpu_tflops_alg": F / ts / 1e12,
            "opt5_peak_level_bytes_sample": peak,
            "gpu_same_sample": {"ms": g_ms, "nnz_per_s": len(v) / (g_ms * 1e-3),
                                "speedup": ts * 1e3 / g_ms},
            "s3ttmc_error_vs_reference": errs,
            "s3ttmc_error_note": "W-weighted relative Frobenius error of the GPU's packed Y "
                                 "against the reference kronecker_product_opt5 output on the "
                                 "sample, same factor"}


def oracle_coeff(R, k):
    """symtens_coeff (src/dense/symtensor.cpp:53-59): k!/prod f! per packed entry."""
    from math import factorial
    out = []
    idx = [0] * k

    def rec(pos, start):
        if pos == k:
            d = 1
            run = 1
            for q in range(1, k + 1):
                if q < k and idx[q] == idx[q - 1]:
                    run += 1
                else:
                    d *= factorial(run)
                    run = 1
            out.append(factorial(k) / d)
            return
        for j in range(start, R):
            idx[pos] = j
            rec(pos + 1, j)
    rec(0, 0)
    return out


def cpu_hooi_estimate(order, dim, nnz, R, s3ttmc_full_s, threads, basis):
    """Where the literal CPU HOOI cannot run (its unpacked n x R^(N-1) Y and
    the SVD of it), an estimate with the same arithmetic: the reference's own
    S3TTMc (opt5) time for the whole tensor extrapolated from the bounded
    sample (`basis` says how), plus a packed-Gram HOOI iteration (P x P Gram,
    eigh, U = Y V, core) timed with numpy's multi-threaded LAPACK on an n x P
    matrix of the same shape."""
    P = binom(R + order - 2, order - 1)
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((dim, P))
    w = rng.random(P) + 0.5
    t0 = time.perf_counter()
    Z = Y * np.sqrt(w)
    G = Z.T @ Z
    lam, V = np.linalg.eigh(G)
    U = Z @ (V[:, -R:] / np.sqrt(np.maximum(lam[-R:], 1e-300)))
    core = U.T @ Y
    t_dense = time.perf_counter() - t0
    t_s3 = s3ttmc_full_s
    return {"kind": "estimate", "cores": threads,
            "literal_infeasible": "the reference HOOI unpacks Y to %d x %d FP64 (%.1f GB) and "
                                  "takes its SVD each iteration" % (dim, R ** (order - 1),
                                                                    dim * R ** (order - 1) * 8e-9),
            "s_per_iter": t_s3 + t_dense, "s3ttmc_s": t_s3, "dense_s": t_dense,
            "sample": "reference opt5 time for all %d nonzeros extrapolated (%s), plus a "
                      "numpy packed-Gram iteration at n=%d, P=%d" % (nnz, basis, dim, P),
            "core_shape": list(core.shape)}


def opt5_level_bytes(t, N, R):
    """FP64 bytes of the level arrays the reference's kronecker_product_opt5
    materialises (sum over levels 1..N-2 of L_l * P_{l+1} * 8), from the
    level sizes of the tree built on the GPU."""
    return sum(t.level_size(l) * binom(R + l, l + 1) * 8 for l in range(1, N - 1))


def host_memory_bytes():
    try:
        return os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES")
    except (ValueError, OSError, AttributeError):  # pragma: no cover
        return 0


def cpu_hoqri_estimate(order, dim, nnz, R, s3ttmc_full_s, threads, level_bytes, basis):
    """CPU HOQRI per iteration where the reference's S3TTMc cannot run on the
    host (its opt5 level arrays exceed host memory): the reference's opt5
    time for the whole tensor extrapolated from the bounded sample, plus the dense part
    of a TuckerHOQRI iteration (core = U^T Y, U = Y diag(W) core^T, thin QR;
    spsymtensor.cpp:254-299) timed with numpy's multi-threaded BLAS/LAPACK on
    matrices of the same shape."""
    P = binom(R + order - 2, order - 1)
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((dim, P))
    w = rng.random(P) + 0.5
    U = np.linalg.qr(rng.standard_normal((dim, R)))[0]
    t0 = time.perf_counter()
    core = U.T @ Y
    Un = Y @ (w[:, None] * core.T)
    U, _ = np.linalg.qr(Un)
    t_dense = time.perf_counter() - t0
    t_s3 = s3ttmc_full_s
    return {"kind": "estimate", "cores": threads,
            "literal_infeasible": "one reference S3TTMc of the whole tensor takes about %.0f s "
                                  "on the host cores and materialises %.0f GB of FP64 level "
                                  "arrays (host memory %.0f GB)"
                                  % (t_s3, level_bytes * 1e-9, host_memory_bytes() * 1e-9),
            "s_per_iter": t_s3 + t_dense, "s3ttmc_s": t_s3, "dense_s": t_dense,
            "sample": "reference opt5 time for all %d nonzeros extrapolated (%s), plus a "
                      "numpy HOQRI dense step at n=%d, P=%d" % (nnz, basis, dim, P)}


def cpu_hoqri_baseline(inds, vals, order, dim, R, max_iter, delta, budget_s=40.0):
    """HOQRI on the host cores: the oracle's restatement of TuckerHOQRI
    (opt5 S3TTMc with OpenMP, numpy BLAS/LAPACK for core, Y W core^T and the
    QR), same tensor, rank and random start, bounded like cpu_hooi_baseline."""
    try:
        import oracle
        from oracle import tucker
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    threads = host_threads()
    t0 = time.perf_counter()
    T = oracle.Tree(order, dim, inds, vals)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    tucker.hoqri(T, R, 1, rel_tol=0.0, delta=delta, seed=0, nthreads=threads)
    t_one = time.perf_counter() - t0
    iters = max(1, min(max_iter, int(budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    _, errs, _ = tucker.hoqri(T, R, iters, rel_tol=0.0, delta=delta, seed=0,
                              nthreads=threads)
    t_run = time.perf_counter() - t0
    return {"kind": "port", "cores": threads, "construction_s": t_build,
            "iters": len(errs), "seconds": t_run, "s_per_iter": t_run / len(errs),
            "final_error": float(errs[-1]),
            "sample": "oracle TuckerHOQRI restatement (opt5 S3TTMc, OpenMP; numpy QR), "
                      "%d iterations" % len(errs)}


def cpu_hooi_baseline(inds, vals, order, dim, R, max_iter, delta, budget_s=40.0,
                      s3ttmc_full_s=None, basis=""):
    """HOOI on the host cores: the oracle's restatement of TuckerHOOI
    (S3TTMc in the reference's opt5 schedule with OpenMP, numpy LAPACK for
    the SVDs), same tensor, rank and random start.  Bounded: if one
    iteration exceeds the budget share, fewer iterations are run and the
    per-iteration time is reported."""
    try:
        import oracle
        from oracle import tucker
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    threads = host_threads()
    if dim * R ** (order - 1) * 8 > 2e9:
        if not s3ttmc_full_s:
            return {"error": "literal CPU HOOI infeasible and no CPU S3TTMc rate measured"}
        return cpu_hooi_estimate(order, dim, len(vals), R, s3ttmc_full_s, threads, basis)
    t0 = time.perf_counter()
    T = oracle.Tree(order, dim, inds, vals)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    tucker.hooi(T, R, 1, rel_tol=0.0, delta=delta, seed=0, nthreads=threads)
    t_one = time.perf_counter() - t0
    iters = max(1, min(max_iter, int(budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    _, errs, _ = tucker.hooi(T, R, iters, rel_tol=0.0, delta=delta, seed=0, nthreads=threads)
    t_run = time.perf_counter() - t0
    return {"kind": "port", "cores": threads, "construction_s": t_build,
            "iters": len(errs), "seconds": t_run, "s_per_iter": t_run / len(errs),
            "final_error": float(errs[-1]), "errors": [float(e) for e in errs],
            "sample": "oracle TuckerHOOI restatement (opt5 S3TTMc, %d OpenMP threads; "
                      "numpy LAPACK SVD of the unpacked n x R^(N-1) Y), the first %d "
                      "iterations from the reference's random start" % (threads, len(errs))}


def cpu_hooi_chunked(inds, vals, order, dim, R, iters, delta, chunk_nnz):
    """HOOI on the host cores, measured end to end, for a tensor whose opt5
    level arrays do not fit host memory in one piece (the order-6 planted
    tensor's need about 300 GB).  The nonzeros (sorted) are cut into
    contiguous chunks of chunk_nnz; each chunk's tree is built once by the
    reference's own DTree + CompressedDTree (oracle/_ref), and each iteration
    runs the reference's HOOI loop (spsymtensor.cpp:201-252): regularize,
    Y = sum over the chunks of the reference's kronecker_product_opt5 (S3TTMc
    is linear in the tensor) on all host cores, then the leading left
    singular vectors of the unpacked Y through the packed Gram (numpy eigh of
    G = Z^T Z, Z = Y diag(sqrt W): the n x 32,768 unpacked Y of the literal
    dgesvd would need 13 GB per iteration), the core and the error.  Same
    random start as the GPU run (mt19937 + normal, seed 0)."""
    sys.path.insert(0, ROOT)
    import oracle
    from oracle import ref
    from oracle import tucker
    if not ref.available():
        return {"error": "oracle/_ref not built"}
    threads = host_threads()
    P = binom(R + order - 2, order - 1)
    W = np.array(oracle_coeff(R, order - 1))
    sw = np.sqrt(W)
    norm = oracle.norm(order, inds, vals)
    t0 = time.perf_counter()
    trees = []
    for c0 in range(0, len(vals), chunk_nnz):
        a = np.ascontiguousarray(inds[c0:c0 + chunk_nnz], dtype=np.uint32)
        v = np.ascontiguousarray(vals[c0:c0 + chunk_nnz])
        tb = time.perf_counter()
        trees.append(ref.lib().ref_build_quiet(order, dim, len(v), a.ctypes.data_as(ref.u32p),
                                               v.ctypes.data_as(ref.f64p)))
        log("cpu chunked HOOI: chunk %d (%d nonzeros) built in %.1f s"
            % (len(trees), len(v), time.perf_counter() - tb))
    build_s = time.perf_counter() - t0
    U = oracle.random_factor(0, dim, R)
    errs, times, s3_times = [], [], []
    try:
        for it in range(iters):
            ti = time.perf_counter()
            U = tucker.regularize(U, delta)
            Uc = np.ascontiguousarray(U)
            Y = np.zeros((dim, P))
            Yc = np.zeros((dim, P))
            ts = time.perf_counter()
            for h in trees:
                ref.lib().ref_opt5(h, Uc.ctypes.data_as(ref.f64p), R,
                                   Yc.ctypes.data_as(ref.f64p), P, threads)
                Y += Yc
            s3_times.append(time.perf_counter() - ts)
            Z = Y * sw
            lam, V = np.linalg.eigh(Z.T @ Z)
            V, lam = V[:, ::-1][:, :R], lam[::-1][:R]
            U = Z @ V / np.sqrt(lam)
            core = U.T @ Y
            err = tucker.fit_error(norm, np.sqrt((core ** 2 * W).sum()))
            errs.append(float(err))
            times.append(time.perf_counter() - ti)
            log("cpu chunked HOOI: iteration %d %.1f s (S3TTMc %.1f s), error %.12f"
                % (it + 1, times[-1], s3_times[-1], err))
    finally:
        for h in trees:
            ref.lib().ref_free(h)
    return {"kind": "reference", "cores": threads, "chunks": len(trees),
            "chunk_nnz": chunk_nnz, "construction_s": build_s, "iters": len(errs),
            "s_per_iter": float(np.mean(times)), "iteration_s": times,
            "s3ttmc_s": s3_times, "errors": errs,
            "sample": "whole tensor in %d contiguous chunks of %d sorted nonzeros: the "
                      "reference's tree per chunk (built once) and kronecker_product_opt5 "
                      "summed over the chunks with %d OpenMP threads per iteration; numpy "
                      "packed-Gram eigen-step" % (len(trees), chunk_nnz, threads)}


def dense_y_side(t, N, n, R, delta, iters=3):
    """One HOOI iteration's device work, op by op (HIP events on the library
    stream), and the MFMA utilisation of the Y-side contractions: the Gram
    G = diag(sqrt W) Y^T Y diag(sqrt W) (64 x 64 upper tiles, FP64 MFMA),
    U^T Y and Y M^T."""
    P = binom(R + N - 2, N - 1)
    # one untimed iteration: the first rocSOLVER call of a process also
    # loads its kernels and sizes its workspace
    t.regularize(delta)
    t.s3ttmc_dev()
    t.leftsvd_y()
    t.gemmtn()
    t.sync()
    t.timing(True)
    t.timing_reset()
    for _ in range(iters):
        t.regularize(delta)
        t.s3ttmc_dev()
        t.leftsvd_y()
        t.gemmtn()
    t.sync()
    ops = {}
    for name in ("s3ttmc_mfma64", "s3ttmc_generic", "reduce_split_rows", "gram", "syevd",
                 "syevj", "eig_warm", "sytrd", "tridiag_top", "ormtr", "y_times_v", "gemmtn",
                 "gram_u", "qr"):
        ms, cnt = t.timing_get(name)
        if cnt:
            ops[name] = ms / iters
    warm = {"attempts": t.timing_get("eig_warm")[1], "taken": t.timing_get("eig_warm_ok")[1],
            "iterations": iters}
    t.timing(False)
    nt = (P + 63) // 64
    tiles = nt * (nt + 1) // 2
    g_ms = ops.get("gram", 0.0)
    out = {"ms_per_iteration": ops, "P": P, "warm_eigen_step": warm}
    if g_ms > 0:
        executed = tiles * 64 * 64 * 2.0 * n   # MFMA flops issued (upper tiles, padded)
        algorithmic = 2.0 * n * P * P          # the full P x P product
        out["gram"] = {"kernel_ms": g_ms,
                       "mfma_tflops": executed / (g_ms * 1e-3) / 1e12,
                       "mfma_utilisation": executed / (g_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                       "alg_tflops": algorithmic / (g_ms * 1e-3) / 1e12,
                       "peak_tflops": FP64_PEAK_TFLOPS}
    y_bytes = 8.0 * n * P
    for name in ("gemmtn", "y_times_v"):
        if ops.get(name):
            out[name] = {"kernel_ms": ops[name],
                         "hbm_gbs": y_bytes / (ops[name] * 1e-3) / 1e9,
                         "hbm_frac": y_bytes / (ops[name] * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return out


def fit_study(t, alg, R, delta, long_iters, sub_tol=1e-7):
    """HOOI/HOQRI time to fit in the FP64, FP32-mixed and BF16-mixed forms.

    reference_rule: every form stops on the reference loop's own test
      |error - previous| < 1e-6 (spsymtensor.cpp:235-238, 283-286), up to 500
      iterations.  On slowly converging tensors this stops while the error
      still drifts, at a point that depends on rounding in early iterations.
    converged: every form runs until the factor's subspace settles (sine of
      the largest principal angle between consecutive FP64 factors below
      sub_tol) or long_iters iterations.  The FP64 run's last error e* and
      captured energy |C*|^2 define the fit; for each form, time_to_fit_error
      is the time to the first iteration with |e - e*| <= 1e-6 e* (the
      metric's relative fit error) and time_to_fit_energy the time to the
      first iteration with |C|^2 >= (1 - 1e-6) |C*|^2 (the same bar on the
      captured energy, which stays informative when e is 1 - O(1e-7)).
    Times include the initial factor and every iteration up to that point."""
    from symprop_amd import learned
    forms = (("fp64", sp.PREC_FP64), ("fp32_mixed", sp.PREC_FP32),
             ("bf16_mixed", sp.PREC_BF16))
    base = dict(alg=alg, rank=R, delta=delta, seed=0, rel_tol=1e-6)
    ref_rule = {}
    for name, prec in forms:
        r = t.tucker(max_iter=500, precision=prec, **base)
        ref_rule[name] = {"iters": r["iters"], "fp64_sweeps": r["fp64_sweeps"],
                          "seconds": r["seconds"], "final_error": float(r["errors"][-1]),
                          "captured_energy": float(r["core_norms"][-1] ** 2)}
        log("%s reference rule: %.3f s, %d iters, error %.12f"
            % (name, r["seconds"], r["iters"], r["errors"][-1]))
    e64 = ref_rule["fp64"]["final_error"]
    c64 = ref_rule["fp64"]["captured_energy"]
    for v in ref_rule.values():
        v["rel_fit_error_vs_fp64"] = abs(v["final_error"] - e64) / e64
        v["rel_energy_diff_vs_fp64"] = (v["captured_energy"] - c64) / c64
    runs = {}
    for name, prec in forms:
        runs[name] = t.tucker(max_iter=long_iters, precision=prec, subspace_tol=sub_tol, **base)
        log("%s converged run: %.3f s, %d iters, error %.12f"
            % (name, runs[name]["seconds"], runs[name]["iters"], runs[name]["errors"][-1]))
    r64 = runs["fp64"]
    e_star = float(r64["errors"][-1])
    c_star = float(r64["core_norms"][-1] ** 2)
    conv = {}
    for name, r in runs.items():
        e = r["errors"]
        c2 = r["core_norms"] ** 2
        ke = np.nonzero(np.abs(e - e_star) <= 1e-6 * e_star)[0]
        kc = np.nonzero(c2 >= (1 - 1e-6) * c_star)[0]
        conv[name] = {
            "iters": r["iters"], "settled": bool(r["iters"] < long_iters),
            "fp64_sweeps": r["fp64_sweeps"], "seconds": r["seconds"],
            "final_error": float(e[-1]), "captured_energy": float(c2[-1]),
            "rel_fit_error_vs_fp64": abs(float(e[-1]) - e_star) / e_star,
            "rel_energy_diff_vs_fp64": (float(c2[-1]) - c_star) / c_star,
            "sin_to_fp64_factor": learned.sin_angle(r["U"], r64["U"]),
            "time_to_fit_error": {"iters": int(ke[0]) + 1, "seconds": float(r["times"][ke[0]])}
            if len(ke) else None,
            "time_to_fit_energy": {"iters": int(kc[0]) + 1, "seconds": float(r["times"][kc[0]])}
            if len(kc) else None}
    return {"reference_rule": ref_rule,
            "converged": {"criterion": "sin angle between consecutive FP64 factors < %g, at "
                                       "most %d iterations" % (sub_tol, long_iters),
                          "fp64_error": e_star, "fp64_captured_energy": c_star,
                          "forms": conv}}


def hoevd_timing(t, R):
    """The HOEVD initial factor (top-R eigenvectors of X_(1) X_(1)^T,
    spsymtensor.cpp:363-382) on the workload tensor, with the solver the
    library picks for its dim, timed after a first call (which also sets up
    the operator's buffers)."""
    t0 = time.perf_counter()
    t.hoevd_init(R)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, w = t.hoevd_init(R)
    sec = time.perf_counter() - t0
    return {"solver": ("block Krylov on the implicit X(1) X(1)^T" if t.dim > 4096 else
                       "dense X(1) X(1)^T and the top-R eigen-step"),
            "first_s": first, "seconds": sec,
            "top_eigenvalues": [float(x) for x in w[::-1]]}


def planted_basis(n, K):
    E = np.zeros((n, K))
    b = n // K
    for c in range(K):
        E[c * b:(c + 1) * b, c] = 1.0
    return np.linalg.qr(E)[0]


def learned_study(dim, nnz, pin, ntrain, ntest, iters, steps):
    """Learned warm start on seeded order-5 planted-partition tensors
    (8 communities, binary values, rank 8, TuckerHOOI with delta 0.5).

    Training targets: the factors plain HOOI converges to from the
    reference's own random start (Config seed 0) on `ntrain` seeds.  Held-out
    seeds run HOOI from that random start, HOEVD (its fallback), the
    hypergraph sketch, the non-learned feature start (top pair-multiplicity
    nodes) and the learned start, until the subspace moves by less than 1e-6
    (sine of the largest principal angle between consecutive factors) or
    `iters` iterations; each start reports iterations, converged fit error,
    captured energy |C|^2, angle to the planted communities and time
    including the cost of producing the start."""
    from symprop_amd import learned
    N, K, R, delta = 5, 8, 8, 0.5
    E = planted_basis(dim, K)
    kw = dict(alg=0, rank=R, max_iter=iters, rel_tol=1e-6, delta=delta, seed=0,
              subspace_tol=1e-6)

    def tensor(seed):
        inds, vals = sp.generate(sp.GEN_PLANTED, N, dim, nnz, pin, K, seed=seed)
        t = sp.SpSymtensor(dim, N, inds, vals)
        t.sync()
        t.inds_host = inds
        return t

    feats, facs, teacher_its = [], [], []
    t0 = time.perf_counter()
    for s in range(1, ntrain + 1):
        t = tensor(s)
        r = t.tucker(**kw)
        teacher_its.append(r["iters"])
        feats.append(learned.features(t))
        facs.append(r["U"])
        del t
        if s % 10 == 0 or s == ntrain:
            log("learned: teacher run %d/%d, %.1f s" % (s, ntrain, time.perf_counter() - t0))
    t_teacher = time.perf_counter() - t0
    model = learned.WarmStartModel(R, seed=0)
    t0 = time.perf_counter()
    losses = model.fit(feats, facs, steps=steps, batch=4096, lr=1e-3)
    t_fit = time.perf_counter() - t0
    log("learned: %d teacher runs %.1f s, fit %.1f s, loss %.4f -> %.4f"
        % (ntrain, t_teacher, t_fit, float(losses[:5].mean()), float(losses[-10:].mean())))
    del feats, facs

    starts = ("random", "hoevd", "sketch", "feature", "learned")
    rows = {k: [] for k in starts}
    for s in range(10001, 10001 + ntest):
        t = tensor(s)
        for k in starts:
            t0 = time.perf_counter()
            U0 = None
            if k == "hoevd":
                U0, _ = t.hoevd_init(R)
            elif k != "random":
                F = learned.features(t)
                U0 = (learned.spectral_start(F, R) if k == "sketch" else
                      learned.feature_start(F, R, t.inds_host) if k == "feature" else
                      model.predict(F, t.inds_host))
            t_init = time.perf_counter() - t0
            r = t.tucker(U0=U0, **kw) if U0 is not None else t.tucker(**kw)
            rows[k].append({"seed": s, "iters": r["iters"],
                            "converged": bool(r["iters"] < iters),
                            "final_error": float(r["errors"][-1]),
                            "captured_energy": float(r["core_norms"][-1] ** 2),
                            "sin_to_planted": learned.sin_angle(r["U"], E),
                            "init_s": t_init, "seconds": t_init + r["seconds"]})
        log("learned seed %d: " % s + ", ".join(
            "%s %d it |C|^2 %.1f" % (k, rows[k][-1]["iters"], rows[k][-1]["captured_energy"])
            for k in starts))
        del t

    def summary(rs):
        conv = [r["iters"] for r in rs if r["converged"]]
        return {"converged": "%d/%d" % (len(conv), len(rs)),
                "median_iters": float(np.median([r["iters"] for r in rs])),
                "mean_final_error": float(np.mean([r["final_error"] for r in rs])),
                "mean_captured_energy": float(np.mean([r["captured_energy"] for r in rs])),
                "mean_sin_to_planted": float(np.mean([r["sin_to_planted"] for r in rs])),
                "median_seconds": float(np.median([r["seconds"] for r in rs])),
                "runs": rs}

    return {"config": {"order": N, "dim": dim, "nnz": nnz, "p_in": pin, "communities": K,
                       "rank": R, "delta": delta, "max_iters": iters,
                       "train_tensors": ntrain, "held_out": ntest,
                       "criterion": "sin(U_k, U_k-1) < 1e-6"},
            "train": {"teacher": "HOOI from the reference's random start (Config seed 0)",
                      "teacher_iters": teacher_its, "teacher_s": t_teacher,
                      "fit_s": t_fit, "steps": steps,
                      "loss_first": float(losses[:5].mean()),
                      "loss_last": float(losses[-10:].mean())},
            "held_out": {k: summary(rows[k]) for k in starts}}


def learned_study_hoqri(nnz, ntrain, ntest, steps):
    """Learned warm start for TuckerHOQRI on seeded order-7 Zipf(1.2) tensors
    (n = 100,000, U(0,1) values, rank 6, delta 0.5).

    Training targets: the factors plain HOQRI converges to from the
    reference's random start on `ntrain` seeds.  Held-out seeds run HOQRI
    under the reference stopping rule |error - previous| < 1e-6 from the
    random start, the hypergraph sketch start and the learned start; each
    reports iterations, final error, its fit error relative to the
    random-start run, and time including the cost of producing the start."""
    from symprop_amd import learned
    N, n, R, delta = 7, 100000, 6, 0.5

    def tensor(seed):
        inds, vals = sp.generate(sp.GEN_ZIPF, N, n, nnz, 1.2, 1, seed=seed)
        t = sp.SpSymtensor(n, N, inds, vals)
        t.sync()
        return t

    def run(t, U0):
        kw = dict(alg=1, rank=R, max_iter=500, rel_tol=1e-6, delta=delta, seed=0)
        return t.tucker(U0=U0, **kw) if U0 is not None else t.tucker(**kw)

    feats, facs, teacher_its = [], [], []
    t0 = time.perf_counter()
    for s in range(1, ntrain + 1):
        t = tensor(s)
        r = run(t, None)
        teacher_its.append(r["iters"])
        feats.append(learned.features(t))
        facs.append(r["U"])
        del t
        log("learned hoqri: teacher run %d/%d (%d iterations), %.1f s"
            % (s, ntrain, r["iters"], time.perf_counter() - t0))
    t_teacher = time.perf_counter() - t0
    model = learned.WarmStartModel(R, seed=0)
    t0 = time.perf_counter()
    losses = model.fit(feats, facs, steps=steps, batch=4096, lr=1e-3)
    t_fit = time.perf_counter() - t0
    del feats, facs
    log("learned hoqri: fit %.1f s, loss %.4f -> %.4f"
        % (t_fit, float(losses[:5].mean()), float(losses[-10:].mean())))
    starts = ("random", "sketch", "learned")
    rows = {k: [] for k in starts}
    for s in range(10001, 10001 + ntest):
        t = tensor(s)
        e_rand = None
        for k in starts:
            t0 = time.perf_counter()
            U0 = None
            if k != "random":
                F = learned.features(t)
                U0 = learned.spectral_start(F, R) if k == "sketch" else model.predict(F)
            t_init = time.perf_counter() - t0
            r = run(t, U0)
            e = float(r["errors"][-1])
            if k == "random":
                e_rand = e
            rows[k].append({"seed": s, "iters": r["iters"], "final_error": e,
                            "rel_fit_error_vs_random": abs(e - e_rand) / e_rand,
                            "init_s": t_init, "seconds": t_init + r["seconds"]})
        log("learned hoqri seed %d: " % s + ", ".join(
            "%s %d it err %.10f" % (k, rows[k][-1]["iters"], rows[k][-1]["final_error"])
            for k in starts))
        del t

    def summary(rs):
        return {"median_iters": float(np.median([r["iters"] for r in rs])),
                "mean_final_error": float(np.mean([r["final_error"] for r in rs])),
                "max_rel_fit_error_vs_random": float(max(r["rel_fit_error_vs_random"]
                                                         for r in rs)),
                "median_seconds": float(np.median([r["seconds"] for r in rs])),
                "runs": rs}

    return {"config": {"order": N, "dim": n, "nnz": nnz, "zipf_s": 1.2, "rank": R,
                       "delta": delta, "algorithm": "HOQRI", "train_tensors": ntrain,
                       "held_out": ntest, "criterion": "|error - previous| < 1e-6"},
            "train": {"teacher": "HOQRI from the reference's random start",
                      "teacher_iters": teacher_its, "teacher_s": t_teacher, "fit_s": t_fit,
                      "steps": steps, "loss_first": float(losses[:5].mean()),
                      "loss_last": float(losses[-10:].mean())},
            "held_out": {k: summary(rows[k]) for k in starts}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="planted6")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--s3ttmc-only", action="store_true",
                    help="construction and the timed S3TTMc steps only (for profiling)")
    ap.add_argument("--no-hooi", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-only", action="store_true")
    ap.add_argument("--learned-dim", type=int, default=20000)
    ap.add_argument("--learned-nnz", type=int, default=2_000_000)
    ap.add_argument("--learned-pin", type=float, default=0.8)
    ap.add_argument("--learned-train", type=int, default=8)
    ap.add_argument("--learned-test", type=int, default=4)
    ap.add_argument("--learned-iters", type=int, default=40)
    ap.add_argument("--learned-steps", type=int, default=800)
    ap.add_argument("--learned-hoqri-only", action="store_true",
                    help="only run the HOQRI learned warm-start study on order-7 Zipf tensors")
    ap.add_argument("--learned-hoqri-nnz", type=int, default=10_000_000)
    ap.add_argument("--learned-hoqri-train", type=int, default=16)
    ap.add_argument("--learned-hoqri-test", type=int, default=4)
    ap.add_argument("--cpu-hooi-chunked", type=int, default=0, metavar="ITERS",
                    help="only measure ITERS iterations of the host-core HOOI with the "
                         "reference's opt5 over contiguous chunks of the workload, next to "
                         "the same iterations on the GPU, and exit")
    ap.add_argument("--chunk-nnz", type=int, default=600_000)
    ap.add_argument("--hooi-only", type=int, default=0, metavar="ITERS",
                    help="only run ITERS FP64 iterations of the workload's Tucker loop "
                         "(after the tree build; used for kernel traces) and exit")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB "
                         "(reference for HBM counter scaling) and exit")
    args = ap.parse_args()
    if args.calib_copy_mb:
        import torch
        n = args.calib_copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": n * 4}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    learned_args = (args.learned_dim, args.learned_nnz, args.learned_pin, args.learned_train,
                    args.learned_test, args.learned_iters, args.learned_steps)
    if args.learned_only:
        print(json.dumps({"learned_warm_start": learned_study(*learned_args)}), flush=True)
        return
    if args.learned_hoqri_only:
        print(json.dumps({"learned_warm_start_hoqri": learned_study_hoqri(
            args.learned_hoqri_nnz, args.learned_hoqri_train, args.learned_hoqri_test,
            args.learned_steps)}), flush=True)
        return

    w = workload(args.workload)
    N, n, R = w["order"], w["dim"], w["rank"]
    log("generating", w["desc"])
    inds, vals = sp.generate(w["kind"], N, n, w["nnz"], w["param"], w["communities"], 0)
    nnz = len(vals)
    log("nnz", nnz)
    if args.hooi_only:
        tg = sp.SpSymtensor(n, N, inds, vals)
        r = tg.tucker(alg=w.get("alg", 0), rank=R, max_iter=args.hooi_only, rel_tol=0.0,
                      delta=w["delta"], seed=0)
        print(json.dumps({"workload": args.workload, "iters": r["iters"],
                          "seconds": r["seconds"],
                          "iteration_times": [float(x) for x in np.diff(r["times"],
                                                                        prepend=0.0)],
                          "errors": [float(e) for e in r["errors"]]}), flush=True)
        return
    if args.cpu_hooi_chunked:
        tg = sp.SpSymtensor(n, N, inds, vals)
        r = tg.tucker(alg=0, rank=R, max_iter=args.cpu_hooi_chunked, rel_tol=0.0,
                      delta=w["delta"], seed=0)
        del tg
        cb = cpu_hooi_chunked(inds, vals, N, n, R, args.cpu_hooi_chunked, w["delta"],
                              args.chunk_nnz)
        if "errors" in cb:
            cb["gpu_fp64_s_per_iter"] = r["seconds"] / max(1, r["iters"])
            cb["gpu_errors"] = [float(e) for e in r["errors"]]
            cb["error_max_diff_vs_gpu"] = float(np.max(np.abs(
                np.array(cb["errors"]) - r["errors"][:len(cb["errors"])])))
            cb["gpu_speedup_per_iteration"] = cb["s_per_iter"] / cb["gpu_fp64_s_per_iter"]
        print(json.dumps({"workload": args.workload, "cpu_hooi_chunked": cb}), flush=True)
        return

    def build_parts():
        """(upload + device checks + norm, S3TTMc-ready build, interior levels)."""
        t0 = time.perf_counter()
        tt = sp.SpSymtensor(n, N, inds, vals, build=False)
        tt.sync()
        t1 = time.perf_counter()
        tt.build()
        tt.sync()
        t2 = time.perf_counter()
        tt.level_size(0)   # interior levels and their CSR links, built on first use
        tt.sync()
        t3 = time.perf_counter()
        return tt, (t1 - t0, t2 - t1, t3 - t2)

    t, first = build_parts()
    insertion_s = first[0] + first[1]
    # a second build of the same tensor: the first one also pays the
    # process's one-time HIP module loading and allocator warm-up
    t2, steady = build_parts()
    del t2
    steady_s = steady[1]
    # algorithmic bytes of the whole tree (every level): one read and one
    # write of every sub-multiset key (k indices of 4 bytes, k = 1..N-1) and
    # of every owner record (row index + value, 12 bytes)
    cons_bytes = float(sum(nnz * binom(N, k) * 4 * k * 2 for k in range(1, N)) +
                       2 * N * nnz * 12)
    full_s = steady[1] + steady[2]
    construction = {"first_s": insertion_s, "steady_s": steady_s,
                    "steady_parts_s": {"upload_checks_norm": steady[0],
                                       "leaf_owners_worklist": steady[1],
                                       "interior_levels": steady[2]},
                    "note": "steady_s is sp_cdtree_build: the leaf level, owners and the "
                            "rem-sorted S3TTMc work list; the interior levels with their CSR "
                            "links (and the exclude-one lists) are built on first use",
                    "full_tree_s": full_s, "alg_bytes_full_tree": cons_bytes,
                    "full_tree_gbs": cons_bytes / full_s / 1e9,
                    "full_tree_hbm_frac": cons_bytes / full_s / 1e9 / HBM_PEAK_GBS}
    log("construction %.3f s (steady build %.3f s, parts %s)" % (insertion_s, steady_s, steady))

    rng = np.random.default_rng(0)
    U = np.linalg.qr(rng.standard_normal((n, R)))[0]
    t.set_factor(U)
    for _ in range(args.warmup):
        t.s3ttmc_dev()
    t.sync()
    t.timing(True)
    t.timing_reset()
    t.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.s3ttmc_dev()
    t.sync()
    wall = time.perf_counter() - t0
    t.timing(False)
    ms_step = wall * 1e3 / args.steps
    kname = "s3ttmc_generic"
    for name in ("s3ttmc_mfma64", "s3ttmc_generic"):
        k_ms, k_cnt = t.timing_get(name)
        if k_cnt:
            kname = name
            break
    if not k_cnt:
        raise SystemExit("no S3TTMc kernel launch was timed")
    r_ms, r_cnt = t.timing_get("reduce_split_rows")
    k_avg = k_ms / max(1, k_cnt)
    Y = t.y()
    if not np.all(np.isfinite(Y)):
        raise SystemExit("non-finite S3TTMc output")
    # the same step with the n x R factor uploaded from host memory first
    # (the boundary's host-buffer form)
    t.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t.set_factor(U)
        t.s3ttmc_dev()
    t.sync()
    ms_step_upload = (time.perf_counter() - t0) * 1e3 / args.steps
    F = f_alg(t, N, R, nnz)
    achieved = F / (k_avg * 1e-3) / 1e12
    traffic, traffic_src = measured_traffic(kname, args.workload)
    log("S3TTMc %.3f ms/step (kernel %.3f ms), %.3e nnz/s, %.2f TFLOP/s alg"
        % (ms_step, k_avg, nnz / (ms_step * 1e-3), achieved))

    out = {
        "metric": "S3TTMc nnz/s + HOOI time-to-fit (mixed/learned vs FP64 GPU) at 1e-6 rel err",
        "value": nnz / (ms_step * 1e-3),
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded generator, %s)" % w["desc"],
        "config": {"workload": args.workload, "order": N, "dim": n, "nnz": nnz,
                   "rank": R, "packed_width": binom(R + N - 2, N - 1),
                   "construction_s": insertion_s},
        "with_factor_upload": {"ms_per_step": ms_step_upload,
                               "nnz_per_s": nnz / (ms_step_upload * 1e-3),
                               "note": "each step first copies the n x R FP64 factor from "
                                       "host memory (sp_factor_upload)"},
        "construction": construction,
        "roofline": {"bound": "mfma", "roof": "FP64 arithmetic (MI355X FP64 vector = "
                     "matrix peak)", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "frac_basis": "the reference schedule's tree-shared FP64 flops "
                                   "(alg_flops) / average kernel time (HIP events); above 1 "
                                   "where the kernel's GEMM formulation needs fewer flops "
                                   "than that schedule (its own arithmetic: executed)",
                     "traffic": traffic, "traffic_unit": "bytes per launch",
                     "traffic_source": traffic_src, "kernel": kname,
                     "kernel_ms": k_avg, "alg_flops": F,
                     "split_reduce_ms": r_ms / max(1, r_cnt)},
    }
    if kname == "s3ttmc_mfma64":
        ex = gemm_work(t, N, R)
        if ex:
            sec = k_avg * 1e-3
            ex["useful_tflops"] = ex["useful_flops"] / sec / 1e12
            ex["issued_tflops"] = ex["issued_flops"] / sec / 1e12
            ex["useful_frac"] = ex["useful_tflops"] / FP64_PEAK_TFLOPS
            ex["issued_frac"] = ex["issued_tflops"] / FP64_PEAK_TFLOPS
            out["roofline"]["executed"] = ex
    if traffic:
        gbs = traffic / (k_avg * 1e-3) / 1e9
        out["roofline"]["hbm_achieved_gbs"] = gbs
        out["roofline"]["hbm_frac"] = gbs / HBM_PEAK_GBS
    if args.s3ttmc_only:
        print(json.dumps(out), flush=True)
        return
    out["dense_y_side"] = dense_y_side(t, N, n, R, w["delta"])
    out["hoevd_init"] = hoevd_timing(t, R)
    log("HOEVD init %.3f s (%s)" % (out["hoevd_init"]["seconds"], out["hoevd_init"]["solver"]))

    alg = w.get("alg", 0)
    aname = ("hooi", "hoqri")[alg]
    level_bytes = opt5_level_bytes(t, N, R)
    hooi_fixed = None
    if not args.no_hooi:
        # the configuration's fixed iteration count, FP64 (compared with the
        # CPU baseline over the same iterations)
        r = t.tucker(alg=alg, rank=R, max_iter=w["max_iter"], rel_tol=0.0, delta=w["delta"],
                     seed=0, precision=sp.PREC_FP64)
        hooi_fixed = {"iters": r["iters"], "seconds": r["seconds"],
                      "s_per_iter": r["seconds"] / max(1, r["iters"]),
                      "final_error": float(r["errors"][-1]),
                      "errors": [float(e) for e in r["errors"]],
                      "iteration_times": [float(x) for x in np.diff(r["times"], prepend=0.0)]}
        out["%s_fixed_iterations" % aname] = hooi_fixed
        if alg != 1:
            # the same ops at the run's last factor, where the eigen-step's
            # warm-started Rayleigh-Ritz form is taken when it converges
            t.set_factor(r["U"])
            out["dense_y_side_at_fixed_point"] = dense_y_side(t, N, n, R, w["delta"])
        out["%s_time_to_fit" % aname] = fit_study(t, alg, R, w["delta"], w["long_iters"])

    if not args.no_learned:
        del t
        log("learned warm start study")
        out["learned_warm_start"] = learned_study(*learned_args)

    if not args.no_cpu:
        log("cpu baseline (reference build on host cores)")
        cb0 = cpu_baseline(inds, vals, N, n, R)
        out["cpu_baseline"] = cb0
        if not args.no_hooi and "value" in cb0:
            log("cpu %s baseline (host cores)" % aname.upper())
            threads = cb0["cores"]
            # the reference's opt5 time for the whole tensor from its flop
            # rate on the sample (the tree-shared flops of the whole tensor
            # over the sample's flops per second)
            s3_full = F / (cb0["cpu_tflops_alg"] * 1e12)
            basis = ("whole-tensor tree-shared flops %.3e / the %.3e flop/s the reference's "
                     "opt5 reached on the %d-nonzero uniform sample"
                     % (F, cb0["cpu_tflops_alg"] * 1e12, cb0["sample_nnz"]))
            if alg == 0:
                cb = cpu_hooi_baseline(inds, vals, N, n, R, w["max_iter"], w["delta"],
                                       s3ttmc_full_s=s3_full, basis=basis)
            elif level_bytes > 0.5 * host_memory_bytes() or s3_full > 20.0:
                cb = cpu_hoqri_estimate(N, n, nnz, R, s3_full, threads, level_bytes, basis)
            else:
                cb = cpu_hoqri_baseline(inds, vals, N, n, R, w["max_iter"], w["delta"])
            out["cpu_%s_baseline" % aname] = cb
            if "s_per_iter" in cb and hooi_fixed:
                cb["gpu_fp64_speedup_per_iteration"] = cb["s_per_iter"] / hooi_fixed["s_per_iter"]
                if cb.get("kind") == "port" and cb.get("iters") == hooi_fixed["iters"]:
                    cb["gpu_fp64_speedup_fixed_iterations"] = (cb["seconds"] /
                                                               hooi_fixed["seconds"])
                    cb["error_trajectory_max_diff_vs_gpu"] = float(np.max(np.abs(
                        np.array(cb["errors"]) - np.array(hooi_fixed["errors"]))))
                # time to fit on the CPU: its measured (or estimated) time per
                # iteration times the iterations the FP64 GPU run needed on the
                # same trajectory (same start, same arithmetic up to rounding)
                fit = out.get("%s_time_to_fit" % aname, {})
                est = {}
                f64r = fit.get("reference_rule", {}).get("fp64")
                if f64r:
                    est["reference_rule_s"] = cb["s_per_iter"] * f64r["iters"]
                f64c = fit.get("converged", {}).get("forms", {}).get("fp64", {})
                for key in ("time_to_fit_error", "time_to_fit_energy"):
                    if f64c.get(key):
                        est[key + "_s"] = cb["s_per_iter"] * f64c[key]["iters"]
                cb["time_to_fit_estimate"] = est
                cb["time_to_fit_estimate_note"] = (
                    "estimate: CPU seconds per iteration x the iteration count of the FP64 "
                    "GPU run on the same trajectory; not measured end to end")
                for sec in ("reference_rule",):
                    for v in fit.get(sec, {}).values():
                        if "seconds" in v and "reference_rule_s" in est:
                            v["speedup_vs_cpu_estimate"] = est["reference_rule_s"] / v["seconds"]
                for v in fit.get("converged", {}).get("forms", {}).values():
                    for key in ("time_to_fit_error", "time_to_fit_energy"):
                        if v.get(key) and est.get(key + "_s"):
                            v[key]["speedup_vs_cpu_estimate"] = est[key + "_s"] / v[key]["seconds"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
