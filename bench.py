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


def workload(name):
    if name == "zipf4":
        return dict(kind=sp.GEN_ZIPF, order=4, dim=10000, nnz=1_000_000, param=1.1,
                    communities=1, rank=10, max_iter=10, delta=0.5,
                    desc="order-4 n=10000 1M unique Zipf(1.1) tuples U(0,1) values, rank 10")
    if name == "planted6":
        return dict(kind=sp.GEN_PLANTED, order=6, dim=50000, nnz=5_000_000, param=0.8,
                    communities=8, rank=8, max_iter=20, delta=0.5,
                    desc="order-6 n=50000 5M unique planted-8-community tuples, rank 8")
    if name == "zipf7":
        return dict(kind=sp.GEN_ZIPF, order=7, dim=100000, nnz=10_000_000, param=1.2,
                    communities=1, rank=6, max_iter=20, delta=0.5, alg=1,
                    desc="order-7 n=100000 10M unique Zipf(1.2) tuples U(0,1) values, rank 6, "
                         "HOQRI")
    if name == "uniform5":
        return dict(kind=sp.GEN_UNIFORM, order=5, dim=327, nnz=8000, param=0.0,
                    communities=1, rank=8, max_iter=20, delta=0.1,
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


def cpu_baseline(inds, vals, order, dim, R, budget_s=25.0):
    """Reference construction + opt5 on the host cores (oracle/_ref)."""
    try:
        sys.path.insert(0, ROOT)
        from oracle import ref
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    if not ref.available():
        return {"error": "oracle/_ref not built"}
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    threads = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    rng = np.random.default_rng(0)
    U = rng.standard_normal((dim, R))
    P = binom(R + order - 2, order - 1)

    def run(n):
        a = np.ascontiguousarray(inds[:n])
        v = np.ascontiguousarray(vals[:n])
        t0 = time.perf_counter()
        h = ref.lib().ref_build(order, dim, n, a.ctypes.data_as(ref.u32p),
                                v.ctypes.data_as(ref.f64p))
        tb = time.perf_counter() - t0
        Y = np.zeros((dim, P))
        t0 = time.perf_counter()
        ref.lib().ref_opt5(h, U.ctypes.data_as(ref.f64p), R, Y.ctypes.data_as(ref.f64p),
                           P, threads)
        ts = time.perf_counter() - t0
        ref.lib().ref_free(h)
        return tb, ts

    # bounded sample: a trial on a prefix sizes the measured prefix so that
    # construction + one opt5 call stay within the budget
    n = min(len(vals), 100_000)
    tb, ts = run(n)
    if n < len(vals):
        scale = budget_s / max(tb + ts, 1e-3)
        n = min(len(vals), int(n * scale))
        tb, ts = run(n)
    t_build, t_s3 = tb, ts
    return {"value": n / t_s3, "unit": "nnz/s", "cores": threads, "kind": "reference",
            "sample": "first %d of %d nonzeros; reference DTree+CompressedDTree build "
                      "%.3f s, one kronecker_product_opt5 %.3f s with %d OpenMP threads"
                      % (n, len(vals), t_build, t_s3, threads),
            "insertion_s": t_build, "s3ttmc_s": t_s3}


def cpu_hooi_estimate(order, dim, nnz, R, s3ttmc_rate, threads):
    """Where the literal CPU HOOI cannot run (its unpacked n x R^(N-1) Y and
    the SVD of it), the strongest CPU form with the same arithmetic: the
    reference's own S3TTMc (opt5) at the rate measured on a bounded sample of
    this tensor, plus a packed-Gram HOOI iteration (P x P Gram, eigh,
    U = Y V, core) timed with numpy's multi-threaded LAPACK on an n x P
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
    t_s3 = nnz / s3ttmc_rate
    return {"kind": "estimate", "cores": threads,
            "literal_infeasible": "the reference HOOI unpacks Y to %d x %d FP64 (%.1f GB) and "
                                  "takes its SVD each iteration" % (dim, R ** (order - 1),
                                                                    dim * R ** (order - 1) * 8e-9),
            "s_per_iter": t_s3 + t_dense, "s3ttmc_s": t_s3, "dense_s": t_dense,
            "sample": "reference opt5 rate %.3e nnz/s from the bounded cpu_baseline sample "
                      "x %d nnz, plus a numpy packed-Gram iteration at n=%d, P=%d"
                      % (s3ttmc_rate, nnz, dim, P), "core_shape": list(core.shape)}


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


def cpu_hoqri_estimate(order, dim, nnz, R, s3ttmc_rate, threads, level_bytes):
    """CPU HOQRI per iteration where the reference's S3TTMc cannot run on the
    host (its opt5 level arrays exceed host memory): the reference's opt5 at
    the rate measured on a bounded prefix of this tensor, plus the dense part
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
    t_s3 = nnz / s3ttmc_rate
    return {"kind": "estimate", "cores": threads,
            "literal_infeasible": "one reference S3TTMc of the whole tensor takes %.0f s on "
                                  "the host cores and materialises %.0f GB of FP64 level arrays "
                                  "(host memory %.0f GB)"
                                  % (nnz / s3ttmc_rate, level_bytes * 1e-9,
                                     host_memory_bytes() * 1e-9),
            "s_per_iter": t_s3 + t_dense, "s3ttmc_s": t_s3, "dense_s": t_dense,
            "sample": "reference opt5 rate %.3e nnz/s from the bounded cpu_baseline sample "
                      "x %d nnz, plus a numpy HOQRI dense step at n=%d, P=%d"
                      % (s3ttmc_rate, nnz, dim, P)}


def cpu_hoqri_baseline(inds, vals, order, dim, R, max_iter, delta, budget_s=40.0):
    """HOQRI on the host cores: the oracle's restatement of TuckerHOQRI
    (opt5 S3TTMc with OpenMP, numpy BLAS/LAPACK for core, Y W core^T and the
    QR), same tensor, rank and random start, bounded like cpu_hooi_baseline."""
    try:
        import oracle
        from oracle import tucker
    except Exception as e:  # pragma: no cover
        return {"error": "oracle unavailable: %s" % e}
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1
    t0 = time.perf_counter()
    T = oracle.Tree(order, dim, inds, vals)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    tucker.hoqri(T, R, 1, rel_tol=0.0, delta=delta, seed=0)
    t_one = time.perf_counter() - t0
    iters = max(1, min(max_iter, int(budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    _, errs, _ = tucker.hoqri(T, R, iters, rel_tol=1e-6, delta=delta, seed=0)
    t_run = time.perf_counter() - t0
    return {"kind": "port", "cores": threads, "construction_s": t_build,
            "iters": len(errs), "seconds": t_run, "s_per_iter": t_run / len(errs),
            "final_error": float(errs[-1]),
            "sample": "oracle TuckerHOQRI restatement (opt5 S3TTMc, OpenMP; numpy QR), "
                      "%d iterations" % len(errs)}


def cpu_hooi_baseline(inds, vals, order, dim, R, max_iter, delta, budget_s=40.0,
                      s3ttmc_rate=None):
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
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1
    if dim * R ** (order - 1) * 8 > 2e9:
        if not s3ttmc_rate:
            return {"error": "literal CPU HOOI infeasible and no CPU S3TTMc rate measured"}
        return cpu_hooi_estimate(order, dim, len(vals), R, s3ttmc_rate, threads)
    t0 = time.perf_counter()
    T = oracle.Tree(order, dim, inds, vals)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, e1, _ = tucker.hooi(T, R, 1, rel_tol=0.0, delta=delta, seed=0)
    t_one = time.perf_counter() - t0
    iters = max(1, min(max_iter, int(budget_s / max(t_one, 1e-3))))
    t0 = time.perf_counter()
    _, errs, _ = tucker.hooi(T, R, iters, rel_tol=1e-6, delta=delta, seed=0)
    t_run = time.perf_counter() - t0
    return {"kind": "port", "cores": threads, "construction_s": t_build,
            "iters": len(errs), "seconds": t_run, "s_per_iter": t_run / len(errs),
            "final_error": float(errs[-1]),
            "sample": "oracle TuckerHOOI restatement (opt5 S3TTMc, %d OpenMP threads; "
                      "numpy LAPACK SVD of the unpacked n x R^(N-1) Y), %d iterations"
                      % (threads, len(errs))}


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
                 "y_times_v", "gemmtn", "gram_u", "qr"):
        ms, cnt = t.timing_get(name)
        if cnt:
            ops[name] = ms / iters
    t.timing(False)
    nt = (P + 63) // 64
    tiles = nt * (nt + 1) // 2
    g_ms = ops.get("gram", 0.0)
    out = {"ms_per_iteration": ops, "P": P}
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


def planted_basis(n, K):
    E = np.zeros((n, K))
    b = n // K
    for c in range(K):
        E[c * b:(c + 1) * b, c] = 1.0
    return np.linalg.qr(E)[0]


def learned_study(dim, nnz, pin, ntrain, ntest, iters, steps):
    """Learned warm start on seeded order-5 planted-partition tensors
    (8 communities, binary values, rank 8, TuckerHOOI with delta 0.5).

    Training targets: the factors plain HOOI converges to on `ntrain` seeds,
    started from the hypergraph sketch (learned.spectral_start).  Held-out
    seeds run HOOI from random, HOEVD, sketch and learned starts until the
    subspace moves by less than 1e-6 (sin of the largest principal angle) or
    `iters` iterations; each start reports iterations, converged fit error,
    angle to the planted communities and time to converge including the cost
    of producing the start."""
    from symprop_amd import learned
    N, K, R, delta = 5, 8, 8, 0.5
    E = planted_basis(dim, K)

    def tensor(seed):
        inds, vals = sp.generate(sp.GEN_PLANTED, N, dim, nnz, pin, K, seed=seed)
        t = sp.SpSymtensor(dim, N, inds, vals)
        t.sync()
        return t

    feats, facs, teacher_its = [], [], []
    t0 = time.perf_counter()
    for s in range(1, ntrain + 1):
        t = tensor(s)
        F = learned.features(t)
        Us, _ = learned.hooi_trace(t, learned.spectral_start(F, R), iters, delta,
                                   stop_tol=1e-6)
        teacher_its.append(learned.converged_at(Us))
        feats.append(F)
        facs.append(Us[-1])
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

    starts = ("random", "hoevd", "sketch", "learned")
    rows = {k: [] for k in starts}
    for s in range(10001, 10001 + ntest):
        t = tensor(s)
        for k in starts:
            t0 = time.perf_counter()
            if k == "random":
                U0 = np.random.default_rng(s).standard_normal((dim, R))
            elif k == "hoevd":
                U0, _ = t.hoevd_init(R)
            else:
                F = learned.features(t)
                U0 = learned.spectral_start(F, R) if k == "sketch" else model.predict(F)
            t_init = time.perf_counter() - t0
            t0 = time.perf_counter()
            Us, errs = learned.hooi_trace(t, U0, iters, delta, stop_tol=1e-6)
            t_run = time.perf_counter() - t0
            kc = learned.converged_at(Us)
            rows[k].append({"seed": s, "converged_at": kc, "iters_run": len(Us),
                            "final_error": float(errs[-1]),
                            "sin_to_planted": learned.sin_angle(Us[-1], E),
                            "init_s": t_init, "seconds": t_init + t_run})
        log("learned seed %d: " % s + ", ".join(
            "%s %s it err %.9f" % (k, rows[k][-1]["converged_at"], rows[k][-1]["final_error"])
            for k in starts))
        del t

    def summary(rs):
        conv = [r["converged_at"] for r in rs if r["converged_at"] is not None]
        return {"converged": "%d/%d" % (len(conv), len(rs)),
                "median_iters_to_1e-6": float(np.median(conv)) if conv else None,
                "mean_final_error": float(np.mean([r["final_error"] for r in rs])),
                "mean_sin_to_planted": float(np.mean([r["sin_to_planted"] for r in rs])),
                "median_seconds": float(np.median([r["seconds"] for r in rs])),
                "runs": rs}

    return {"config": {"order": N, "dim": dim, "nnz": nnz, "p_in": pin, "communities": K,
                       "rank": R, "delta": delta, "max_iters": iters,
                       "train_tensors": ntrain, "held_out": ntest,
                       "criterion": "sin(U_k, U_k-1) < 1e-6"},
            "train": {"teacher": "HOOI from the sketch start",
                      "teacher_converged_at": teacher_its, "teacher_s": t_teacher,
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

    t0 = time.perf_counter()
    t = sp.SpSymtensor(n, N, inds, vals)
    t.sync()
    insertion_s = time.perf_counter() - t0
    # a second build of the same tensor: the first one also pays the
    # process's one-time HIP module loading and allocator warm-up
    t0 = time.perf_counter()
    t2 = sp.SpSymtensor(n, N, inds, vals)
    t2.sync()
    steady_s = time.perf_counter() - t0
    del t2
    # algorithmic bytes of the construction: one read and one write of every
    # sub-multiset key (k indices of 4 bytes, k = 1..N-1) and of every owner
    # record (row index + value, 12 bytes)
    cons_bytes = float(sum(nnz * binom(N, k) * 4 * k * 2 for k in range(1, N)) +
                       2 * N * nnz * 12)
    construction = {"first_s": insertion_s, "steady_s": steady_s, "alg_bytes": cons_bytes,
                    "achieved_gbs": cons_bytes / steady_s / 1e9,
                    "hbm_frac": cons_bytes / steady_s / 1e9 / HBM_PEAK_GBS}
    log("construction %.3f s (steady %.3f s)" % (insertion_s, steady_s))

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
        "construction": construction,
        "roofline": {"bound": "mfma", "roof": "FP64 arithmetic (MI355X FP64 vector = "
                     "matrix peak)", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": traffic, "traffic_unit": "bytes per launch",
                     "traffic_source": traffic_src, "kernel": kname,
                     "kernel_ms": k_avg, "alg_flops": F,
                     "split_reduce_ms": r_ms / max(1, r_cnt)},
    }

    if traffic:
        gbs = traffic / (k_avg * 1e-3) / 1e9
        out["roofline"]["hbm_achieved_gbs"] = gbs
        out["roofline"]["hbm_frac"] = gbs / HBM_PEAK_GBS
    out["dense_y_side"] = dense_y_side(t, N, n, R, w["delta"])

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
                      "final_error": float(r["errors"][-1])}
        out["%s_fixed_iterations" % aname] = hooi_fixed
        # time to fit: every form runs the reference loop's own stopping rule
        # |error - previous| < 1e-6 (spsymtensor.cpp:235-238, 283-286), up to
        # 500 iterations; the mixed forms switch to FP64 sweeps on the way
        runs = {}
        for name, prec in (("fp64", sp.PREC_FP64), ("fp32_mixed", sp.PREC_FP32),
                           ("bf16_mixed", sp.PREC_BF16)):
            try:
                r = t.tucker(alg=alg, rank=R, max_iter=500, rel_tol=1e-6,
                             delta=w["delta"], seed=0, precision=prec)
            except sp.SpError as e:
                runs[name] = {"error": str(e)}
                continue
            runs[name] = {"seconds": r["seconds"], "iters": r["iters"],
                          "fp64_sweeps": r["fp64_sweeps"],
                          "final_error": float(r["errors"][-1])}
            log("%s %s: %.3f s, %d iters, error %.12e" % (aname.upper(), name, r["seconds"],
                                                         r["iters"], r["errors"][-1]))
        if "final_error" in runs.get("fp64", {}):
            e64 = runs["fp64"]["final_error"]
            for k, v in runs.items():
                if "final_error" in v:
                    v["rel_fit_error_vs_fp64"] = abs(v["final_error"] - e64) / e64
        out["%s_time_to_fit" % aname] = runs

    if not args.no_learned:
        del t
        log("learned warm start study")
        out["learned_warm_start"] = learned_study(*learned_args)

    if not args.no_cpu:
        log("cpu baseline (reference build on host cores)")
        out["cpu_baseline"] = cpu_baseline(inds, vals, N, n, R)
        if not args.no_hooi:
            log("cpu %s baseline (host cores)" % aname.upper())
            rate = out["cpu_baseline"].get("value")
            threads = out["cpu_baseline"].get("cores", 1)
            if alg == 0:
                cb = cpu_hooi_baseline(inds, vals, N, n, R, w["max_iter"], w["delta"],
                                       s3ttmc_rate=rate)
            elif rate and (level_bytes > 0.5 * host_memory_bytes() or nnz / rate > 20.0):
                # one reference S3TTMc of the whole tensor does not fit host
                # memory or the CPU time budget
                cb = cpu_hoqri_estimate(N, n, nnz, R, rate, threads, level_bytes)
            else:
                cb = cpu_hoqri_baseline(inds, vals, N, n, R, w["max_iter"], w["delta"])
            out["cpu_%s_baseline" % aname] = cb
            if "s_per_iter" in cb and hooi_fixed:
                cb["gpu_fp64_speedup_per_iteration"] = cb["s_per_iter"] / hooi_fixed["s_per_iter"]
                f64 = out.get("%s_time_to_fit" % aname, {}).get("fp64", {})
                if "iters" in f64:
                    # the CPU loop runs the same iteration sequence (same start,
                    # same stopping rule), so its time to fit is its
                    # per-iteration time times the FP64 iteration count
                    cb["time_to_fit_estimate_s"] = cb["s_per_iter"] * f64["iters"]
                    cb["time_to_fit_estimate_note"] = (
                        "CPU s/iteration x the %d iterations the FP64 run needed" % f64["iters"])
                    for k, v in out["%s_time_to_fit" % aname].items():
                        if "seconds" in v:
                            v["speedup_vs_cpu_time_to_fit"] = (cb["time_to_fit_estimate_s"] /
                                                               v["seconds"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
