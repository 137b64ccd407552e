#!/usr/bin/env python3
"""Benchmark of the pattern-evaluator fit on one MI355X.

Workload (one "step" = one complete fit): 5,000,000 positions of uniformly
random legal play at 24 empties, the 8-pattern set {L02X, B5, Q3, D8, L1, L2,
Ep, D7} (k = 34 entries per row, 139,644 columns), targets b = A x_true + noise
(x_true ~ 1.5 N(0,1), noise ~ 4 N(0,1), rounded to even integers, clipped to
+-64).  A step solves min |A x - b| from x0 = 0 with the mixed-precision
refined LSQR (FP32 products, FP64 recurrences and FP64 refinement residual) to
|A^T(b - A x)| / |A^T b| <= 1e-10, inputs resident in HBM.

Printed: one JSON line with the time-to-fit (value, seconds), the FP64 LSQR
time-to-fit on the same GPU, the SpMV / scatter kernel bandwidth against the
8 TB/s HBM roofline (HIP events on the library's stream), and the CPU
baseline (the oracle's FP64 LSQR on the host cores, a bounded sample of
iterations scaled to the iteration count the fit needs).

Four more workloads ride along in the same line (--legs):
  reference_fit  the reference's own fit: 100k positions at 20 empties,
             {L02X, B5, Q3, D8}, Jacobi-PCG 100 iterations (FitWeights) on the
             GPU, the oracle on the host cores and the compiled reference.
  buckets    50 independent systems, one per empty count 10..59, rows drawn
             log-uniform in [20k, 800k] (seed 1003), pattern set {L02X, B5, Q3,
             D8}; all fitted in the same launches (mixed LSQR to 1e-10 per
             system) against the same 50 fits run one after another.
  evaluator  the differentiable evaluator: 5M positions at 30 empties, pattern
             set {L02X, B5, Q3, D8}, Huber(4) Adam on minibatches of 65,536;
             positions/s of a training epoch and the gather / scatter kernel
             bandwidth of one minibatch.
  matrix_free 20M positions spread over 15..45 empties, 12 patterns (k = 54):
             the mixed LSQR with indices recomputed from the positions in every
             product against the stored-index FP64 LSQR, both to 1e-10.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-fit at rel. residual 1e-10 (mixed+refine vs FP64 LSQR); SpMV HBM GB/s"
HBM_PEAK_GBPS = 8000.0
# 32-bit integer VALU issue of one MI355X: 256 CUs x 4 SIMDs x 16 lanes at
# 2.4 GHz (a wave64 instruction takes a SIMD 4 cycles), in lane operations/s
VALU_PEAK_OPS = 256 * 64 * 2.4e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    """The host CPU: model name, physical cores and logical CPUs of the
    machine, and the CPUs this process may run on."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "cpus_allowed": allowed}


def omp_threads():
    """OpenMP threads of the CPU baseline: OMP_NUM_THREADS, else the CPUs this
    process may run on."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v)
    return cpu_info()["cpus_allowed"]


def event_time_ms(ctx, fn, reps):
    """Average GPU time of fn() (which enqueues on the library stream)."""
    import torch

    s = ctx.torch_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_baseline(A, b, fp64_iterations, budget_s, x_gpu, pos_host=None, masks=None):
    """The oracle's FP64 LSQR (column-scaled, OpenMP over all host threads) on
    the same matrix.  When the iterations the GPU FP64 fit needed fit in the
    budget they all run, the time is measured whole, and the CPU solution is
    the FP64 reference for the GPU weights; otherwise a bounded sample of
    iterations is timed and scaled.  Beside it, timed on the same cores: the
    oracle's index build of every position (GroupIndexer, compared with the
    device matrix entry for entry) and its Jacobi count pass."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    idx = A.download()
    bh = b.cpu().numpy()
    steps = {}
    if pos_host is not None:
        G = O.Group(masks)
        t0 = time.perf_counter()
        ih = G.indices(pos_host)
        steps["index_build_s"] = time.perf_counter() - t0
        steps["index_build_matches_gpu"] = bool(np.array_equal(ih, idx))
        del ih
    t0 = time.perf_counter()
    d = O.jacobi(idx, A.cols).astype(np.float64)
    steps["jacobi_s"] = time.perf_counter() - t0
    ok = np.isfinite(d) & (d > 0)
    D = np.where(ok, 1.0 / np.sqrt(np.rint(1.0 / np.where(ok, d, 1.0))), 1.0)
    t0 = time.perf_counter()
    O.lsqr(idx, A.cols, bh, 2, colscale=D)
    per = (time.perf_counter() - t0) / 2
    cores = omp_threads()
    out = {"unit": "s", "cores": cores, "kind": "port", **cpu_info(), **steps}
    if per * fp64_iterations <= budget_s:
        t0 = time.perf_counter()
        xc, _, _ = O.lsqr(idx, A.cols, bh, fp64_iterations, colscale=D)
        secs = time.perf_counter() - t0
        rel = O.rel_normal_residual(idx, A.cols, bh, xc)[0]
        xg = x_gpu.cpu().numpy()
        pc, pg = O.spmv(idx, xc), O.spmv(idx, xg)
        out.update(value=secs, sample=f"the whole fit: {fp64_iterations} iterations of the oracle's column-scaled "
                                      f"FP64 LSQR on the full {A.rows}x{A.cols} matrix",
                   rel_normal_residual=rel,
                   gpu_mixed_weight_error=float(np.linalg.norm(xg - xc) / np.linalg.norm(xc)),
                   gpu_mixed_prediction_error=float(np.linalg.norm(pg - pc) / np.linalg.norm(pc)))
    else:
        n = int(max(2, min(200, budget_s / max(per, 1e-6))))
        t0 = time.perf_counter()
        O.lsqr(idx, A.cols, bh, n, colscale=D)
        per = (time.perf_counter() - t0) / n
        out.update(value=per * fp64_iterations,
                   sample=f"{n} iterations of the oracle's column-scaled FP64 LSQR on the full {A.rows}x{A.cols} "
                          f"matrix ({per:.3f} s/iteration), scaled to the {fp64_iterations} iterations the FP64 "
                          f"fit needs to reach 1e-10")
    R = O.ref()
    if R is not None:
        m = R.ref_csr_create(idx.ctypes.data_as(O.u32p), A.rows, A.k, A.cols)
        try:
            x = np.zeros(A.cols, np.float32)
            bf = bh.astype(np.float32)
            t0 = time.perf_counter()
            R.ref_lsqr(m, bf.ctypes.data_as(O.f32p), A.rows, 3, x.ctypes.data_as(O.f32p), None)
            out["reference_fp32_lsqr_s_per_iteration"] = (time.perf_counter() - t0) / 3
        finally:
            R.ref_csr_free(m)
    return out


def leg_reference_fit(ctx):
    """The reference's own fit: 100,000 positions at 20 empties, {L02X, B5, Q3,
    D8}, Jacobi-PCG on the normal equations for 100 iterations (FitWeights,
    PatternFit/PatternFit.cpp:7-39) -- on the GPU in FP64 and FP32, the
    oracle's FP64 restatement on the host cores, and the compiled reference
    (float, its own OpenMP loops) where it was built."""
    import cassandra_amd.capi as capi
    from cassandra_amd import names as N
    from cassandra_amd.fit import Patterns

    P = Patterns(ctx, N.FIT4)
    pos = ctx.random_play(1001, 100_000, 20)
    A = P.matrix(pos)
    _, b = A.make_targets(2001, 3001)
    out = {"rows": A.rows, "k": A.k, "cols": A.cols, "iterations": 100}
    for name, prec in (("gpu_fp64", capi.PREC_FP64), ("gpu_fp32", capi.PREC_FP32)):
        opts = dict(solver=capi.SOLVER_PCG, precision=prec, tol=0.0, max_iterations=100, check_every=100)
        A.solve(b, **opts)
        ctx.synchronize()
        t0 = time.perf_counter()
        x, _ = A.solve(b, **opts)
        ctx.synchronize()
        out[name + "_seconds"] = time.perf_counter() - t0
        if prec == capi.PREC_FP64:
            x64 = x
    out.update(cpu_baseline_pcg(A.download(), A.cols, b.cpu().numpy(), x64.cpu().numpy(), 100))
    log(f"reference fit: {out}")
    return out


def cpu_baseline_pcg(idx, cols, bh, x_gpu, iterations):
    """CPU baseline of the reference fit: the oracle's FP64 PCG and the
    compiled reference's float FitWeights PCG, timed on the host cores; the
    oracle's iterate is also the FP64 check of the GPU's."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    out = {"cores": omp_threads()}
    t0 = time.perf_counter()
    xo, _ = O.pcg(idx, cols, bh, iterations)
    out["cpu_oracle_fp64_seconds"] = time.perf_counter() - t0
    po = O.spmv(idx, xo)
    out["gpu_fp64_prediction_error"] = float(np.linalg.norm(O.spmv(idx, x_gpu) - po) / np.linalg.norm(po))
    R = O.ref()
    if R is not None:
        m = R.ref_csr_create(idx.ctypes.data_as(O.u32p), idx.shape[0], idx.shape[1], cols)
        try:
            xr = np.zeros(cols, np.float32)
            bf = bh.astype(np.float32)
            t0 = time.perf_counter()
            R.ref_fit_pcg(m, bf.ctypes.data_as(O.f32p), idx.shape[0], iterations, xr.ctypes.data_as(O.f32p), None)
            out["cpu_reference_fp32_seconds"] = time.perf_counter() - t0
            out["reference_prediction_error"] = float(np.linalg.norm(O.spmv(idx, xr.astype(np.float64)) - po)
                                                      / np.linalg.norm(po))
        finally:
            R.ref_csr_free(m)
    return out


def leg_buckets(ctx):
    """50 empty-count systems in one batch against 50 separate fits."""
    import numpy as np
    import torch

    import cassandra_amd.capi as capi
    from cassandra_amd import names as N
    from cassandra_amd.fit import Patterns

    rng = np.random.default_rng(1003)
    empties = list(range(10, 60))
    rows = np.rint(np.exp(rng.uniform(np.log(20000), np.log(800000), size=len(empties)))).astype(np.int64)
    P = Patterns(ctx, N.FIT4)
    parts = [ctx.random_play(1003 + 7919 * i, int(n), e) for i, (e, n) in enumerate(zip(empties, rows))]
    rb = np.concatenate([[0], np.cumsum(rows)])
    A = P.matrix_systems(torch.cat(parts), rb)
    _, b = A.make_targets(2003, 3003)
    log(f"buckets: {A.nsys} systems, {A.rows} rows, k={A.k}, {A.cols} columns")
    # the small buckets (fewer rows than columns) are ill-conditioned and need
    # thousands of iterations; both runs get the same generous budget
    opts = dict(precision=capi.PREC_MIXED, tol=1e-10, max_iterations=20000)
    A.solve_systems(b, **dict(opts, max_iterations=50))  # warm: workspace, graph
    ctx.synchronize()
    t0 = time.perf_counter()
    x, rep, rel, its = A.solve_systems(b, **opts)
    ctx.synchronize()
    t_batch = time.perf_counter() - t0
    log(f"buckets batched: {t_batch:.4f} s, {rep}, {int(np.sum(rel <= 1e-10))} buckets at 1e-10")
    off_tol = [{"system": int(i), "empties": empties[i], "rows": int(rows[i]), "rel": float(rel[i]),
                "iterations": int(its[i])} for i in np.flatnonzero(rel > 1e-10)]
    if off_tol:
        log(f"buckets above 1e-10: {off_tol}")
    # the same fits one system at a time (each warmed first)
    mats = [P.matrix(p_) for p_ in parts]
    del parts
    bs = [b[int(rb[i]):int(rb[i + 1])].contiguous() for i in range(len(mats))]
    for M, bi in zip(mats, bs):
        M.solve(bi, **dict(opts, max_iterations=50))
    ctx.synchronize()
    log("buckets: one-by-one warm done")
    t0 = time.perf_counter()
    seq_rel = []
    for i, (M, bi) in enumerate(zip(mats, bs)):
        _, r = M.solve(bi, **opts)
        seq_rel.append(r["rel_normal_residual"])
        if i % 10 == 9:
            log(f"buckets one by one: {i + 1} done")
    ctx.synchronize()
    t_seq = time.perf_counter() - t0
    seq_rel = np.array(seq_rel)
    log(f"buckets one by one: {t_seq:.4f} s, {int(np.sum(seq_rel <= 1e-10))} buckets at 1e-10")
    return {"systems": int(A.nsys), "rows": int(A.rows), "k": int(A.k), "cols_per_system": int(A.cols_per_system),
            "max_iterations": opts["max_iterations"], "seconds_batched": t_batch, "seconds_one_by_one": t_seq,
            "speedup": t_seq / t_batch, "iterations": rep["iterations"],
            "buckets_at_tol": int(np.sum(rel <= 1e-10)), "max_rel_normal_residual": float(np.max(rel)),
            "median_rel_normal_residual": float(np.median(rel)), "median_iterations": float(np.median(its)),
            "buckets_at_tol_one_by_one": int(np.sum(seq_rel <= 1e-10)),
            "max_rel_one_by_one": float(np.max(seq_rel)), "above_tol": off_tol}


def mf_valu_ops(P):
    """Algorithmic 32-bit VALU lane operations per row of the matrix-free
    products: a bit-field extract and a multiply-add per square of every
    slot's pattern variation (its index, sum_j 3^j digit_j), plus, per slot,
    the column's LDS address and the accumulation (A x) or the atomic's
    address (A^T y, whose LDS atomic is not a VALU operation)."""
    squares = sum(bin(int(v)).count("1") for v in P.variations())
    return {"squares_per_row": squares, "spmv": 2 * squares + 2 * P.k, "spmtv": 2 * squares + P.k}


def mf_launches(ctx, F):
    """Three launches each of the matrix-free FP32 A x and A^T y (for counter
    runs: SQ_INSTS_VALU and the like per dispatch)."""
    import ctypes as C_

    import torch

    import cassandra_amd.capi as capi

    L = capi.lib()
    hp = lambda t_: C_.c_void_p(t_.data_ptr())  # noqa: E731
    x32 = torch.randn(F.cols, dtype=torch.float32, device="cuda")
    y32 = torch.randn(F.rows, dtype=torch.float32, device="cuda")
    z32 = torch.empty(F.cols, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        capi.check(L.cad_spmv_f32(F.h, hp(x32), hp(y32)))
    for _ in range(3):
        capi.check(L.cad_spmtv_f32(F.h, hp(y32), hp(z32)))
    ctx.synchronize()
    log(f"matrix-free launches done: A x and A^T y on {F.rows}x{F.cols} (k={F.k})")


def leg_matrix_free(ctx, traffic=False):
    """20M positions over 15..45 empties, 12 patterns: the matrix-free mixed
    LSQR (indices recomputed in every product) against the stored-index FP64
    LSQR on the same GPU, and the products against the VALU issue rate (with
    traffic set: only launch the products, for counter runs)."""
    import ctypes as C_

    import torch

    import cassandra_amd.capi as capi
    from cassandra_amd import names as N
    from cassandra_amd.fit import Patterns

    n = 20_000_000
    P = Patterns(ctx, N.FIT12)
    pos = ctx.random_play(1005, n, 15, 45)
    F = P.matrix_free(pos)
    if traffic:
        mf_launches(ctx, F)
        return None
    _, b = F.make_targets(2005, 3005)
    log(f"matrix-free: {F.rows} rows, k={F.k}, cols={F.cols}")
    F.solve(b, precision=capi.PREC_MIXED, tol=1e-10)  # warm
    ctx.synchronize()
    t0 = time.perf_counter()
    _, rep_f = F.solve(b, precision=capi.PREC_MIXED, tol=1e-10)
    ctx.synchronize()
    t_mf = time.perf_counter() - t0
    log(f"matrix-free mixed: {t_mf:.3f} s {rep_f}")
    L = capi.lib()
    hp = lambda t_: C_.c_void_p(t_.data_ptr())  # noqa: E731
    x32 = torch.randn(F.cols, dtype=torch.float32, device="cuda")
    y32 = torch.empty(F.rows, dtype=torch.float32, device="cuda")
    z32 = torch.empty(F.cols, dtype=torch.float32, device="cuda")
    mf_spmv = event_time_ms(ctx, lambda: L.cad_spmv_f32(F.h, hp(x32), hp(y32)), 10)
    mf_spmtv = event_time_ms(ctx, lambda: L.cad_spmtv_f32(F.h, hp(y32), hp(z32)), 10)
    del F
    S = P.matrix(pos)
    del pos
    S.solve(b, precision=capi.PREC_FP64, tol=1e-10)  # warm
    ctx.synchronize()
    t0 = time.perf_counter()
    _, rep_s = S.solve(b, precision=capi.PREC_FP64, tol=1e-10)
    ctx.synchronize()
    t_s = time.perf_counter() - t0
    log(f"stored fp64: {t_s:.3f} s {rep_s}")
    st_spmv = event_time_ms(ctx, lambda: L.cad_spmv_f32(S.h, hp(x32), hp(y32)), 10)
    csr_u32 = S.rows * (4 * S.k + 4) + 4 * S.cols
    ops = mf_valu_ops(P)
    valu = {}
    for name, ms in (("spmv", mf_spmv), ("spmtv", mf_spmtv)):
        achieved = ops[name] * S.rows / (ms * 1e-3)
        valu[name] = {"bound": "valu", "ops_per_row": ops[name], "achieved": achieved / 1e12,
                      "peak": VALU_PEAK_OPS / 1e12, "unit": "T lane-ops/s", "frac": achieved / VALU_PEAK_OPS}
    log(f"matrix-free VALU roofline ({ops['squares_per_row']} squares a row): {valu}")
    # what a row keeps resident in HBM: its indices (or its position and
    # ternary words), plus the solver's row vectors (u in FP32, b and r in
    # FP64) -- the capacity the matrix-free matrix trades its speed for
    row_vectors = 4 + 8 + 8
    resident = {"stored": S.index_bytes * S.k + row_vectors, "matrix_free": 16 + 16 + row_vectors}
    return {"rows": S.rows, "k": S.k, "cols": S.cols, "seconds_matrix_free_mixed": t_mf,
            "resident_bytes_per_row": resident,
            "rows_per_288GB": {key: int(288e9 // val) for key, val in resident.items()},
            "iterations_matrix_free": rep_f["iterations"], "rel_matrix_free": rep_f["rel_normal_residual"],
            "seconds_stored_fp64": t_s, "iterations_stored_fp64": rep_s["iterations"],
            "rel_stored_fp64": rep_s["rel_normal_residual"], "speedup": t_s / t_mf,
            "mf_spmv_us": mf_spmv * 1e3, "mf_spmtv_us": mf_spmtv * 1e3, "stored_spmv_us": st_spmv * 1e3,
            "mf_spmv_equiv_u32_csr_GBps": csr_u32 / (mf_spmv * 1e-3) / 1e9,
            "index_bytes_stored": S.index_bytes, "squares_per_row": ops["squares_per_row"],
            "valu_roofline": valu}


def leg_evaluator(ctx):
    """The differentiable evaluator: Huber/Adam training throughput, the
    forward gather and backward scatter kernels of one minibatch (HIP events
    around the C-ABI calls alone), the resident evaluator's batch latency, and
    the Huber fit against the least-squares fit of the same data."""
    import ctypes as C_

    import torch

    import cassandra_amd.capi as capi
    from cassandra_amd import names as N
    from cassandra_amd.evaluator import huber_fit, huber_fit_fused
    from cassandra_amd.fit import Patterns

    P = Patterns(ctx, N.FIT4)
    n, B = 5_000_000, 65536
    pos = ctx.random_play(1004, n, 30)
    A = P.matrix(pos)
    xt, b = A.make_targets(2004, 3004)
    truth = A.spmv(xt)  # the noise-free targets A x_true
    idx = P.extract(pos, capi.LAYOUT_ROW_MAJOR)
    target = b.float()
    huber_fit(P, idx[: 4 * B], target[: 4 * B], epochs=1, batch=B)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, losses = huber_fit(P, idx, target, epochs=1, batch=B)
    torch.cuda.synchronize()
    t_epoch = time.perf_counter() - t0
    log(f"evaluator: epoch {t_epoch:.3f} s, loss {losses}")
    # the same epoch as one library call (per step: the forward with the loss
    # and its gradient reading the minibatch's rows through the order, the
    # backward, Adam)
    huber_fit_fused(P, idx[: 4 * B], target[: 4 * B], epochs=1, batch=B)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, losses_f = huber_fit_fused(P, idx, target, epochs=1, batch=B)
    torch.cuda.synchronize()
    t_epoch_f = time.perf_counter() - t0
    log(f"evaluator: fused epoch {t_epoch_f:.4f} s, loss {losses_f}")

    # one minibatch through the C-ABI on the library's stream, kernels alone
    L = capi.lib()
    hp = lambda t_: C_.c_void_p(t_.data_ptr())  # noqa: E731
    w = torch.randn(P.cols, dtype=torch.float32, device="cuda")
    sub = idx[:B].contiguous()
    g = torch.randn(B, dtype=torch.float32, device="cuda")
    out = torch.empty(B, dtype=torch.float32, device="cuda")
    grad = torch.empty(P.cols, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    reps = 50
    fwd_ms = event_time_ms(ctx, lambda: L.cad_eval_forward_f32(P.h, hp(sub), B, hp(w), hp(out)), reps)
    bwd_ms = event_time_ms(ctx, lambda: L.cad_eval_backward_f32(P.h, hp(sub), B, hp(g), hp(grad)), reps)
    bwd_na_ms = event_time_ms(ctx, lambda: L.cad_eval_backward_f32_noalloc(P.h, hp(sub), B, hp(g), hp(grad)), reps)
    nbytes = B * (4 * P.k + 4) + 4 * P.cols
    # the resident evaluator (expanded tables kept on the device): positions
    # in, scores out -- on device positions, and host batches with the copies
    E = P.evaluator(w)
    pos_b = pos[:B].contiguous()
    res_ms = event_time_ms(ctx, lambda: L.cad_evaluator_eval(E.h, hp(pos_b), B, hp(out)), reps)
    host_pos = pos_b.cpu().numpy().view("uint64")
    host_lat = {}
    for nb in (64, 4096, B):
        E.eval_host(host_pos[:nb])
        t0 = time.perf_counter()
        for _ in range(20):
            E.eval_host(host_pos[:nb])
        host_lat[str(nb)] = (time.perf_counter() - t0) / 20 * 1e6
    del E, pos_b

    # the Huber(4) fit of several epochs against the least-squares fit (mixed
    # LSQR to 1e-10) of the same rows: errors against the targets and against
    # the noise-free A x_true
    epochs = 8
    model, hl = huber_fit(P, idx, target, epochs=epochs, batch=B)
    with torch.no_grad():
        p_h = model(idx).double()
    x_ls, rep_ls = A.solve(b, precision=capi.PREC_MIXED, tol=1e-10)
    p_ls = A.spmv(x_ls)

    def quality(p):
        r = b - p
        e = truth - p
        hub = torch.nn.functional.huber_loss(p, b, delta=4.0)
        return {"rmse_vs_targets": float(torch.sqrt(torch.mean(r * r))),
                "mae_vs_targets": float(torch.mean(torch.abs(r))), "huber4_vs_targets": float(hub),
                "rmse_vs_noise_free": float(torch.sqrt(torch.mean(e * e)))}

    fit_cmp = {"huber_adam": dict(quality(p_h), epochs=epochs, lr=0.05, minibatch=B, final_epoch_loss=hl[-1]),
               "least_squares": dict(quality(p_ls), solver="mixed LSQR", iterations=rep_ls["iterations"],
                                     rel_normal_residual=rep_ls["rel_normal_residual"])}
    log(f"evaluator fits: {fit_cmp}")
    return {"positions": n, "minibatch": B, "epoch_seconds": t_epoch, "positions_per_s": n / t_epoch,
            "epoch_seconds_fused_step": t_epoch_f, "positions_per_s_fused_step": n / t_epoch_f,
            "loss_epoch_fused_step": losses_f[0],
            "loss_epoch": losses[0], "bytes_per_minibatch": nbytes,
            "kernel_us": {"forward": fwd_ms * 1e3, "backward": bwd_ms * 1e3, "backward_noalloc": bwd_na_ms * 1e3,
                          "resident_eval_positions": res_ms * 1e3},
            "forward_GBps": nbytes / (fwd_ms * 1e-3) / 1e9, "backward_GBps": nbytes / (bwd_ms * 1e-3) / 1e9,
            "resident_eval_positions_per_s": B / (res_ms * 1e-3),
            "resident_eval_host_batch_latency_us": host_lat,
            "fit_comparison": fit_cmp}


def traffic_launches(ctx, A):
    """Three launches each of the FP32 A x and A^T y products, then three
    elementwise copies of 1 GiB (the scale for the FETCH_SIZE / WRITE_SIZE
    counters)."""
    import ctypes as C_

    import torch

    import cassandra_amd.capi as capi

    L = capi.lib()
    x32 = torch.randn(A.cols, dtype=torch.float32, device="cuda")
    y32 = torch.empty(A.rows, dtype=torch.float32, device="cuda")
    u32 = torch.randn(A.rows, dtype=torch.float32, device="cuda")
    z32 = torch.empty(A.cols, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    hp = lambda t_: C_.c_void_p(t_.data_ptr())  # noqa: E731
    for _ in range(3):
        capi.check(L.cad_spmv_f32(A.h, hp(x32), hp(y32)))
    for _ in range(3):
        capi.check(L.cad_spmtv_f32(A.h, hp(u32), hp(z32)))
    ctx.synchronize()
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)  # an elementwise kernel: reads and writes exactly 1 GiB
    torch.cuda.synchronize()
    log(f"traffic launches done: A x and A^T y on {A.rows}x{A.cols} (k={A.k}, {A.index_bytes} B/index), "
        f"copy of {src.numel() * 4} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=5_000_000)
    ap.add_argument("--empties", type=int, default=24)
    ap.add_argument("--reps", type=int, default=50, help="launches per kernel timing")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0, help="seconds of CPU fit before sampling")
    ap.add_argument("--legs", default="reference_fit,buckets,evaluator,matrix_free",
                    help="extra workloads: reference_fit,buckets,evaluator,matrix_free (or empty)")
    ap.add_argument("--traffic", action="store_true",
                    help="only launch the measured kernels a few times and a 1 GiB device copy (for counter runs; "
                         "with matrix_free among the legs: the matrix-free products instead)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the default product schedules instead of timing the candidates")
    ap.add_argument("--schedule", default="",
                    help="pin product schedule fields after the autotuner, e.g. scatter_order=1,scatter_rpt=4")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the fit is built and measured "
              f"on a single MI355X; multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cassandra_amd.capi as capi
    from cassandra_amd import names as N
    from cassandra_amd.fit import Context, Patterns

    ctx = Context(0)
    legs = [x for x in args.legs.split(",") if x]
    if args.traffic and "matrix_free" in legs:
        leg_matrix_free(ctx, traffic=True)
        ctx.close()
        return
    t = time.perf_counter()
    pos = ctx.random_play(1002, args.rows, args.empties)
    P = Patterns(ctx, N.FIT8)
    A = P.matrix(pos)
    xt, b = A.make_targets(2002, 3002)
    log(f"data: {A.rows} rows, k={A.k}, cols={A.cols}, index bytes {A.index_bytes} ({time.perf_counter() - t:.1f} s)")
    # index extraction (GroupIndexer::InsertIndices over every row,
    # PatternFit/PatternFit.cpp:19-21): positions -> u32 slot-major indices
    extract = None
    if not args.traffic:
        L = capi.lib()
        out_idx = torch.empty(A.rows * A.k, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        ex_ms = event_time_ms(ctx, lambda: L.cad_index_extract(P.h, C.c_void_p(pos.data_ptr()), A.rows,
                                                             C.c_void_p(out_idx.data_ptr()),
                                                             capi.LAYOUT_SLOT_MAJOR), 10)
        ex_bytes = A.rows * (16 + 4 * A.k)
        extract = {"kernel": "k_index_extract (u32 slot-major)", "launch_us": ex_ms * 1e3,
                   "algorithmic_bytes": ex_bytes, "achieved_GBps": ex_bytes / (ex_ms * 1e-3) / 1e9,
                   "frac": ex_bytes / (ex_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS}
        log(f"index extraction: {extract}")
        del out_idx
    pos_host = None if (args.traffic or args.no_cpu) else pos.cpu().numpy()
    del pos
    t = time.perf_counter()
    tune_log = A.autotune() if args.autotune else ""
    log(f"autotune ({time.perf_counter() - t:.2f} s):\n{tune_log}")
    if args.schedule:
        knobs = {}
        for kv in args.schedule.split(","):
            key, val = kv.split("=")
            knobs[key.strip()] = float(val) if "." in val else int(val)
        A.set_schedule(**knobs)
        tune_log = (tune_log or "") + f"pinned: {A.schedule()}\n"
        log(f"schedule pinned: {A.schedule()}")
    if args.traffic:
        traffic_launches(ctx, A)
        ctx.close()
        return
    opts = dict(precision=capi.PREC_MIXED, tol=1e-10, verbose=int(args.verbose))
    for i in range(args.warmup):
        _, rep = A.solve(b, **opts)
        log(f"warmup {i}: {rep}")
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    reps = []
    for i in range(args.steps):
        x, rep = A.solve(b, **opts)
        reps.append(rep)
    ctx.synchronize()
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / max(1, args.steps)
    log(f"mixed: {t_mixed:.4f} s/fit, last {reps[-1]}")
    if not all(r["converged"] for r in reps):
        log("WARNING: a mixed fit did not converge")

    fp64 = None
    if not args.no_fp64:
        A.solve(b, precision=capi.PREC_FP64, tol=1e-10)  # warm workspace and graph
        ctx.synchronize()
        t0 = time.perf_counter()
        _, rep64 = A.solve(b, precision=capi.PREC_FP64, tol=1e-10, verbose=int(args.verbose))
        ctx.synchronize()
        fp64 = dict(rep64, seconds=time.perf_counter() - t0)
        log(f"fp64: {fp64}")

    # kernel bandwidth (HIP events on the library stream)
    x32 = torch.randn(A.cols, dtype=torch.float32, device="cuda")
    y32 = torch.empty(A.rows, dtype=torch.float32, device="cuda")
    u32 = torch.randn(A.rows, dtype=torch.float32, device="cuda")
    z32 = torch.empty(A.cols, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    L = capi.lib()
    hp = lambda t_: C.c_void_p(t_.data_ptr())  # noqa: E731
    spmv_ms = event_time_ms(ctx, lambda: L.cad_spmv_f32(A.h, hp(x32), hp(y32)), args.reps)
    spmtv_ms = event_time_ms(ctx, lambda: L.cad_spmtv_f32(A.h, hp(u32), hp(z32)), args.reps)
    ib = A.index_bytes
    bytes_pass = A.rows * (ib * A.k + 4) + 4 * A.cols  # index stream + row vector + column vector
    bytes_u32 = A.rows * (4 * A.k + 4) + 4 * A.cols
    spmv_gbps = bytes_pass / (spmv_ms * 1e-3) / 1e9
    spmtv_gbps = bytes_pass / (spmtv_ms * 1e-3) / 1e9
    log(f"spmv {spmv_ms * 1e3:.1f} us {spmv_gbps:.0f} GB/s; spmtv {spmtv_ms * 1e3:.1f} us {spmtv_gbps:.0f} GB/s")
    # Each product kernel timed on its own (HIP events on the library's
    # stream): the roofline is the slower of the A x gather and the A^T y
    # window tiles, each charged the whole pass's algorithmic bytes (index
    # stream, row vector, column vector; the tiles' partial slices and the
    # reduction are overhead, not algorithm).
    kt = A.kernel_times(args.reps)
    log(f"kernel times (us): {kt}")
    # the Jacobi count pass (JacobiPreconditionerOfATA): the index stream once
    jac_bytes = A.rows * ib * A.k + 4 * A.cols
    jacobi = {"kernel": "k_jacobi_part", "launch_us": kt["jacobi_part"], "algorithmic_bytes": jac_bytes,
              "achieved_GBps": jac_bytes / (kt["jacobi_part"] * 1e-6) / 1e9,
              "frac": jac_bytes / (kt["jacobi_part"] * 1e-6) / 1e9 / HBM_PEAK_GBPS}
    # the measured ceiling: a device-to-device copy of 1 GiB (torch's copy
    # kernel on the library's stream), read + write bytes over its time
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    with torch.cuda.stream(ctx.torch_stream()):
        copy_ms = event_time_ms(ctx, lambda: dst.copy_(src), 10)
    copy_gbps = 2 * src.numel() * 4 / (copy_ms * 1e-3) / 1e9
    del src, dst
    log(f"device copy of 1 GiB: {copy_ms * 1e3:.0f} us, {copy_gbps:.0f} GB/s")
    cands = [("k_gather_tile (A x)", kt["gather"]), ("k_scatter_tile (A^T y windows)", kt["scatter_tiles"])]
    dom_name, dom_us = max(cands, key=lambda c: c[1])
    dom_gbps = bytes_pass / (dom_us * 1e-6) / 1e9
    # HBM bytes per launch of that kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes over `bench.py --traffic` (scaled by an
    # elementwise copy of 1 GiB; see profiles/traffic.json)
    # (the scatter's traffic depends on its work order: measured for both)
    # The figure is used only when it was measured on the kernel source this
    # run uses (the SHA-256 of csrc/kernels.hpp recorded beside it); a
    # kernel changed since reports null rather than a stale count.
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        try:
            import hashlib

            tj = json.load(open(tpath))
            src = os.path.join(ROOT, "cassandra_amd", "csrc", "kernels.hpp")
            sha = hashlib.sha256(open(src, "rb").read()).hexdigest()
            key = dom_name.split()[0]
            if key == "k_scatter_tile":
                key = f"{key}/order{A.schedule()['scatter_order']}"
            if tj.get("kernels_sha256") == sha:
                traffic = tj.get(key, {}).get("total_bytes")
            else:
                log("traffic: profiles/traffic.json was measured on other kernel source; reporting null")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu and fp64 is not None:
        cpu = cpu_baseline(A, b, fp64["iterations"], args.cpu_budget, x, pos_host, N.FIT8)
        log(f"cpu: {cpu}")
    shape = dict(rows=A.rows, k=A.k, cols=A.cols)
    del A, b, x
    legs = set(args.legs.split(",")) if args.legs else set()
    reference_fit = leg_reference_fit(ctx) if "reference_fit" in legs else None
    buckets = leg_buckets(ctx) if "buckets" in legs else None
    evaluator = leg_evaluator(ctx) if "evaluator" in legs else None
    matrix_free = leg_matrix_free(ctx) if "matrix_free" in legs else None

    out = {
        "metric": METRIC, "value": t_mixed, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 products, fp64 recurrences", "data": "synthetic (random legal play, seeded)",
        "config": {"workload": f"LSQR fit, {shape['rows']} random-play positions at {args.empties} empties, "
                               f"8 patterns (k={shape['k']}, cols={shape['cols']}), tol 1e-10", **shape,
                   "parallelism": "single GPU"},
        "mixed": {"iterations": reps[-1]["iterations"], "refinements": reps[-1]["refinements"],
                  "rel_normal_residual": reps[-1]["rel_normal_residual"]},
        "fp64_lsqr": None if fp64 is None else {"seconds": fp64["seconds"], "iterations": fp64["iterations"],
                                                 "rel_normal_residual": fp64["rel_normal_residual"]},
        "speedup_vs_fp64": None if fp64 is None else fp64["seconds"] / t_mixed,
        "index_extract": extract,
        "jacobi": jacobi,
        "kernels": {"spmv_us": spmv_ms * 1e3, "spmv_GBps": spmv_gbps, "spmtv_us": spmtv_ms * 1e3,
                    "spmtv_GBps": spmtv_gbps, "bytes_per_pass": bytes_pass, "index_bytes": ib,
                    "kernel_us": kt, "device_copy_GBps": copy_gbps,
                    # the same pass in the reference's u32 CSR layout, over the same time
                    "spmv_equiv_u32_GBps": bytes_u32 / (spmv_ms * 1e-3) / 1e9,
                    "spmtv_equiv_u32_GBps": bytes_u32 / (spmtv_ms * 1e-3) / 1e9},
        "roofline": {"kernel": dom_name, "launch_us": dom_us, "bound": "hbm", "achieved": dom_gbps,
                     "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": dom_gbps / HBM_PEAK_GBPS, "traffic": traffic,
                     "measured_copy_GBps": copy_gbps},
        "schedule": tune_log.strip().splitlines()[-1] if tune_log else "default",
        "cpu_baseline": cpu,
        "reference_fit": reference_fit,
        "buckets": buckets,
        "evaluator": evaluator,
        "matrix_free": matrix_free,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
