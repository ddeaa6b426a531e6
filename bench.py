#!/usr/bin/env python3
"""Benchmark of the batched LU engine on one MI355X.

Metric: solves/s for n=256, batch 4096, 8 right-hand sides, reaching FP64
backward error with FP32-MFMA factorisation + FP64 iterative refinement
(rocsolver_amd_dsgesv_strided_batched), compared with plain FP64 dgesv on the
same GPU.  A step is one batched solve of the whole batch with A and B
resident in HBM.  Inputs: A = U diag(s) V^T (U, V orthogonal from QR of
N(0,1) matrices, s log-spaced in [1, kappa], kappa log-uniform in
[1e2, 1e6] per matrix), B ~ N(0,1), seeded, generated on the GPU.

Prints one JSON line (rank 0).  Extra fields: plain FP64 dgesv throughput
(this library, and the image's rocSOLVER when loadable), accuracy of the
timed solves, the dominant kernel's roofline and the host-LAPACK CPU baseline.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N, BC, NRHS = 256, 4096, 8
FP32_MFMA_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
HBM_PEAK_TBS = 8.0  # MI355X HBM3E peak
EPS64 = 2.0 ** -52


def lu_flops(n: int) -> float:
    """LAPACK getrf operation count (2n^3/3 - n^2/2 - n/6), the reference's
    GFLOP convention (scripts/perf/postprocess.py:14-18)."""
    return 2.0 * n ** 3 / 3.0 - n ** 2 / 2.0 - n / 6.0


def gesv_flops(n: int, nrhs: int) -> float:
    return lu_flops(n) + nrhs * (2.0 * n * n - n)


def event_time(torch, fn, reps: int, stream=None) -> float:
    """Mean duration (s) of fn() measured with HIP events on the stream."""
    s = stream or torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    total = 0.0
    for _ in range(reps):
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        total += e0.elapsed_time(e1) * 1e-3
    return total / reps


def cpu_baseline(torch, A, B, X_gpu, budget_s: float):
    """Host LAPACK dgesv_ over a bounded sample of the same batch (the
    reference clients' CPU path, cpu_gesv -> dgesv_), single thread and all
    host cores; also the forward error of the GPU solutions on that sample."""
    import numpy as np

    from oracle import lapack
    if not lapack.available():
        return None
    lapack.set_num_threads(1)
    sample = min(BC, 1024)
    Ah = A[:sample].cpu().numpy()
    Bh = B[:sample].cpu().numpy()
    Xg = X_gpu[:sample].cpu().numpy()
    _, _, desc = lapack.load()

    def solve_range(lo, hi, out):
        ipiv = np.zeros(N, dtype=np.int32)
        for i in range(lo, hi):
            a = Ah[i].copy()
            x = Bh[i].copy()
            lapack.gesv_one(N, NRHS, a.ctypes.data, N, ipiv.ctypes.data, x.ctypes.data, N)
            out[i] = x

    Xc = np.zeros_like(Bh)
    # single thread, until the budget or the sample is spent
    t0 = time.perf_counter()
    done = 0
    while done < sample and time.perf_counter() - t0 < budget_s:
        hi = min(sample, done + 16)
        solve_range(done, hi, Xc)
        done = hi
    t1 = time.perf_counter()
    rate1 = done / (t1 - t0)
    cores = len(os.sched_getaffinity(0))
    threads = min(cores, 64)
    chunks = [(i, min(sample, i + 8)) for i in range(0, sample, 8)]
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda c: solve_range(c[0], c[1], Xc), chunks))
    t1 = time.perf_counter()
    rate_all = sample / (t1 - t0)
    fe = np.abs(Xg - Xc).reshape(sample, -1).max(axis=1) / np.abs(Xc).reshape(sample, -1).max(axis=1)
    return {
        "value": round(rate_all, 1), "unit": "solves/s", "cores": threads, "kind": "reference",
        "sample": (f"dgesv_ of {sample} matrices of the same n=256/8-RHS batch through host "
                   f"LAPACK ({desc}), the LAPACK call the reference clients time (cpu_gesv); "
                   f"{threads} threads each calling single-threaded LAPACK; one thread: "
                   f"{rate1:.1f} solves/s over {done} matrices"),
        "single_thread_value": round(rate1, 1),
        "gpu_forward_error_max_vs_lapack": float(fe.max()),
    }


def rocsolver_reference_gpu(torch, handle, A, B, reps: int):
    """Plain FP64 dgesv of the image's rocSOLVER (comparison only)."""
    path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "librocsolver.so")
    try:
        rs = ctypes.CDLL(path)
    except OSError:
        return None
    P, I, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
    f = rs.rocsolver_dgesv_strided_batched
    f.argtypes = [P, I, I, P, I, S, P, S, P, I, S, P, I]
    Aw, Bw = A.clone(), B.clone()
    ipiv = torch.zeros((BC, N), dtype=torch.int32, device="cuda")
    info = torch.zeros(BC, dtype=torch.int32, device="cuda")

    def run():
        f(handle.value, N, NRHS, Aw.data_ptr(), N, N * N, ipiv.data_ptr(), N, Bw.data_ptr(), N,
          N * NRHS, info.data_ptr(), BC)
    total = 0.0
    for _ in range(reps):
        Aw.copy_(A)
        Bw.copy_(B)
        torch.cuda.synchronize()
        total += event_time(torch, run, 1)
    return BC / (total / reps)


def bench_vbatched(args) -> int:
    """Variable-size FP64 gesv: 10000 matrices, n log-uniform integers in
    [8, 512], U(-1,1) + n I, 1..4 RHS.  A step is one vbatched gesv of the
    whole batch (plan + persistent kernel) replayed from a captured graph;
    compared with a loop of plain dgesv_batched launches, one per (n, nrhs)
    group, on the same data."""
    import numpy as np
    import torch

    import rocsolver_amd as ra
    from rocsolver_amd import autotune
    from rocsolver_amd import workloads as wl

    torch.cuda.set_device(0)
    n, nrhs = wl.config4_sizes(10000)
    vb = ra.VBatch(n, nrhs, torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(wl.BASE_SEED + 4)
    vb.A.uniform_(-1.0, 1.0, generator=gen)
    vb.B.uniform_(-1.0, 1.0, generator=gen)
    diag = np.concatenate([vb.offA[b] + np.arange(n[b]) * (n[b] + 1) for b in range(len(n))])
    dval = np.repeat(n.astype(np.float64), n)
    vb.A[torch.from_numpy(diag).cuda()] += torch.from_numpy(dval).cuda()
    A0, B0 = vb.A.clone(), vb.B.clone()
    handle = ra.Handle()
    tuned = None
    if args.tune:
        table, ms, log = autotune.tune_vbatched(handle, vb, "gesv", reps=3)
        tuned = {"table": table, "ms": round(ms, 3), "candidates": len(log)}
    flops = float(sum(gesv_flops(int(k), int(r)) for k, r in zip(n, nrhs)))

    def restore():
        vb.A.copy_(A0)
        vb.B.copy_(B0)

    # eager persistent launch
    restore()
    vb.gesv(handle)
    torch.cuda.synchronize()
    eager = []
    for _ in range(args.warmup + args.steps):
        restore()
        eager.append(event_time(torch, lambda: vb.gesv(handle), 1))
    eager_t = sum(eager[args.warmup:]) / args.steps
    X_vb = vb.B.clone()
    LU_vb = vb.A.clone()

    # the same call captured in a graph (workspace already allocated)
    graph_t, graph_err = None, None
    try:
        restore()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            handle.set_stream(torch.cuda.current_stream())
            vb.gesv(handle)
        handle.set_stream(torch.cuda.current_stream())
        times = []
        for _ in range(args.warmup + args.steps):
            restore()
            times.append(event_time(torch, g.replay, 1))
        graph_t = sum(times[args.warmup:]) / args.steps
        graph_same = bool(torch.equal(vb.B, X_vb) and torch.equal(vb.A, LU_vb))
    except Exception as e:  # report, do not hide, a capture failure
        graph_err = f"{type(e).__name__}: {e}"
        graph_same = None
        handle.set_stream(torch.cuda.current_stream())

    # per-size loop of plain batched launches, one per (n, nrhs) group
    groups = {}
    for b in range(len(n)):
        groups.setdefault((int(n[b]), int(nrhs[b])), []).append(b)
    plan = []
    for (k, r), idx in groups.items():
        it = torch.tensor(idx, dtype=torch.int64, device="cuda")
        plan.append((k, r, len(idx), vb.pA[it].contiguous(), vb.pB[it].contiguous(),
                     torch.zeros((len(idx), k), dtype=torch.int32, device="cuda"),
                     torch.zeros(len(idx), dtype=torch.int32, device="cuda")))

    def loop():
        for k, r, cnt, pa, pb, ipv, inf in plan:
            ra.check("rocsolver_dgesv_batched", handle.value, k, r, pa.data_ptr(), k,
                     ipv.data_ptr(), k, pb.data_ptr(), k, inf.data_ptr(), cnt)
    restore()
    loop()
    torch.cuda.synchronize()
    lt = []
    for _ in range(args.warmup + args.steps):
        restore()
        lt.append(event_time(torch, loop, 1))
    loop_t = sum(lt[args.warmup:]) / args.steps
    # orders above 64 factor with substitution block solves in the persistent
    # kernel and with inverted diagonal blocks in the fixed-size one, so the
    # two agree to rounding there: report the largest solution difference
    # relative to the solution's size
    loop_same = bool(torch.equal(vb.B, X_vb) and torch.equal(vb.A, LU_vb))
    loop_rel = float((vb.B - X_vb).abs().max() / X_vb.abs().max().clamp_min(1e-300))

    best = graph_t if graph_t is not None else eager_t
    out = {
        "metric": "matrices/s variable-size FP64 gesv, 10000 matrices n in [8,512], 1-4 RHS",
        "value": round(len(n) / best, 1),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": ("variable-size gesv: 10000 matrices, n log-uniform integers in "
                                "[8, 512], U(-1,1) + n I, nrhs uniform in 1..4"),
                   "matrices": len(n), "sum_n3": int((n.astype(np.int64) ** 3).sum())},
        "gflops": round(flops / best / 1e9, 1),
        "graph_replay_ms": None if graph_t is None else round(graph_t * 1e3, 4),
        "graph_error": graph_err,
        "graph_result_identical": graph_same,
        "eager_ms": round(eager_t * 1e3, 4),
        "per_size_loop_ms": round(loop_t * 1e3, 4),
        "per_size_loop_launches": len(plan),
        "per_size_loop_result_identical": loop_same,
        "per_size_loop_max_rel_solution_diff": loop_rel,
        "speedup_vs_per_size_loop": round(loop_t / best, 3),
        "tuning_table": ra.vbatched_get_tuning(),
        "autotune": tuned,
        "nonzero_info": int((vb.info != 0).sum()),
    }
    print(json.dumps(out))
    return 0


def measured_traffic(kernel: str, summary: str = "traffic_mixed_gesv_n256_batch4096_nrhs8.md"):
    """HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated against a 1 GiB
    copy) of `kernel`, from the counter summary `summary` committed under
    profiles/; the counters need their own profiler passes, so they are not
    collected here."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", summary)
    try:
        text = open(path).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        rec = json.loads(block)[kernel]
    except (OSError, IndexError, KeyError, ValueError):
        return None
    return int(rec["fetch_bytes"] + rec["write_bytes"])


GETRF32_N, GETRF32_BATCH = 32, 131072


def measure_getrf32(args, with_cpu: bool):
    """Plain FP64 rocsolver_dgetrf_strided_batched at n = 32 (lda 32, stride
    1024, entries U(-1,1)), batch 131072 (2.2 GB of HBM traffic per call, so
    launch latency does not enter the roofline figure).  A step is one call;
    the input is restored before each step with a device copy queued ahead of
    the start event, so the GPU never waits for the host inside the timed
    region.  Roofline: HBM bytes per matrix = A read + A written + ipiv + info.
    Returns (seconds per call, GB/s, 1 GiB device-copy GB/s, nonzero info,
    host-LAPACK baseline or None)."""
    import numpy as np
    import torch

    import rocsolver_amd as ra
    n, bc = GETRF32_N, GETRF32_BATCH
    g = torch.Generator(device="cuda").manual_seed(11)
    A0 = torch.rand((bc, n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    A = torch.empty_like(A0)
    ipiv = torch.zeros((bc, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(bc, dtype=torch.int32, device="cuda")
    handle = ra.Handle()

    def step():
        ra.check("rocsolver_dgetrf_strided_batched", handle.value, n, n, A.data_ptr(), n, n * n,
                 ipiv.data_ptr(), n, info.data_ptr(), bc)
    s = torch.cuda.current_stream()
    for _ in range(max(args.warmup, 1)):
        A.copy_(A0)
        step()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    times = []
    for _ in range(max(args.steps, 3)):
        A.copy_(A0)
        e0.record(s)
        step()
        e1.record(s)
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    t = float(np.mean(times))
    per = 2 * n * n * 8 + n * 4 + 4
    gbs = bc * per / t / 1e9
    nonzero = int((info != 0).sum())
    del A
    # achievable HBM rate on this box, for context: a 1 GiB device-to-device
    # copy (1 GiB read + 1 GiB written), best of 5
    src = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    dst.copy_(src)
    copy_best = float("inf")
    for _ in range(5):
        e0.record(s)
        dst.copy_(src)
        e1.record(s)
        torch.cuda.synchronize()
        copy_best = min(copy_best, e0.elapsed_time(e1) * 1e-3)
    copy_gbs = 2 * src.numel() * 8 / copy_best / 1e9
    del src, dst
    cpu = None
    if with_cpu:
        from oracle import lapack
        if lapack.available():
            lapack.set_num_threads(1)
            sample = 16384
            Ah = A0[:sample].cpu().numpy().copy()
            t0 = time.perf_counter()
            lapack.getrf(Ah, n)
            t1 = time.perf_counter()
            cpu = {"value": round(sample / (t1 - t0), 1), "unit": "matrices/s", "cores": 1,
                   "kind": "reference",
                   "sample": (f"dgetrf_ of {sample} matrices of the same batch through host "
                              f"LAPACK ({lapack.load()[2]}), one thread")}
    del A0
    handle.close()
    return t, gbs, copy_gbs, nonzero, cpu


def getrf32_roofline(t, gbs, copy_gbs):
    n = GETRF32_N
    per = 2 * n * n * 8 + n * 4 + 4
    return {"kernel": "getrf_small_kernel<double,32,32,pivot> (team of 32 lanes per matrix)",
            "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_TBS * 1e3,
            "unit": "GB/s", "frac": round(gbs / 1e3 / HBM_PEAK_TBS, 4),
            "traffic": measured_traffic("getrf_small_kernel<double,32,32,pivot>",
                                        "traffic_getrf_n32_batch131072.md"),
            "bytes_per_matrix": per, "launch_ms": round(t * 1e3, 4),
            "device_copy_gbs": round(copy_gbs, 1),
            "frac_of_device_copy": round(gbs / copy_gbs, 4)}


def bench_getrf32(args) -> int:
    """--workload getrf32: the n=32 dgetrf HBM roofline as its own JSON line."""
    import torch
    torch.cuda.set_device(0)
    n, bc = GETRF32_N, GETRF32_BATCH
    t, gbs, copy_gbs, nonzero, cpu = measure_getrf32(args, not args.no_cpu)
    out = {
        "metric": "n=32 dgetrf_strided_batched HBM roofline fraction",
        "value": round(gbs / 1e3 / HBM_PEAK_TBS, 4),
        "unit": "fraction of 8 TB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded U(-1,1), generated on the GPU)",
        "config": {"workload": "plain FP64 dgetrf_strided_batched, n=32, lda=32, stride 1024",
                   "n": n, "batch": bc},
        "matrices_per_s": round(bc / t, 1),
        "nonzero_info": nonzero,
        "roofline": getrf32_roofline(t, gbs, copy_gbs),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def event_ms(torch, fn, reps: int, pre=None) -> float:
    """Mean duration (ms) of fn() over reps calls, HIP events on the current
    stream; pre() (restoring inputs) runs before each start event."""
    s = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    tot = 0.0
    for _ in range(reps):
        if pre:
            pre()
        e0.record(s)
        fn()
        e1.record(s)
        torch.cuda.synchronize()
        tot += e0.elapsed_time(e1)
    return tot / reps


def measure_config1(torch, ra, handle, reps: int):
    """C1: plain FP64 dgetrf + dgetrs (1 RHS) of 1024 matrices of order 32,
    U(-1,1), as captured-graph replays.  The configured batch is 17 MB, a few
    microseconds of HBM time, below launch latency: one graph holds 16 factor
    + solve pairs, each on its own pristine copy of the batch (restored
    outside the timed region), and a replay is timed with HIP events."""
    n, bc, K = 32, 1024, 16
    g = torch.Generator(device="cuda").manual_seed(20261203)
    A0 = torch.rand((bc, n, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    B0 = torch.rand((bc, 1, n), dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    As = [A0.clone() for _ in range(K)]
    Bs = [B0.clone() for _ in range(K)]
    ipiv = torch.zeros((K, bc, n), dtype=torch.int32, device="cuda")
    info = torch.zeros((K, bc), dtype=torch.int32, device="cuda")

    def pair(k):
        ra.check("rocsolver_dgetrf_strided_batched", handle.value, n, n, As[k].data_ptr(), n,
                 n * n, ipiv[k].data_ptr(), n, info[k].data_ptr(), bc)
        ra.check("rocsolver_dgetrs_strided_batched", handle.value, ra.OP_NONE, n, 1,
                 As[k].data_ptr(), n, n * n, ipiv[k].data_ptr(), n, Bs[k].data_ptr(), n, n, bc)

    def restore():
        for k in range(K):
            As[k].copy_(A0)
            Bs[k].copy_(B0)
    pair(0)
    torch.cuda.synchronize()
    cs = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    restore()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=cs):
        handle.set_stream(cs)
        for k in range(K):
            pair(k)
    handle.set_stream(torch.cuda.current_stream())
    ms = event_ms(torch, graph.replay, reps, restore) / K
    # eager pairs for comparison (launch overhead included)
    eager = event_ms(torch, lambda: pair(0), reps, restore)
    X = Bs[0][:, 0, :].double()
    A = A0.transpose(1, 2)
    R = (torch.matmul(A, X.unsqueeze(2)).squeeze(2) - B0[:, 0, :]).abs().amax(dim=1)
    berr = (R / (A.abs().sum(dim=2).amax(dim=1) * X.abs().amax(dim=1))).amax().item()
    per_getrf, per_getrs = 2 * n * n * 8 + n * 4 + 4, n * n * 8 + n * 4 + 2 * n * 8
    gbs = bc * (per_getrf + per_getrs) / (ms * 1e-3) / 1e9
    return {"workload": "dgetrf + dgetrs, n=32, batch 1024, 1 RHS, U(-1,1)",
            "ms_per_factor_and_solve_graph": round(ms, 5), "ms_eager": round(eager, 4),
            "systems_per_s": round(bc / (ms * 1e-3), 1),
            "max_backward_error": berr,
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_TBS * 1e3,
                         "unit": "GB/s", "frac": round(gbs / 1e3 / HBM_PEAK_TBS, 4),
                         "bytes_per_system": per_getrf + per_getrs,
                         "note": "17 MB per call: MALL-resident, launch-latency bound"}}


def measure_config2(torch, ra, handle, reps: int):
    """C2: plain FP64 dgetrf of 16384 matrices of order 128 (N(0,1) + 12.8 I)
    and dgetrs with 4 N(0,1) right-hand sides; A restored outside the timed
    region before each factorisation."""
    n, bc, nrhs = 128, 16384, 4
    g = torch.Generator(device="cuda").manual_seed(20261204)
    A0 = torch.randn((bc, n, n), dtype=torch.float64, device="cuda", generator=g)
    A0.diagonal(dim1=1, dim2=2).add_(0.1 * n)
    B0 = torch.randn((bc, nrhs, n), dtype=torch.float64, device="cuda", generator=g)
    A, B = torch.empty_like(A0), torch.empty_like(B0)
    ipiv = torch.zeros((bc, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(bc, dtype=torch.int32, device="cuda")

    def getrf():
        ra.check("rocsolver_dgetrf_strided_batched", handle.value, n, n, A.data_ptr(), n, n * n,
                 ipiv.data_ptr(), n, info.data_ptr(), bc)

    def getrs():
        ra.check("rocsolver_dgetrs_strided_batched", handle.value, ra.OP_NONE, n, nrhs,
                 A.data_ptr(), n, n * n, ipiv.data_ptr(), n, B.data_ptr(), n, n * nrhs, bc)
    A.copy_(A0)
    getrf()
    tf = event_ms(torch, getrf, reps, lambda: A.copy_(A0))
    ts = event_ms(torch, getrs, reps, lambda: B.copy_(B0))
    fl_f, fl_s = lu_flops(n), nrhs * (2.0 * n * n - n)
    by_f, by_s = 2 * n * n * 8 + n * 4 + 4, n * n * 8 + n * 4 + 2 * nrhs * n * 8
    del A0, A
    return {"workload": "dgetrf + dgetrs, n=128, batch 16384, 4 RHS, N(0,1) + 12.8 I",
            "dgetrf_ms": round(tf, 4), "dgetrs_ms": round(ts, 4),
            "dgetrf_gbs": round(bc * by_f / (tf * 1e-3) / 1e9, 1),
            "dgetrf_tflops": round(bc * fl_f / (tf * 1e-3) / 1e12, 3),
            "dgetrs_gbs": round(bc * by_s / (ts * 1e-3) / 1e9, 1),
            "nonzero_info": int((info != 0).sum()),
            "roofline": {"kernel": "lu_fused_kernel<double,double,128>", "bound": "hbm",
                         "achieved": round(bc * by_f / (tf * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_TBS * 1e3, "unit": "GB/s",
                         "frac": round(bc * by_f / (tf * 1e-3) / 1e12 / HBM_PEAK_TBS, 4),
                         "fp64_frac": round(bc * fl_f / (tf * 1e-3) / 1e12 / 78.6, 4),
                         "bytes_per_matrix": by_f, "flops_per_matrix": fl_f}}


def measure_config4(torch, ra, handle, reps: int):
    """C4: variable-size FP64 gesv of 10000 matrices (n log-uniform integers in
    [8, 512], U(-1,1) + n I, 1..4 RHS) in one persistent launch, replayed from a
    captured graph; inputs restored outside the timed region."""
    import numpy as np

    from rocsolver_amd import workloads as wl
    n, nrhs = wl.config4_sizes(10000)
    vb = ra.VBatch(n, nrhs, torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(wl.BASE_SEED + 4)
    vb.A.uniform_(-1.0, 1.0, generator=gen)
    vb.B.uniform_(-1.0, 1.0, generator=gen)
    diag = np.concatenate([vb.offA[b] + np.arange(n[b]) * (n[b] + 1) for b in range(len(n))])
    vb.A[torch.from_numpy(diag).cuda()] += torch.from_numpy(np.repeat(n.astype(np.float64), n)).cuda()
    A0, B0 = vb.A.clone(), vb.B.clone()

    def restore():
        vb.A.copy_(A0)
        vb.B.copy_(B0)
    restore()
    vb.gesv(handle)
    torch.cuda.synchronize()
    cs = torch.cuda.Stream()
    graph = torch.cuda.CUDAGraph()
    restore()
    torch.cuda.synchronize()
    with torch.cuda.graph(graph, stream=cs):
        handle.set_stream(cs)
        vb.gesv(handle)
    handle.set_stream(torch.cuda.current_stream())
    ms = event_ms(torch, graph.replay, reps, restore)
    flops = float(sum(gesv_flops(int(k), int(r)) for k, r in zip(n, nrhs)))
    return {"workload": "variable-size dgesv, 10000 matrices, n log-uniform in [8, 512], 1-4 RHS",
            "graph_replay_ms": round(ms, 4), "matrices_per_s": round(len(n) / (ms * 1e-3), 1),
            "tflops": round(flops / (ms * 1e-3) / 1e12, 3),
            "fp64_frac": round(flops / (ms * 1e-3) / 1e12 / 78.6, 4),
            "nonzero_info": int((vb.info != 0).sum())}


def measure_config5(torch, ra, handle, reps: int):
    """C5: adjoint of X = A^-1 B, FP64, 2048 systems of order 96 (N(0,1) + 10 I),
    16 RHS: forward gesv, then the backward dB = A^-T dX (transposed solve
    with the stored LU and pivots) and dA = -dB X^T."""
    n, bc, nrhs = 96, 2048, 16
    g = torch.Generator(device="cuda").manual_seed(20261205)
    A0 = torch.randn((bc, n, n), dtype=torch.float64, device="cuda", generator=g)
    A0.diagonal(dim1=1, dim2=2).add_(10.0)
    B0 = torch.randn((bc, nrhs, n), dtype=torch.float64, device="cuda", generator=g)
    dX = torch.randn((bc, nrhs, n), dtype=torch.float64, device="cuda", generator=g)
    LU, X = torch.empty_like(A0), torch.empty_like(B0)
    dB, dA = torch.empty_like(B0), torch.empty_like(A0)
    ipiv = torch.zeros((bc, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(bc, dtype=torch.int32, device="cuda")

    def fwd():
        ra.check("rocsolver_dgesv_strided_batched", handle.value, n, nrhs, LU.data_ptr(), n, n * n,
                 ipiv.data_ptr(), n, X.data_ptr(), n, n * nrhs, info.data_ptr(), bc)

    def bwd():
        ra.check("rocsolver_amd_dgesv_adjoint_strided_batched", handle.value, n, nrhs,
                 LU.data_ptr(), n, n * n, ipiv.data_ptr(), n, X.data_ptr(), n, n * nrhs,
                 dX.data_ptr(), n, n * nrhs, dB.data_ptr(), n, n * nrhs, dA.data_ptr(), n, n * n,
                 bc)

    def restore():
        LU.copy_(A0)
        X.copy_(B0)
    tf = event_ms(torch, fwd, reps, restore)
    restore()
    fwd()
    tb = event_ms(torch, bwd, reps)
    per = 184704  # LU read + X, dX read + dB, dA written (bytes per system)
    return {"workload": "adjoint of X = A^-1 B, FP64, n=96, batch 2048, 16 RHS",
            "forward_gesv_ms": round(tf, 4), "backward_adjoint_ms": round(tb, 4),
            "roofline": {"kernel": "adjoint (transposed getrs + outer product)", "bound": "hbm",
                         "achieved": round(bc * per / (tb * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_TBS * 1e3, "unit": "GB/s",
                         "frac": round(bc * per / (tb * 1e-3) / 1e12 / HBM_PEAK_TBS, 4),
                         "bytes_per_system": per}}


def measure_half_width(torch, ra, handle, A, B, reps: int):
    """The metric workload solved with equilibrated FP16 / BF16 matrix-core
    operands in the FP32 LU's updates (rocsolver_amd_dxgesv_strided_batched):
    time per batch, iteration histogram, FP64 fallbacks, backward error."""
    out = {}
    bc, n, lda = A.shape
    _, nrhs, ldb = B.shape
    X = torch.zeros((bc, nrhs, n), dtype=torch.float64, device="cuda")
    it = torch.zeros(bc, dtype=torch.int32, device="cuda")
    info = torch.zeros(bc, dtype=torch.int32, device="cuda")
    for name in ("fp16", "bf16"):
        def run():
            ra.check("rocsolver_amd_dxgesv_strided_batched", handle.value, ra.FACTOR[name], n, nrhs,
                     A.data_ptr(), lda, n * lda, B.data_ptr(), ldb, nrhs * ldb, X.data_ptr(), n,
                     nrhs * n, it.data_ptr(), info.data_ptr(), None, bc)
        run()
        ms = event_ms(torch, run, reps)
        Arm = A.transpose(1, 2)
        R = B.transpose(1, 2) - torch.matmul(Arm, X.transpose(1, 2))
        bwd = (R.abs().amax(dim=1) / (Arm.abs().sum(dim=2).amax(dim=1)[:, None]
                                      * X.abs().amax(dim=2) + B.abs().amax(dim=2))).amax().item()
        itc = it.cpu()
        hist = {int(k): int(v) for k, v in zip(*torch.unique(itc, return_counts=True))}
        out[name] = {"ms": round(ms, 4), "solves_per_s": round(bc / (ms * 1e-3), 1),
                     "iter_histogram": hist, "fallback_matrices": int((itc < 0).sum()),
                     "max_backward_error": bwd, "target_10_n_eps64": 10 * n * EPS64}
    return out


def measure_blocked(torch, ra, handle, reps: int):
    """Blocked path (orders beyond the fused kernels, tall matrices): plain
    FP64 dgetrf of 64 matrices of order 1024 and of 64 matrices of 4096 x 128,
    and dgetrs of the square ones with 8 right-hand sides, beside the image's
    rocSOLVER dgetrf on the same inputs (comparison only)."""
    bc, nrhs = 64, 8
    g = torch.Generator(device="cuda").manual_seed(20261205)
    out = {"workload": "dgetrf n=1024 and 4096x128, batch 64 each; dgetrs n=1024, 8 RHS; N(0,1)"}
    rs = None
    try:
        rs = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                      "librocsolver.so"))
        P, I, S = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64
        rs.rocsolver_dgetrf_strided_batched.argtypes = [P, I, I, P, I, S, P, S, P, I]
    except OSError:
        rs = None
    for key, m, n in (("square_n1024", 1024, 1024), ("tall_4096x128", 4096, 128)):
        A0 = torch.randn((bc, n, m), dtype=torch.float64, device="cuda", generator=g)
        A = torch.empty_like(A0)
        k = min(m, n)
        ipiv = torch.zeros((bc, k), dtype=torch.int32, device="cuda")
        info = torch.zeros(bc, dtype=torch.int32, device="cuda")

        def getrf():
            ra.check("rocsolver_dgetrf_strided_batched", handle.value, m, n, A.data_ptr(), m,
                     m * n, ipiv.data_ptr(), k, info.data_ptr(), bc)

        def ref():
            rs.rocsolver_dgetrf_strided_batched(handle.value, m, n, A.data_ptr(), m, m * n,
                                                ipiv.data_ptr(), k, info.data_ptr(), bc)
        A.copy_(A0)
        getrf()
        t = event_ms(torch, getrf, reps, lambda: A.copy_(A0))
        fl = (m * n * n - n ** 3 / 3.0 - n * n / 2.0) if m >= n else lu_flops(n)
        r = {"dgetrf_ms": round(t, 4), "tflops": round(bc * fl / (t * 1e-3) / 1e12, 3),
             "fp64_frac": round(bc * fl / (t * 1e-3) / 1e12 / 78.6, 4),
             "nonzero_info": int((info != 0).sum())}
        if key == "square_n1024":
            B0 = torch.randn((bc, nrhs, n), dtype=torch.float64, device="cuda", generator=g)
            B = torch.empty_like(B0)

            def getrs():
                ra.check("rocsolver_dgetrs_strided_batched", handle.value, ra.OP_NONE, n, nrhs,
                         A.data_ptr(), n, n * n, ipiv.data_ptr(), n, B.data_ptr(), n, n * nrhs,
                         bc)
            A.copy_(A0)
            getrf()
            r["dgetrs_8rhs_ms"] = round(event_ms(torch, getrs, reps, lambda: B.copy_(B0)), 4)
            del B0, B
        if rs is not None:
            A.copy_(A0)
            ref()
            tr = event_ms(torch, ref, reps, lambda: A.copy_(A0))
            r["rocsolver_dgetrf_ms"] = round(tr, 4)
            r["speedup_vs_rocsolver"] = round(tr / t, 2)
        out[key] = r
        del A0, A
    return out


def other_configs(args, torch, ra, A, B):
    """Every other configuration, measured in the same run (each guarded: a
    failure is reported, not hidden)."""
    handle = ra.Handle()
    reps = max(3, min(args.steps, 10))
    res = {}
    for key, fn in (("c1_getrf_getrs_n32_b1024", lambda: measure_config1(torch, ra, handle, reps)),
                    ("c2_getrf_getrs_n128_b16384", lambda: measure_config2(torch, ra, handle, reps)),
                    ("c4_vbatched_gesv_10000", lambda: measure_config4(torch, ra, handle, reps)),
                    ("c5_adjoint_n96_b2048", lambda: measure_config5(torch, ra, handle, reps)),
                    ("blocked_getrf_n1024_b64", lambda: measure_blocked(torch, ra, handle, reps)),
                    ("c3_half_width_factors",
                     lambda: measure_half_width(torch, ra, handle, A, B, reps))):
        try:
            res[key] = fn()
        except Exception as e:  # report, do not hide
            res[key] = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.synchronize()
    try:
        t, gbs, copy_gbs, nonzero, _ = measure_getrf32(args, False)
        res["getrf_n32_b131072_hbm"] = dict(getrf32_roofline(t, gbs, copy_gbs),
                                            nonzero_info=nonzero)
    except Exception as e:
        res["getrf_n32_b131072_hbm"] = {"error": f"{type(e).__name__}: {e}"}
    handle.close()
    return res


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-budget", type=float, default=8.0,
                    help="seconds of single-thread host LAPACK in the CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other configurations (C1, C2, C4, C5, half-width factors, "
                         "n=32 HBM roofline) that the default line also reports")
    ap.add_argument("--workload", choices=["mixed", "vbatched", "getrf32"], default="mixed",
                    help="mixed: the headline n=256 mixed-precision solve (default); vbatched: "
                         "10000 matrices of orders 8..512 in one persistent launch; getrf32: "
                         "HBM roofline of plain n=32 dgetrf_strided_batched")
    ap.add_argument("--tune", action="store_true",
                    help="vbatched: run the schedule autotuner on the workload first")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every matrix is "
              "independent, so the batch would be split across GPUs with no exchange)",
              file=sys.stderr)
        return 2
    if args.workload == "vbatched":
        return bench_vbatched(args)
    if args.workload == "getrf32":
        return bench_getrf32(args)

    import torch

    import rocsolver_amd as ra
    from rocsolver_amd import workloads as wl

    torch.cuda.set_device(0)
    A, B, kappa = wl.config3_torch(BC, N, NRHS, device="cuda")
    torch.cuda.synchronize()
    handle = ra.Handle()
    X = torch.zeros((BC, NRHS, N), dtype=torch.float64, device="cuda")
    it = torch.zeros(BC, dtype=torch.int32, device="cuda")
    info = torch.zeros(BC, dtype=torch.int32, device="cuda")
    berr = torch.zeros(BC, dtype=torch.float64, device="cuda")

    def mixed_step():
        ra.check("rocsolver_amd_dsgesv_strided_batched", handle.value, N, NRHS, A.data_ptr(), N,
                 N * N, B.data_ptr(), N, N * NRHS, X.data_ptr(), N, N * NRHS, it.data_ptr(),
                 info.data_ptr(), berr.data_ptr(), BC)

    for _ in range(args.warmup):
        mixed_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mixed_step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = BC * args.steps / elapsed

    # accuracy of the timed solves, measured on the GPU in FP64
    Arm = A.transpose(1, 2)  # row-major view of each column-major matrix
    R = B.transpose(1, 2) - torch.matmul(Arm, X.transpose(1, 2))
    anorm = Arm.abs().sum(dim=2).amax(dim=1)
    rn = R.abs().amax(dim=1)
    xn = X.abs().amax(dim=2)
    bn = B.abs().amax(dim=2)
    bwd = (rn / (anorm[:, None] * xn + bn)).amax(dim=1)
    iters = it.cpu()
    hist = {int(k): int(v) for k, v in zip(*torch.unique(iters, return_counts=True))}

    # plain FP64 dgesv of this library (A is overwritten: refreshed outside the timing)
    Aw, Bw = A.clone(), B.clone()
    ipiv = torch.zeros((BC, N), dtype=torch.int32, device="cuda")
    info64 = torch.zeros(BC, dtype=torch.int32, device="cuda")

    def plain_step():
        ra.check("rocsolver_dgesv_strided_batched", handle.value, N, NRHS, Aw.data_ptr(), N, N * N,
                 ipiv.data_ptr(), N, Bw.data_ptr(), N, N * NRHS, info64.data_ptr(), BC)
    plain_reps = max(2, min(args.steps, 5))
    tot = 0.0
    for r in range(plain_reps + 1):
        Aw.copy_(A)
        Bw.copy_(B)
        torch.cuda.synchronize()
        t = event_time(torch, plain_step, 1)
        if r > 0:  # first call is a cold call
            tot += t
    plain = BC / (tot / plain_reps)
    ref_gpu = rocsolver_reference_gpu(torch, handle, A, B, 2)
    del Aw, Bw

    # dominant kernel: the FP32 factorisation of the FP64 batch, timed alone
    LU = torch.empty((BC, N, N), dtype=torch.float32, device="cuda")
    ipiv32 = torch.zeros((BC, N), dtype=torch.int32, device="cuda")
    info32 = torch.zeros(BC, dtype=torch.int32, device="cuda")

    def lu_step():
        ra.check("rocsolver_amd_dsgetrf_strided_batched", handle.value, N, A.data_ptr(), N, N * N,
                 LU.data_ptr(), N, N * N, ipiv32.data_ptr(), N, info32.data_ptr(), BC)
    lu_step()
    torch.cuda.synchronize()
    lu_t = event_time(torch, lu_step, max(3, args.steps))
    achieved = BC * lu_flops(N) / lu_t / 1e12
    del LU

    cpu = None if args.no_cpu else cpu_baseline(torch, A, B, X, args.cpu_budget)
    others = None if args.no_others else other_configs(args, torch, ra, A, B)

    out = {
        "metric": "solves/s n=256 batch 4096 8 RHS: FP32-MFMA+IR vs plain FP64 dgesv at eps64",
        "value": round(value, 1),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 LU + f64 refinement",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": ("mixed-precision batched gesv: n=256, batch 4096, 8 RHS, "
                                "A = U diag(s) V^T, kappa log-uniform in [1e2, 1e6]"),
                   "n": N, "batch": BC, "nrhs": NRHS},
        "plain_fp64_dgesv_solves_per_s": round(plain, 1),
        "speedup_vs_plain_fp64": round(value / plain, 3),
        "rocsolver_dgesv_solves_per_s": None if ref_gpu is None else round(ref_gpu, 1),
        "accuracy": {
            "max_backward_error": float(bwd.max()),
            "target_10_n_eps64": 10 * N * EPS64,
            "iter_histogram": hist,
            "fallback_matrices": int((iters < 0).sum()),
            "nonzero_info": int((info != 0).sum()),
        },
        "roofline": {
            "kernel": "lu_fused_kernel<double,float,256> (FP32 LU of the FP64 batch)",
            "bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_MFMA_PEAK_TFLOPS, 4),
            "traffic": measured_traffic("lu_fused_kernel<double,float,256>"),
            "launch_ms": round(lu_t * 1e3, 4),
        },
        "cpu_baseline": cpu,
        "other_configs": others,
    }
    if others and "getrf_n32_b131072_hbm" in others and "frac" in others["getrf_n32_b131072_hbm"]:
        out["getrf_n32_hbm_fraction"] = others["getrf_n32_b131072_hbm"]["frac"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
