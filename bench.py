"""SGEMM 8192^3 on one MI355X: BF16x3 split-precision MFMA vs plain FP32 MFMA,
each with its error against an FP64 gemm of the same inputs.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--variants NN,NT,TN,TT]
  python bench.py --configs cfg1,cfg2,cfg2b,cfg3a,cfg3b,cfg4,l3

The second form measures the other workloads (mojoblas_amd/bench_configs.py
prints their GPU lines) and times the restated mojoBLAS CPU path of each
beside it on the same host, one JSON line per CPU leg.

A step is one SGEMM C := A*B (alpha=1, beta=0, column-major, m=n=k=8192,
A/B iid N(0,1)) with operands resident in HBM.  The headline value is the
BF16x3 N,N throughput; the plain-FP32 kernel and the other transpose
variants are timed the same way and reported alongside.  Kernel time comes
from HIP events recorded on the stream the kernels are launched on
(torch's current stream, which the library is handed).  One JSON line is
printed on stdout; progress goes to stderr.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

N = 8192
FLOPS = 2.0 * N * N * N
PEAK_BF16 = 2516.6  # TFLOP/s, dense bf16 MFMA, MI355X
PEAK_FP32 = 157.3   # TFLOP/s, fp32 (matrix = vector), MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_gemm(mb, torch, ta, tb, A, B, C, prec, steps, warmup):
    """Whole-step time (events on torch's current stream, which is the launch
    stream the library is handed) and, for the split-precision paths, the
    library's own event timing of the main kernel and of the pack kernels."""
    import ctypes

    from mojoblas_amd import _lib

    L = _lib.lib()
    for _ in range(warmup):
        mb.gemm(ta, tb, N, N, N, 1.0, A, N, B, N, 0.0, C, N, precision=prec)
    torch.cuda.synchronize()
    L.mbh_timing_enable(1)
    tm = np.zeros(4)
    L.mbh_timing_read(ctypes.c_void_p(tm.ctypes.data), 1)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        rc = mb.gemm(ta, tb, N, N, N, 1.0, A, N, B, N, 0.0, C, N, precision=prec)
        assert rc == 0
    e.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    L.mbh_timing_read(ctypes.c_void_p(tm.ctypes.data), 1)
    L.mbh_timing_enable(0)
    step_ms = s.elapsed_time(e) / steps
    if tm[1] > 0:
        main_ms, pack_ms = tm[0] / tm[1], tm[2] / tm[3]
    else:  # single-kernel path (fp32): the step is the kernel
        main_ms, pack_ms = step_ms, 0.0
    return step_ms, wall / steps * 1e3, main_ms, pack_ms


def errors(torch, C, C64, ABabs):
    d = (C.double() - C64).abs()
    return {
        "max_abs": float(d.max()),
        "rel_frobenius": float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(C64)),
        "max_normwise": float((d / ABabs).max()),  # max |dC_ij| / (|A||B|)_ij
    }


def host_info():
    """CPU model, the logical CPUs this process may use, and the NUMA layout."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nodes = {}
    base = "/sys/devices/system/node"
    try:
        for d in sorted(os.listdir(base)):
            if d.startswith("node") and d[4:].isdigit():
                with open(os.path.join(base, d, "cpulist")) as f:
                    nodes[d] = f.read().strip()
    except OSError:
        pass
    return {"cpu_model": model, "cpus_usable": len(os.sched_getaffinity(0)), "cpus_online": os.cpu_count(),
            "numa_nodes": nodes}


def _repeat(fn, seconds, min_reps=1):
    """fn() repeated until `seconds` have passed (at least min_reps times);
    returns (reps, elapsed seconds), mirroring the reference's harness
    (benchmarks/benchmark_level3.mojo:29,38: warm-up, then repeat for >= 1 s)."""
    fn()  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if reps >= min_reps and el >= seconds:
            return reps, el


def _cpu_leg(config, routine, flops_per_rep, reps, el, cores, sample):
    rec = {"config": config, "routine": routine,
           "cpu_baseline": {"value": round(reps * flops_per_rep / el / 1e12, 5), "unit": "TFLOP/s",
                            "cores": cores, "kind": "port", "sample": sample, "seconds": round(el, 2)}}
    rec["cpu_baseline"].update(host_info())
    print(json.dumps(rec), flush=True)


def cpu_configs(names, seconds):
    """The restated mojoBLAS CPU path (oracle/baseline.c) of each workload,
    timed on this host on a bounded sample of the workload (scaled to
    TFLOP/s): the column sample keeps the per-column work of the full size,
    and columns are independent in every restated loop nest."""
    import oracle

    rng = np.random.default_rng(11)
    thr = oracle.base_threads()
    f32 = np.float32
    if "cfg1" in names:
        n = 512
        a, b = rng.uniform(-1, 1, n * n).astype(f32), rng.uniform(-1, 1, n * n).astype(f32)
        c = rng.uniform(-1, 1, n * n).astype(f32)
        reps, el = _repeat(lambda: oracle.base_sgemm("N", "N", n, n, n, 1.5, a, n, b, n, 0.5, c, n), seconds)
        _cpu_leg("cfg1", "sgemm NN 512^3 alpha=1.5 beta=0.5 (gemm_v11 blocking)", 2.0 * n ** 3, reps, el, thr,
                 "%d x the full workload" % reps)
    if "cfg2" in names:
        n, cols = 8192, 64
        a, b = rng.standard_normal(n * n).astype(f32), rng.standard_normal(n * n).astype(f32)
        c = np.zeros(n * cols, f32)
        for t in ("NT", "TN", "TT"):
            reps, el = _repeat(lambda: oracle.base_sgemm(t[0], t[1], n, cols, n, 1.0, a, n, b, n, 0.0, c, n),
                               seconds / 3)
            _cpu_leg("cfg2", "sgemm %s 8192^3 (_gemm_naive)" % t, 2.0 * n * n * cols, reps, el, thr,
                     "%d x the first %d columns of C (m=k=8192): the column-parallel naive sweep does the same "
                     "work per column as at n=8192" % (reps, cols))
        del a, b, c
    if "cfg2b" in names:
        m, n, k = 4096, 4096, 16384
        a, b = rng.standard_normal(m * k).astype(f32), rng.standard_normal(k * n).astype(f32)
        c = np.zeros(m * n, f32)
        reps, el = _repeat(lambda: oracle.base_sgemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m), seconds)
        _cpu_leg("cfg2b", "sgemm NN 4096x4096x16384 (gemm_v11 blocking)", 2.0 * m * n * k, reps, el, thr,
                 "%d x the full workload" % reps)
        del a, b, c
    if "cfg3a" in names:
        m, cols = 4096, 128
        lo = np.tril(rng.uniform(-1, 1, (m, m)) / m, -1) + np.diag(rng.uniform(1, 2, m))
        a = np.asfortranarray(lo.astype(f32)).ravel(order="F")
        b0 = rng.uniform(-1, 1, m * cols).astype(f32)
        b = b0.copy()

        def run():
            b[:] = b0
            oracle.base_strsm_left("L", "N", "N", m, cols, 1.0, a, m, b, m)
        reps, el = _repeat(run, seconds)
        _cpu_leg("cfg3a", "strsm L,L,N,N m=4096 (trsm.mojo:110-148, serial)", float(m) * m * cols, reps, el, 1,
                 "%d x the first %d of the 4096 right-hand sides (B reset each time): the reference's trsm "
                 "solves column after column" % (reps, cols))
        del a, b, b0
    if "cfg3b" in names:
        n, k = 8192, 2048
        a = rng.standard_normal(n * k).astype(f32)
        c = np.zeros(n * n, f32)
        reps, el = _repeat(lambda: oracle.base_ssyrk_n("L", n, k, 1.0, a, n, 0.0, c, n), seconds)
        _cpu_leg("cfg3b", "ssyrk L,N n=8192 k=2048 (syrk.mojo:136-166, column-parallel)", float(n) * (n + 1) * k,
                 reps, el, thr, "%d x the full workload" % reps)
        del a, c
    if "cfg4" in names:
        from mojoblas_amd import batched

        m, n, k = batched.config4_shapes()
        ops = [(int(mi), int(ni), int(ki), rng.uniform(-1, 1, mi * ki).astype(f32),
                rng.uniform(-1, 1, ki * ni).astype(f32), np.zeros(mi * ni, f32)) for mi, ni, ki in zip(m, n, k)]

        def run():
            for mi, ni, ki, a, b, c in ops:
                oracle.base_sgemm("N", "N", mi, ni, ki, 1.0, a, mi, b, ki, 0.0, c, mi)
        reps, el = _repeat(run, seconds)
        _cpu_leg("cfg4", "4096 variable-size NN sgemms, one gemm call each (gemm_v11 blocking)",
                 float(np.sum(2.0 * m * n * k)), reps, el, thr, "%d x the whole batch" % reps)
        del ops
    if "l3" in names:
        m, cols = 8192, 256
        s = rng.uniform(-1, 1, (m, m)).astype(f32)
        a = np.asfortranarray(np.tril(s)).ravel(order="F")
        b = rng.standard_normal(m * cols).astype(f32)
        c = np.zeros(m * cols, f32)
        reps, el = _repeat(lambda: oracle.base_ssymm("L", "L", m, cols, 1.0, a, m, b, m, 0.0, c, m), seconds)
        _cpu_leg("l3", "ssymm L,L m=8192 (symm.mojo:96-158, column-parallel)", 2.0 * m * m * cols, reps, el, thr,
                 "%d x the first %d of the 8192 columns of B and C" % (reps, cols))
        del s, a, b, c
        m, cols = 4096, 128
        a = np.asfortranarray(np.tril(rng.uniform(-1, 1, (m, m))).astype(f32)).ravel(order="F")
        b0 = rng.standard_normal(m * cols).astype(f32)
        b = b0.copy()

        def run():
            b[:] = b0
            oracle.base_strmm("L", "L", "N", "N", m, cols, 1.0, a, m, b, m)
        reps, el = _repeat(run, seconds)
        _cpu_leg("l3", "strmm L,L,N,N m=4096 (trmm.mojo:128-147, serial)", float(m) * m * cols, reps, el, 1,
                 "%d x the first %d of the 4096 columns of B (B reset each time)" % (reps, cols))


def cpu_baseline(args):
    """Restated mojoBLAS gemm_v11 (oracle/baseline.c) on the host cores, on the
    same 8192^3 N,N workload, repeated until >= args.cpu_seconds."""
    import oracle

    a = np.random.default_rng(1).standard_normal(N * N).astype(np.float32)
    b = np.random.default_rng(2).standard_normal(N * N).astype(np.float32)
    c = np.zeros(N * N, np.float32)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.base_sgemm("N", "N", N, N, N, 1.0, a, N, b, N, 0.0, c, N)
        reps += 1
        el = time.perf_counter() - t0
        log("cpu baseline rep %d: %.1f s" % (reps, el))
        if el >= args.cpu_seconds:
            break
    tf = reps * FLOPS / el / 1e12
    rec = {
        "value": round(tf, 4),
        "unit": "TFLOP/s",
        "cores": oracle.base_threads(),
        "kind": "port",
        "sample": "%d x SGEMM N,N 8192^3 (alpha=1, beta=0) through the restated mojoBLAS gemm_v11 blocking "
                  "(TK=1024, MC=256, 16x16 SIMD outer-product micro-kernel), %.1f s" % (reps, el),
    }
    rec.update(host_info())
    return rec



def committed_traffic(root, name="gemm_packed_bf16x3_8192_traffic.txt"):
    """HBM bytes per launch of the BF16x3 main kernel at 8192^3 from the
    committed profile (two rocprofv3 PMC passes, FETCH_SIZE and WRITE_SIZE,
    calibrated against a 1 GiB copy), and None in its place when the file is
    missing or the kernel's sources no longer hash to what was profiled.
    Returns (bytes or None, note or None)."""
    tpath = os.path.join(root, "profiles", name)
    if not os.path.exists(tpath):
        return None, None
    rec = {}
    with open(tpath) as f:
        for ln in f:
            key, sep, val = ln.partition(":")
            if sep and not ln.startswith(" "):
                rec[key.strip()] = val.strip()
    srcs = rec.get("kernel_sources", "").split()
    h = hashlib.sha256()
    for src in srcs:
        path = os.path.join(root, src)
        if not os.path.exists(path):
            return None, "stale: %s is missing" % src
        with open(path, "rb") as f:
            h.update(f.read())
    if srcs and h.hexdigest() == rec.get("kernel_sources_sha256") and "traffic_bytes_per_launch" in rec:
        return int(rec["traffic_bytes_per_launch"]), None
    return None, "stale: the kernel's sources changed since the counters were collected"

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--variants", default="NN,NT,TN,TT")
    ap.add_argument("--precisions", default="bf16x3,fp32,fp8",
                    help="subset to time (profiler passes); bf16x3 is always included")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-errors", action="store_true")
    ap.add_argument("--settle", type=float, default=0.5,
                    help="seconds of untimed gemms (every precision timed) before the first timed variant")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 1 GiB device copy (known HBM bytes) to calibrate PMC byte counters")
    ap.add_argument("--configs", default="",
                    help="instead of the headline: these workloads (GPU lines from mojoblas_amd.bench_configs, "
                         "then the restated CPU path of each); cfg2 is CPU-only here (the GPU side of its "
                         "transposes is in the headline's variants)")
    ap.add_argument("--reps", type=int, default=10, help="GPU repetitions per workload with --configs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the level-3 path runs on a single MI355X; "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.configs:
        names = [c.strip() for c in args.configs.split(",") if c.strip()]
        import torch

        import mojoblas_amd as mb
        from mojoblas_amd import bench_configs

        torch.cuda.set_device(0)
        for name in names:
            if name in bench_configs.CONFIGS:
                log("== %s (GPU)" % name)
                bench_configs.CONFIGS[name](torch, mb, args.reps)
                torch.cuda.empty_cache()
        log("== CPU legs")
        cpu_configs(set(names), max(1.0, args.cpu_seconds / 4))
        return

    import torch
    import mojoblas_amd as mb

    torch.cuda.set_device(0)
    g = torch.Generator(device="cuda").manual_seed(1234)
    A = torch.randn(N * N, device="cuda", generator=g)
    B = torch.randn(N * N, device="cuda", generator=g)
    C = torch.empty(N * N, device="cuda")
    if args.copy_calibration:
        src = torch.empty(1 << 28, device="cuda")  # 1 GiB read, 1 GiB written per copy
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    variants = [v.strip().upper() for v in args.variants.split(",") if v.strip()]
    precs = ["bf16x3"] + [p for p in ("fp32", "fp8") if p in args.precisions.split(",")]
    if "NN" not in variants:
        variants = ["NN"] + variants
    # Clock settle: every precision runs, untimed, for --settle seconds before
    # anything is timed.  Without it the first variant timed (the headline
    # BF16x3 N,N) ran its main kernel about 3 % slower than the identical
    # kernel timed moments later for N,T / T,N / T,T, in every run: the GPU
    # comes out of the host-side setup still ramping its clocks.  The timed
    # regions below are unchanged (W warm-up steps, then exactly K steps).
    t_end = time.perf_counter() + args.settle
    while time.perf_counter() < t_end:
        for prec in precs:
            mb.gemm("N", "N", N, N, N, 1.0, A, N, B, N, 0.0, C, N, precision=prec)
        torch.cuda.synchronize()
    results, kernel_ms = {}, {}
    for prec in precs:
        for v in variants:
            ms, wall_ms, main_ms, pack_ms = time_gemm(mb, torch, v[0], v[1], A, B, C, prec, args.steps, args.warmup)
            results[(prec, v)] = ms
            kernel_ms[(prec, v)] = (main_ms, pack_ms)
            log("%-6s %s: %.3f ms  %.1f TFLOP/s (main kernel %.3f ms = %.1f TFLOP/s, pack %.3f ms; wall %.3f ms)"
                % (prec, v, ms, FLOPS / ms / 1e9, main_ms, FLOPS / main_ms / 1e9, pack_ms, wall_ms))

    err = {}
    if not args.no_errors:
        # FP64 gemm of the same inputs (torch fp64 on the GPU), itself checked
        # against the fp64 oracle on sampled entries.
        A64 = A.view(N, N).t().double()  # column-major buffer -> logical matrix
        B64 = B.view(N, N).t().double()
        C64 = A64 @ B64
        ABabs = A64.abs() @ B64.abs()
        import oracle

        rng = np.random.default_rng(5)
        rows, cols = rng.integers(0, N, 256), rng.integers(0, N, 256)
        ref = oracle.sgemm_sample("N", "N", N, 1.0, A.cpu().numpy(), N, B.cpu().numpy(), N, 0.0, None, rows, cols)
        chk = C64[torch.from_numpy(rows).cuda(), torch.from_numpy(cols).cuda()].cpu().numpy()
        assert np.allclose(chk, ref, rtol=1e-12, atol=1e-9), "fp64 reference disagrees with the oracle"
        for prec in precs:
            mb.gemm("N", "N", N, N, N, 1.0, A, N, B, N, 0.0, C, N, precision=prec)
            torch.cuda.synchronize()
            err[prec] = errors(torch, C.view(N, N).t(), C64, ABabs)
            log(prec, "error vs fp64:", err[prec])
        del A64, B64, C64, ABabs

    ms_bf = results[("bf16x3", "NN")]
    tf_bf = FLOPS / ms_bf / 1e9
    tf_fp = FLOPS / results[("fp32", "NN")] / 1e9 if ("fp32", "NN") in results else None
    main_ms, pack_ms = kernel_ms[("bf16x3", "NN")]
    tf_main = FLOPS / main_ms / 1e9
    peak = PEAK_BF16 / 3.0
    traffic, traffic_note = committed_traffic(os.path.dirname(os.path.abspath(__file__)))
    line = {
        "metric": "SGEMM TFLOP/s @8192^3: BF16x3-MFMA vs plain FP32 at stated FP64-oracle error",
        "value": round(tf_bf, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_bf, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (bf16x3 split on bf16 MFMA, fp32 accumulate)",
        "data": "synthetic: A, B iid N(0,1) fp32 (torch cuda Generator seed 1234), column-major, resident in HBM",
        "config": {"workload": "SGEMM N,N m=n=k=8192 lda=ldb=ldc=8192 alpha=1 beta=0", "m": N, "n": N, "k": N},
        "fp32_tflops": round(tf_fp, 3) if tf_fp else None,
        "fp8_tflops": round(FLOPS / results[("fp8", "NN")] / 1e9, 3) if ("fp8", "NN") in results else None,
        "bf16x3_over_fp32": round(tf_bf / tf_fp, 3) if tf_fp else None,
        "variants_tflops": {"%s_%s" % (p, v): round(FLOPS / ms / 1e9, 3) for (p, v), ms in results.items()},
        "errors_vs_fp64": err,
        "error_ratio_bf16x3_over_fp32": (
            {k: round(err["bf16x3"][k] / err["fp32"][k], 3) for k in err["fp32"]} if "fp32" in err else None),
        "roofline": {
            "bound": "mfma",
            "achieved": round(tf_main, 3),
            "peak": round(peak, 1),
            "peak_basis": "MI355X dense bf16 MFMA 2516.6 TFLOP/s / 3 bf16 products per fp32 product",
            "unit": "TFLOP/s",
            "frac": round(tf_main / peak, 4),
            "traffic": traffic,
            "traffic_source": ("profiles/gemm_packed_bf16x3_8192_traffic.txt: rocprofv3 --pmc FETCH_SIZE and "
                               "WRITE_SIZE passes of this kernel at this size (FETCH_SIZE x2.000, WRITE_SIZE x1.000 "
                               "from a 1 GiB device copy in the same process), kernel sources checked by sha256; "
                               "this run reads no counters" if traffic is not None else traffic_note),
            "algorithmic_bytes": 2 * 2 * N * N * 2 + N * N * 4,
            "kernel": "gemm_packed_kernel<3,256,256,2,4>",
            "kernel_ms": round(main_ms, 4),
            "pack_ms_per_step": round(pack_ms, 4),
            "fp32_kernel_frac": round(tf_fp / PEAK_FP32, 4) if tf_fp else None,
        },
    }
    if not args.no_cpu:
        del A, B, C
        torch.cuda.empty_cache()
        line["cpu_baseline"] = cpu_baseline(args)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
