"""Autotuner for the variable-size batched gemm.

    python -m mojoblas_amd.tune_vbatched --out DIR [--reps 10]

1. Tile class per shape bucket: for every (bucket(m), bucket(n)) pair of the
   6x6 grid (buckets <=16, 32, 64, 128, 256, 512) a synthetic batch of 256
   problems is drawn (m, n uniform inside their buckets, k log-uniform in
   [16, 512]) and timed with every problem on 64x64 tiles, then on 128x128
   tiles; the faster class wins the bucket.
2. Wave schedule: the 4096-problem config batch (m, n, k log-uniform in
   [16, 512]) is timed with the winning table, with all-64x64 and
   all-128x128 tables and with threshold tables (128x128 tiles once both m
   and n exceed 64, or 128), each with 1..8 persistent workgroups per CU
   (capped per tile class at what the kernel keeps resident); the fastest
   (table, count) pair wins.  The per-bucket step times each bucket alone,
   so it cannot see how classes share the GPU inside one mixed batch; the
   whole-batch tables cover that.
Both steps run for fp32 and for bf16-input mode.  Every measurement is
appended to DIR/vbatched_tuning.jsonl (one JSON record per line: precision,
bucket, class or workgroups, median ms) and the winners are written as
DIR/vbatched_tuning.hpp, the header compiled into the library.
"""
import argparse
import json
import os
import time

import numpy as np

BUCKET_EDGES = [(1, 16), (17, 32), (33, 64), (65, 128), (129, 256), (257, 512)]


def timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    return float(np.median(ts))


def run_batch(torch, batched, m, n, k, prec, reps, gen):
    bufs, offs, ptrs = batched.pack_problems(torch, m, n, k, gen=gen)
    plan = batched.VBatchPlan(m, n, k, precision=prec)
    ms = timed(torch, lambda: plan.exec(1.0, ptrs[0], ptrs[1], 0.0, ptrs[2]), reps)
    flops = plan.info()["flops"]
    plan.close()
    del bufs
    return ms, flops


def header_text(tables, wgs):
    rows = []
    for name, t in (("FP32", tables["fp32"]), ("BF16", tables["bf16"])):
        body = ",\n".join("    {" + ", ".join(str(int(v)) for v in r) + "}" for r in t)
        rows.append("{  // %s\n%s,\n}" % (name, body))
    return (
        "// Tile-class table and persistent-grid size of the variable-size batched\n"
        "// gemm, written by mojoblas_amd/tune_vbatched.py from measurements on an\n"
        "// MI355X.  Indexed [bucket(m)][bucket(n)], bucket 0: <=16, 1: <=32, 2: <=64,\n"
        "// 3: <=128, 4: <=256, 5: larger; class 0 = 64x64 tiles, 1 = 128x128 tiles.\n"
        "#pragma once\n\n#define VB_BUCKETS 6\n\n"
        "// [0] = MBH_PREC_FP32, [1] = MBH_PREC_BF16\n"
        "static const int VB_DEFAULT_WGS_PER_CU[2] = {%d, %d};\n\n"
        "static const int VB_DEFAULT_CLASS[2][VB_BUCKETS][VB_BUCKETS] = {%s};\n"
        % (wgs["fp32"], wgs["bf16"], ", ".join(rows)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=256)
    args = ap.parse_args()
    import torch

    from mojoblas_amd import batched

    os.makedirs(args.out, exist_ok=True)
    log = open(os.path.join(args.out, "vbatched_tuning.jsonl"), "a")
    rng = np.random.default_rng(99)
    gen = torch.Generator(device="cuda").manual_seed(99)
    tables, wgs = {}, {}
    for prec in ("fp32", "bf16"):
        table = np.zeros((6, 6), np.int32)
        for bm, (mlo, mhi) in enumerate(BUCKET_EDGES):
            for bn, (nlo, nhi) in enumerate(BUCKET_EDGES):
                m = rng.integers(mlo, mhi + 1, args.batch)
                n = rng.integers(nlo, nhi + 1, args.batch)
                k = np.floor(np.exp(rng.uniform(np.log(16), np.log(512), args.batch))).astype(np.int64)
                best = None
                for cls in (0, 1):
                    batched.set_tuning(np.full((6, 6), cls), 4, prec)
                    ms, flops = run_batch(torch, batched, m, n, k, prec, args.reps, gen)
                    rec = {"ts": time.time(), "step": "class", "precision": prec, "bucket_m": bm, "bucket_n": bn,
                           "class": cls, "ms": ms, "tflops": flops / ms / 1e9}
                    log.write(json.dumps(rec) + "\n")
                    log.flush()
                    if best is None or ms < best[0]:
                        best = (ms, cls)
                table[bm, bn] = best[1]
                print("%s bucket (%d,%d): class %d" % (prec, bm, bn, best[1]), flush=True)
        m, n, k = batched.config4_shapes()
        best = None
        gt64 = np.zeros((6, 6), np.int32)
        gt64[3:, 3:] = 1
        gt128 = np.zeros((6, 6), np.int32)
        gt128[4:, 4:] = 1
        for name, cand in (("tuned", table), ("all64", np.zeros((6, 6), np.int32)), ("mn>64", gt64),
                           ("mn>128", gt128), ("all128", np.ones((6, 6), np.int32))):
            for w in (1, 2, 3, 4, 5, 6, 8):
                batched.set_tuning(cand, w, prec)
                ms, flops = run_batch(torch, batched, m, n, k, prec, args.reps, gen)
                rec = {"ts": time.time(), "step": "wgs_per_cu", "precision": prec, "table": name, "wgs_per_cu": w,
                       "ms": ms, "tflops": flops / ms / 1e9}
                log.write(json.dumps(rec) + "\n")
                log.flush()
                print("%s config batch, %s table, %d workgroups/CU: %.3f ms (%.1f TFLOP/s)"
                      % (prec, name, w, ms, flops / ms / 1e9), flush=True)
                if best is None or ms < best[0]:
                    best = (ms, w, cand)
        tables[prec] = best[2]
        wgs[prec] = best[1]
        batched.set_tuning(None, 0, prec)
    with open(os.path.join(args.out, "vbatched_tuning.hpp"), "w") as f:
        f.write(header_text(tables, wgs))
    print("wrote", os.path.join(args.out, "vbatched_tuning.hpp"))


if __name__ == "__main__":
    main()
