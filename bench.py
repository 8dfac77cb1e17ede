#!/usr/bin/env python3
"""Benchmark: batched MIRK4 solves of 16,384 seeded Bratu problems on one MI355X.

A step is one full batched solve (Newton + defect control + mesh refinement,
persistent kernel) of all instances, inputs resident in device memory.  Prints
one JSON line with the throughput in BVP solves/s (successful solves with
defect <= 1e-6), the roofline of the solve kernel, the oracle timed on the
host cores, and the plain one-instance-at-a-time GPU rate.
"""
import argparse
import json
import os
import pathlib
import re
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
from __graft_entry__ import load_package  # noqa: E402

METRIC = "BVP solves/s at defect≤1e-6: batched+learned+mixed vs plain FP64 GPU MIRK"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(M, counters, n0_nodes, nf_nodes, p_j=8):
    """Bytes of the streaming model: per interval and Newton iteration
    40M + 16 + 4 M^2 p_J; per convergence-check pass 16M + 16; per defect
    interval 8M + 24; per transfer node 8(M+1); plus guess in and solution out."""
    it, dfct, tr, chk = counters
    b = it * (40 * M + 16 + 4 * M * M * p_j)
    b += chk * (16 * M + 16)
    b += dfct * (8 * M + 24)
    b += tr * 8 * (M + 1)
    b += 8 * M * n0_nodes + 8 * (M + 1) * nf_nodes
    return b


def read_traffic(workload):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    for p in sorted((ROOT / "profiles").glob("*.md")):
        txt = p.read_text()
        if workload in txt:
            m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", txt)
            if m:
                return float(m.group(1))
    return None


def lintanh_flops(M, counters, refine, s=5):
    """FLOPs of the block path per launch: per interval-iteration the tangent GEMMs of
    stages 2..s-1 (4 M^3 each), the pivoted elimination of the 2M x (3M+1) stacked
    rows and the FP64 residual; per refinement step a J*v product and two sweeps."""
    it, dfct, tr, chk = counters
    elim = sum(2 * (2 * M - j - 1) * (3 * M + 1 - j) for j in range(M))
    per_it = 4 * M ** 3 * (s - 2) + elim + 2 * s * 2 * M * M
    per_ref = 2 * s * 2 * M * M * 2 + 2 * 2 * M * 2 * M + 2 * 3 * M * M
    per_chk = 2 * s * 2 * M * M
    per_def = 2 * 9 * 2 * M * M
    return it * (per_it + refine * per_ref) + chk * per_chk + dfct * per_def


def bench_lintanh(args, _abi, _lib, synthetic):
    n = args.instances if args.instances != 16384 else 2048
    M = 32
    specs, Ns = synthetic.lintanh_batch(n, M=M)
    opts = _abi.default_opts(_abi.MIRK6)
    opts.abstol = 1e-6
    opts.max_num_subintervals = 8192
    opts.precision = _abi.PREC_FP32_REFINE64
    opts.refine_iters = 2
    ctx = _lib.context(0)
    log(f"[bench] lintanh: uploading {n} instances, M={M} (sum N = {int(Ns.sum())})")
    batch = _lib.Batch(ctx, opts, specs)
    log(f"[bench] persistent waves {batch.info()[0]}")
    for w in range(args.warmup):
        batch.run()
        log(f"[bench] warmup {w}: kernel {batch.sync():.1f} ms")
    kms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.run()
        kms.append(batch.sync())
        log(f"[bench] step {k}: kernel {kms[-1]:.1f} ms")
    t1 = time.perf_counter()
    res = batch.results()
    counters = batch.counters()
    good = sum(1 for r in res if r.retcode == 0 and r.defect_norm <= 1e-6)
    wall = (t1 - t0) / args.steps
    kmean = sum(kms) / len(kms)
    flops = lintanh_flops(M, counters, opts.refine_iters)
    tfs = flops / (kmean * 1e-3) / 1e12
    log(f"[bench] {good}/{n} solved; {good / wall:.1f} solves/s; {tfs:.2f} TFLOP/s; counters {counters}; "
        f"newton iters mean {sum(r.newton_iters for r in res) / n:.2f}")
    print(json.dumps({"workload": f"lintanh{n}: {n} BVPs u'=Au+0.1 tanh(u), n=32, MIRK6, FP32 factors + FP64 "
                                  "refinement, abstol 1e-6", "value": good / wall, "unit": "solves/s",
                      "ms_per_step": wall * 1e3, "kernel_ms": kmean, "solved": good, "tflops": tfs,
                      "peak_fp32_mfma_tflops": 157.3, "counters": counters}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=32)
    ap.add_argument("--workload", default="bratu", choices=["bratu", "lintanh"],
                    help="bratu: 16,384 Bratu BVPs (MIRK4, FP64); lintanh: 2,048 n=32 linear+tanh BVPs "
                         "(MIRK6, FP32 factors + FP64 refinement)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); instances shard across "
              "GPUs with no exchange, so the whole-job rate is the per-GPU rate times the GPU count",
              file=sys.stderr)
        sys.exit(2)

    bvp = load_package()
    from boundaryvaluediffeq_amd import _abi, _lib, synthetic

    if args.copy_probe:
        ms = _lib.copy_probe(_lib.context(0), args.copy_probe, repeats=args.steps)
        gbs = 2 * args.copy_probe / (ms * 1e-3) / 1e9
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "ms": ms, "GB_per_s_read_plus_write": gbs}))
        return

    if args.workload == "lintanh":
        return bench_lintanh(args, _abi, _lib, synthetic)
    n = args.instances
    abstol = 1e-8
    specs, lams, Ns = synthetic.bratu_batch(n)
    workload = f"bratu{n // 1024}k"
    opts = _abi.default_opts(_abi.MIRK4)
    opts.abstol = abstol
    ctx = _lib.context(0)
    log(f"[bench] uploading {n} instances (sum N = {int(Ns.sum())})")
    batch = _lib.Batch(ctx, opts, specs)
    info = batch.info()
    log(f"[bench] persistent waves {info[0]}, node capacity {info[1]}")
    for w in range(args.warmup):
        batch.run()
        ms = batch.sync()
        log(f"[bench] warmup {w}: kernel {ms:.2f} ms")
    kms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        batch.run()
        kms.append(batch.sync())
    t1 = time.perf_counter()
    log(f"[bench] timed {args.steps} steps: kernel ms {['%.2f' % x for x in kms]}")
    res = batch.results()
    counters = batch.counters()
    good = sum(1 for r in res if r.retcode == 0 and r.defect_norm <= min(1e-6, abstol))
    wall = (t1 - t0) / args.steps
    value = good / wall
    kmean = sum(kms) / len(kms)
    n0_nodes = int(sum(s.N + 1 for s in specs))
    nf_nodes = int(sum(r.N + 1 for r in res))
    abytes = algorithmic_bytes(2, counters, n0_nodes, nf_nodes)
    achieved = abytes / (kmean * 1e-3) / 1e9
    log(f"[bench] {good}/{n} solved; {value:.1f} solves/s; kernel {kmean:.2f} ms; "
        f"algorithmic {abytes / 1e9:.3f} GB -> {achieved:.1f} GB/s; counters {counters}")

    # plain form: one instance per launch, FP64, constant guess
    plain = None
    if args.plain_sample > 0:
        stride = max(1, n // args.plain_sample)
        idx = list(range(0, n, stride))[:args.plain_sample]
        tot = 0.0
        ok = 0
        for i in idx:
            b1 = _lib.Batch(ctx, opts, [specs[i]])
            b1.run()
            b1.sync()
            ta = time.perf_counter()
            b1.run()
            b1.sync()
            tot += time.perf_counter() - ta
            r = b1.results()[0]
            ok += int(r.retcode == 0 and r.defect_norm <= min(1e-6, abstol))
            b1.close()
        plain = {"value": ok / tot, "unit": "solves/s", "sample": f"{len(idx)} instances (every {stride}th), "
                 "each solved alone in its own launch", "speedup_batched_over_plain": value / (ok / tot)}
        log(f"[bench] plain single-instance GPU: {ok / tot:.1f} solves/s")

    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as O
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        oo = O.opts(4, abstol=abstol)
        log(f"[bench] oracle on {threads} threads, {n} instances")
        tc = time.perf_counter()
        rets = O.solve_batch(oo, specs, nthreads=threads, keep_solutions=False)
        tcpu = time.perf_counter() - tc
        okc = sum(1 for r in rets if r == 0)
        cpu = {"value": okc / tcpu, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": f"the same {n} instances, one per OpenMP thread (dynamic schedule), {tcpu:.1f} s wall"}
        log(f"[bench] oracle: {okc / tcpu:.1f} solves/s")

    traffic = read_traffic("k_solve")
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{workload}: {n} Bratu BVPs u''+lambda e^u=0, lambda~U(0.1,3.4), "
                               "N~U{64..1024}, zero guess, MIRK4, abstol 1e-8, adaptive",
                   "instances": n, "method": "MIRK4", "precision": "fp64", "variant": "batched persistent",
                   "solved": good},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_solve<2,4>", "kernel_ms": kmean, "algorithmic_bytes": abytes},
        "cpu_baseline": cpu,
        "plain_gpu": plain,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
