#!/usr/bin/env python3
"""Benchmark: batched MIRK4 solves of 16,384 seeded Bratu problems on one MI355X.

A step is one full batched solve (learned guess of every node, then Newton +
defect control + mesh refinement in one persistent kernel, replayed as a
hipGraph) of all instances, inputs resident in device memory, with the chunk
elimination factors stored in FP32 (mixed precision; residuals, elimination and
the Newton test in FP64).  The warm-start network is trained first (untimed) on
solves of the Bratu family with boundary values u(0)=a, u(1)=b that the plain
method itself generates.  Prints one JSON line with the throughput in BVP
solves/s (successful solves with defect <= 1e-6), the same batch in FP64 with
and without the network, the plain one-instance-at-a-time FP64 GPU rate, the
roofline of the solve kernel and the oracle timed on the host cores.  Other
workloads (--workload) print their own lines.
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
    """Bytes of the streaming model, per launch.  Per interval and Newton iteration:
    the dual pass reads y and the mesh (8M + 8) and writes the elimination record, the
    back-substitution reads the record, y and the mesh (the residual of the corrected
    state -- the next convergence check -- and its defect are evaluated there) and
    writes y (16M + 8): 24M + 16 + 2 p_J (2M^2 + M), p_J bytes per record value (8
    FP64, 4 FP32 records).  Per interval of a separate residual pass (line-search and
    trust-region stages) 16M + 16; per interval of a separate defect pass (when mesh
    selection follows) 8M + 16 (shat, 8 more bytes, not counted); per transfer node
    8(M+1); plus guess in and solution out."""
    it, dfct, tr, chk = counters
    b = it * (24 * M + 16 + 2 * p_j * (2 * M * M + M))
    b += chk * (16 * M + 16)
    b += dfct * (8 * M + 16)
    b += tr * 8 * (M + 1)
    b += 8 * M * n0_nodes + 8 * (M + 1) * nf_nodes
    return b


def read_traffic(workload, key="traffic_bytes_per_launch"):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    for p in sorted((ROOT / "profiles").glob("*.md")):
        txt = p.read_text()
        if workload in txt:
            m = re.search(key + r":\s*([0-9.eE+]+)", txt)
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
    batch.set_graph(not args.no_graph)
    log(f"[bench] persistent waves {batch.info()[0]}")
    tune = None
    if not args.no_tune:
        cand_ms, (variant, blocks) = batch.autotune(repeats=1)
        tune = {"candidates_ms": cand_ms, "persistent_waves": blocks}
        log(f"[bench] autotune: {['%.0f' % x for x in cand_ms]} ms -> {blocks} waves")
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
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        k = min(n, args.cpu_sample)
        sample = specs[:: max(1, n // k)][:k]
        cpu = cpu_baseline(sample, opts.abstol, threads, method=6, max_num_subintervals=opts.max_num_subintervals)
    print(json.dumps({"workload": f"lintanh{n}: {n} BVPs u'=Au+0.1 tanh(u), n=32, MIRK6, FP32 factors + FP64 "
                                  "refinement, abstol 1e-6", "value": good / wall, "unit": "solves/s",
                      "ms_per_step": wall * 1e3, "kernel_ms": kmean, "solved": good, "tflops": tfs,
                      "peak_fp32_mfma_tflops": 157.3, "counters": counters, "autotune": tune,
                      "cpu_baseline": cpu}), flush=True)


def timed_runs(batch, steps, warmup, tag):
    for w in range(warmup):
        batch.run()
        log(f"[bench] {tag} warmup {w}: {batch.sync():.2f} ms")
    kms, gms, sms = [], [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.run()
        kms.append(batch.sync())
        g, s_ = batch.timing()
        gms.append(g)
        sms.append(s_)
    wall = (time.perf_counter() - t0) / steps
    log(f"[bench] {tag}: run ms {['%.3f' % x for x in kms]} (guess {sum(gms) / steps:.3f}, "
        f"solve {sum(sms) / steps:.3f}); wall {wall * 1e3:.3f} ms/step")
    return dict(run_ms=sum(kms) / steps, guess_ms=sum(gms) / steps, solve_ms=sum(sms) / steps, wall=wall)


def n_solved(res, abstol):
    return sum(1 for r in res if r.retcode == 0 and r.defect_norm <= min(1e-6, abstol))


def train_warm_start(ctx, args, opts, _abi, synthetic, learned):
    """The warm-start network, trained on plain solves of the Bratu family with
    u(0)=a, u(1)=b (lambda ~ U(0.1, 3.4), a, b ~ U(-1, 1)); untimed."""
    specs, _, _ = synthetic.bratu_ab_batch(args.train_solves, seed=synthetic.SEED_BRATU_AB_TRAIN)
    hopts = _abi.Opts.from_buffer_copy(opts)
    hopts.max_newton_iters = 50  # the box has instances without a solution; they are left out anyway
    t0 = time.perf_counter()
    x, y, sols = learned.harvest(ctx, specs, hopts, per_solve=args.train_nodes)
    th = time.perf_counter() - t0
    solved = sum(1 for r in sols if r.retcode == 0)
    log(f"[bench] training data: {len(specs)} plain solves ({solved} solved) -> {len(x)} samples in {th:.1f} s")
    t0 = time.perf_counter()
    net, loss, ms = learned.train(ctx, x, y, epochs=args.train_epochs, batch=8192, lr=args.train_lr, lr_min=2e-5,
                                  warmup=200, use_graph=not args.no_graph, chunk_epochs=5,
                                  progress=lambda e, l: log(f"[bench] training: epoch {e}, loss {l:.3e}"))
    log(f"[bench] training: {len(loss)} Adam steps, loss {loss[0]:.3e} -> {loss[-1]:.3e}, device {ms:.0f} ms, "
        f"wall {time.perf_counter() - t0:.1f} s")
    return net, {"train_solves": len(specs), "train_solved": solved, "samples": int(len(x)),
                 "nodes_per_solve": args.train_nodes, "adam_steps": int(len(loss)), "epochs": args.train_epochs,
                 "final_loss": float(loss[-1]), "train_device_ms": ms, "harvest_s": th}


def plain_single(ctx, opts, specs, sample, abstol, _lib):
    """The plain form: one instance per launch, FP64 factors, constant guess."""
    n = len(specs)
    stride = max(1, n // sample)
    idx = list(range(0, n, stride))[:sample]
    tot, ok = 0.0, 0
    for i in idx:
        b1 = _lib.Batch(ctx, opts, [specs[i]])
        b1.set_graph(False)
        b1.run()
        b1.sync()
        ta = time.perf_counter()
        b1.run()
        b1.sync()
        tot += time.perf_counter() - ta
        ok += n_solved(b1.results(), abstol)
        b1.close()
    log(f"[bench] plain single-instance GPU: {ok / tot:.1f} solves/s")
    return {"value": ok / tot, "unit": "solves/s",
            "sample": f"{len(idx)} instances (every {stride}th), each solved alone in its own launch"}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(specs, abstol, threads, method=4, **kw):
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    oo = O.opts(method, abstol=abstol, **kw)
    log(f"[bench] oracle on {threads} threads, {len(specs)} instances")
    tc = time.perf_counter()
    rets = O.solve_batch(oo, specs, nthreads=threads, keep_solutions=False)
    tcpu = time.perf_counter() - tc
    okc = sum(1 for r in rets if r == 0)
    log(f"[bench] oracle: {okc / tcpu:.1f} solves/s")
    return {"value": okc / tcpu, "unit": "solves/s", "cores": threads, "cpu_model": cpu_model(), "kind": "port",
            "sample": f"{len(specs)} of the same instances, one per OpenMP thread (dynamic schedule), "
                      f"{tcpu:.1f} s wall"}


def error_vs_oracle(specs, gpu_res, abstol, threads, method=4, chunk=2048, **kw):
    """Untimed: the oracle's solutions of the same instances (same tolerances), compared
    with the GPU's at every mesh point, relative to max(1, |u|).  Where the final meshes
    differ, the distance is the oracle's own Newton correction at the GPU's state on the
    GPU's mesh (the distance to the discrete solution there)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import pyoracle as O
    oo = O.opts(method, abstol=abstol, **kw)
    emax, same, other, retc_diff = 0.0, 0, 0, 0
    for c0 in range(0, len(specs), chunk):
        part = specs[c0:c0 + chunk]
        rs = O.solve_batch(oo, part, nthreads=threads)
        for s, r, g in zip(part, rs, gpu_res[c0:c0 + chunk]):
            if r.retcode != g.retcode:
                retc_diff += 1
            if r.retcode != 0 or g.retcode != 0:
                continue
            if r.N == g.N and np.max(np.abs(r.t - g.t)) <= 1e-12:
                e = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
                same += 1
            else:
                d = O.newton_step(method, s.with_state(g.t, g.u.reshape(-1))).reshape(-1, s.M)
                e = float(np.max(np.abs(d) / np.maximum(1.0, np.abs(g.u))))
                other += 1
            emax = max(emax, e)
    return {"max_rel_err_vs_oracle": emax, "compared_same_mesh": same, "compared_on_gpu_mesh": other,
            "retcode_mismatches": retc_diff, "tolerance": 1e-9, "within_tolerance": emax <= 1e-9}


def pendulum_config1(ctx, _abi, _lib, runs=200):
    """The README pendulum (MIRK4, dt = 0.05, abstol 1e-6), one instance: GPU latency per
    solve (plan replay, HIP events and wall clock, median of `runs`) and the oracle's
    single-thread median over `runs` solves."""
    import statistics
    import math
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    bvp = load_package()
    prob = bvp.BVProblem(bvp.PendulumFunction(), [(math.pi / 4, 0, -math.pi / 2), (math.pi / 2, 0, math.pi / 2)],
                         [math.pi / 2, math.pi / 2], (0.0, math.pi / 2), [9.81])
    spec = bvp.extract(prob, dt=0.05)
    o = _abi.default_opts(_abi.MIRK4)
    b = _lib.Batch(ctx, o, [spec])
    b.run()
    b.sync()
    dev, wall = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        b.run()
        dev.append(b.sync())
        wall.append(time.perf_counter() - t0)
    r = b.results()[0]
    b.close()
    oo = O.opts(4)
    O.solve(oo, spec)
    cpu = []
    for _ in range(runs):
        t0 = time.perf_counter()
        rc = O.solve(oo, spec)
        cpu.append(time.perf_counter() - t0)
    return {"workload": "README pendulum, MIRK4, dt=0.05 (N=32), abstol 1e-6, single instance",
            "gpu_us_per_solve_device": statistics.median(dev) * 1e3,
            "gpu_us_per_solve_wall": statistics.median(wall) * 1e6, "gpu_retcode": r.retcode, "gpu_N": r.N,
            "cpu_oracle_us_per_solve_1thread": statistics.median(cpu) * 1e6, "cpu_retcode": rc.retcode,
            "runs": runs}


def bench_learned(args, _abi, _lib, synthetic, learned):
    """Learned warm start on 16,384 unseen Bratu (a, b) instances: solves/s,
    Newton iterations and fallback rate with and without the network."""
    ctx = _lib.context(0)
    opts = _abi.default_opts(_abi.MIRK4)
    opts.abstol = 1e-8
    net, tinfo = train_warm_start(ctx, args, opts, _abi, synthetic, learned)
    specs, _, _ = synthetic.bratu_ab_batch(args.instances, seed=synthetic.SEED_BRATU_AB_EVAL)
    batch = _lib.Batch(ctx, opts, specs)
    tp = timed_runs(batch, args.steps, args.warmup, "plain guess")
    rp = batch.results()
    batch.set_warm_start(net, _abi.FEAT_PARAM, args.anchors)
    tw = timed_runs(batch, args.steps, args.warmup, "learned guess")
    rw = batch.results()
    sp, sw = learned.summary(rp), learned.summary(rw)
    log(f"[bench] plain {sp}\n[bench] learned {sw}")
    print(json.dumps({"workload": f"bratu_ab{args.instances}: {args.instances} unseen Bratu BVPs with u(0)=a, "
                                  "u(1)=b, lambda~U(0.1,3.4), a,b~U(-1,1), MIRK4, abstol 1e-8",
                      "plain_guess": {"solves_per_s": n_solved(rp, 1e-8) / tp["wall"], **tp, **sp},
                      "learned_guess": {"solves_per_s": n_solved(rw, 1e-8) / tw["wall"], **tw, **sw},
                      "newton_iters_saved_per_solved_instance":
                          sp["mean_newton_iters_solved"] - sw["mean_newton_iters_solved"],
                      "training": tinfo}), flush=True)


def bench_singpert(args, _abi, _lib, synthetic, learned):
    """Adaptive batch: 8,192 problems eps u'' + t u' - u = 0 on [-1, 1], u(-1)=1, u(1)=2,
    eps ~ LogUniform(1e-4, 1e-1), initial mesh 32 intervals, refinement to <= 16,384
    intervals at abstol 1e-6, one persistent launch; with and without a warm-start
    network trained on solves of the same family (features log10 eps, t)."""
    n = args.instances if args.instances != 16384 else 8192
    ctx = _lib.context(0)
    opts = _abi.default_opts(_abi.MIRK4)
    opts.abstol = 1e-6
    opts.max_num_subintervals = 16384
    specs, eps = synthetic.singpert_batch(n)
    batch = _lib.Batch(ctx, opts, specs)
    batch.set_graph(not args.no_graph)
    tune = None
    if not args.no_tune:
        cand_ms, (variant, blocks) = batch.autotune(repeats=1)
        tune = {"candidates_ms": cand_ms, "variant": variant, "persistent_waves": blocks}
        log(f"[bench] autotune: {['%.1f' % x for x in cand_ms]} ms -> variant {variant}, {blocks} waves")
    tp = timed_runs(batch, args.steps, args.warmup, "constant guess")
    rp = batch.results()
    cp = batch.counters()
    out = {"workload": f"singpert{n}: {n} BVPs eps u''+t u'-u=0, eps~LogUniform(1e-4,1e-1), N0=32, adaptive to "
                       "<=16384 intervals, MIRK4, abstol 1e-6, FP64",
           "constant_guess": {"solves_per_s": n_solved(rp, 1e-6) / tp["wall"], **tp, **learned.summary(rp),
                              "final_N_mean": float(sum(r.N for r in rp) / n),
                              "final_N_max": int(max(r.N for r in rp)), "counters": cp},
           "autotune": tune}
    if not args.no_learned:
        tspecs, _ = synthetic.singpert_batch(args.train_solves // 10, seed=synthetic.SEED_SINGPERT_TRAIN)
        t0 = time.perf_counter()
        x, y, sols = learned.harvest(ctx, tspecs, opts, per_solve=args.train_nodes, kind=_abi.FEAT_LOG10_PARAM,
                                     chunk=4096)
        log(f"[bench] training data: {len(tspecs)} solves -> {len(x)} samples in {time.perf_counter() - t0:.1f} s")
        net, loss, ms = learned.train(ctx, x, y, epochs=args.train_epochs, batch=8192, lr=args.train_lr,
                                      lr_min=2e-5, warmup=200, use_graph=not args.no_graph, chunk_epochs=5,
                                      progress=lambda e, l: log(f"[bench] training: epoch {e}, loss {l:.3e}"))
        log(f"[bench] training: {len(loss)} steps, loss {loss[0]:.3e} -> {loss[-1]:.3e}, device {ms:.0f} ms")
        batch.set_warm_start(net, _abi.FEAT_LOG10_PARAM, 0)  # 33 nodes per instance: the network at each
        tw = timed_runs(batch, args.steps, args.warmup, "learned guess")
        rw = batch.results()
        out["learned_guess"] = {"solves_per_s": n_solved(rw, 1e-6) / tw["wall"], **tw, **learned.summary(rw),
                                "final_N_mean": float(sum(r.N for r in rw) / n), "counters": batch.counters()}
        out["training"] = {"solves": len(tspecs), "samples": int(len(x)), "adam_steps": int(len(loss)),
                           "final_loss": float(loss[-1]), "device_ms": ms}
    # the same batch with FP32 elimination records (FP64 residual, defect and Newton
    # test): constant guess, then with the network (the combination this workload names)
    if not args.no_mixed:
        omx = _abi.default_opts(_abi.MIRK4)
        omx.abstol = 1e-6
        omx.max_num_subintervals = 16384
        omx.precision = _abi.PREC_FP32_REFINE64
        batch.close()
        batch = _lib.Batch(ctx, omx, specs)
        batch.set_graph(not args.no_graph)
        if not args.no_tune:
            cand_ms, (variant, blocks) = batch.autotune(repeats=1)
            out["autotune_mixed"] = {"candidates_ms": cand_ms, "variant": variant,
                                     "records": "fp32" if batch.precision() == _abi.PREC_FP32_REFINE64 else "fp64"}
            log(f"[bench] autotune (mixed): {['%.1f' % x for x in cand_ms]} ms -> variant {variant}, "
                f"{out['autotune_mixed']['records']} records")
        tm = timed_runs(batch, args.steps, args.warmup, "mixed, constant guess")
        rm = batch.results()
        out["mixed_constant_guess"] = {"solves_per_s": n_solved(rm, 1e-6) / tm["wall"], **tm, **learned.summary(rm),
                                       "final_N_mean": float(sum(r.N for r in rm) / n), "counters": batch.counters()}
        if not args.no_learned:
            batch.set_warm_start(net, _abi.FEAT_LOG10_PARAM, 0)
            tl = timed_runs(batch, args.steps, args.warmup, "mixed, learned guess")
            rl = batch.results()
            out["mixed_learned_guess"] = {"solves_per_s": n_solved(rl, 1e-6) / tl["wall"], **tl,
                                          **learned.summary(rl), "final_N_mean": float(sum(r.N for r in rl) / n),
                                          "counters": batch.counters()}
    print(json.dumps(out), flush=True)


def bench_firk(args, _abi, _lib, synthetic, learned):
    """FIRK collocation on the headline's 16,384 Bratu instances (the expanded form of
    lib/BoundaryValueDiffEqFIRK; RadauIIa3 unless --firk-method), abstol 1e-8, the
    headline's Newton tolerance, constant zero guess, FP64, one persistent launch; the
    oracle's FIRK restatement on the host cores as the CPU baseline and the checker."""
    import numpy as np
    n = args.instances
    code = args.firk_method
    s_ = code % 100
    abstol = 1e-8
    specs, lams, Ns = synthetic.bratu_batch(n)
    ctx = _lib.context(0)
    opts = _abi.default_opts(code)
    opts.abstol = abstol
    opts.newton_abstol = args.newton_abstol
    opts.nlsolve_stages = _abi.STAGE_NEWTON | _abi.STAGE_BACKTRACKING
    batch = _lib.Batch(ctx, opts, specs)
    batch.set_graph(not args.no_graph)
    tune = None
    if not args.no_tune:
        cand_ms, (variant, blocks) = batch.autotune(repeats=1)
        tune = {"candidates_ms": cand_ms, "persistent_waves": blocks}
        log(f"[bench] autotune: {['%.2f' % x for x in cand_ms]} ms -> {blocks} waves")
    t = timed_runs(batch, args.steps, args.warmup, f"firk {code}")
    res = batch.results()
    cnt = batch.counters()
    batch.close()
    ok = n_solved(res, abstol)
    value = ok / (t["solve_ms"] * 1e-3)
    out = {"workload": f"bratu{n // 1024}k, FIRK method {code} ({s_} stages, {'nested' if code >= _abi.FIRK_NESTED else 'expanded'} "
                       f"form): {n} Bratu BVPs, "
                       f"lambda~U(0.1,3.4), N~U{{64..1024}}, abstol {abstol}, Newton tolerance {args.newton_abstol}, "
                       "constant zero guess, FP64",
           "method": code, "value": value, "unit": "solves/s (device solve time)", "solve_ms": t["solve_ms"],
           "wall_ms_per_step": t["wall"] * 1e3, "solved": ok, **learned.summary(res),
           "final_N_mean": float(sum(r.N for r in res) / n), "counters": cnt, "autotune": tune}
    log(f"[bench] firk {code}: {ok}/{n} solved; {value:.1f} solves/s; solve {t['solve_ms']:.3f} ms")
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as O
        threads = args.cpu_threads or os.cpu_count() or 1
        oo = O.opts(code, abstol=abstol, newton_abstol=args.newton_abstol, nlsolve_stages=3)
        tc = time.perf_counter()
        rs = O.solve_batch(oo, specs, nthreads=threads, keep_solutions=not args.no_accuracy)
        tcpu = time.perf_counter() - tc
        okc = sum(1 for r in rs if (r if isinstance(r, int) else r.retcode) == 0)
        out["cpu_baseline"] = {"value": okc / tcpu, "unit": "solves/s", "cores": threads, "cpu_model": cpu_model(),
                               "kind": "port", "sample": f"the same {n} instances, {tcpu:.1f} s wall"}
        log(f"[bench] oracle: {okc / tcpu:.1f} solves/s on {threads} threads")
        if not args.no_accuracy:
            emax, same, other, rdiff = 0.0, 0, 0, 0
            for sp, r, g in zip(specs, rs, res):
                if r.retcode != g.retcode:
                    rdiff += 1
                if r.retcode != 0 or g.retcode != 0:
                    continue
                if r.N == g.N and np.max(np.abs(r.t - g.t)) <= 1e-12:
                    e = float(np.max(np.abs(g.u - r.u) / np.maximum(1.0, np.abs(r.u))))
                    same += 1
                else:
                    st = sp.with_state(g.t, g.u.reshape(-1))
                    d = O.firk_newton_step(code, st, g.stages)
                    e = float(np.max(np.abs(d)) / max(1.0, float(np.max(np.abs(g.u)))))
                    other += 1
                emax = max(emax, e)
            out["accuracy"] = {"max_rel_err_vs_oracle": emax, "compared_same_mesh": same,
                               "compared_on_gpu_mesh": other, "retcode_mismatches": rdiff, "tolerance": 1e-9,
                               "within_tolerance": emax <= 1e-9}
            log(f"[bench] accuracy vs oracle: {out['accuracy']}")
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=16384)
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all host cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-sample", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=128, help="instances of the oracle run (lintanh workload)")
    ap.add_argument("--no-learned", action="store_true", help="skip the warm-start network (constant guess only)")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (variant 0 at full residency)")
    ap.add_argument("--no-mixed", action="store_true", help="headline run with FP64 factor storage")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch kernels directly instead of replaying captured graphs (for counter profiling)")
    ap.add_argument("--train-solves", type=int, default=200000)
    ap.add_argument("--train-nodes", type=int, default=64, help="training samples per solve")
    ap.add_argument("--train-epochs", type=int, default=40)
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--save-net", default="", help="write the trained network to this .npz")
    ap.add_argument("--load-net", default="", help="use the network in this .npz instead of training one")
    ap.add_argument("--anchors", type=int, default=4,
                    help="network evaluations per instance (Hermite-interpolated to the nodes); 0: every node")
    ap.add_argument("--firk-method", type=int, default=103,
                    help="FIRK method code of --workload firk (100 + s Radau IIA, 200/300/400 + s Lobatto IIIA/B/C; "
                         "+ 1000: the nested form)")
    ap.add_argument("--workload", default="bratu", choices=["bratu", "lintanh", "learned", "singpert", "firk"],
                    help="bratu: 16,384 Bratu BVPs (MIRK4, FP64, learned warm start); lintanh: 2,048 n=32 "
                         "linear+tanh BVPs (MIRK6, FP32 factors + FP64 refinement); learned: warm start on "
                         "16,384 unseen Bratu (a, b) BVPs")
    ap.add_argument("--newton-abstol", type=float, default=None,
                    help="Newton tolerance max|F| of every leg (GPU and oracle alike); 0: abstol, the reference's "
                         "default")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the untimed accuracy legs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); instances shard across GPUs with no "
              "exchange, so the whole-job rate is the per-GPU rate times the GPU count", file=sys.stderr)
        sys.exit(2)

    pkg = load_package()
    from importlib import import_module
    _abi = import_module(pkg.__name__ + "._abi")
    _lib = import_module(pkg.__name__ + "._lib")
    synthetic = import_module(pkg.__name__ + ".synthetic")
    learned = import_module(pkg.__name__ + ".learned")
    if args.newton_abstol is None:
        # every leg, GPU and oracle alike, stops Newton there (tighter than the reference's
        # default, abstol), so that two solves of one instance agree to 1e-9 relative
        args.newton_abstol = synthetic.HEADLINE_NEWTON_ABSTOL

    if args.copy_probe:
        ms = _lib.copy_probe(_lib.context(0), args.copy_probe, repeats=args.steps)
        gbs = 2 * args.copy_probe / (ms * 1e-3) / 1e9
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "ms": ms, "GB_per_s_read_plus_write": gbs}))
        return
    if args.workload == "lintanh":
        return bench_lintanh(args, _abi, _lib, synthetic)
    if args.workload == "learned":
        return bench_learned(args, _abi, _lib, synthetic, learned)
    if args.workload == "singpert":
        return bench_singpert(args, _abi, _lib, synthetic, learned)
    if args.workload == "firk":
        return bench_firk(args, _abi, _lib, synthetic, learned)

    n = args.instances
    abstol = 1e-8
    specs, lams, Ns = synthetic.bratu_batch(n)
    workload = f"bratu{n // 1024}k"
    opts = _abi.default_opts(_abi.MIRK4)
    opts.abstol = abstol
    opts.newton_abstol = args.newton_abstol
    ctx = _lib.context(0)
    log(f"[bench] uploading {n} instances (sum N = {int(Ns.sum())})")
    batch = _lib.Batch(ctx, opts, specs)
    batch.set_graph(not args.no_graph)
    info = batch.info()
    log(f"[bench] persistent waves {info[0]}, node capacity {info[1]}")
    n0_nodes = int(sum(s.N + 1 for s in specs))
    tune = None
    if not args.no_tune:  # untimed: pick the launch configuration by measurement
        cand_ms, (variant, blocks) = batch.autotune(repeats=3)
        tune = {"candidates_ms": cand_ms, "variant": variant, "persistent_waves": blocks,
                "variants": "solve kernel register allocation held to 1, 2, 3 waves per SIMD, then at least 2, 4, "
                            "8 intervals per lane for the fastest", "min_intervals_per_lane": batch.chunking()}
        log(f"[bench] autotune: candidates {['%.3f' % x for x in cand_ms]} ms -> variant {variant}, {blocks} waves")

    # batched, constant guess
    ta = timed_runs(batch, args.steps, args.warmup, "batched fp64, constant guess")
    res_a = batch.results()
    cnt_a = batch.counters()
    good_a = n_solved(res_a, abstol)
    # batched, learned guess (the headline form)
    tinfo = None
    tb, res_b, cnt_b, good_b = ta, res_a, cnt_a, good_a
    if not args.no_learned:
        if args.load_net:  # a network saved by an earlier run (profiling runs skip the training)
            import numpy as np
            z = np.load(args.load_net)
            net = _lib.MLP(ctx)
            net.set_params(z["params"])
            net.set_normalization(z["in_shift"], z["in_scale"], z["out_shift"], z["out_scale"])
            tinfo = {"loaded": args.load_net}
        else:
            net, tinfo = train_warm_start(ctx, args, opts, _abi, synthetic, learned)
        if args.save_net:
            import numpy as np
            np.savez(args.save_net, params=net.params(), in_shift=net.norm[0], in_scale=net.norm[1],
                     out_shift=net.norm[2], out_scale=net.norm[3])
        batch.set_warm_start(net, _abi.FEAT_PARAM, args.anchors)
        tb = timed_runs(batch, args.steps, args.warmup, "batched fp64, learned guess")
        res_b = batch.results()
        cnt_b = batch.counters()
        good_b = n_solved(res_b, abstol)
    fp64_learned = None if args.no_learned else {"value": good_b / tb["wall"], "ms_per_step": tb["wall"] * 1e3,
                                                 "guess_ms": tb["guess_ms"], "solve_ms": tb["solve_ms"],
                                                 "solved": good_b}
    # batched + learned + mixed (the headline form): the same plan with the chunk
    # factors stored in FP32; residuals and the convergence test stay FP64
    p_j = 8
    tune_mx = None
    if not args.no_mixed:
        omx = _abi.default_opts(_abi.MIRK4)
        omx.abstol = abstol
        omx.newton_abstol = args.newton_abstol
        omx.precision = _abi.PREC_FP32_REFINE64
        batch.close()
        batch = _lib.Batch(ctx, omx, specs)
        batch.set_graph(not args.no_graph)
        if not args.no_tune:
            cand_ms, (variant, blocks) = batch.autotune(repeats=3)
            tune_mx = {"candidates_ms": cand_ms, "variant": variant, "persistent_waves": blocks,
                       "candidates": "variants 0-2 with FP32 records, then 0-2 with FP64 records, then at least "
                                     "2, 4, 8 intervals per lane for the fastest",
                       "records": "fp32" if batch.precision() == _abi.PREC_FP32_REFINE64 else "fp64",
                       "min_intervals_per_lane": batch.chunking()}
            log(f"[bench] autotune (mixed): candidates {['%.3f' % x for x in cand_ms]} ms -> variant {variant}, "
                f"{tune_mx['records']} records")
        if not args.no_learned:
            batch.set_warm_start(net, _abi.FEAT_PARAM, args.anchors)
        tag = "batched mixed, " + ("constant" if args.no_learned else "learned") + " guess"
        tb = timed_runs(batch, args.steps, args.warmup, tag)
        res_b = batch.results()
        cnt_b = batch.counters()
        good_b = n_solved(res_b, abstol)
        p_j = 4
    value = good_b / tb["wall"]
    nf_nodes = int(sum(r.N + 1 for r in res_b))
    abytes = algorithmic_bytes(2, cnt_b, n0_nodes, nf_nodes, p_j=p_j)
    achieved = abytes / (tb["solve_ms"] * 1e-3) / 1e9
    sa, sb = learned.summary(res_a), learned.summary(res_b)
    log(f"[bench] {good_b}/{n} solved; {value:.1f} solves/s; solve kernel {tb['solve_ms']:.3f} ms; algorithmic "
        f"{abytes / 1e9:.3f} GB -> {achieved:.1f} GB/s; counters {cnt_b}; newton iters/instance "
        f"{sa['mean_newton_iters']:.2f} -> {sb['mean_newton_iters']:.2f}")

    # untimed accuracy of the headline results: the same plan configuration with
    # report_err (one more FP64 Newton correction per instance; the solution bits are
    # unchanged, which is checked), then the oracle's solutions of the same instances
    accuracy = None
    if not args.no_accuracy:
        import numpy as np
        oe = _abi.Opts.from_buffer_copy(batch.opts)
        oe.report_err = 1
        be = _lib.Batch(ctx, oe, specs)
        be.set_config(*batch.config())
        be.set_chunking(batch.chunking())
        if not args.no_learned:
            be.set_warm_start(net, _abi.FEAT_PARAM, args.anchors)
        be.run()
        be.sync()
        res_e = be.results()
        be.close()
        ok_e = [r.err_vs_fp64 for r in res_e if r.retcode == 0]
        accuracy = {"newton_abstol": args.newton_abstol or abstol,
                    "same_bits_as_timed_run": all(np.array_equal(a.u, c.u) for a, c in zip(res_b, res_e)),
                    "err_vs_fp64_max": float(max(ok_e)) if ok_e else None,
                    "err_vs_fp64_p99": float(np.percentile(ok_e, 99)) if ok_e else None,
                    "err_vs_fp64": "max_i |d_i|/max(1,|u_i|) of the FP64 Newton correction d at each returned "
                                   "solution, on the device"}
        log(f"[bench] accuracy: err_vs_fp64 max {accuracy['err_vs_fp64_max']:.3e}, same bits "
            f"{accuracy['same_bits_as_timed_run']}")

    # the same plan with Newton stopped where the reference stops it (max|F| <= abstol,
    # MIRK/mirk.jl:51): fewer iterations, but two solves of one instance then differ by up
    # to norm(J^-1) * abstol, so this figure cannot carry the 1e-9 agreement (timed the
    # same way, outside the headline's timed region)
    at_ref = None
    if not args.no_accuracy and args.newton_abstol and args.newton_abstol != abstol:
        orf = _abi.Opts.from_buffer_copy(batch.opts)
        orf.newton_abstol = abstol
        br = _lib.Batch(ctx, orf, specs)
        br.set_graph(not args.no_graph)
        br.set_config(*batch.config())
        br.set_chunking(batch.chunking())
        if not args.no_learned:
            br.set_warm_start(net, _abi.FEAT_PARAM, args.anchors)
        tr = timed_runs(br, args.steps, args.warmup, "same plan at the reference's Newton tolerance")
        res_r = br.results()
        br.close()
        sr = learned.summary(res_r)
        at_ref = {"newton_abstol": abstol, "value": n_solved(res_r, abstol) / tr["wall"],
                  "ms_per_step": tr["wall"] * 1e3, "solve_ms": tr["solve_ms"],
                  "mean_newton_iters": sr["mean_newton_iters"]}
        log(f"[bench] at the reference's Newton tolerance: {at_ref}")

    plain = plain_single(ctx, opts, specs, args.plain_sample, abstol, _lib) if args.plain_sample > 0 else None
    if plain:
        plain["speedup_of_value"] = value / plain["value"]
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
        cpu = cpu_baseline(specs, abstol, threads, newton_abstol=args.newton_abstol)
        if accuracy is not None:
            accuracy.update(error_vs_oracle(specs, res_b, abstol, threads, newton_abstol=args.newton_abstol))
            log(f"[bench] accuracy vs oracle: {accuracy}")
    pend = None
    if not args.no_accuracy:
        pend = pendulum_config1(ctx, _abi, _lib)
        log(f"[bench] config 1: {pend}")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tb["wall"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.no_mixed or batch.precision() != _abi.PREC_FP32_REFINE64 else "f64+f32-records",
        "data": "synthetic",
        "timed_region": "device solve of the resident batch (learned guess + persistent solve); H2D upload of "
                        "the batch and D2H of the results are outside it",
        "config": {"workload": f"{workload}: {n} Bratu BVPs u''+lambda e^u=0, lambda~U(0.1,3.4), "
                               "N~U{64..1024}, MIRK4, abstol 1e-8, adaptive; learned warm start "
                               f"(BF16 MLP at {args.anchors} anchors per instance + Hermite, trained untimed on plain "
                               "solves of the Bratu (a, b) family) with fallback to the zero guess",
                   "instances": n, "method": "MIRK4",
                   "precision": "fp64 records" if args.no_mixed else
                   ("mixed: chunk elimination records stored in FP32, FP64 elimination, residual and Newton test"
                    if batch.precision() == _abi.PREC_FP32_REFINE64 else
                    "mixed requested, FP64 records chosen by the autotuner"),
                   "variant": "batched persistent, hipGraph replay" + ("" if args.no_learned else
                                                                       " + learned warm start") +
                   ("" if args.no_mixed else " + mixed precision"),
                   "solved": good_b,
                   "batched_constant_guess": {"value": good_a / ta["wall"], "ms_per_step": ta["wall"] * 1e3,
                                              "solved": good_a, **sa},
                   "batched_learned_guess_fp64": fp64_learned,
                   "headline_run": {"value": value, "guess_ms": tb["guess_ms"], "solve_ms": tb["solve_ms"], **sb},
                   "same_plan_at_reference_newton_tol": at_ref,
                   "training": tinfo, "autotune": tune, "autotune_mixed": tune_mx},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic("k_solve", "traffic_bytes_per_launch" +
                                                                  ("_fp64" if args.no_mixed else "")),
                     "kernel": "k_solve<2,4,variant,BRATU," + ("false>" if args.no_mixed else "true>"),
                     "kernel_ms": tb["solve_ms"], "algorithmic_bytes": abytes},
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "accuracy": accuracy,
        "config1_pendulum": pend,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
