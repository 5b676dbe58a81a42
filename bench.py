#!/usr/bin/env python3
"""Throughput of the LPS-ABC time loop on one MI355X, in node-steps/s, with the absorbing boundary's
quality gate: reflected energy E_int(T)/E0 = 1 + dE/E0 <= 1e-4 (src/main.cpp:986-1043, E0 at step 30).

Default workload (cfg4): the learned ABC closure against the least-squares (LSQ) ABC on the held-out
runs of the closure's workload - R = 125 mm domains (dx = 0.5 mm, delta = 2 mm, 196,321 nodes and
6,268 ABC rows each), 1-4 smooth P/S Gaussian pulses, nu on the 0.025 grid over [0.15, 0.45], all runs
advanced together as one batch. Two legs on the same batch and inputs:
  * LSQ: G_u / G_v / G_theta (src/main.cpp:792-974);
  * learned: the same plus the network's correction at every ABC row (amplitude gate, energy guard
    with per-instance fallback to LSQ).
A step is one time step of every run (dilatation, force with the fused Verlet update, ABC rows, and
in the learned leg the closure). Each leg times exactly --steps steps with inputs resident on the GPU,
the window placed so that it contains an output step (src/main.cpp:981-1123 runs every 30 steps);
value = sum of nodes x steps / time of the learned leg. An untimed pass of each leg over the full
run length (--eval-steps) gives every run's reflected energy; config reports both legs' rates, the
closure's overhead and the gate counts.

roofline: the learned leg's dominant kernel, the closure's network (k_mlp_fwd_wide, BF16 MFMA), its
average duration from HIP events on its own stream inside the timed window (timing mode 3), against
its algorithmic FLOPs (1,083,392 per ABC row); the LSQ leg's force kernel (k_force_rb, HBM, 105 B
per node) is in config. cpu_baseline: the reference solver compiled from its own sources
(oracle/_ref, when present) on one run of the workload for a bounded number of steps, compute-only
(output block off) with the I/O-inclusive figure beside it, and the oracle restatement on the same
sample.

Other workloads: --workload cfg5 (256 domains R 50-500 mm, learned vs LSQ), cfg2 (R = 500 mm single
domain, the force pass's HBM roofline), cfg1 (Example 1), cfg3 (64-domain LSQ batch).

  python bench.py [--steps K] [--warmup W] [--workload cfg4|cfg5|cfg2|cfg1|cfg3] [--runs R]
                  [--eval-steps S] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time
from dataclasses import replace

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
METRIC = "node-steps/s at reflected dE/E0≤1e-4: learned vs LSQ ABC, 1 MI355X"
MLP_FLOPS_PER_ROW = 2.0 * (1600 * 256 + 2 * 256 * 256 + 256 * 4)
GATE = 1e-4


def schedule(args):
    """The step schedule (lps_options.concurrent): --serial is 0, else --schedule, else the
    library's default."""
    return 0 if args.serial else args.schedule


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def window_start(k):
    """Untimed steps before a K-step timed window so that the window holds an output step (a step
    that is a multiple of 30): none when K >= 30, else the window ends at step 30."""
    return 0 if k >= 30 else 30 - k


def measured_traffic(kernel, workload):
    """HBM bytes per launch of kernel on workload from the committed counter summary (FETCH_SIZE
    and WRITE_SIZE passes of rocprofv3, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    e = table.get(f"{kernel}@{workload}")
    if not e:
        return None
    return {"bytes_per_launch": e["bytes_per_launch"],
            "source": f"profiles/hbm_traffic.json ({e['command']})"}


def force_roofline(path, ti, n_nodes, force_ms, workload, passes=None):
    """Roofline of the force pass's main kernel. Its compulsory HBM traffic, each element once:
    reads u, theta, v, a (8 B each component) and kind (1 B), writes a, v and the next prediction:
    105 B per node of the full tiles it covers (the tile kernels leave U^{n+1} of an interior node
    in the prediction ring; the exact kernel also writes u: 121 B). The halo reads of the
    LDS-staged windows are L2 hits and not counted. A CSR-gather formulation of the same pass
    would move 68 B per node plus 4 B per bond (csr_gather_bytes)."""
    fk = path.split("force ")[1].split(" ")[0].rstrip(",") if "force " in path else "exact"
    kernel = {"blocked": "k_force_rb", "unrolled": "k_force_t68<true>",
              "generic": "k_force_tile", "exact": "k_force_exact"}.get(fk, fk)
    units = ti["full_tile_nodes"] if fk in ("blocked", "unrolled") else n_nodes
    per_unit = 121 if fk == "exact" else 105
    force_bytes = per_unit * units
    achieved = force_bytes / (force_ms * 1e-3) / 1e9
    traffic = measured_traffic(kernel, workload)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": traffic["bytes_per_launch"] if traffic else None,
           "kernel": kernel, "units": units, "unit_name": "nodes of the full tiles",
           "bytes_per_unit": per_unit, "algorithmic_bytes": force_bytes,
           "avg_ms": round(force_ms, 5), "csr_gather_bytes": 68 * units + 4 * 68 * units,
           "traffic_source": traffic["source"] if traffic else None}
    if passes:
        out["passes_ms"] = {k: round(v, 5) for k, v in passes.items()}
    return out


def mlp_roofline(n_rows, mlp_ms, workload):
    flops = n_rows * MLP_FLOPS_PER_ROW
    achieved = flops / (mlp_ms * 1e-3) / 1e12
    traffic = measured_traffic("k_mlp_fwd_wide", workload)
    return {"bound": "mfma", "achieved": round(achieved, 1), "peak": BF16_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / BF16_PEAK_TFLOPS, 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": "k_mlp_fwd_wide<true>", "units": n_rows, "unit_name": "ABC rows",
            "flops_per_unit": MLP_FLOPS_PER_ROW, "algorithmic_flops": flops,
            "avg_ms": round(mlp_ms, 5),
            "traffic_source": traffic["source"] if traffic else None}


def pass_breakdown(ctx, steps):
    """Per-pass times (dilatation, force, rest, step) from a separate run of min(steps, 600) steps
    with all four events per step, after the timed region. Leaves timing off and the context
    reset."""
    n = min(steps, 600)
    ctx.reset()
    ctx.set_timing(True)
    ctx.run(n)
    p = ctx.pass_times()
    ctx.set_timing(False)
    ctx.reset()
    return dict(p, steps=n)


def timed_window(ctx, steps, mode):
    """From the reset state: window_start(steps) untimed steps, then exactly `steps` timed steps
    (wall clock around a blocking run, the device synchronized before and after) with timing mode
    `mode` events. Returns (seconds, pass times, records of the window)."""
    pre = window_start(steps)
    if pre:
        ctx.run(pre)
    ctx.set_timing(mode)
    t0 = time.perf_counter()
    recs = ctx.run(steps)
    t1 = time.perf_counter()
    pt = ctx.pass_times()
    ctx.set_timing(False)
    return t1 - t0, pt, recs


REF_MAX_NODES = 1_000_000


def run_streamed(cmd, cwd, env, timeout_s):
    """Runs cmd, passing each line it prints on to the log as it comes (the reference reports its
    setup phases and steps), and returns (exit status, its whole stdout). Killed after timeout_s."""
    import threading
    p = subprocess.Popen(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True)
    timer = threading.Timer(timeout_s, p.kill)
    timer.start()
    lines = []
    try:
        for line in p.stdout:
            lines.append(line)
            if line.strip():
                log("[reference] " + line.rstrip())
        rc = p.wait()
    finally:
        timer.cancel()
    return rc, "".join(lines)


def parse_marching(stdout):
    m = re.findall(r"Time marching took ([0-9.eE+-]+) seconds", stdout)
    return float(m[-1]) if m else None


def cpu_baseline_for(prob, steps, threads, note):
    """The reference solver built from its own sources in oracle/_ref (absent: the oracle alone) on
    `prob` for `steps` steps on `threads` host threads: compute-only (its output block switched off,
    _ref/PD2_LPS_ABC_noout) as the value, the as-is binary (output block and VTK every 30 steps)
    beside it, and the oracle restatement on the same steps. Times are the reference's own "Time
    marching took" line (src/main.cpp:716-1166); the oracle's is its loop."""
    import pyoracle
    from lps_amd import files
    sample = replace(prob, steps=steps)
    N = prob.N
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    bins = {k: os.path.join(ref_dir, b) for k, b in
            (("compute", "PD2_LPS_ABC_noout"), ("io", "PD2_LPS_ABC"))}
    out = {"unit": "node-steps/s", "cores": threads}
    t = {}
    # the reference reads ASCII job files and builds its horizons with a serial kd-tree: minutes
    # of setup on a million-node domain, so there the oracle alone is timed
    use_ref = N <= REF_MAX_NODES and all(os.path.exists(b) for b in bins.values())
    if use_ref:
        tmp = tempfile.mkdtemp(prefix="lps_cpu_")
        try:
            log(f"[bench] CPU baseline: writing the job files ({N} nodes)")
            files.write_job(sample, os.path.join(tmp, "job"))
            for k, b in bins.items():
                log(f"[bench] CPU baseline: reference solver ({k}), {steps} steps")
                rc, stdout = run_streamed([b, "_out", "job"], tmp, env, 600)
                t[k] = parse_marching(stdout) if rc == 0 else None
                log(f"[bench] reference ({k}) rc {rc}: {t[k]} s for {steps} steps")
        finally:
            subprocess.run(["rm", "-rf", tmp])
    log(f"[bench] CPU baseline: oracle, {steps} steps")
    o = pyoracle.Oracle(sample, threads=threads)
    tc = time.perf_counter()
    o.run(steps)
    t["port"] = time.perf_counter() - tc
    if t.get("compute"):
        out.update(value=N * steps / t["compute"], kind="reference",
                   io_inclusive_value=N * steps / t["io"] if t.get("io") else None,
                   port_value=N * steps / t["port"])
        out["sample"] = (f"{note}: {N} nodes x {steps} steps; the reference solver compiled from "
                         f"its own sources (-O3 OpenMP), compute-only ({t['compute']:.2f} s); "
                         f"as-is with output block and VTK every 30 steps {t.get('io')} s; the "
                         f"oracle restatement {t['port']:.2f} s")
    else:
        out.update(value=N * steps / t["port"], kind="port")
        why = (f"the reference's own setup (ASCII files, serial kd-tree) takes minutes above "
               f"{REF_MAX_NODES} nodes" if N > REF_MAX_NODES else "reference binaries not built here")
        out["sample"] = (f"{note}: {N} nodes x {steps} steps, oracle restatement of "
                         f"src/main.cpp (OpenMP, -O2 no FMA), {t['port']:.2f} s ({why})")
    return out


def run_logged(ctx, steps, what, chunk=1000):
    """ctx.run(steps) in chunks of `chunk` steps with a progress line after each (a long untimed
    pass stays visibly alive); the records of every instance, concatenated."""
    out, done = None, 0
    while done < steps:
        n = min(chunk, steps - done)
        recs = ctx.run(n)
        out = recs if out is None else [a + b for a, b in zip(out, recs)]
        done += n
        log(f"[bench]   {what}: {done}/{steps} steps")
    return out


def summary(refl):
    import numpy as np
    r = np.asarray(refl, dtype=float)
    return {"median": float(np.median(r)), "max": float(np.max(r)),
            "gate_le_1e-4": int(np.sum(r <= GATE))}


def bench_learned(args):
    """Learned closure vs LSQ on one batch (cfg4: held-out R = 125 mm runs; cfg5: 256 domains of
    R 50-500 mm)."""
    import numpy as np
    from lps_amd import closure as cl, configs, learned as L, solver
    workers = args.workers or min(16, os.cpu_count() or 1)
    t = time.time()
    cached = bool(args.cache) and os.path.exists(args.cache)
    if args.workload == "cfg4":
        eval_steps = args.eval_steps or 6000
        if cached:
            probs = configs.load_batch(args.cache)
        else:
            runs = L.draw_runs(3000, args.runs)
            probs = [L.run_problem(r, eval_steps, workers) for r in runs]
        desc = (f"cfg4: {len(probs)} held-out runs of the learned-closure workload, R=125 mm, "
                f"dx=0.5 mm, delta=2 mm, 1-4 P/S Gaussian pulses, nu on the 0.025 grid, one batch; "
                f"gate over {eval_steps} steps")
    else:
        w = configs.BATCH_WORKLOADS["cfg5"]
        eval_steps = args.eval_steps or w.steps
        if cached:
            probs = configs.load_batch(args.cache)
        else:
            probs = configs.make_batch("cfg5", workers=workers, steps=eval_steps,
                                       progress=lambda i, p: log(f"[bench]   instance {i}: "
                                                                 f"R={p.info['R']} nu={p.info['nu']} "
                                                                 f"N={p.N} C={p.C}"))
        desc = f"cfg5: {w.description}; gate over {eval_steps} steps"
    if args.cache and not cached:
        configs.save_batch(probs, args.cache)
    elif cached:
        log(f"[bench] loaded {len(probs)} instances from {args.cache}")
    n_nodes = sum(p.N for p in probs)
    n_rows = sum(p.C for p in probs)
    log(f"[bench] {len(probs)} instances, {n_nodes} nodes, {n_rows} ABC rows, generated in "
        f"{time.time() - t:.1f}s")
    t = time.time()
    steps = args.steps if args.steps is not None else 60
    wpath = os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd", "lps_amd", "data",
                         "closure_r125.npz")
    with cl.Closure() as clo, solver.Batch(probs, variant=args.variant,
                                           concurrent=schedule(args)) as b:
        meta = L.load_into(clo, wpath)
        log(f"[bench] GPU setup {time.time() - t:.1f}s, path: {b.path}")
        gates = [args.amp_min * float(np.max(np.abs(p.ic))) for p in probs]

        def learned_on():
            b.attach_closure(clo)
            b.closure_guard(args.guard)
            b.set_closure_gain([1.0] * len(probs))
            b.set_closure_gate(gates)

        # LSQ leg: warm both graphs (untimed, and with timing events), then the timed window
        if args.warmup > 0:
            b.run(args.warmup)
            b.set_timing(2)
            b.run(args.warmup)
            b.set_timing(False)
        b.reset()
        t_lsq, pt_lsq, _ = timed_window(b, steps, 2)
        log(f"[bench] LSQ: {steps} steps in {t_lsq:.4f}s -> {n_nodes * steps / t_lsq:.4e} "
            f"node-steps/s; force kernel {pt_lsq['force_ms']:.4f} ms")
        ti = b.tile_info()
        passes = pass_breakdown(b, steps)
        # learned leg
        learned_on()
        if args.warmup > 0:
            b.run(args.warmup)
            b.set_timing(3)
            b.run(args.warmup)
            b.set_timing(False)
        b.reset()
        t_clo, pt_clo, _ = timed_window(b, steps, 3)
        log(f"[bench] learned: {steps} steps in {t_clo:.4f}s -> {n_nodes * steps / t_clo:.4e} "
            f"node-steps/s; network kernel {pt_clo['force_ms']:.4f} ms")
        # untimed full-length passes: the gate
        b.reset()
        tg = time.perf_counter()
        rl = run_logged(b, eval_steps, "learned, full run")
        t_full_clo = time.perf_counter() - tg
        fb = b.closure_guard()
        b.detach_closure()
        b.reset()
        tg = time.perf_counter()
        rq = run_logged(b, eval_steps, "LSQ, full run")
        t_full_lsq = time.perf_counter() - tg
        path = b.path
    refl_l = [1.0 + r[-1]["de_rel"] for r in rl]
    refl_q = [1.0 + r[-1]["de_rel"] for r in rq]
    selected = [min(a, c) for a, c in zip(refl_l, refl_q)]
    log(f"[bench] full {eval_steps} steps: LSQ {t_full_lsq:.1f}s, learned {t_full_clo:.1f}s; "
        f"reflected median LSQ {np.median(refl_q):.3e}, learned {np.median(refl_l):.3e}")
    value = n_nodes * steps / t_clo
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_for(probs[0], args.cpu_steps, min(16, os.cpu_count() or 1),
                               "instance 0 of the workload")
    out = {
        "metric": METRIC, "value": value, "unit": "node-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": t_clo / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (state), bf16 MFMA with f32 accumulation (closure)", "data": "synthetic",
        "config": {
            "workload": desc, "instances": len(probs), "nodes": n_nodes, "abc_rows": n_rows,
            "kernels": path, "window_start_step": window_start(steps),
            "lsq_node_steps_per_s": n_nodes * steps / t_lsq,
            "closure_overhead": t_clo / t_lsq - 1.0,
            "full_run": {"steps": eval_steps,
                         "lsq_node_steps_per_s": n_nodes * eval_steps / t_full_lsq,
                         "learned_node_steps_per_s": n_nodes * eval_steps / t_full_clo,
                         "overhead": t_full_clo / t_full_lsq - 1.0},
            "reflected": {"lsq": summary(refl_q), "learned": summary(refl_l),
                          "learned_with_lsq_fallback": summary(selected),
                          "learned_le_lsq": int(sum(a <= c for a, c in zip(refl_l, refl_q))),
                          "guard_fallbacks": int(sum(int(f) >= 0 for f in fb))},
            "closure": {"weights": "lps_amd/data/closure_r125.npz", "meta": meta,
                        "gate_rel": args.amp_min, "guard": args.guard},
            "lsq_force_roofline": force_roofline(path, ti, n_nodes, pt_lsq["force_ms"],
                                                 args.workload, passes),
        },
        "roofline": mlp_roofline(n_rows, pt_clo["force_ms"], args.workload),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out, default=float), flush=True)


def bench_batch(args):
    """An LSQ batch (cfg3): every instance advances the same steps in the same launches."""
    from lps_amd import configs, solver
    w = configs.BATCH_WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else 60
    workers = args.workers or min(16, os.cpu_count() or 1)
    t = time.time()
    probs = configs.make_batch(w.name, workers=workers, steps=w.steps)
    n_nodes = sum(p.N for p in probs)
    log(f"[bench] {len(probs)} instances, {n_nodes} nodes, ready in {time.time() - t:.1f}s")
    with solver.Batch(probs, variant=args.variant, concurrent=schedule(args)) as b:
        if args.warmup > 0:
            b.run(args.warmup)
            b.set_timing(2)
            b.run(args.warmup)
            b.set_timing(False)
        b.reset()
        elapsed, pt, _ = timed_window(b, steps, 2)
        ti = b.tile_info()
        passes = pass_breakdown(b, steps)
        b.reset()
        tg = time.perf_counter()
        full_steps = args.eval_steps or w.steps
        recs = b.run(full_steps)
        full_s = time.perf_counter() - tg
        path = b.path
    refl = [1.0 + r[-1]["de_rel"] for r in recs]
    cpu = None if args.no_cpu_baseline else cpu_baseline_for(
        probs[0], args.cpu_steps, min(16, os.cpu_count() or 1), "instance 0 of the workload")
    out = {
        "metric": METRIC, "value": n_nodes * steps / elapsed, "unit": "node-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "instances": len(probs),
                   "nodes": n_nodes, "kernels": path, "abc": "least-squares only",
                   "full_run": {"steps": full_steps, "node_steps_per_s": n_nodes * full_steps / full_s},
                   "reflected": summary(refl)},
        "roofline": force_roofline(path, ti, n_nodes, pt["force_ms"], w.name, passes),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_single(args):
    """One domain (cfg2: the force pass's HBM roofline on R = 500 mm; cfg1: Example 1)."""
    from lps_amd import configs, solver
    w = configs.WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else 60
    workers = args.workers or min(16, os.cpu_count() or 1)
    t = time.time()
    if args.cache and os.path.exists(args.cache):
        prob = configs.load(args.cache)
        log(f"[bench] loaded {w.name} from {args.cache}")
    else:
        prob = configs.make(w.name, workers=workers)
        if args.cache:
            configs.save(prob, args.cache)
    log(f"[bench] N={prob.N} C={prob.C} dt={prob.dt:.3e} ready in {time.time() - t:.1f}s")
    with solver.Solver(prob, exact=args.exact, use_graph=True, variant=args.variant,
                       concurrent=schedule(args)) as s:
        rp, _ = s.horizons()
        nnz = int(rp[-1]) - prob.N
        if args.warmup > 0:
            s.run(args.warmup)
            s.set_timing(2)
            s.run(args.warmup)
            s.set_timing(False)
        s.reset()
        elapsed, pt, _ = timed_window(s, steps, 2)
        ti = s.tile_info()
        passes = pass_breakdown(s, steps)
        tg = time.perf_counter()
        full_steps = args.eval_steps or w.steps
        recs = s.run(full_steps)
        full_s = time.perf_counter() - tg
        path = s.path
    refl = 1.0 + recs[-1]["de_rel"]
    cpu = None if args.no_cpu_baseline else cpu_baseline_for(
        prob, args.cpu_steps, min(16, os.cpu_count() or 1), f"{w.name}")
    out = {
        "metric": METRIC, "value": prob.N * steps / elapsed, "unit": "node-steps/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "nodes": prob.N, "abc_nodes": prob.C,
                   "bonds": nnz, "dt": prob.dt, "exact": bool(args.exact), "kernels": path,
                   "abc": "least-squares only", "window_start_step": window_start(steps),
                   "full_run": {"steps": full_steps, "node_steps_per_s": prob.N * full_steps / full_s,
                                "reflected_energy_final": refl,
                                "gate_reflected_le_1e-4": bool(refl <= GATE)}},
        "roofline": force_roofline(path, ti, prob.N, pt["force_ms"], w.name, passes),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 60)")
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="cfg4",
                    help="cfg4 (default: learned vs LSQ, held-out runs), cfg5 (256 domains, "
                         "learned vs LSQ), cfg2 (R=500 mm), cfg1 (Example 1), cfg3 (64 domains)")
    ap.add_argument("--runs", type=int, default=256, help="cfg4: held-out runs (one batch)")
    ap.add_argument("--eval-steps", type=int, default=None,
                    help="length of the untimed full pass (the gate of the learned workloads; "
                         "default: the workload's, 6000 for cfg4)")
    ap.add_argument("--amp-min", type=float, default=1e-4,
                    help="closure amplitude gate, relative to the run's largest initial |u|")
    ap.add_argument("--guard", type=float, default=1e-3, help="closure energy guard (units of E0)")
    ap.add_argument("--exact", action="store_true")
    ap.add_argument("--variant", type=int, default=0,
                    help="fast-path node kernels: 0 autotune, 1 generic tile, 2 unrolled stencil, "
                         "3 register-blocked stencil")
    ap.add_argument("--serial", action="store_true",
                    help="run the ABC rows' kernels on the main stream instead of beside the "
                         "tile kernels on a second stream")
    ap.add_argument("--schedule", type=int, default=None, choices=(0, 1, 2, 3, 4),
                    help="step schedule: 0 ABC rows after the tile kernels, 1 on a second stream "
                         "beside them, 2 also the partial tiles on a third, 3 the ABC rows inside "
                         "the dilatation launch, 4 as 1 with the deep tiles' dilatation and force "
                         "in one launch (default: the library's, 2 for batches of small domains, "
                         "else 1)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=90)
    ap.add_argument("--workers", type=int, default=None)
    ap.add_argument("--cache", default=None, metavar="NPZ",
                    help="cfg1/cfg2/cfg4/cfg5: load the generated problems from NPZ, or generate "
                         "them and save them there (profiled runs generate outside the profiler)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="run only a plain copy kernel over BYTES (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.add(src, 1.0, out=dst)  # reads and writes args.copy_probe bytes each
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8,
                          "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the path is measured on a "
              f"single MI355X; sharding over several GPUs is a design note, not a measured "
              f"configuration.", file=sys.stderr)
        sys.exit(2)
    if args.workload in ("cfg4", "cfg5"):
        bench_learned(args)
    elif args.workload == "cfg3":
        bench_batch(args)
    else:
        bench_single(args)


if __name__ == "__main__":
    main()
