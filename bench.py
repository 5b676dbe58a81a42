#!/usr/bin/env python3
"""Throughput of the LPS-ABC time loop on one MI355X, in node-steps/s.

Workload (default): the large single domain - R = 500 mm lattice (dx = 0.5 mm, 3,141,549 nodes,
25,116 ABC nodes, 213,195,828 bonds), delta = 2 mm, nu = 0.33, Gaussian pulse of radius 40 mm at
(150, -100) mm, least-squares ABC with G_theta, 6000 velocity-Verlet steps at the swept dt. A step is
one time step of the loop over the whole domain (dilatation pass, force pass with the fused Verlet/ABC
epilogue); the timed region is the loop as the reference times it (src/main.cpp:716-1166): all steps
plus the output block every 30 steps, with inputs resident on the GPU and no VTK output.

Printed: one JSON line with the metric, the roofline of the force pass's main kernel (its average
duration from HIP events on its stream inside the timed region, which records only that kernel's
two events per step, against its compulsory HBM bytes, 105 per node; the other passes' times come
from a separate instrumented run after the timed region) and the CPU oracle timed on a bounded
sample of the same workload on the host's cores.

  python bench.py [--steps K] [--warmup W] [--workload cfg2|cfg1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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
    return {"bytes_per_launch": e["bytes_per_launch"], "source": f"profiles/hbm_traffic.json ({e['command']})"}


def force_roofline(path, ti, n_nodes, pt, workload, passes=None):
    """Roofline of the force pass's main kernel (the longest launch of a step). Its compulsory HBM
    traffic, each element once: reads u, theta, v, a (8 B each component) and kind (1 B), writes
    a, v and the next prediction: 105 B per node (the tile kernels leave U^{n+1} of an interior
    node in the prediction ring; the exact kernel also writes u: 121 B). The halo reads of the
    LDS-staged windows are L2 hits and not counted. A CSR-gather formulation of the same pass would move 68 B per node plus
    4 B per bond (coordinates, m, row pointer, column index; csr_gather_bytes)."""
    fk = path.split("force ")[1].split(" ")[0].rstrip(",") if "force " in path else "exact"
    kernel = {"blocked": "k_force_rb", "unrolled": "k_force_t68<true>",
              "generic": "k_force_tile", "exact": "k_force_exact"}.get(fk, fk)
    units = ti["full_tile_nodes"] if fk in ("blocked", "unrolled") else n_nodes
    per_unit = 121 if fk == "exact" else 105
    force_bytes = per_unit * units
    achieved = force_bytes / (pt["force_ms"] * 1e-3) / 1e9
    traffic = measured_traffic(kernel, workload)
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": kernel, "units": units, "unit_name": "nodes of the full tiles",
            "bytes_per_unit": per_unit, "algorithmic_bytes": force_bytes,
            "avg_ms": round(pt["force_ms"], 5),
            "csr_gather_bytes": 68 * units + 4 * 68 * units,
            "traffic_source": traffic["source"] if traffic else None,
            "passes_ms": {k: round(v, 5) for k, v in (passes or pt).items()}}


def pass_breakdown(ctx, steps):
    """Per-pass times (dilatation, force, rest, step) from a separate run of min(steps, 600) steps
    with all four events per step, after the timed region (whose steps carry only the force kernel's
    two events). Leaves timing off and the context reset."""
    n = min(steps, 600)
    ctx.reset()
    ctx.set_timing(True)
    ctx.run(n)
    p = ctx.pass_times()
    ctx.set_timing(False)
    ctx.reset()
    return dict(p, steps=n)


def bench_batch(args, configs, solver):
    """A batched workload: every instance advances the same number of steps in the same launches;
    value = sum of the instances' nodes x steps / time."""
    w = configs.BATCH_WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else w.steps
    workers = args.workers or min(16, os.cpu_count() or 1)
    t = time.time()
    if args.cache and os.path.exists(args.cache):
        probs = configs.load_batch(args.cache)
        for p in probs:
            p.steps = steps
        log(f"[bench] loaded {w.name} from {args.cache}")
    else:
        log(f"[bench] generating {w.name} ({w.description}) with {workers} workers")
        probs = configs.make_batch(
            w.name, workers=workers, steps=steps,
            progress=lambda i, p: log(f"[bench]   instance {i}: R={p.info['R']} nu={p.info['nu']} "
                                      f"N={p.N} C={p.C} dt={p.dt:.3e}"))
        if args.cache:
            configs.save_batch(probs, args.cache)
    n_nodes = sum(p.N for p in probs)
    n_abc = sum(p.C for p in probs)
    log(f"[bench] {len(probs)} instances, {n_nodes} nodes, {n_abc} ABC rows, ready in "
        f"{time.time() - t:.1f}s")
    t = time.time()
    b = solver.Batch(probs, variant=args.variant, concurrent=not args.serial)
    path = b.path
    log(f"[bench] GPU setup {time.time() - t:.1f}s, path: {path}")
    if args.warmup > 0:  # both graphs (untimed, and timed with the event nodes) get built here
        b.run(args.warmup)
        b.set_timing(2)
        b.run(args.warmup)
        b.set_timing(False)
        b.reset()
    b.set_timing(2)  # events around the force kernel only
    t0 = time.perf_counter()
    recs = b.run(steps)
    elapsed = time.perf_counter() - t0
    pt = b.pass_times()
    value = n_nodes * steps / elapsed
    log(f"[bench] {steps} steps in {elapsed:.3f}s -> {value:.4e} node-steps/s; passes {pt}")
    b.reset()
    b.set_timing(False)
    tg = time.perf_counter()
    b.run(steps)
    untimed_s = time.perf_counter() - tg
    log(f"[bench] untimed: {steps} steps in {untimed_s:.3f}s -> {n_nodes * steps / untimed_s:.4e} "
        f"node-steps/s")
    passes = pass_breakdown(b, steps)
    roofline = force_roofline(path, b.tile_info(), n_nodes, pt, w.name, passes)
    b.close()
    refl = [1.0 + r[-1]["de_rel"] for r in recs if r]
    cpu = None
    if not args.no_cpu_baseline:
        import pyoracle
        threads = min(16, os.cpu_count() or 1)
        p0 = probs[0]
        o = pyoracle.Oracle(p0, threads=threads)
        tc = time.perf_counter()
        o.run(args.cpu_steps)
        tcpu = time.perf_counter() - tc
        cpu = {"value": p0.N * args.cpu_steps / tcpu, "unit": "node-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"instance 0 ({p0.N} nodes), first {args.cpu_steps} steps (oracle "
                         f"restatement of src/main.cpp, OpenMP, -O2 no FMA), {tcpu:.1f}s"}
    out = {
        "metric": "node-steps/s at reflected dE/E0\u22641e-4: learned vs LSQ ABC, 1 MI355X",
        "value": value, "unit": "node-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "instances": len(probs),
                   "nodes": n_nodes, "abc_nodes": n_abc, "kernels": path,
                   "abc": "least-squares G_u/G_v/G_theta (learned closure: --workload cfg4)",
                   "streams": "serial" if args.serial else "concurrent",
                   "untimed_node_steps_per_s": n_nodes * steps / untimed_s,
                   "reflected_energy_final_max": max(refl) if refl else None,
                   "instances_gate_reflected_le_1e-4": sum(r <= 1e-4 for r in refl)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_learned(args):
    """Learned ABC closure against the least-squares ABC on held-out runs of the R = 125 mm domain
    (1-4 smooth P/S Gaussian pulses, nu on the 0.025 grid), all advanced in one batch: the
    least-squares batch, then the same batch with the closure (amplitude gate at the training's
    1e-4, energy guard 1e-3 E0 per output record with per-instance fallback). value = node-steps/s
    of the batch with the closure; the least-squares batch's rate and both reflected energies
    (E_int(T)/E0 at the last output record) are in config."""
    from lps_amd import closure as cl, learned as L
    steps = args.steps if args.steps is not None else 6000
    runs = L.draw_runs(3000, args.runs)
    workers = args.workers or min(16, os.cpu_count() or 1)
    path = os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd", "lps_amd", "data",
                        "closure_r125.npz")
    with cl.Closure() as clo:
        meta = L.load_into(clo, path)
        log(f"[bench] closure {path}: trained on {meta.get('train_runs')} runs, teacher-forced "
            f"error left vs LSQ {meta['fit'].get('val_ratio')}")
        ev = L.evaluate(clo, runs, steps, chunk=args.runs, workers=workers, guard=1e-3,
                        amp_min_rel=1e-4, gain=1.0)
    sm = ev["summary"]
    n_rows = sum(L.truncated_problem(r.nu).C for r in runs)
    v_l = ev["node_steps"] / ev["t_learned"]
    v_q = ev["node_steps"] / ev["t_lsq"]
    # the closure's share of a step, from the two loops' wall times (history shift + network)
    t_clo = (ev["t_learned"] - ev["t_lsq"]) / steps
    flops = n_rows * 2.0 * (1600 * 256 + 2 * 256 * 256 + 256 * 4)
    log(f"[bench] {len(runs)} runs x {steps} steps: LSQ {v_q:.4e}, learned {v_l:.4e} node-steps/s; "
        f"closure {t_clo * 1e3:.3f} ms/step for {n_rows} rows")
    out = {
        "metric": "node-steps/s at reflected dE/E0\u22641e-4: learned vs LSQ ABC, 1 MI355X",
        "value": v_l, "unit": "node-steps/s", "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": ev["t_learned"] / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 (state), bf16 MFMA / f32 accumulation (closure)",
        "data": "synthetic",
        "config": {"workload": f"cfg4: {len(runs)} held-out runs, R=125 mm, 1-4 P/S pulses, nu on "
                               f"the 0.025 grid, {steps} steps, one batch",
                   "abc_rows": n_rows, "lsq_node_steps_per_s": v_q,
                   "closure_overhead": sm["overhead"],
                   "reflected_median_lsq": sm["median_lsq"],
                   "reflected_median_learned": sm["median_learned"],
                   "learned_le_lsq": sm["learned_le_lsq"], "fallbacks": sm["fallbacks"],
                   "gate_reflected_le_1e-4": {"lsq": sm["gate_lsq"], "learned": sm["gate_learned"]},
                   "closure_weights": meta},
        "roofline": {"bound": "mfma", "achieved": round(flops / t_clo / 1e12, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(flops / t_clo / 1e12 / 2500.0, 4),
                     "traffic": None, "kernel": "k_mlp_fwd_wide<true> (+ k_hist_shift)",
                     "flops_per_row": flops / n_rows,
                     "note": "closure time = (learned - LSQ loop wall time) / steps"},
        "cpu_baseline": None,
    }
    print(json.dumps(out, default=float), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--workload", default="cfg2",
                    help="cfg2 (default), cfg1, cfg3 (64-domain batch), cfg4 (learned vs LSQ ABC)")
    ap.add_argument("--runs", type=int, default=64, help="cfg4: held-out runs (one batch)")
    ap.add_argument("--exact", action="store_true")
    ap.add_argument("--variant", type=int, default=0,
                    help="fast-path node kernels: 0 autotune, 1 generic tile, 2 unrolled stencil, "
                         "3 register-blocked stencil")
    ap.add_argument("--serial", action="store_true",
                    help="run the ABC rows' kernels on the main stream instead of beside the "
                         "tile kernels on a second stream")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--workers", type=int, default=None)
    ap.add_argument("--cache", default=None, metavar="NPZ",
                    help="load the generated workload from NPZ, or generate and save it there")
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
              f"single MI355X; multi-GPU sharding is a design note, not a measured configuration.",
              file=sys.stderr)
        sys.exit(2)

    from lps_amd import configs, solver
    if args.workload == "cfg4":
        bench_learned(args)
        return
    if args.workload in configs.BATCH_WORKLOADS:
        bench_batch(args, configs, solver)
        return
    w = configs.WORKLOADS[args.workload]
    steps = args.steps if args.steps is not None else w.steps
    workers = args.workers or min(16, os.cpu_count() or 1)
    t = time.time()
    if args.cache and os.path.exists(args.cache):
        prob = configs.load(args.cache)
        prob.steps = steps
        log(f"[bench] loaded {w.name} from {args.cache}")
    else:
        log(f"[bench] generating {w.name} ({w.description}) with {workers} workers")
        prob = configs.make(w.name, workers=workers, steps=steps)
        if args.cache:
            configs.save(prob, args.cache)
    log(f"[bench] N={prob.N} C={prob.C} dt={prob.dt:.3e} ready in {time.time() - t:.1f}s")

    t = time.time()
    s = solver.Solver(prob, exact=args.exact, use_graph=True, variant=args.variant,
                      concurrent=not args.serial)
    rp, _ = s.horizons()
    path = s.path
    nnz = int(rp[-1]) - prob.N  # bonds, self excluded
    log(f"[bench] GPU setup (horizons, m_i, theta history) {time.time() - t:.1f}s, bonds={nnz}, "
        f"path: {s.path}")
    if args.warmup > 0:  # both graphs (untimed, and timed with the event nodes) get built here
        s.run(args.warmup)
        s.set_timing(2)
        s.run(args.warmup)
        s.set_timing(False)
        s.reset()
    log("[bench] warmup done; timed run")

    s.set_timing(2)  # events around the force kernel only
    t0 = time.perf_counter()
    recs = s.run(steps)
    t1 = time.perf_counter()
    pt = s.pass_times()
    elapsed = t1 - t0
    value = prob.N * steps / elapsed
    last = recs[-1] if recs else None
    refl = (1.0 + last["de_rel"]) if last else None
    log(f"[bench] {steps} steps in {elapsed:.3f}s -> {value:.4e} node-steps/s; passes {pt}")
    # the same steps without the per-kernel timing events
    s.reset()
    s.set_timing(False)
    tg = time.perf_counter()
    s.run(steps)
    untimed_s = time.perf_counter() - tg
    log(f"[bench] untimed: {steps} steps in {untimed_s:.3f}s -> {prob.N * steps / untimed_s:.4e} "
        f"node-steps/s")
    passes = pass_breakdown(s, steps)
    roofline = force_roofline(path, s.tile_info(), prob.N, pt, w.name, passes)
    s.close()

    cpu = None
    if not args.no_cpu_baseline:
        import pyoracle
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] CPU oracle: setup then {args.cpu_steps} steps on {threads} threads")
        o = pyoracle.Oracle(prob, threads=threads)
        tc = time.perf_counter()
        o.run(args.cpu_steps)
        tcpu = time.perf_counter() - tc
        cpu = {"value": prob.N * args.cpu_steps / tcpu, "unit": "node-steps/s", "cores": threads,
               "kind": "port",
               "sample": f"first {args.cpu_steps} steps of the same workload (oracle restatement of "
                         f"src/main.cpp, OpenMP, -O2 no FMA), {tcpu:.1f}s"}
        log(f"[bench] CPU {cpu['value']:.4e} node-steps/s")

    out = {
        "metric": "node-steps/s at reflected dE/E0\u22641e-4: learned vs LSQ ABC, 1 MI355X",
        "value": value, "unit": "node-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": elapsed / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{w.name}: {w.description}", "nodes": prob.N, "abc_nodes": prob.C,
                   "bonds": nnz, "dt": prob.dt, "exact": bool(args.exact), "kernels": path,
                   "abc": "least-squares G_u/G_v/G_theta (learned closure: --workload cfg4)",
                   "streams": "serial" if args.serial else "concurrent",
                   "untimed_node_steps_per_s": prob.N * steps / untimed_s,
                   "reflected_energy_final": refl,
                   "gate_reflected_le_1e-4": (refl is not None and refl <= 1e-4)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
