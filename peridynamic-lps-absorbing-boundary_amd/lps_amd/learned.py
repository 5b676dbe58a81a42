"""Learned ABC closure: training data from reflection-free runs, training, and closed-loop evaluation
against the least-squares ABC.

The reference has no learned closure; its ABC is the least-squares G_u / G_v update
(src/main.cpp:883-974, operators from preprocessor/main.cpp:415-637). Here a network (lps_amd.closure)
adds a correction to that update at every ABC row, and the least-squares ABC stays both the
fallback (gain 0) and the standard it is measured against: a closure is kept for an instance only
where its reflected energy is at or below the least-squares ABC's.

Workload (truncated domain R = 125 mm, truth R = 375 mm):
  * initial conditions: 1-4 Gaussian pulses, radius U(5, 40) mm, centres uniform in the disk of
    radius 0.6 R, amplitude N(0, 1e-3), radial (P) or tangential (S) polarisation;
  * nu uniform on the 0.025 grid over [0.15, 0.45] (configs.NU_GRID), each with the stability
    sweep's dt (configs.sweep_pair);
  * truth: the same pulses on an R = 375 mm domain without ABC. Its field at the truncated domain's
    ABC rows is reflection-free until a wave from the pulses has gone out to 375 mm and back to the
    rows' clouds (2 * 375 - 121 - max(|c| + r) mm at c_P, 10% margin);
  * samples (teacher forcing): inputs from truth's history, target = truth's U^{n+1}, V^{n+1} minus
    the truncated rows' G_u / G_v on truth's U^n, U^{n-1}, V^n, V^{n-1} (lps_closure_collect).
Reflected energy of a run: E_int(T) / E0 = 1 + dE/E0 of the last output record
(src/main.cpp:986-1043, E0 at step 30).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field, replace
from typing import Dict, List, Optional

import numpy as np

from . import configs, preprocess as pp

R_TRUNC = 125.0
R_TRUTH = 375.0
DX = pp.DX_DEFAULT
DELTA = pp.DELTA_DEFAULT


@dataclass
class Run:
    seed: int
    nu: float
    pulses: List[pp.Pulse]
    d_max: float       # max over pulses of |centre| + radius (mm)
    truth_steps: int   # reflection-free window of the truth run, capped at the waves' passage
    dt: float
    dk: float


def wave_speeds(nu: float):
    mc = pp.material(nu=nu)
    return math.sqrt((mc.lam + 2.0 * mc.mu) / mc.rho), math.sqrt(mc.mu / mc.rho)


def draw_runs(seed: int, count: int, R: float = R_TRUNC, centre_frac: float = 0.6) -> List[Run]:
    rng = np.random.default_rng(seed)
    runs = []
    for k in range(count):
        nu = configs.NU_GRID[int(rng.integers(len(configs.NU_GRID)))]
        pulses = []
        for _ in range(int(rng.integers(1, 5))):
            r = rng.uniform(5.0, 40.0)
            rho = centre_frac * R * math.sqrt(rng.uniform())
            phi = rng.uniform(0.0, 2.0 * math.pi)
            amp = rng.normal(0.0, 1e-3)
            pol = "P" if rng.uniform() < 0.5 else "S"
            pulses.append(pp.Pulse(rho * math.cos(phi), rho * math.sin(phi), r, amp, (1.0, 0.0),
                                   polarization=pol))
        d_max = max(math.hypot(p.cx, p.cy) + p.radius for p in pulses)
        dt, dk = configs.sweep_pair(nu)
        cP, cS = wave_speeds(nu)
        band = R - 2.0 * DELTA  # innermost cloud node of the ABC rows
        t_valid = 0.9 * (2.0 * R_TRUTH - band - d_max) / cP
        t_pass = (R + d_max) / cS + 5e-6
        runs.append(Run(seed=int(splitmix(seed, k)), nu=nu, pulses=pulses, d_max=d_max,
                        truth_steps=int(min(t_valid, t_pass) / dt), dt=dt, dk=dk))
    return runs


def splitmix(a: int, b: int) -> int:
    x = (a * 0x9E3779B97F4A7C15 + b + 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


_TRUNC_CACHE: Dict[float, pp.Problem] = {}


def truncated_problem(nu: float, workers: Optional[int] = None) -> pp.Problem:
    """The R = 125 mm domain with its least-squares ABC operators for nu (zero initial state)."""
    if nu not in _TRUNC_CACHE:
        dt, dk = configs.sweep_pair(nu)
        _TRUNC_CACHE[nu] = pp.make_problem(R_TRUNC, nu, pulses=[], steps=1, dt=dt, dk_factor=dk,
                                           workers=workers, robust_ls=True, symmetric=True)
    return _TRUNC_CACHE[nu]


def run_problem(run: Run, steps: int, workers: Optional[int] = None) -> pp.Problem:
    base = truncated_problem(run.nu, workers)
    return replace(base, ic=pp.gaussian_ic(base.nodes, run.pulses), steps=steps,
                   info=dict(base.info, run_seed=run.seed))


def log(*a):
    print("[learned]", *a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------
# Samples
# ---------------------------------------------------------------------------------------------
def collect(clo, runs: List[Run], keep: float, sample_every: int, amp_min_rel: float,
            y_max: float = 1.0, workers: Optional[int] = None, device: int = 0) -> dict:
    """Teacher-forced samples of every run appended to clo's store, grouped by nu."""
    from .solver import Solver
    stats = {"runs": 0, "truth_node_steps": 0, "truth_s": 0.0, "prep_s": 0.0}
    by_nu: Dict[float, List[Run]] = {}
    for r in runs:
        by_nu.setdefault(r.nu, []).append(r)
    for nu in sorted(by_nu):
        t0 = time.time()
        trunc = truncated_problem(nu, workers)
        dt, dk = configs.sweep_pair(nu)
        truth_p = pp.make_problem(R_TRUTH, nu, pulses=[], steps=1, dt=dt, with_abc=False)
        stats["prep_s"] += time.time() - t0
        with Solver(truth_p, abc_enabled=False, device=device) as truth:
            for r in by_nu[nu]:
                t1 = time.time()
                ic = pp.gaussian_ic(truth_p.nodes, r.pulses)
                amp = float(np.max(np.abs(ic)))
                truth.set_initial(ic)
                n = clo.collect(truth, trunc, r.truth_steps, first_step=8,
                                sample_every=sample_every, keep=keep, amp_min=amp_min_rel * amp,
                                y_max=y_max, seed=r.seed)
                dtr = time.time() - t1
                stats["runs"] += 1
                stats["truth_node_steps"] += truth_p.N * r.truth_steps
                stats["truth_s"] += dtr
                log(f"nu {nu:.3f} run {stats['runs']}/{len(runs)}: {r.truth_steps} truth steps, "
                    f"{len(r.pulses)} pulses, samples {n} ({dtr:.2f} s)")
    return stats


def collect_closed(clo, runs: List[Run], keep: float, sample_every: int, amp_min_rel: float,
                   y_max: float = 1.0, gain: float = 1.0, chunk: int = 60, blowup: float = 1.5,
                   workers: Optional[int] = None, device: int = 0) -> dict:
    """Closed-loop samples (DAgger): every run on the truncated domain with the closure attached
    (gain, no guard) beside its reflection-free truth run; inputs are the closure's own history,
    targets the correction that would put the truncated run back on truth
    (lps_closure_collect_closed). A run stops contributing once its energy exceeds `blowup` times
    its energy after the first chunk (a diverged run's samples would teach nothing)."""
    from .solver import Solver
    stats = {"runs": 0, "steps": 0, "stopped": 0, "s": 0.0}
    by_nu: Dict[float, List[Run]] = {}
    for r in runs:
        by_nu.setdefault(r.nu, []).append(r)
    for nu in sorted(by_nu):
        dt, dk = configs.sweep_pair(nu)
        truth_p = pp.make_problem(R_TRUTH, nu, pulses=[], steps=1, dt=dt, with_abc=False)
        with Solver(truth_p, abc_enabled=False, device=device) as truth:
            for r in by_nu[nu]:
                t1 = time.time()
                prob = run_problem(r, r.truth_steps, workers)
                amp = float(np.max(np.abs(prob.ic)))
                truth.set_initial(pp.gaussian_ic(truth_p.nodes, r.pulses))
                with Solver(prob, device=device) as trunc:
                    trunc.attach_closure(clo)
                    trunc.closure_guard(0.0)
                    trunc.set_closure_gain([gain])
                    trunc.set_closure_gate([amp_min_rel * amp])
                    e_first, done = None, 0
                    while done < r.truth_steps:
                        n = min(chunk, r.truth_steps - done)
                        clo.collect_closed(truth, trunc, n, first_step=8, sample_every=sample_every,
                                           keep=keep, amp_min=amp_min_rel * amp, y_max=y_max,
                                           seed=r.seed)
                        done += n
                        e = trunc.diagnostics()["energy"]
                        e_first = e if e_first is None else e_first
                        if not (e <= blowup * e_first):
                            stats["stopped"] += 1
                            break
                stats["runs"] += 1
                stats["steps"] += done
                stats["s"] += time.time() - t1
                log(f"closed-loop nu {nu:.3f} run {stats['runs']}/{len(runs)}: {done} of "
                    f"{r.truth_steps} steps, samples {clo.count()} ({time.time() - t1:.1f} s)")
    return stats


def dagger(clo, train_runs: List[Run], val_runs: List[Run], iters: int, runs_per_iter: int,
           gains: List[float], keep: float, sample_every: int, amp_min_rel: float, y_max: float,
           steps: int, batch: int, lr0: float, lr1: float, eval_steps: int, seed: int = 0,
           workers: Optional[int] = None, out: Optional[str] = None) -> dict:
    """DAgger-style rounds on a closure fitted on teacher-forced samples: each round adds
    closed-loop samples from runs_per_iter training runs driven by the current network, fine-tunes
    on every sample so far (the output scales stay those of the first fit), and evaluates the
    closed loop on val_runs against the least-squares ABC. Keeps the round whose median reflected
    energy is lowest (weights written to `out`)."""
    rounds, best = [], None
    for it in range(iters):
        part = train_runs[(it * runs_per_iter) % len(train_runs):][:runs_per_iter]
        g = gains[min(it, len(gains) - 1)]
        st = collect_closed(clo, part, keep, sample_every, amp_min_rel, y_max=y_max, gain=g,
                            workers=workers)
        n = min(clo.count(), clo.capacity())
        t0 = time.time()
        curve = clo.train(steps, batch, lr0, lr1, 0, n, seed=seed + 17 * (it + 1))
        ev = evaluate(clo, val_runs, eval_steps, workers=workers, guard=1e-3,
                      amp_min_rel=amp_min_rel, gain=1.0)
        sm = ev["summary"]
        rec = {"round": it, "gain_collect": g, "collect": st, "samples": n,
               "train_s": time.time() - t0, "loss_last": float(np.mean(curve[-50:])),
               "eval": sm}
        rounds.append(rec)
        log(f"DAgger round {it}: samples {n}, closed loop median LSQ {sm['median_lsq']:.3e} "
            f"learned {sm['median_learned']:.3e}, learned <= LSQ on {sm['learned_le_lsq']}/"
            f"{sm['runs']}, fallbacks {sm['fallbacks']}")
        if best is None or sm["median_learned"] < best[0]:
            best = (sm["median_learned"], it, clo.get_params())
            if out:
                save(clo, out, {"dagger_round": it, "eval": sm, "samples": n})
    if best is not None:
        clo.set_params(best[2])
    return {"rounds": rounds, "best_round": best[1] if best else None}


def target_scales(clo, first: int, count: int, chunk: int = 65536, pieces: int = 16):
    """Weighted RMS of the u and v targets over a spread sample of [first, first + count)."""
    su = sv = sw = 0.0
    for k in range(pieces):
        a = first + (count * k) // pieces
        n = min(chunk, first + count - a)
        if n <= 0:
            continue
        _, aux = clo.read(a, n)
        w = aux[:, 4].astype(np.float64)
        su += float(np.sum(w * (aux[:, 0] ** 2 + aux[:, 1] ** 2)))
        sv += float(np.sum(w * (aux[:, 2] ** 2 + aux[:, 3] ** 2)))
        sw += float(np.sum(w))
    return math.sqrt(su / (2.0 * sw)), math.sqrt(sv / (2.0 * sw))


def fit(clo, n_train: int, n_val: int, steps: int, batch: int, lr0: float, lr1: float,
        seed: int = 0, noise: float = 0.0) -> dict:
    s_u, s_v = target_scales(clo, 0, n_train)
    clo.set_scales(s_u, s_v)
    log(f"target scales sigma_u {s_u:.4g} sigma_v {s_v:.4g}")
    t0 = time.time()
    curve = clo.train(steps, batch, lr0, lr1, 0, n_train, seed=seed, input_noise=noise)
    t_train = time.time() - t0
    out = {"sigma": [s_u, s_v], "steps": steps, "batch": batch, "noise": noise, "train_s": t_train,
           "loss_first": float(np.mean(curve[:50])), "loss_last": float(np.mean(curve[-50:]))}
    if n_val > 0:
        l, b, _ = clo.evaluate(n_train, n_val)
        out["val_ratio"] = l / b  # remaining fraction of the least-squares error (teacher forced)
    l, b, _ = clo.evaluate(0, min(n_train, 1 << 20))
    out["train_ratio"] = l / b
    log(f"trained {steps} steps x {batch} in {t_train:.1f} s: loss {out['loss_first']:.4g} -> "
        f"{out['loss_last']:.4g}; error left vs LSQ: train {out['train_ratio']:.3f}, "
        f"validation {out.get('val_ratio', float('nan')):.3f}")
    return out


def save(clo, path: str, meta: dict):
    s_u, s_v = clo.get_scales()
    np.savez(path, params=clo.get_params(), sigma=np.array([s_u, s_v], np.float32),
             meta=np.array(json.dumps(meta)))


def export_bin(params: np.ndarray, sigma, path: str):
    """The weight file of PD2_LPS_ABC --abc=learned=<path>: the float32 parameters, then sigma_u,
    sigma_v, little endian."""
    out = np.concatenate([np.asarray(params, np.float32).ravel(), np.asarray(sigma, np.float32)])
    out.astype("<f4").tofile(path)


def load_into(clo, path: str) -> dict:
    z = np.load(path, allow_pickle=False)
    clo.set_params(z["params"])
    clo.set_scales(float(z["sigma"][0]), float(z["sigma"][1]))
    return json.loads(str(z["meta"]))


# ---------------------------------------------------------------------------------------------
# Closed loop
# ---------------------------------------------------------------------------------------------
def reflected(recs) -> float:
    return 1.0 + recs[-1]["de_rel"]


def energy_rise(recs) -> float:
    """Largest rise of the energy between consecutive output records, in units of E0."""
    e = np.array([r["energy"] for r in recs])
    return float(np.max(np.diff(e)) / e[0]) if len(e) > 1 and e[0] > 0 else 0.0


def evaluate(clo, runs: List[Run], steps: int, chunk: int = 64, device: int = 0,
             workers: Optional[int] = None, guard: float = 1e-3, amp_min_rel: float = 0.0,
             gain: float = 1.0) -> dict:
    """Every run on the truncated domain with the least-squares ABC and with the closure (same
    batch, reset between); per-instance reflected energy and the time of both loops."""
    from .solver import Batch
    res = {"lsq": [], "learned": [], "nu": [], "fallback_step": [], "lsq_rise": [],
           "learned_rise": [], "t_lsq": 0.0, "t_learned": 0.0, "node_steps": 0, "gain": gain}
    for a in range(0, len(runs), chunk):
        part = runs[a:a + chunk]
        probs = [run_problem(r, steps, workers) for r in part]
        with Batch(probs, device=device) as b:
            t0 = time.time()
            rl = b.run(steps)
            t1 = time.time()
            b.reset()
            b.attach_closure(clo)
            b.closure_guard(guard)
            b.set_closure_gain([gain] * len(probs))
            b.set_closure_gate([amp_min_rel * float(np.max(np.abs(p.ic))) for p in probs])
            t2 = time.time()
            rn = b.run(steps)
            t3 = time.time()
            fb = b.closure_guard()
        res["t_lsq"] += t1 - t0
        res["t_learned"] += t3 - t2
        res["node_steps"] += sum(p.N for p in probs) * steps
        for i, r in enumerate(part):
            res["lsq"].append(reflected(rl[i]))
            res["learned"].append(reflected(rn[i]))
            res["nu"].append(r.nu)
            res["fallback_step"].append(int(fb[i]))
            res["lsq_rise"].append(energy_rise(rl[i]))
            res["learned_rise"].append(energy_rise(rn[i]))
        lq = np.array(res["lsq"][-len(part):])
        ln = np.array(res["learned"][-len(part):])
        log(f"eval runs {a + 1}-{a + len(part)}: reflected energy median LSQ {np.median(lq):.3e} "
            f"learned {np.median(ln):.3e}; learned <= LSQ on {int(np.sum(ln <= lq))}/{len(part)}")
    lq, ln = np.array(res["lsq"]), np.array(res["learned"])
    res["summary"] = {
        "runs": len(runs), "steps": steps,
        "median_lsq": float(np.median(lq)), "median_learned": float(np.median(ln)),
        "mean_lsq": float(np.mean(lq)), "mean_learned": float(np.mean(ln)),
        "learned_le_lsq": int(np.sum(ln <= lq)),
        "gate_lsq": int(np.sum(lq <= 1e-4)), "gate_learned": int(np.sum(ln <= 1e-4)),
        "fallbacks": int(np.sum(np.array(res["fallback_step"]) >= 0)),
        "lsq_max_rise": float(np.max(res["lsq_rise"])),
        "overhead": res["t_learned"] / res["t_lsq"] - 1.0,
    }
    return res


def run_eval(clo, a, report, tag):
    for gain in a.gains:
        for gate in (a.gates or [a.amp_min]):
            ev = evaluate(clo, draw_runs(3000 + a.seed, a.eval_runs), a.eval_steps,
                          workers=a.workers, guard=a.guard, amp_min_rel=gate, gain=gain)
            report[f"eval_{tag}_gain{gain:g}_gate{gate:g}"] = ev
            log(f"closed loop, {tag}, gain {gain:g}, gate {gate:g}:", json.dumps(ev["summary"]))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train-runs", type=int, default=512)
    ap.add_argument("--val-runs", type=int, default=32)
    ap.add_argument("--eval-runs", type=int, default=64)
    ap.add_argument("--eval-steps", type=int, default=6000)
    ap.add_argument("--keep", type=float, default=0.01)
    ap.add_argument("--sample-every", type=int, default=8)
    ap.add_argument("--amp-min", type=float, default=1e-4)
    ap.add_argument("--y-max", type=float, default=1.0)
    ap.add_argument("--guard", type=float, default=1e-3)
    ap.add_argument("--gains", type=float, nargs="+", default=[1.0])
    ap.add_argument("--gates", type=float, nargs="+", default=None,
                    help="amplitude gates (relative to the IC amplitude) of the evaluation; "
                         "default the training's --amp-min")
    ap.add_argument("--noise", type=float, nargs="+", default=[0.0],
                    help="input-noise levels; one network trained and evaluated per level")
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr0", type=float, default=1e-3)
    ap.add_argument("--lr1", type=float, default=1e-5)
    ap.add_argument("--capacity", type=int, default=8_000_000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="closure_r125.npz")
    ap.add_argument("--report", default=None)
    ap.add_argument("--load", default=None, help="skip training: load this parameter file")
    ap.add_argument("--workers", type=int, default=None)
    ap.add_argument("--dagger", type=int, default=0, metavar="ROUNDS",
                    help="closed-loop (DAgger) rounds after the teacher-forced fit (or --load)")
    ap.add_argument("--dagger-runs", type=int, default=32, help="training runs per DAgger round")
    ap.add_argument("--dagger-gains", type=float, nargs="+", default=[0.5, 1.0],
                    help="closure gain while collecting, per round (the last repeats)")
    ap.add_argument("--dagger-steps", type=int, default=5000, help="Adam steps per round")
    ap.add_argument("--dagger-eval-runs", type=int, default=32)
    ap.add_argument("--dagger-keep", type=float, default=0.05)
    ap.add_argument("--tf-runs", type=int, default=0,
                    help="with --load and --dagger: teacher-forced samples of this many runs "
                         "first, so the rounds train on both kinds")
    ap.add_argument("--export-bin", default=None, metavar="BIN",
                    help="with --load: write the weight file of PD2_LPS_ABC --abc=learned=BIN and stop")
    a = ap.parse_args(argv)
    if a.export_bin:
        z = np.load(a.load, allow_pickle=False)
        export_bin(z["params"], z["sigma"], a.export_bin)
        return {"exported": a.export_bin}
    from .closure import Closure
    report = {"args": vars(a)}
    with Closure(seed=a.seed) as clo:
        if a.load:
            report["loaded"] = load_into(clo, a.load)
            if a.dagger > 0:
                clo.reserve(a.capacity)
                if a.tf_runs > 0:
                    report["collect_tf"] = collect(clo, draw_runs(1000 + a.seed, a.tf_runs),
                                                   a.keep, a.sample_every, a.amp_min,
                                                   y_max=a.y_max, workers=a.workers)
                report["dagger"] = dagger(
                    clo, draw_runs(4000 + a.seed, a.dagger * a.dagger_runs),
                    draw_runs(2000 + a.seed, a.dagger_eval_runs), a.dagger, a.dagger_runs,
                    a.dagger_gains, a.dagger_keep, a.sample_every, a.amp_min, a.y_max,
                    a.dagger_steps, a.batch, a.lr0 * 0.3, a.lr1, a.eval_steps, a.seed,
                    a.workers, out=a.out)
            if a.eval_runs > 0:
                run_eval(clo, a, report, "loaded")
        else:
            train_runs = draw_runs(1000 + a.seed, a.train_runs)
            val_runs = draw_runs(2000 + a.seed, a.val_runs)
            clo.reserve(a.capacity)
            t0 = time.time()
            report["collect_train"] = collect(clo, train_runs, a.keep, a.sample_every, a.amp_min,
                                              y_max=a.y_max, workers=a.workers)
            n_train = min(clo.count(), a.capacity)
            report["collect_val"] = collect(clo, val_runs, a.keep, a.sample_every, a.amp_min,
                                            y_max=a.y_max, workers=a.workers)
            n_all = min(clo.count(), a.capacity)
            report["samples"] = {"train": n_train, "val": n_all - n_train,
                                 "dropped": clo.count() - n_all}
            report["collect_s"] = time.time() - t0
            log(f"samples: train {n_train}, validation {n_all - n_train} "
                f"({report['collect_s']:.0f} s)")
            init = clo.get_params()
            for noise in a.noise:
                clo.set_params(init)
                tag = f"noise{noise:g}"
                report["fit_" + tag] = fit(clo, n_train, n_all - n_train, a.train_steps, a.batch,
                                           a.lr0, a.lr1, a.seed, noise)
                out = a.out if len(a.noise) == 1 else a.out.replace(".npz", f"_{tag}.npz")
                save(clo, out, {"fit": report["fit_" + tag], "samples": report["samples"],
                                "train_runs": a.train_runs, "seed": a.seed})
                log("saved", out)
                if a.eval_runs > 0:
                    run_eval(clo, a, report, tag)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(report, f, indent=1, default=float)
    return report


if __name__ == "__main__":
    main()
