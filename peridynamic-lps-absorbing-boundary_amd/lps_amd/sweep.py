"""Poisson-ratio sweep driver and energy plots.

The reference README reports a nu sweep of its Example 1 (README.md: the table of nu, lambda_PD,
final dE/E0 and absorption, and one energy plot per nu) made by a run_sweep.sh and a
plot_energy.py that the repository does not hold (README.md:84, .gitignore:3-4). This module
regenerates them on the GPU:

  * per nu: the Example 1 domain (R = 125 mm, dx = 0.5 mm, delta = 2 mm, one ABC layer, a centred
    Gaussian pulse of radius 40 mm along (1, 1): preprocessor/main.cpp:850-902, 1280-1289) with
    its own least-squares ABC operators at the (dt, delta_kappa) of the preprocessor's stability
    sweep (:1047-1174), run for StepNo = ceil(210 us / dt) steps (:1177-1179);
  * every nu advanced together as one batch (lps_create_batch); each keeps the records up to its
    own StepNo, the same bits as a run alone;
  * outputs: <out>/nu<ddd>.energy and .verify in the solver's format, <out>/sweep.md (the README
    table plus the run times) and, with --plots, <out>/energy_plot_nu<ddd>.png.

    python -m lps_amd.sweep --out sweep_out --plots
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import configs, files, preprocess as pp

R_EXAMPLE1 = 125.0
T_FINAL = pp.T_FINAL  # 210 us (preprocessor/main.cpp:1177-1179)


def log(*a):
    print("[sweep]", *a, file=sys.stderr, flush=True)


def build(nus: Sequence[float], R: float = R_EXAMPLE1, ls: str = "cpu",
          run_sweep: bool = False, workers: Optional[int] = None) -> List[pp.Problem]:
    """One Example-1 problem per nu. (dt, delta_kappa) from the stability sweep: recomputed with
    run_sweep, else the pinned results of configs.NU_SWEEP / NU_SWEEP_MID (the same sweep run once,
    every pair stable)."""
    probs = []
    for nu in nus:
        pulses = [pp.Pulse(0.0, 0.0, 40.0, 1.0, (1.0, 1.0))]
        t = time.time()
        if run_sweep:
            p = pp.make_problem(R, nu, pulses=pulses, workers=workers, ls=ls)
        else:
            dt, dk = configs.sweep_pair(nu)
            p = pp.make_problem(R, nu, pulses=pulses, dt=dt, dk_factor=dk, workers=workers, ls=ls)
        p.steps = int(math.ceil(T_FINAL / p.dt))
        log(f"nu={nu}: N={p.N} C={p.C} dt={p.dt:.3e} StepNo={p.steps} built in "
            f"{time.time() - t:.1f}s")
        probs.append(p)
    return probs


def run(probs: List[pp.Problem]) -> Dict:
    """Advances all problems as one batch; returns each one's records up to its own StepNo."""
    from . import solver
    t = time.time()
    with solver.Batch(probs) as b:
        setup = time.time() - t
        steps = max(p.steps for p in probs)
        t = time.time()
        recs = b.run(steps)
        loop = time.time() - t
    out = [r[: p.steps // files.SUBSTEP_NO] for r, p in zip(recs, probs)]
    node_steps = sum(p.N * p.steps for p in probs)
    return dict(records=out, setup_s=setup, loop_s=loop, batch_steps=steps, node_steps=node_steps)


def table(nus: Sequence[float], probs: List[pp.Problem], records) -> List[Dict]:
    rows = []
    for nu, p, recs in zip(nus, probs, records):
        mc = pp.material(nu=nu, delta=float(p.data2[7]), dx=float(p.data2[8]), dy=float(p.data2[9]))
        last = recs[-1]
        rows.append(dict(nu=nu, lambda_pd=mc.lambda_pd, dt=p.dt, steps=p.steps,
                         final_de=last["de_rel"], absorption=-100.0 * last["de_rel"],
                         reflected=1.0 + last["de_rel"]))
    return rows


def markdown(rows: List[Dict], meta: Dict) -> str:
    out = ["| nu   | lambda_PD  | Final dE/E0 | Absorption | 1 + dE/E0 | dt (ns) | StepNo |",
           "|------|------------|-------------|------------|-----------|---------|--------|"]
    for r in rows:
        out.append(f"| {r['nu']:.2f} | {r['lambda_pd']:+.5f}   | {r['final_de']:+.3f}      | "
                   f"{r['absorption']:6.2f}%    | {r['reflected']:.2e}  | {r['dt'] * 1e9:.1f}    | "
                   f"{r['steps']} |")
    out.append("")
    out.append(f"{len(rows)} runs as one batch on the GPU: {meta['node_steps']:.3e} node-steps in "
               f"{meta['loop_s']:.2f} s ({meta['node_steps'] / meta['loop_s']:.3e} node-steps/s), "
               f"setup {meta['setup_s']:.1f} s.")
    return "\n".join(out) + "\n"


def plot(path: str, nu: float, recs: List[Dict[str, float]]):
    """Energy plot of one run: total, kinetic and potential energy against time (left), and
    dE/E0 (right)."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    t = np.array([r["t"] for r in recs]) * 1e6
    fig, ax = plt.subplots(1, 2, figsize=(11, 4))
    for key, lab in (("energy", "E total"), ("energy_k", "kinetic"), ("energy_p", "potential")):
        ax[0].plot(t, [r[key] for r in recs], label=lab)
    ax[0].set_xlabel("t (us)")
    ax[0].set_ylabel("energy (interior)")
    ax[0].legend()
    ax[0].set_title(f"nu = {nu:.2f}")
    ax[1].plot(t, [r["de_rel"] for r in recs])
    ax[1].set_xlabel("t (us)")
    ax[1].set_ylabel("dE / E0")
    ax[1].axhline(-1.0, color="gray", lw=0.5)
    fig.tight_layout()
    fig.savefig(path, dpi=110)
    plt.close(fig)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--nus", type=float, nargs="+", default=list(configs.NU_SET))
    ap.add_argument("--R", type=float, default=R_EXAMPLE1)
    ap.add_argument("--ls", choices=("cpu", "gpu"), default="cpu",
                    help="least-squares ABC operators on the CPU (LAPACK) or the GPU")
    ap.add_argument("--run-sweep", action="store_true",
                    help="recompute each nu's (dt, delta_kappa) sweep instead of the pinned pairs")
    ap.add_argument("--steps", type=int, default=None, help="override StepNo (all runs)")
    ap.add_argument("--workers", type=int, default=None)
    ap.add_argument("--out", default="sweep_out")
    ap.add_argument("--plots", action="store_true")
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)
    probs = build(a.nus, a.R, a.ls, a.run_sweep, a.workers)
    if a.steps is not None:
        for p in probs:
            p.steps = a.steps
    res = run(probs)
    rows = table(a.nus, probs, res["records"])
    for nu, recs in zip(a.nus, res["records"]):
        tag = f"nu{int(round(nu * 100)):03d}"
        files.write_series(os.path.join(a.out, tag + ".energy"), recs, "energy")
        files.write_series(os.path.join(a.out, tag + ".verify"), recs, "verify")
        if a.plots:
            plot(os.path.join(a.out, f"energy_plot_{tag}.png"), nu, recs)
    meta = {k: res[k] for k in ("setup_s", "loop_s", "batch_steps", "node_steps")}
    with open(os.path.join(a.out, "sweep.md"), "w") as f:
        f.write(markdown(rows, meta))
    print(json.dumps(dict(rows=rows, **meta)))


if __name__ == "__main__":
    main()
