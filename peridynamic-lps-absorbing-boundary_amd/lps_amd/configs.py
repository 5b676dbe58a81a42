"""The named workloads of the benchmark and the tests, as generator calls.

The (dt, delta_kappa) pairs are the results of the preprocessor's stability sweep
(preprocessor/main.cpp:1047-1174, restated in preprocess.make_problem with dt=None) on each domain,
computed once and pinned here so that a run does not repeat the sweep's ~10^5 least-squares solves.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

from . import preprocess as pp


@dataclass(frozen=True)
class Workload:
    name: str
    R: float
    nu: float
    steps: int
    pulse: pp.Pulse
    dt: float            # sweep result
    dk_factor: float     # sweep result
    description: str
    # ABC operators from the symmetric builder (one octant of rows solved, the others their exact
    # lattice images, as for the batched workloads) instead of one solve per row
    symmetric: bool = False


WORKLOADS = {
    # Example 1 of the reference: R = 125 mm, centred pulse of radius 40 mm, nu = 0.25.
    # The sweep on this domain selects dt = 17.5 ns, delta_kappa factor 0.9 (max ||G|| = 0.99891).
    "cfg1": Workload("cfg1", 125.0, 0.25, 4000, pp.Pulse(0.0, 0.0, 40.0, 1.0, (1.0, 1.0)),
                     17.5e-9, 0.9,
                     "R=125 mm, dx=0.5 mm, delta=2 mm, nu=0.25, centred Gaussian pulse r=40 mm, "
                     "LSQ ABC (G_theta), 4000 steps"),
    # Large single domain: R = 500 mm, nu = 0.33, pulse at (0.3 R, -0.2 R).
    "cfg2": Workload("cfg2", 500.0, 0.33, 6000, pp.Pulse(150.0, -100.0, 40.0, 1.0, (1.0, 1.0)),
                     17.0e-9, 0.7,
                     "R=500 mm, dx=0.5 mm, delta=2 mm, nu=0.33, Gaussian pulse r=40 mm at "
                     "(150,-100) mm, LSQ ABC (G_theta, symmetric builder), 6000 steps",
                     symmetric=True),
}


_ARRAYS = ("nodes", "data2", "bctypes", "bcvalues", "ic", "icd", "abc", "clouds", "Gu", "Gv", "Gth")


def save(p: pp.Problem, path: str) -> None:
    """Binary copy of a generated problem (numpy .npz, no pickled objects)."""
    import numpy as np
    np.savez(path, steps=np.int64(p.steps), **{k: getattr(p, k) for k in _ARRAYS})


def load(path: str) -> pp.Problem:
    import numpy as np
    z = np.load(path, allow_pickle=False)
    return pp.Problem(steps=int(z["steps"]), **{k: z[k] for k in _ARRAYS})


def make(name: str, workers: Optional[int] = None, steps: Optional[int] = None) -> pp.Problem:
    w = WORKLOADS[name]
    p = pp.make_problem(w.R, w.nu, pulses=[w.pulse], steps=steps or w.steps, dt=w.dt,
                        dk_factor=w.dk_factor, workers=workers, symmetric=w.symmetric)
    p.info["workload"] = w.name
    return p


# ---------------------------------------------------------------------------------------------
# Batched workloads: many independent domains advanced together (lps_create_batch).
# ---------------------------------------------------------------------------------------------
NU_SET = (0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45)

# (dt, delta_kappa factor) of the preprocessor's stability sweep (preprocess.make_problem with
# dt=None) on an R = 100 mm domain, per nu (every pair stable: max ||G|| between 0.9961 and
# 0.99996); each batch instance takes the pair of its nu and builds its own ABC operators for its
# own radius.
NU_SWEEP = {
    0.15: (17.0e-9, 0.7),
    0.20: (17.0e-9, 0.7),
    0.25: (17.5e-9, 0.9),
    0.30: (15.0e-9, 0.5),
    0.35: (17.5e-9, 0.9),
    0.40: (17.0e-9, 0.7),
    0.45: (16.0e-9, 0.7),
}
# The same sweep at the midpoints (max ||G|| between 0.9985 and 0.9998, all stable): with NU_SWEEP
# a 0.025 grid over [0.15, 0.45], the nu values of the learned-closure workloads.
NU_SWEEP_MID = {
    0.175: (17.0e-9, 0.7),
    0.225: (17.5e-9, 1.0),
    0.275: (17.5e-9, 0.9),
    0.325: (17.0e-9, 0.7),
    0.375: (17.5e-9, 0.9),
    0.425: (16.0e-9, 0.7),
}
NU_GRID = tuple(sorted(list(NU_SWEEP) + list(NU_SWEEP_MID)))


def sweep_pair(nu: float):
    """(dt, delta_kappa factor) of a nu on the 0.025 grid."""
    key = round(nu, 3)
    return NU_SWEEP[key] if key in NU_SWEEP else NU_SWEEP_MID[key]


@dataclass(frozen=True)
class BatchWorkload:
    name: str
    count: int
    r_min: float
    r_max: float
    steps: int
    seed: int
    description: str
    nu_grid: bool = False  # nu ~ U(0.15, 0.45) snapped to NU_GRID (else drawn from NU_SET)


BATCH_WORKLOADS = {
    # 64 instances, R ~ U(50, 250) mm snapped to dx, nu from NU_SET, 1-3 Gaussian pulses with
    # r ~ U(10, 40) mm, centres uniform in the disk of radius R/2, amplitude N(0, 1e-3), direction at
    # a uniform angle; 3000 steps.
    "cfg3": BatchWorkload("cfg3", 64, 50.0, 250.0, 3000, 0,
                          "64 domains, R ~ U(50,250) mm, nu in {0.15..0.45}, 1-3 pulses, LSQ ABC "
                          "(G_theta), 3000 steps"),
    # 256 instances, R ~ U(50, 500) mm snapped to dx, nu ~ U(0.15, 0.45) on the 0.025 grid (each
    # with its sweep pair), pulses as in cfg3; 5000 steps. About 3e8 nodes and 3.5e6 ABC rows.
    "cfg5": BatchWorkload("cfg5", 256, 50.0, 500.0, 5000, 5,
                          "256 domains, R ~ U(50,500) mm, nu ~ U(0.15,0.45) on the 0.025 grid, "
                          "1-3 pulses, LSQ ABC (G_theta), 5000 steps", nu_grid=True),
}


def batch_specs(name: str, count: Optional[int] = None):
    """The instances of a batched workload (radius, nu, pulses, dt, delta_kappa factor), drawn from
    the workload's seed."""
    import numpy as np
    w = BATCH_WORKLOADS[name]
    rng = np.random.default_rng(w.seed)
    dx = pp.DX_DEFAULT
    specs = []
    for _ in range(count or w.count):
        R = round(rng.uniform(w.r_min, w.r_max) / dx) * dx
        if w.nu_grid:
            nu = min(NU_GRID, key=lambda g: abs(g - rng.uniform(0.15, 0.45)))
        else:
            nu = NU_SET[int(rng.integers(len(NU_SET)))]
        pulses = []
        for _ in range(int(rng.integers(1, 4))):
            r = rng.uniform(10.0, 40.0)
            rho = 0.5 * R * math.sqrt(rng.uniform())
            phi = rng.uniform(0.0, 2.0 * math.pi)
            ang = rng.uniform(0.0, 2.0 * math.pi)
            amp = rng.normal(0.0, 1e-3)
            pulses.append(pp.Pulse(rho * math.cos(phi), rho * math.sin(phi), r, amp,
                                   (math.cos(ang), math.sin(ang))))
        dt, dk = sweep_pair(nu)
        specs.append(dict(R=R, nu=nu, pulses=pulses, dt=dt, dk_factor=dk))
    return specs


def make_batch(name: str, workers: Optional[int] = None, steps: Optional[int] = None,
               count: Optional[int] = None, progress=None, only=None):
    """The problems of a batched workload; progress(i, problem) is called after each one. The ABC
    operators come from the symmetric builder (one octant solved, preprocess.lattice_symmetry).
    only: indices of the instances to make (default all)."""
    w = BATCH_WORKLOADS[name]
    probs = []
    for i, sp in enumerate(batch_specs(name, count)):
        if only is not None and i not in only:
            continue
        p = pp.make_problem(sp["R"], sp["nu"], pulses=sp["pulses"], steps=steps or w.steps,
                            dt=sp["dt"], dk_factor=sp["dk_factor"], workers=workers,
                            symmetric=True)
        p.info["workload"] = name
        probs.append(p)
        if progress:
            progress(i, p)
    return probs


def save_batch(probs, path: str) -> None:
    import numpy as np
    arrays = {"count": np.int64(len(probs))}
    for i, p in enumerate(probs):
        arrays[f"i{i}_steps"] = np.int64(p.steps)
        for k in _ARRAYS:
            arrays[f"i{i}_{k}"] = getattr(p, k)
    np.savez(path, **arrays)


def load_batch(path: str):
    import numpy as np
    z = np.load(path, allow_pickle=False)
    return [pp.Problem(steps=int(z[f"i{i}_steps"]), **{k: z[f"i{i}_{k}"] for k in _ARRAYS})
            for i in range(int(z["count"]))]
