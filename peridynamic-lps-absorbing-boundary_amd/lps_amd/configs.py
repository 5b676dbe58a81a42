"""The named workloads of the benchmark and the tests, as generator calls.

The (dt, delta_kappa) pairs are the results of the preprocessor's stability sweep
(preprocessor/main.cpp:1047-1174, restated in preprocess.make_problem with dt=None) on each domain,
computed once and pinned here so that a run does not repeat the sweep's ~10^5 least-squares solves.
"""
from __future__ import annotations

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
                     "(150,-100) mm, LSQ ABC (G_theta), 6000 steps"),
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
                        dk_factor=w.dk_factor, workers=workers)
    p.info["workload"] = w.name
    return p
