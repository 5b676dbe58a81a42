import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def load_golden(name):
    """A golden case as (Problem, dict of reference outputs). ref["theta_method"] is the ABC
    dilatation the reference ran (1 G_theta, 0 IDW); a fixture with a `base` holds only outputs,
    for the inputs of that other fixture."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    if "base" in z:
        prob, _ = load_golden(str(z["base"]))
        prob.info = {"golden": name}
        if "steps" in z:
            prob.steps = int(z["steps"])
        ref = {k: z[k] for k in ("energy", "verify", "vtk_points", "vtk_u", "vtk_v", "vtk_theta")}
        ref["vtk_step"] = int(z["vtk_step"])
        ref["theta_method"] = int(z["theta_method"])
        return prob, ref
    return _load_full(z, name)


def _load_full(z, name):
    from lps_amd.preprocess import Problem
    if "Gu32" in z:
        Gu, Gv, Gth = (z[k].astype(np.float64) for k in ("Gu32", "Gv32", "Gth32"))
    elif "Gu" in z:
        Gu, Gv, Gth = z["Gu"], z["Gv"], z["Gth"]
    else:
        Gu = Gv = Gth = np.zeros((0, 0))
    prob = Problem(nodes=z["nodes"], steps=int(z["steps"]), data2=z["data2"],
                   bctypes=z["bctypes"].astype(np.int64), bcvalues=np.zeros(2 * z["nodes"].shape[0]),
                   ic=z["ic"], icd=z["icd"], abc=z["abc"].astype(np.int64),
                   clouds=z["clouds"].astype(np.int64), Gu=Gu, Gv=Gv, Gth=Gth,
                   info={"golden": name})
    if "prenotch" in z:
        prob.prenotch = z["prenotch"]
    ref = {k: z[k] for k in ("energy", "verify", "vtk_points", "vtk_u", "vtk_v", "vtk_theta")}
    ref["vtk_step"] = int(z["vtk_step"])
    ref["theta_method"] = int(z["theta_method"]) if "theta_method" in z else 1
    return prob, ref


def vtk_index(nodes, points):
    """Row of each VTK point in the node array (exact coordinate match)."""
    lookup = {(float(x), float(y)): i for i, (x, y) in enumerate(nodes)}
    return np.array([lookup[(float(x), float(y))] for x, y in points])


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


ENERGY_KEYS = ["t", "energy", "energy_k", "energy_p", "px", "py", "max_u", "theta_max", "ub_rms", "vb_rms"]
VERIFY_KEYS = ["t", "energy", "de_rel", "px", "py", "theta_max", "max_u", "max_v", "max_a"]


def records_table(recs, keys):
    return np.array([[r[k] for k in keys] for r in recs], dtype=np.float64)


def assert_series_close(mine, ref, keys, rtol, prob, what=""):
    """Column-wise relative L2 check of an output series. The momentum columns are conserved
    quantities that sit at round-off level (|P| ~ 1e-21 while |p| <= sqrt(M 2K)); they are compared
    against that Cauchy-Schwarz scale sqrt(rho dV N * 2 E_k) instead of their own size. dE/E0 is
    E/E0 - 1 and is compared on the scale of E/E0."""
    mine = np.asarray(mine)
    ref = np.asarray(ref)
    assert mine.shape == ref.shape, (what, mine.shape, ref.shape)
    rho, dV = float(prob.data2[2]), float(prob.data2[8] * prob.data2[9])
    ek = ref[:, keys.index("energy")] if "energy_k" not in keys else ref[:, keys.index("energy_k")]
    pscale = np.sqrt(np.abs(rho * dV * prob.N * 2.0 * ek)) + 1e-300
    for j, k in enumerate(keys):
        if k in ("px", "py"):
            err = float(np.max(np.abs(mine[:, j] - ref[:, j]) / pscale))
        elif k == "de_rel":  # E/E0 - 1: compared on the scale of E/E0
            err = float(np.max(np.abs(mine[:, j] - ref[:, j]) / (1.0 + np.abs(ref[:, j]))))
        else:
            err = rel_l2(mine[:, j], ref[:, j])
        assert err <= rtol, (what, k, err)


@pytest.fixture(scope="session")
def golden():
    return load_golden
