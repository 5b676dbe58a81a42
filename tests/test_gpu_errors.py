"""Error paths of the C ABI, as the reference's FATAL exits (src/main.cpp): a horizon of more than
128 nodes (kd_tree_interface.h:96-99, the oracle refuses the same input), a failed material
consistency check (src/main.cpp:310-322) and malformed problems, each reported with its error code
before any step runs."""
import numpy as np
import pytest
from conftest import load_golden

import pyoracle
from lps_amd import preprocess as pp, solver

pytestmark = pytest.mark.gpu


def test_horizon_overflow_is_reported():
    """A search radius of 6.6 dx on the lattice holds 137 nodes (> 128) in an interior horizon; 6 dx
    holds 113."""
    print("horizon overflow", flush=True)
    prob, _ = load_golden("abc_r10_nu025")
    r = 6.6 * 0.5
    with pytest.raises(solver.LpsError) as e:
        solver.Solver(prob, exact=True, horizon_radius=r)
    assert e.value.code == solver.LPS_ERR_HORIZON
    # the oracle's build refuses the same input (nnz < 0), and accepts a radius just inside 128
    L = pyoracle.lib()
    rp = np.zeros(prob.N + 1, dtype=np.int64)
    nodes = np.ascontiguousarray(prob.nodes)
    assert L.oracle_build_horizons(prob.N, nodes.ctypes.data, r, 128, rp.ctypes.data, None) < 0
    r_ok = 6.0 * 0.5  # 113 nodes
    assert L.oracle_build_horizons(prob.N, nodes.ctypes.data, r_ok, 128, rp.ctypes.data, None) > 0
    with solver.Solver(prob, exact=True, horizon_radius=r_ok) as s:
        rg, _ = s.horizons()
    assert np.array_equal(rg, rp)


def test_inconsistent_material_is_fatal():
    """nu = 0.5 makes lambda infinite: the Lame round trip fails (src/main.cpp:310-322)."""
    print("material check", flush=True)
    prob, _ = load_golden("abc_r10_nu025")
    prob.data2 = prob.data2.copy()
    prob.data2[1] = 0.5
    with pytest.raises(solver.LpsError) as e:
        solver.Solver(prob)
    assert e.value.code == solver.LPS_ERR_CONSISTENCY


def test_malformed_problems_are_refused():
    print("malformed", flush=True)
    prob, _ = load_golden("abc_r10_nu025")
    cases = []
    empty = pp.Problem(**{**prob.__dict__})
    empty.nodes = np.zeros((0, 2))
    empty.ic = np.zeros(0)
    empty.icd = np.zeros(0)
    empty.bctypes = np.zeros(0, np.int64)
    empty.abc = np.zeros(0, np.int64)
    empty.clouds = np.zeros((0, 0), np.int64)
    cases.append(empty)
    zero_dx = pp.Problem(**{**prob.__dict__})
    zero_dx.data2 = prob.data2.copy()
    zero_dx.data2[8] = 0.0
    cases.append(zero_dx)
    for p in cases:
        with pytest.raises(solver.LpsError) as e:
            solver.Solver(p)
        assert e.value.code == solver.LPS_ERR_ARG
    # the G_theta method without G_theta operators
    with pytest.raises(solver.LpsError) as e:
        no_gth = pp.Problem(**{**prob.__dict__})
        no_gth.Gth = np.zeros((0, 0))
        solver.Solver(no_gth, theta_method=solver.LPS_THETA_GTHETA)
    assert e.value.code == solver.LPS_ERR_ARG
    # ... which the IDW method does not need
    no_gth = pp.Problem(**{**prob.__dict__})
    no_gth.Gth = np.zeros((0, 0))
    with solver.Solver(no_gth, theta_method=solver.LPS_THETA_IDW) as s:
        s.run(30)
