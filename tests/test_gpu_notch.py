"""Prenotch visibility (src/main.cpp:119-140, 386-402) on the GPU: the horizons built with the
notches equal the oracle's (sets and order), IDW runs match the oracle, the drop-in executable
reproduces the reference solver's outputs on the notched job, and a batch refuses a notched problem.
The time loop's parity on the notched golden job is in test_gpu_parity.py (exact path bit for bit
against the oracle, fast path against the reference's outputs)."""
import os
import subprocess

import numpy as np
import pytest
from conftest import (ENERGY_KEYS, PKG, VERIFY_KEYS, assert_series_close, load_golden,
                      records_table, rel_l2, vtk_index)

import pyoracle
from lps_amd import files, solver

pytestmark = pytest.mark.gpu

NAME = "notch_r12_nu025"


@pytest.mark.parametrize("exact", [True, False])
def test_notched_horizons_equal_oracle(exact):
    print("notch horizons", exact, flush=True)
    prob, _ = load_golden(NAME)
    o = pyoracle.Oracle(prob)
    orp, ocols = o.horizons()
    with solver.Solver(prob, exact=exact) as s:
        rp, cols = s.horizons()
        path = s.path
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    prob.prenotch = np.zeros((0, 4))
    with solver.Solver(prob, exact=exact) as s:
        assert s.horizons()[0][-1] > rp[-1] + 1000  # the notches cut bonds
    if not exact:  # cut rows differ from the lattice stencil: the per-bond kernels run
        assert "horizons differ" in path, path


@pytest.mark.parametrize("exact", [True, False])
def test_notched_idw_matches_reference(exact):
    """A notched job (nu = 0.33) with the IDW ABC dilatation against the reference solver built
    with it (series and last fields within 1e-10), and on the exact path the oracle bit for bit."""
    print("notch idw", exact, flush=True)
    prob, ref = load_golden("notch_r12_nu033_idw")
    assert ref["theta_method"] == solver.LPS_THETA_IDW and prob.prenotch.shape == (2, 4)
    with solver.Solver(prob, theta_method=solver.LPS_THETA_IDW, exact=exact) as s:
        recs = s.run(prob.steps)
        gf = s.fetch()
    assert_series_close(records_table(recs, ENERGY_KEYS), ref["energy"], ENERGY_KEYS, 1e-10,
                        prob, "notch idw")
    assert_series_close(records_table(recs, VERIFY_KEYS), ref["verify"], VERIFY_KEYS, 1e-10,
                        prob, "notch idw")
    idx = vtk_index(prob.nodes, ref["vtk_points"])
    assert rel_l2(gf["u"].reshape(-1, 2)[idx], ref["vtk_u"]) <= 1e-10
    assert rel_l2(gf["theta"][idx], ref["vtk_theta"]) <= 1e-10
    if exact:
        o = pyoracle.Oracle(prob, theta_method=0)
        o.run(prob.steps)
        of = o.fetch()
        for k in ("u", "v", "theta"):
            assert np.array_equal(gf[k], of[k]), k


def test_cli_notched_job_reproduces_the_reference(tmp_path):
    print("notch cli", flush=True)
    prob, ref = load_golden(NAME)
    job = str(tmp_path / "job")
    files.write_job(prob, job, exact_text=True)
    r = subprocess.run([os.path.join(PKG, "PD2_LPS_ABC"), "_out", "job"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert_series_close(files.read_series(job + "_out.energy"), ref["energy"], ENERGY_KEYS, 1e-10,
                        prob, NAME)
    assert_series_close(files.read_series(job + "_out.verify"), ref["verify"], VERIFY_KEYS, 1e-10,
                        prob, NAME)
    u = files.read_vtk_field(f"{job}_out-{ref['vtk_step']}.vtk")
    assert np.array_equal(u["points"], ref["vtk_points"])
    assert rel_l2(u["vectors"], ref["vtk_u"]) <= 1e-10


def test_batch_refuses_a_notched_problem():
    print("notch batch", flush=True)
    prob, _ = load_golden(NAME)
    other, _ = load_golden("abc_r12_nu033")
    with pytest.raises(solver.LpsError):
        solver.Batch([other, prob])
