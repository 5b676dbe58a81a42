"""The nu-sweep driver (lps_amd.sweep) and the .energy/.verify writer it uses."""
import os

import numpy as np
import pytest

from lps_amd import files, preprocess as pp, sweep


def _records(n, de_last=-0.5):
    recs = []
    for k in range(1, n + 1):
        r = {c: float(k) * 0.1 + j for j, c in enumerate(files.ENERGY_COLUMNS)}
        r.update({c: float(k) * 0.2 + j for j, c in enumerate(files.VERIFY_COLUMNS) if c not in r})
        r["t"] = k * 30 * 1.7e-8
        r["de_rel"] = de_last * k / n
        recs.append(r)
    return recs


def test_series_files_read_back(tmp_path):
    recs = _records(5)
    for kind, cols in (("energy", files.ENERGY_COLUMNS), ("verify", files.VERIFY_COLUMNS)):
        path = str(tmp_path / ("x." + kind))
        files.write_series(path, recs, kind)
        head = open(path).readline().strip()
        assert head == (files.ENERGY_HEADER if kind == "energy" else files.VERIFY_HEADER)
        back = files.read_series(path)
        want = np.array([[r[c] for c in cols] for r in recs])
        assert np.array_equal(back, want)  # 16 significant digits, as the solver writes


def test_table_and_markdown():
    nus = [0.25, 0.33]
    probs = []
    for nu in nus:
        p = pp.Problem(nodes=np.zeros((1, 2)), steps=600, data2=np.array(
            [0.19, nu, 8e-15, 0, 0, 0, 1.7e-8, 2.0, 0.5, 0.5, 0, 0, 0, 0, 0]),
            bctypes=np.zeros(2, np.int64), bcvalues=np.zeros(2), ic=np.zeros(2), icd=np.zeros(2),
            abc=np.zeros(0, np.int64), clouds=np.zeros((0, 0), np.int64), Gu=np.zeros((0, 0)),
            Gv=np.zeros((0, 0)), Gth=np.zeros((0, 0)))
        probs.append(p)
    rows = sweep.table(nus, probs, [_records(20, -0.999), _records(20, -0.9)])
    assert rows[0]["lambda_pd"] == 0.0  # lambda = mu at nu = 1/4
    assert rows[1]["lambda_pd"] > 0.0
    assert abs(rows[0]["absorption"] - 99.9) < 1e-9
    md = sweep.markdown(rows, dict(node_steps=1e9, loop_s=0.5, setup_s=1.0, batch_steps=600))
    assert "| 0.25 |" in md and "99.90%" in md and "2.000e+09 node-steps/s" in md


@pytest.mark.gpu
def test_sweep_runs_match_single_runs(tmp_path):
    """Two nu on a small domain through the driver (one batch, plots on): the records equal those
    of each problem run alone, and the files and table are written."""
    print("sweep", flush=True)
    from lps_amd import solver
    nus = [0.25, 0.35]
    probs = sweep.build(nus, R=20.0)
    probs[0].steps, probs[1].steps = 300, 240  # runs of different lengths in one batch
    res = sweep.run(probs)
    for p, recs in zip(probs, res["records"]):
        with solver.Solver(p, variant=3) as s:
            alone = s.run(p.steps)
        assert len(recs) == len(alone) == p.steps // 30
        for a_, b_ in zip(recs, alone):
            assert a_ == b_
    out = str(tmp_path / "out")
    sweep.main(["--nus", "0.25", "--R", "16", "--steps", "120", "--out", out, "--plots"])
    for name in ("nu025.energy", "nu025.verify", "sweep.md", "energy_plot_nu025.png"):
        assert os.path.getsize(os.path.join(out, name)) > 0
