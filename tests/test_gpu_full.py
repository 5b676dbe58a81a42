"""Example 1 at full size (R = 125 mm, nu = 0.25, 196,321 nodes, 6,268 ABC rows, 4000 steps): the
fast path against the oracle run on the same problem (u, v, theta and the output series within 1e-10
relative L2 after the full run), and against the reference solver's own outputs for this workload
(tests/golden/cfg1_reference.npz: series and a 4096-point sample of the last VTK snapshot) whenever
the regenerated ABC operators are bit-identical to those the reference ran on."""
import os
import sys

import numpy as np
import pytest
from conftest import (ENERGY_KEYS, GOLDEN, VERIFY_KEYS, assert_series_close, records_table,
                      rel_l2)

import pyoracle
from lps_amd import configs, solver

pytestmark = pytest.mark.gpu

sys.path.insert(0, GOLDEN)
from make_cfg1 import operator_sha  # noqa: E402


@pytest.fixture(scope="module")
def example1():
    print("example 1 full size: generating", flush=True)
    threads = min(16, os.cpu_count() or 1)
    prob = configs.make("cfg1", workers=threads)
    last = (prob.steps // 30) * 30  # the reference's last VTK snapshot
    with solver.Solver(prob) as s:
        print("gpu", s.path, flush=True)
        recs = s.run(last)
        f_last = s.fetch()
        s.run(prob.steps - last)
        f = s.fetch()
    return prob, last, recs, f_last, f


def test_example1_full_run_matches_oracle(example1):
    prob, last, recs, f_last, f = example1
    print("oracle", flush=True)
    o = pyoracle.Oracle(prob, threads=min(16, os.cpu_count() or 1))
    orecs = o.run(last)
    of_last = o.fetch()
    o.run(prob.steps - last)
    of = o.fetch()
    for k in ("u", "v", "theta"):
        assert rel_l2(f_last[k], of_last[k]) <= 1e-10, k
        assert rel_l2(f[k], of[k]) <= 1e-10, k
    assert_series_close(records_table(recs, ENERGY_KEYS), records_table(orecs, ENERGY_KEYS),
                        ENERGY_KEYS, 1e-10, prob, "cfg1 vs oracle")


def test_example1_full_run_matches_reference(example1):
    prob, last, recs, f_last, f = example1
    z = np.load(os.path.join(GOLDEN, "cfg1_reference.npz"), allow_pickle=False)
    if operator_sha(prob) != str(z["g_sha"]):
        pytest.skip("the regenerated ABC operators differ in bits from those the reference ran on")
    assert_series_close(records_table(recs, ENERGY_KEYS), z["energy"], ENERGY_KEYS, 1e-10, prob,
                        "cfg1 energy vs reference")
    assert_series_close(records_table(recs, VERIFY_KEYS), z["verify"], VERIFY_KEYS, 1e-10, prob,
                        "cfg1 verify vs reference")
    idx = z["sample"]
    assert int(z["vtk_step"]) == last
    assert rel_l2(f_last["u"].reshape(-1, 2)[idx], z["u"]) <= 1e-10
    assert rel_l2(f_last["v"].reshape(-1, 2)[idx], z["v"]) <= 1e-10
    assert rel_l2(f_last["theta"][idx], z["theta"]) <= 1e-10
