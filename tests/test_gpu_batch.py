"""Batches (lps_create_batch): instances of different radius, nu, dt, ABC operators and initial
conditions advanced together give the same bits as each instance alone on the same kernels, and the
reference's outputs for the golden jobs."""
import numpy as np
import pytest
from conftest import (ENERGY_KEYS, VERIFY_KEYS, assert_series_close, load_golden, records_table,
                      rel_l2, vtk_index)

from lps_amd import preprocess as pp, solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def instances():
    g10, ref10 = load_golden("abc_r10_nu025")
    g12, ref12 = load_golden("abc_r12_nu033")
    r40 = pp.make_problem(40.0, 0.3, steps=600, dt=1.7e-8, dk_factor=0.7,
                          pulses=[pp.Pulse(8.0, -5.0, 12.0, 1.0, (1.0, 1.0))], workers=8)
    return [g10, r40, g12], [ref10, None, ref12]


@pytest.mark.parametrize("variant", [3, 2])
def test_batch_equals_instances_alone(variant, instances):
    print("batch vs alone", variant, flush=True)
    probs, _ = instances
    steps = 300
    with solver.Batch(probs, variant=variant) as b:
        assert b.count == 3 and b.path.startswith("batch of 3"), b.path
        recs = b.run(steps)
        fields = [b.fetch(i) for i in range(3)]
    for i, p in enumerate(probs):
        with solver.Solver(p, variant=variant) as s:
            srecs = s.run(steps)
            sf = s.fetch()
        for k in ("u", "v", "a", "theta"):
            assert np.array_equal(fields[i][k], sf[k]), (i, k)
        assert len(recs[i]) == len(srecs)
        for r, q in zip(recs[i], srecs):
            assert r == q, (i, r, q)


def test_instances_with_the_same_operators_share_them(instances):
    """Instances whose ABC inputs are the same (one geometry and material, other initial
    conditions) keep one copy of the operator blocks in the batch; every instance still gives the
    same bits as alone."""
    print("batch shared operators", flush=True)
    probs, _ = instances
    g10 = probs[0]
    other = pp.Problem(**{**g10.__dict__})
    other.ic = -0.7 * g10.ic
    batch = [g10, probs[1], other, g10]
    steps = 150
    with solver.Batch(batch, variant=3) as b:
        assert "2 instances share an earlier one's ABC operators" in b.path, b.path
        recs = b.run(steps)
        fields = [b.fetch(i) for i in range(len(batch))]
    for i, p in enumerate(batch):
        with solver.Solver(p, variant=3) as s:
            srecs = s.run(steps)
            sf = s.fetch()
        for k in ("u", "v", "a", "theta"):
            assert np.array_equal(fields[i][k], sf[k]), (i, k)
        assert recs[i] == srecs, i
    assert not np.array_equal(fields[0]["u"], fields[2]["u"])


def test_batch_matches_the_reference_golden_outputs(instances):
    print("batch vs golden", flush=True)
    probs, refs = instances
    steps = max(p.steps for p, r in zip(probs, refs) if r is not None)
    with solver.Batch(probs) as b:
        recs = b.run(steps)
        fields = [b.fetch(i) for i in range(3)]
    for i, (p, ref) in enumerate(zip(probs, refs)):
        if ref is None:
            continue
        n = p.steps // 30
        assert_series_close(records_table(recs[i][:n], ENERGY_KEYS), ref["energy"], ENERGY_KEYS,
                            1e-10, p, f"batch {i}")
        assert_series_close(records_table(recs[i][:n], VERIFY_KEYS), ref["verify"], VERIFY_KEYS,
                            1e-10, p, f"batch {i}")
        if p.steps == steps:
            idx = vtk_index(p.nodes, ref["vtk_points"])
            assert rel_l2(fields[i]["u"].reshape(-1, 2)[idx], ref["vtk_u"]) <= 1e-10


def test_batch_reports_a_nonfinite_instance(instances):
    print("batch nonfinite", flush=True)
    probs, _ = instances
    bad = pp.Problem(**{**probs[1].__dict__})
    bad.ic = bad.ic.copy()
    bad.ic[7] = np.nan
    with solver.Batch([probs[0], bad]) as b:
        with pytest.raises(solver.LpsError) as e:
            b.run(60)
        assert e.value.code == solver.LPS_ERR_NONFINITE and "instance 1" in str(e.value)


def test_single_problem_calls_are_refused_on_a_batch(instances):
    probs, _ = instances
    with solver.Batch(probs[:2]) as b:
        with pytest.raises(solver.LpsError) as e:
            solver.Solver.run(b, 30)
        assert e.value.code == solver.LPS_ERR_ARG
