"""The CPU oracle against the reference's own outputs (golden fixtures made by
tests/golden/make_golden.py from the reference solver and kd-tree compiled from its sources)."""
import numpy as np
import pytest
from conftest import (ENERGY_KEYS, VERIFY_KEYS, assert_series_close, load_golden, records_table,
                      rel_l2, vtk_index)

import pyoracle

CASES = ["noabc_r30", "abc_r10_nu025", "abc_r12_nu033", "notch_r12_nu025",
         # the reference solver built with the IDW ABC dilatation (src/main.cpp:765-790)
         "abc_r12_nu033_idw", "notch_r12_nu033_idw", "abc_r10_nu045_idw"]


@pytest.fixture(scope="module", params=CASES)
def oracle_run(request):
    prob, ref = load_golden(request.param)
    o = pyoracle.Oracle(prob, theta_method=ref["theta_method"], threads=4)
    recs = o.run(prob.steps)
    return request.param, prob, ref, o, recs


def test_series_match_reference(oracle_run):
    # tolerance 1e-12: the oracle sums each node's bonds in (distance, index) order, the reference in
    # its kd-tree tie order; nothing else differs
    name, prob, ref, o, recs = oracle_run
    assert_series_close(records_table(recs, ENERGY_KEYS), ref["energy"], ENERGY_KEYS, 1e-12, prob, name)
    assert_series_close(records_table(recs, VERIFY_KEYS), ref["verify"], VERIFY_KEYS, 1e-12, prob, name)


def test_final_fields_match_reference_vtk(oracle_run):
    # the VTK snapshot holds 12 significant digits: tolerance 1e-10 relative L2
    name, prob, ref, o, recs = oracle_run
    assert o.step == ref["vtk_step"] == prob.steps
    idx = vtk_index(prob.nodes, ref["vtk_points"])
    f = o.fetch()
    assert rel_l2(f["u"].reshape(-1, 2)[idx], ref["vtk_u"]) <= 1e-10
    assert rel_l2(f["v"].reshape(-1, 2)[idx], ref["vtk_v"]) <= 1e-10
    assert rel_l2(f["theta"][idx], ref["vtk_theta"]) <= 1e-10


@pytest.mark.parametrize("name", ["lattice", "jitter", "ties"])
def test_horizon_sets_match_reference_kdtree(name):
    z = np.load(__import__("conftest").GOLDEN + "/kdtree_sets.npz")
    nodes, r = z[name + "_nodes"], float(z[name + "_r"])
    rrp, rcols = z[name + "_row_ptr"], z[name + "_cols"]
    rp, cols = pyoracle.build_horizons(nodes, r, 128)
    assert np.array_equal(rp, rrp)
    for i in range(nodes.shape[0]):
        mine, theirs = cols[rp[i]:rp[i + 1]], rcols[rrp[i]:rrp[i + 1]]
        assert mine[0] == i
        assert np.array_equal(np.sort(mine), np.sort(theirs)), i
        d = np.hypot(*(nodes[mine] - nodes[i]).T)
        assert np.all(np.diff(d) >= 0)


def test_prenotch_fixture_is_not_vacuous():
    """The notched case's two segments cut bonds (src/main.cpp:386-402), and without them the
    oracle's series would miss the reference's by far more than the tolerance."""
    prob, ref = load_golden("notch_r12_nu025")
    assert prob.prenotch.shape == (2, 4)
    cut = pyoracle.Oracle(prob, threads=4)
    prob.prenotch = np.zeros((0, 4))
    whole = pyoracle.Oracle(prob, threads=4)
    assert whole.horizons()[0][-1] - cut.horizons()[0][-1] > 1000
    recs = whole.run(prob.steps)
    assert rel_l2(records_table(recs, ["energy"])[:, 0], ref["energy"][:, 1]) > 1e-2


@pytest.mark.parametrize("name", ["abc_r12_nu033_idw", "notch_r12_nu033_idw", "abc_r10_nu045_idw"])
def test_idw_fixtures_are_not_vacuous(name):
    """The IDW runs differ from the G_theta runs of the same inputs by far more than the tolerance,
    and an IDW with another weight (1/d^2 for 1/d) misses the reference's series: the fixtures pin
    the IDW weights, not only which method ran."""
    prob, ref = load_golden(name)
    assert ref["theta_method"] == 0
    gth = pyoracle.Oracle(prob, theta_method=1, threads=4).run(prob.steps)
    e_ref = ref["energy"][:, 1]
    assert rel_l2(records_table(gth, ["energy"])[:, 0], e_ref) > 1e-6
    sq = pyoracle.Oracle(prob, theta_method=0, threads=4, plant=1)
    e_sq = records_table(sq.run(prob.steps), ["energy"])[:, 0]
    assert rel_l2(e_sq, e_ref) > 1e-9
