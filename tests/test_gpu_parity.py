"""The HIP path against the oracle and the reference's golden outputs, through the C ABI."""
import re
import sys

import numpy as np
import pytest
from conftest import (ENERGY_KEYS, VERIFY_KEYS, assert_series_close, load_golden, records_table,
                      rel_l2, vtk_index)

import pyoracle
from lps_amd import preprocess as pp, solver

pytestmark = pytest.mark.gpu

CASES = ["noabc_r30", "abc_r10_nu025", "abc_r12_nu033", "notch_r12_nu025",
         # the reference solver built with the IDW ABC dilatation (src/main.cpp:765-790)
         "abc_r12_nu033_idw", "notch_r12_nu033_idw", "abc_r10_nu045_idw"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["lattice", "jitter", "ties"])
def test_gpu_horizons_equal_oracle_and_reference_sets(name):
    say("horizons", name)
    z = np.load(__import__("conftest").GOLDEN + "/kdtree_sets.npz")
    nodes, r = z[name + "_nodes"], float(z[name + "_r"])
    n = nodes.shape[0]
    prob = pp.Problem(nodes=nodes, steps=0, data2=np.array([0.19, 0.25, 8e-15, 0, 0, 0, 1e-8, 2.0,
                                                            0.5, 0.5, 0, 0, 0, 0, 0]),
                      bctypes=-np.ones(2 * n, dtype=np.int64), bcvalues=np.zeros(2 * n),
                      ic=np.zeros(2 * n), icd=np.zeros(2 * n), abc=np.zeros(0, np.int64),
                      clouds=np.zeros((0, 0), np.int64), Gu=np.zeros((0, 0)), Gv=np.zeros((0, 0)),
                      Gth=np.zeros((0, 0)))
    with solver.Solver(prob, exact=True, horizon_radius=r) as s:
        rp, cols = s.horizons()
    orp, ocols = pyoracle.build_horizons(nodes, r, 128)
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)  # same sets, same order
    rrp, rcols = z[name + "_row_ptr"], z[name + "_cols"]
    assert np.array_equal(rp, rrp)
    for i in range(n):
        assert np.array_equal(np.sort(cols[rp[i]:rp[i + 1]]), np.sort(rcols[rrp[i]:rrp[i + 1]]))


@pytest.mark.parametrize("name", CASES)
def test_exact_mode_matches_oracle_bitwise(name):
    """Exact mode keeps the reference's operation order and neighbour order: the fields after the
    run equal the oracle's to the last bit."""
    say("exact vs oracle", name)
    prob, ref = load_golden(name)
    o = pyoracle.Oracle(prob, theta_method=ref["theta_method"])
    orecs = o.run(prob.steps)
    of = o.fetch()
    with solver.Solver(prob, exact=True, theta_method=ref["theta_method"]) as s:
        grecs = s.run(prob.steps)
        gf = s.fetch()
    for k in ("u", "v", "a", "theta", "m"):
        assert np.array_equal(gf[k], of[k]), (k, np.max(np.abs(gf[k] - of[k])))
    # output block: same per-node terms, reduction order differs (1e-14)
    assert_series_close(records_table(grecs, ENERGY_KEYS), records_table(orecs, ENERGY_KEYS),
                        ENERGY_KEYS, 1e-14, prob, name)


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("graph,variant,concurrent", [(True, 0, 1), (False, 0, 1), (True, 3, 1),
                                                      (True, 2, 0), (True, 0, 3), (False, 3, 3),
                                                      (True, 3, 4), (False, 0, 4), (True, 3, 2),
                                                      (False, 0, 2)])
def test_fast_mode_matches_reference_golden(name, graph, variant, concurrent):
    """Fast path (tile kernels, FMA) against the reference's outputs: 1e-10 relative L2; autotuned,
    register-blocked and unrolled node kernels, ABC rows beside or after the tile kernels."""
    say("fast vs reference golden", name, graph, variant, concurrent)
    prob, ref = load_golden(name)
    with solver.Solver(prob, exact=False, use_graph=graph, variant=variant,
                       concurrent=concurrent, theta_method=ref["theta_method"]) as s:
        recs = s.run(prob.steps)
        f = s.fetch()
    assert_series_close(records_table(recs, ENERGY_KEYS), ref["energy"], ENERGY_KEYS, 1e-10, prob, name)
    assert_series_close(records_table(recs, VERIFY_KEYS), ref["verify"], VERIFY_KEYS, 1e-10, prob, name)
    idx = vtk_index(prob.nodes, ref["vtk_points"])
    assert rel_l2(f["u"].reshape(-1, 2)[idx], ref["vtk_u"]) <= 1e-10
    assert rel_l2(f["v"].reshape(-1, 2)[idx], ref["vtk_v"]) <= 1e-10
    assert rel_l2(f["theta"][idx], ref["vtk_theta"]) <= 1e-10


@pytest.mark.parametrize("name", ["abc_r12_nu033_idw", "notch_r12_nu033_idw",
                                  "abc_r10_nu045_idw"])
def test_exact_idw_matches_reference_golden(name):
    """The exact path with the IDW ABC dilatation against the reference solver built with it:
    series and last fields within 1e-10 (the fast path is in test_fast_mode_matches_reference_golden,
    the exact path against the oracle bit for bit in test_exact_mode_matches_oracle_bitwise)."""
    say("exact idw vs reference golden", name)
    prob, ref = load_golden(name)
    assert ref["theta_method"] == solver.LPS_THETA_IDW
    with solver.Solver(prob, exact=True, theta_method=solver.LPS_THETA_IDW) as s:
        recs = s.run(prob.steps)
        f = s.fetch()
    assert_series_close(records_table(recs, ENERGY_KEYS), ref["energy"], ENERGY_KEYS, 1e-10, prob, name)
    assert_series_close(records_table(recs, VERIFY_KEYS), ref["verify"], VERIFY_KEYS, 1e-10, prob, name)
    idx = vtk_index(prob.nodes, ref["vtk_points"])
    assert rel_l2(f["u"].reshape(-1, 2)[idx], ref["vtk_u"]) <= 1e-10
    assert rel_l2(f["v"].reshape(-1, 2)[idx], ref["vtk_v"]) <= 1e-10
    assert rel_l2(f["theta"][idx], ref["vtk_theta"]) <= 1e-10


def test_nonfinite_state_is_reported_at_first_output_step():
    say("nonfinite")
    prob, _ = load_golden("abc_r10_nu025")
    prob.ic = prob.ic.copy()
    prob.ic[10] = np.nan
    with solver.Solver(prob) as s:
        with pytest.raises(solver.LpsError) as e:
            s.run(90)
        assert e.value.code == solver.LPS_ERR_NONFINITE
        assert s.step_index == 30


def test_lattice_inputs_take_the_tile_path_and_others_fall_back():
    say("path selection")
    prob, _ = load_golden("abc_r10_nu025")
    with solver.Solver(prob) as s:
        assert s.path.startswith("tile"), s.path
    with solver.Solver(prob, exact=True) as s:
        assert s.path.startswith("exact"), s.path
    jit = pp.Problem(**{**prob.__dict__})
    rng = np.random.default_rng(5)
    jit.nodes = prob.nodes + rng.uniform(-0.05, 0.05, size=prob.nodes.shape)
    with solver.Solver(jit) as s:
        assert s.path.startswith("exact"), s.path
        recs = s.run(60)
    o = pyoracle.Oracle(jit)
    orecs = o.run(60)
    assert abs(recs[-1]["energy"] - orecs[-1]["energy"]) <= 1e-13 * abs(orecs[-1]["energy"])


@pytest.fixture(scope="module")
def r40_problem():
    return pp.make_problem(40.0, 0.3, steps=600, dt=1.7e-8, dk_factor=0.7,
                           pulses=[pp.Pulse(8.0, -5.0, 12.0, 1.0, (1.0, 1.0))], workers=8)


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_tile_path_matches_oracle_on_a_larger_domain(variant, r40_problem):
    """R = 40 mm (20k nodes, 121 tiles, full and partial ones), nu = 0.3, off-centre pulse,
    600 steps; autotuned, generic, unrolled and register-blocked node kernels."""
    say("tile R=40 variant", variant)
    prob = r40_problem
    names = {1: "generic", 2: "unrolled", 3: "blocked"}
    with solver.Solver(prob, variant=variant) as s:
        assert s.path.startswith("tile")
        if variant:
            want = names[variant]
            assert f"dilatation {want}, force {want}" in s.path, s.path
        recs = s.run(600)
        f = s.fetch()
    o = pyoracle.Oracle(prob)
    orecs = o.run(600)
    of = o.fetch()
    for k in ("u", "v", "theta"):
        assert rel_l2(f[k], of[k]) <= 1e-10, k
    assert_series_close(records_table(recs, ENERGY_KEYS), records_table(orecs, ENERGY_KEYS),
                        ENERGY_KEYS, 1e-10, prob, "R40")


@pytest.mark.parametrize("exact", [False, True])
def test_members_that_add_nothing_change_no_bit(exact):
    """The ABC rows are ragged: the tile path drops each cloud's trailing members that add nothing
    (pads, -1, and nodes whose G_u, G_v and G_theta coefficients are all zero, as the reference's
    reader turns -1 into node 0). A problem with seven such members appended to every cloud runs
    to the same bits as the problem itself."""
    say("padded clouds", exact)
    import dataclasses
    prob, _ = load_golden("abc_r12_nu033")
    C = prob.C
    extra = np.array([-1, 0, -1, 0, 0, -1, 0] * C).reshape(C, 7)
    pad = lambda G, w: np.concatenate([G, np.zeros((G.shape[0], w * 7))], axis=1)
    big = dataclasses.replace(prob, clouds=np.concatenate([prob.clouds, extra], axis=1),
                              Gu=pad(prob.Gu, 4), Gv=pad(prob.Gv, 4), Gth=pad(prob.Gth, 2))
    out = []
    for p in (prob, big):
        with solver.Solver(p, exact=exact, variant=0 if exact else 3) as s:
            recs = s.run(150)
            out.append((s.fetch(), recs))
    for k in ("u", "v", "a", "theta"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("graph", [True, False])
def test_abc_prediction_folded_or_apart_gives_the_same_bits(graph, monkeypatch, r40_problem):
    """The ABC nodes' prediction folded into the partial tiles' force pass, or a kernel of its own
    after G_u / G_v with the partial tiles beside them (the autotuner times both): the same bits,
    on one domain and on a batch."""
    say("prediction fold/apart", graph)
    g12, _ = load_golden("abc_r12_nu033")
    out = []
    for mode in ("fold", "apart"):
        monkeypatch.setenv("LPS_ABC_PREDICT", mode)
        # schedule 1: the fold applies to the partial tiles' force pass on the main stream
        with solver.Solver(r40_problem, variant=3, use_graph=graph, concurrent=1) as s:
            assert ("ABC prediction " + ("in" if mode == "fold" else "apart")) in s.path, s.path
            recs = s.run(95)
            one = (s.fetch(), recs)
        with solver.Batch([g12, r40_problem], variant=3, use_graph=graph, concurrent=1) as b:
            brecs = b.run(95)
            many = ([b.fetch(i) for i in range(2)], brecs)
        out.append((one, many))
    (a1, am), (b1, bm) = out
    for k in ("u", "v", "a", "theta"):
        assert np.array_equal(a1[0][k], b1[0][k]), k
        for i in range(2):
            assert np.array_equal(am[0][i][k], bm[0][i][k]), (i, k)
    assert a1[1] == b1[1] and am[1] == bm[1]


@pytest.mark.parametrize("graph", [True, False])
def test_step_schedules_give_the_same_bits(graph, r40_problem):
    """The ABC rows after the tile kernels (0), on a second stream (1), with the partial tiles on a
    third stream in whole-tile blocks (2), as extra workgroups of the dilatation launch (3), or on a
    second stream with the deep tiles' dilatation and force in one launch (4): the same node
    arithmetic in every schedule, so the same bits, on one domain and on a batch."""
    say("schedules", graph)
    g12, _ = load_golden("abc_r12_nu033")
    out = []
    for sched in (0, 1, 2, 3, 4):
        with solver.Solver(r40_problem, variant=3, use_graph=graph, concurrent=sched) as s:
            # the R = 40 mm disc has deep tiles (the fused kernel runs) and shallow ones
            deep = int(re.search(r"(\d+) deep", s.path).group(1))
            full = int(re.search(r"(\d+) full", s.path).group(1))
            assert 0 < deep < full, s.path
            recs = s.run(95)
            one = (s.fetch(), recs)
        with solver.Batch([g12, r40_problem], variant=3, use_graph=graph, concurrent=sched) as b:
            brecs = b.run(95)
            many = ([b.fetch(i) for i in range(2)], brecs)
        out.append((one, many))
    (a1, am) = out[0]
    for b1, bm in out[1:]:
        for k in ("u", "v", "a", "theta"):
            assert np.array_equal(a1[0][k], b1[0][k]), k
            for i in range(2):
                assert np.array_equal(am[0][i][k], bm[0][i][k]), (i, k)
        assert a1[1] == b1[1] and am[1] == bm[1]


@pytest.mark.parametrize("graph", [True, False])
def test_two_wavefronts_per_tile_give_the_same_bits(graph, monkeypatch, r40_problem):
    """The register-blocked dilatation and force kernels with 4 nodes per lane (a wavefront per
    tile) or 2 (two wavefronts per tile sharing its window), in every combination: the same node
    arithmetic, so the same bits, on one domain and on a batch, in schedules 1, 3 and 4."""
    say("nodes per lane", graph)
    g12, _ = load_golden("abc_r12_nu033")
    out = []
    for nodes in ("4,4", "2,2", "2,4", "4,2"):
        monkeypatch.setenv("LPS_RB_NODES", nodes)
        for sched in (1, 3, 4):
            with solver.Solver(r40_problem, variant=3, use_graph=graph, concurrent=sched) as s:
                recs = s.run(95)
                one = (s.fetch(), recs)
            with solver.Batch([g12, r40_problem], variant=3, use_graph=graph, concurrent=sched) as b:
                brecs = b.run(95)
                many = ([b.fetch(i) for i in range(2)], brecs)
            out.append((one, many))
    (a1, am) = out[0]
    for b1, bm in out[1:]:
        for k in ("u", "v", "a", "theta"):
            assert np.array_equal(a1[0][k], b1[0][k]), k
            for i in range(2):
                assert np.array_equal(am[0][i][k], bm[0][i][k]), (i, k)
        assert a1[1] == b1[1] and am[1] == bm[1]


@pytest.mark.parametrize("graph", [True, False])
def test_timing_mode_reports_pass_times_and_changes_no_result(graph, r40_problem):
    """lps_set_timing: per-pass events (mode 1) or the force kernel's two events only (mode 2),
    inside the 30-step graphs when graphs are on; the state afterwards is the same bits as an
    untimed run, and switching modes in between changes nothing."""
    say("timing", graph)
    prob = r40_problem
    with solver.Solver(prob, use_graph=graph) as s:
        s.run(60)
        plain = s.fetch()
        s.reset()
        s.set_timing(True)
        s.run(60)
        pt = s.pass_times()
        s.set_timing(False)
        timed = s.fetch()
        s.reset()
        s.set_timing(2)
        s.run(30)
        s.set_timing(True)
        s.run(1)
        s.set_timing(2)
        s.run(29)
        pf = s.pass_times()
        s.set_timing(False)
        force_only = s.fetch()
    for k in ("u", "v", "theta"):
        assert np.array_equal(plain[k], timed[k]), k
        assert np.array_equal(plain[k], force_only[k]), k
    assert all(pt[k] > 0 for k in ("theta_ms", "force_ms", "step_ms")), pt
    assert pt["step_ms"] >= pt["theta_ms"] + pt["force_ms"], pt
    assert pf["force_ms"] > 0 and pf["theta_ms"] == pf["abc_ms"] == pf["step_ms"] == 0.0, pf


@pytest.mark.parametrize("name", ["abc_r12_nu033", "noabc_r30"])
def test_fields_after_any_step_count_match_oracle(name):
    """The tile path keeps U^{n+1} of the interior nodes only in its six-entry prediction ring and
    copies it into U before a fetch or an output record. Step counts that land on every ring
    phase, with a 30-step graph replay in between, give the oracle's fields and energy."""
    say("ring phases", name)
    prob, _ = load_golden(name)
    o = pyoracle.Oracle(prob)
    with solver.Solver(prob) as s:
        assert s.path.startswith("tile"), s.path
        for k in (1, 1, 3, 1, 7, 17, 31, 2):
            s.step(k)
            o.run(k)
            assert s.step_index == o.step
            f, of = s.fetch(), o.fetch()
            for key in ("u", "v", "a", "theta"):
                assert rel_l2(f[key], of[key]) <= 1e-10, (s.step_index, key)
            e, oe = s.diagnostics()["energy"], o.diagnostics()["energy"]
            assert abs(e - oe) <= 1e-10 * abs(oe), (s.step_index, e, oe)
