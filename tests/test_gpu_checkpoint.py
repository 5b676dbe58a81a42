"""Checkpoint and restart (lps_checkpoint_save / lps_checkpoint_load): a run stopped at any step,
written out, restored into a fresh context of the same problem and continued, gives the same bits
(fields and output records) as the run that was never stopped, on the fast path, the exact path
and a batch. Files of another problem, truncated files and other files are refused, and a refused
file leaves the context as it was."""
import numpy as np
import pytest
from conftest import load_golden

from lps_amd import preprocess as pp, solver

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "a", "theta")


@pytest.fixture(scope="module")
def r40():
    return pp.make_problem(40.0, 0.3, steps=600, dt=1.7e-8, dk_factor=0.7,
                           pulses=[pp.Pulse(8.0, -5.0, 12.0, 1.0, (1.0, 1.0))], workers=8)


def assert_same(a_fields, a_recs, b_fields, b_recs):
    for k in FIELDS:
        assert np.array_equal(a_fields[k], b_fields[k]), k
    assert len(a_recs) == len(b_recs)
    for r, q in zip(a_recs, b_recs):
        assert r == q, (r, q)


@pytest.mark.parametrize("exact,stop", [(False, 270), (False, 137), (True, 137)])
def test_restart_equals_uninterrupted_run(tmp_path, r40, exact, stop):
    print("restart", exact, stop, flush=True)
    total = 600
    kw = dict(exact=exact, variant=0 if exact else 3)
    with solver.Solver(r40, **kw) as s:
        full = s.run(total)
        full_f = s.fetch()
    path = str(tmp_path / "r40.ckpt")
    with solver.Solver(r40, **kw) as s:
        first = s.run(stop)
        s.checkpoint(path)
    with solver.Solver(r40, **kw) as s:
        s.restore(path)
        assert s.step_index == stop
        rest = s.run(total - stop)
        rest_f = s.fetch()
    assert_same(full_f, full, rest_f, first + rest)


def test_batch_restart(tmp_path, r40):
    print("batch restart", flush=True)
    g12, _ = load_golden("abc_r12_nu033")
    probs = [g12, r40]
    path = str(tmp_path / "batch.ckpt")
    with solver.Batch(probs, variant=3) as b:
        full = b.run(300)
        full_f = [b.fetch(i) for i in range(2)]
    with solver.Batch(probs, variant=3) as b:
        first = b.run(120)
        b.checkpoint(path)
    with solver.Batch(probs, variant=3) as b:
        b.restore(path)
        rest = b.run(180)
        rest_f = [b.fetch(i) for i in range(2)]
    for i in range(2):
        assert_same(full_f[i], full[i], rest_f[i], first[i] + rest[i])


def test_refused_files_leave_the_context_unchanged(tmp_path, r40):
    print("refusals", flush=True)
    g12, _ = load_golden("abc_r12_nu033")
    other = str(tmp_path / "other.ckpt")
    with solver.Solver(g12, variant=3) as s:
        s.run(60)
        s.checkpoint(other)
    good = str(tmp_path / "good.ckpt")
    with solver.Solver(r40, variant=3) as s:
        s.run(90)
        s.checkpoint(good)
    blob = open(good, "rb").read()
    truncated = str(tmp_path / "truncated.ckpt")
    open(truncated, "wb").write(blob[: len(blob) // 2])
    garbage = str(tmp_path / "garbage.ckpt")
    open(garbage, "wb").write(b"not a checkpoint at all" * 4)
    with solver.Solver(r40, variant=3) as ref:
        ref.run(150)
        want = ref.fetch()
    with solver.Solver(r40, variant=3) as s:
        s.run(60)
        for bad, text in ((other, "another problem"), (truncated, "truncated"),
                          (garbage, "not a checkpoint"), (str(tmp_path / "missing"), "cannot open")):
            with pytest.raises(solver.LpsError, match=text):
                s.restore(bad)
        assert s.step_index == 60
        s.run(90)
        got = s.fetch()
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k


def test_checkpoint_of_other_abc_operators_is_refused(tmp_path, r40):
    """Two problems that differ only in their ABC operators (the same lattice, nu and dt, the
    operators built with another delta_kappa, as the (dt, delta_kappa) sweep varies them): the
    checkpoint of one is refused by the other, which stays as it was."""
    print("refuse other operators", flush=True)
    other_dk = pp.make_problem(40.0, 0.3, steps=600, dt=1.7e-8, dk_factor=0.9,
                               pulses=[pp.Pulse(8.0, -5.0, 12.0, 1.0, (1.0, 1.0))], workers=8)
    assert np.array_equal(other_dk.nodes, r40.nodes) and np.array_equal(other_dk.ic, r40.ic)
    assert not np.array_equal(other_dk.Gu, r40.Gu)
    path = str(tmp_path / "dk07.ckpt")
    with solver.Solver(r40, variant=3) as s:
        s.run(90)
        s.checkpoint(path)
    with solver.Solver(other_dk, variant=3) as ref:
        ref.run(60)
        want = ref.fetch()
    with solver.Solver(other_dk, variant=3) as s:
        s.run(30)
        with pytest.raises(solver.LpsError, match="another problem"):
            s.restore(path)
        assert s.step_index == 30
        s.run(30)
        got = s.fetch()
    for k in FIELDS:
        assert np.array_equal(got[k], want[k]), k
