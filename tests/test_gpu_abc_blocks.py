"""ABC operator blocks shared by lattice images. On an origin-centred lattice domain whose operators
were built once per canonical row and mapped by the lattice symmetries (make_problem(symmetric=True)),
every ABC row's G_u, G_v and G_theta are exact images of its canonical row's, and the library stores
only the canonical rows' operators. The run equals the run with an operator block per row
(LPS_ABC_SYMMETRY=off) up to the order of each image row's member sums (the row kernels add an image
row's members in its block's order): 1e-12; operators that are images only up to rounding, or one
changed coefficient, keep a block per row."""
import dataclasses
import sys

import numpy as np
import pytest

from lps_amd import configs, preprocess as pp, solver

pytestmark = pytest.mark.gpu

SHARED = "lattice images"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sym40():
    return pp.make_problem(40.0, 0.3, steps=300, dt=1.7e-8, dk_factor=0.7,
                           pulses=[pp.Pulse(8.0, -5.0, 12.0, 1.0, (1.0, 1.0))], workers=8,
                           symmetric=True)


def run(prob, n, batch_with=None, **kw):
    if batch_with is None:
        with solver.Solver(prob, variant=3, **kw) as s:
            recs = s.run(n)
            return s.path, [(s.fetch(), recs)]
    with solver.Batch([batch_with, prob], variant=3, **kw) as b:
        recs = b.run(n)
        return b.path, [(b.fetch(i), recs[i]) for i in range(2)]


def assert_same(a, b, tol=1e-12):
    """Fields within tol relative L2 and records within tol relative (the same operators, another
    order of each image row's member sums)."""
    assert len(a) == len(b)
    for (fa, ra), (fb, rb) in zip(a, b):
        for k in ("u", "v", "a", "theta"):
            d = np.linalg.norm(fa[k] - fb[k])
            assert d <= tol * np.linalg.norm(fb[k]), (k, d / np.linalg.norm(fb[k]))
        assert len(ra) == len(rb)
        for x, y in zip(ra, rb):
            assert x["step"] == y["step"]
            for key in ("energy", "energy_k", "energy_p", "max_u", "max_v", "max_a", "theta_max",
                        "ub_rms", "vb_rms"):
                assert abs(x[key] - y[key]) <= tol * max(abs(y[key]), 1e-300), (key, x[key], y[key])


@pytest.mark.parametrize("graph", [True, False])
def test_shared_blocks_give_the_bits_of_blocks_per_row(sym40, graph, monkeypatch):
    say("shared ABC blocks", graph)
    path, shared = run(sym40, 150, use_graph=graph)
    assert SHARED in path, path
    monkeypatch.setenv("LPS_ABC_SYMMETRY", "off")
    path, own = run(sym40, 150, use_graph=graph)
    assert SHARED not in path, path
    assert_same(shared, own)


def test_batch_of_shared_blocks(sym40, monkeypatch):
    say("shared ABC blocks, batch")
    dt, dk = configs.sweep_pair(0.4)
    other = pp.make_problem(20.0, 0.4, steps=300, dt=dt, dk_factor=dk,
                            pulses=[pp.Pulse(3.0, 2.0, 6.0, 1.0, (0.0, 1.0))], workers=8,
                            symmetric=True)
    path, shared = run(sym40, 120, batch_with=other)
    assert SHARED in path, path
    monkeypatch.setenv("LPS_ABC_SYMMETRY", "off")
    path, own = run(sym40, 120, batch_with=other)
    assert SHARED not in path, path
    assert_same(shared, own)


def test_rows_that_are_not_exact_images_keep_their_own_blocks(sym40):
    """One coefficient of one non-canonical row changed by one ulp: no shared blocks, and the run
    is the run of those operators (the tile path against the exact path, 1e-10)."""
    say("inexact images")
    Gu = sym40.Gu.copy()
    nodes = sym40.nodes[sym40.abc]
    r = int(np.nonzero(nodes[:, 0] < -10.0)[0][0])  # a row outside the canonical octant
    j = int(np.nonzero(Gu[2 * r] != 0.0)[0][0])
    Gu[2 * r, j] = np.nextafter(Gu[2 * r, j], np.inf)
    bent = dataclasses.replace(sym40, Gu=Gu)
    path, got = run(bent, 150)
    assert SHARED not in path, path
    with solver.Solver(bent, exact=True) as s:
        s.run(150)
        want = s.fetch()
    for k in ("u", "v"):
        d = np.linalg.norm(got[0][0][k] - want[k]) / np.linalg.norm(want[k])
        assert d <= 1e-10, (k, d)


def test_row_by_row_operators_keep_their_own_blocks():
    """Operators solved row by row (as the reference's preprocessor does) are lattice images only up
    to rounding: every row keeps its own block."""
    say("row-by-row operators")
    prob = pp.make_problem(20.0, 0.3, steps=60, dt=1.7e-8, dk_factor=0.7,
                           pulses=[pp.Pulse(0.0, 0.0, 6.0, 1.0, (1.0, 0.0))], workers=8)
    with solver.Solver(prob, variant=3) as s:
        assert SHARED not in s.path, s.path
