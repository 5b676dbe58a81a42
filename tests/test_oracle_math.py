"""Closed forms and invariants that pin the oracle independently of the reference binary:
horizon counts and symmetry, brute-force neighbour sets, the weighted volume, the dilatation and force
of affine fields and rigid motions, the plane-wave dispersion matrix, momentum balance, the strain
energy of an isotropic strain, and energy conservation without ABCs."""
import math

import numpy as np
import pytest

import pyoracle
from lps_amd import preprocess as pp

SEARCH_R = 2.0 + 0.5 / 2.0 + 1e-5


def no_abc_problem(R, nu, u0=None, v0=None, dt=1e-8):
    mc = pp.material(nu)
    nodes = pp.lattice_domain(R)
    N = nodes.shape[0]
    data2 = np.array([mc.E, nu, mc.rho, 0, 0, 0, dt, 2.0, 0.5, 0.5, mc.s0, 0, 0, 0, 0])
    return pp.Problem(nodes=nodes, steps=0, data2=data2, bctypes=-np.ones(2 * N, dtype=np.int64),
                      bcvalues=np.zeros(2 * N), ic=np.zeros(2 * N) if u0 is None else u0,
                      icd=np.zeros(2 * N) if v0 is None else v0, abc=np.zeros(0, np.int64),
                      clouds=np.zeros((0, 0), np.int64), Gu=np.zeros((0, 0)), Gv=np.zeros((0, 0)),
                      Gth=np.zeros((0, 0)))


def interior(nodes, R, depth):
    return np.hypot(nodes[:, 0], nodes[:, 1]) <= R - depth


def test_interior_horizon_is_the_68_offset_template():
    nodes = pp.lattice_domain(8.0)
    rp, cols = pyoracle.build_horizons(nodes, SEARCH_R)
    lens = np.diff(rp)
    inner = interior(nodes, 8.0, 2.5)
    assert np.all(lens[inner] == 69)  # self + 68 bonds
    i = int(np.argmin(np.hypot(nodes[:, 0], nodes[:, 1])))
    d = np.hypot(*(nodes[cols[rp[i] + 1:rp[i + 1]]] - nodes[i]).T)
    vals, counts = np.unique(np.round(d, 12), return_counts=True)
    assert len(vals) == 12
    assert list(counts) == [4, 4, 4, 8, 4, 4, 8, 8, 4, 8, 4, 8]


def test_horizons_are_symmetric_and_equal_brute_force():
    rng = np.random.default_rng(7)
    nodes = rng.uniform(-4, 4, size=(300, 2))
    rp, cols = pyoracle.build_horizons(nodes, 1.3)
    for i in range(len(nodes)):
        dxy = nodes - nodes[i]
        d = np.sqrt(dxy[:, 0] * dxy[:, 0] + dxy[:, 1] * dxy[:, 1])
        want = np.nonzero(d <= 1.3)[0]
        got = cols[rp[i]:rp[i + 1]]
        assert got[0] == i
        assert np.array_equal(np.sort(got), want)
        for j in got[1:]:
            assert i in cols[rp[j]:rp[j + 1]]


def test_horizon_overflow_is_an_error():
    nodes = pp.lattice_domain(3.0, 0.1, 0.1)
    with pytest.raises(RuntimeError):
        pyoracle.build_horizons(nodes, 1.0, 128)


def test_weighted_volume_of_interior_nodes_is_the_template_value():
    o = pyoracle.Oracle(no_abc_problem(8.0, 0.25))
    m = o.fetch()["m"]
    nodes = pp.lattice_domain(8.0)
    mc = pp.material(0.25)
    inner = interior(nodes, 8.0, 2.5)
    assert np.allclose(m[inner], 25.95427624970138, rtol=1e-14, atol=0)
    assert abs(o.m_ref - 25.95427624970138) <= 1e-13
    assert abs(mc.m_ref - 25.95427624970138) <= 1e-13
    # continuum value pi delta^4 / 2 within the reference's 5 % warning band (src/main.cpp:453)
    assert abs(o.m_ref - math.pi * 16 / 2) / (math.pi * 16 / 2) < 0.05


@pytest.mark.parametrize("nu", [0.25, 0.33])
def test_dilatation_and_force_of_affine_fields(nu):
    R = 10.0
    nodes = pp.lattice_domain(R)
    o = pyoracle.Oracle(no_abc_problem(R, nu))
    eps, gam, om = 1e-3, 2e-3, 3e-3
    x, y = nodes[:, 0], nodes[:, 1]
    inner = interior(nodes, R, 2.5)
    deep = interior(nodes, R, 5.0)
    cases = {
        "isotropic": (np.stack([eps * x, eps * y], 1), 2 * eps),
        "shear": (np.stack([gam * y, gam * x], 1), 0.0),
        "translation": (np.tile([0.3, -0.7], (len(x), 1)), 0.0),
        "rotation": (np.stack([-om * y, om * x], 1), 0.0),
    }
    for name, (u, theta_exp) in cases.items():
        th = o.dilatation(u.reshape(-1))
        assert np.allclose(th[inner], theta_exp, atol=1e-15, rtol=1e-12), name
        acc = o.force(u.reshape(-1), th).reshape(-1, 2)
        scale = 8.0 * pp.material(nu).mu / 25.95 * 2.0 * 1e-3 / pp.material(nu).rho
        assert np.max(np.abs(acc[deep])) <= 1e-10 * scale, name
        if name in ("translation", "rotation"):
            assert np.max(np.abs(th)) <= 1e-15, name


def test_force_balance_sums_to_zero():
    R = 8.0
    nodes = pp.lattice_domain(R)
    o = pyoracle.Oracle(no_abc_problem(R, 0.33))
    rng = np.random.default_rng(3)
    u = rng.normal(size=2 * len(nodes)) * 1e-3
    th = o.dilatation(u)
    acc = o.force(u, th).reshape(-1, 2)
    assert np.all(np.abs(acc.sum(0)) <= 1e-12 * np.abs(acc).sum(0))


@pytest.mark.parametrize("nu", [0.25, 0.33])
def test_plane_wave_force_equals_dispersion_matrix(nu):
    """For u = g cos(a.x) the interior force is M g cos(a.x) with
    M = (2 lambda_PD / m) S S^T + 2 alpha N (preprocessor/main.cpp:199-246): S, N are the
    template sums of xi (e^{i a.xi} - 1) and xi xi^T/|xi|^2 (e^{i a.xi} - 1)."""
    R = 14.0
    mc = pp.material(nu)
    nodes = pp.lattice_domain(R)
    o = pyoracle.Oracle(no_abc_problem(R, nu))
    tg = pp.template_grid(2.0, 0.5, 0.5)
    a = np.array([0.31, -0.17])
    g = np.array([0.6, 0.8]) * 1e-3
    xin = np.hypot(tg[:, 0], tg[:, 1])
    b = np.array([pp.beta(r, 2.0, 0.5) for r in xin]) * 0.25
    ef = np.exp(1j * (tg @ a)) - 1.0
    S = (tg * (ef * b)[:, None]).sum(0)
    Nm = np.einsum("kp,kq,k->pq", tg / xin[:, None], tg / xin[:, None], ef * b)
    M = (2.0 * mc.lambda_pd / mc.m_ref) * np.outer(S, S) + 2.0 * mc.alpha_lps * Nm
    assert np.max(np.abs(M.imag)) <= 1e-12 * np.max(np.abs(M.real))
    phase = nodes @ a
    u = (g[None, :] * np.cos(phase)[:, None]).reshape(-1)
    th = o.dilatation(u)
    acc = o.force(u, th).reshape(-1, 2)
    want = (M.real @ g)[None, :] * np.cos(phase)[:, None] / mc.rho
    deep = interior(nodes, R, 5.0)
    err = np.max(np.abs(acc[deep] - want[deep])) / np.max(np.abs(want[deep]))
    assert err <= 1e-11
    # theta of the plane wave: (2/m) e^{i a.x} g.S, real part
    th_want = np.real((2.0 / mc.m_ref) * (g @ S) * np.exp(1j * phase))
    inner = interior(nodes, R, 2.5)
    assert np.max(np.abs(th[inner] - th_want[inner])) <= 1e-12 * np.max(np.abs(th_want))


def test_isotropic_strain_energy_is_two_kappa_eps_squared():
    R = 10.0
    nodes = pp.lattice_domain(R)
    eps = 1e-3
    u = np.stack([eps * nodes[:, 0], eps * nodes[:, 1]], 1).reshape(-1)
    mc = pp.material(0.33)
    # zero time step: the state after one step is the initial field itself
    o = pyoracle.Oracle(no_abc_problem(R, 0.33, u0=u, dt=0.0))
    o.run(1)
    d = o.diagnostics()
    f = o.fetch()
    inner = interior(nodes, R, 2.5)
    # W_i = 2 kappa eps^2 at interior nodes: energy = sum W_i dV over all nodes; check the interior part
    w_expected = 2.0 * mc.kappa * eps * eps
    assert np.allclose(f["theta"][inner], 2 * eps, rtol=1e-12)
    total_lower = w_expected * 0.25 * inner.sum()
    assert d["energy_p"] >= total_lower * (1 - 1e-12)
    assert d["energy_k"] == 0.0


def test_energy_is_conserved_without_abc():
    """Velocity Verlet without ABCs: the energy oscillates without drift."""
    R = 12.0
    nodes = pp.lattice_domain(R)
    r2 = nodes[:, 0] ** 2 + nodes[:, 1] ** 2
    val = np.where(np.sqrt(r2) < 6.0, np.exp(-r2 / 9.0), 0.0) * 1e-3
    u = np.stack([val, val], 1).reshape(-1)
    o = pyoracle.Oracle(no_abc_problem(R, 0.3, u0=u, dt=1.5e-8))
    recs = o.run(600)
    e = np.array([r["energy"] for r in recs])
    assert np.max(np.abs(e / e[0] - 1.0)) < 1e-3
    p = np.array([[r["px"], r["py"]] for r in recs])
    assert np.max(np.abs(p)) <= 1e-12 * math.sqrt(2 * e[0] * 8e-15 * 0.25 * len(nodes))


def test_idw_abc_dilatation_of_an_isotropic_strain():
    """IDW at the ABC nodes (src/main.cpp:765-790) for u = eps x: every non-ABC node has
    theta = 2 eps (its m_i is the sum over its own horizon), so the weighted mean over an ABC
    node's non-ABC neighbours is exactly 2 eps, while the direct value, biased by the m_i := m_ref
    override on the truncated horizons of the outer ring, is not."""
    R, eps, delta = 10.0, 1e-3, 2.0
    mc = pp.material(0.25)
    nodes = pp.lattice_domain(R)
    N = nodes.shape[0]
    abc = np.nonzero(np.hypot(nodes[:, 0], nodes[:, 1]) > R - delta)[0]
    C, M = len(abc), 49
    bctypes = -np.ones(2 * N, dtype=np.int64)
    bctypes[2 * abc] = 1
    bctypes[2 * abc + 1] = 1
    x, y = nodes[:, 0], nodes[:, 1]
    data2 = np.array([mc.E, 0.25, mc.rho, 0, 0, 0, 1.7e-8, delta, 0.5, 0.5, mc.s0, 0, 0, 0, 0])
    prob = pp.Problem(nodes=nodes, steps=1, data2=data2, bctypes=bctypes, bcvalues=np.zeros(2 * N),
                      ic=np.stack([eps * x, eps * y], 1).reshape(-1), icd=np.zeros(2 * N),
                      abc=abc.astype(np.int64), clouds=np.zeros((C, M), np.int64),
                      Gu=np.zeros((2 * C, 4 * M)), Gv=np.zeros((2 * C, 4 * M)), Gth=np.zeros((C, 2 * M)))
    o = pyoracle.Oracle(prob, theta_method=0)
    direct = o.dilatation(prob.ic)
    o.run(1)
    th = o.fetch()["theta"]
    is_abc = np.zeros(N, bool)
    is_abc[abc] = True
    assert np.allclose(th, 2 * eps, rtol=1e-12, atol=0)
    assert np.max(np.abs(direct[is_abc] / (2 * eps) - 1)) > 1e-2


def _exact_intersect(a, b, c, d):
    """Closed-segment intersection in exact rational arithmetic, by another method than the
    oracle's: solve a + t (b - a) = c + s (d - c) for t, s in [0, 1]; parallel segments intersect
    when they are collinear and their projections on the common line overlap."""
    from fractions import Fraction as F
    a, b, c, d = ([F(v) for v in p] for p in (a, b, c, d))
    r = (b[0] - a[0], b[1] - a[1])
    q = (d[0] - c[0], d[1] - c[1])
    w = (c[0] - a[0], c[1] - a[1])
    den = r[0] * q[1] - r[1] * q[0]
    if den != 0:
        t = (w[0] * q[1] - w[1] * q[0]) / den
        s = (w[0] * r[1] - w[1] * r[0]) / den
        return 0 <= t <= 1 and 0 <= s <= 1
    if w[0] * r[1] - w[1] * r[0] != 0 or (r == (0, 0) and q == (0, 0) and w != (0, 0)):
        return False  # parallel on different lines, or two distinct points
    # collinear: overlap of the projections on the line's dominant axis
    k = 0 if max(abs(r[0]), abs(q[0]), abs(w[0])) > 0 else 1
    lo1, hi1 = sorted((a[k], b[k]))
    lo2, hi2 = sorted((c[k], d[k]))
    return max(lo1, lo2) <= min(hi1, hi2)


def test_prenotch_intersection_exact_on_grid_points():
    """On coordinates that are multiples of 0.5 every product in the test is exact, so the
    visibility test (src/main.cpp:119-140) must equal the exact predicate, including collinear,
    touching and overlapping segments."""
    rng = np.random.default_rng(7)
    pts = rng.integers(-4, 5, size=(20000, 8)) * 0.5
    # force many degenerate configurations: shared end points and collinear pieces
    pts[::7, 4:6] = pts[::7, 0:2]
    pts[1::7, 4:6] = 2 * pts[1::7, 2:4] - pts[1::7, 0:2]
    pts[2::7, 6:8] = (pts[2::7, 0:2] + pts[2::7, 2:4]) / 2
    pts[3::7, 1] = pts[3::7, 3] = pts[3::7, 5] = pts[3::7, 7] = 1.0
    agree, hits = 0, 0
    for row in pts:
        a, b, c, d = row[0:2], row[2:4], row[4:6], row[6:8]
        if np.array_equal(a, b) or np.array_equal(c, d):
            continue  # zero-length segments: not a bond or a notch
        want = _exact_intersect(a, b, c, d)
        assert pyoracle.segments_intersect(a, b, c, d) == want, (a, b, c, d, want)
        agree += 1
        hits += want
    assert agree > 15000 and 0.2 < hits / agree < 0.8


def test_prenotch_intersection_random_segments():
    rng = np.random.default_rng(8)
    for row in rng.uniform(-3.0, 3.0, size=(5000, 8)):
        a, b, c, d = row[0:2], row[2:4], row[4:6], row[6:8]
        assert pyoracle.segments_intersect(a, b, c, d) == _exact_intersect(a, b, c, d)
        assert pyoracle.segments_intersect(b, a, d, c) == _exact_intersect(a, b, c, d)
