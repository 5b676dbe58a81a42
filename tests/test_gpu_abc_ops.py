"""Least-squares ABC operators built on the GPU (lps_abc_operators, lps_abc_system,
lps_lstsq_batched) against closed forms and the CPU builder (lps_amd.preprocess, LAPACK dgelsd as
preprocessor/main.cpp:368-402).

The systems are numerically rank-deficient (about 20 of up to 196 singular values above
1e-14 s_1, the smallest kept ones near 1e-13 s_1), so the solution along the smallest kept
directions is determined only to about eps s_1 / s_r ~ 1e-3, by any backward-stable solver: LAPACK's
dgelsd and dgelss differ on these systems by 1.7e-3 (median over rows) and up to 0.19 on single
rows, with residuals of 1e-15 |b| for both. The operators are therefore compared through what fixes
them: the systems themselves (to rounding), the solver on problems with known minimum-norm
solutions, the least-squares residual, the kept rank, the operators' norms and the runs they
give."""
import math

import numpy as np
import pytest
from conftest import rel_l2

from lps_amd import abc_gpu, preprocess as pp, solver

pytestmark = pytest.mark.gpu


def _orth(rng, n, k):
    q, _ = np.linalg.qr(rng.standard_normal((n, k)))
    return q


@pytest.mark.parametrize("rcond", [-1.0, 0.0, 1.0, 2.0])
def test_lstsq_nonpositive_rcond_means_machine_epsilon_as_in_dgelsd(rcond):
    """dgelsd (LAPACK dlalsd) replaces an rcond <= 0 or >= 1 by the machine epsilon 2^-53 and drops
    s_i <= eps s_1. Diagonal systems with singular values just above and just below that cut: the
    GPU solver keeps the same rank and gives the same solution as scipy's dgelsd."""
    from scipy.linalg import lapack
    print("lstsq rcond", rcond, flush=True)
    eps = 2.0 ** -53
    cases = [[1.0, 1.02 * eps, 0.5], [1.0, 0.98 * eps, 0.5], [2.0, 0.3, 1e-3]]
    A = np.stack([np.diag(c) for c in cases])
    B = np.stack([np.array([[1.0], [1e-16], [2.0]]) for _ in cases])
    X, rank = abc_gpu.lstsq(A, B, rcond)
    for p, c in enumerate(cases):
        lw = lapack.dgelsd_lwork(3, 3, 1, rcond)
        x, _s, r, info = lapack.dgelsd(A[p], B[p].copy(), int(lw[0]), int(lw[1]), cond=rcond)
        assert info == 0
        assert rank[p] == r, (p, rank[p], r)
        assert np.allclose(X[p][:, 0], x[:3, 0], rtol=1e-14, atol=0), (p, X[p], x)
    assert list(rank) == [3, 2, 3]


@pytest.mark.parametrize("m,n,low", [(30, 12, 0), (12, 30, 0), (224, 196, 0), (224, 196, 24)])
def test_lstsq_known_minimum_norm_solutions(m, n, low):
    """A = U diag(s) V' with prescribed singular values, some below rcond s_1: the solution is
    V diag(1 / s_i, s_i > rcond s_1) U' B. low > 0: numerical rank low (the range-compressed path
    of large systems); otherwise most singular values are kept (the full SVD)."""
    print("lstsq", m, n, low, flush=True)
    rng = np.random.default_rng(m * 1000 + n + low)
    k = min(m, n)
    batch, nrhs = 5, 2
    A = np.zeros((batch, m, n))
    B = rng.standard_normal((batch, m, nrhs))
    want = np.zeros((batch, n, nrhs))
    want_rank = np.zeros(batch, dtype=np.int32)
    for p in range(batch):
        s = np.logspace(0, -8, k) if p % 2 == 0 else np.logspace(0, -20, k)
        if low:
            s = np.concatenate([np.logspace(0, -8 if p % 2 == 0 else -12, low),
                                np.full(k - low, 1e-18)])
        s[(s > 1e-16) & (s < 1e-12)] = 1e-18  # none near the cut: rounding of A moves them
        U, V = _orth(rng, m, k), _orth(rng, n, k)
        A[p] = (U * s) @ V.T
        keep = s > 1e-14 * s[0]
        want[p] = V[:, keep] @ ((U[:, keep].T @ B[p]) / s[keep, None])
        want_rank[p] = keep.sum()
    X, rank = abc_gpu.lstsq(A, B, 1e-14)
    for p in range(batch):
        # kept directions down to s = 1e-8 (p even) or 1e-12 (p odd): error ~ eps / s_min
        tol = 1e-6 if p % 2 == 0 else 2e-2
        assert rel_l2(X[p], want[p]) <= tol, (p, rel_l2(X[p], want[p]))
        assert rank[p] == want_rank[p], (p, rank[p], want_rank[p])


def test_lstsq_full_rank_matches_numpy():
    print("lstsq full rank", flush=True)
    rng = np.random.default_rng(3)
    A = rng.standard_normal((64, 40, 16))
    B = rng.standard_normal((64, 40, 2))
    X, rank = abc_gpu.lstsq(A, B, 1e-14)
    for p in range(64):
        want = np.linalg.lstsq(A[p], B[p], rcond=None)[0]
        assert rel_l2(X[p], want) <= 1e-12
    assert np.all(rank == 16)


@pytest.fixture(scope="module")
def r16():
    """An R = 16 mm problem built on the CPU, with what the builder needs. (dt, delta_kappa) are
    the reference sweep's choice for this domain (max ||G|| = 0.9989)."""
    R, nu, dt, dkf = 16.0, 0.25, 1.75e-8, 0.9
    prob = pp.make_problem(R, nu, steps=600, dt=dt, dk_factor=dkf,
                           pulses=[pp.Pulse(3.0, -2.0, 6.0, 1.0, (1.0, 1.0))], workers=8)
    mc = pp.material(nu=nu, delta=2.0, dx=0.5, dy=0.5)
    tg = pp.template_grid(2.0, 0.5, 0.5)
    nodes = prob.nodes
    abc_idx, tau = pp.abc_nodes(nodes, R, 2.0)
    clouds = pp.abc_clouds(nodes, abc_idx, 2.0)
    modes = pp.compute_all_modes(mc, tg, dkf / math.sqrt(mc.c_P)).sorted_by_kappa()
    assert np.array_equal(abc_idx, prob.abc)
    return dict(R=R, nu=nu, dt=dt, dkf=dkf, prob=prob, mc=mc, nodes=nodes, abc=abc_idx, tau=tau,
                clouds=clouds, M=prob.max_cloud, modes=modes)


def _cpu_systems(P, i):
    """The CPU builder's own matrices of row i (captured from preprocess._dgelsd)."""
    cap = []
    orig = pp._dgelsd

    def spy(AT, Bm, rcond=1e-14):
        cap.append((AT.copy(), Bm.copy()))
        return orig(AT, Bm, rcond)

    pp._dgelsd = spy
    try:
        a = P["abc"][i]
        pp.abc_node_uv(P["nodes"][a, 0], P["nodes"][a, 1], P["tau"][i], P["nodes"][P["clouds"][i]],
                       P["modes"], P["M"], P["dt"])
        pp.abc_node_theta(P["nodes"][a, 0], P["nodes"][a, 1], P["tau"][i],
                          P["nodes"][P["clouds"][i]], P["modes"], P["mc"], P["M"], P["dt"])
    finally:
        pp._dgelsd = orig
    return cap  # [(ATu, tu), (ATv, tv), (ATth, tth)]


def test_systems_equal_the_cpu_builder(r16):
    """Same matrices and right-hand sides up to rounding, with the GPU's columns (tl, k, c) at
    tl 2 Mc + 2 k + c against the CPU's tl 2 nC + 2 k + c, and zero columns past the cloud."""
    print("abc systems", flush=True)
    P = r16
    M = P["M"]
    rows = [0, 7, len(P["abc"]) // 2, len(P["abc"]) - 1]
    for kind in (abc_gpu.KIND_U, abc_gpu.KIND_V, abc_gpu.KIND_THETA):
        A, B = abc_gpu.system(P["nodes"], P["abc"], P["tau"], P["clouds"], M, P["modes"], P["dt"],
                              P["mc"].m_ref, kind)
        for i in rows:
            AT, Bc = _cpu_systems(P, i)[kind]
            nC = len(P["clouds"][i])
            per = 2 if kind != abc_gpu.KIND_THETA else 1
            cols = [tl * per * M + per * k + c for tl in range(2) for k in range(nC)
                    for c in range(per)]
            assert np.max(np.abs(A[i][:, cols] - AT)) <= 1e-13 * np.max(np.abs(AT)), (kind, i)
            rest = np.setdiff1d(np.arange(A.shape[2]), cols)
            assert np.all(A[i][:, rest] == 0.0)
            assert np.max(np.abs(B[i] - Bc)) <= 1e-13 * np.max(np.abs(Bc)), (kind, i)


def _x_of_row(G, i, nC, per):
    """Row i's solution in the CPU system's column order (tl per nC + per k + c)."""
    if per == 2:
        x = np.zeros((4 * nC, 2))
        for tl in range(2):
            for k in range(nC):
                for c in range(2):
                    x[tl * 2 * nC + 2 * k + c] = G[2 * i:2 * i + 2, 4 * k + 2 * tl + c]
        return x
    x = np.zeros((2 * nC, 1))
    for tl in range(2):
        for k in range(nC):
            x[tl * nC + k, 0] = G[i, 2 * k + tl]
    return x


def test_operators_match_the_cpu_builder(r16):
    print("abc operators", flush=True)
    P = r16
    prob = P["prob"]
    Gu, Gv, Gth, rank = abc_gpu.operators(P["nodes"], P["abc"], P["tau"], P["clouds"], P["M"],
                                          P["modes"], P["dt"], P["mc"].m_ref)
    assert np.all(rank > 0)  # every SVD converged
    C = len(P["abc"])
    # per row: a least-squares solution (residual at the minimum), the rank LAPACK's SVD sees, and
    # a norm no larger than the CPU's minimum-norm solution beyond the level it is determined to
    same_rank, rows = 0, list(range(0, C, max(1, C // 40)))
    for i in rows:
        systems = _cpu_systems(P, i)
        nC = len(P["clouds"][i])
        for kind, G, cpuG, per in ((0, Gu, prob.Gu, 2), (1, Gv, prob.Gv, 2), (2, Gth, prob.Gth, 1)):
            AT, b = systems[kind]
            x = _x_of_row(G, i, nC, per)
            xc = _x_of_row(cpuG, i, nC, per)
            assert np.linalg.norm(AT @ x - b) <= 1e-12 * np.linalg.norm(b), (i, kind)
            assert np.linalg.norm(x) <= 1.25 * np.linalg.norm(xc), (i, kind)
            s = np.linalg.svd(AT, compute_uv=False)
            same_rank += int(rank[i, kind] == np.sum(s > 1e-14 * s[0]))
    assert same_rank >= 0.9 * 3 * len(rows), same_rank
    # the stability diagnostic of the sweep (preprocessor/main.cpp:946-958): max over rows of ||G||
    n_gpu = max(pp.g_norm(Gu[2 * i:2 * i + 2], Gv[2 * i:2 * i + 2]) for i in range(C))
    n_cpu = max(pp.g_norm(prob.Gu[2 * i:2 * i + 2], prob.Gv[2 * i:2 * i + 2]) for i in range(C))
    assert abs(n_gpu - n_cpu) <= 1e-2 * n_cpu, (n_gpu, n_cpu)
    # the operators as a whole, at the level two LAPACK solvers agree to on these systems
    for mine, ref in ((Gu, prob.Gu), (Gv, prob.Gv), (Gth, prob.Gth)):
        assert rel_l2(mine, ref) <= 0.1, rel_l2(mine, ref)


def test_runs_with_gpu_operators_match_cpu_operators(r16):
    """The time loop on the GPU-built operators against the CPU-built ones: the same series to
    the level the operators agree."""
    print("abc runs", flush=True)
    P = r16
    cpu = P["prob"]
    gpu = pp.make_problem(P["R"], P["nu"], steps=600, dt=P["dt"], dk_factor=P["dkf"],
                          pulses=[pp.Pulse(3.0, -2.0, 6.0, 1.0, (1.0, 1.0))], ls="gpu")
    assert gpu.info["ls"] == "gpu" and gpu.info["ls_unconverged"] == 0
    out = []
    for prob in (cpu, gpu):
        with solver.Solver(prob, variant=3) as s:
            recs = s.run(prob.steps)
        out.append(np.array([[r["energy"], r["de_rel"]] for r in recs]))
    e_cpu, e_gpu = out
    assert np.all(e_cpu[:, 0] <= 1.001 * e_cpu[0, 0])  # a stable run: the ABC takes energy out
    assert rel_l2(e_gpu[:, 0], e_cpu[:, 0]) <= 1e-3, rel_l2(e_gpu[:, 0], e_cpu[:, 0])
    # reflected energy 1 + dE/E0 at the end
    assert abs((1 + e_gpu[-1, 1]) - (1 + e_cpu[-1, 1])) <= 0.05 * (1 + e_cpu[-1, 1])


def test_sweep_on_the_gpu_picks_a_stable_pair():
    print("abc sweep", flush=True)
    prob = pp.make_problem(12.0, 0.25, steps=300, ls="gpu")
    assert prob.info["sweep_stable"] and prob.info["sweep_maxG"] <= 1.0
    assert prob.info["maxG"] <= 1.0 + 1e-6
