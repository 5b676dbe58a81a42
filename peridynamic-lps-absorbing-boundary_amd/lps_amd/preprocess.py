"""Input generator: the LPS-ABC preprocessor restated with the domain as a parameter.

This is the caller side of the solver's time loop. It builds everything the solver reads: the
circular lattice domain, the absorbing-boundary (ABC) nodes and their interpolation clouds, the
discrete LPS dispersion relation, the least-squares ABC operators G_u, G_v and G_theta, the
(dt, delta_kappa) stability sweep and the Gaussian-pulse initial conditions.

It follows preprocessor/main.cpp of the reference:
  material constants           preprocessor/main.cpp:107-129, nu override :810-817
  beta                         preprocessor/main.cpp:134-141
  template grid, m_ref         preprocessor/main.cpp:148-182
  dispersion (2x2 eigen)       preprocessor/main.cpp:199-312
  mode grid                    preprocessor/main.cpp:319-356
  LS solve (dgelsd, 1e-14)     preprocessor/main.cpp:368-402
  G_u / G_v per ABC node       preprocessor/main.cpp:415-637
  G_theta per ABC node         preprocessor/main.cpp:657-785
  domain, ABC nodes, clouds    preprocessor/main.cpp:850-902
  (dt, dk) sweep, N_STEPS      preprocessor/main.cpp:1047-1179
  Gaussian IC                  preprocessor/main.cpp:1280-1289

The reference hard-codes R = 125 mm, nu = 0.25 and one centred pulse; here R, nu, the pulses, the
step count and (optionally) dt are parameters, and everything else keeps the reference's rules.
The least-squares solves use LAPACK dgelsd through scipy (the same routine and rcond the reference
links). The operators are inputs of the solver, and no test of the reference pins their values.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

# Constants of preprocessor/main.cpp:59-92.
MAT_E = 190.0 / 1000.0
MAT_NU = 0.25
MAT_RHO = 8000.0 / 1e18
MAT_G0 = 22170.0 * 1e-9
DX_DEFAULT = 0.5
DELTA_DEFAULT = 4.0 * DX_DEFAULT
ABCS_STEP = 2
MODE_DELTA_PHI = 0.0392
MODE_N_SUB = 5
SVD_RCOND = 1e-14
DK_FACTORS = (0.3, 0.5, 0.7, 0.8, 0.9, 1.0, 1.1, 1.2, 1.5, 2.0)
DT_VALUES = (5.0e-9, 8.0e-9, 10.0e-9, 12.0e-9, 14.0e-9, 15.0e-9, 16.0e-9, 17.0e-9, 17.5e-9,
             18.0e-9, 20.0e-9, 25.0e-9, 30.0e-9, 35.0e-9)
T_FINAL = 6000.0 * 35.0e-9


@dataclass
class Material:
    E: float
    nu: float
    rho: float
    G0: float
    delta: float
    dx: float
    dy: float
    lam: float = 0.0
    mu: float = 0.0
    kappa: float = 0.0
    s0: float = 0.0
    c_P: float = 0.0
    c_S: float = 0.0
    m_ref: float = 0.0
    alpha_lps: float = 0.0
    lambda_pd: float = 0.0


def material(nu: float = MAT_NU, E: float = MAT_E, rho: float = MAT_RHO, G0: float = MAT_G0,
             delta: float = DELTA_DEFAULT, dx: float = DX_DEFAULT, dy: float = DX_DEFAULT) -> Material:
    """Material and LPS constants (preprocessor/main.cpp:107-129, 810-817, 831-835)."""
    mc = Material(E=E, nu=nu, rho=rho, G0=G0, delta=delta, dx=dx, dy=dy)
    mc.lam = mc.E * mc.nu / ((1.0 + mc.nu) * (1.0 - 2.0 * mc.nu))
    mc.mu = mc.E / (2.0 * (1.0 + mc.nu))
    mc.kappa = mc.lam + mc.mu
    mc.s0 = math.sqrt(5.0 * math.pi * mc.G0 / (12.0 * mc.E * delta))
    mc.c_P = math.sqrt((mc.lam + 2.0 * mc.mu) / mc.rho)
    mc.c_S = math.sqrt(mc.mu / mc.rho)
    tg = template_grid(delta, dx, dy)
    mc.m_ref = template_m_ref(tg, delta, dx, dy)
    mc.alpha_lps = 8.0 * mc.mu / mc.m_ref
    mc.lambda_pd = 2.0 * (mc.lam - mc.mu) / mc.m_ref
    return mc


def beta(xi: float, delta: float, dx: float) -> float:
    """Partial-volume factor (preprocessor/main.cpp:134-141, src/main.cpp:100-107)."""
    if xi <= delta - dx / 2.0:
        return 1.0
    if xi <= delta + dx / 2.0:
        return (delta + dx / 2.0 - xi) / dx
    return 0.0


def _accumulated_range(lo: float, hi: float, step: float) -> List[float]:
    """The values of `for (double v = lo; v <= hi; v += step)` with the same rounding."""
    out = []
    v = lo
    while v <= hi:
        out.append(v)
        v += step
    return out


def template_grid(delta: float, dx: float, dy: float) -> np.ndarray:
    """Bond offsets of an interior node, |xi| in (1e-12, delta+dx/2+1e-10]
    (preprocessor/main.cpp:148-169). Returns (68, 2) for delta = 4 dx."""
    rmax = delta + dx / 2.0
    loop_max = delta + dx
    grid = []
    for xx in _accumulated_range(-loop_max, loop_max + 1e-10, dx):
        for yy in _accumulated_range(-loop_max, loop_max + 1e-10, dy):
            r = math.sqrt(xx * xx + yy * yy)
            if 1e-12 < r <= rmax + 1e-10:
                grid.append((xx, yy))
    return np.asarray(grid, dtype=np.float64)


def template_m_ref(tg: np.ndarray, delta: float, dx: float, dy: float) -> float:
    """Weighted volume of a full horizon (preprocessor/main.cpp:174-182)."""
    m = 0.0
    for xx, yy in tg:
        r = math.sqrt(xx * xx + yy * yy)
        m += r * r * beta(r, delta, dx) * dx * dy
    return m


# ---------------------------------------------------------------------------------------------
# Dispersion relation and mode grid
# ---------------------------------------------------------------------------------------------
@dataclass
class Modes:
    a1: np.ndarray
    a2: np.ndarray
    omP: np.ndarray
    omS: np.ndarray
    gP: np.ndarray  # (M, 2)
    gS: np.ndarray  # (M, 2)
    S_r: np.ndarray  # (M, 2)
    S_i: np.ndarray  # (M, 2)

    def __len__(self):
        return len(self.a1)

    def sorted_by_kappa(self) -> "Modes":
        """Ascending |kappa|^2 (preprocessor/main.cpp:982-986); ties keep the grid order."""
        k2 = self.a1 * self.a1 + self.a2 * self.a2
        o = np.argsort(k2, kind="stable")
        return Modes(*(getattr(self, f)[o] for f in
                       ("a1", "a2", "omP", "omS", "gP", "gS", "S_r", "S_i")))


def solve_dispersion(a1: float, a2: float, mc: Material, tg: np.ndarray):
    """LPS dispersion matrix M^LPS and its 2x2 eigen-decomposition (preprocessor/main.cpp:199-312).

    Returns (omega_P, omega_S, gamma_P, gamma_S, S_real, S_imag)."""
    dV = mc.dx * mc.dy
    S = [0j, 0j]
    N = [[0j, 0j], [0j, 0j]]
    for xx, yy in tg:
        xi2 = xx * xx + yy * yy
        xn = math.sqrt(xi2)
        b = beta(xn, mc.delta, mc.dx)
        ph = a1 * xx + a2 * yy
        ef = complex(math.cos(ph) - 1.0, math.sin(ph))
        S[0] += xx * ef * b * dV
        S[1] += yy * ef * b * dV
        comp = (xx, yy)
        for p in range(2):
            for k in range(2):
                N[p][k] += (comp[p] * comp[k] / xi2) * ef * b * dV
    cS = 2.0 * mc.lambda_pd / mc.m_ref
    cN = 2.0 * mc.alpha_lps
    M = [[0.0, 0.0], [0.0, 0.0]]
    for p in range(2):
        for k in range(2):
            M[p][k] = (cS * S[p] * S[k] + cN * N[p][k]).real
    tr = M[0][0] + M[1][1]
    det = M[0][0] * M[1][1] - M[0][1] * M[1][0]
    disc = tr * tr - 4.0 * det
    if disc < 0.0:
        disc = 0.0
    eig1 = (tr + math.sqrt(disc)) / 2.0
    eig2 = (tr - math.sqrt(disc)) / 2.0
    omP2 = -eig2 / mc.rho
    omS2 = -eig1 / mc.rho
    omP = math.sqrt(omP2) if omP2 > 0.0 else 0.0
    omS = math.sqrt(omS2) if omS2 > 0.0 else 0.0

    def eigvec(eig):
        a00, a01 = M[0][0] - eig, M[0][1]
        a10, a11 = M[1][0], M[1][1] - eig
        if abs(a00) + abs(a01) >= abs(a10) + abs(a11):
            g = [-a01, a00]
        else:
            g = [a11, -a10]
        nrm = math.sqrt(g[0] * g[0] + g[1] * g[1])
        if nrm > 1e-30:
            g = [g[0] / nrm, g[1] / nrm]
        return g

    mnorm = abs(M[0][0]) + abs(M[0][1]) + abs(M[1][0]) + abs(M[1][1])
    if mnorm < 1e-30:
        gP, gS = [1.0, 0.0], [0.0, 1.0]
    else:
        gP, gS = eigvec(eig2), eigvec(eig1)
    return omP, omS, gP, gS, [S[0].real, S[1].real], [S[0].imag, S[1].imag]


def compute_all_modes(mc: Material, tg: np.ndarray, delta_kappa: float) -> Modes:
    """Mode grid theta in dphi*[-1,1], |kappa| in dk*[0,1], duplicates removed
    (preprocessor/main.cpp:319-356)."""
    rows = []
    for it in range(-MODE_N_SUB, MODE_N_SUB + 1):
        th = MODE_DELTA_PHI * float(it) / MODE_N_SUB
        for ik in range(0, MODE_N_SUB + 1):
            kv = delta_kappa * float(ik) / MODE_N_SUB
            a1 = kv * math.cos(th)
            a2 = kv * math.sin(th)
            if any(math.sqrt((r[0] - a1) ** 2 + (r[1] - a2) ** 2) < 1e-12 for r in rows):
                continue
            rows.append((a1, a2) + solve_dispersion(a1, a2, mc, tg))
    return Modes(
        a1=np.array([r[0] for r in rows]), a2=np.array([r[1] for r in rows]),
        omP=np.array([r[2] for r in rows]), omS=np.array([r[3] for r in rows]),
        gP=np.array([r[4] for r in rows]), gS=np.array([r[5] for r in rows]),
        S_r=np.array([r[6] for r in rows]), S_i=np.array([r[7] for r in rows]))


# ---------------------------------------------------------------------------------------------
# Least squares (dgelsd) and the per-node ABC operators
# ---------------------------------------------------------------------------------------------
_LWORK_CACHE = {}


def _dgelsd(AT: np.ndarray, B: np.ndarray, rcond: float = SVD_RCOND) -> np.ndarray:
    """Minimum-norm least squares min ||AT x - B|| with singular values below rcond*s_max dropped
    (LAPACK dgelsd, as preprocessor/main.cpp:368-402). Returns x with AT.shape[1] rows."""
    from scipy.linalg import lapack
    m, n = AT.shape
    nrhs = B.shape[1]
    ldb = max(m, n)
    key = (m, n, nrhs)
    if key not in _LWORK_CACHE:
        lw = lapack.dgelsd_lwork(m, n, nrhs, rcond)
        _LWORK_CACHE[key] = (int(lw[0]), int(lw[1]))
    lwork, liwork = _LWORK_CACHE[key]
    Bp = np.zeros((ldb, nrhs), dtype=np.float64, order="F")
    Bp[:m] = B
    x, _s, _rank, info = lapack.dgelsd(np.asfortranarray(AT), Bp, lwork, liwork, cond=rcond)
    if info != 0:
        # the reference warns and keeps what dgelsd left in B (preprocessor/main.cpp:605-611,
        # 769-772); the stability sweep then judges the resulting operator by its norm
        import warnings
        warnings.warn(f"dgelsd did not converge (info={info})")
        if _ROBUST_LS[0]:
            # same minimum-norm solution from the QR-iteration SVD (dgelss)
            Bq = np.zeros((ldb, nrhs), dtype=np.float64, order="F")
            Bq[:m] = B
            out = lapack.dgelss(np.asfortranarray(AT), Bq, cond=rcond)
            if out[-1] == 0:
                return out[1][:n]
    return x[:n]


# make_problem(robust_ls=True): a dgelsd that does not converge is solved again with dgelss
# instead of keeping dgelsd's output (a deviation from the reference, for generated workloads).
_ROBUST_LS = [False]


def _local_frame(cloud_xy: np.ndarray, ax: float, ay: float, tau: float):
    cosT, sinT = math.cos(tau), math.sin(tau)
    dxg = cloud_xy[:, 0] - ax
    dyg = cloud_xy[:, 1] - ay
    xl = cosT * dxg + sinT * dyg
    yl = -sinT * dxg + cosT * dyg
    return cosT, sinT, xl, yl


def abc_node_uv(ax: float, ay: float, tau: float, cloud_xy: np.ndarray, modes: Modes,
                max_cloud: int, dt: float) -> Tuple[np.ndarray, np.ndarray]:
    """G_u and G_v rows of one ABC node (preprocessor/main.cpp:415-637).

    Returns two (2, 4*max_cloud) arrays; columns [4k..4k+3] weight (U_x, U_y, Unm1_x, Unm1_y) of
    cloud node k and row 0/1 give the x/y component."""
    nC = cloud_xy.shape[0]
    ATu, tu, ATv, tvv = abc_node_uv_system(ax, ay, tau, cloud_xy, modes, dt)
    xu = _dgelsd(ATu, tu)
    xv = _dgelsd(ATv, tvv)
    Gu = np.zeros((2, 4 * max_cloud))
    Gv = np.zeros((2, 4 * max_cloud))
    for tl in range(ABCS_STEP):
        src = tl * 2 * nC + 2 * np.arange(nC)
        dst = 4 * np.arange(nC) + 2 * tl
        for r in range(2):
            Gu[r, dst] = xu[src, r]
            Gu[r, dst + 1] = xu[src + 1, r]
            Gv[r, dst] = xv[src, r]
            Gv[r, dst + 1] = xv[src + 1, r]
    return Gu, Gv


def abc_node_uv_system(ax: float, ay: float, tau: float, cloud_xy: np.ndarray, modes: Modes,
                       dt: float):
    """The least-squares systems of abc_node_uv: (ATu, tu) for G_u and (ATv, tv) for G_v, rows 4
    per mode, columns (time level, cloud node, component) = tl * 2 nC + 2 k + c."""
    nC = cloud_xy.shape[0]
    M = len(modes)
    cosT, sinT, xl, yl = _local_frame(cloud_xy, ax, ay, tau)
    ATu = np.zeros((4 * M, 2 * nC * ABCS_STEP))
    ATv = np.zeros_like(ATu)
    phase = modes.a1[:, None] * xl[None, :] + modes.a2[:, None] * yl[None, :]
    gP, gS, omP, omS = modes.gP, modes.gS, modes.omP[:, None], modes.omS[:, None]
    for tl in range(ABCS_STEP):
        tv = -tl * dt
        pP = phase - omP * tv
        pS = phase - omS * tv
        cP, sP, cS, sS = np.cos(pP), np.sin(pP), np.cos(pS), np.sin(pS)
        disp = [(gP[:, 0:1] * cP, gP[:, 1:2] * cP), (gS[:, 0:1] * cS, gS[:, 1:2] * cS),
                (gP[:, 0:1] * sP, gP[:, 1:2] * sP), (gS[:, 0:1] * sS, gS[:, 1:2] * sS)]
        vel = [(gP[:, 0:1] * omP * sP, gP[:, 1:2] * omP * sP),
               (gS[:, 0:1] * omS * sS, gS[:, 1:2] * omS * sS),
               (-gP[:, 0:1] * omP * cP, -gP[:, 1:2] * omP * cP),
               (-gS[:, 0:1] * omS * cS, -gS[:, 1:2] * omS * cS)]
        cols = tl * 2 * nC + 2 * np.arange(nC)
        for b in range(4):
            lx, ly = disp[b]
            ATu[b::4, cols] = cosT * lx - sinT * ly
            ATu[b::4, cols + 1] = sinT * lx + cosT * ly
            lx, ly = vel[b]
            ATv[b::4, cols] = cosT * lx - sinT * ly
            ATv[b::4, cols + 1] = sinT * lx + cosT * ly
    om = np.stack([modes.omP, modes.omS])  # (2, M)
    cdt, sdt = np.cos(-om * dt), np.sin(-om * dt)
    tu = np.zeros((4 * M, 2))
    tvv = np.zeros((4 * M, 2))
    g = (gP, gS)
    for b in range(4):
        w = b % 2  # 0: P, 1: S
        gg = g[w]
        if b < 2:
            lx, ly = gg[:, 0] * cdt[w], gg[:, 1] * cdt[w]
            vx, vy = gg[:, 0] * om[w] * sdt[w], gg[:, 1] * om[w] * sdt[w]
        else:
            lx, ly = gg[:, 0] * sdt[w], gg[:, 1] * sdt[w]
            vx, vy = -gg[:, 0] * om[w] * cdt[w], -gg[:, 1] * om[w] * cdt[w]
        tu[b::4, 0] = cosT * lx - sinT * ly
        tu[b::4, 1] = sinT * lx + cosT * ly
        tvv[b::4, 0] = cosT * vx - sinT * vy
        tvv[b::4, 1] = sinT * vx + cosT * vy
    return ATu, tu, ATv, tvv


def abc_node_theta(ax: float, ay: float, tau: float, cloud_xy: np.ndarray, modes: Modes,
                   mc: Material, max_cloud: int, dt: float) -> np.ndarray:
    """G_theta row of one ABC node (preprocessor/main.cpp:657-785).

    Returns (2*max_cloud,); columns [2k, 2k+1] weight (ThetaN, ThetaNm1) of cloud node k."""
    nC = cloud_xy.shape[0]
    M = len(modes)
    cosT, sinT, xl, yl = _local_frame(cloud_xy, ax, ay, tau)
    coeff = 2.0 / mc.m_ref
    gP, gS = modes.gP, modes.gS
    thP_r = coeff * (gP[:, 0] * modes.S_r[:, 0] + gP[:, 1] * modes.S_r[:, 1])
    thP_i = coeff * (gP[:, 0] * modes.S_i[:, 0] + gP[:, 1] * modes.S_i[:, 1])
    thS_r = coeff * (gS[:, 0] * modes.S_r[:, 0] + gS[:, 1] * modes.S_r[:, 1])
    thS_i = coeff * (gS[:, 0] * modes.S_i[:, 0] + gS[:, 1] * modes.S_i[:, 1])
    AT = np.zeros((4 * M, nC * ABCS_STEP))
    phase = modes.a1[:, None] * xl[None, :] + modes.a2[:, None] * yl[None, :]
    for tl in range(ABCS_STEP):
        tv = -tl * dt
        pP = phase - modes.omP[:, None] * tv
        pS = phase - modes.omS[:, None] * tv
        cP, sP, cS, sS = np.cos(pP), np.sin(pP), np.cos(pS), np.sin(pS)
        cols = tl * nC + np.arange(nC)
        AT[0::4, cols] = thP_r[:, None] * cP - thP_i[:, None] * sP
        AT[1::4, cols] = thS_r[:, None] * cS - thS_i[:, None] * sS
        AT[2::4, cols] = thP_r[:, None] * sP + thP_i[:, None] * cP
        AT[3::4, cols] = thS_r[:, None] * sS + thS_i[:, None] * cS
    cPd, sPd = np.cos(-modes.omP * dt), np.sin(-modes.omP * dt)
    cSd, sSd = np.cos(-modes.omS * dt), np.sin(-modes.omS * dt)
    t = np.zeros((4 * M, 1))
    t[0::4, 0] = thP_r * cPd - thP_i * sPd
    t[1::4, 0] = thS_r * cSd - thS_i * sSd
    t[2::4, 0] = thP_r * sPd + thP_i * cPd
    t[3::4, 0] = thS_r * sSd + thS_i * cSd
    x = _dgelsd(AT, t)[:, 0]
    G = np.zeros(2 * max_cloud)
    for tl in range(ABCS_STEP):
        G[2 * np.arange(nC) + tl] = x[tl * nC + np.arange(nC)]
    return G


def g_norm(Gu: np.ndarray, Gv: np.ndarray) -> float:
    """max(||G_u||_F, ||G_v||_F) of one node (preprocessor/main.cpp:946-958)."""
    return max(math.sqrt(float(np.sum(Gu * Gu))), math.sqrt(float(np.sum(Gv * Gv))))


# ---------------------------------------------------------------------------------------------
# Domain, ABC nodes, clouds, initial conditions
# ---------------------------------------------------------------------------------------------
def lattice_domain(R: float, dx: float = DX_DEFAULT, dy: float = DX_DEFAULT) -> np.ndarray:
    """Lattice nodes with sqrt(x^2+y^2) <= R+1e-10, x-major order (preprocessor/main.cpp:850-860)."""
    xs = np.array(_accumulated_range(-R, R + 1e-10, dx))
    ys = np.array(_accumulated_range(-R, R + 1e-10, dy))
    X = np.repeat(xs, len(ys))
    Y = np.tile(ys, len(xs))
    keep = np.sqrt(X * X + Y * Y) <= R + 1e-10
    return np.stack([X[keep], Y[keep]], axis=1)


def abc_nodes(nodes: np.ndarray, R: float, delta: float):
    """ABC nodes r > R - delta with their outward angle (preprocessor/main.cpp:865-875)."""
    r = np.sqrt(nodes[:, 0] ** 2 + nodes[:, 1] ** 2)
    idx = np.nonzero(r > R - delta)[0]
    tau = np.arctan2(nodes[idx, 1], nodes[idx, 0])
    return idx.astype(np.int64), tau


def _lattice_clouds(nodes: np.ndarray, abc_idx: np.ndarray, radius: float):
    """abc_clouds on a lattice of spacing h = the smallest nonzero |x| step: candidate offsets
    (i h, j h) with the same distance test as the k-d tree path, looked up by integer
    coordinates. None when the nodes are not lattice points of one spacing."""
    if len(abc_idx) == 0 or nodes.shape[0] < 2:
        return None
    xs = np.unique(np.abs(nodes[:, 0]))
    h = float(xs[xs > 0].min()) if np.any(xs > 0) else 0.0
    if h <= 0.0:
        return None
    ix = np.rint(nodes[:, 0] / h).astype(np.int64)
    iy = np.rint(nodes[:, 1] / h).astype(np.int64)
    if not (np.array_equal(ix * h, nodes[:, 0]) and np.array_equal(iy * h, nodes[:, 1])):
        return None
    off = int(max(np.abs(ix).max(), np.abs(iy).max())) + 1
    width = 2 * off + 1
    keys = (ix + off) * width + (iy + off)
    order = np.argsort(keys, kind="stable")
    skeys = keys[order]
    m = int(radius / h) + 1
    oi, oj = np.meshgrid(np.arange(-m, m + 1), np.arange(-m, m + 1), indexing="ij")
    oi, oj = oi.ravel(), oj.ravel()
    a = nodes[abc_idx]
    cx = ix[abc_idx][:, None] + oi[None, :]
    cy = iy[abc_idx][:, None] + oj[None, :]
    k = (cx + off) * width + (cy + off)
    pos = np.clip(np.searchsorted(skeys, k), 0, len(skeys) - 1)
    hit = (skeys[pos] == k) & (np.abs(cx) < off) & (np.abs(cy) < off)
    cand = np.where(hit, order[pos], -1)
    # the k-d tree path's test, on the candidates' own coordinates
    d = np.sqrt((nodes[np.maximum(cand, 0), 0] - a[:, 0:1]) ** 2 +
                (nodes[np.maximum(cand, 0), 1] - a[:, 1:2]) ** 2)
    keep = hit & (d <= radius + 1e-10)
    big = np.where(keep, cand, np.iinfo(np.int64).max)
    big.sort(axis=1)
    counts = keep.sum(axis=1)
    return [big[r, :counts[r]] for r in range(len(abc_idx))]


def abc_clouds(nodes: np.ndarray, abc_idx: np.ndarray, radius: float) -> List[np.ndarray]:
    """Nodes within radius+1e-10 of each ABC node, ascending index (preprocessor/main.cpp:884-902).
    Lattice inputs (coordinates exact multiples of one spacing) take the lattice offsets within
    the radius, all rows at once; other inputs a k-d tree per row."""
    fast = _lattice_clouds(nodes, abc_idx, radius)
    if fast is not None:
        return fast
    from scipy.spatial import cKDTree
    tree = cKDTree(nodes)
    out = []
    for a in abc_idx:
        cand = np.asarray(tree.query_ball_point(nodes[a], radius + 1e-6), dtype=np.int64)
        d = np.sqrt((nodes[cand, 0] - nodes[a, 0]) ** 2 + (nodes[cand, 1] - nodes[a, 1]) ** 2)
        out.append(np.sort(cand[d <= radius + 1e-10]))
    return out


@dataclass
class Pulse:
    """u = amp * dir * exp(-|x-c|^2 / (r^2/16)) for |x-c| < r. The reference pulse
    (preprocessor/main.cpp:1280-1289) is Pulse(0, 0, 40, 1, (1, 1)).

    polarization "P": u = amp * (x - c) / a * exp(1/2 - |x-c|^2 / (2 a^2)), a = r / (4 sqrt 2): the
    gradient of the Gaussian (curl-free, a pure dilatational pulse), |u| peaking at amp at
    |x - c| = a. "S": the same rotated by 90 degrees (divergence-free, a pure shear pulse). Both are
    smooth at the centre."""
    cx: float
    cy: float
    radius: float
    amp: float = 1.0
    direction: Tuple[float, float] = (1.0, 1.0)
    polarization: Optional[str] = None  # None: fixed direction; "P": radial; "S": tangential


def gaussian_ic(nodes: np.ndarray, pulses: Sequence[Pulse]) -> np.ndarray:
    u = np.zeros(2 * nodes.shape[0])
    for p in pulses:
        dxp = nodes[:, 0] - p.cx
        dyp = nodes[:, 1] - p.cy
        d2 = dxp * dxp + dyp * dyp
        inside = np.sqrt(d2) < p.radius
        val = np.exp(-d2[inside] / (p.radius * p.radius / 16.0))
        if p.polarization is None:
            ex = np.full(val.shape, p.direction[0])
            ey = np.full(val.shape, p.direction[1])
        else:
            a = p.radius / (4.0 * math.sqrt(2.0))
            k = math.exp(0.5) / a
            ex, ey = k * dxp[inside], k * dyp[inside]
            if p.polarization == "S":
                ex, ey = -ey, ex
        u[0::2][inside] += p.amp * ex * val
        u[1::2][inside] += p.amp * ey * val
    return u


# ---------------------------------------------------------------------------------------------
# The problem (everything the solver reads) and its generation
# ---------------------------------------------------------------------------------------------
@dataclass
class Problem:
    nodes: np.ndarray                 # (N, 2)
    steps: int
    data2: np.ndarray                 # 15 doubles: E nu rho wf t0 t1 dt delta dx dy s0 m kappa alpha lpd
    bctypes: np.ndarray               # (2N,) int
    bcvalues: np.ndarray              # (2N,)
    ic: np.ndarray                    # (2N,) u0 interleaved
    icd: np.ndarray                   # (2N,) v0 interleaved
    abc: np.ndarray                   # (C,) ABC node ids
    clouds: np.ndarray                # (C, M) cloud ids, -1 padded
    Gu: np.ndarray                    # (2C, 4M)
    Gv: np.ndarray                    # (2C, 4M)
    Gth: np.ndarray                   # (C, 2M)
    prenotch: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    vinodes: np.ndarray = field(default_factory=lambda: np.zeros((0,), dtype=np.int64))
    info: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.nodes.shape[0]

    @property
    def C(self) -> int:
        return self.abc.shape[0]

    @property
    def max_cloud(self) -> int:
        return self.clouds.shape[1] if self.clouds.ndim == 2 else 0

    @property
    def dt(self) -> float:
        return float(self.data2[6])


def _pool_map(fn, items, workers, shared=None):
    """[fn(item) for item in items], on `workers` processes when workers > 1. fn reads the map's
    common data (modes, material, sample rows) from _W, installed from `shared` before each chunk
    of items. The workers come from a 'forkserver': they start from a fresh interpreter and are
    never a copy of this process, so the map is safe after this process has loaded the HIP runtime
    (a forked copy of a process holding GPU state must not run). The pool is kept for later maps."""
    if workers <= 1 or len(items) <= 1:
        _install(shared)
        _limit_threads()
        try:
            return [fn(it) for it in items]
        finally:
            _restore_threads()
    pool = _get_pool(workers)
    cs = max(1, len(items) // (workers * 8))
    chunks = [(fn, shared, items[i:i + cs]) for i in range(0, len(items), cs)]
    return [r for part in pool.map(_run_chunk, chunks, chunksize=1) for r in part]


_POOL = {}


def _get_pool(workers):
    if _POOL.get("n") != workers:
        _close_pool()
        import multiprocessing as mp
        import sys
        import types
        ctx = mp.get_context("forkserver")
        ctx.set_forkserver_preload([__name__])
        # Workers start with the caller's main module hidden: they run only this module's
        # functions, and a new interpreter would otherwise re-run a caller's script that has no
        # `if __name__ == "__main__"` guard.
        main = sys.modules["__main__"]
        sys.modules["__main__"] = types.ModuleType("__main__")
        try:
            _POOL.update(n=workers, pool=ctx.Pool(workers, initializer=_limit_threads))
        finally:
            sys.modules["__main__"] = main
    return _POOL["pool"]


def _close_pool():
    if _POOL.get("pool") is not None:
        _POOL["pool"].terminate()
        _POOL["pool"].join()
    _POOL.clear()


import atexit  # noqa: E402
atexit.register(_close_pool)


def _run_chunk(args):
    fn, shared, items = args
    _install(shared)
    return [fn(it) for it in items]


def _install(shared):
    _W.clear()
    if shared:
        _W.update(shared)
    _ROBUST_LS[0] = bool(_W.get("robust_ls", False))


_THREAD_LIMIT = []


def _limit_threads():
    """BLAS/LAPACK on one thread: the per-row solves are small (up to 196 columns) and LAPACK's
    thread pool makes them about 200 times slower, not faster. The limit reaches only libraries
    already loaded, so scipy's LAPACK is loaded first."""
    try:
        import scipy.linalg.lapack  # noqa: F401 - load its BLAS before limiting
        from threadpoolctl import threadpool_limits
        _THREAD_LIMIT.append(threadpool_limits(1))
    except Exception:  # noqa: BLE001 - thread limiting is an optimisation only
        pass


def _restore_threads():
    while _THREAD_LIMIT:
        _THREAD_LIMIT.pop().restore_original_limits()


# The pool's common data, in every worker (and in this process for a serial map).
_W = {}


def _row_items(nodes, abc_idx, tau, clouds, rows):
    """Per-row inputs of the operator solves: (x, y, outward angle, cloud coordinates)."""
    return [(float(nodes[abc_idx[i], 0]), float(nodes[abc_idx[i], 1]), float(tau[i]),
             np.ascontiguousarray(nodes[clouds[i]])) for i in rows]


def _w_uv(item):
    x, y, tau, cxy = item
    return abc_node_uv(x, y, tau, cxy, _W["modes"], _W["M"], _W["dt"])


def _w_theta(item):
    x, y, tau, cxy = item
    return abc_node_theta(x, y, tau, cxy, _W["modes"], _W["mc"], _W["M"], _W["dt"])


def _w_sweep(args):
    idk, dt_try = args
    modes = _W["sweep_modes"][idk]
    max_G = 0.0
    for x, y, tau, cxy in _W["sample"]:
        Gu, Gv = abc_node_uv(x, y, tau, cxy, modes, _W["M"], dt_try)
        max_G = max(max_G, g_norm(Gu, Gv))
        if max_G > 1.5:
            break
    return max_G


def lattice_symmetry(nodes: np.ndarray, abc_idx: np.ndarray, clouds: Sequence[np.ndarray],
                     dx: float, dy: float):
    """The 8 symmetries of a square lattice about the origin (x -> -x, y -> -y, x <-> y) applied to
    the ABC rows of an origin-centred domain: every row is the image T c of a canonical row c in the
    octant 0 <= y <= x, under a signed permutation T. Returns (canon, T, pos): canon[a] the row
    index of row a's canonical row (canon[c] == c for canonical rows), T[a] (2 x 2, canonical ->
    row a), pos[a, k] the position in row a's cloud of the image of cloud node k of its canonical
    row (-1 past the cloud). None when the nodes are not lattice points (x / dx not an exact
    integer) or a row's cloud is not the image of its canonical row's cloud.

    Why it holds: a row's least-squares systems (preprocessor/main.cpp:415-785) are built in the
    frame of its outward normal tau = atan2(y, x), over a mode grid symmetric in the incidence
    angle (theta on -dphi .. dphi). Rotating or mirroring the row and its cloud by T therefore maps
    its systems onto the canonical row's, with the cloud's (x, y) components transformed by T. The
    minimum-norm solution of the mapped system is the mapped solution: each 2 x 2 block of G_u
    and G_v becomes T G T' and each G_theta weight stays."""
    ix = np.rint(nodes[:, 0] / dx).astype(np.int64)
    iy = np.rint(nodes[:, 1] / dy).astype(np.int64)
    if not (np.array_equal(ix * dx, nodes[:, 0]) and np.array_equal(iy * dy, nodes[:, 1])):
        return None
    C = len(abc_idx)
    if C == 0:
        return None
    off = int(max(np.abs(ix).max(), np.abs(iy).max())) + 1
    width = 2 * off + 1
    keys = (ix + off) * width + (iy + off)
    order = np.argsort(keys, kind="stable")
    skeys = keys[order]

    def lookup(kx, ky):
        k = (kx + off) * width + (ky + off)
        p = np.clip(np.searchsorted(skeys, k), 0, len(skeys) - 1)
        ok = (skeys[p] == k) & (np.abs(kx) < off) & (np.abs(ky) < off)
        return np.where(ok, order[p], -1)

    ia, ja = ix[abc_idx], iy[abc_idx]
    c1 = np.maximum(np.abs(ia), np.abs(ja))
    c2 = np.minimum(np.abs(ia), np.abs(ja))
    row_of = -np.ones(nodes.shape[0], dtype=np.int64)
    row_of[abc_idx] = np.arange(C)
    cnode = lookup(c1, c2)
    if np.any(cnode < 0):
        return None
    canon = row_of[cnode]
    if np.any(canon < 0):
        return None
    s1 = np.where(ia < 0, -1, 1)
    s2 = np.where(ja < 0, -1, 1)
    swap = np.abs(ia) < np.abs(ja)
    T = np.zeros((C, 2, 2), dtype=np.int64)
    T[:, 0, 0] = np.where(swap, 0, s1)
    T[:, 0, 1] = np.where(swap, s1, 0)
    T[:, 1, 0] = np.where(swap, s2, 0)
    T[:, 1, 1] = np.where(swap, 0, s2)
    M = max(len(c) for c in clouds)
    lens = np.array([len(c) for c in clouds])
    CL = -np.ones((C, M), dtype=np.int64)
    for a, c in enumerate(clouds):
        CL[a, :len(c)] = c
    if not np.array_equal(lens, lens[canon]):
        return None
    src = CL[canon]                                   # canonical cloud nodes, per row
    valid = src >= 0
    ox = np.where(valid, ix[np.maximum(src, 0)] - c1[:, None], 0)
    oy = np.where(valid, iy[np.maximum(src, 0)] - c2[:, None], 0)
    imx = ia[:, None] + T[:, 0, 0, None] * ox + T[:, 0, 1, None] * oy
    imy = ja[:, None] + T[:, 1, 0, None] * ox + T[:, 1, 1, None] * oy
    img = np.where(valid, lookup(imx, imy), -1)
    if np.any(valid & (img < 0)):
        return None
    # position of each image node in row a's (ascending) cloud
    big = nodes.shape[0] + 1
    flat = np.where(CL >= 0, np.arange(C)[:, None] * big + CL, np.iinfo(np.int64).max)
    fo = np.argsort(flat, axis=None, kind="stable")
    fs = flat.ravel()[fo]
    q = np.arange(C)[:, None] * big + np.maximum(img, 0)
    p = np.clip(np.searchsorted(fs, q), 0, fs.size - 1)
    found = fs[p] == q
    if np.any(valid & ~found):
        return None
    pos = np.where(valid, fo[p] % M, -1)
    return canon, T, pos


def symmetric_operators(Gu_c: np.ndarray, Gv_c: np.ndarray, Gth_c: np.ndarray,
                        canon_rows: np.ndarray, canon: np.ndarray, T: np.ndarray, pos: np.ndarray):
    """G_u, G_v, G_theta of every row from those of the canonical rows (lattice_symmetry):
    Gu_c, Gv_c (2 Cc x 4M), Gth_c (Cc x 2M) in the order of canon_rows (the canonical rows'
    indices). Block (t, k) of row a is T[a] G_c(t, k) T[a]' at cloud position pos[a, k]."""
    C, M = pos.shape
    slot = -np.ones(C, dtype=np.int64)
    slot[canon_rows] = np.arange(len(canon_rows))
    ci = slot[canon]
    Tf = T.astype(np.float64)
    a_idx, k_idx = np.nonzero(pos >= 0)
    dst = pos[a_idx, k_idx]

    def blocks(G):
        B = G.reshape(-1, 2, M, 2, 2)[ci]             # (C, out, k, t, in)
        B = np.einsum("aij,ajktl,aml->aiktm", Tf, B, Tf)
        out = np.zeros_like(B)
        out[a_idx, :, dst] = B[a_idx, :, k_idx]
        return out.reshape(2 * C, 4 * M)

    Gu, Gv = blocks(Gu_c), blocks(Gv_c)
    Gth = np.zeros((C, 2 * M))
    if Gth_c is not None and Gth_c.size:
        th = Gth_c.reshape(-1, M, 2)[ci]
        o = np.zeros_like(th)
        o[a_idx, dst] = th[a_idx, k_idx]
        Gth = o.reshape(C, 2 * M)
    return Gu, Gv, Gth


def default_workers() -> int:
    n = os.cpu_count() or 1
    return max(1, min(n, 16))


def make_problem(R: float, nu: float = MAT_NU, pulses: Optional[Sequence[Pulse]] = None,
                 steps: Optional[int] = None, dt: Optional[float] = None,
                 dx: float = DX_DEFAULT, delta: Optional[float] = None,
                 workers: Optional[int] = None, with_abc: bool = True,
                 dk_factor: Optional[float] = None, robust_ls: bool = False,
                 ls: str = "cpu", device: int = 0, symmetric: bool = False) -> Problem:
    """Generate a full solver input set for a circular lattice domain of radius R.

    dt=None runs the reference's (dt, delta_kappa) sweep (preprocessor/main.cpp:1047-1174);
    otherwise dt and dk_factor (delta_kappa = dk_factor / sqrt(c_P), default 1) are taken as given,
    e.g. a sweep result computed once. steps=None takes ceil(210 us / dt)
    (preprocessor/main.cpp:1177-1179). robust_ls=True re-solves a least-squares problem on which
    dgelsd does not converge with dgelss (the reference keeps dgelsd's output). ls="gpu" builds
    the operators (and the sweep's) on the GPU (lps_amd.abc_gpu) instead of with LAPACK.
    symmetric=True (CPU builder) solves only the rows of one octant and maps them onto the others
    (lattice_symmetry), 8 times fewer solves; the operators are least-squares solutions of every
    row's own systems, equal to direct solves up to the systems' conditioning."""
    if ls not in ("cpu", "gpu"):
        raise ValueError("ls must be 'cpu' or 'gpu'")
    _ROBUST_LS[0] = robust_ls
    try:
        return _make_problem(R, nu, pulses, steps, dt, dx, delta, workers, with_abc, dk_factor,
                             ls, device, symmetric)
    finally:
        _ROBUST_LS[0] = False


def _sweep_gpu(nodes, abc_idx, tau, clouds, M, sample, sweep_modes, m_ref, device):
    """max ||G|| over the sample rows for every (delta_kappa, dt) of the sweep, on the GPU."""
    from . import abc_gpu
    sub = np.asarray(sample, dtype=np.int64)
    cl = [clouds[i] for i in sub]
    res = []
    for idk in range(len(DK_FACTORS)):
        for dtv in DT_VALUES:
            Gu, Gv, _, _ = abc_gpu.operators(nodes, abc_idx[sub], tau[sub], cl, M, sweep_modes[idk],
                                             dtv, m_ref, with_theta=False, device=device)
            res.append(max(g_norm(Gu[2 * i:2 * i + 2], Gv[2 * i:2 * i + 2])
                           for i in range(len(sub))))
    return res


def _make_problem(R, nu, pulses, steps, dt, dx, delta, workers, with_abc, dk_factor, ls="cpu",
                  device=0, symmetric=False) -> Problem:
    delta = 4.0 * dx if delta is None else delta
    workers = default_workers() if workers is None else workers
    mc = material(nu=nu, delta=delta, dx=dx, dy=dx)
    tg = template_grid(delta, dx, dx)
    nodes = lattice_domain(R, dx, dx)
    N = nodes.shape[0]
    if pulses is None:
        pulses = [Pulse(0.0, 0.0, 40.0, 1.0, (1.0, 1.0))]
    info = {"R": R, "nu": nu, "dx": dx, "delta": delta}
    if with_abc:
        abc_idx, tau = abc_nodes(nodes, R, delta)
        clouds = abc_clouds(nodes, abc_idx, delta)
        M = max(len(c) for c in clouds)
        dk0 = 1.0 / math.sqrt(mc.c_P)
        robust = _ROBUST_LS[0]
        if dt is None:
            C = len(abc_idx)
            sample = [i for i in range(C) if len(clouds[i]) <= 26]
            sample += list(range(min(20, C)))
            sample += list(range(max(0, C - 20), C))
            sample = sorted(set(sample))
            sweep_modes = [compute_all_modes(mc, tg, dk0 * f).sorted_by_kappa()
                           for f in DK_FACTORS]
            jobs = [(idk, dtv) for idk in range(len(DK_FACTORS)) for dtv in DT_VALUES]
            if ls == "gpu":
                res = _sweep_gpu(nodes, abc_idx, tau, clouds, M, sample, sweep_modes,
                                 mc.m_ref, device)
            else:
                res = _pool_map(_w_sweep, jobs, workers, shared=dict(
                    M=M, sweep_modes=sweep_modes, robust_ls=robust,
                    sample=_row_items(nodes, abc_idx, tau, clouds, sample)))
            best = None
            for (idk, dtv), mg in zip(jobs, res):
                if mg <= 1.0 and (best is None or mg > best[2]):
                    best = (idk, dtv, mg)
            if best is None:
                for (idk, dtv), mg in zip(jobs, res):
                    if best is None or mg < best[2]:
                        best = (idk, dtv, mg)
                info["sweep_stable"] = False
            else:
                info["sweep_stable"] = True
            dk = dk0 * DK_FACTORS[best[0]]
            dt = best[1]
            info.update(sweep_dk_factor=DK_FACTORS[best[0]], sweep_maxG=best[2],
                        sweep_samples=len(sample))
        else:
            dk = dk0 * (1.0 if dk_factor is None else dk_factor)
        modes = compute_all_modes(mc, tg, dk).sorted_by_kappa()
        C = len(abc_idx)
        sym = lattice_symmetry(nodes, abc_idx, clouds, dx, dx) if symmetric else None
        rows = np.arange(C) if sym is None else np.nonzero(sym[0] == np.arange(C))[0]
        if ls == "gpu":
            from . import abc_gpu
            Gu, Gv, Gth, rank = abc_gpu.operators(nodes, abc_idx[rows], tau[rows],
                                                  [clouds[i] for i in rows], M, modes, dt,
                                                  mc.m_ref, device=device)
            info["ls"] = "gpu"
            info["ls_unconverged"] = int(np.sum(rank < 0))
            if sym is not None:
                Gu, Gv, Gth = symmetric_operators(Gu, Gv, Gth, rows, *sym)
                info["ls"] = f"gpu, symmetric ({len(rows)} of {C} rows solved)"
        else:
            items = _row_items(nodes, abc_idx, tau, clouds, rows)
            shared = dict(modes=modes, M=M, dt=dt, mc=mc, robust_ls=robust)
            uv = _pool_map(_w_uv, items, workers, shared)
            th = _pool_map(_w_theta, items, workers, shared)
            n = len(rows)
            Gu = np.zeros((2 * n, 4 * M))
            Gv = np.zeros((2 * n, 4 * M))
            for i, (gu, gv) in enumerate(uv):
                Gu[2 * i:2 * i + 2] = gu
                Gv[2 * i:2 * i + 2] = gv
            Gth = np.stack(th) if n else np.zeros((0, 2 * M))
            if sym is not None:
                Gu, Gv, Gth = symmetric_operators(Gu, Gv, Gth, rows, *sym)
                info["ls"] = f"cpu, symmetric ({n} of {C} rows solved)"
        cl = -np.ones((C, M), dtype=np.int64)
        for i, c in enumerate(clouds):
            cl[i, :len(c)] = c
        info["maxG"] = max((g_norm(Gu[2 * i:2 * i + 2], Gv[2 * i:2 * i + 2]) for i in range(C)),
                           default=0.0)
        _W.clear()
    else:
        if dt is None:
            raise ValueError("dt is required without ABCs")
        abc_idx = np.zeros((0,), dtype=np.int64)
        cl = np.zeros((0, 0), dtype=np.int64)
        Gu = np.zeros((0, 0))
        Gv = np.zeros((0, 0))
        Gth = np.zeros((0, 0))
    if steps is None:
        steps = int(math.ceil(T_FINAL / dt))
    bctypes = -np.ones(2 * N, dtype=np.int64)
    bctypes[2 * abc_idx] = 1
    bctypes[2 * abc_idx + 1] = 1
    ic = gaussian_ic(nodes, pulses)
    data2 = np.array([mc.E, mc.nu, mc.rho, 0.0, 0.0, steps * dt, dt, delta, dx, dx, mc.s0,
                      mc.m_ref, mc.kappa, mc.alpha_lps, mc.lambda_pd])
    info["dt"] = dt
    return Problem(nodes=nodes, steps=steps, data2=data2, bctypes=bctypes,
                   bcvalues=np.zeros(2 * N), ic=ic, icd=np.zeros(2 * N), abc=abc_idx,
                   clouds=cl, Gu=Gu, Gv=Gv, Gth=Gth, info=info)


def reference_test_problem(R: float = 30.0, steps: int = 300, dt: float = 35e-9,
                           nu: float = 0.25) -> Problem:
    """The no-ABC case of tests/generate_test_data.cpp:27-73 (pulse exp(-r^2/100) for r < 15)."""
    mc = material(nu=nu)
    nodes = lattice_domain(R)
    N = nodes.shape[0]
    r2 = nodes[:, 0] ** 2 + nodes[:, 1] ** 2
    ic = np.zeros(2 * N)
    inside = np.sqrt(r2) < 15.0
    val = np.exp(-r2[inside] / 100.0)
    ic[0::2][inside] = val
    ic[1::2][inside] = val
    data2 = np.array([mc.E, nu, mc.rho, 0.0, 0.0, steps * dt, dt, 2.0, 0.5, 0.5, mc.s0])
    return Problem(nodes=nodes, steps=steps, data2=data2, bctypes=-np.ones(2 * N, dtype=np.int64),
                   bcvalues=np.zeros(2 * N), ic=ic, icd=np.zeros(2 * N),
                   abc=np.zeros((0,), dtype=np.int64), clouds=np.zeros((0, 0), dtype=np.int64),
                   Gu=np.zeros((0, 0)), Gv=np.zeros((0, 0)), Gth=np.zeros((0, 0)),
                   info={"R": R, "nu": nu})
