"""Least-squares ABC operators built on the GPU (lps_abc_operators and friends in the C ABI).

The reference's preprocessor solves one small least-squares problem per ABC row and operator
(preprocessor/main.cpp:415-637 for G_u / G_v, :657-785 for G_theta) with LAPACK dgelsd at
rcond 1e-14 (:368-402). lps_amd.preprocess restates that on the CPU. Here the systems are built on
the device, solved as a batch (SVD, minimum-norm solution with the same rcond rule) and scattered
into the .ciNodesU / .ciNodesV / .ciNodesTheta layouts. The wave modes (the 2x2 dispersion
problems, :199-356) and the clouds stay on the CPU: they are small.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import solver

KIND_U, KIND_V, KIND_THETA = 0, 1, 2


def _structs(nodes: np.ndarray, abc_idx: np.ndarray, tau: np.ndarray,
             clouds: Sequence[np.ndarray], max_cloud: int, modes, keep: List[np.ndarray]):
    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data

    C_ = len(abc_idx)
    cl_len = np.array([len(c) for c in clouds], dtype=np.int32)
    cxy = np.zeros((C_, max_cloud, 2))
    for i, c in enumerate(clouds):
        cxy[i, :len(c)] = nodes[c]
    r = solver.LpsAbcRows()
    r.count, r.max_cloud = C_, max_cloud
    r.xy = arr(nodes[abc_idx], np.float64)
    r.tau = arr(tau, np.float64)
    r.cloud_len = arr(cl_len, np.int32)
    r.cloud_xy = arr(cxy, np.float64)
    m = solver.LpsAbcModes()
    m.count = len(modes)
    for name, src in (("a1", modes.a1), ("a2", modes.a2), ("om_p", modes.omP), ("om_s", modes.omS),
                      ("g_p", modes.gP), ("g_s", modes.gS), ("s_r", modes.S_r),
                      ("s_i", modes.S_i)):
        setattr(m, name, arr(src, np.float64))
    return r, m


def _check(rc: int):
    if rc != 0:
        raise solver.LpsError(rc, solver.load_library().lps_global_error().decode())


def operators(nodes: np.ndarray, abc_idx: np.ndarray, tau: np.ndarray,
              clouds: Sequence[np.ndarray], max_cloud: int, modes, dt: float, m_ref: float,
              rcond: float = 1e-14, with_theta: bool = True, device: int = 0
              ) -> Tuple[np.ndarray, np.ndarray, Optional[np.ndarray], np.ndarray]:
    """(G_u, G_v, G_theta or None, rank (C x 3)) for the given rows, modes and dt."""
    L = solver.load_library()
    keep: List[np.ndarray] = []
    r, m = _structs(nodes, abc_idx, tau, clouds, max_cloud, modes, keep)
    Cn = len(abc_idx)
    Gu = np.zeros((2 * Cn, 4 * max_cloud))
    Gv = np.zeros((2 * Cn, 4 * max_cloud))
    Gth = np.zeros((Cn, 2 * max_cloud)) if with_theta else None
    rank = np.zeros((Cn, 3), dtype=np.int32)
    _check(L.lps_abc_operators(device, C.byref(r), C.byref(m), dt, 2.0 / m_ref, rcond,
                               Gu.ctypes.data, Gv.ctypes.data,
                               Gth.ctypes.data if with_theta else None, rank.ctypes.data))
    return Gu, Gv, Gth, rank


def system(nodes: np.ndarray, abc_idx: np.ndarray, tau: np.ndarray, clouds: Sequence[np.ndarray],
           max_cloud: int, modes, dt: float, m_ref: float, kind: int, first: int = 0,
           count: Optional[int] = None, device: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """The least-squares systems (A: count x m x n, B: count x m x nrhs) of one operator."""
    L = solver.load_library()
    keep: List[np.ndarray] = []
    r, md = _structs(nodes, abc_idx, tau, clouds, max_cloud, modes, keep)
    count = len(abc_idx) - first if count is None else count
    m = 4 * len(modes)
    n = (2 if kind == KIND_THETA else 4) * max_cloud
    nrhs = 1 if kind == KIND_THETA else 2
    A = np.zeros((count, n, m))  # column-major per problem: A[p].T is the m x n matrix
    B = np.zeros((count, nrhs, m))
    _check(L.lps_abc_system(device, C.byref(r), C.byref(md), dt, 2.0 / m_ref, kind, first, count,
                            A.ctypes.data, B.ctypes.data))
    return A.transpose(0, 2, 1), B.transpose(0, 2, 1)


def lstsq(A: np.ndarray, B: np.ndarray, rcond: float = 1e-14, device: int = 0
          ) -> Tuple[np.ndarray, np.ndarray]:
    """Minimum-norm least-squares solutions of a batch: A (batch x m x n), B (batch x m x nrhs)
    -> X (batch x n x nrhs), rank (batch)."""
    L = solver.load_library()
    batch, m, n = A.shape
    nrhs = B.shape[2]
    Af = np.ascontiguousarray(A.transpose(0, 2, 1))  # column-major per problem
    Bf = np.ascontiguousarray(B.transpose(0, 2, 1))
    X = np.zeros((batch, nrhs, n))
    rank = np.zeros(batch, dtype=np.int32)
    _check(L.lps_lstsq_batched(device, m, n, nrhs, batch, Af.ctypes.data, Bf.ctypes.data, rcond,
                               X.ctypes.data, rank.ctypes.data))
    return X.transpose(0, 2, 1), rank
