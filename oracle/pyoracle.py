"""ctypes binding of the CPU oracle (liblps_oracle.so). TEST INFRASTRUCTURE ONLY: tests, the smoke
entry and the bench's CPU-baseline leg use it as the checker; the product path never imports it."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblps_oracle.so")


class OracleProblem(C.Structure):
    _fields_ = [("N", C.c_int32), ("C", C.c_int32), ("M", C.c_int32),
                ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double), ("dt", C.c_double),
                ("delta", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nodes", C.c_void_p), ("bctypes", C.c_void_p), ("ic", C.c_void_p),
                ("icd", C.c_void_p), ("abc", C.c_void_p), ("clouds", C.c_void_p),
                ("Gu", C.c_void_p), ("Gv", C.c_void_p), ("Gth", C.c_void_p),
                ("theta_method", C.c_int32), ("abc_enabled", C.c_int32),
                ("n_notch", C.c_int32), ("notch", C.c_void_p), ("plant", C.c_int32)]


class OracleDiag(C.Structure):
    _fields_ = [("step", C.c_int64)] + [(n, C.c_double) for n in (
        "t", "energy", "energy_k", "energy_p", "px", "py", "max_u", "max_v", "max_a",
        "theta_max", "theta_min", "ub_rms", "vb_rms", "de_rel")] + [
        ("max_u_node", C.c_int32), ("max_v_node", C.c_int32)]


DIAG_FIELDS = [f[0] for f in OracleDiag._fields_]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} missing: build it with `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        L.oracle_build_horizons.restype = C.c_int64
        L.oracle_build_horizons.argtypes = [C.c_int32, C.c_void_p, C.c_double, C.c_int32,
                                            C.c_void_p, C.c_void_p]
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(OracleProblem), C.c_int32]
        L.oracle_destroy.argtypes = [C.c_void_p]
        L.oracle_error.restype = C.c_char_p
        L.oracle_run.restype = C.c_int32
        L.oracle_run.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
        L.oracle_diagnostics.argtypes = [C.c_void_p, C.c_void_p]
        L.oracle_fetch.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.oracle_step_index.restype = C.c_int64
        L.oracle_step_index.argtypes = [C.c_void_p]
        L.oracle_m_ref.restype = C.c_double
        L.oracle_m_ref.argtypes = [C.c_void_p]
        L.oracle_horizons.restype = C.c_int64
        L.oracle_horizons.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_dilatation.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_force.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.oracle_segments_intersect.restype = C.c_int
        L.oracle_segments_intersect.argtypes = [C.c_double] * 8
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def build_horizons(nodes: np.ndarray, r: float, max_len: int = 128):
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    N = nodes.shape[0]
    rp = np.zeros(N + 1, dtype=np.int64)
    nnz = lib().oracle_build_horizons(N, _ptr(nodes), r, max_len, _ptr(rp), None)
    if nnz < 0:
        raise RuntimeError(lib().oracle_error().decode())
    cols = np.zeros(max(nnz, 1), dtype=np.int32)
    lib().oracle_build_horizons(N, _ptr(nodes), r, max_len, _ptr(rp), _ptr(cols))
    return rp, cols[:nnz]


def segments_intersect(a, b, c, d) -> bool:
    """Prenotch visibility test of segment a-b against segment c-d (points as (x, y))."""
    return bool(lib().oracle_segments_intersect(float(a[0]), float(a[1]), float(b[0]), float(b[1]),
                                                float(c[0]), float(c[1]), float(d[0]), float(d[1])))


def diag_to_dict(d: OracleDiag) -> dict:
    return {k: getattr(d, k) for k in DIAG_FIELDS}


class Oracle:
    """The CPU oracle on one problem (lps_amd.preprocess.Problem)."""

    def __init__(self, prob, theta_method: int = 1, abc_enabled: bool = True, threads: int = 0,
                 plant: int = 0):
        """plant (tests only): 1 swaps the IDW weight 1/d for 1/d^2, a mistake the reference's
        IDW fixtures must catch."""
        self._keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        N, Cn, M = prob.N, prob.C, prob.max_cloud
        d2 = prob.data2
        p = OracleProblem()
        p.N, p.C, p.M = N, Cn, M
        p.E, p.nu, p.rho, p.dt = d2[0], d2[1], d2[2], d2[6]
        p.delta, p.dx, p.dy = d2[7], d2[8], d2[9]
        p.nodes = _ptr(arr(prob.nodes, np.float64))
        p.bctypes = _ptr(arr(prob.bctypes, np.int32))
        p.ic = _ptr(arr(prob.ic, np.float64))
        p.icd = _ptr(arr(prob.icd, np.float64))
        if Cn > 0:
            cl = np.trunc(prob.clouds + 0.5).astype(np.int32) if prob.clouds.dtype.kind == "f" else \
                np.where(prob.clouds < 0, 0, prob.clouds).astype(np.int32)
            p.abc = _ptr(arr(prob.abc, np.int32))
            p.clouds = _ptr(arr(cl, np.int32))
            p.Gu = _ptr(arr(prob.Gu, np.float64))
            p.Gv = _ptr(arr(prob.Gv, np.float64))
            p.Gth = _ptr(arr(prob.Gth, np.float64)) if prob.Gth.size else None
        p.theta_method = theta_method
        p.plant = plant
        p.abc_enabled = 1 if (abc_enabled and Cn > 0) else 0
        pre = np.asarray(getattr(prob, "prenotch", np.zeros((0, 4))), dtype=np.float64)
        if pre.shape[0] > 0:
            p.n_notch = pre.shape[0]
            p.notch = _ptr(arr(pre.reshape(-1, 4), np.float64))
        self.N = N
        self.dt = float(d2[6])
        self._c = lib().oracle_create(C.byref(p), threads)
        if not self._c:
            raise RuntimeError(lib().oracle_error().decode())

    def __del__(self):
        if getattr(self, "_c", None):
            lib().oracle_destroy(self._c)
            self._c = None

    def run(self, nsteps: int):
        recs = (OracleDiag * max(1, nsteps // 30 + 1))()
        n = lib().oracle_run(self._c, nsteps, recs, len(recs))
        if n < 0:
            raise FloatingPointError("NaN/Inf in energy or displacement")
        return [diag_to_dict(recs[i]) for i in range(n)]

    def diagnostics(self) -> dict:
        d = OracleDiag()
        lib().oracle_diagnostics(self._c, C.byref(d))
        return diag_to_dict(d)

    def fetch(self):
        N = self.N
        u, v, a = (np.zeros(2 * N) for _ in range(3))
        th, m = np.zeros(N), np.zeros(N)
        lib().oracle_fetch(self._c, _ptr(u), _ptr(v), _ptr(a), _ptr(th), _ptr(m))
        return {"u": u, "v": v, "a": a, "theta": th, "m": m}

    @property
    def step(self) -> int:
        return lib().oracle_step_index(self._c)

    @property
    def m_ref(self) -> float:
        return lib().oracle_m_ref(self._c)

    def horizons(self):
        nnz = lib().oracle_horizons(self._c, None, None)
        rp = np.zeros(self.N + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int32)
        lib().oracle_horizons(self._c, _ptr(rp), _ptr(cols))
        return rp, cols

    def dilatation(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        th = np.zeros(self.N)
        lib().oracle_dilatation(self._c, _ptr(u), _ptr(th))
        return th

    def force(self, u, theta):
        u = np.ascontiguousarray(u, dtype=np.float64)
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        acc = np.zeros(2 * self.N)
        lib().oracle_force(self._c, _ptr(u), _ptr(theta), _ptr(acc))
        return acc
