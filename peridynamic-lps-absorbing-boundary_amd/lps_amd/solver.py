"""ctypes binding of liblps_amd.so (the C ABI declared in include/lps_amd.h).

This is the Python face of the drop-in: a `Solver` is one lps_ctx, created from a
`preprocess.Problem` (the contents of the job files). There is no fallback: if the HIP library is
missing or no GPU is present, creating a solver raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, List, Optional

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "liblps_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(PKG_DIR), "include", "lps_amd.h")

LPS_OK = 0
LPS_ERR_ARG = -1
LPS_ERR_HIP = -2
LPS_ERR_HORIZON = -3
LPS_ERR_NONFINITE = -4
LPS_ERR_CONSISTENCY = -5
LPS_THETA_IDW = 0
LPS_THETA_GTHETA = 1


class LpsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lps error {code}: {msg}")
        self.code = code


class LpsProblem(C.Structure):
    _fields_ = [("node_no", C.c_int32), ("constrains_no", C.c_int32),
                ("max_cloud_size", C.c_int32),
                ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double), ("dt", C.c_double),
                ("delta", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
                ("nodes", C.c_void_p), ("bctypes", C.c_void_p), ("icvalues", C.c_void_p),
                ("icdvalues", C.c_void_p), ("inf_nodes", C.c_void_p), ("iclouds", C.c_void_p),
                ("ci_nodes_u", C.c_void_p), ("ci_nodes_v", C.c_void_p),
                ("ci_nodes_theta", C.c_void_p), ("prenotch_no", C.c_int32),
                ("prenotches", C.c_void_p)]


class LpsAbcModes(C.Structure):
    _fields_ = [("count", C.c_int32)] + [(n, C.c_void_p) for n in (
        "a1", "a2", "om_p", "om_s", "g_p", "g_s", "s_r", "s_i")]


class LpsAbcRows(C.Structure):
    _fields_ = [("count", C.c_int32), ("max_cloud", C.c_int32), ("xy", C.c_void_p),
                ("tau", C.c_void_p), ("cloud_len", C.c_void_p), ("cloud_xy", C.c_void_p)]


class LpsOptions(C.Structure):
    _fields_ = [("theta_method", C.c_int32), ("abc_enabled", C.c_int32), ("exact", C.c_int32),
                ("use_graph", C.c_int32), ("device", C.c_int32), ("horizon_radius", C.c_double),
                ("variant", C.c_int32), ("concurrent", C.c_int32)]


class LpsDiag(C.Structure):
    _fields_ = [("step", C.c_int64)] + [(n, C.c_double) for n in (
        "t", "energy", "energy_k", "energy_p", "px", "py", "max_u", "max_v", "max_a",
        "theta_max", "theta_min", "ub_rms", "vb_rms", "de_rel")] + [
        ("max_u_node", C.c_int32), ("max_v_node", C.c_int32)]


DIAG_FIELDS = [f[0] for f in LpsDiag._fields_]

# Every function include/lps_amd.h declares, with its ctypes signature.
SIGNATURES = {
    "lps_default_options": (None, [C.POINTER(LpsOptions)]),
    "lps_create": (C.c_int32, [C.POINTER(LpsProblem), C.POINTER(LpsOptions), C.POINTER(C.c_void_p)]),
    "lps_reset": (C.c_int32, [C.c_void_p]),
    "lps_set_initial": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "lps_abc_operators": (C.c_int32, [C.c_int32, C.POINTER(LpsAbcRows), C.POINTER(LpsAbcModes),
                                      C.c_double, C.c_double, C.c_double] + [C.c_void_p] * 4),
    "lps_abc_system": (C.c_int32, [C.c_int32, C.POINTER(LpsAbcRows), C.POINTER(LpsAbcModes),
                                   C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_int32,
                                   C.c_void_p, C.c_void_p]),
    "lps_lstsq_batched": (C.c_int32, [C.c_int32] * 5 + [C.c_void_p, C.c_void_p, C.c_double,
                                                        C.c_void_p, C.c_void_p]),
    "lps_checkpoint_save": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "lps_checkpoint_load": (C.c_int32, [C.c_void_p, C.c_char_p]),
    "lps_create_batch": (C.c_int32, [C.c_void_p, C.c_int32, C.POINTER(LpsOptions),
                                     C.POINTER(C.c_void_p)]),
    "lps_instance_count": (C.c_int32, [C.c_void_p]),
    "lps_run_batch": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    "lps_diagnostics_batch": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_fetch_instance": (C.c_int32, [C.c_void_p, C.c_int32] + [C.c_void_p] * 5),
    "lps_step": (C.c_int32, [C.c_void_p, C.c_int32]),
    "lps_run": (C.c_int32, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]),
    "lps_diagnostics": (C.c_int32, [C.c_void_p, C.POINTER(LpsDiag)]),
    "lps_fetch": (C.c_int32, [C.c_void_p] + [C.c_void_p] * 5),
    "lps_horizons": (C.c_int64, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "lps_step_index": (C.c_int64, [C.c_void_p]),
    "lps_m_ref": (C.c_double, [C.c_void_p]),
    "lps_set_timing": (C.c_int32, [C.c_void_p, C.c_int32]),
    "lps_pass_times": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_describe": (C.c_char_p, [C.c_void_p]),
    "lps_tile_info": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_last_error": (C.c_char_p, [C.c_void_p]),
    "lps_global_error": (C.c_char_p, []),
    "lps_destroy": (None, [C.c_void_p]),
    # learned ABC closure
    "lps_closure_param_count": (C.c_int64, []),
    "lps_closure_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "lps_closure_destroy": (None, [C.c_void_p]),
    "lps_closure_error": (C.c_char_p, [C.c_void_p]),
    "lps_closure_get_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_closure_set_params": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_closure_set_scales": (C.c_int32, [C.c_void_p, C.c_float, C.c_float]),
    "lps_closure_get_scales": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_closure_forward": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "lps_closure_forward_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "lps_closure_train_batch": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.c_int32, C.c_float, C.c_void_p, C.c_void_p]),
    "lps_closure_dataset_reserve": (C.c_int32, [C.c_void_p, C.c_int64]),
    "lps_closure_dataset_count": (C.c_int64, [C.c_void_p]),
    "lps_closure_dataset_clear": (C.c_int32, [C.c_void_p]),
    "lps_closure_dataset_append": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]),
    "lps_closure_dataset_read": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                             C.c_void_p]),
    "lps_closure_train": (C.c_int32, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                      C.c_int64, C.c_int64, C.c_uint64, C.c_float, C.c_void_p]),
    "lps_closure_evaluate": (C.c_int32, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "lps_closure_collect": (C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(LpsProblem), C.c_int32,
                                        C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                        C.c_uint64]),
    "lps_closure_collect_closed": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                               C.c_int32, C.c_int32, C.c_double, C.c_double,
                                               C.c_double, C.c_uint64]),
    "lps_closure_attach": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_closure_detach": (C.c_int32, [C.c_void_p]),
    "lps_closure_set_gain": (C.c_int32, [C.c_void_p, C.c_void_p]),
    "lps_closure_inspect": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "lps_closure_guard": (C.c_int32, [C.c_void_p, C.c_double, C.c_void_p]),
    "lps_closure_set_gate": (C.c_int32, [C.c_void_p, C.c_void_p]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load liblps_amd.so; raises OSError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing; build it with __graft_entry__.build() or `make -C "
                      f"{PKG_DIR}`")
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


def diag_dict(d: LpsDiag) -> Dict[str, float]:
    return {k: getattr(d, k) for k in DIAG_FIELDS}


def _lps_problem(prob, keep: List[np.ndarray]) -> "LpsProblem":
    """The C struct of a problem; the arrays it points to are appended to keep."""
    def arr(a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a

    N, Cn, M = prob.N, prob.C, prob.max_cloud
    d2 = prob.data2
    p = LpsProblem()
    p.node_no, p.constrains_no, p.max_cloud_size = N, Cn, M
    p.E, p.nu, p.rho, p.dt = d2[0], d2[1], d2[2], d2[6]
    p.delta, p.dx, p.dy = d2[7], d2[8], d2[9]
    p.nodes = _ptr(arr(prob.nodes, np.float64))
    p.bctypes = _ptr(arr(prob.bctypes, np.int32))
    p.icvalues = _ptr(arr(prob.ic, np.float64))
    p.icdvalues = _ptr(arr(prob.icd, np.float64))
    if Cn > 0:
        cl = np.asarray(prob.clouds)
        # the file reader's (int)(x + 0.5): a -1 pad becomes node 0
        cl = np.trunc(cl.astype(np.float64) + 0.5).astype(np.int32)
        p.inf_nodes = _ptr(arr(prob.abc, np.int32))
        p.iclouds = _ptr(arr(cl, np.int32))
        p.ci_nodes_u = _ptr(arr(prob.Gu, np.float64))
        p.ci_nodes_v = _ptr(arr(prob.Gv, np.float64))
        if prob.Gth is not None and prob.Gth.size:
            p.ci_nodes_theta = _ptr(arr(prob.Gth, np.float64))
    pre = np.asarray(getattr(prob, "prenotch", np.zeros((0, 4))), dtype=np.float64)
    if pre.shape[0] > 0:
        p.prenotch_no = pre.shape[0]
        p.prenotches = _ptr(arr(pre.reshape(-1, 4), np.float64))
    return p


def _lps_options(L, theta_method, abc_enabled, exact, use_graph, device, horizon_radius, variant,
                 concurrent) -> "LpsOptions":
    o = LpsOptions()
    L.lps_default_options(C.byref(o))
    o.theta_method = theta_method
    o.abc_enabled = 1 if abc_enabled else 0
    o.exact = 1 if exact else 0
    o.use_graph = 1 if use_graph else 0
    o.device = device
    o.horizon_radius = horizon_radius
    o.variant = variant
    if concurrent is not None:
        # 0 the ABC rows after the tiles, 1 beside them (second stream), 2 also the partial tiles
        # beside the full ones (third stream), 3 the ABC rows inside the dilatation launch
        o.concurrent = int(concurrent)
    return o


class Solver:
    """One lps_ctx on the GPU for a problem (lps_amd.preprocess.Problem)."""

    def __init__(self, prob, theta_method: int = LPS_THETA_GTHETA, abc_enabled: bool = True,
                 exact: bool = False, use_graph: bool = True, device: int = 0,
                 horizon_radius: float = 0.0, variant: int = 0, concurrent: Optional[int] = None):
        L = load_library()
        self._keep: List[np.ndarray] = []
        N = prob.N
        d2 = prob.data2
        p = _lps_problem(prob, self._keep)
        o = _lps_options(L, theta_method, abc_enabled, exact, use_graph, device, horizon_radius,
                         variant, concurrent)
        h = C.c_void_p()
        rc = L.lps_create(C.byref(p), C.byref(o), C.byref(h))
        if rc != LPS_OK:
            raise LpsError(rc, L.lps_global_error().decode())
        self._h = h
        self._L = L
        self._keep.clear()
        self.N = N
        self.dt = float(d2[6])

    def close(self):
        if getattr(self, "_h", None):
            self._L.lps_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise LpsError(rc, self._L.lps_last_error(self._h).decode())
        return rc

    def reset(self):
        self._check(self._L.lps_reset(self._h))

    def set_initial(self, ic: np.ndarray, icd: Optional[np.ndarray] = None):
        """New initial conditions (interleaved, problem node order), then back to step 0."""
        ic = np.ascontiguousarray(ic, dtype=np.float64)
        icd = np.zeros_like(ic) if icd is None else np.ascontiguousarray(icd, dtype=np.float64)
        assert ic.size == 2 * self.N and icd.size == 2 * self.N
        self._check(self._L.lps_set_initial(self._h, _ptr(ic), _ptr(icd)))

    def checkpoint(self, path: str):
        """Writes the time-loop state, step count and E0 to path (lps_checkpoint_save)."""
        self._check(self._L.lps_checkpoint_save(self._h, os.fsencode(path)))

    def restore(self, path: str):
        """Restores a checkpoint of the same problem and options (lps_checkpoint_load)."""
        self._check(self._L.lps_checkpoint_load(self._h, os.fsencode(path)))

    def step(self, nsteps: int):
        self._check(self._L.lps_step(self._h, nsteps))

    def run(self, nsteps: int) -> List[Dict[str, float]]:
        recs = (LpsDiag * max(1, nsteps // 30 + 1))()
        n = self._check(self._L.lps_run(self._h, nsteps, recs, len(recs)))
        return [diag_dict(recs[i]) for i in range(n)]

    def diagnostics(self) -> Dict[str, float]:
        d = LpsDiag()
        self._check(self._L.lps_diagnostics(self._h, C.byref(d)))
        return diag_dict(d)

    def fetch(self) -> Dict[str, np.ndarray]:
        N = self.N
        u, v, a = (np.zeros(2 * N) for _ in range(3))
        th, m = np.zeros(N), np.zeros(N)
        self._check(self._L.lps_fetch(self._h, _ptr(u), _ptr(v), _ptr(a), _ptr(th), _ptr(m)))
        return {"u": u, "v": v, "a": a, "theta": th, "m": m}

    def horizons(self):
        nnz = self._check(self._L.lps_horizons(self._h, None, None))
        rp = np.zeros(self.N + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        self._check(self._L.lps_horizons(self._h, _ptr(rp), _ptr(cols)))
        return rp, cols[:nnz]

    @property
    def path(self) -> str:
        return self._L.lps_describe(self._h).decode()

    @property
    def step_index(self) -> int:
        return self._L.lps_step_index(self._h)

    @property
    def m_ref(self) -> float:
        return self._L.lps_m_ref(self._h)

    def set_timing(self, on):
        """True / 1: the four pass events per step; 2: only the force kernel's two; 3: only the two
        around the learned closure's network kernel (a closure attached); False / 0: off. In modes
        2 and 3 pass_times()["force_ms"] holds that kernel's time."""
        mode = int(on) if (not isinstance(on, bool) and on in (2, 3)) else (1 if on else 0)
        self._check(self._L.lps_set_timing(self._h, mode))

    def pass_times(self):
        out = np.zeros(4)
        self._check(self._L.lps_pass_times(self._h, _ptr(out)))
        return {"theta_ms": float(out[0]), "force_ms": float(out[1]), "abc_ms": float(out[2]),
                "step_ms": float(out[3])}

    # learned ABC closure (lps_amd.closure.Closure)
    def attach_closure(self, clo):
        """Apply the learned closure clo at every ABC row from the next step on."""
        self._check(self._L.lps_closure_attach(self._h, clo._h))
        self._closure = clo  # kept alive while attached

    def detach_closure(self):
        self._check(self._L.lps_closure_detach(self._h))
        self._closure = None

    def set_closure_gain(self, gain):
        g = np.ascontiguousarray(gain, dtype=np.float32)
        self._check(self._L.lps_closure_set_gain(self._h, _ptr(g)))

    def set_closure_gate(self, min_su):
        g = np.ascontiguousarray(min_su, dtype=np.float32)
        self._check(self._L.lps_closure_set_gate(self._h, _ptr(g)))

    def closure_guard(self, tol: float = -1.0):
        """Set the fallback guard's tolerance (tol >= 0) and return each instance's fallback step
        (-1: the closure is still on)."""
        n = getattr(self, "count", 1)
        out = np.zeros(n, dtype=np.int64)
        self._check(self._L.lps_closure_guard(self._h, float(tol), _ptr(out)))
        return out

    def closure_inspect(self, n_rows: int):
        """Inputs (n_rows x 1600 BF16 bits), scales (n_rows x 2) and applied outputs (n_rows x 8)
        of the last completed step."""
        x = np.zeros((n_rows, 1600), dtype=np.uint16)
        sc = np.zeros((n_rows, 2), dtype=np.float32)
        ap = np.zeros((n_rows, 8), dtype=np.float32)
        self._check(self._L.lps_closure_inspect(self._h, _ptr(x), _ptr(sc), _ptr(ap)))
        return x, sc, ap

    def tile_info(self):
        """Tile-path layout: tiles, full tiles, partial tiles, nodes in full tiles, nodes (zeros on
        the exact path)."""
        out = np.zeros(5, dtype=np.int64)
        self._check(self._L.lps_tile_info(self._h, _ptr(out)))
        return {"tiles": int(out[0]), "full_tiles": int(out[1]), "partial_tiles": int(out[2]),
                "full_tile_nodes": int(out[3]), "nodes": int(out[4])}


class Batch(Solver):
    """Many problems advanced together in one lps_ctx (lps_create_batch): the same dx, dy and
    delta, each with its own radius, nu, dt, ABC operators and initial conditions."""

    def __init__(self, probs, theta_method: int = LPS_THETA_GTHETA, use_graph: bool = True,
                 device: int = 0, variant: int = 0, concurrent: Optional[int] = None):
        L = load_library()
        self._keep: List[np.ndarray] = []
        structs = (LpsProblem * len(probs))()
        for i, p in enumerate(probs):
            structs[i] = _lps_problem(p, self._keep)
        o = _lps_options(L, theta_method, True, False, use_graph, device, 0.0, variant,
                         concurrent)
        h = C.c_void_p()
        rc = L.lps_create_batch(structs, len(probs), C.byref(o), C.byref(h))
        if rc != LPS_OK:
            raise LpsError(rc, L.lps_global_error().decode())
        self._h = h
        self._L = L
        self._keep.clear()
        self.sizes = [p.N for p in probs]
        self.count = L.lps_instance_count(h)

    def run(self, nsteps: int) -> List[List[Dict[str, float]]]:
        """Output records of every instance: result[i] is instance i's list of records."""
        nrec = max(1, nsteps // 30 + 1)
        recs = (LpsDiag * (nrec * self.count))()
        n = self._check(self._L.lps_run_batch(self._h, nsteps, recs, nrec))
        return [[diag_dict(recs[k * self.count + i]) for k in range(n)] for i in range(self.count)]

    def diagnostics(self) -> List[Dict[str, float]]:
        d = (LpsDiag * self.count)()
        self._check(self._L.lps_diagnostics_batch(self._h, d))
        return [diag_dict(d[i]) for i in range(self.count)]

    def fetch(self, instance: int = 0) -> Dict[str, np.ndarray]:
        N = self.sizes[instance]
        u, v, a = (np.zeros(2 * N) for _ in range(3))
        th, m = np.zeros(N), np.zeros(N)
        self._check(self._L.lps_fetch_instance(self._h, instance, _ptr(u), _ptr(v), _ptr(a),
                                               _ptr(th), _ptr(m)))
        return {"u": u, "v": v, "a": a, "theta": th, "m": m}
