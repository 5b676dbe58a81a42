"""The solver's file interface: the job input files and the .energy/.verify outputs.

Writers follow the preprocessor's export (preprocessor/main.cpp:1294-1453); readers follow the
solver's reads (src/main.cpp:178-293) with the semantics of src/helpers/utils.h:24-80:
whitespace-separated text, integer index arrays read as double and rounded with (int)(x + 0.5),
so a -1 pad in .iClouds becomes node 0.
"""
from __future__ import annotations

import os
from typing import Dict, List

import numpy as np

from .preprocess import Problem

SUBSTEP_NO = 30  # src/main.cpp:34

ENERGY_HEADER = "# t  E_total  E_kinetic  E_potential  P_x  P_y  max_u  max_theta"   # src/main.cpp:697
VERIFY_HEADER = "# t  E_total  dE/E0  P_x  P_y  max_theta  max_u  max_v  max_a"     # src/main.cpp:705


def fmt16(v: float) -> str:
    """A double as C++ iostream prints it with setprecision(16)."""
    return format(float(v), ".16g")


def _write_lines(path: str, rows: List[str]) -> None:
    with open(path, "w") as f:
        f.write("\n".join(rows))
        if rows:
            f.write("\n")


def _write_matrix(path: str, mat: np.ndarray, fmt) -> None:
    with open(path, "w") as f:
        for row in mat:
            f.write(" ".join(fmt(v) for v in row))
            f.write("\n")


def fmt17(v: float) -> str:
    """A double with 17 significant digits: reads back to the same double."""
    return format(float(v), ".17g")


def write_job(p: Problem, job: str, exact_text: bool = False) -> None:
    """Write the input files <job>.* (preprocessor/main.cpp:1294-1453).

    The preprocessor prints 16 significant digits; exact_text=True prints 17 so that every double
    reads back unchanged (used for test fixtures whose binary and text forms must agree)."""
    f = fmt17 if exact_text else fmt16
    d = os.path.dirname(job)
    if d:
        os.makedirs(d, exist_ok=True)
    _write_lines(job + ".data1", [str(p.N), str(p.steps), str(p.prenotch.shape[0]),
                                  str(p.vinodes.shape[0]), str(p.C), str(p.max_cloud)])
    _write_lines(job + ".data2", [f(v) for v in p.data2])
    _write_matrix(job + ".nodes", p.nodes, f)
    _write_lines(job + ".bctypes", [str(int(v)) for v in p.bctypes])
    _write_lines(job + ".bcvalues", [f(v) for v in p.bcvalues])
    _write_lines(job + ".icvalues", [f(v) for v in p.ic])
    _write_lines(job + ".icdvalues", [f(v) for v in p.icd])
    _write_matrix(job + ".prenotch", p.prenotch, f)
    _write_lines(job + ".vinodes", [str(int(v)) for v in p.vinodes])
    if p.C > 0:
        _write_lines(job + ".infNodesIndex", [str(int(v)) for v in p.abc])
        _write_matrix(job + ".iClouds", p.clouds, lambda v: str(int(v)))
        _write_matrix(job + ".ciNodesU", p.Gu, f)
        _write_matrix(job + ".ciNodesV", p.Gv, f)
        _write_matrix(job + ".ciNodesTheta", p.Gth, f)


def _read_numbers(path: str, count: int) -> np.ndarray:
    """`count` whitespace-separated doubles; missing trailing values read as 0 like a failed
    iostream extraction (utils.h:24-34)."""
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    with open(path) as f:
        toks = f.read().split()
    out = np.zeros(count)
    n = min(count, len(toks))
    if n:
        out[:n] = np.array(toks[:n], dtype=np.float64)
    return out


def _round_int(a: np.ndarray) -> np.ndarray:
    """static_cast<int>(x + 0.5) (utils.h:61, :76): truncation toward zero."""
    return np.trunc(a + 0.5).astype(np.int64)


def read_job(job: str) -> Problem:
    """Read <job>.* with the solver's semantics (src/main.cpp:178-293)."""
    d1 = _read_numbers(job + ".data1", 6).astype(np.int64)
    N, steps, npre, nvi, C, M = (int(v) for v in d1)
    data2 = _read_numbers(job + ".data2", 15)
    nodes = _read_numbers(job + ".nodes", 2 * N).reshape(N, 2)
    bctypes = _read_numbers(job + ".bctypes", 2 * N).astype(np.int64)
    bcvalues = _read_numbers(job + ".bcvalues", 2 * N)
    icd = _read_numbers(job + ".icdvalues", 2 * N)
    ic = _read_numbers(job + ".icvalues", 2 * N)
    pre = _read_numbers(job + ".prenotch", 4 * npre).reshape(npre, 4) if npre > 0 else np.zeros((0, 4))
    vin = _read_numbers(job + ".vinodes", nvi).astype(np.int64) if nvi > 0 else np.zeros(0, np.int64)
    if C > 0:
        Gu = _read_numbers(job + ".ciNodesU", 2 * C * 4 * M).reshape(2 * C, 4 * M)
        Gv = _read_numbers(job + ".ciNodesV", 2 * C * 4 * M).reshape(2 * C, 4 * M)
        clouds = _round_int(_read_numbers(job + ".iClouds", C * M)).reshape(C, M)
        abc = _round_int(_read_numbers(job + ".infNodesIndex", C))
        Gth = _read_numbers(job + ".ciNodesTheta", C * 2 * M).reshape(C, 2 * M)
    else:
        Gu = Gv = Gth = np.zeros((0, 0))
        clouds = np.zeros((0, 0), dtype=np.int64)
        abc = np.zeros(0, dtype=np.int64)
    return Problem(nodes=nodes, steps=steps, data2=data2, bctypes=bctypes, bcvalues=bcvalues,
                   ic=ic, icd=icd, abc=abc, clouds=clouds, Gu=Gu, Gv=Gv, Gth=Gth,
                   prenotch=pre, vinodes=vin, info={"job": job})


ENERGY_COLUMNS = ("t", "energy", "energy_k", "energy_p", "px", "py", "max_u", "theta_max",
                  "ub_rms", "vb_rms")
VERIFY_COLUMNS = ("t", "energy", "de_rel", "px", "py", "theta_max", "max_u", "max_v", "max_a")


def write_series(path: str, records: List[Dict[str, float]], kind: str) -> None:
    """Output records (solver.Solver.run dicts) as a .energy or .verify file: header, then one
    row per record, precision 16, as src/main.cpp:694-707, 1106-1123 writes them."""
    header, cols = (ENERGY_HEADER, ENERGY_COLUMNS) if kind == "energy" else \
        (VERIFY_HEADER, VERIFY_COLUMNS)
    _write_lines(path, [header] + [" ".join(fmt16(r[c]) for c in cols) for r in records])


def read_series(path: str) -> np.ndarray:
    """Rows of a .energy or .verify file (header skipped) as a float array."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            rows.append([float(t) for t in line.split()])
    return np.array(rows, dtype=np.float64)


def read_vtk_field(path: str) -> Dict[str, np.ndarray]:
    """Points and the first data block of a legacy VTK file written by the solver
    (src/helpers/utils.h:94-224): ASCII, or BINARY (big-endian doubles, int32 vertices)."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw.split(b"\n", 3)[2].strip() == b"BINARY":
        return _read_vtk_binary(raw)
    toks = raw.decode().split()
    out: Dict[str, np.ndarray] = {}
    i = toks.index("POINTS")
    n = int(toks[i + 1])
    pts = np.array(toks[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
    out["points"] = pts[:, :2]
    if "VECTORS" in toks:
        j = toks.index("VECTORS")
        out["vectors"] = np.array(toks[j + 3:j + 3 + 3 * n], dtype=np.float64).reshape(n, 3)[:, :2]
    else:
        j = toks.index("LOOKUP_TABLE")
        out["scalars"] = np.array(toks[j + 2:j + 2 + n], dtype=np.float64)
    return out


def _read_vtk_binary(raw: bytes) -> Dict[str, np.ndarray]:
    def after(key: bytes, start: int = 0):
        k = raw.index(key, start)
        e = raw.index(b"\n", k)
        return raw[k:e].split(), e + 1

    out: Dict[str, np.ndarray] = {}
    head, p = after(b"POINTS ")
    n = int(head[1])
    out["points"] = np.frombuffer(raw, ">f8", 3 * n, p).reshape(n, 3)[:, :2].astype(np.float64)
    head, p = after(b"VERTICES ", p + 24 * n)
    head, p = after(b"POINT_DATA ", p + 4 * int(head[2]))
    if raw.startswith(b"VECTORS", p):
        head, p = after(b"VECTORS ", p)
        out["vectors"] = np.frombuffer(raw, ">f8", 3 * n, p).reshape(n, 3)[:, :2].astype(np.float64)
    else:
        head, p = after(b"LOOKUP_TABLE ", p)
        out["scalars"] = np.frombuffer(raw, ">f8", n, p).astype(np.float64)
    return out
