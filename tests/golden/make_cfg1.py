"""Writes tests/golden/cfg1_reference.npz: the reference solver's outputs for the full-size Example 1
workload (R = 125 mm, nu = 0.25, 196,321 nodes, 6,268 ABC rows, 4000 steps).

Run in a container where the reference build exists (`make -C oracle`):
    python tests/golden/make_cfg1.py [--run-dir DIR]

It generates the workload (lps_amd.configs "cfg1"), writes the job files with 17 digits, runs
oracle/_ref/PD2_LPS_ABC (built from src/main.cpp) with every VTK snapshot but the last suppressed
(a directory where a snapshot file would go makes the reference's open fail; src/helpers/utils.h:112-113)
and stores:
  energy, verify       the reference's .energy / .verify series (133 records)
  sample, u, v, theta  a seeded sample of 4096 points of its last VTK snapshot (step 3990):
                       node ids and the 12-digit field values
  norms                L2 norms of u, v, theta over all snapshot points
  g_sha                SHA-256 of the generated G_u, G_v, G_theta and nodes: a consumer that
                       regenerates the workload compares only when its operators are the same bits
The whole run takes several minutes on 8 cores. With --run-dir, an existing run directory (job files
and outputs of a previous run of this script) is reused.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "peridynamic-lps-absorbing-boundary_amd"))
from lps_amd import configs, files  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "PD2_LPS_ABC")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg1_reference.npz")


def operator_sha(p) -> str:
    h = hashlib.sha256()
    for a in (p.nodes, p.Gu, p.Gv, p.Gth):
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--run-dir", default=None)
    args = ap.parse_args()
    d = args.run_dir or os.path.join("/tmp", "cfg1_reference_run")
    job = os.path.join(d, "job")
    if not os.path.exists(job + "_out.verify"):
        os.makedirs(d, exist_ok=True)
        prob = configs.make("cfg1", workers=os.cpu_count())
        files.write_job(prob, job, exact_text=True)
        last = (prob.steps // 30) * 30
        for s in range(0, prob.steps + 1, 30):
            if s != last:
                for pre in ("-", "-vel-", "-theta-"):
                    os.makedirs(f"{job}_out{pre}{s}.vtk", exist_ok=True)
        t = time.time()
        r = subprocess.run([REF_BIN, "_out", "job"], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        print(f"reference run {time.time() - t:.0f} s")
    q = files.read_job(job)  # exactly what the reference read
    last = (q.steps // 30) * 30
    u = files.read_vtk_field(f"{job}_out-{last}.vtk")
    v = files.read_vtk_field(f"{job}_out-vel-{last}.vtk")
    th = files.read_vtk_field(f"{job}_out-theta-{last}.vtk")
    lookup = {(float(x), float(y)): i for i, (x, y) in enumerate(q.nodes)}
    ids = np.array([lookup[(float(x), float(y))] for x, y in u["points"]])
    rng = np.random.default_rng(12345)
    pick = np.sort(rng.choice(len(ids), size=4096, replace=False))
    np.savez_compressed(
        OUT, energy=files.read_series(job + "_out.energy"),
        verify=files.read_series(job + "_out.verify"), vtk_step=np.int64(last),
        sample=ids[pick].astype(np.int32), u=u["vectors"][pick], v=v["vectors"][pick],
        theta=th["scalars"][pick],
        norms=np.array([np.linalg.norm(u["vectors"]), np.linalg.norm(v["vectors"]),
                        np.linalg.norm(th["scalars"])]),
        n_points=np.int64(len(ids)), g_sha=np.array(operator_sha(q)))
    print(OUT, os.path.getsize(OUT))


if __name__ == "__main__":
    main()
