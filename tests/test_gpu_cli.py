"""The drop-in executable PD2_LPS_ABC on golden jobs written in the reference's file format: the
.energy/.verify files and the last VTK snapshot match the reference solver's within 1e-10, and a
non-finite state ends the run with exit code 1 (src/main.cpp:1126-1133)."""
import os
import subprocess
import sys

import numpy as np
import pytest
from conftest import (ENERGY_KEYS, PKG, VERIFY_KEYS, assert_series_close, load_golden, rel_l2,
                      vtk_index)

from lps_amd import files

pytestmark = pytest.mark.gpu

EXE = os.path.join(PKG, "PD2_LPS_ABC")


def run_cli(tmp_path, prob, *opts):
    job = str(tmp_path / "job")
    files.write_job(prob, job, exact_text=True)
    r = subprocess.run([EXE, "_out", "job", *opts], cwd=tmp_path, capture_output=True, text=True,
                       timeout=240)
    return job, r


@pytest.mark.parametrize("name", ["abc_r10_nu025", "noabc_r30"])
def test_cli_reproduces_the_reference_outputs(name, tmp_path):
    print("cli", name, flush=True)
    sys.stdout.flush()
    prob, ref = load_golden(name)
    job, r = run_cli(tmp_path, prob)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert_series_close(files.read_series(job + "_out.energy"), ref["energy"], ENERGY_KEYS, 1e-10,
                        prob, name)
    assert_series_close(files.read_series(job + "_out.verify"), ref["verify"], VERIFY_KEYS, 1e-10,
                        prob, name)
    s = ref["vtk_step"]
    u = files.read_vtk_field(f"{job}_out-{s}.vtk")
    v = files.read_vtk_field(f"{job}_out-vel-{s}.vtk")
    th = files.read_vtk_field(f"{job}_out-theta-{s}.vtk")
    assert np.array_equal(u["points"], ref["vtk_points"])
    assert rel_l2(u["vectors"], ref["vtk_u"]) <= 1e-10
    assert rel_l2(v["vectors"], ref["vtk_v"]) <= 1e-10
    assert rel_l2(th["scalars"], ref["vtk_theta"]) <= 1e-10
    vtk_index(prob.nodes, u["points"])  # every point is a node


def test_cli_exits_1_on_a_nonfinite_state(tmp_path):
    print("cli nonfinite", flush=True)
    prob, _ = load_golden("abc_r10_nu025")
    prob.ic = prob.ic.copy()
    prob.ic[10] = np.nan
    job, r = run_cli(tmp_path, prob, "--vtk=off")
    assert r.returncode == 1, r.stdout[-2000:]
    assert "FATAL" in r.stdout + r.stderr


def test_cli_batch_manifest_runs_several_jobs_in_one_process(tmp_path):
    """--batch=<manifest>: the jobs advance together and each gets the reference's outputs."""
    print("cli batch", flush=True)
    names = ["abc_r10_nu025", "abc_r12_nu033"]
    steps = 300
    lines = []
    for k, name in enumerate(names):
        prob, _ = load_golden(name)
        prob.steps = steps
        files.write_job(prob, str(tmp_path / f"job{k}"), exact_text=True)
        lines.append(f"_out job{k}")
    (tmp_path / "jobs.txt").write_text("# two jobs\n" + "\n".join(lines) + "\n")
    r = subprocess.run([EXE, "--batch=jobs.txt", "--vtk=off"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for k, name in enumerate(names):
        prob, ref = load_golden(name)
        n = steps // 30
        job = str(tmp_path / f"job{k}")
        assert_series_close(files.read_series(job + "_out.energy"), ref["energy"][:n], ENERGY_KEYS,
                            1e-10, prob, name)
        assert_series_close(files.read_series(job + "_out.verify"), ref["verify"][:n], VERIFY_KEYS,
                            1e-10, prob, name)


def test_cli_learned_closure(tmp_path):
    """--abc=learned=<weights.bin>: a closure whose output layer is zero gives the least-squares
    ABC's outputs (to 1e-12: each run autotunes its own node kernels; the bitwise identity is tested
    through the C ABI with the kernels fixed); a trained closure runs; a malformed file is refused."""
    print("cli learned", flush=True)
    from lps_amd import closure as cl, learned
    prob, _ = load_golden("abc_r10_nu025")
    prob.steps = 120
    d = cl.split_params(np.random.default_rng(3).normal(0.0, 0.05, 542468).astype(np.float32))
    d["W4"][:] = 0.0
    d["b4"][:] = 0.0
    learned.export_bin(cl.join_params(d), [0.03, 0.03], str(tmp_path / "zero.bin"))
    job, r = run_cli(tmp_path, prob, "--vtk=off")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    lsq = files.read_series(job + "_out.energy")
    job, r = run_cli(tmp_path, prob, "--vtk=off", "--abc=learned=zero.bin")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "closure" in r.stdout
    assert_series_close(files.read_series(job + "_out.energy"), lsq, ENERGY_KEYS, 1e-12, prob)
    z = np.load(os.path.join(PKG, "lps_amd", "data", "closure_r125.npz"), allow_pickle=False)
    learned.export_bin(z["params"], z["sigma"], str(tmp_path / "trained.bin"))
    job, r = run_cli(tmp_path, prob, "--vtk=off", "--abc=learned=trained.bin")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    (tmp_path / "bad.bin").write_bytes(b"1234")
    job, r = run_cli(tmp_path, prob, "--vtk=off", "--abc=learned=bad.bin")
    assert r.returncode == 1 and "not a closure weight file" in r.stderr


def test_cli_checkpoint_restart_writes_the_same_files(tmp_path):
    """--stop-at + --checkpoint, then --restart: the .energy/.verify files equal, byte for byte,
    those of the run that was never stopped (--exact: the same kernels in every process)."""
    print("cli checkpoint", flush=True)
    prob, _ = load_golden("abc_r12_nu033")
    job = str(tmp_path / "job")
    files.write_job(prob, job, exact_text=True)

    def cli(*opts):
        r = subprocess.run([EXE, *opts], cwd=tmp_path, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        return r

    cli("_full", "job", "--vtk=off", "--exact")
    cli("_split", "job", "--vtk=off", "--exact", "--stop-at=150", "--checkpoint=job.ckpt")
    r = cli("_split", "job", "--vtk=off", "--exact", "--restart=job.ckpt")
    assert "at step 150" in r.stdout
    for ext in (".energy", ".verify"):
        full = open(job + "_full" + ext).read()
        split = open(job + "_split" + ext).read()
        assert len(full.splitlines()) == prob.steps // 30 + 1
        assert split == full, ext
    # a checkpoint of another problem is refused with exit code 1
    other, _ = load_golden("abc_r10_nu025")
    files.write_job(other, str(tmp_path / "other"), exact_text=True)
    r = subprocess.run([EXE, "_x", "other", "--vtk=off", "--exact", "--restart=job.ckpt"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=240)
    assert r.returncode == 1 and "another problem" in r.stderr, r.stderr[-2000:]


def test_cli_vtk_async_and_binary(tmp_path):
    """Snapshots written on background threads equal the inline ones byte for byte; --vtk=binary
    holds the same points and fields unrounded (the ASCII values within their 12 digits)."""
    print("cli vtk", flush=True)
    prob, ref = load_golden("abc_r10_nu025")
    prob.steps = 120
    job = str(tmp_path / "job")
    files.write_job(prob, job, exact_text=True)

    def cli(*opts):
        r = subprocess.run([EXE, *opts], cwd=tmp_path, capture_output=True, text=True, timeout=240)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]

    # --exact: the same kernels in every process (the default autotunes per process)
    cli("_sync", "job", "--exact", "--vtk-threads=0")
    cli("_async", "job", "--exact", "--vtk-threads=4")
    cli("_bin", "job", "--exact", "--vtk=binary")
    names = ["-0.vtk"] + [f"{k}-{s}.vtk" for s in range(30, 121, 30) for k in ("", "-vel", "-theta")]
    for nm in names:
        a = open(f"{job}_sync{nm}", "rb").read()
        assert a == open(f"{job}_async{nm}", "rb").read(), nm
        asc = files.read_vtk_field(f"{job}_sync{nm}")
        b = files.read_vtk_field(f"{job}_bin{nm}")
        assert np.array_equal(asc["points"], b["points"])
        key = "vectors" if "vectors" in b else "scalars"
        assert np.all(np.abs(asc[key] - b[key]) <= 5e-12 * np.abs(b[key]) + 1e-300), nm
