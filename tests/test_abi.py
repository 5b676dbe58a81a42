"""The C-ABI library: it loads, exports every entry point include/lps_amd.h declares, and fails
loudly (no silent fallback) where no GPU is present."""
import re

import pytest

from lps_amd import preprocess as pp, solver


def declared_functions():
    text = open(solver.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lps_[a-z_0-9]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(solver.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = solver.load_library()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_create_without_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(solver.LpsError) as e:
        solver.Solver(pp.reference_test_problem(R=3.0, steps=30))
    assert e.value.code == solver.LPS_ERR_HIP


def test_cli_usage_and_missing_input_exit_codes(tmp_path):
    """PD2_LPS_ABC keeps the reference's exit codes (src/main.cpp:154-160, 178-185): 0 with the
    usage message, 1 with a FATAL message when a job file is missing. Neither touches the GPU."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(solver.LIB_PATH), "PD2_LPS_ABC")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage" in r.stdout
    r = subprocess.run([exe, "_out", "nojob"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "FATAL: Cannot read nojob.data1" in r.stderr


def test_lstsq_refuses_a_nan_rcond_before_touching_the_device():
    """lps_lstsq_batched / lps_abc_operators refuse rcond = NaN at the ABI (LPS_ERR_ARG): the
    check runs before any HIP call, so it holds without a GPU too."""
    import numpy as np
    from lps_amd import abc_gpu
    A = np.eye(3)[None]
    B = np.ones((1, 3, 1))
    with pytest.raises(solver.LpsError) as e:
        abc_gpu.lstsq(A, B, float("nan"))
    assert e.value.code == solver.LPS_ERR_ARG and "rcond" in str(e.value)
