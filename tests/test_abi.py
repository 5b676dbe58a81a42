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
