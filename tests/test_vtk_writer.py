"""The CLI's VTK writers (host/job_io.cpp): the BINARY encoding holds the same points and fields
as the reference's ASCII one (src/helpers/utils.h:94-224), unrounded, with the same exclusion of
points beyond the radius. Built with g++ from the sources; no GPU."""
import os
import subprocess

import numpy as np
import pytest
from conftest import PKG

from lps_amd import files


@pytest.fixture(scope="module")
def writer(tmp_path_factory):
    d = tmp_path_factory.mktemp("vtk")
    exe = str(d / "vtk_writer")
    src = os.path.join(os.path.dirname(__file__), "vtk_writer_main.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(PKG, "host"), src,
                    os.path.join(PKG, "host", "job_io.cpp"), "-o", exe], check=True)
    return exe, d


def test_binary_equals_ascii(writer):
    exe, d = writer
    for fmt in ("ascii", "binary"):
        subprocess.run([exe, str(d / fmt), fmt], check=True)
    n = 300
    i = np.arange(n)
    nodes = np.stack([0.5 * (i % 20) - 5.0, 0.5 * (i // 20) - 3.5], axis=1)
    keep = np.sqrt((nodes ** 2).sum(1)) <= 5.2
    u = np.stack([np.sin(0.1 * i) * 1e-3, np.cos(0.37 * i) / 3.0], axis=1)
    th = np.exp(-0.01 * i) - 0.5
    a = files.read_vtk_field(str(d / "ascii-vec.vtk"))
    b = files.read_vtk_field(str(d / "binary-vec.vtk"))
    assert np.array_equal(b["points"], nodes[keep]) and np.array_equal(a["points"], nodes[keep])
    # binary: the doubles themselves (numpy's sin/cos/exp may differ from the C library's in the
    # last bit); ASCII: 12 significant digits, relative rounding up to 5e-12
    assert np.allclose(b["vectors"], u[keep], rtol=1e-15, atol=0)
    assert np.all(np.abs(a["vectors"] - b["vectors"]) <= 5e-12 * np.abs(b["vectors"]))
    sa = files.read_vtk_field(str(d / "ascii-sca.vtk"))
    sb = files.read_vtk_field(str(d / "binary-sca.vtk"))
    assert np.allclose(sb["scalars"], th[keep], rtol=1e-15, atol=1e-15)
    assert np.all(np.abs(sa["scalars"] - sb["scalars"]) <= 5e-12 * np.abs(sb["scalars"]))
    head = open(d / "binary-vec.vtk", "rb").read(200)
    assert b"BINARY\nDATASET POLYDATA\nPOINTS " in head
