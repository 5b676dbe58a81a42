"""lps_closure_forward_noalloc (no device allocation, a fixed per-GPU staging area) against
lps_closure_forward, through the C ABI: the same bits on two calls in a row, the second with fewer
rows, and the size limit refused with an error."""
import numpy as np
import pytest

from lps_amd import closure as cl, solver

pytestmark = pytest.mark.gpu


def test_noalloc_forward_equals_forward_on_two_calls():
    print("noalloc forward", flush=True)
    rng = np.random.default_rng(11)
    with cl.Closure(seed=12) as c:
        for n in (2048, 333):  # the second call smaller: stale rows of the first must not leak
            xb = cl.bf16_bits(rng.normal(0.0, 1.0, (n, cl.N_IN)).astype(np.float32))
            want = c.forward(xb)
            got = c.forward_noalloc(xb)
            assert got.shape == want.shape == (n, cl.N_OUT)
            assert np.array_equal(got, want), n
            assert np.all(np.isfinite(got)) and np.any(got != 0)


def test_noalloc_forward_refuses_more_rows_than_its_area():
    print("noalloc limit", flush=True)
    with cl.Closure(seed=13) as c:
        xb = np.zeros((2049, cl.N_IN), dtype=np.uint16)
        with pytest.raises(solver.LpsError, match="2048"):
            c.forward_noalloc(xb)
        assert c.forward_noalloc(xb[:0]).shape == (0, cl.N_OUT)
