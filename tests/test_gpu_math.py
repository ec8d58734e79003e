"""The fast correctly rounded FP32 primitives of the solver kernels equal the
general IEEE sequences on this device: reciprocal and square root over every
float of their ranges, division over 2^28 pseudo-random operand pairs."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def test_fast_primitives_match_ieee(gpu_lib):
    out = (C.c_uint64 * 6)()
    from pbf_amd import _lib
    _lib.check(gpu_lib.pbf_selftest_math(1 << 28, 12345, out))
    rcp_bad, sqrt_bad, div_bad, n_rcp, n_sqrt, n_div = list(out)
    print(f"reciprocal {rcp_bad}/{n_rcp}, sqrt {sqrt_bad}/{n_sqrt}, division {div_bad}/{n_div}", flush=True)
    assert n_rcp == 240 << 23 and n_sqrt == 104 << 23 and n_div == 1 << 28
    assert (rcp_bad, sqrt_bad, div_bad) == (0, 0, 0)
