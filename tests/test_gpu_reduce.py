"""The solvers' global reductions are compensated (double-double): sums of
per-workgroup partials with large cancellations come out as the exactly
rounded sum (math.fsum), where plain FP64 summation in any order loses the
small terms.  Every consuming workgroup must also get the same bits, since
they all take the CG decisions from their own copy of the sum.  The
reference sums with MPI_Allreduce (lib/blas_qkxTM.cpp:24)."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def dev_sum(q, vals, mode):
    v = np.ascontiguousarray(vals, dtype=np.float64)
    res = ctypes.c_double()
    q._check(q.lib().qkx_reduce_sum(v.ctypes.data_as(ctypes.c_void_p), v.size, mode, ctypes.byref(res)),
             "qkx_reduce_sum")
    return res.value


@pytest.mark.parametrize("n", [8, 1000, 5000, 40000])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_cancelling_partials_sum_exactly(q, mode, n):
    # +2^60, 1, -2^60, 1, ...: every 1 is lost by plain FP64 once it meets 2^60
    v = np.zeros(n)
    v[0::4], v[1::4], v[2::4], v[3::4] = 2.0 ** 60, 1.0, -(2.0 ** 60), 1.0
    exact = math.fsum(v)
    assert float(np.sum(v)) != exact or n <= 4     # plain pairwise summation misses it
    assert dev_sum(q, v, mode) == exact


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ill_conditioned_random_sums(q, mode):
    rng = np.random.default_rng(5)
    for n in (300, 3000, 30000):
        big = rng.standard_normal(n // 2) * 10.0 ** rng.integers(0, 14, n // 2)
        small = rng.standard_normal(n - 2 * (n // 2) + n // 2)
        v = np.concatenate([big, -big, small])
        rng.shuffle(v)
        exact = math.fsum(v)
        got = dev_sum(q, v, mode)
        # condition number sum|v| / |sum v| ~ 1e15: plain FP64 would be off by O(1)
        assert abs(got - exact) <= 2e-16 * abs(exact) + 1e-300, (n, got, exact)
