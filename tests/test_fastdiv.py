"""The multiply-shift division the hop kernels use for site coordinates
(qkx_internal.h: make_fastdiv / fdiv): exact for every index below 2^30 and
every lattice extent a field may have (1..8192), checked by brute force
against integer division at the edges and at random points."""
import random

import pytest


def make_fastdiv(d):
    # the C++ make_fastdiv, step for step
    l = 0
    while (1 << l) < d:
        l += 1
    s = 30 + l
    m = (1 << s) // d + 1
    assert m < (1 << 32)
    return m, s


@pytest.mark.parametrize("lo,hi", [(1, 1025), (1025, 8193)])
def test_fastdiv_exact(lo, hi):
    rng = random.Random(lo)
    for d in range(lo, hi):
        m, s = make_fastdiv(d)
        ns = [0, 1, d - 1, d, d + 1, 2 * d - 1, (1 << 30) - 1, (1 << 30) - d, (1 << 27) - 1]
        ns += [rng.randrange(1 << 30) for _ in range(40)]
        for n in ns:
            if n >= 0:
                assert (n * m) >> s == n // d, (d, n)
