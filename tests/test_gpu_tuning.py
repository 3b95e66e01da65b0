"""Launch shapes of the batched hops: FP32 and BF16 batches that share their
links through LDS agree with per-thread links and with single fields (the
same arithmetic; the compiler may contract a product and a sum into one FMA
differently in the two kernels, so the last FP32 bit, or the BF16 rounding
of a few elements, can differ), and the autotuner (which picks workgroup
size, link sharing and pipeline depth per launch kind by timing them)
changes no output bit of the FP16 and FP64 operators and at most that
last-bit difference of the FP32 one."""
import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu


def _srcs(q, L, n, seed):
    return np.stack([q.gen_gaussian(L, seed + r) for r in range(n)])


@pytest.mark.parametrize("prec,tol,same", [("SINGLE", 2e-6, 1e-6), ("BF16", 2e-2, 2e-3)])
@pytest.mark.parametrize("L,nrhs", [((4, 4, 4, 8), 12), ((8, 6, 4, 8), 3), ((8, 8, 8, 16), 12), ((8, 8, 8, 16), 5)])
def test_fp32_bf16_link_sharing_matches(q, prec, tol, same, L, nrhs):
    """FP32 / BF16 batches of 3-16 RHS on 12-real links: links rebuilt once
    per workgroup into LDS agree with every thread fetching its own (within
    one rounding of the storage format), on the even-odd normal operator and
    the full-lattice tm / tmDag; both meet the oracle within the storage
    precision."""
    print(f"{prec} link sharing L={L} nrhs={nrhs}", flush=True)
    P = getattr(q, prec)
    kappa, mu = 0.1632, 0.004
    U = q.gen_gauge(L, 0.3, 23)
    u = q.Gauge(L, boundaryT=-1, host=U)
    ul = q.Gauge(L, P, q.RECON_12, -1)
    ul.convert_from(u)
    p = q.params(L, kappa, mu, 1, -1)
    src = _srcs(q, L, nrhs, 300)
    be, oe = q.Spinor(L, P, q.EVEN, host=src, nrhs=nrhs), q.Spinor(L, P, q.EVEN, nrhs=nrhs)
    bf, of = q.Spinor(L, P, host=src, nrhs=nrhs), q.Spinor(L, P, nrhs=nrhs)
    got = {}
    try:
        q.set_autotune(0)
        for share in (0, 2):
            q.set_batch_link_sharing(share)
            oe.tm_xxDagTm_xx(be, ul, p, q.EE)
            r1 = oe.download()
            of.tm(bf, ul, p)
            r2 = of.download()
            of.tmDag(bf, ul, p)
            got[share] = (r1, r2, of.download())
    finally:
        q.set_batch_link_sharing(1)
        q.set_autotune(1)
    for k, (a, b) in enumerate(zip(got[0], got[2])):
        assert rel(b, a) <= same, k
    Ub = orc.apply_bc(U, L, -1)
    for r in (0, nrhs - 1):
        xs, ys = q.Spinor(L, P, q.EVEN, host=src[r]), q.Spinor(L, P, q.EVEN)
        ys.tm_xxDagTm_xx(xs, ul, p, q.EE)
        assert rel(got[2][0][r], ys.download()) <= same, r
        assert rel(got[2][1][r], orc.tm(src[r], Ub, L, kappa, mu, 1)) <= tol, r


@pytest.mark.parametrize("prec", ["HALF", "SINGLE", "DOUBLE"])
def test_autotuned_shapes_change_no_bit(q, prec):
    """A batched operator launched with the tuner's shapes equals the same
    operator with the fixed defaults, bit for bit; the tuner reports each
    launch kind it timed."""
    L, nrhs = (16, 16, 16, 16), 12
    P = getattr(q, prec)
    U = q.gen_gauge(L, 0.3, 5)
    u = q.Gauge(L, boundaryT=-1, host=U)
    if P == q.DOUBLE:
        ul = u
    else:
        ul = q.Gauge(L, P, q.RECON_12, -1)
        ul.convert_from(u)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    src = _srcs(q, L, nrhs, 500)
    x, y = q.Spinor(L, P, q.EVEN, host=src, nrhs=nrhs), q.Spinor(L, P, q.EVEN, nrhs=nrhs)
    xf, yf = q.Spinor(L, P, host=src, nrhs=nrhs), q.Spinor(L, P, nrhs=nrhs)
    out = {}
    try:
        for mode in (0, 1):
            q.set_autotune(mode)
            y.tm_xxDagTm_xx(x, ul, p, q.EE)
            yf.tm(xf, ul, p)
            out[mode] = (y.download(), yf.download())
        report = q.autotune_report()
    finally:
        q.set_autotune(1)
    print(report, flush=True)
    if P == q.SINGLE:
        assert rel(out[1][0], out[0][0]) <= 1e-6 and rel(out[1][1], out[0][1]) <= 1e-6
    else:
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert f"nrhs={nrhs}" in report and "block=" in report
