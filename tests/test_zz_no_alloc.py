"""qkx_tm_xx_noalloc through the C ABI against qkx_tm_xx: the same Schur
operator (and normal form) with its intermediate fields in the library's
static scratch buffer instead of its field pool.  Two calls in a row, the
second on a smaller lattice, must each reproduce qkx_tm_xx bit for bit (the
same kernels in the same order); a field whose intermediates exceed the
buffer is refused with a clear error."""
import ctypes

import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu


def noalloc(q, out, inp, u, p, parity, normal, basis=0):
    rc = q.lib().qkx_tm_xx_noalloc(out.h, inp.h, u.h, ctypes.byref(p), basis, parity, normal)
    if rc != 0:
        raise q.QkxError(q.lib().qkx_last_error().decode())


@pytest.mark.parametrize("prec,subset", [("DOUBLE", "FULL"), ("DOUBLE", "EVEN"), ("HALF", "EVEN"),
                                         ("SINGLE", "ODD")])
def test_noalloc_schur_matches_tm_xx(q, prec, subset):
    P, S = getattr(q, prec), getattr(q, subset)
    parity = q.OO if subset == "ODD" else q.EE
    for L, seed in (((16, 16, 16, 32), 3), ((8, 8, 8, 16), 5)):   # the second call is smaller
        print(f"noalloc tm_xx {prec} {subset} L={L}", flush=True)
        kappa, mu = 0.1632, 0.004
        U = q.gen_gauge(L, 0.3, seed)
        u64 = q.Gauge(L, boundaryT=-1, host=U)
        if P == q.DOUBLE:
            u = u64
        else:
            u = q.Gauge(L, P, q.RECON_12, -1)
            u.convert_from(u64)
        p = q.params(L, kappa, mu, 1, -1)
        psi = q.gen_gaussian(L, seed + 1)
        x = q.Spinor(L, P, S, host=psi)
        for normal in (0, 1):
            ref, got = q.Spinor(L, P, S), q.Spinor(L, P, S)
            q._check(q.lib().qkx_tm_xx(ref.h, x.h, u.h, ctypes.byref(p), 0, parity, normal), "qkx_tm_xx")
            noalloc(q, got, x, u, p, parity, normal)
            assert np.array_equal(got.download(), ref.download()), (L, normal)
            if P == q.DOUBLE:
                Ub = orc.apply_bc(U, L, -1)
                oparity = orc.OO if parity == q.OO else orc.EE
                want = orc.tm_xx(psi, Ub, L, kappa, mu, 1, orc.UKQCD, oparity, normal)
                assert rel(got.download(), want) <= 1e-13


def test_noalloc_batch_and_refusal(q):
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.3, 7)
    u = q.Gauge(L, boundaryT=-1, host=U)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    src = np.stack([q.gen_gaussian(L, 40 + r) for r in range(4)])
    x = q.Spinor(L, subset=q.EVEN, host=src, nrhs=4)
    ref, got = q.Spinor(L, subset=q.EVEN, nrhs=4), q.Spinor(L, subset=q.EVEN, nrhs=4)
    ref.tm_xxDagTm_xx(x, u, p, q.EE)
    noalloc(q, got, x, u, p, q.EE, 1)
    assert np.array_equal(got.download(), ref.download())
    # FP64 normal form at 24^3x48 needs 2 x 331776 x 192 bytes, beyond the buffer
    Lb = (24, 24, 24, 48)
    ub = q.Gauge(Lb, boundaryT=-1, host=q.gen_gauge(Lb, 0.3, 1))
    xb, yb = q.Spinor(Lb, subset=q.EVEN), q.Spinor(Lb, subset=q.EVEN)
    with pytest.raises(q.QkxError, match="scratch buffer"):
        noalloc(q, yb, xb, ub, q.params(Lb, 0.1632, 0.004, 1, -1), q.EE, 1)
