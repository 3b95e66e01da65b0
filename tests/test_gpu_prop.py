"""GPU propagators: layout (upload/download, absorbVector) and the rotations
to the physical basis against the reference's golden vectors bit for bit,
rescale, and the drivers' chain for up and down quarks
(bin/test_seq_source.cpp:81-117, 143-144): smeared point sources -> batched
solve -> rescale(2 kappa) -> rotation, every stage against the oracle."""
import os

import numpy as np
import pytest

import oracle as orc
from golden_util import probes
from refops import rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def columns(psi, phi):
    M = np.zeros(psi.shape[:1] + (4, 4, 3, 3), np.complex128)
    for nu in range(4):
        for c2 in range(3):
            M[:, :, nu, :, c2] = (nu + 1) * psi + (c2 + 1) * 1j * phi
    return M


def test_prop_layout_rotation_rescale_match_reference(q):
    g = np.load(os.path.join(GOLD, "prop_r4448.npz"), allow_pickle=False)
    L = tuple(int(v) for v in g["L"])
    kappa, sp, sf = g["params"]
    psi, phi = q.gen_gaussian(L, int(sp)), q.gen_gaussian(L, int(sf))
    _, sites = probes(L)
    print(f"propagator L={L}", flush=True)
    P = q.Propagator(L)
    P.zero()
    for nu in range(4):
        for c2 in range(3):
            P.absorbVector(q.Spinor(L, host=(nu + 1) * psi + (c2 + 1) * 1j * phi), nu, c2)
    M = P.download_prop()
    assert np.array_equal(M, columns(psi, phi))
    assert np.array_equal(M[sites], g["site_prop_in"])
    assert abs(P.norm2() - float(g["prop_norm2"][0])) <= 1e-13 * float(g["prop_norm2"][0])
    P2 = q.Propagator(L, host=M)
    assert np.array_equal(P2.download_prop(), M)
    P.rotateToPhysicalBasePlus()
    P2.rotateToPhysicalBaseMinus()
    plus, minus = P.download_prop(), P2.download_prop()
    assert np.array_equal(plus[sites], g["site_prop_plus"])
    assert np.array_equal(minus[sites], g["site_prop_minus"])
    assert np.array_equal(plus, orc.prop_rotate(M, +1))
    assert np.array_equal(minus, orc.prop_rotate(M, -1))
    v = q.Spinor(L, host=psi)
    v.rescale(2 * kappa)
    assert np.array_equal(v.download()[sites], g["site_rescale"])
    with pytest.raises(q.QkxError):
        P.absorbVector(q.Spinor(L), 4, 0)


@pytest.mark.parametrize("sign", [+1, -1])
def test_propagator_chain(q, sign):
    """Up (twistSign +1, rotated with +) and down (-1, rotated with -) quark
    propagators from smeared point sources at config-1 (8^3x16)."""
    L = (8, 8, 8, 16)
    kappa, mu = 0.156, 0.01
    U = q.gen_gauge(L, 0.4, 1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    ape = q.Gauge(L, boundaryT=-1)
    ape.APE_smearing(u, 5, 0.5)
    A = orc.apply_bc(ape.download(), L, -1)
    x0 = (1, 2, 3, 4)
    print(f"propagator chain twistSign={sign}", flush=True)
    src = q.point_sources(L, x0)
    b = q.Spinor(L, nrhs=12)
    b.Gaussian_smearing(src, ape, 10, 4.0, t=x0[3])
    bs = b.download()
    P = q.Propagator(L)
    p = q.params(L, kappa, mu, sign, -1)
    st = q.solve_batch(P, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-11)
    cols = P.download()
    V = int(np.prod(L))
    iv = ((x0[3] * L[2] + x0[2]) * L[1] + x0[1]) * L[0] + x0[0]
    for r in range(12):
        pt = np.zeros((V, 4, 3), complex)
        pt[iv, r // 3, r % 3] = 1.0
        assert rel(bs[r], orc.gauss_smear(pt, A, L, 10, 4.0)) <= 1e-13, r
        assert st[r].converged and rel(orc.tm(cols[r], Ub, L, kappa, mu, sign), bs[r]) <= 1e-11, r
    P.rescale(2 * kappa)
    M = P.download_prop()
    assert rel(M, columns_of(cols) * (2 * kappa)) <= 1e-15
    if sign > 0:
        P.rotateToPhysicalBasePlus()
    else:
        P.rotateToPhysicalBaseMinus()
    assert np.array_equal(P.download_prop(), orc.prop_rotate(M, sign))


def columns_of(cols):
    """Propagator [V, 4, 4, 3, 3] from its 12 columns [12, V, 4, 3]."""
    M = np.zeros(cols.shape[1:2] + (4, 4, 3, 3), np.complex128)
    for r in range(12):
        M[:, :, r // 3, :, r % 3] = cols[r]
    return M
