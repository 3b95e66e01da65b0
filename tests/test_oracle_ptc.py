"""Pins of the CPU oracle's parallel-transport convolution (the learned
preconditioner's layer) by what the mathematics fixes: finite differences
for the weight gradients, the adjoint identity, gauge equivariance, and the
identity / pure-shift special cases.  Each is chosen so that a plausible
mistake (a transposed or un-daggered link, a swapped path direction, a
missing conjugate in the gradient, a lost boundary sign) fails it."""
import numpy as np
import pytest

import oracle as orc
import qkxtm_mpi_amd as q

L = (2, 4, 2, 4)
V = int(np.prod(L))
rng = np.random.default_rng(5)


def crand(*shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def su3_field(n, seed):
    """n random SU(3) matrices (QR of a complex Gaussian, phase-fixed)."""
    r = np.random.default_rng(seed)
    out = np.zeros((n, 3, 3), complex)
    for i in range(n):
        A = r.standard_normal((3, 3)) + 1j * r.standard_normal((3, 3))
        Q, R = np.linalg.qr(A)
        Q = Q * (np.diag(R) / np.abs(np.diag(R)))
        out[i] = Q / np.linalg.det(Q) ** (1 / 3)
    return out


def shift(iv, mu, d):
    x = [iv % L[0], (iv // L[0]) % L[1], (iv // (L[0] * L[1])) % L[2], iv // (L[0] * L[1] * L[2])]
    x[mu] = (x[mu] + d) % L[mu]
    return ((x[3] * L[2] + x[2]) * L[1] + x[1]) * L[0] + x[0]


@pytest.fixture(scope="module")
def U():
    return orc.apply_bc(q.gen_gauge(L, 0.6, 9), L, -1)


def test_identity_path():
    W = np.zeros((9, 3, 3), complex)
    W[0] = np.eye(3)
    f = crand(V, 3, 3)
    assert np.allclose(orc.ptc(f, q.gen_gauge(L, 0.6, 9), L, W), f, rtol=0, atol=1e-15)


@pytest.mark.parametrize("mu", range(4))
def test_unit_links_give_pure_shifts(mu):
    """U = 1: forward path mu reads f(x+mu), backward path f(x-mu)."""
    U1 = np.zeros((V, 4, 3, 3), complex)
    U1[:, :] = np.eye(3)
    f = crand(V, 2, 3)
    Wf = np.zeros((9, 2, 2), complex)
    Wf[1 + 2 * mu] = np.eye(2)
    Wb = np.zeros((9, 2, 2), complex)
    Wb[2 + 2 * mu] = np.eye(2)
    iv = np.arange(V)
    assert np.allclose(orc.ptc(f, U1, L, Wf), f[shift(iv, mu, 1)], atol=1e-15)
    assert np.allclose(orc.ptc(f, U1, L, Wb), f[shift(iv, mu, -1)], atol=1e-15)


def test_adjoint_identity(U):
    """<a, P b> = <P^dag a, b> for a PTC layer P with cin != cout."""
    W = crand(9, 5, 3)
    b, a = crand(V, 3, 3), crand(V, 5, 3)
    lhs = np.vdot(a, orc.ptc(b, U, L, W))
    rhs = np.vdot(orc.ptc_adj(a, U, L, W), b)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_gauge_equivariance(U):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag, f(x) -> g(x) f(x) maps the layer
    output to g(x) out(x), boundary sign included."""
    g = su3_field(V, 3)
    Ug = np.empty_like(U)
    for iv in range(V):
        for mu in range(4):
            Ug[iv, mu] = g[iv] @ U[iv, mu] @ g[shift(iv, mu, 1)].conj().T
    W = crand(9, 4, 2)
    f = crand(V, 2, 3)
    fg = np.einsum("xab,xcb->xca", g, f)
    out = orc.ptc(f, U, L, W)
    outg = orc.ptc(fg, Ug, L, W)
    assert np.allclose(outg, np.einsum("xab,xcb->xca", g, out), atol=1e-12)


def test_network_gradient_finite_differences(U):
    """dL/dRe W + i dL/dIm W of L = |M r - e|^2 against central differences."""
    ch = 3
    W1, W2 = 0.3 * crand(9, ch, 4), 0.3 * crand(9, 4, ch)
    r, e = crand(V, 4, 3), crand(V, 4, 3)
    loss, G1, G2 = orc.net_loss_grad(r, e, U, L, W1, W2)
    assert np.isclose(loss, np.sum(np.abs(orc.net_apply(r, U, L, W1, W2) - e) ** 2), rtol=1e-13)
    h = 1e-6
    picks = [(0, (0, 0, 0)), (0, (3, 2, 1)), (0, (8, 1, 3)), (1, (0, 0, 0)), (1, (5, 3, 2)), (1, (2, 1, 1))]
    for layer, idx in picks:
        for part, unit in ((0, 1.0), (1, 1j)):
            Wp = [W1.copy(), W2.copy()]
            Wm = [W1.copy(), W2.copy()]
            Wp[layer][idx] += h * unit
            Wm[layer][idx] -= h * unit
            lp = orc.net_loss_grad(r, e, U, L, *Wp)[0]
            lm = orc.net_loss_grad(r, e, U, L, *Wm)[0]
            fd = (lp - lm) / (2 * h)
            G = (G1, G2)[layer][idx]
            an = G.real if part == 0 else G.imag
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (layer, idx, part, fd, an)
