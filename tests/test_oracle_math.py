"""Pins of the CPU oracle by what the mathematics fixes, independent of the
reference's code: each check is chosen so that a plausible mistake (a wrong
projector sign, a swapped neighbour, a transposed link, a lost boundary sign
or twist sign) fails it."""
import numpy as np
import pytest

import oracle as orc
import qkxtm_mpi_amd as q
from refops import rel

# UKQCD gamma matrices written out here from the basis definition (rows),
# gamma_1..gamma_4 and gamma_5 = gamma_1 gamma_2 gamma_3 gamma_4.
I = 1j
G_UKQCD = [
    np.array([[0, 0, 0, I], [0, 0, I, 0], [0, -I, 0, 0], [-I, 0, 0, 0]]),
    np.array([[0, 0, 0, 1], [0, 0, -1, 0], [0, -1, 0, 0], [1, 0, 0, 0]]),
    np.array([[0, 0, I, 0], [0, 0, 0, -I], [-I, 0, 0, 0], [0, I, 0, 0]]),
    np.diag([1, 1, -1, -1]).astype(complex),
]
G5_UKQCD = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)


def test_gamma_algebra():
    """Clifford algebra {g_mu, g_nu} = 2 delta, hermiticity, g5 = g1 g2 g3 g4."""
    for m in range(4):
        assert np.allclose(G_UKQCD[m], G_UKQCD[m].conj().T)
        for n in range(4):
            ac = G_UKQCD[m] @ G_UKQCD[n] + G_UKQCD[n] @ G_UKQCD[m]
            assert np.allclose(ac, 2 * np.eye(4) * (m == n))
    assert np.allclose(G_UKQCD[0] @ G_UKQCD[1] @ G_UKQCD[2] @ G_UKQCD[3], G5_UKQCD)


def coords(L):
    V = int(np.prod(L))
    iv = np.arange(V)
    x = iv % L[0]
    y = (iv // L[0]) % L[1]
    z = (iv // (L[0] * L[1])) % L[2]
    t = iv // (L[0] * L[1] * L[2])
    return np.stack([x, y, z, t], 1)


@pytest.mark.parametrize("bt", [-1, +1])
def test_free_field_plane_wave(bt):
    """Unit gauge: D e^{ipx} u = [1 - 2k sum cos p + 2ik sum g_mu sin p + i a g5] e^{ipx} u
    with p_t = (2n+1) pi / Lt for the antiperiodic boundary."""
    L = (4, 6, 4, 8)
    kappa, mu, s = 0.13, 0.07, -1
    U = np.zeros((int(np.prod(L)), 4, 3, 3), complex)
    U[:] = np.eye(3)
    Ub = orc.apply_bc(U, L, bt)
    n = (1, 2, 1, 3)
    pv = np.array([2 * np.pi * n[m] / L[m] for m in range(3)] +
                  [(2 * n[3] + (1 if bt == -1 else 0)) * np.pi / L[3]])
    rng = np.random.default_rng(3)
    u = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    phase = np.exp(1j * coords(L) @ pv)
    psi = phase[:, None, None] * u[None]
    a = 2 * kappa * mu * s
    M = (1 - 2 * kappa * np.cos(pv).sum()) * np.eye(4) + 2j * kappa * sum(
        G_UKQCD[m] * np.sin(pv[m]) for m in range(4)) + 1j * a * G5_UKQCD
    expect = phase[:, None, None] * (M @ u)[None]
    assert rel(orc.tm(psi, Ub, L, kappa, mu, s), expect) < 1e-13
    # H^dag: H^dag = g5 H g5 -> sign of the sin term flips
    Md = (1 - 2 * kappa * np.cos(pv).sum()) * np.eye(4) - 2j * kappa * sum(
        G_UKQCD[m] * np.sin(pv[m]) for m in range(4)) - 1j * a * G5_UKQCD
    assert rel(orc.tm(psi, Ub, L, kappa, mu, s, orc.UKQCD, orc.TM_DAG), phase[:, None, None] * (Md @ u)[None]) < 1e-13


def _fields(L, seed, eps=0.5):
    U = q.gen_gauge(L, eps, seed)
    return U, q.gen_gaussian(L, seed + 1), q.gen_gaussian(L, seed + 2)


def test_adjointness_and_gamma5_hermiticity():
    L, kappa, mu = (4, 4, 6, 4), 0.17, 0.03
    U, psi, phi = _fields(L, 5)
    Ub = orc.apply_bc(U, L, -1)
    Dpsi = orc.tm(psi, Ub, L, kappa, mu, 1)
    Ddphi = orc.tm(phi, Ub, L, kappa, mu, 1, orc.UKQCD, orc.TM_DAG)
    lhs, rhs = np.vdot(phi.ravel(), Dpsi.ravel()), np.vdot(Ddphi.ravel(), psi.ravel())
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    # g5 D(s) g5 = D(-s)^dag  ->  tmDag(s) psi = g5 tm(-s) g5 psi
    for b in (orc.UKQCD, orc.CHIRAL):
        g = orc.gamma5(orc.tm(orc.gamma5(psi, L, b), Ub, L, kappa, mu, -1, b), L, b)
        if b == orc.UKQCD:
            assert rel(orc.tm(psi, Ub, L, kappa, mu, 1, b, orc.TM_DAG), g) < 1e-14
        # the hopping term alone is gamma5-hermitian in both bases
        h = orc.gamma5(orc.hop(orc.gamma5(psi, L, b), Ub, L, b), L, b)
        hd = orc.hop(psi, Ub, L, b, 1) if b == orc.UKQCD else None
        if hd is not None:
            assert rel(hd, h) < 1e-14
        assert abs(np.vdot(phi.ravel(), orc.hop(psi, Ub, L, b).ravel()) -
                   np.vdot(h.ravel() * 0 + orc.gamma5(orc.hop(orc.gamma5(phi, L, b), Ub, L, b), L, b).ravel(),
                           psi.ravel())) < 1e-12 * np.linalg.norm(phi) * np.linalg.norm(psi) * 16


def random_su3(n, rng):
    A = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    Q, R = np.linalg.qr(A)
    d = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (d / abs(d))[:, None, :]
    det = np.linalg.det(Q)
    Q[:, :, 0] /= det[:, None]
    return Q


@pytest.mark.parametrize("basis", [orc.UKQCD, orc.CHIRAL])
def test_gauge_covariance(basis):
    """D[U^g] (g psi) = g D[U] psi with U^g_mu(x) = g(x) U_mu(x) g(x+mu)^dag."""
    L, kappa, mu = (4, 6, 4, 4), 0.15, 0.05
    U, psi, _ = _fields(L, 9)
    Ub = orc.apply_bc(U, L, -1)
    V = int(np.prod(L))
    g = random_su3(V, np.random.default_rng(1))
    c = coords(L)
    Ug = np.empty_like(Ub)
    for m in range(4):
        cp = c.copy()
        cp[:, m] = (cp[:, m] + 1) % L[m]
        ivp = ((cp[:, 3] * L[2] + cp[:, 2]) * L[1] + cp[:, 1]) * L[0] + cp[:, 0]
        Ug[:, m] = g @ Ub[:, m] @ g[ivp].conj().transpose(0, 2, 1)
    gpsi = np.einsum("vab,vsb->vsa", g, psi)
    lhs = orc.tm(np.ascontiguousarray(gpsi), np.ascontiguousarray(Ug), L, kappa, mu, 1, basis)
    rhs = np.einsum("vab,vsb->vsa", g, orc.tm(psi, Ub, L, kappa, mu, 1, basis))
    assert rel(lhs, rhs) < 1e-13


@pytest.mark.parametrize("parity", [orc.EE, orc.OO])
def test_schur_complement_identity(parity):
    """tm_xx(EE) psi = even part of D (psi_e + chi_o) with chi_o chosen so that the
    odd part vanishes: chi_o = kappa A^-1 H_oe psi_e."""
    L, kappa, mu, s = (4, 4, 4, 6), 0.16, 0.02, 1
    U, psi, _ = _fields(L, 13)
    Ub = orc.apply_bc(U, L, -1)
    odd = (coords(L).sum(1) & 1).astype(bool)
    keep = ~odd if parity == orc.EE else odd
    pe = psi * keep[:, None, None]
    hop = orc.hop(pe, Ub, L, orc.UKQCD, 0, orc.OE if parity == orc.EE else orc.EO)
    chi = kappa * orc.twist(hop, L, kappa, mu, s, orc.UKQCD, orc.TWIST_INV)
    full = orc.tm(np.ascontiguousarray(pe + chi), Ub, L, kappa, mu, s)
    assert np.max(np.abs(full[~keep])) < 1e-13 * np.max(np.abs(full))
    got = orc.tm_xx(psi, Ub, L, kappa, mu, s, orc.UKQCD, parity)
    assert rel(got, full * keep[:, None, None]) < 1e-13
    # the normal form is M^dag M: <phi, M^dag M psi> = <M phi, M psi>
    phi = q.gen_gaussian(L, 77) * keep[:, None, None]
    mm = orc.tm_xx(psi, Ub, L, kappa, mu, s, orc.UKQCD, parity, 1)
    m_psi = orc.tm_xx(psi, Ub, L, kappa, mu, s, orc.UKQCD, parity)
    m_phi = orc.tm_xx(np.ascontiguousarray(phi), Ub, L, kappa, mu, s, orc.UKQCD, parity)
    assert abs(np.vdot(phi.ravel(), mm.ravel()) - np.vdot(m_phi.ravel(), m_psi.ravel())) < 1e-11 * np.linalg.norm(
        m_phi) * np.linalg.norm(m_psi)


def test_chiral_basis_is_a_rotation_of_ukqcd():
    """The chiral hopping term is the UKQCD one in the basis of the reference's
    rotateFromChiralToUKQCD (lib/lattice_util.cpp:1834-1864), psi_UKQCD = T psi_chiral.
    The chiral gamma5 of the reference, diag(1,1,-1,-1) (:1800-1820), is
    -T^-1 g5_UKQCD T, so the chiral twisted-mass operator at twistSign s is the
    rotated UKQCD operator at -s."""
    L, kappa, mu, s = (4, 4, 4, 4), 0.14, 0.03, 1
    U, psi, _ = _fields(L, 21)
    Ub = orc.apply_bc(U, L, -1)
    r = 1 / np.sqrt(2)
    T = np.array([[-r, 0, r, 0], [0, -r, 0, r], [r, 0, r, 0], [0, r, 0, r]])
    Ti = np.linalg.inv(T)
    rot = lambda M, v: np.ascontiguousarray(np.einsum("st,vtc->vsc", M, v))  # noqa: E731
    assert np.allclose(Ti @ G5_UKQCD @ T, -np.diag([1, 1, -1, -1]))
    assert rel(orc.hop(psi, Ub, L, orc.CHIRAL), rot(Ti, orc.hop(rot(T, psi), Ub, L, orc.UKQCD))) < 1e-14
    lhs = orc.tm(psi, Ub, L, kappa, mu, s, orc.CHIRAL)
    rhs = rot(Ti, orc.tm(rot(T, psi), Ub, L, kappa, mu, -s, orc.UKQCD))
    assert rel(lhs, rhs) < 1e-13
    for parity in (orc.EE, orc.OO):
        lhs = orc.tm_xx(psi, Ub, L, kappa, mu, s, orc.CHIRAL, parity)
        rhs = rot(Ti, orc.tm_xx(rot(T, psi), Ub, L, kappa, mu, -s, orc.UKQCD, parity))
        assert rel(lhs, rhs) < 1e-13


def test_blas_against_numpy():
    L = (4, 4, 4, 4)
    a, b = q.gen_gaussian(L, 1), q.gen_gaussian(L, 2)
    assert abs(orc.redot(a, b) - np.vdot(a.ravel(), b.ravel()).real) < 1e-12 * np.linalg.norm(a) * np.linalg.norm(b)
    assert abs(orc.norm2(a) - np.vdot(a.ravel(), a.ravel()).real) < 1e-12 * orc.norm2(a)
    assert rel(orc.axpy(a, 0.3, b), a + 0.3 * b) < 1e-15
    assert rel(orc.xpay(a, 0.3, b), 0.3 * a + b) < 1e-15


def test_oracle_cg_solves_normal_equations():
    L, kappa, mu = (4, 4, 4, 8), 0.15, 0.02
    U, psi, _ = _fields(L, 31)
    Ub = orc.apply_bc(U, L, -1)
    x, k, r2 = orc.cg(psi, Ub, L, kappa, mu, 1, 1e-12, 2000)
    ax = orc.tm(x, Ub, L, kappa, mu, 1, orc.UKQCD, orc.TM_DAG_TM)
    assert rel(ax, psi) < 1e-11 and 0 < k < 2000


def test_generator_sources_and_su3():
    L = (4, 4, 4, 8)
    V = int(np.prod(L))
    z = q.gen_z2(L, 5)
    assert np.vdot(z.ravel(), z.ravel()).real == pytest.approx(12 * V, rel=1e-15)
    assert np.allclose(np.abs(z.real), 1 / np.sqrt(2)) and np.allclose(np.abs(z.imag), 1 / np.sqrt(2))
    U = q.gen_gauge(L, 0.3, 7).reshape(-1, 3, 3)
    assert np.allclose(U @ U.conj().transpose(0, 2, 1), np.eye(3), atol=1e-14)
    assert np.allclose(np.linalg.det(U), 1, atol=1e-14)
    U0 = q.gen_gauge(L, 0.0, 7).reshape(-1, 3, 3)
    assert np.array_equal(U0, np.broadcast_to(np.eye(3), U0.shape))
    # thread-order independence: same field regardless of lattice split position
    assert np.array_equal(q.gen_gauge(L, 0.3, 7).reshape(-1, 3, 3), U)
    pt = q.gen_point(L, [1, 2, 3, 4], 2, 1)
    assert pt.sum() == 1 and pt[((4 * 4 + 3) * 4 + 2) * 4 + 1, 2, 1] == 1
