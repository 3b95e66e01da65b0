"""The oracle's source preparation (SU(3) projection, APE and Gaussian
smearing, plaquette) against golden vectors written by the reference's own
code (tests/golden/make_golden.py smear), and against what the mathematics
fixes: the projection is the unitary polar factor with the determinant phase
removed, smearing is gauge covariant, the unit field is a fixed point of APE,
and on the unit field Gaussian smearing is a known Fourier multiplier."""
import glob
import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as orc
import qkxtm_mpi_amd as q
from golden_util import (gauge_probes, generic_matrices, input_hash, probes, summarize, summarize_gauge)
from refops import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(path):
    return np.load(path, allow_pickle=False)


def polar_su3(M):
    P, _ = sl.polar(M)
    return P * np.exp(-1j * np.angle(np.linalg.det(M)) / 3)


def random_su3(rng, n):
    Z = rng.standard_normal((n, 3, 3)) + 1j * rng.standard_normal((n, 3, 3))
    Q, R = np.linalg.qr(Z)
    Q = Q * (np.diagonal(R, axis1=1, axis2=2) / np.abs(np.diagonal(R, axis1=1, axis2=2)))[:, None, :]
    return Q / np.linalg.det(Q)[:, None, None] ** (1 / 3)


def shift(L, mu, d):
    """Site index of x + d*mu for every lexicographic site x."""
    V = int(np.prod(L))
    c = np.array(np.unravel_index(np.arange(V), tuple(reversed(L)))).T[:, ::-1].copy()   # x, y, z, t
    c[:, mu] = (c[:, mu] + d) % L[mu]
    return ((c[:, 3] * L[2] + c[:, 2]) * L[1] + c[:, 1]) * L[0] + c[:, 0]


def gauge_transform(U, g, L):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dag."""
    out = np.empty_like(U)
    for mu in range(4):
        gp = g[shift(L, mu, +1)]
        out[:, mu] = g @ U[:, mu] @ np.conj(np.transpose(gp, (0, 2, 1)))
    return out


# ---- against the reference's own outputs --------------------------------
@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "smear_*.npz"))))
def test_oracle_smearing_matches_reference(path):
    g = _load(path)
    L = tuple(int(v) for v in g["L"])
    eps, bt, sg, sp, na, aa, ng, ag = g["params"]
    U = q.gen_gauge(L, eps, int(sg))
    psi = q.gen_gaussian(L, int(sp))
    assert str(g["input_hash"]) == input_hash(U, psi), "generator output drifted"
    Ub = orc.apply_bc(U, L, int(bt))
    assert abs(orc.plaquette(Ub, L) - float(g["plaq_in"][0])) <= 1e-14
    ape = orc.ape_smear(Ub, L, int(na), float(aa))
    assert abs(orc.plaquette(ape, L) - float(g["plaq_ape"][0])) <= 1e-14
    GR, gsites = gauge_probes(L)
    s = summarize_gauge(ape, GR, gsites)
    n2 = float(g["gnorm2_ape"])
    assert abs(s["norm2"] - n2) <= 1e-13 * n2
    # a probe sums V*36 products: rounding of order eps * sqrt(norm2 * terms)
    assert np.max(np.abs(s["proj"] - g["gproj_ape"])) <= 1e-15 * np.sqrt(n2 * ape.size)
    assert rel(s["sites"], g["gsite_ape"]) <= 1e-14
    R, sites = probes(L)
    for name, links in (("gauss", Ub), ("gauss_ape", ape)):
        y = orc.gauss_smear(psi, links, L, int(ng), float(ag))
        s = summarize(y, R, sites)
        n2 = float(g["norm2_" + name])
        assert abs(s["norm2"] - n2) <= 1e-13 * n2, name
        assert np.max(np.abs(s["proj"] - g["proj_" + name])) <= 1e-15 * np.sqrt(n2 * y.size), name
        assert rel(s["sites"], g["site_" + name]) <= 1e-14, name


def test_oracle_projection_matches_reference():
    g = _load(os.path.join(GOLD, "proj_p4444.npz"))
    L = tuple(int(v) for v in g["L"])
    M = generic_matrices(L, int(g["seed"]))
    assert str(g["input_hash"]) == input_hash(M)
    P = orc.project_su3_field(M, L)
    GR, gsites = gauge_probes(L)
    gsites = np.union1d(gsites, [0, 1])
    assert np.array_equal(gsites, g["gsites"])
    s = summarize_gauge(P, GR, gsites)
    n2 = float(g["gnorm2_proj"])
    assert abs(s["norm2"] - n2) <= 1e-13 * n2
    assert np.max(np.abs(s["proj"] - g["gproj_proj"])) <= 1e-15 * np.sqrt(n2 * P.size)
    assert rel(s["sites"], g["gsite_proj"]) <= 1e-14
    # temporal links pass through the projection untouched
    assert np.array_equal(P[:, 3], M[:, 3])


# ---- what the mathematics fixes ------------------------------------------
def test_projection_is_polar_factor_with_unit_determinant():
    rng = np.random.default_rng(3)
    for t in range(300):
        M = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
        if t % 3 == 1:   # near SU(3), as the APE sums are after smoothing
            M = 4 * random_su3(rng, 1)[0] + 0.05 * M
        U = orc.su3_project(M)
        assert np.max(np.abs(U - polar_su3(M))) <= 1e-10
        assert np.max(np.abs(U.conj().T @ U - np.eye(3))) <= 1e-13
        assert abs(np.linalg.det(U) - 1) <= 1e-13


def test_projection_diagonal_shortcut():
    """H = M^dag M diagonal: row i scaled by 1/|column i| and no phase removal
    (lib/blas.cpp:99-123)."""
    U = orc.su3_project(np.diag([2.0, 3.0j, 1.0]))
    assert np.allclose(U, np.diag([1.0, 1.0j, 1.0]), atol=1e-15)
    assert np.allclose(orc.su3_project(2.5 * np.eye(3)), np.eye(3), atol=1e-15)
    W = random_su3(np.random.default_rng(5), 1)[0]
    assert np.max(np.abs(orc.su3_project(1.7 * W) - W)) <= 1e-13   # generic SU(3): fixed point


def test_ape_fixed_point_and_temporal_links():
    L = (4, 4, 4, 4)
    V = int(np.prod(L))
    one = np.tile(np.eye(3, dtype=np.complex128), (V, 4, 1, 1))
    assert np.max(np.abs(orc.ape_smear(one, L, 3, 0.5) - one)) <= 1e-14
    U = q.gen_gauge(L, 0.4, 7)
    A = orc.ape_smear(U, L, 2, 0.5)
    assert np.array_equal(A[:, 3], U[:, 3])
    assert np.array_equal(orc.ape_smear(U, L, 0, 0.5), U)
    # smearing raises the plaquette of a rough field
    assert orc.plaquette(A, L) > orc.plaquette(U, L) + 0.1


def test_smearing_is_gauge_covariant():
    L = (4, 4, 6, 4)
    V = int(np.prod(L))
    rng = np.random.default_rng(9)
    U = q.gen_gauge(L, 0.4, 17)
    g = random_su3(rng, V)
    Ug = gauge_transform(U, g, L)
    A, Ag = orc.ape_smear(U, L, 2, 0.5), orc.ape_smear(Ug, L, 2, 0.5)
    assert np.max(np.abs(Ag - gauge_transform(A, g, L))) <= 1e-12
    assert abs(orc.plaquette(Ug, L) - orc.plaquette(U, L)) <= 1e-13
    psi = q.gen_gaussian(L, 19)
    s, sg = orc.gauss_smear(psi, U, L, 3, 4.0), orc.gauss_smear(np.einsum("xab,xsb->xsa", g, psi), Ug, L, 3, 4.0)
    assert rel(sg, np.einsum("xab,xsb->xsa", g, s)) <= 1e-13


def test_gauss_smearing_unit_field_fourier_multiplier():
    """On U = 1, one step multiplies momentum mode p by
    (1 + 2 alpha sum_{i<3} cos p_i) / (1 + 6 alpha); the time direction is untouched."""
    L = (4, 6, 8, 4)
    V = int(np.prod(L))
    one = np.tile(np.eye(3, dtype=np.complex128), (V, 4, 1, 1))
    psi = q.gen_gaussian(L, 23)
    n, a = 3, 2.5
    got = orc.gauss_smear(psi, one, L, n, a)
    f = psi.reshape(L[3], L[2], L[1], L[0], 12)
    F = np.fft.fftn(f, axes=(0, 1, 2, 3))
    kt, kz, ky, kx = np.meshgrid(*[2 * np.pi * np.fft.fftfreq(L[m]) for m in (3, 2, 1, 0)], indexing="ij")
    mult = ((1 + 2 * a * (np.cos(kx) + np.cos(ky) + np.cos(kz))) / (1 + 6 * a)) ** n
    ref = np.fft.ifftn(F * mult[..., None], axes=(0, 1, 2, 3)).reshape(V, 4, 3)
    assert rel(got, ref) <= 1e-13
