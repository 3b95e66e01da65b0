"""The oracle's propagator rotation against golden vectors written by the
reference's own Propagator code (tests/golden/make_golden.py prop), and
against what the mathematics fixes: with R = 1 + s i g5 and g5^2 = 1,
R(-s) R(s) = 2, so the minus rotation undoes the plus one; the plus rotation
of a propagator equals (1 + i g5) S (1 + i g5) / 2 formed as dense 4x4 spin
products; and g5 commutes with R(s), so rotating g5 S g5 gives g5 (rotated
S) g5."""
import os

import numpy as np

import oracle as orc
import qkxtm_mpi_amd as q
from golden_util import input_hash, probes

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G5 = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)


def columns(psi, phi):
    M = np.zeros(psi.shape[:1] + (4, 4, 3, 3), np.complex128)
    for nu in range(4):
        for c2 in range(3):
            M[:, :, nu, :, c2] = (nu + 1) * psi + (c2 + 1) * 1j * phi
    return M


def test_oracle_rotation_matches_reference_bit_for_bit():
    g = np.load(os.path.join(GOLD, "prop_r4448.npz"), allow_pickle=False)
    L = tuple(int(v) for v in g["L"])
    kappa, sp, sf = g["params"]
    psi, phi = q.gen_gaussian(L, int(sp)), q.gen_gaussian(L, int(sf))
    assert str(g["input_hash"]) == input_hash(psi, phi)
    _, sites = probes(L)
    M = columns(psi, phi)
    assert np.array_equal(M[sites], g["site_prop_in"])
    assert abs(np.vdot(M.ravel(), M.ravel()).real - float(g["prop_norm2"][0])) <= 1e-13 * float(g["prop_norm2"][0])
    P, Mi = orc.prop_rotate(M, +1), orc.prop_rotate(M, -1)
    assert np.array_equal(P[sites], g["site_prop_plus"])
    assert np.array_equal(Mi[sites], g["site_prop_minus"])
    assert abs(np.vdot(P.ravel(), P.ravel()).real - float(g["norm2_prop_plus"])) <= 1e-13 * float(g["norm2_prop_plus"])
    assert np.array_equal((psi * (2 * kappa))[sites], g["site_rescale"])


def test_rotation_algebra():
    L = (4, 4, 2, 2)
    rng = np.random.default_rng(3)
    V = int(np.prod(L))
    M = rng.standard_normal((V, 4, 4, 3, 3)) + 1j * rng.standard_normal((V, 4, 4, 3, 3))
    P = orc.prop_rotate(M, +1)
    R = np.eye(4) + 1j * G5
    dense = 0.5 * np.einsum("ab,xbcij,cd->xadij", R, M, R)
    assert np.max(np.abs(P - dense)) <= 1e-14
    # the minus rotation undoes the plus one: R(-1) R(+1) = 2
    assert np.max(np.abs(orc.prop_rotate(P, -1) - M)) <= 1e-14
    # g5 commutes with R(s): rotating g5 S g5 gives g5 (rotated S) g5
    g5M = np.einsum("ab,xbcij,cd->xadij", G5, M, G5)
    assert np.max(np.abs(orc.prop_rotate(g5M, +1) - np.einsum("ab,xbcij,cd->xadij", G5, P, G5))) <= 1e-14
