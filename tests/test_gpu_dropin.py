"""GPU: the reference-driver counterparts, built on the C++98 drop-in classes,
against the CPU oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu
HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qkxtm_mpi_amd", "host")


def test_driver_test_invert(q):
    """bin/test_invert counterpart: CG through qkxTM::Vector / CG on config-1."""
    print("driver test_invert", flush=True)
    r = subprocess.run([os.path.join(HOST, "test_invert")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    norm = float(re.search(r"the norm of vector is (\S+)", r.stdout).group(1))
    res = float(re.search(r"true residual (\S+)", r.stdout).group(1))
    iters = int(re.search(r"iterations (\d+)", r.stdout).group(1))
    L = (8, 8, 8, 16)
    U = orc.apply_bc(q.gen_gauge(L, 0.4, 1), L, -1)
    src = q.gen_z2(L, 2)
    b = orc.tm(src, U, L, 0.156, 0.01, 1, orc.UKQCD, orc.TM_DAG)
    x, k, _ = orc.cg(b, U, L, 0.156, 0.01, 1, 1e-10, 5000)
    assert abs(iters - k) <= 2
    assert abs(norm - orc.norm2(x)) <= 1e-8 * orc.norm2(x)
    assert res <= 1e-9


def test_driver_test_applyTM(q, tmp_path):
    """bin/test_applyTM counterpart: tm_xxDagTm_xx(OO) through qkxTM::Vector."""
    print("driver test_applyTM", flush=True)
    dump = tmp_path / "dump.txt"
    r = subprocess.run([os.path.join(HOST, "test_applyTM"), str(dump)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a = np.loadtxt(dump)
    L = (8, 8, 8, 8)
    V = int(np.prod(L))
    vin = (a[:, 4] + 1j * a[:, 5]).reshape(V, 4, 3)
    vout = (a[:, 6] + 1j * a[:, 7]).reshape(V, 4, 3)
    assert np.array_equal(vin, q.gen_z2(L, 12))
    U = q.gen_gauge(L, 0.4, 11)       # boundaryT = +1 in this driver
    ref = orc.tm_xx(np.ascontiguousarray(vin), U, L, 0.156, 0.1, 1, orc.UKQCD, orc.OO, 1)
    assert rel(vout, ref) <= 1e-13
    # Vector::normalize after the output was written
    assert abs(float(re.search(r"norm of the normalized vector is (\S+)", r.stdout).group(1)) - 1) <= 1e-14


def test_driver_test_gauss_smearing(q, tmp_path):
    """bin/test_gauss_smearing counterpart: plaquette, APE smearing, Gaussian
    smearing of a point source through qkxTM::Gauge / qkxTM::Vector, against
    the oracle's chain on the same links."""
    print("driver test_gauss_smearing", flush=True)
    dump = tmp_path / "slice.txt"
    lime = tmp_path / "conf.lime"
    r = subprocess.run([os.path.join(HOST, "test_gauss_smearing"), str(dump), str(lime)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    plaq = float(re.search(r"plaquette is (\S+)", r.stdout).group(1))
    plaq_ape = float(re.search(r"plaquette smeared is (\S+)", r.stdout).group(1))
    norm = float(re.search(r"the norm of vector is (\S+)", r.stdout).group(1))
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.4, 1)
    A = orc.ape_smear(U, L, 20, 0.5)
    pt = np.zeros((int(np.prod(L)), 4, 3), complex)
    pt[0, 0, 0] = 1.0
    ref = orc.gauss_smear(pt, A, L, 50, 4.0)
    assert abs(plaq - orc.plaquette(U, L)) <= 1e-14
    assert abs(plaq_ape - orc.plaquette(A, L)) <= 1e-12
    assert abs(norm - orc.norm2(ref)) <= 1e-10 * orc.norm2(ref)
    a = np.loadtxt(dump)
    got = (a[:, 5] + 1j * a[:, 6]).reshape(-1, 4, 3)
    assert rel(got, ref[:L[0] * L[1] * L[2]]) <= 1e-10
    assert np.array_equal(q.read_lime(lime, L), U)   # the file the driver wrote and read back
    # the reference alternates between the vector and vec_tmp: after an even
    # step count vec_tmp holds the result, after an odd one the iterate before
    # the last step (lib/lattice_util.cpp:1097-1143)
    assert abs(float(re.search(r"the norm of vec_tmp is (\S+)", r.stdout).group(1)) - orc.norm2(ref)) <= \
        1e-10 * orc.norm2(ref)
    for n in (1, 7):
        r = subprocess.run([os.path.join(HOST, "test_gauss_smearing"), str(dump), str(lime), str(n)],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        prev = orc.gauss_smear(pt, A, L, n - 1, 4.0) if n > 1 else pt
        last = orc.gauss_smear(pt, A, L, n, 4.0)
        got_tmp = float(re.search(r"the norm of vec_tmp is (\S+)", r.stdout).group(1))
        got_vec = float(re.search(r"the norm of vector is (\S+)", r.stdout).group(1))
        assert abs(got_tmp - orc.norm2(prev)) <= 1e-10 * orc.norm2(prev), n
        assert abs(got_vec - orc.norm2(last)) <= 1e-10 * orc.norm2(last), n


def test_driver_test_prop(q, tmp_path):
    """bin/test_seq_source propagator counterpart (reference CG per column
    through qkxTM::Vector / CG / Propagator) against the library's batched
    path on the same inputs: point sources, smeared on their time slice,
    solve_batch, rescale, rotation."""
    print("driver test_prop", flush=True)
    dump = tmp_path / "site0.txt"
    r = subprocess.run([os.path.join(HOST, "test_prop"), str(dump)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    norms = [float(re.search(rf"{k} norm (\S+)", r.stdout).group(1)) for k in ("up", "down")]
    a = np.loadtxt(dump)
    got = (a[:, 5] + 1j * a[:, 6]).reshape(2, 4, 4, 3, 3)
    L = (8, 8, 8, 16)
    kappa, mu = 0.156, 0.01
    U = q.gen_gauge(L, 0.4, 1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    ape = q.Gauge(L, boundaryT=-1)
    ape.APE_smearing(q.Gauge(L, boundaryT=-1, host=U), 5, 0.5)
    for quark, sign in enumerate((+1, -1)):
        b = q.Spinor(L, nrhs=12)
        b.Gaussian_smearing(q.point_sources(L, (0, 0, 0, 0)), ape, 10, 4.0, t=0)
        P = q.Propagator(L)
        st = q.solve_batch(P, b, u, q.params(L, kappa, mu, sign, -1), eo=True, inner_prec=q.HALF, delta=0.05,
                           tol=1e-12)
        assert all(s.converged for s in st)
        P.rescale(2 * kappa)
        if sign > 0:
            P.rotateToPhysicalBasePlus()
        else:
            P.rotateToPhysicalBaseMinus()
        M = P.download_prop()
        assert abs(norms[quark] - P.norm2()) <= 1e-8 * norms[quark], quark
        assert rel(got[quark], M[0]) <= 1e-8, quark
