"""GPU solvers: the reference CG through the C ABI against the oracle CG, and
the true-residual solvers (full / even-odd, FP64 / FP32 / FP16 / BF16 inner
with FP64 reliable updates) against the FP64 reference solution."""
import os

import numpy as np
import pytest

import oracle as orc
from golden_util import probes, summarize
from refops import rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# config-1: 8^3x16, eps 0.4, kappa 0.156, mu 0.01, antiperiodic t, Z2 source
L1 = (8, 8, 8, 16)
K1, M1 = 0.156, 0.01


@pytest.fixture(scope="module")
def cfg1(q):
    U = q.gen_gauge(L1, 0.4, 1)
    src = q.gen_z2(L1, 2)
    Ub = orc.apply_bc(U, L1, -1)
    return U, Ub, src


def test_reference_cg_matches_oracle(q, cfg1):
    """CG::operator() semantics: D^dag D x = D^dag src, iterated-residual stop."""
    U, Ub, src = cfg1
    b = orc.tm(src, Ub, L1, K1, M1, 1, orc.UKQCD, orc.TM_DAG)
    x_ref, k_ref, _ = orc.cg(b, Ub, L1, K1, M1, 1, 1e-10, 5000)
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1, tol=1e-10, maxIter=5000)
    x, bd = q.Spinor(L1), q.Spinor(L1, host=b)
    st = q.CG(p, u)(x, bd)
    print("iterations gpu", st.iterations, "oracle", k_ref, flush=True)
    assert st.converged and abs(st.iterations - k_ref) <= 2
    assert rel(x.download(), x_ref) <= 1e-9


def test_reference_cg_maxiter_status(q, cfg1):
    U, Ub, src = cfg1
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1, tol=1e-12, maxIter=7)
    x, bd = q.Spinor(L1), q.Spinor(L1, host=src)
    st = q.CG(p, u)(x, bd)
    assert st.iterations == 7 and not st.converged


@pytest.mark.parametrize("eo,prec", [(False, "DOUBLE"), (True, "DOUBLE"), (True, "SINGLE"), (True, "HALF"),
                                     (True, "BF16"), (False, "SINGLE")])
def test_solve_true_residual_and_reference_solution(q, cfg1, eo, prec):
    """True residual <= 1e-10 (recomputed by the oracle in FP64) and the
    solution within 1e-9 of the reference CG solution at tol 1e-14."""
    U, Ub, src = cfg1
    g = np.load(os.path.join(GOLD, "cg_cfg1.npz"), allow_pickle=False)
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1)
    x, b = q.Spinor(L1), q.Spinor(L1, host=src)
    st = q.solve(x, b, u, p, eo=eo, inner_prec=getattr(q, prec), tol=1e-10, maxIter=20000)
    xs = x.download()
    res = rel(orc.tm(xs, Ub, L1, K1, M1, 1), src)
    print(f"eo={eo} prec={prec} iters={st.iterations} rel_updates={st.reliable_updates} res={res:.2e}",
          flush=True)
    assert st.converged and res <= 1e-10 and abs(st.true_residual - res) <= 1e-3 * res + 1e-14
    R, sites = probes(L1)
    s = summarize(xs, R, sites)
    n = np.sqrt(float(g["norm2_cg_x"]))
    assert abs(np.sqrt(s["norm2"]) - n) <= 1e-9 * n
    assert rel(s["sites"], g["site_cg_x"]) <= 1e-9


def test_solve_nonfinite_source(q):
    L = (4, 4, 4, 4)
    u = q.Gauge(L, host=q.gen_gauge(L, 0.3, 3))
    src = q.gen_z2(L, 4)
    src[5, 1, 2] = np.nan
    x, b = q.Spinor(L), q.Spinor(L, host=src)
    with pytest.raises(q.QkxError):
        q.solve(x, b, u, q.params(L, 0.15, 0.01), tol=1e-10)


def test_solve_zero_source(q):
    L = (4, 4, 4, 4)
    u = q.Gauge(L, host=q.gen_gauge(L, 0.3, 3))
    x, b = q.Spinor(L, host=q.gen_z2(L, 1)), q.Spinor(L)
    st = q.solve(x, b, u, q.params(L, 0.15, 0.01), tol=1e-10)
    assert st.converged and not np.any(x.download())


def test_half_spinor_tiny_sites_store_as_zero(q):
    """A site whose largest component is an FP32 denormal is stored as zero in
    FP16 (its per-site scale would overflow on inversion), not as NaN."""
    L = (4, 4, 4, 4)
    psi = np.zeros((256, 4, 3), complex)
    psi[0] = 1.0
    psi[5] = 1e-39
    psi[9, 2, 1] = 3e-38
    h = q.Spinor(L, q.HALF, host=psi)
    out = h.download()
    assert np.all(np.isfinite(out))
    assert not np.any(out[5])
    assert np.allclose(out[0], 1.0, rtol=1e-3) and abs(out[9, 2, 1] - 3e-38) <= 1e-3 * 3e-38


def test_half_inner_point_source(q):
    """FP16 inner solve from a point source: the residual far from the source
    falls below the smallest normal float without breaking the recurrence."""
    L = (16, 16, 16, 32)
    U = q.gen_gauge(L, 0.3, 1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    src = q.gen_point(L, q.gen_site(L, 5), 0, 0)
    x, b = q.Spinor(L), q.Spinor(L, host=src)
    st = q.solve(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.01, tol=1e-10, maxIter=20000)
    res = rel(orc.tm(x.download(), orc.apply_bc(U, L, -1), L, 0.1632, 0.004, 1), src)
    print(f"fp16 point source: iterations {st.iterations}, residual {res:.2e}", flush=True)
    assert st.converged and res <= 1e-10


@pytest.mark.parametrize("prec", ["DOUBLE", "SINGLE", "HALF", "BF16"])
def test_solve_with_separate_update_kernel(q, cfg1, prec):
    """The unfused CG (separate update kernel and Ap field) reaches the same
    solution as the default fused one (checked at FP64 and FP16; at FP32 and
    BF16 it stores Ap at p's precision and only converges alike); moving
    x += alpha p from the fused hop into the direction kernel (mode 2) gives
    the same bits and iterations at every inner precision."""
    U, Ub, src = cfg1
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1)
    b = q.Spinor(L1, host=src)
    out, its = [], []
    try:
        for fuse in (1, 0, 2):
            q.set_cg_fuse(fuse)
            x = q.Spinor(L1)
            st = q.solve(x, b, u, p, eo=True, inner_prec=getattr(q, prec), delta=0.05, tol=1e-10)
            assert st.converged
            out.append(x.download())
            its.append((st.iterations, st.reliable_updates))
            res = rel(orc.tm(out[-1], Ub, L1, K1, M1, 1), src)
            assert res <= 1e-10
    finally:
        q.set_cg_fuse(2)
    if prec in ("DOUBLE", "HALF"):
        assert rel(out[0], out[1]) <= 1e-9
    assert its[2] == its[0] and np.array_equal(out[2], out[0])
    x = q.Spinor(L1)
    try:
        q.set_cg_fuse(False)
        bn = q.Spinor(L1, host=orc.tm(src, Ub, L1, K1, M1, 1, orc.UKQCD, orc.TM_DAG))
        st = q.CG(q.params(L1, K1, M1, 1, -1, tol=1e-10, maxIter=5000), u)(x, bn)
        assert st.converged
    finally:
        q.set_cg_fuse(2)


def test_graph_replay_identical_to_eager(q):
    """CG iterations replayed from captured HIP graphs give the same solution,
    bit for bit, as the same iterations launched one by one (reference CG and
    the mixed-precision even-odd solver, two solves in a row each)."""
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.4, 1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    src = q.gen_z2(L, 2)
    p = q.params(L, 0.156, 0.01, 1, -1)
    b = q.Spinor(L, host=src)
    out = {}
    try:
        for on in (False, True):
            q.set_cg_graph(3 if on else 0)
            for rep in range(2):
                print(f"graph={on} rep={rep}", flush=True)
                x = q.Spinor(L)
                st = q.solve(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-10)
                y = q.Spinor(L)
                cg = q.CG(p, u)
                sc = cg(y, b)
                out[(on, rep)] = (x.download(), st.iterations, st.reliable_updates, y.download(), sc.iterations)
    finally:
        q.set_cg_graph(0)
    ref = out[(False, 0)]
    for key, v in out.items():
        assert np.array_equal(v[0], ref[0]) and v[1:3] == ref[1:3], key
        assert np.array_equal(v[3], ref[3]) and v[4] == ref[4], key


def test_fp64_operator_copy_only_for_su3_links(q):
    """The solvers' 12-real copies (the FP64 check operator, the FP32/FP16
    inner fields) rebuild row 2 as conj(row0 x row1), so they carry a field
    only while its links are SU(3): to 1e-14 for the FP64 copy, 1e-6 below.
    Links that carry the antiperiodic sign (boundaryT = +1) are copied with
    the last time slice's t-links negated and boundaryT = -1, the same
    operator; APE-smeared links (SU(3) to about 1e-10) keep the 18-real FP64
    operator; links that are not SU(3) at all are refused for mixed precision
    and solved in FP64 on their 18 reals.  Every path reaches the same
    solution."""
    L, kappa, mu = (8, 8, 8, 16), 0.1632, 0.004
    U = q.gen_gauge(L, 0.3, 3)
    u = q.Gauge(L, boundaryT=-1, host=U)
    assert u.su3_deviation() <= 1e-14
    us = q.Gauge(L, boundaryT=+1, host=orc.apply_bc(U, L, -1))
    assert us.su3_deviation() > 1.0
    ape = q.Gauge(L, boundaryT=-1)
    ape.APE_smearing(u, 5, 0.5)
    assert 1e-14 < ape.su3_deviation() < 1e-6
    b = q.gen_z2(L, 4)

    def solve(g, bt, prec):
        x = q.Spinor(L)
        print(f"solve boundaryT={bt} prec={prec} deviation={g.su3_deviation():.2e}", flush=True)
        st = q.solve(x, q.Spinor(L, host=b), g, q.params(L, kappa, mu, 1, bt), eo=True, inner_prec=prec,
                     delta=0.01, tol=1e-10)
        assert st.converged and st.true_residual <= 1e-10
        return x.download()

    ref = solve(u, -1, q.DOUBLE)
    for g, bt, prec in ((us, +1, q.DOUBLE), (u, -1, q.HALF), (us, +1, q.HALF), (us, +1, q.SINGLE)):
        assert rel(solve(g, bt, prec), ref) <= 1e-8, (bt, prec)
    solve(ape, -1, q.HALF)
    # links that are not SU(3): refused for mixed precision, solved in FP64
    wide = q.Gauge(L, boundaryT=-1, host=1.01 * U)
    assert wide.su3_deviation() > 1e-3
    with pytest.raises(q.QkxError, match="not SU\\(3\\)"):
        q.solve(q.Spinor(L), q.Spinor(L, host=b), wide, q.params(L, kappa, mu, 1, -1), eo=True,
                inner_prec=q.HALF, delta=0.01, tol=1e-10)
    solve(wide, -1, q.DOUBLE)
