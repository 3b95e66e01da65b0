"""GPU batches of right-hand sides: every operator applied to a batch field
against the oracle applied to each right-hand side, batch results against the
same operator on single fields, and the lockstep batched solver (per-rhs
twisted mass, frozen converged right-hand sides, zero sources) against the
oracle's true residual and the reference CG solution."""
import os

import numpy as np
import pytest

import oracle as orc
from golden_util import probes, summarize
from refops import rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

LS = (4, 4, 4, 8)
KS, MS = 0.153, 0.07


def _srcs(q, L, n, seed=100):
    return np.stack([q.gen_gaussian(L, seed + r) for r in range(n)])


@pytest.fixture(scope="module")
def small(q):
    U = q.gen_gauge(LS, 0.4, 21)
    return U, orc.apply_bc(U, LS, -1)


@pytest.mark.parametrize("nrhs", [2, 5, 16])
def test_batch_operators_match_oracle(q, small, nrhs):
    U, Ub = small
    u = q.Gauge(LS, boundaryT=-1, host=U)
    p = q.params(LS, KS, MS, 1, -1)
    src = _srcs(q, LS, nrhs)
    print(f"batch ops L={LS} nrhs={nrhs}", flush=True)
    b = q.Spinor(LS, host=src, nrhs=nrhs)
    out = q.Spinor(LS, nrhs=nrhs)
    cases = [
        ("tm", lambda: out.tm(b, u, p), lambda s: orc.tm(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.TM)),
        ("tm_chiral", lambda: out.tm(b, u, p, q.CHIRAL), lambda s: orc.tm(s, Ub, LS, KS, MS, 1, orc.CHIRAL, orc.TM)),
        ("tmDag", lambda: out.tmDag(b, u, p), lambda s: orc.tm(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.TM_DAG)),
        ("tmDagtm", lambda: out.tmDagtm(b, u, p), lambda s: orc.tm(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.TM_DAG_TM)),
        ("Wilson", lambda: out.Wilson(b, u, p), lambda s: orc.tm(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.WILSON)),
        ("dslash", lambda: out.applyDslash(b, u), lambda s: orc.hop(s, Ub, LS, orc.UKQCD, 0, orc.ALL)),
        ("tm_xx_EE", lambda: out.tm_xx(b, u, p, q.EE), lambda s: orc.tm_xx(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.EE, 0)),
        ("tm_xxDagTm_xx_OO", lambda: out.tm_xxDagTm_xx(b, u, p, q.OO),
         lambda s: orc.tm_xx(s, Ub, LS, KS, MS, 1, orc.UKQCD, orc.OO, 1)),
        ("twistInv", lambda: out.twist(b, p, q.TWIST_INV), lambda s: orc.twist(s, LS, KS, MS, 1, orc.UKQCD,
                                                                              orc.TWIST_INV)),
    ]
    for name, run, ref in cases:
        run()
        got = out.download()
        for r in range(nrhs):
            e = rel(got[r], ref(src[r]))
            assert e <= 1e-13, f"{name} rhs {r}: rel err {e:.2e}"


@pytest.mark.parametrize("prec", ["SINGLE", "HALF", "BF16"])
def test_batch_normal_op_equals_single_fields(q, small, prec):
    """A batch with one right-hand side per thread and per-thread links gives
    bit-identical results to single fields (each right-hand side sees the
    same arithmetic); the RHS-tiled kernels (2, 3 right-hand sides per
    thread) and the shared-link kernels agree to the storage precision."""
    U, _ = small
    P = getattr(q, prec)
    u = q.Gauge(LS, boundaryT=-1, host=U)
    ul = q.Gauge(LS, P, q.RECON_12, -1)
    ul.convert_from(u)
    p = q.params(LS, KS, MS, 1, -1)
    nrhs = 6
    src = _srcs(q, LS, nrhs, 300)
    b = q.Spinor(LS, P, q.EVEN, host=src, nrhs=nrhs)
    out = q.Spinor(LS, P, q.EVEN, nrhs=nrhs)
    single = []
    for r in range(nrhs):
        bs = q.Spinor(LS, P, q.EVEN, host=src[r])
        os_ = q.Spinor(LS, P, q.EVEN)
        os_.tm_xxDagTm_xx(bs, ul, p, q.EE)
        single.append(os_.download())
    tol = {"SINGLE": 1e-6, "HALF": 3e-3, "BF16": 2e-2}[prec]
    try:
        q.set_batch_link_sharing(0)
        for rt in (1, 2, 3):
            q.set_rhs_tile(rt)
            out.tm_xxDagTm_xx(b, ul, p, q.EE)
            got = out.download()
            for r in range(nrhs):
                if rt == 1:
                    assert np.array_equal(got[r], single[r]), f"{prec} rhs {r}"
                else:
                    assert rel(got[r], single[r]) <= tol, f"{prec} tile {rt} rhs {r}"
        q.set_rhs_tile(0)
        q.set_batch_link_sharing(2)
        out.tm_xxDagTm_xx(b, ul, p, q.EE)
        got = out.download()
        for r in range(nrhs):
            assert rel(got[r], single[r]) <= tol, f"{prec} shared links rhs {r}"
    finally:
        q.set_rhs_tile(0)
        q.set_batch_link_sharing(1)


def test_batch_rejects_mismatched_fields(q, small):
    U, _ = small
    u = q.Gauge(LS, boundaryT=-1, host=U)
    p = q.params(LS, KS, MS, 1, -1)
    a = q.Spinor(LS, nrhs=3)
    c = q.Spinor(LS, nrhs=2)
    with pytest.raises(q.QkxError):
        c.tm(a, u, p)
    with pytest.raises(q.QkxError):
        q.Spinor(LS, nrhs=q.MAX_RHS + 1)


# config-1 lattice: 8^3x16, eps 0.4, kappa 0.156, antiperiodic t
L1 = (8, 8, 8, 16)
K1, M1 = 0.156, 0.01


@pytest.fixture(scope="module")
def cfg1(q):
    U = q.gen_gauge(L1, 0.4, 1)
    return U, orc.apply_bc(U, L1, -1)


@pytest.mark.parametrize("eo,prec", [(True, "DOUBLE"), (False, "DOUBLE"), (True, "SINGLE"), (True, "HALF"),
                                     (True, "BF16")])
def test_solve_batch_per_rhs_mu(q, cfg1, eo, prec):
    """Four right-hand sides, four twisted masses on one gauge field: each
    solution meets 1e-10 against the oracle operator with its own mu; the
    first (config-1 source and mu) also matches the reference CG solution."""
    U, Ub = cfg1
    g = np.load(os.path.join(GOLD, "cg_cfg1.npz"), allow_pickle=False)
    mus = np.array([M1, 0.02, 0.05, 0.1])
    src = np.stack([q.gen_z2(L1, 2)] + [q.gen_z2(L1, 40 + r) for r in range(3)])
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1)
    x, b = q.Spinor(L1, nrhs=4), q.Spinor(L1, host=src, nrhs=4)
    delta = {"DOUBLE": 0.1, "SINGLE": 0.01, "HALF": 0.05, "BF16": 0.1}[prec]
    print(f"solve_batch eo={eo} prec={prec}", flush=True)
    st = q.solve_batch(x, b, u, p, mu=mus, eo=eo, inner_prec=getattr(q, prec), delta=delta, tol=1e-10,
                       maxIter=20000)
    xs = x.download()
    for r in range(4):
        res = rel(orc.tm(xs[r], Ub, L1, K1, mus[r], 1), src[r])
        print(f"  rhs {r} mu={mus[r]} res={res:.2e} st={st[r].true_residual:.2e} it={st[r].iterations}"
              f" rel={st[r].reliable_updates}", flush=True)
        assert st[r].converged and res <= 1e-10
        assert abs(st[r].true_residual - res) <= 1e-3 * res + 1e-14
    R, sites = probes(L1)
    s = summarize(xs[0], R, sites)
    n = np.sqrt(float(g["norm2_cg_x"]))
    assert abs(np.sqrt(s["norm2"]) - n) <= 1e-9 * n
    assert rel(s["sites"], g["site_cg_x"]) <= 1e-9


def test_solve_batch_deterministic_and_zero_source(q, cfg1):
    """A zero right-hand side gives x = 0 without stalling the others, and two
    runs give bit-identical solutions (fixed-order reductions)."""
    U, Ub = cfg1
    src = np.stack([q.gen_z2(L1, 7), np.zeros((int(np.prod(L1)), 4, 3), complex), q.gen_z2(L1, 8)])
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1)
    b = q.Spinor(L1, host=src, nrhs=3)
    outs = []
    for _ in range(2):
        x = q.Spinor(L1, nrhs=3)
        st = q.solve_batch(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-10)
        outs.append(x.download())
        assert all(s.converged for s in st)
    assert np.array_equal(outs[0], outs[1])
    assert not np.any(outs[0][1])
    for r in (0, 2):
        assert rel(orc.tm(outs[0][r], Ub, L1, K1, M1, 1), src[r]) <= 1e-10


def test_solve_batch_propagator_point_sources(q, cfg1):
    """The 12 spin-colour point sources of a propagator in one batch (the
    reference's test_invert_prop loop) at FP16 inner precision."""
    U, Ub = cfg1
    V = int(np.prod(L1))
    src = np.zeros((12, V, 4, 3), complex)
    site = 3 + 8 * (5 + 8 * (2 + 8 * 9))
    for sc in range(12):
        src[sc, site, sc // 3, sc % 3] = 1.0
    u = q.Gauge(L1, boundaryT=-1, host=U)
    p = q.params(L1, K1, M1, 1, -1)
    x, b = q.Spinor(L1, nrhs=12), q.Spinor(L1, host=src, nrhs=12)
    st = q.solve_batch(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-10)
    xs = x.download()
    for sc in range(12):
        res = rel(orc.tm(xs[sc], Ub, L1, K1, M1, 1), src[sc])
        assert st[sc].converged and res <= 1e-10, f"spin-colour {sc}: {res:.2e}"


@pytest.mark.parametrize("L,nrhs", [(LS, 12), (LS, 8), ((8, 6, 4, 8), 9), ((8, 8, 8, 16), 12), ((8, 8, 8, 16), 16),
                                    (LS, 3), ((8, 6, 4, 8), 5), ((8, 8, 8, 16), 7)])
def test_batch_link_sharing_identical(q, L, nrhs):
    """FP16 batches of 3-16 right-hand sides fetch and rebuild each site's
    links once per workgroup, in LDS: the same bits as every (site, RHS)
    thread doing it itself, and as the same operator on single fields.  The
    small lattices leave the last workgroup partly beyond the sites; the
    8^3x16 ones take the XCD-remapped grid."""
    print(f"link sharing L={L} nrhs={nrhs}", flush=True)
    U = q.gen_gauge(L, 0.4, 23)
    u = q.Gauge(L, boundaryT=-1, host=U)
    uh = q.Gauge(L, q.HALF, q.RECON_12, -1)
    uh.convert_from(u)
    p = q.params(L, KS, MS, 1, -1)
    src = _srcs(q, L, nrhs, 300)
    be, oe = q.Spinor(L, q.HALF, q.EVEN, host=src, nrhs=nrhs), q.Spinor(L, q.HALF, q.EVEN, nrhs=nrhs)
    bf, of = q.Spinor(L, q.HALF, host=src, nrhs=nrhs), q.Spinor(L, q.HALF, nrhs=nrhs)
    got = {}
    try:
        for share in (0, 2):
            q.set_batch_link_sharing(share)
            oe.tm_xxDagTm_xx(be, uh, p, q.EE)
            r1 = oe.download()
            of.tm(bf, uh, p)
            r2 = of.download()
            of.tmDag(bf, uh, p)
            got[share] = (r1, r2, of.download())
    finally:
        q.set_batch_link_sharing(1)
    for k, (a, b) in enumerate(zip(got[0], got[2])):
        assert np.array_equal(a, b), k
    for r in (0, nrhs - 1):
        xs, ys = q.Spinor(L, q.HALF, q.EVEN, host=src[r]), q.Spinor(L, q.HALF, q.EVEN)
        ys.tm_xxDagTm_xx(xs, uh, p, q.EE)
        assert np.array_equal(ys.download(), got[2][0][r]), r


def test_batch_with_sign_carrying_links(q):
    """Links that carry the antiperiodic sign (boundaryT = +1, the drop-in
    drivers' links) through qkx_solve_batch on the shared-link fused path:
    its 12-real copies are stored with the last slice's t-links negated and
    boundaryT = -1, which must give the solution of the unsigned links."""
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.3, 7)
    Ub = orc.apply_bc(U, L, -1)
    x0 = q.gen_site(L, 3)
    print(f"batch with sign-carrying links, 12 point sources at {x0}", flush=True)
    b = q.point_sources(L, x0)
    sols = []
    try:
        q.set_batch_link_sharing(2)   # this lattice is below the automatic threshold
        for u, bt in ((q.Gauge(L, boundaryT=-1, host=U), -1), (q.Gauge(L, boundaryT=+1, host=Ub), +1)):
            x = q.Spinor(L, nrhs=12)
            st = q.solve_batch(x, b, u, q.params(L, 0.1632, 0.004, 1, bt), eo=True, inner_prec=q.HALF,
                               delta=0.01, tol=1e-10)
            assert all(s.converged and s.true_residual <= 1e-10 for s in st), bt
            sols.append(x.download())
    finally:
        q.set_batch_link_sharing(1)
    V = int(np.prod(L))
    iv = ((x0[3] * L[2] + x0[2]) * L[1] + x0[1]) * L[0] + x0[0]
    for r in (0, 7, 11):
        pt = np.zeros((V, 4, 3), complex)
        pt[iv, r // 3, r % 3] = 1.0
        assert rel(orc.tm(sols[1][r], Ub, L, 0.1632, 0.004, 1), pt) <= 1e-10, r
        assert rel(sols[1][r], sols[0][r]) <= 1e-8, r


def test_batched_fused_update_matches_unfused(q):
    """Even-odd FP16 batches of 8-16 RHS fuse x += alpha p, r -= alpha Ap
    into the last hop (Ap never stored): both paths reach every column's
    true residual, in the same number of iterations, with solutions that
    agree to the tolerance.  With x += alpha p moved into the direction
    kernel (mode 2) the fused path gives the same bits and iterations."""
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.3, 7)
    u = q.Gauge(L, boundaryT=-1, host=U)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    x0 = q.gen_site(L, 3)
    print(f"fused batch update, 12 point sources at {x0}", flush=True)
    b = q.point_sources(L, x0)
    sols, its = [], []
    try:
        q.set_batch_link_sharing(2)   # this lattice is below the automatic threshold
        for fuse in (0, 1, 2):
            q.set_cg_fuse(fuse)
            x = q.Spinor(L, nrhs=12)
            st = q.solve_batch(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.01, tol=1e-10)
            assert all(s.converged and s.true_residual <= 1e-10 for s in st), fuse
            sols.append(x.download())
            its.append([s.iterations for s in st])
    finally:
        q.set_cg_fuse(2)
        q.set_batch_link_sharing(1)
    assert abs(its[0][0] - its[1][0]) <= 2
    assert its[2] == its[1] and np.array_equal(sols[2], sols[1])
    for r in range(12):
        assert rel(sols[1][r], sols[0][r]) <= 1e-8, r


@pytest.mark.parametrize("mus", [[0.004, 0.006, 0.01, 0.015, 0.02, 0.03, 0.05, 0.08], [0.004, 0.01, 0.02]])
def test_multi_mu_batch_on_the_fused_path(q, mus):
    """Eight (or three) twisted masses on one gauge field: shared links and
    the fused CG update; every member meets its true residual and agrees with
    the single-RHS solve of its own mass."""
    L, kappa = (8, 8, 8, 16), 0.1632
    n = len(mus)
    U = q.gen_gauge(L, 0.3, 9)
    u = q.Gauge(L, boundaryT=-1, host=U)
    src = q.gen_z2(L, 6)
    print(f"multi-mu batch of {n} on the fused path", flush=True)
    b = q.Spinor(L, host=np.stack([src] * n), nrhs=n)
    x = q.Spinor(L, nrhs=n)
    try:
        q.set_batch_link_sharing(2)   # this lattice is below the automatic threshold
        st = q.solve_batch(x, b, u, q.params(L, kappa, mus[0], 1, -1), mu=mus, eo=True, inner_prec=q.HALF,
                           delta=0.01, tol=1e-10)
    finally:
        q.set_batch_link_sharing(1)
    assert all(s.converged and s.true_residual <= 1e-10 for s in st)
    got = x.download()
    for r in (0, n // 2, n - 1):
        xs, bs = q.Spinor(L), q.Spinor(L, host=src)
        s1 = q.solve(xs, bs, u, q.params(L, kappa, mus[r], 1, -1), eo=True, inner_prec=q.HALF, delta=0.01,
                     tol=1e-10)
        assert s1.converged
        assert rel(got[r], xs.download()) <= 1e-8, r


@pytest.mark.parametrize("share", [2, 0])
def test_solves_are_deterministic(q, share):
    """Every reduction sums in a fixed order (no atomics), so the same solve
    run twice gives the same bits: the single-RHS mixed solver, and the
    batched solver on the shared-link fused path (share 2) and on the
    per-thread path (share 0)."""
    L = (8, 8, 8, 16)
    U = q.gen_gauge(L, 0.3, 11)
    u = q.Gauge(L, boundaryT=-1, host=U)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    b1 = q.Spinor(L, host=q.gen_z2(L, 8))
    bb = q.point_sources(L, q.gen_site(L, 2))
    print(f"determinism, link sharing mode {share}", flush=True)
    runs = []
    try:
        q.set_batch_link_sharing(share)
        for _ in range(2):
            x1, xb = q.Spinor(L), q.Spinor(L, nrhs=12)
            s1 = q.solve(x1, b1, u, p, eo=True, inner_prec=q.HALF, delta=0.01, tol=1e-10)
            sb = q.solve_batch(xb, bb, u, p, eo=True, inner_prec=q.HALF, delta=0.01, tol=1e-10)
            runs.append((x1.download(), xb.download(), s1.iterations, [s.iterations for s in sb]))
    finally:
        q.set_batch_link_sharing(1)
    (a1, ab, ia, iba), (c1, cb, ic, ibc) = runs
    assert ia == ic and iba == ibc
    assert np.array_equal(a1, c1) and np.array_equal(ab, cb)


def test_batched_graph_replay_identical_to_eager(q):
    """Batched iterations replayed from captured HIP graphs give the same solutions, bit for bit, and the same
    iteration counts as the same iterations launched one by one: two solves in
    a row each, the plain batch and the learned solver (whose iterations add
    the network's launches)."""
    L, nrhs = (8, 8, 8, 16), 4
    U = q.gen_gauge(L, 0.3, 3)
    u = q.Gauge(L, boundaryT=-1, host=U)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(u)
    p = q.params(L, 0.1632, 0.004, 1, -1)
    b = q.Spinor(L, host=np.stack([q.gen_z2(L, 40 + r) for r in range(nrhs)]), nrhs=nrhs)
    net = q.Net(4, seed=2)      # close to the identity
    out = {}
    try:
        for mode in (0, 2, 0, 2):
            q.set_cg_graph(mode)
            x, y = q.Spinor(L, nrhs=nrhs), q.Spinor(L, nrhs=nrhs)
            sb = q.solve_batch(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-10)
            sl = q.solve_learned(y, b, u, p, net, u32, inner_prec=q.SINGLE, delta=0.1, tol=1e-10, stall=0.5)
            print(f"graph mode {mode}: batch {sb[0].iterations} iterations, learned {sl[0].iterations}", flush=True)
            out.setdefault(mode, []).append((x.download(), [s.iterations for s in sb], y.download(),
                                             [s.iterations for s in sl]))
    finally:
        q.set_cg_graph(0)
    ref = out[0][0]
    for mode, runs in out.items():
        for v in runs:
            assert np.array_equal(v[0], ref[0]) and v[1] == ref[1], mode
            assert np.array_equal(v[2], ref[2]) and v[3] == ref[3], mode
