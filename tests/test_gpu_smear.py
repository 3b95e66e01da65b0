"""GPU source preparation against the oracle and the reference's golden
vectors: SU(3) projection, APE smearing, plaquette, Gaussian smearing (single
fields, batches, one time slice), the drivers' chain of smeared source ->
solve -> smeared sink, and the refresh of the solvers' cached low-precision
links after the links change."""
import os

import numpy as np
import pytest

import oracle as orc
from golden_util import gauge_probes, generic_matrices, probes, summarize, summarize_gauge
from refops import rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_projection_matches_oracle_and_reference(q):
    g = _load("proj_p4444.npz")
    L = tuple(int(v) for v in g["L"])
    M = generic_matrices(L, int(g["seed"]))
    print(f"projection L={L}", flush=True)
    u = q.Gauge(L, boundaryT=1, host=M)
    u.projectSU3()
    got = u.download()
    ref = orc.project_su3_field(M, L)
    assert np.max(np.abs(got - ref)) <= 1e-12
    assert np.array_equal(got[:, 3], M[:, 3])
    GR, gsites = gauge_probes(L)
    gsites = np.union1d(gsites, [0, 1])
    s = summarize_gauge(got, GR, gsites)
    assert rel(s["sites"], g["gsite_proj"]) <= 1e-12
    assert np.max(np.abs(s["proj"] - g["gproj_proj"])) <= 1e-10


@pytest.mark.parametrize("case", ["smear_s4648.npz", "smear_s6448.npz"])
def test_ape_gauss_plaquette_match_oracle_and_reference(q, case):
    g = _load(case)
    L = tuple(int(v) for v in g["L"])
    eps, bt, sg, sp, na, aa, ng, ag = g["params"]
    bt, na, ng = int(bt), int(na), int(ng)
    U = q.gen_gauge(L, eps, int(sg))
    psi = q.gen_gaussian(L, int(sp))
    Ub = orc.apply_bc(U, L, bt)
    print(f"smearing {case} L={L} nAPE={na} nGauss={ng}", flush=True)
    u = q.Gauge(L, boundaryT=bt, host=U)
    assert abs(u.calculatePlaquette() - float(g["plaq_in"][0])) <= 1e-13
    ape = q.Gauge(L, boundaryT=bt)
    ape.APE_smearing(u, na, float(aa))
    A = ape.download()
    Ao = orc.ape_smear(U, L, na, float(aa))     # stored links: no boundary sign
    assert np.max(np.abs(A - Ao)) <= 1e-12
    assert np.array_equal(A[:, 3], U[:, 3])
    assert np.array_equal(u.download(), U)      # the input is not modified
    assert abs(ape.calculatePlaquette() - float(g["plaq_ape"][0])) <= 1e-13
    GR, gsites = gauge_probes(L)
    s = summarize_gauge(orc.apply_bc(A, L, bt), GR, gsites)
    assert rel(s["sites"], g["gsite_ape"]) <= 1e-12
    R, sites = probes(L)
    for name, links in (("gauss", u), ("gauss_ape", ape)):
        out = q.Spinor(L)
        out.Gaussian_smearing(q.Spinor(L, host=psi), links, ng, float(ag))
        y = out.download()
        ref = orc.gauss_smear(psi, Ub if name == "gauss" else orc.apply_bc(Ao, L, bt), L, ng, float(ag))
        assert rel(y, ref) <= 1e-13, name
        s = summarize(y, R, sites)
        assert rel(s["sites"], g["site_" + name]) <= 1e-12, name


def test_gauss_batch_slice_and_noop(q):
    """A batch smears every member as a single field would (bit for bit); a
    point source smeared on its time slice only equals the whole-lattice
    smearing; nsmear = 0 leaves the output untouched."""
    L = (6, 4, 8, 6)
    U = q.gen_gauge(L, 0.3, 71)
    u = q.Gauge(L, boundaryT=-1, host=U)
    V = int(np.prod(L))
    srcs = np.stack([q.gen_gaussian(L, 72 + r) for r in range(3)])
    print("gauss batch/slice", flush=True)
    b = q.Spinor(L, host=srcs, nrhs=3)
    ob = q.Spinor(L, nrhs=3)
    ob.Gaussian_smearing(b, u, 5, 4.0)
    got = ob.download()
    for r in range(3):
        o1 = q.Spinor(L)
        o1.Gaussian_smearing(q.Spinor(L, host=srcs[r]), u, 5, 4.0)
        assert np.array_equal(got[r], o1.download()), r
        assert rel(got[r], orc.gauss_smear(srcs[r], orc.apply_bc(U, L, -1), L, 5, 4.0)) <= 1e-13
    # point source on time slice 4: slice-only smearing == whole-lattice smearing
    x0 = (1, 2, 3, 4)
    pt = q.gen_point(L, x0, 2, 1)
    full, sl = q.Spinor(L), q.Spinor(L, host=np.ones((V, 4, 3), complex))
    full.Gaussian_smearing(q.Spinor(L, host=pt), u, 6, 4.0)
    sl.Gaussian_smearing(q.Spinor(L, host=pt), u, 6, 4.0, t=x0[3])
    assert np.array_equal(full.download(), sl.download())
    assert np.count_nonzero(full.download().reshape(L[3], -1)[np.arange(L[3]) != x0[3]]) == 0
    keep = q.Spinor(L, host=srcs[0])
    keep.Gaussian_smearing(q.Spinor(L, host=srcs[1]), u, 0, 4.0)
    assert np.array_equal(keep.download(), srcs[0])
    with pytest.raises(q.QkxError):
        keep.Gaussian_smearing(keep, u, 2, 4.0)


def test_gauss_batch_of_12_whole_lattice_and_slice(q):
    """The 12-field batch of a propagator's sink (whole lattice) and of its
    point sources (one time slice): every member against the oracle, and the
    slice form equal to the whole-lattice one on fields that live on it."""
    L = (4, 6, 4, 8)
    U = q.gen_gauge(L, 0.3, 81)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    V = int(np.prod(L))
    print("gauss batch of 12", flush=True)
    srcs = np.stack([q.gen_gaussian(L, 90 + r) for r in range(12)])
    ob = q.Spinor(L, nrhs=12)
    ob.Gaussian_smearing(q.Spinor(L, host=srcs, nrhs=12), u, 7, 2.5)
    got = ob.download()
    for r in range(12):
        assert rel(got[r], orc.gauss_smear(srcs[r], Ub, L, 7, 2.5)) <= 1e-13, r
    t0 = 5
    on_slice = srcs.copy()
    on_slice[:, np.arange(V) // (V // L[3]) != t0] = 0
    full, sl = q.Spinor(L, nrhs=12), q.Spinor(L, nrhs=12)
    full.Gaussian_smearing(q.Spinor(L, host=on_slice, nrhs=12), u, 7, 2.5)
    sl.Gaussian_smearing(q.Spinor(L, host=on_slice, nrhs=12), u, 7, 2.5, t=t0)
    assert np.array_equal(full.download(), sl.download())


def test_smeared_source_solve_smeared_sink(q):
    """The drivers' chain (bin/test_invert.cpp:77-117): APE-smear the links,
    Gaussian-smear a point source with them, solve D x = b on the unsmeared
    links, Gaussian-smear the solution; every stage against the oracle, the
    later ones on the GPU's own input to that stage.

    20 APE steps: the reference's closed-form projection is ill-conditioned
    where M^dag M is nearly degenerate, as it becomes on a smooth field.  On
    this field one link (relative eigenvalue gaps 7e-5, 8e-4) moves by 1.3e-10
    when the oracle's input is perturbed by 1e-16 relative noise, while 99.9%
    of the entries move by less than 7e-12: that bounds any two roundings of
    the algorithm, the reference's included."""
    L = (8, 8, 8, 16)
    kappa, mu = 0.156, 0.01
    U = q.gen_gauge(L, 0.4, 1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    ape = q.Gauge(L, boundaryT=-1)
    print("chain: APE 20 x 0.5, Gauss 50 x 4.0", flush=True)
    ape.APE_smearing(u, 20, 0.5)
    Ao = orc.ape_smear(U, L, 20, 0.5)
    A = ape.download()
    d = np.abs(A - Ao)
    assert d.max() <= 1e-9 and np.quantile(d, 0.999) <= 3e-11
    assert abs(ape.calculatePlaquette() - orc.plaquette(Ao, L)) <= 1e-12
    Ab = orc.apply_bc(A, L, -1)
    pt = q.gen_point(L, (0, 0, 0, 0), 0, 0)
    b = q.Spinor(L)
    b.Gaussian_smearing(q.Spinor(L, host=pt), ape, 50, 4.0, t=0)
    src = orc.gauss_smear(pt, Ab, L, 50, 4.0)
    assert rel(b.download(), src) <= 1e-12
    x = q.Spinor(L)
    st = q.solve(x, b, u, q.params(L, kappa, mu, 1, -1), eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-12)
    xs = x.download()
    assert st.converged and rel(orc.tm(xs, Ub, L, kappa, mu, 1), src) <= 1e-12
    sink = q.Spinor(L)
    sink.Gaussian_smearing(x, ape, 50, 4.0)
    assert rel(sink.download(), orc.gauss_smear(xs, Ab, L, 50, 4.0)) <= 1e-12


def test_cached_inner_links_follow_link_changes(q):
    """A mixed-precision solve caches FP16/FP32 copies of the links; after the
    links change (APE smearing into the field, or a new upload) the next solve
    uses the new links, not the cached old ones."""
    L = (8, 8, 8, 8)
    kappa, mu = 0.15, 0.05
    U1, U2 = q.gen_gauge(L, 0.4, 81), q.gen_gauge(L, 0.4, 82)
    src = q.gen_z2(L, 83)
    p = q.params(L, kappa, mu, 1, -1)
    u = q.Gauge(L, boundaryT=-1, host=U1)
    b = q.Spinor(L, host=src)
    print("cache refresh", flush=True)
    for links in ("first", "upload", "ape"):
        if links == "upload":
            u.upload(U2)
            cur = U2
        elif links == "ape":
            tmp = q.Gauge(L, boundaryT=-1, host=U1)
            u.APE_smearing(tmp, 2, 0.5)
            cur = u.download()
        else:
            cur = U1
        x = q.Spinor(L)
        st = q.solve(x, b, u, p, eo=True, inner_prec=q.HALF, delta=0.05, tol=1e-10)
        res = rel(orc.tm(x.download(), orc.apply_bc(cur, L, -1), L, kappa, mu, 1), src)
        print(f"  {links}: iterations {st.iterations} residual {res:.2e}", flush=True)
        assert st.converged and res <= 1e-10, links
        # the same solve on a fresh field holding these links: identical run
        xf = q.Spinor(L)
        stf = q.solve(xf, b, q.Gauge(L, boundaryT=-1, host=cur), p, eo=True, inner_prec=q.HALF, delta=0.05,
                      tol=1e-10)
        assert stf.iterations == st.iterations and np.array_equal(xf.download(), x.download()), links
