"""GPU parity: every reference operator through the C ABI against the CPU
oracle on the same seeded inputs.  FP64 bar: relative L2 <= 1e-13."""
import numpy as np
import pytest

import oracle as orc
from refops import ALL_OPS, gpu_op, oracle_op, rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-13


def _setup(q, L, eps, seed, bt):
    U = q.gen_gauge(L, eps, seed)
    psi = q.gen_gaussian(L, seed + 1)
    phi = q.gen_gaussian(L, seed + 2)
    return U, psi, phi


@pytest.mark.parametrize("L,eps,kappa,mu,ts,bt", [
    ((4, 6, 4, 8), 0.4, 0.156, 0.01, +1, -1),
    ((8, 4, 6, 4), 0.3, 0.1632, 0.004, -1, +1),
    ((8, 8, 8, 16), 0.4, 0.156, 0.01, +1, -1),
])
def test_all_ops_fp64_in_kernel_bc(q, L, eps, kappa, mu, ts, bt):
    """Gauge uploaded without the sign; the kernel applies boundaryT at the t wrap."""
    print(f"case L={L} bt={bt} ts={ts}", flush=True)
    U, psi, phi = _setup(q, L, eps, 5, bt)
    Ub = orc.apply_bc(U, L, bt)
    u = q.Gauge(L, q.DOUBLE, q.RECON_18, boundaryT=bt, host=U)
    p = q.params(L, kappa, mu, ts, bt)
    ps, ph = q.Spinor(L, host=psi), q.Spinor(L, host=phi)
    for name in ALL_OPS:
        ref = oracle_op(name, Ub, psi, phi, L, kappa, mu, ts)
        got = gpu_op(q, name, u, ps, ph, p)
        assert rel(got, ref) <= TOL64, (name, rel(got, ref))


def test_ops_with_signed_links_match(q):
    """Links that already carry the boundary sign (reference drivers) with
    boundaryT=+1 in the kernel give the same operator."""
    L, kappa, mu, ts = (4, 4, 6, 8), 0.156, 0.01, 1
    U, psi, phi = _setup(q, L, 0.4, 9, -1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=+1, host=Ub)
    p = q.params(L, kappa, mu, ts, -1)
    ps, ph = q.Spinor(L, host=psi), q.Spinor(L, host=phi)
    for name in ("tm", "tmDag", "tm_chiral", "tm_xx_EE", "tm_xxDagTm_xx_OO", "applyDslash_xx_OE"):
        ref = oracle_op(name, Ub, psi, phi, L, kappa, mu, ts)
        assert rel(gpu_op(q, name, u, ps, ph, p), ref) <= TOL64, name


def test_recon12_fp64(q):
    L, kappa, mu = (4, 6, 4, 8), 0.1632, 0.004
    U, psi, phi = _setup(q, L, 0.3, 13, -1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, q.DOUBLE, q.RECON_12, -1, host=U)
    p = q.params(L, kappa, mu, 1, -1)
    ps, ph = q.Spinor(L, host=psi), q.Spinor(L, host=phi)
    for name in ("tm", "tmDag", "tm_xxDagTm_xx_EE"):
        ref = oracle_op(name, Ub, psi, phi, L, kappa, mu, 1)
        assert rel(gpu_op(q, name, u, ps, ph, p), ref) <= 1e-13, name


def test_parity_subset_fields(q):
    """tm_xx on even-only fields equals the full-field version."""
    L, kappa, mu = (8, 4, 4, 8), 0.16, 0.02
    U, psi, phi = _setup(q, L, 0.4, 17, -1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    p = q.params(L, kappa, mu, 1, -1)
    for parity, sub in ((q.EE, q.EVEN), (q.OO, q.ODD)):
        for normal in (0, 1):
            ref = orc.tm_xx(psi, Ub, L, kappa, mu, 1, orc.UKQCD, parity, normal)
            x = q.Spinor(L, subset=sub, host=psi)
            y = q.Spinor(L, subset=sub)
            (y.tm_xxDagTm_xx if normal else y.tm_xx)(x, u, p, parity)
            assert rel(y.download(), ref) <= TOL64


@pytest.mark.parametrize("L", [(4, 6, 4, 8), (12, 6, 4, 10), (16, 8, 8, 16)])
def test_blocked_site_order_identical(q, L):
    """The blocked walk (y blocks, t blocks, then z) changes which thread takes
    which site, not what a site computes: every operator, the 16-bit Schur
    operator and a batch give the same bits as the checkerboard-index order,
    and the FP64 results still meet the oracle."""
    print(f"case L={L}", flush=True)
    kappa, mu = 0.1632, 0.004
    U, psi, phi = _setup(q, L, 0.3, 31, -1)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    uh = q.Gauge(L, q.HALF, q.RECON_12, -1)
    uh.convert_from(u)
    p = q.params(L, kappa, mu, 1, -1)
    ps, ph = q.Spinor(L, host=psi), q.Spinor(L, host=phi)
    xh, yh = q.Spinor(L, q.HALF, q.EVEN, host=psi), q.Spinor(L, q.HALF, q.EVEN)
    srcs = np.stack([psi, phi, psi - 2 * phi])
    xb, yb = q.Spinor(L, subset=q.EVEN, host=srcs, nrhs=3), q.Spinor(L, subset=q.EVEN, nrhs=3)
    names = ("tm", "tmDag", "tm_chiral", "tm_xx_EE", "tm_xxDagTm_xx_OO", "applyDslash_xx_OE", "applyDslash")
    got = {}
    try:
        for order in (0, 1, 2):
            q.set_site_order(order)
            out = [gpu_op(q, n, u, ps, ph, p) for n in names]
            yh.tm_xxDagTm_xx(xh, uh, p, q.EE)
            yb.tm_xxDagTm_xx(xb, u, p, q.EE)
            got[order] = out + [yh.download(), yb.download()]
    finally:
        q.set_site_order(2)
    for order in (1, 2):
        for k, (a, b) in enumerate(zip(got[0], got[order])):
            assert np.array_equal(a, b), (order, k)
    for n, g in zip(names, got[1]):
        assert rel(g, oracle_op(n, Ub, psi, phi, L, kappa, mu, 1)) <= TOL64, n


@pytest.mark.parametrize("L", [(8, 8, 4, 8), (8, 80, 4, 4), (128, 4, 4, 4), (6, 10, 6, 6)])
def test_fp16_neighbour_staging_identical(q, L):
    """Single-RHS FP16 hops on one parity with the x and y neighbours staged
    in LDS, per workgroup or per wave, give the same bits as the per-thread
    gathers: the hop into either
    parity of a full field, the even-odd twisted operator and its normal
    operator (both hop directions) on parity fields, and the FP16 mixed-precision
    solve (iterations and solution).  The lattices put y-boundary neighbours
    inside and outside a workgroup's staged range (an xy slice of 320 sites),
    stage the largest x row (64 sites) and leave the last workgroup partly
    empty (540 sites of a parity)."""
    print(f"case L={L}", flush=True)
    kappa, mu = 0.1632, 0.004
    U, psi, _ = _setup(q, L, 0.3, 41, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    uh = q.Gauge(L, q.HALF, q.RECON_12, -1)
    uh.convert_from(u)
    p = q.params(L, kappa, mu, 1, -1)
    b = q.Spinor(L, host=psi)
    xf = q.Spinor(L, q.HALF, host=psi)
    runs = {}
    try:
        for on in (0, 1, 2):
            q.set_neighbour_staging(on)
            out = []
            for par in (q.EO, q.OE):
                y = q.Spinor(L, q.HALF, host=np.zeros_like(psi))
                y.applyDslash_xx(xf, uh, par)
                out.append(y.download())
            for parity, sub in ((q.EE, q.EVEN), (q.OO, q.ODD)):
                x = q.Spinor(L, q.HALF, sub, host=psi)
                for normal in (0, 1):
                    y = q.Spinor(L, q.HALF, sub)
                    (y.tm_xxDagTm_xx if normal else y.tm_xx)(x, uh, p, parity)
                    out.append(y.download())
            xs = q.Spinor(L)
            st = q.solve(xs, b, u, p, eo=True, inner_prec=q.HALF, delta=0.01, tol=1e-10)
            assert st.converged
            runs[on] = (out, xs.download(), st.iterations)
    finally:
        q.set_neighbour_staging(0)
    a, xa, ia = runs[0]
    for mode in (1, 2):
        c, xc, ic = runs[mode]
        assert ia == ic, mode
        assert np.array_equal(xa, xc), mode
        for k, (g0, g1) in enumerate(zip(a, c)):
            assert np.array_equal(g0, g1), (mode, k)


@pytest.mark.parametrize("prec,recon,tol", [
    ("SINGLE", 18, 2e-6), ("SINGLE", 12, 2e-6), ("HALF", 12, 2e-3), ("BF16", 12, 2e-2)])
def test_low_precision_tm(q, prec, recon, tol):
    """Low-precision storage: tm and the Schur normal operator within the
    precision's rounding (FP32 ~1e-7, FP16 ~5e-4, BF16 ~4e-3 per element)."""
    L, kappa, mu = (8, 8, 4, 8), 0.1632, 0.004
    P = getattr(q, prec)
    U, psi, phi = _setup(q, L, 0.3, 21, -1)
    Ub = orc.apply_bc(U, L, -1)
    u64 = q.Gauge(L, boundaryT=-1, host=U)
    u = q.Gauge(L, P, recon, -1)
    u.convert_from(u64)
    p = q.params(L, kappa, mu, 1, -1)
    x = q.Spinor(L, P, host=psi)
    y = q.Spinor(L, P)
    y.tm(x, u, p)
    assert rel(y.download(), orc.tm(psi, Ub, L, kappa, mu, 1)) <= tol
    xe = q.Spinor(L, P, q.EVEN, host=psi)
    ye = q.Spinor(L, P, q.EVEN)
    ye.tm_xxDagTm_xx(xe, u, p, q.EE)
    assert rel(ye.download(), orc.tm_xx(psi, Ub, L, kappa, mu, 1, orc.UKQCD, orc.EE, 1)) <= 4 * tol


@pytest.mark.parametrize("prec,tol", [("DOUBLE", 1e-13), ("SINGLE", 2e-6), ("HALF", 2e-3), ("BF16", 2e-2)])
def test_sign_carrying_links_in_12_reals(q, prec, tol):
    """Links that carry the antiperiodic sign (boundaryT = +1), converted or
    uploaded into 12-real storage: row 2 rebuilt from the signed rows would
    have the wrong sign on the last time slice, so the field is stored with
    those t-links negated and takes boundaryT = -1 -- the same operator as the
    18-real field.  Links that are not SU(3) are refused."""
    L, kappa, mu = (8, 8, 4, 8), 0.1632, 0.004
    P = getattr(q, prec)
    U, psi, _ = _setup(q, L, 0.3, 25, -1)
    Ub = orc.apply_bc(U, L, -1)
    us = q.Gauge(L, boundaryT=+1, host=Ub)
    p = q.params(L, kappa, mu, 1, -1)
    ref = orc.tm(psi, Ub, L, kappa, mu, 1)
    ref_e = orc.tm_xx(psi, Ub, L, kappa, mu, 1, orc.UKQCD, orc.EE, 1)
    conv = q.Gauge(L, P, q.RECON_12, +1)
    conv.convert_from(us)
    up = q.Gauge(L, P, q.RECON_12, +1, host=Ub)
    for name, u in (("convert", conv), ("upload", up)):
        x, y = q.Spinor(L, P, host=psi), q.Spinor(L, P)
        y.tm(x, u, p)
        assert rel(y.download(), ref) <= tol, name
        xe, ye = q.Spinor(L, P, q.EVEN, host=psi), q.Spinor(L, P, q.EVEN)
        ye.tm_xxDagTm_xx(xe, u, p, q.EE)
        assert rel(ye.download(), ref_e) <= 4 * tol, name
    wide = q.Gauge(L, boundaryT=+1, host=1.01 * Ub)
    with pytest.raises(q.QkxError, match="not SU\\(3\\)"):
        q.Gauge(L, P, q.RECON_12, +1).convert_from(wide)
    with pytest.raises(q.QkxError, match="not SU\\(3\\)"):
        q.Gauge(L, P, q.RECON_12, -1, host=1.01 * U)


def test_blas(q):
    L = (4, 4, 4, 8)
    a = q.gen_gaussian(L, 31)
    b = q.gen_gaussian(L, 32)
    x, y = q.Spinor(L, host=a), q.Spinor(L, host=b)
    assert abs(q.reCdotXdagY(x, y) - orc.redot(a, b)) <= 1e-13 * np.sqrt(orc.norm2(a) * orc.norm2(b))
    assert abs(x.norm2() - orc.norm2(a)) <= 1e-14 * orc.norm2(a)
    q.X_eq_X_p_aY(x, 0.37, y)
    assert rel(x.download(), orc.axpy(a, 0.37, b)) <= 1e-15
    x.upload(a)
    q.X_eq_aX_p_Y(x, -1.25, y)
    assert rel(x.download(), orc.xpay(a, -1.25, b)) <= 1e-15


def test_upload_download_roundtrip_and_subsets(q):
    L = (4, 6, 8, 4)
    a = q.gen_gaussian(L, 41)
    for prec in (q.DOUBLE,):
        f = q.Spinor(L, prec, host=a)
        assert np.array_equal(f.download(), a)
    e = q.Spinor(L, q.DOUBLE, q.EVEN, host=a)
    got = e.download()
    V = np.prod(L)
    x = np.arange(V)
    c = np.stack(np.unravel_index(x, (L[3], L[2], L[1], L[0])), 1)
    odd = (c.sum(1) & 1).astype(bool)
    assert np.array_equal(got[~odd], a[~odd]) and not np.any(got[odd])


def test_mem_info_counts_fields(q):
    """qkx_mem_info: bytes in use grow by at least a new field's size."""
    used0, total = q.mem_info()
    assert 0 < used0 <= total
    L = (32, 32, 16, 32)      # a size no other test allocates
    f = q.Spinor(L, nrhs=5)   # 5 x 524288 sites x 192 B = 503 MB
    used1, _ = q.mem_info()
    assert used1 - used0 >= 0.9 * 5 * 524288 * 192
    del f


def test_argument_errors(q):
    L = (4, 4, 4, 4)
    with pytest.raises(q.QkxError):
        q.Spinor((3, 4, 4, 4))
    u = q.Gauge(L, host=q.gen_gauge(L, 0.3, 1))
    a, b = q.Spinor(L), q.Spinor(L, subset=q.EVEN)
    with pytest.raises(q.QkxError):
        b.applyDslash(a, u)                     # full-lattice hop needs full fields
    with pytest.raises(q.QkxError):
        a.tm(a, u, q.params(L, 0.1, 0.1))       # out must differ from in
    with pytest.raises(q.QkxError):
        a.applyDslash_xx(a.__class__(L), u, q.EE)  # EE is not a hopping parity
