"""The stated workloads beyond the headline one, at their own sizes:

  * 32^3x64, kappa 0.1632, mu 0.004: the 12 spin-colour point sources as one
    batch with the FP32 inner solver the workload states;
  * 48^3x96, eps 0.25: the production 12-source propagator with the BF16
    inner solver, every column checked by an FP64 residual formed outside
    the solver;
  * 16^3x32, kappa 0.1625, mu 0.004, eps 0.3: the learned preconditioner
    trained on plain-CG pairs (fewer pairs than the full workload, same
    recipe) and used by flexible GCR on an unseen gauge field, with and
    without the network, each solution checked by the oracle in FP64.

The reference has one solve per call (lib/solvers.cpp:16-81); each column
here must meet the residual the reference CG would be asked for."""
import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu


def point_residuals(q, P, u, p, L, x0, cols):
    """|D x_r - delta_r| for the given columns of a 12-member solution, with
    the FP64 operator and device BLAS (the point source has norm 1)."""
    V = int(np.prod(L))
    iv = ((x0[3] * L[2] + x0[2]) * L[1] + x0[1]) * L[0] + x0[0]
    pt = np.zeros((V, 4, 3), complex)
    xr, br, Dx = q.Spinor(L), q.Spinor(L), q.Spinor(L)
    out = []
    for r in cols:
        xr.upload(P.download_rhs(r))
        pt[iv, r // 3, r % 3] = 1.0
        br.upload(pt)
        pt[iv, r // 3, r % 3] = 0.0
        Dx.tm(xr, u, p)
        q.X_eq_aX_p_Y(Dx, -1.0, br)
        out.append(np.sqrt(Dx.norm2()))
    return out


def test_fp32_inner_batch_32x32x32x64(q):
    L = (32, 32, 32, 64)
    kappa, mu = 0.1632, 0.004
    U = q.gen_gauge(L, 0.3, 1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    del U
    p = q.params(L, kappa, mu, 1, -1)
    x0 = q.gen_site(L, 5)
    print(f"32^3x64: 12 point sources at {x0}, FP32 inner", flush=True)
    b = q.point_sources(L, x0)
    P = q.Spinor(L, nrhs=12)
    st = q.solve_batch(P, b, u, p, eo=True, inner_prec=q.SINGLE, delta=0.003, tol=1e-10)
    print(f"  iterations {st[0].iterations} reliable updates {st[0].reliable_updates}", flush=True)
    assert all(s.converged and s.true_residual <= 1e-10 for s in st)
    res = point_residuals(q, P, u, p, L, x0, range(12))
    print("  residuals " + " ".join(f"{r:.1e}" for r in res), flush=True)
    assert max(res) <= 1e-10


def test_bf16_inner_batch_48x48x48x96(q):
    L = (48, 48, 48, 96)
    kappa, mu = 0.1632, 0.004
    print("48^3x96: generating links", flush=True)
    U = q.gen_gauge(L, 0.25, 1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    del U
    p = q.params(L, kappa, mu, 1, -1)
    x0 = q.gen_site(L, 5)
    print(f"48^3x96: 12 point sources at {x0}, BF16 inner", flush=True)
    b = q.point_sources(L, x0)
    P = q.Spinor(L, nrhs=12)
    st = q.solve_batch(P, b, u, p, eo=True, inner_prec=q.BF16, delta=0.03, tol=1e-10)
    print(f"  iterations {st[0].iterations} reliable updates {st[0].reliable_updates}", flush=True)
    assert all(s.converged and s.true_residual <= 1e-10 for s in st)
    del b
    res = point_residuals(q, P, u, p, L, x0, range(12))
    print("  residuals " + " ".join(f"{r:.1e}" for r in res), flush=True)
    assert max(res) <= 1e-10


def test_learned_preconditioner_16x16x16x32(q):
    from qkxtm_mpi_amd import learned
    L, kappa, mu, eps = (16, 16, 16, 32), 0.1625, 0.004, 0.3
    print("16^3x32: pairs from plain-CG iterates", flush=True)
    pairs = learned.make_pairs(L, kappa, mu, eps, gauge_seeds=[101, 102], sources_per_gauge=2, every=4,
                               max_pairs=512)
    net = q.Net(16, seed=1)
    loss0 = learned.pair_loss(net, pairs, kappa, mu)
    hist = learned.train(net, pairs, kappa, mu, epochs=2, batch=16, lr=2e-3, seed=2)
    print(f"  {len(pairs)} pairs, loss {loss0:.3f} -> {hist[-1]:.3f}", flush=True)
    assert hist[-1] < loss0
    del pairs
    U = q.gen_gauge(L, eps, 201)   # not among the training fields
    Ub = orc.apply_bc(U, L, -1)
    u64 = q.Gauge(L, boundaryT=-1, host=U)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(u64)
    p = q.params(L, kappa, mu, 1, -1)
    src = q.gen_z2(L, 208)
    b = q.Spinor(L, host=src)
    sols = {}
    for name, kw in (("gcr", {}), ("gcr_learned", {"net": net, "u_net": u32})):
        x = q.Spinor(L)
        st = q.solve_gcr(x, b, u64, p, restart=8, tol=1e-10, maxIter=20000, stall=0.9, **kw)
        xs = x.download()
        res = rel(orc.tm(xs, Ub, L, kappa, mu, 1), src)
        print(f"  {name}: iterations {st.iterations} (fallback CG {st.fallback_iterations}) residual {res:.2e}",
              flush=True)
        assert st.converged and res <= 1e-10
        sols[name] = xs
    assert rel(sols["gcr_learned"], sols["gcr"]) <= 1e-8
