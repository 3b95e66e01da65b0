"""Several independent systems -- uneven lattices, their own gauge fields,
twisted masses and kappas -- solved in lockstep by qkx_solve_multi, every CG
launch serving all of them.  Each instance must meet its own true residual
(recomputed by the oracle in FP64, the reference operator
lib/lattice_util.cpp:1925-1928) and agree with its own single solve
(qkx_solve, lib/solvers.cpp:16-81 is the reference's one-system CG)."""
import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu


def _instances(q, specs):
    out = []
    for L, eps, seed, kappa, mu in specs:
        U = q.gen_gauge(L, eps, seed)
        src = q.gen_z2(L, seed + 100)
        out.append(dict(L=L, U=U, Ub=orc.apply_bc(U, L, -1), src=src, kappa=kappa, mu=mu,
                        u=q.Gauge(L, boundaryT=-1, host=U), b=q.Spinor(L, host=src), x=q.Spinor(L),
                        p=q.params(L, kappa, mu, 1, -1)))
    return out


@pytest.mark.parametrize("prec,delta", [("DOUBLE", 0.1), ("SINGLE", 0.003), ("HALF", 0.01), ("BF16", 0.03)])
def test_uneven_instances_in_lockstep(q, prec, delta):
    specs = [((4, 4, 4, 8), 0.3, 1, 0.1632, 0.004), ((8, 8, 8, 16), 0.3, 2, 0.1632, 0.01),
             ((8, 6, 4, 8), 0.4, 3, 0.156, 0.02), ((8, 8, 8, 16), 0.3, 4, 0.16, 0.004)]
    I = _instances(q, specs)
    print(f"multi-instance solve, {prec} inner, {len(I)} instances", flush=True)
    st = q.solve_multi([c["x"] for c in I], [c["b"] for c in I], [c["u"] for c in I], [c["p"] for c in I],
                       inner_prec=getattr(q, prec), delta=delta, tol=1e-10)
    for c, s in zip(I, st):
        xs = c["x"].download()
        res = rel(orc.tm(xs, c["Ub"], c["L"], c["kappa"], c["mu"], 1), c["src"])
        print(f"  L={c['L']} mu={c['mu']} iterations {s.iterations} updates {s.reliable_updates} residual {res:.2e}",
              flush=True)
        assert s.converged and res <= 1e-10 and abs(s.true_residual - res) <= 1e-3 * res + 1e-15
        x1 = q.Spinor(c["L"])
        s1 = q.solve(x1, c["b"], c["u"], c["p"], eo=True, inner_prec=getattr(q, prec), delta=delta, tol=1e-10)
        assert s1.converged and rel(xs, x1.download()) <= 1e-8
        assert abs(s.iterations - s1.iterations) <= max(3, s1.iterations // 10)


def test_config3_second_batch_16x16x16x32_24x24x24x48(q):
    """The workload's uneven batch: 16^3x32 and 24^3x48 lattices, each at
    mu = 0.004, 0.01, 0.02 (six instances in one lockstep solve), FP32 inner
    with FP64 reliable updates."""
    specs = []
    for L, seed in (((16, 16, 16, 32), 11), ((24, 24, 24, 48), 12)):
        for mu in (0.004, 0.01, 0.02):
            specs.append((L, 0.3, seed, 0.1632, mu))
    I = []
    gauges = {}
    for L, eps, seed, kappa, mu in specs:   # one gauge field per lattice, shared by its three masses
        if L not in gauges:
            U = q.gen_gauge(L, eps, seed)
            gauges[L] = (U, orc.apply_bc(U, L, -1), q.Gauge(L, boundaryT=-1, host=U))
        U, Ub, u = gauges[L]
        src = q.gen_z2(L, seed + 100)
        I.append(dict(L=L, Ub=Ub, src=src, kappa=kappa, mu=mu, u=u, b=q.Spinor(L, host=src), x=q.Spinor(L),
                      p=q.params(L, kappa, mu, 1, -1)))
    print("uneven batch: 16^3x32 and 24^3x48 at three masses each, FP32 inner", flush=True)
    st = q.solve_multi([c["x"] for c in I], [c["b"] for c in I], [c["u"] for c in I], [c["p"] for c in I],
                       inner_prec=q.SINGLE, delta=0.003, tol=1e-10)
    for c, s in zip(I, st):
        res = rel(orc.tm(c["x"].download(), c["Ub"], c["L"], c["kappa"], c["mu"], 1), c["src"])
        print(f"  L={c['L']} mu={c['mu']} iterations {s.iterations} residual {res:.2e}", flush=True)
        assert s.converged and res <= 1e-10


def test_multi_zero_source_and_argument_checks(q):
    I = _instances(q, [((4, 4, 4, 8), 0.3, 5, 0.15, 0.01), ((8, 8, 8, 16), 0.3, 6, 0.15, 0.01)])
    I[0]["b"].zero()
    st = q.solve_multi([c["x"] for c in I], [c["b"] for c in I], [c["u"] for c in I], [c["p"] for c in I],
                       inner_prec=q.HALF, delta=0.05, tol=1e-10)
    assert st[0].converged and not np.any(I[0]["x"].download())
    assert st[1].converged and st[1].true_residual <= 1e-10
    with pytest.raises(q.QkxError):
        q.solve_multi([I[0]["x"]], [I[1]["b"]], [I[0]["u"]], [I[0]["p"]])   # lattices differ
