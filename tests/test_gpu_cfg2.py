"""The headline workload (24^3x48, eps 0.3, kappa 0.1632, mu 0.004,
antiperiodic t, Z2 source; gauge seed 1, source seed 2) against the
reference itself: tests/golden/cg_cfg2.npz holds summaries (norm, four
seeded Gaussian projections, 64 sampled sites) of the reference's own
operators (lib/lattice_util.cpp:1925-2071) and of its CG solution
(lib/solvers.cpp:16-81, D^dag D x = D^dag b, iterated tolerance 1e-14) on
these inputs.

Operators: every summary to 1e-13, and the whole output against the CPU
oracle to 1e-13 relative L2 (the oracle is pinned to the same fixture in
tests/test_oracle_golden.py).  Solvers: every variant of qkx_solve reaches
true residual <= 1e-10, recomputed by the oracle in FP64, and agrees with
the reference solution to <= 1e-9: in the norm, on the 64 sites, and in
the relative L2 error estimated from the four Gaussian projections
(E |<R, e>|^2 = |e|^2 E|R_i|^2 for a Gaussian probe R)."""
import os

import numpy as np
import pytest

import oracle as orc
from golden_util import input_hash, probes, summarize
from refops import gpu_op, oracle_op, rel

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def cfg2(q):
    g = np.load(os.path.join(GOLD, "cg_cfg2.npz"), allow_pickle=False)
    L = tuple(int(v) for v in g["L"])
    eps, kappa, mu, ts, bt, sg, sp, _ = g["params"]
    U = q.gen_gauge(L, eps, int(sg))
    src = q.gen_z2(L, int(sp))
    assert str(g["input_hash"]) == input_hash(U, src), "generator output drifted"
    R, sites = probes(L, len(g["sites"]))
    assert np.array_equal(sites, g["sites"])
    Ub = orc.apply_bc(U, L, int(bt))
    u = q.Gauge(L, boundaryT=int(bt), host=U)
    p = q.params(L, kappa, mu, int(ts), int(bt))
    return dict(g=g, L=L, kappa=kappa, mu=mu, ts=int(ts), U=U, Ub=Ub, src=src, R=R, sites=sites, u=u, p=p)


def projected_error(s, g, name, R):
    """Relative L2 error estimated from the Gaussian projections: each
    <R_k, x - x_ref> has variance |x - x_ref|^2 E|R_i|^2."""
    var = np.mean([np.vdot(r.ravel(), r.ravel()).real / r.size for r in R])
    d = s["proj"] - g["proj_" + name]
    return np.sqrt(np.mean(np.abs(d) ** 2) / var / float(g["norm2_" + name]))


@pytest.mark.parametrize("name", ["tm", "tmDag", "tm_chiral", "tm_xx_EE", "tm_xxDagTm_xx_EE",
                                  "tm_xxDagTm_xx_OO"])
def test_operator_matches_reference_at_24x24x24x48(q, cfg2, name):
    c = cfg2
    print(f"24^3x48 operator {name}", flush=True)
    src_d = q.Spinor(c["L"], host=c["src"])
    y = gpu_op(q, name, c["u"], src_d, src_d, c["p"])
    s = summarize(y, c["R"], c["sites"])
    g = c["g"]
    n2 = float(g["norm2_" + name])
    assert abs(s["norm2"] - n2) <= 1e-13 * n2
    assert projected_error(s, g, name, c["R"]) <= 1e-13
    assert rel(s["sites"], g["site_" + name]) <= 1e-13
    ref = oracle_op(name, c["Ub"], c["src"], c["src"], c["L"], c["kappa"], c["mu"], c["ts"])
    assert rel(y, ref) <= 1e-13


# (eo, inner precision, reliable-update delta): the tuned deltas of the
# benchmark and the delta = 0.1 the workload states
SOLVES = [(False, "DOUBLE", 0.1), (True, "DOUBLE", 0.1), (True, "SINGLE", 0.003), (True, "SINGLE", 0.1),
          (True, "HALF", 0.01), (True, "HALF", 0.1), (True, "BF16", 0.03), (True, "BF16", 0.1)]


@pytest.mark.parametrize("eo,prec,delta", SOLVES)
def test_solve_matches_reference_solution_at_24x24x24x48(q, cfg2, eo, prec, delta):
    c = cfg2
    L = c["L"]
    print(f"24^3x48 solve eo={eo} {prec} delta={delta}", flush=True)
    x, b = q.Spinor(L), q.Spinor(L, host=c["src"])
    st = q.solve(x, b, c["u"], c["p"], eo=eo, inner_prec=getattr(q, prec), delta=delta, tol=1e-10)
    xs = x.download()
    res = rel(orc.tm(xs, c["Ub"], L, c["kappa"], c["mu"], c["ts"]), c["src"])
    s = summarize(xs, c["R"], c["sites"])
    g = c["g"]
    n = np.sqrt(float(g["norm2_cg_x"]))
    err_norm = abs(np.sqrt(s["norm2"]) - n) / n
    err_proj = projected_error(s, g, "cg_x", c["R"])
    err_sites = rel(s["sites"], g["site_cg_x"])
    print(f"  iterations {st.iterations} reliable updates {st.reliable_updates} residual {res:.2e} "
          f"error vs reference: norm {err_norm:.1e} projections {err_proj:.1e} sites {err_sites:.1e}", flush=True)
    assert st.converged and res <= 1e-10 and abs(st.true_residual - res) <= 1e-3 * res + 1e-15
    assert err_norm <= 1e-9 and err_proj <= 1e-9 and err_sites <= 1e-9


def test_operators_match_oracle_at_32x32x32x64(q):
    """The config-3 lattice: tm and the even-odd normal operator against the
    oracle, whole outputs, on a Gaussian field."""
    L = (32, 32, 32, 64)
    kappa, mu = 0.1632, 0.004
    U = q.gen_gauge(L, 0.3, 1)
    psi = q.gen_gaussian(L, 3)
    Ub = orc.apply_bc(U, L, -1)
    u = q.Gauge(L, boundaryT=-1, host=U)
    del U
    p = q.params(L, kappa, mu, 1, -1)
    X = q.Spinor(L, host=psi)
    for name in ("tm", "tm_xxDagTm_xx_EE"):
        print(f"32^3x64 operator {name}", flush=True)
        y = gpu_op(q, name, u, X, X, p)
        assert rel(y, oracle_op(name, Ub, psi, psi, L, kappa, mu, 1)) <= 1e-13, name
