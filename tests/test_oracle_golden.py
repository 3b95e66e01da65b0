"""The CPU oracle against golden vectors written by the reference's own code
(tests/golden/make_golden.py runs the compiled reference)."""
import glob
import os

import numpy as np
import pytest

import oracle as orc
import qkxtm_mpi_amd as q
from refops import oracle_op, rel
from golden_util import input_hash, probes, summarize

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(path):
    return np.load(path, allow_pickle=False)


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ops_*.npz"))))
def test_oracle_matches_reference_ops(path):
    g = _load(path)
    L = tuple(int(v) for v in g["L"])
    eps, kappa, mu, ts, bt, sg, sp, sf = g["params"]
    ts, bt = int(ts), int(bt)
    U = q.gen_gauge(L, eps, int(sg))
    psi = q.gen_gaussian(L, int(sp))
    phi = q.gen_gaussian(L, int(sf))
    assert str(g["input_hash"]) == input_hash(U, psi, phi), "generator output drifted"
    Ub = orc.apply_bc(U, L, bt)
    R, sites = probes(L)
    assert np.array_equal(sites, g["sites"])
    for name in g["names"]:
        y = oracle_op(str(name), Ub, psi, phi, L, kappa, mu, ts)
        s = summarize(y, R, sites)
        n2 = float(g["norm2_" + name])
        assert abs(s["norm2"] - n2) <= 1e-13 * max(n2, 1e-300), name
        scale = np.sqrt(max(n2, 1e-300)) * np.sqrt(12 * np.prod(L))
        assert np.max(np.abs(s["proj"] - g["proj_" + name])) <= 1e-13 * scale, name
        assert rel(s["sites"], g["site_" + name]) <= 1e-14, name
        if "full_" + name in g:
            assert rel(y, g["full_" + name]) <= 1e-15, name
    # BLAS scalars
    assert abs(orc.redot(psi, phi) - float(g["scalar_reCdotXdagY"][0])) <= 1e-13 * abs(float(g["scalar_reCdotXdagY"][0]))
    assert abs(orc.norm2(psi) - float(g["scalar_norm2"][0])) <= 1e-13 * float(g["scalar_norm2"][0])


def test_oracle_cg_matches_reference_cg():
    g = _load(os.path.join(GOLD, "cg_cfg1.npz"))
    L = tuple(int(v) for v in g["L"])
    eps, kappa, mu, ts, bt, sg, sp, sf = g["params"]
    U = q.gen_gauge(L, eps, int(sg))
    src = q.gen_z2(L, int(sp))
    assert str(g["input_hash"]) == input_hash(U, src)
    Ub = orc.apply_bc(U, L, int(bt))
    R, sites = probes(L)
    for name in ("tm", "tmDag", "tmDagtm", "tm_xxDagTm_xx_EE"):
        y = oracle_op(name, Ub, src, src, L, kappa, mu, int(ts))
        s = summarize(y, R, sites)
        assert abs(s["norm2"] - float(g["norm2_" + name])) <= 1e-13 * float(g["norm2_" + name])
        assert rel(s["sites"], g["site_" + name]) <= 1e-14
    b = orc.tm(src, Ub, L, kappa, mu, int(ts), orc.UKQCD, orc.TM_DAG)
    x, k, r2 = orc.cg(b, Ub, L, kappa, mu, int(ts), 1e-14, 20000)
    s = summarize(x, R, sites)
    # same algorithm, same operation order: agreement far below the 1e-9 bar
    assert abs(s["norm2"] - float(g["norm2_cg_x"])) <= 1e-12 * float(g["norm2_cg_x"])
    assert rel(s["sites"], g["site_cg_x"]) <= 1e-12
    res = rel(orc.tm(x, Ub, L, kappa, mu, int(ts)), src)
    assert res <= 1e-12 and float(g["cg_true_residual"]) <= 1e-12


def test_oracle_matches_reference_at_24x24x24x48():
    """Config 2 (the headline solve): the reference's own operators on the
    24^3x48 Z2 source, and the reference CG solution's true residual (the
    fixture that pins every solver of the library at this size)."""
    g = _load(os.path.join(GOLD, "cg_cfg2.npz"))
    L = tuple(int(v) for v in g["L"])
    eps, kappa, mu, ts, bt, sg, sp, sf = g["params"]
    U = q.gen_gauge(L, eps, int(sg))
    src = q.gen_z2(L, int(sp))
    assert str(g["input_hash"]) == input_hash(U, src), "generator output drifted"
    Ub = orc.apply_bc(U, L, int(bt))
    R, sites = probes(L, len(g["sites"]))
    assert np.array_equal(sites, g["sites"])
    for name in ("tm", "tmDag", "tm_xx_EE", "tm_xxDagTm_xx_EE", "tm_chiral", "tm_xxDagTm_xx_OO"):
        y = oracle_op(name, Ub, src, src, L, kappa, mu, int(ts))
        s = summarize(y, R, sites)
        n2 = float(g["norm2_" + name])
        assert abs(s["norm2"] - n2) <= 1e-13 * n2, name
        assert np.max(np.abs(s["proj"] - g["proj_" + name])) <= 1e-13 * np.sqrt(n2 * 12 * np.prod(L)), name
        assert rel(s["sites"], g["site_" + name]) <= 1e-14, name
    assert float(g["cg_true_residual"]) <= 1e-13
