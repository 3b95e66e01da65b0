"""Mapping of the reference operator names (as written by the golden-fixture
harness) onto the CPU oracle and onto the HIP library.

Inputs follow the harness (oracle/ref/ref_harness.cpp): psi, phi random
spinors; the gauge carries the boundary sign (Gauge::applyBoundaryCondition,
reference lib/lattice_util.cpp:708-724) for the oracle, while the HIP path
applies it inside the kernels (or receives already-signed links).
"""
import numpy as np

import oracle as orc

PAR = {"EO": orc.EO, "OE": orc.OE, "OO": orc.OO, "EE": orc.EE}


def oracle_op(name, Ub, psi, phi, L, kappa, mu, ts):
    """Oracle output of reference operator `name` (Ub already has the BC sign)."""
    b = orc.CHIRAL if "_chiral" in name else orc.UKQCD
    par = PAR.get(name.rsplit("_", 1)[-1])
    base = name.replace("_chiral", "")
    if par is not None:
        base = base.rsplit("_", 1)[0]
    if base == "applyDslash":
        return orc.hop(psi, Ub, L, b, 0, orc.ALL)
    if base == "applyDslashDag":
        return orc.hop(psi, Ub, L, b, 1, orc.ALL)
    if base == "applyDslash_xx":
        return orc.hop(psi, Ub, L, b, 0, par)
    if base == "applyTwistAddDslash":
        return orc.twist_add_dslash(phi, psi, L, kappa, mu, ts, b, 0)
    if base == "applyTwistDagAddDslashDag":
        return orc.twist_add_dslash(phi, psi, L, kappa, mu, ts, b, 1)
    if base == "applyTwistAddDslash_PC":
        return orc.twist_add_dslash_pc(phi, psi, L, kappa, mu, ts, b, par)
    if base == "applyDeltaAddDslash":
        return orc.delta_add_dslash(phi, psi, L, kappa)
    kinds = {"applyTwist": orc.TWIST, "applyTwistDag": orc.TWIST_DAG, "applyTwistInv": orc.TWIST_INV,
             "applyTwistDagInv": orc.TWIST_DAG_INV}
    if base in kinds:
        return orc.twist(psi, L, kappa, mu, ts, b, kinds[base])
    if base == "applyGamma5":
        return orc.gamma5(psi, L, b)
    ops = {"tm": orc.TM, "tmDag": orc.TM_DAG, "tmDagtm": orc.TM_DAG_TM, "Wilson": orc.WILSON}
    if base in ops:
        return orc.tm(psi, Ub, L, kappa, mu, ts, b, ops[base])
    if base == "tm_xx":
        return orc.tm_xx(psi, Ub, L, kappa, mu, ts, b, par, 0)
    if base == "tm_xxDagTm_xx":
        return orc.tm_xx(psi, Ub, L, kappa, mu, ts, b, par, 1)
    if base == "X_eq_X_p_aY":
        return orc.axpy(psi, 0.37, phi)
    if base == "X_eq_aX_p_Y":
        return orc.xpay(psi, -1.25, phi)
    raise KeyError(name)


def gpu_op(q, name, u, psi_d, phi_d, p):
    """Run reference operator `name` through the C ABI; returns host array."""
    L = p.L[:]
    out = q.Spinor(L)
    b = q.CHIRAL if "_chiral" in name else q.UKQCD
    par = {"EO": q.EO, "OE": q.OE, "OO": q.OO, "EE": q.EE}.get(name.rsplit("_", 1)[-1])
    base = name.replace("_chiral", "")
    if par is not None:
        base = base.rsplit("_", 1)[0]
    if base == "applyDslash":
        out.applyDslash(psi_d, u, b)
    elif base == "applyDslashDag":
        out.applyDslashDag(psi_d, u, b)
    elif base == "applyDslash_xx":
        out.applyDslash_xx(psi_d, u, par, b)
    elif base == "applyTwistAddDslash":
        out.twist_add_dslash(phi_d, psi_d, p, q.XPAY_TM, b)
    elif base == "applyTwistDagAddDslashDag":
        out.twist_add_dslash(phi_d, psi_d, p, q.XPAY_TM_DAG, b)
    elif base == "applyTwistAddDslash_PC":
        out.twist_add_dslash(phi_d, psi_d, p, q.XPAY_PC, b, par)
    elif base == "applyDeltaAddDslash":
        out.twist_add_dslash(phi_d, psi_d, p, q.XPAY_WILSON, b)
    elif base in ("applyTwist", "applyTwistDag", "applyTwistInv", "applyTwistDagInv", "applyGamma5"):
        kind = {"applyTwist": q.TWIST, "applyTwistDag": q.TWIST_DAG, "applyTwistInv": q.TWIST_INV,
                "applyTwistDagInv": q.TWIST_DAG_INV, "applyGamma5": q.GAMMA5}[base]
        out.twist(psi_d, p, kind, b)
    elif base == "tm":
        out.tm(psi_d, u, p, b)
    elif base == "tmDag":
        out.tmDag(psi_d, u, p)
    elif base == "tmDagtm":
        out.tmDagtm(psi_d, u, p)
    elif base == "Wilson":
        out.Wilson(psi_d, u, p, b)
    elif base == "tm_xx":
        out.tm_xx(psi_d, u, p, par, b)
    elif base == "tm_xxDagTm_xx":
        out.tm_xxDagTm_xx(psi_d, u, p, par, b)
    elif base == "X_eq_X_p_aY":
        out.copy(psi_d)
        q.X_eq_X_p_aY(out, 0.37, phi_d)
    elif base == "X_eq_aX_p_Y":
        out.copy(psi_d)
        q.X_eq_aX_p_Y(out, -1.25, phi_d)
    else:
        raise KeyError(name)
    return out.download()


ALL_OPS = [
    "applyDslash", "applyDslash_chiral", "applyDslashDag", "applyDslash_xx_EO", "applyDslash_xx_OE",
    "applyDslash_xx_chiral_EO", "applyDslash_xx_chiral_OE", "applyTwistAddDslash", "applyTwistAddDslash_chiral",
    "applyTwistAddDslash_PC_OO", "applyTwistAddDslash_PC_EE", "applyTwistAddDslash_PC_chiral_OO",
    "applyTwistAddDslash_PC_chiral_EE", "applyDeltaAddDslash", "applyTwistDagAddDslashDag", "applyTwist",
    "applyTwistDag", "applyTwistInv", "applyTwistInv_chiral", "applyTwistDagInv", "applyGamma5",
    "applyGamma5_chiral", "tm", "tm_chiral", "tmDag", "tmDagtm", "Wilson", "tm_xx_EE", "tm_xx_OO",
    "tm_xxDagTm_xx_EE", "tm_xxDagTm_xx_OO", "tm_xx_chiral_EE", "tm_xx_chiral_OO", "tm_xxDagTm_xx_chiral_EE",
    "tm_xxDagTm_xx_chiral_OO", "X_eq_X_p_aY", "X_eq_aX_p_Y",
]


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
