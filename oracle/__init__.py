"""ctypes bindings of the CPU oracle (oracle/qkx_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product (qkxtm_mpi_amd) never imports this.
Arrays are numpy complex128 in the reference layout: spinor [V,4,3], gauge
[V,4,3,3] (reference include/qkxTM.h:117-134).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

EO, OE, OO, EE, ALL = 0, 1, 2, 3, -1
UKQCD, CHIRAL = 0, 1
TWIST, TWIST_DAG, TWIST_INV, TWIST_DAG_INV = 0, 1, 2, 3
TM, TM_DAG, TM_DAG_TM, WILSON = 0, 1, 2, 3

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        IP = ctypes.POINTER(ctypes.c_int)
        sig = {
            "orc_hop": ([P, P, P, IP, I, I, I], None),
            "orc_twist_add_dslash": ([P, P, P, IP, D, D, I, I, I], None),
            "orc_twist_add_dslash_pc": ([P, P, P, IP, D, D, I, I, I], None),
            "orc_delta_add_dslash": ([P, P, P, IP, D], None),
            "orc_twist": ([P, P, IP, D, D, I, I, I], None),
            "orc_gamma5": ([P, P, IP, I], None),
            "orc_tm": ([P, P, P, IP, D, D, I, I, I], None),
            "orc_tm_xx": ([P, P, P, IP, D, D, I, I, I, I], None),
            "orc_redot": ([P, P, ctypes.c_long], D),
            "orc_axpy": ([P, D, P, ctypes.c_long], None),
            "orc_xpay": ([P, D, P, ctypes.c_long], None),
            "orc_norm2": ([P, ctypes.c_long], D),
            "orc_cg": ([P, P, P, IP, D, D, I, D, I, ctypes.POINTER(ctypes.c_double)], I),
            "orc_apply_bc": ([P, IP, I], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _L(L):
    return (ctypes.c_int * 4)(*L)


def _vol(L):
    return int(np.prod(L))


def _new(L):
    return np.zeros((_vol(L), 4, 3), dtype=np.complex128)


def apply_bc(U, L, boundaryT):
    U = np.array(U, dtype=np.complex128, copy=True)
    lib().orc_apply_bc(_p(U), _L(L), boundaryT)
    return U


def hop(psi, U, L, basis=UKQCD, dagger=0, parity=ALL):
    out = _new(L)
    lib().orc_hop(_p(out), _p(psi), _p(U), _L(L), basis, dagger, parity)
    return out


def twist_add_dslash(d, psi, L, kappa, mu, s, basis=UKQCD, dagger=0):
    out = _new(L)
    lib().orc_twist_add_dslash(_p(out), _p(d), _p(psi), _L(L), kappa, mu, s, basis, dagger)
    return out


def twist_add_dslash_pc(d, psi, L, kappa, mu, s, basis, parity):
    out = _new(L)
    lib().orc_twist_add_dslash_pc(_p(out), _p(d), _p(psi), _L(L), kappa, mu, s, basis, parity)
    return out


def delta_add_dslash(d, psi, L, kappa):
    out = _new(L)
    lib().orc_delta_add_dslash(_p(out), _p(d), _p(psi), _L(L), kappa)
    return out


def twist(psi, L, kappa, mu, s, basis, kind):
    out = _new(L)
    lib().orc_twist(_p(out), _p(psi), _L(L), kappa, mu, s, basis, kind)
    return out


def gamma5(psi, L, basis=UKQCD):
    out = _new(L)
    lib().orc_gamma5(_p(out), _p(psi), _L(L), basis)
    return out


def tm(psi, U, L, kappa, mu, s, basis=UKQCD, op=TM):
    out = _new(L)
    lib().orc_tm(_p(out), _p(psi), _p(U), _L(L), kappa, mu, s, basis, op)
    return out


def tm_xx(psi, U, L, kappa, mu, s, basis, parity, normal=0):
    out = _new(L)
    lib().orc_tm_xx(_p(out), _p(psi), _p(U), _L(L), kappa, mu, s, basis, parity, normal)
    return out


def redot(x, y):
    return lib().orc_redot(_p(x), _p(y), x.shape[0])


def norm2(x):
    return lib().orc_norm2(_p(x), x.shape[0])


def axpy(x, a, y):
    x = x.copy()
    lib().orc_axpy(_p(x), a, _p(y), x.shape[0])
    return x


def xpay(x, a, y):
    x = x.copy()
    lib().orc_xpay(_p(x), a, _p(y), x.shape[0])
    return x


def cg(b, U, L, kappa, mu, s, tol, maxIter, x0=None):
    """Reference CG on D^dag D; returns (x, iterations, iterated r2)."""
    x = _new(L) if x0 is None else np.array(x0, dtype=np.complex128, copy=True)
    r2 = ctypes.c_double()
    k = lib().orc_cg(_p(x), _p(b), _p(U), _L(L), kappa, mu, s, tol, maxIter, ctypes.byref(r2))
    return x, k, r2.value
