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
            "orc_ptc": ([P, P, P, IP, I, I, P], None),
            "orc_ptc_adj": ([P, P, P, IP, I, I, P], None),
            "orc_ptc_wgrad": ([P, P, P, P, IP, I, I], None),
            "orc_net_apply": ([P, P, P, IP, I, P, P], None),
            "orc_net_loss_grad": ([P, P, P, IP, I, P, P, P, P], D),
            "orc_su3_project": ([P], None),
            "orc_project_su3_field": ([P, IP], None),
            "orc_ape_smear": ([P, P, IP, I, D], None),
            "orc_gauss_smear": ([P, P, P, IP, I, D], None),
            "orc_plaquette": ([P, IP], D),
            "orc_prop_rotate": ([P, ctypes.c_long, I], None),
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


# ---- learned preconditioner: parallel-transport convolutions (qkx_oracle_ptc.c)
def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def ptc(f, U, L, W):
    """One PTC layer; f [V, cin, 3], W [9, cout, cin] -> [V, cout, 3]."""
    f, U, W = _c(f), _c(U), _c(W)
    cout, cin = W.shape[1], W.shape[2]
    out = np.zeros((_vol(L), cout, 3), dtype=np.complex128)
    lib().orc_ptc(_p(out), _p(f), _p(U), _L(L), cin, cout, _p(W))
    return out


def ptc_adj(g, U, L, W):
    """Adjoint of ptc with weights W [9, cout, cin]; g [V, cout, 3] -> [V, cin, 3]."""
    g, U, W = _c(g), _c(U), _c(W)
    cout, cin = W.shape[1], W.shape[2]
    out = np.zeros((_vol(L), cin, 3), dtype=np.complex128)
    lib().orc_ptc_adj(_p(out), _p(g), _p(U), _L(L), cin, cout, _p(W))
    return out


def ptc_wgrad(g, f, U, L):
    """sum_x g_a(x) . conj(T_p f_b (x)) -> [9, cout, cin]."""
    g, f, U = _c(g), _c(f), _c(U)
    cout, cin = g.shape[1], f.shape[1]
    G = np.zeros((9, cout, cin), dtype=np.complex128)
    lib().orc_ptc_wgrad(_p(G), _p(g), _p(f), _p(U), _L(L), cin, cout)
    return G


def net_apply(psi, U, L, W1, W2):
    """M psi = L2(L1(psi)); psi [V, 4, 3]; W1 [9, ch, 4], W2 [9, 4, ch]."""
    psi, U, W1, W2 = _c(psi), _c(U), _c(W1), _c(W2)
    out = _new(L)
    lib().orc_net_apply(_p(out), _p(psi), _p(U), _L(L), W1.shape[1], _p(W1), _p(W2))
    return out


def net_loss_grad(r, e, U, L, W1, W2):
    """(|M r - e|^2, dL/dW1, dL/dW2) in the dL/dRe + i dL/dIm convention."""
    r, e, U, W1, W2 = _c(r), _c(e), _c(U), _c(W1), _c(W2)
    G1, G2 = np.zeros_like(W1), np.zeros_like(W2)
    loss = lib().orc_net_loss_grad(_p(r), _p(e), _p(U), _L(L), W1.shape[1], _p(W1), _p(W2), _p(G1), _p(G2))
    return loss, G1, G2


# ---- source preparation (qkx_oracle_smear.c)
def su3_project(M):
    """SU(3) projection of one 3x3 complex matrix (the reference's algorithm)."""
    m = np.array(M, dtype=np.complex128, copy=True).reshape(3, 3)
    lib().orc_su3_project(_p(m))
    return m


def project_su3_field(U, L):
    """Projection of the spatial links of a gauge field [V, 4, 3, 3]."""
    U = np.array(U, dtype=np.complex128, copy=True)
    lib().orc_project_su3_field(_p(U), _L(L))
    return U


def ape_smear(U, L, nsmear, alpha):
    """nsmear APE steps on the spatial links; temporal links unchanged."""
    U = _c(U)
    out = U.copy()
    lib().orc_ape_smear(_p(out), _p(U), _L(L), nsmear, alpha)
    return out


def gauss_smear(psi, U, L, nsmear, alpha):
    """nsmear Gaussian smearing steps of a spinor [V, 4, 3] with links U."""
    psi, U = _c(psi), _c(U)
    out = psi.copy()
    lib().orc_gauss_smear(_p(out), _p(psi), _p(U), _L(L), nsmear, alpha)
    return out


def plaquette(U, L):
    return lib().orc_plaquette(_p(_c(U)), _L(L))


def prop_rotate(M, sign):
    """Propagator [V, 4, 4, 3, 3] rotated to the physical basis (sign +1 / -1)."""
    M = np.array(M, dtype=np.complex128, copy=True)
    lib().orc_prop_rotate(_p(M), M.shape[0], sign)
    return M
