"""qkxtm_mpi_amd -- MI355X twisted-mass Wilson operator and CG solvers.

Python host mirror of the C ABI in ``include/qkx_hip.h`` (libqkx_hip.so, built
in-tree for gfx950).  The reference's operators are methods of
``qkxTM::Vector`` acting on host arrays ``Complex M[V][4][3]``
(reference include/lattice_util.h:80-156); here a :class:`Spinor` owns a
device field and converts from/to that same host layout, and every operator
keeps the reference's name and argument meaning.

There is no CPU fallback: importing works without a GPU (the generators and
the library load are host-only), but any compute call needs the HIP library
and a device and raises :class:`QkxError` otherwise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqkx_hip.so")

# enums (include/qkx_hip.h)
SUCCESS, NOT_CONVERGED = 0, 1
PARITY_ALL, EO, OE, OO, EE = -1, 0, 1, 2, 3            # include/lattice_util.h:23
UKQCD, CHIRAL = 0, 1
DOUBLE, SINGLE, HALF, BF16 = 0, 1, 2, 3
RECON_18, RECON_12 = 18, 12
FULL, EVEN, ODD = 0, 1, 2
TWIST, TWIST_DAG, TWIST_INV, TWIST_DAG_INV, GAMMA5 = 0, 1, 2, 3, 4
XPAY_TM, XPAY_TM_DAG, XPAY_WILSON, XPAY_PC = 0, 1, 2, 3
OP_TM, OP_TM_DAG, OP_TM_DAG_TM, OP_WILSON = 0, 1, 2, 3
MAX_RHS = 16                                            # QKX_MAX_RHS


class QkxError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """Hot-path subset of the reference LatticeInfo (include/qkxTM.h:156-204)."""
    _fields_ = [("L", ctypes.c_int * 4), ("kappa", ctypes.c_double), ("mu", ctypes.c_double),
                ("twistSign", ctypes.c_int), ("boundaryT", ctypes.c_int), ("tol", ctypes.c_double),
                ("maxIter", ctypes.c_int), ("reliableDelta", ctypes.c_double)]


class SolverStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("reliable_updates", ctypes.c_int), ("converged", ctypes.c_int),
                ("iterated_r2", ctypes.c_double), ("true_residual", ctypes.c_double), ("seconds", ctypes.c_double),
                ("fallback_iterations", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GcrOpts(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int), ("maxIter", ctypes.c_int), ("tol", ctypes.c_double),
                ("stall", ctypes.c_double)]


class SolveOpts(ctypes.Structure):
    _fields_ = [("eo", ctypes.c_int), ("inner_prec", ctypes.c_int), ("inner_recon", ctypes.c_int),
                ("delta", ctypes.c_double), ("check_every", ctypes.c_int), ("maxIter", ctypes.c_int),
                ("tol", ctypes.c_double)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)

# name -> argtypes (all return int unless listed in _RESTYPES)
SIGNATURES = {
    "qkx_init": [_I], "qkx_finalize": [], "qkx_last_error": [], "qkx_device_count": [_IP], "qkx_sync": [], "qkx_mem_info": [_DP, _DP],
    "qkx_timer_start": [], "qkx_timer_stop": [_DP], "qkx_prof_enable": [_I], "qkx_prof_read": [_IP, _DP, _DP],
    "qkx_prof_read_prec": [_I, _IP, _DP, _DP], "qkx_prof_reset": [], "qkx_set_tuning": [_I, _I, _I, _I], "qkx_set_cg_batch": [_I], "qkx_set_rhs_tile": [_I], "qkx_set_site_order": [_I], "qkx_set_neighbour_staging": [_I], "qkx_gauge_su3_deviation": [_P, _DP], "qkx_set_batch_link_sharing": [_I], "qkx_set_autotune": [_I], "qkx_autotune_report": [ctypes.c_char_p, _I, _IP], "qkx_set_cg_fuse": [_I], "qkx_set_cg_graph": [_I],
    "qkx_gauge_create": [ctypes.POINTER(_P), _IP, _I, _I, _I], "qkx_gauge_destroy": [_P],
    "qkx_gauge_upload": [_P, _P], "qkx_gauge_convert": [_P, _P], "qkx_gauge_download": [_P, _P],
    "qkx_ape_smear": [_P, _P, _I, _D], "qkx_gauge_project_su3": [_P], "qkx_plaquette": [_DP, _P],
    "qkx_gauss_smear": [_P, _P, _P, _I, _D, _I],
    "qkx_prop_upload": [_P, _P], "qkx_prop_download": [_P, _P], "qkx_prop_absorb": [_P, _P, _I, _I],
    "qkx_prop_rotate_physical": [_P, _I], "qkx_spinor_rescale": [_P, _D],
    "qkx_lime_read_gauge": [ctypes.c_char_p, _IP, _P],
    "qkx_lime_write_gauge": [ctypes.c_char_p, _IP, _P, ctypes.c_char_p],
    "qkx_gauge_read_lime": [_P, ctypes.c_char_p], "qkx_gauge_write_lime": [_P, ctypes.c_char_p, ctypes.c_char_p],
    "qkx_spinor_create": [ctypes.POINTER(_P), _IP, _I, _I], "qkx_spinor_destroy": [_P],
    "qkx_spinor_create_batch": [ctypes.POINTER(_P), _IP, _I, _I, _I],
    "qkx_spinor_upload_rhs": [_P, _I, _P], "qkx_spinor_download_rhs": [_P, _I, _P],
    "qkx_spinor_upload": [_P, _P], "qkx_spinor_download": [_P, _P], "qkx_spinor_zero": [_P],
    "qkx_spinor_copy": [_P, _P],
    "qkx_dslash": [_P, _P, _P, _I, _I, _I], "qkx_twist": [_P, _P, ctypes.POINTER(Params), _I, _I],
    "qkx_twist_add_dslash": [_P, _P, _P, ctypes.POINTER(Params), _I, _I, _I],
    "qkx_tm": [_P, _P, _P, ctypes.POINTER(Params), _I, _I],
    "qkx_tm_xx": [_P, _P, _P, ctypes.POINTER(Params), _I, _I, _I],
    "qkx_tm_xx_noalloc": [_P, _P, _P, ctypes.POINTER(Params), _I, _I, _I],
    "qkx_redot": [_DP, _P, _P], "qkx_axpy": [_P, _D, _P], "qkx_xpay": [_P, _D, _P], "qkx_norm2": [_DP, _P],
    "qkx_cg": [_P, _P, _P, ctypes.POINTER(Params), ctypes.POINTER(SolverStats)],
    "qkx_solve": [_P, _P, _P, _P, ctypes.POINTER(Params), ctypes.POINTER(SolveOpts), ctypes.POINTER(SolverStats)],
    "qkx_solve_batch": [_P, _P, _P, _P, ctypes.POINTER(Params), _DP, ctypes.POINTER(SolveOpts),
                        ctypes.POINTER(SolverStats)],
    "qkx_true_residual": [_DP, _P, _P, _P, ctypes.POINTER(Params)],
    "qkx_solve_multi": [_I, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(Params),
                        ctypes.POINTER(SolveOpts), ctypes.POINTER(SolverStats)],
    "qkx_reduce_sum": [_P, _I, _I, _DP],
    "qkx_cg_record": [_P, _P, _P, ctypes.POINTER(Params), _I, ctypes.POINTER(_P), _I, _IP,
                      ctypes.POINTER(SolverStats)],
    "qkx_net_create": [ctypes.POINTER(_P), _I], "qkx_net_destroy": [_P], "qkx_net_nweights": [_P, _IP],
    "qkx_net_set_weights": [_P, _P], "qkx_net_get_weights": [_P, _P], "qkx_net_set_bf16": [_P, _I],
    "qkx_net_apply": [_P, _P, _P, _P], "qkx_net_accumulate_grad": [_P, _P, _P, _P],
    "qkx_net_apply_eo": [_P, _P, _P, _P, _I], "qkx_net_release_buffers": [_P],
    "qkx_net_set_hidden_limit": [_P, ctypes.c_ulonglong],
    "qkx_solve_learned": [_P, _P, _P, _P, _P, ctypes.POINTER(Params), ctypes.POINTER(SolveOpts), _D,
                          ctypes.POINTER(SolverStats)],
    "qkx_spinor_gaussian": [_P, ctypes.c_ulonglong],
    "qkx_net_read_grad": [_P, _P, _DP], "qkx_net_zero_grad": [_P],
    "qkx_net_adam": [_P, _D, _D, _D, _D, _I],
    "qkx_solve_gcr": [_P, _P, _P, _P, _P, ctypes.POINTER(Params), ctypes.POINTER(GcrOpts),
                      ctypes.POINTER(SolverStats)],
    "qkx_gen_gauge": [_P, _IP, _D, ctypes.c_ulonglong], "qkx_gen_z2": [_P, _IP, ctypes.c_ulonglong],
    "qkx_gen_gaussian": [_P, _IP, ctypes.c_ulonglong], "qkx_gen_point": [_P, _IP, _IP, _I, _I],
    "qkx_gen_site": [_IP, _IP, ctypes.c_ulonglong],
}
_RESTYPES = {"qkx_last_error": ctypes.c_char_p}

_lib = None


def lib():
    """Load libqkx_hip.so (in-tree).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QkxError(f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C qkxtm_mpi_amd)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def _check(rc, what=""):
    if rc not in (SUCCESS,):
        msg = lib().qkx_last_error().decode()
        raise QkxError(f"{what} failed (status {rc}): {msg}")
    return rc


def _ivec(vals):
    return (ctypes.c_int * len(vals))(*vals)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def volume(L):
    return int(L[0]) * int(L[1]) * int(L[2]) * int(L[3])


# ---------------------------------------------------------------------------
# synthetic inputs (host, no GPU): reference layout arrays of complex128
def gen_gauge(L, eps, seed):
    """U[V,4,3,3] complex128, U = exp(i eps H) reunitarised to SU(3)."""
    U = np.zeros((volume(L), 4, 3, 3), dtype=np.complex128)
    _check(lib().qkx_gen_gauge(_ptr(U), _ivec(L), eps, seed), "qkx_gen_gauge")
    return U


def gen_z2(L, seed):
    psi = np.zeros((volume(L), 4, 3), dtype=np.complex128)
    _check(lib().qkx_gen_z2(_ptr(psi), _ivec(L), seed), "qkx_gen_z2")
    return psi


def gen_gaussian(L, seed):
    psi = np.zeros((volume(L), 4, 3), dtype=np.complex128)
    _check(lib().qkx_gen_gaussian(_ptr(psi), _ivec(L), seed), "qkx_gen_gaussian")
    return psi


def gen_point(L, x0, spin, colour):
    psi = np.zeros((volume(L), 4, 3), dtype=np.complex128)
    _check(lib().qkx_gen_point(_ptr(psi), _ivec(L), _ivec(x0), spin, colour), "qkx_gen_point")
    return psi


def gen_site(L, seed):
    x0 = (ctypes.c_int * 4)()
    _check(lib().qkx_gen_site(x0, _ivec(L), seed), "qkx_gen_site")
    return [x0[i] for i in range(4)]


# ---------------------------------------------------------------------------
def init(device=0):
    _check(lib().qkx_init(device), "qkx_init")


def device_count():
    n = ctypes.c_int(0)
    lib().qkx_device_count(ctypes.byref(n))
    return n.value


def sync():
    _check(lib().qkx_sync(), "qkx_sync")


def mem_info():
    """(bytes in use, total bytes) of the current device."""
    used, total = ctypes.c_double(), ctypes.c_double()
    _check(lib().qkx_mem_info(ctypes.byref(used), ctypes.byref(total)), "qkx_mem_info")
    return used.value, total.value


def params(L, kappa, mu, twistSign=1, boundaryT=-1, tol=1e-10, maxIter=10000, reliableDelta=0.1):
    p = Params()
    for i in range(4):
        p.L[i] = L[i]
    p.kappa, p.mu, p.twistSign, p.boundaryT = kappa, mu, twistSign, boundaryT
    p.tol, p.maxIter, p.reliableDelta = tol, int(maxIter), reliableDelta
    return p


class Gauge:
    """Device gauge field (reference qkxTM::Gauge, include/lattice_util.h:186-212).

    ``boundaryT=-1`` applies the antiperiodic sign of the last time slice inside
    the kernels; pass ``+1`` for links that already carry it
    (Gauge::applyBoundaryCondition, lib/lattice_util.cpp:708-724)."""

    def __init__(self, L, prec=DOUBLE, recon=RECON_18, boundaryT=-1, host=None):
        self.L = tuple(L)
        self.prec, self.recon = prec, recon
        h = ctypes.c_void_p()
        _check(lib().qkx_gauge_create(ctypes.byref(h), _ivec(L), prec, recon, boundaryT), "qkx_gauge_create")
        self.h = h
        if host is not None:
            self.upload(host)

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        assert U.size == volume(self.L) * 36
        _check(lib().qkx_gauge_upload(self.h, _ptr(U)), "qkx_gauge_upload")

    def convert_from(self, src: "Gauge"):
        _check(lib().qkx_gauge_convert(self.h, src.h), "qkx_gauge_convert")

    def download(self):
        """FP64 recon-18 links as stored, complex128 [V, 4, 3, 3]."""
        U = np.empty((volume(self.L), 4, 3, 3), dtype=np.complex128)
        _check(lib().qkx_gauge_download(self.h, _ptr(U)), "qkx_gauge_download")
        return U

    def su3_deviation(self):
        """max |row2 - conj(row0 x row1)| over the links (FP64 18-real fields)."""
        d = ctypes.c_double()
        _check(lib().qkx_gauge_su3_deviation(self.h, ctypes.byref(d)), "qkx_gauge_su3_deviation")
        return d.value

    def read_lime(self, path):
        """Links from an ILDG/LIME file (qkxTM_MPI_getGaugeLime, lib/read_conf_lime.cpp:111)."""
        _check(lib().qkx_gauge_read_lime(self.h, os.fsencode(path)), "qkx_gauge_read_lime")

    def write_lime(self, path, message=""):
        """ILDG/LIME file of the links (qkxTM_writeGaugeLime, lib/write_conf_lime.cpp:49)."""
        _check(lib().qkx_gauge_write_lime(self.h, os.fsencode(path), message.encode()), "qkx_gauge_write_lime")

    # ---- source preparation, named after the reference's Gauge methods ------
    def APE_smearing(self, U_tmp: "Gauge", nsmear, alpha):
        """Gauge::APE_smearing (lib/lattice_util.cpp:950): this field's spatial
        links become nsmear APE steps of U_tmp's, its temporal links U_tmp's."""
        _check(lib().qkx_ape_smear(self.h, U_tmp.h, int(nsmear), float(alpha)), "qkx_ape_smear")

    def projectSU3(self):
        """qcd_projectSU33d (lib/blas.cpp:6) on the spatial links, in place."""
        _check(lib().qkx_gauge_project_su3(self.h), "qkx_gauge_project_su3")

    def calculatePlaquette(self):
        v = ctypes.c_double()
        _check(lib().qkx_plaquette(ctypes.byref(v), self.h), "qkx_plaquette")
        return v.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qkx_gauge_destroy(self.h)
            self.h = None


class Spinor:
    """Device spinor field (reference qkxTM::Vector, include/lattice_util.h:80-156)."""

    def __init__(self, L, prec=DOUBLE, subset=FULL, host=None, nrhs=1):
        self.L = tuple(L)
        self.prec, self.subset, self.nrhs = prec, subset, int(nrhs)
        h = ctypes.c_void_p()
        if self.nrhs == 1:
            _check(lib().qkx_spinor_create(ctypes.byref(h), _ivec(L), prec, subset), "qkx_spinor_create")
        else:
            _check(lib().qkx_spinor_create_batch(ctypes.byref(h), _ivec(L), prec, subset, self.nrhs),
                   "qkx_spinor_create_batch")
        self.h = h
        if host is not None:
            self.upload(host)

    def upload(self, psi):
        """Whole field from (V, 4, 3) complex -- or (nrhs, V, 4, 3) for a batch."""
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        if self.nrhs == 1:
            assert psi.size == volume(self.L) * 12
            _check(lib().qkx_spinor_upload(self.h, _ptr(psi)), "qkx_spinor_upload")
        else:
            psi = psi.reshape(self.nrhs, volume(self.L) * 12)
            for r in range(self.nrhs):
                self.upload_rhs(r, psi[r])

    def download(self):
        if self.nrhs > 1:
            return np.stack([self.download_rhs(r) for r in range(self.nrhs)])
        out = np.zeros((volume(self.L), 4, 3), dtype=np.complex128)
        _check(lib().qkx_spinor_download(self.h, _ptr(out)), "qkx_spinor_download")
        return out

    def upload_rhs(self, r, psi):
        psi = np.ascontiguousarray(psi, dtype=np.complex128)
        assert psi.size == volume(self.L) * 12
        _check(lib().qkx_spinor_upload_rhs(self.h, int(r), _ptr(psi)), "qkx_spinor_upload_rhs")

    def download_rhs(self, r):
        out = np.zeros((volume(self.L), 4, 3), dtype=np.complex128)
        _check(lib().qkx_spinor_download_rhs(self.h, int(r), _ptr(out)), "qkx_spinor_download_rhs")
        return out

    def zero(self):
        _check(lib().qkx_spinor_zero(self.h), "qkx_spinor_zero")

    def copy(self, src: "Spinor"):
        _check(lib().qkx_spinor_copy(self.h, src.h), "qkx_spinor_copy")

    def gaussian(self, seed):
        """Device-generated Gaussian noise (variance 1 per real component)."""
        _check(lib().qkx_spinor_gaussian(self.h, int(seed)), "qkx_spinor_gaussian")

    def norm2(self):
        v = ctypes.c_double()
        _check(lib().qkx_norm2(ctypes.byref(v), self.h), "qkx_norm2")
        return v.value

    def rescale(self, a):
        """Vector::rescale (lib/lattice_util.cpp:2314): every member times a."""
        _check(lib().qkx_spinor_rescale(self.h, float(a)), "qkx_spinor_rescale")

    def Gaussian_smearing(self, vec_tmp: "Spinor", gauge: Gauge, nsmear, alpha, t=-1):
        """Vector::Gaussian_smearing (lib/lattice_util.cpp:1080): this field
        becomes nsmear smearing steps of vec_tmp; t >= 0 smears only that time
        slice and zeroes the rest."""
        _check(lib().qkx_gauss_smear(self.h, vec_tmp.h, gauge.h, int(nsmear), float(alpha), int(t)),
               "qkx_gauss_smear")

    # ---- operators, named after the reference's Vector methods -------------
    def applyDslash(self, psi, u, basis=UKQCD):
        _check(lib().qkx_dslash(self.h, psi.h, u.h, basis, 0, PARITY_ALL), "applyDslash")

    def applyDslashDag(self, psi, u, basis=UKQCD):
        _check(lib().qkx_dslash(self.h, psi.h, u.h, basis, 1, PARITY_ALL), "applyDslashDag")

    def applyDslash_xx(self, psi, u, parity, basis=UKQCD):
        _check(lib().qkx_dslash(self.h, psi.h, u.h, basis, 0, parity), "applyDslash_xx")

    def twist(self, psi, p, kind, basis=UKQCD):
        _check(lib().qkx_twist(self.h, psi.h, ctypes.byref(p), basis, kind), "qkx_twist")

    def twist_add_dslash(self, dslash, psi, p, kind, basis=UKQCD, parity=PARITY_ALL):
        _check(lib().qkx_twist_add_dslash(self.h, dslash.h, psi.h, ctypes.byref(p), basis, kind, parity),
               "qkx_twist_add_dslash")

    def tm_op(self, psi, u, p, op=OP_TM, basis=UKQCD):
        _check(lib().qkx_tm(self.h, psi.h, u.h, ctypes.byref(p), basis, op), "qkx_tm")

    def tm(self, psi, u, p, basis=UKQCD):
        self.tm_op(psi, u, p, OP_TM, basis)

    def tmDag(self, psi, u, p):
        self.tm_op(psi, u, p, OP_TM_DAG)

    def tmDagtm(self, psi, u, p):
        self.tm_op(psi, u, p, OP_TM_DAG_TM)

    def Wilson(self, psi, u, p, basis=UKQCD):
        self.tm_op(psi, u, p, OP_WILSON, basis)

    def tm_xx(self, psi, u, p, parity, basis=UKQCD):
        _check(lib().qkx_tm_xx(self.h, psi.h, u.h, ctypes.byref(p), basis, parity, 0), "tm_xx")

    def tm_xxDagTm_xx(self, psi, u, p, parity, basis=UKQCD):
        _check(lib().qkx_tm_xx(self.h, psi.h, u.h, ctypes.byref(p), basis, parity, 1), "tm_xxDagTm_xx")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qkx_spinor_destroy(self.h)
            self.h = None


# ---- BLAS (include/blas_qkxTM.h:3-5) ---------------------------------------
def reCdotXdagY(x: Spinor, y: Spinor):
    v = ctypes.c_double()
    _check(lib().qkx_redot(ctypes.byref(v), x.h, y.h), "reCdotXdagY")
    return v.value


def X_eq_X_p_aY(x: Spinor, a: float, y: Spinor):
    _check(lib().qkx_axpy(x.h, a, y.h), "X_eq_X_p_aY")


def X_eq_aX_p_Y(x: Spinor, a: float, y: Spinor):
    _check(lib().qkx_xpay(x.h, a, y.h), "X_eq_aX_p_Y")


# ---- solvers -----------------------------------------------------------------
class Propagator(Spinor):
    """Device propagator (reference qkxTM::Propagator, include/lattice_util.h:158-185):
    a batch of 12 FP64 fields, member 3 nu + c2 = column (source spin nu,
    source colour c2).  The solution of ``solve_batch`` on the 12 spin-colour
    point sources (``point_sources``) is one."""

    def __init__(self, L, host=None):
        super().__init__(L, DOUBLE, FULL, nrhs=12)
        if host is not None:
            self.upload_prop(host)

    def upload_prop(self, M):
        M = np.ascontiguousarray(M, dtype=np.complex128)
        assert M.size == volume(self.L) * 144
        _check(lib().qkx_prop_upload(self.h, _ptr(M)), "qkx_prop_upload")

    def download_prop(self):
        """complex128 [V, 4, 4, 3, 3], the reference's Propagator::M layout."""
        M = np.empty((volume(self.L), 4, 4, 3, 3), dtype=np.complex128)
        _check(lib().qkx_prop_download(self.h, _ptr(M)), "qkx_prop_download")
        return M

    def absorbVector(self, phi: Spinor, nu, c2):
        _check(lib().qkx_prop_absorb(self.h, phi.h, int(nu), int(c2)), "qkx_prop_absorb")

    copyVector = absorbVector

    def rotateToPhysicalBasePlus(self):
        _check(lib().qkx_prop_rotate_physical(self.h, 1), "qkx_prop_rotate_physical")

    def rotateToPhysicalBaseMinus(self):
        _check(lib().qkx_prop_rotate_physical(self.h, -1), "qkx_prop_rotate_physical")


def read_lime(path, L):
    """Host links complex128 [V, 4, 3, 3] from an ILDG/LIME file."""
    U = np.empty((volume(L), 4, 3, 3), dtype=np.complex128)
    _check(lib().qkx_lime_read_gauge(os.fsencode(path), _ivec(L), _ptr(U)), "qkx_lime_read_gauge")
    return U


def write_lime(path, L, U, message=""):
    """ILDG/LIME file (the reference writer's three records) from host links."""
    U = np.ascontiguousarray(U, dtype=np.complex128)
    assert U.size == volume(L) * 36
    _check(lib().qkx_lime_write_gauge(os.fsencode(path), _ivec(L), _ptr(U), message.encode()),
           "qkx_lime_write_gauge")


def point_sources(L, x0, into=None):
    """The 12 spin-colour point sources at site x0 as a 12-member batch (member
    3 spin + colour), the right-hand sides of a propagator."""
    V = volume(L)
    iv = ((x0[3] * L[2] + x0[2]) * L[1] + x0[1]) * L[0] + x0[0]
    b = into if into is not None else Spinor(L, nrhs=12)
    psi = np.zeros((V, 4, 3), np.complex128)
    for r in range(12):
        psi[iv, r // 3, r % 3] = 1.0
        b.upload_rhs(r, psi)
        psi[iv, r // 3, r % 3] = 0.0
    return b


class CG:
    """Reference CG (include/solvers.h:3-22, lib/solvers.cpp:8-81): solves
    D^dag D out = in on the full lattice; ``out`` accumulates onto its value."""

    def __init__(self, p: Params, gauge: Gauge):
        self.p, self.u = p, gauge
        self.stats = SolverStats()

    def __call__(self, out: Spinor, inp: Spinor):
        rc = lib().qkx_cg(out.h, inp.h, self.u.h, ctypes.byref(self.p), ctypes.byref(self.stats))
        if rc == NOT_CONVERGED:
            return self.stats  # the reference only warns (lib/solvers.cpp:68-70)
        _check(rc, "CG")
        return self.stats


def solve(x: Spinor, b: Spinor, u: Gauge, p: Params, eo=True, inner_prec=DOUBLE, inner_recon=RECON_12,
          delta=0.1, tol=1e-10, maxIter=100000, u_inner: Gauge | None = None, allow_not_converged=False):
    """Solve D x = b to true relative residual ``tol``; returns SolverStats."""
    o = SolveOpts(int(eo), inner_prec, inner_recon, delta, 50, int(maxIter), tol)
    st = SolverStats()
    rc = lib().qkx_solve(x.h, b.h, u.h, u_inner.h if u_inner else None, ctypes.byref(p), ctypes.byref(o),
                         ctypes.byref(st))
    if rc == NOT_CONVERGED and allow_not_converged:
        return st
    _check(rc, "qkx_solve")
    return st


def solve_batch(x: Spinor, b: Spinor, u: Gauge, p: Params, mu=None, eo=True, inner_prec=DOUBLE,
                inner_recon=RECON_12, delta=0.1, tol=1e-10, maxIter=100000, u_inner: Gauge | None = None,
                allow_not_converged=False):
    """Solve D_r x_r = b_r for every right-hand side of a batch in lockstep
    (per-rhs twisted mass ``mu`` optional); returns a list of SolverStats."""
    o = SolveOpts(int(eo), inner_prec, inner_recon, delta, 50, int(maxIter), tol)
    st = (SolverStats * b.nrhs)()
    mu_arr = None
    if mu is not None:
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        assert mu.size == b.nrhs
        mu_arr = mu.ctypes.data_as(_DP)
    rc = lib().qkx_solve_batch(x.h, b.h, u.h, u_inner.h if u_inner else None, ctypes.byref(p), mu_arr,
                               ctypes.byref(o), st)
    if not (rc == NOT_CONVERGED and allow_not_converged):
        _check(rc, "qkx_solve_batch")
    return list(st)


def solve_multi(xs, bs, us, ps, inner_prec=DOUBLE, delta=0.1, tol=1e-10, maxIter=100000, allow_not_converged=False):
    """Solve D_i x_i = b_i for independent instances -- each on its own
    lattice, gauge field and parameters -- in lockstep, even-odd, every CG
    launch serving all of them; returns a list of SolverStats."""
    n = len(xs)
    assert len(bs) == n and len(us) == n and len(ps) == n
    o = SolveOpts(1, inner_prec, RECON_12, delta, 50, int(maxIter), tol)
    st = (SolverStats * n)()
    xa = (_P * n)(*[f.h.value for f in xs])
    ba = (_P * n)(*[f.h.value for f in bs])
    ua = (_P * n)(*[g.h.value for g in us])
    pa = (Params * n)(*ps)
    rc = lib().qkx_solve_multi(n, xa, ba, ua, pa, ctypes.byref(o), st)
    if not (rc == NOT_CONVERGED and allow_not_converged):
        _check(rc, "qkx_solve_multi")
    return list(st)


def cg_record(x: Spinor, b: Spinor, u: Gauge, p: Params, every: int, store):
    """Reference CG (as CG.__call__) that copies the iterate into store[i]
    after every `every` iterations; returns (SolverStats, number stored)."""
    st = SolverStats()
    arr = (ctypes.c_void_p * max(1, len(store)))(*[f.h.value for f in store])
    n = ctypes.c_int()
    rc = lib().qkx_cg_record(x.h, b.h, u.h, ctypes.byref(p), int(every), arr, len(store), ctypes.byref(n),
                             ctypes.byref(st))
    if rc != NOT_CONVERGED:
        _check(rc, "qkx_cg_record")
    return st, n.value


class Net:
    """Learned preconditioner M = L2 L1: two gauge-equivariant
    parallel-transport convolutions, 4 -> ch -> 4 colour-vector channels.
    Weights are complex arrays W1 [9, ch, 4] and W2 [9, 4, ch]; path 0 is the
    identity, 1+2mu the forward and 2+2mu the backward transport along mu."""

    def __init__(self, ch=16, W1=None, W2=None, seed=0, noise=1e-2):
        self.ch = ch
        h = ctypes.c_void_p()
        _check(lib().qkx_net_create(ctypes.byref(h), ch), "qkx_net_create")
        self.h = h
        if W1 is None:
            W1, W2 = Net.identity_weights(ch, seed, noise)
        self.set_weights(W1, W2)

    @staticmethod
    def identity_weights(ch, seed=0, noise=1e-2):
        """M = 1 plus small complex noise (hidden channels 0..3 carry the spins)."""
        r = np.random.default_rng(seed)
        W1 = noise * (r.standard_normal((9, ch, 4)) + 1j * r.standard_normal((9, ch, 4)))
        W2 = noise * (r.standard_normal((9, 4, ch)) + 1j * r.standard_normal((9, 4, ch)))
        W1[0, :4, :4] += np.eye(4)
        W2[0, :4, :4] += np.eye(4)
        return W1, W2

    def set_weights(self, W1, W2):
        w = np.concatenate([np.asarray(W1, np.complex64).ravel(), np.asarray(W2, np.complex64).ravel()])
        assert w.size == 2 * 9 * 4 * self.ch
        w = np.ascontiguousarray(w.view(np.float32))
        _check(lib().qkx_net_set_weights(self.h, _ptr(w)), "qkx_net_set_weights")

    def get_weights(self):
        n = 9 * 4 * self.ch
        w = np.zeros(4 * n, np.float32)
        _check(lib().qkx_net_get_weights(self.h, _ptr(w)), "qkx_net_get_weights")
        c = w.view(np.complex64).astype(np.complex128)
        return c[:n].reshape(9, self.ch, 4), c[n:].reshape(9, 4, self.ch)

    def set_bf16(self, on=True):
        _check(lib().qkx_net_set_bf16(self.h, int(on)), "qkx_net_set_bf16")

    def apply(self, out: Spinor, inp: Spinor, u_net: Gauge):
        _check(lib().qkx_net_apply(self.h, out.h, inp.h, u_net.h), "qkx_net_apply")

    def release_buffers(self):
        """Free the work buffers (training features, even-odd hidden channels)."""
        _check(lib().qkx_net_release_buffers(self.h), "qkx_net_release_buffers")

    def set_hidden_limit(self, nbytes):
        """Bound the even-odd hidden buffer (0: the 4 GiB default)."""
        _check(lib().qkx_net_set_hidden_limit(self.h, int(nbytes)), "qkx_net_set_hidden_limit")

    def apply_eo(self, out: Spinor, inp: Spinor, u_net: Gauge, adjoint=False):
        """out = N_e inp (or N_e^dag inp): the network on even-parity fields."""
        _check(lib().qkx_net_apply_eo(self.h, out.h, inp.h, u_net.h, int(adjoint)), "qkx_net_apply_eo")

    def accumulate_grad(self, r: Spinor, e: Spinor, u_net: Gauge):
        _check(lib().qkx_net_accumulate_grad(self.h, r.h, e.h, u_net.h), "qkx_net_accumulate_grad")

    def read_grad(self):
        n = 9 * 4 * self.ch
        g = np.zeros(4 * n, np.float32)
        loss = ctypes.c_double()
        _check(lib().qkx_net_read_grad(self.h, _ptr(g), ctypes.byref(loss)), "qkx_net_read_grad")
        c = g.view(np.complex64).astype(np.complex128)
        return c[:n].reshape(9, self.ch, 4), c[n:].reshape(9, 4, self.ch), loss.value

    def read_loss(self):
        loss = ctypes.c_double()
        _check(lib().qkx_net_read_grad(self.h, None, ctypes.byref(loss)), "qkx_net_read_grad")
        return loss.value

    def zero_grad(self):
        _check(lib().qkx_net_zero_grad(self.h), "qkx_net_zero_grad")

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8, count=1):
        _check(lib().qkx_net_adam(self.h, lr, beta1, beta2, eps, int(count)), "qkx_net_adam")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qkx_net_destroy(self.h)
            self.h = None


def solve_gcr(x: Spinor, b: Spinor, u: Gauge, p: Params, net: "Net | None" = None, u_net: Gauge | None = None,
              restart=8, tol=1e-10, maxIter=10000, stall=0.0, allow_not_converged=False):
    """Flexible GCR on D x = b with the network as right preconditioner."""
    o = GcrOpts(int(restart), int(maxIter), tol, stall)
    st = SolverStats()
    rc = lib().qkx_solve_gcr(x.h, b.h, u.h, u_net.h if u_net else None, net.h if net else None, ctypes.byref(p),
                             ctypes.byref(o), ctypes.byref(st))
    if not (rc == NOT_CONVERGED and allow_not_converged):
        _check(rc, "qkx_solve_gcr")
    return st


def solve_learned(x: Spinor, b: Spinor, u: Gauge, p: Params, net: "Net | None", u_net: Gauge | None,
                  inner_prec=SINGLE, delta=0.1, tol=1e-10, maxIter=100000, stall=0.5, allow_not_converged=False):
    """The learned, mixed-precision, multi-RHS solver: even-odd CG on
    (Mh N)^dag (Mh N) y = (Mh N)^dag bh, x = N y, for every right-hand side of
    the batch, falling back to the plain CG when a reliable-update cycle cuts
    the worst true residual by less than ``stall``; net=None is solve_batch.
    Returns a list of SolverStats."""
    o = SolveOpts(1, inner_prec, RECON_12, delta, 50, int(maxIter), tol)
    st = (SolverStats * b.nrhs)()
    rc = lib().qkx_solve_learned(x.h, b.h, u.h, u_net.h if u_net else None, net.h if net else None,
                                 ctypes.byref(p), ctypes.byref(o), float(stall), st)
    if not (rc == NOT_CONVERGED and allow_not_converged):
        _check(rc, "qkx_solve_learned")
    return list(st)


def true_residual(x: Spinor, b: Spinor, u: Gauge, p: Params):
    v = ctypes.c_double()
    _check(lib().qkx_true_residual(ctypes.byref(v), x.h, b.h, u.h, ctypes.byref(p)), "qkx_true_residual")
    return v.value


class Timer:
    """HIP events on the library stream."""

    def start(self):
        _check(lib().qkx_timer_start(), "qkx_timer_start")

    def stop(self):
        v = ctypes.c_double()
        _check(lib().qkx_timer_stop(ctypes.byref(v)), "qkx_timer_stop")
        return v.value


def prof_enable(on=True):
    _check(lib().qkx_prof_enable(int(on)), "qkx_prof_enable")


def prof_read(prec=-1):
    """(launches, ms, bytes) of the instrumented hopping kernels since the last
    reset, optionally only those on spinors of precision ``prec``."""
    n, ms, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    _check(lib().qkx_prof_read_prec(prec, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)), "qkx_prof_read")
    return n.value, ms.value, by.value


def prof_reset():
    _check(lib().qkx_prof_reset(), "qkx_prof_reset")


def set_cg_batch(iterations=16):
    _check(lib().qkx_set_cg_batch(iterations), "qkx_set_cg_batch")


def set_cg_fuse(mode=2):
    """CG update fused into the normal operator's last hop (1 or True), in a
    separate kernel (0 or False), or fused with x += alpha p moved into the
    direction kernel where x is FP32 (2, the default)."""
    _check(lib().qkx_set_cg_fuse(int(mode)), "qkx_set_cg_fuse")


def set_cg_graph(mode=3):
    """Replay CG iterations from captured HIP graphs: mode bit 0 the
    single-RHS solvers, bit 1 the batched solvers (library default 0);
    True means 3, False 0."""
    mode = 3 if mode is True else 0 if mode is False else int(mode)
    _check(lib().qkx_set_cg_graph(mode), "qkx_set_cg_graph")


def set_rhs_tile(rt=0):
    _check(lib().qkx_set_rhs_tile(int(rt)), "qkx_set_rhs_tile")


def set_autotune(mode=1):
    """Hop autotuner: 0 off, 1 batched launches (default), 2 every launch."""
    _check(lib().qkx_set_autotune(int(mode)), "qkx_set_autotune")


def autotune_report():
    """The tuned launch kinds and their chosen shapes, one line each."""
    n = ctypes.c_int()
    _check(lib().qkx_autotune_report(None, 0, ctypes.byref(n)), "qkx_autotune_report")
    buf = ctypes.create_string_buffer(max(1, n.value))
    _check(lib().qkx_autotune_report(buf, n.value, ctypes.byref(n)), "qkx_autotune_report")
    return buf.value.decode()


def set_batch_link_sharing(mode=1):
    """Links shared between the RHS threads of FP16 batches: 0 never, 1 when the
    launch is large enough (default), 2 always."""
    _check(lib().qkx_set_batch_link_sharing(int(mode)), "qkx_set_batch_link_sharing")


def set_site_order(mode=2):
    """Site order of the hopping kernels: 1 (or True) blocked, 0 (or False)
    checkerboard index order, 2 (default) blocked for the full-lattice FP64
    hops only."""
    _check(lib().qkx_set_site_order(int(mode)), "qkx_set_site_order")


def set_neighbour_staging(mode=0):
    """Single-RHS FP16 hops on one parity stage the x and y neighbour spinors
    in LDS -- 1: per workgroup, 2: per wave -- or gather them per thread (0,
    the default); the results are identical."""
    _check(lib().qkx_set_neighbour_staging(int(mode)), "qkx_set_neighbour_staging")


def set_tuning(xcd_remap=True, block=256, depth64=2, depth32=2):
    _check(lib().qkx_set_tuning(int(xcd_remap), block, depth64, depth32), "qkx_set_tuning")
