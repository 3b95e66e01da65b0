"""The C ABI library loads without a GPU and exports every function that
include/qkx_hip.h declares; the C++ drop-in library exports the reference's
Vector / Gauge / CG / BLAS methods.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "qkx_hip.h")
LIB = os.path.join(ROOT, "qkxtm_mpi_amd", "libqkx_hip.so")
DROPIN = os.path.join(ROOT, "qkxtm_mpi_amd", "host", "libqkxTM.so")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qkx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 35
    for must in ("qkx_tm", "qkx_tm_xx", "qkx_dslash", "qkx_twist", "qkx_cg", "qkx_solve", "qkx_redot",
                 "qkx_axpy", "qkx_xpay", "qkx_norm2"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    import qkxtm_mpi_amd as q
    missing = [n for n in declared_functions() if n not in q.SIGNATURES]
    assert not missing, missing
    q.lib()   # binds every signature


def test_dropin_exports_reference_methods():
    out = subprocess.run(["nm", "-DC", "--defined-only", DROPIN], capture_output=True, text=True, check=True).stdout
    for sym in ("qkxTM::Vector::tm(", "qkxTM::Vector::tmDag(", "qkxTM::Vector::tmDagtm(", "qkxTM::Vector::tm_xx(",
                "qkxTM::Vector::tm_xxDagTm_xx(", "qkxTM::Vector::tm_chiral(", "qkxTM::Vector::applyDslash(",
                "qkxTM::Vector::applyDslash_xx(", "qkxTM::Vector::applyTwistInv(", "qkxTM::Vector::Wilson(",
                "qkxTM::Gauge::applyBoundaryCondition(", "CG::operator()(", "CG::CG(", "reCdotXdagY(",
                "X_eq_X_p_aY(", "X_eq_aX_p_Y("):
        assert sym in out, sym


def test_dropin_headers_compile_as_cxx98(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "qkxtm/qkxTM.h"\n#include "qkxtm/lattice_util.h"\n#include "qkxtm/solvers.h"\n'
                   '#include "qkxtm/blas_qkxTM.h"\nint main(){ LatticeInfo l; Complex c; c.real() = 1.0;'
                   ' return (int)c.real() - 1 + l.P[0] - 1; }\n')
    r = subprocess.run(["g++", "-std=gnu++98", "-fsyntax-only", "-I" + os.path.join(ROOT, "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
