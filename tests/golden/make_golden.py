"""Write the golden fixtures in tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness (the reference's lib/{lattice_util,util,blas,
blas_qkxTM,solvers}.cpp compiled by oracle/ref/Makefile) on synthetic inputs
from the library's generator and stores, per operator output y:
  * |y|^2,
  * <r_k, y> for four seeded Gaussian vectors r_k (a random projection: any
    wrong site, spin, colour or sign changes it),
  * the spinors of 8 seeded sites,
and, for a handful of operators on the 4^4 lattice, the whole output vector.

Inputs are not stored: they are regenerated from (L, eps, seeds) by the
library's counter-hash generator; a hash of them is kept to detect drift.

usage (in the build container, after __graft_entry__.build()):
    python tests/golden/make_golden.py [ops] [cfg2] [smear] [prop]     (default: all)
"""
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
import qkxtm_mpi_amd as q  # noqa: E402  (generator only, host code)
sys.path.insert(0, os.path.join(ROOT, "tests"))
from golden_util import (input_hash, probes, summarize, gauge_probes, summarize_gauge,  # noqa: E402
                         generic_matrices)

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))

FULL_OPS = ["tm", "tmDag", "tm_chiral", "tm_xx_EE", "tm_xxDagTm_xx_OO", "tm_xx_chiral_OO"]

# (tag, L, eps, kappa, mu, twistSign, boundaryT, gauge seed, psi seed, phi seed)
CASES = [
    ("a4444", (4, 4, 4, 4), 0.4, 0.156, 0.01, +1, -1, 11, 12, 13),
    ("b4648", (4, 6, 4, 8), 0.3, 0.1632, 0.004, -1, +1, 21, 22, 23),
    ("c4648", (4, 6, 4, 8), 0.5, 0.17, 0.05, +1, -1, 31, 32, 33),
]
# config-1 of the benchmark set: 8^3x16, eps 0.4, kappa 0.156, mu 0.01, Z2 source
CG_CASE = ("cfg1", (8, 8, 8, 16), 0.4, 0.156, 0.01, +1, -1, 1, 2, 3)
# config-2 (the headline solve): 24^3x48, eps 0.3, kappa 0.1632, mu 0.004, Z2
# source, gauge seed 1 and source seed 2 as in bench.py; 64 sampled sites
CG_CASE2 = ("cfg2", (24, 24, 24, 48), 0.3, 0.1632, 0.004, +1, -1, 1, 2, 3)


def write_input(path, L, kappa, mu, ts, bt, tol, maxit, U, psi, phi):
    with open(path, "wb") as f:
        f.write(struct.pack("<4i", *L))
        f.write(struct.pack("<dd", kappa, mu))
        f.write(struct.pack("<ii", ts, bt))
        f.write(struct.pack("<d", tol))
        f.write(struct.pack("<i", maxit))
        f.write(U.tobytes())
        f.write(psi.tobytes())
        f.write(phi.tobytes())


def read_records(path):
    b = open(path, "rb").read()
    o, recs = 0, {}
    while o < len(b):
        name = b[o:o + 48].split(b"\0")[0].decode()
        n = struct.unpack("<q", b[o + 48:o + 56])[0]
        o += 56
        recs[name] = np.frombuffer(b[o:o + 8 * n], dtype=np.float64).copy()
        o += 8 * n
    return recs


def run_ops_case(case, tmpdir):
    tag, L, eps, kappa, mu, ts, bt, sg, sp, sf = case
    U = q.gen_gauge(L, eps, sg)
    psi = q.gen_gaussian(L, sp)
    phi = q.gen_gaussian(L, sf)
    fin = os.path.join(tmpdir, "in.bin")
    fout = os.path.join(tmpdir, "out.bin")
    write_input(fin, L, kappa, mu, ts, bt, 1e-10, 100, U, psi, phi)
    subprocess.run([HARNESS, "ops", fin, fout], check=True, stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    R, sites = probes(L)
    data = {"L": np.array(L), "params": np.array([eps, kappa, mu, ts, bt, sg, sp, sf]),
            "input_hash": np.array(input_hash(U, psi, phi)), "sites": sites}
    names = []
    for name, v in recs.items():
        if v.size == 1:
            data["scalar_" + name] = v
            continue
        y = v.view(np.complex128)
        s = summarize(y, R, sites)
        data["norm2_" + name] = np.array(s["norm2"])
        data["proj_" + name] = s["proj"]
        data["site_" + name] = s["sites"]
        if tag == "a4444" and name in FULL_OPS:
            data["full_" + name] = y.reshape(-1, 4, 3)
        names.append(name)
    data["names"] = np.array(names)
    np.savez_compressed(os.path.join(OUT, f"ops_{tag}.npz"), **data)
    print(f"ops_{tag}.npz: {len(names)} operator outputs")


def run_cg_case(tmpdir, case=CG_CASE, nsites=8):
    tag, L, eps, kappa, mu, ts, bt, sg, sp, sf = case
    U = q.gen_gauge(L, eps, sg)
    src = q.gen_z2(L, sp)
    phi = np.zeros_like(src)
    fin = os.path.join(tmpdir, "in.bin")
    fout = os.path.join(tmpdir, "out.bin")
    # ops on the config-1 field (tm of the Z2 source)
    write_input(fin, L, kappa, mu, ts, bt, 1e-14, 20000, U, src, phi)
    subprocess.run([HARNESS, "ops", fin, fout], check=True, stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    os.remove(fout)
    R, sites = probes(L, nsites)
    data = {"L": np.array(L), "params": np.array([eps, kappa, mu, ts, bt, sg, sp, sf]),
            "input_hash": np.array(input_hash(U, src)), "sites": sites}
    for name in ("tm", "tmDag", "tmDagtm", "tm_xx_EE", "tm_xxDagTm_xx_EE", "tm_chiral", "tm_xxDagTm_xx_OO"):
        s = summarize(recs[name].view(np.complex128), R, sites)
        data["norm2_" + name] = np.array(s["norm2"])
        data["proj_" + name] = s["proj"]
        data["site_" + name] = s["sites"]
    # reference CG on D^dag D x = D^dag src, tol 1e-14 (bin/test_invert_prop.cpp:91-94)
    subprocess.run([HARNESS, "cg", fin, fout], check=True, stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    x = recs["cg_x"].view(np.complex128)
    Dx = recs["cg_Dx"].view(np.complex128)
    s = summarize(x, R, sites)
    data["norm2_cg_x"] = np.array(s["norm2"])
    data["proj_cg_x"] = s["proj"]
    data["site_cg_x"] = s["sites"]
    res = np.linalg.norm(Dx - src.ravel()) / np.linalg.norm(src.ravel())
    data["cg_true_residual"] = np.array(res)
    np.savez_compressed(os.path.join(OUT, f"cg_{tag}.npz"), **data)
    print(f"cg_{tag}.npz: reference CG true residual {res:.3e}")


# smearing: (tag, L, eps, boundaryT, gauge seed, psi seed, nAPE, alphaAPE, nGauss, alphaGauss)
# with the reference driver's alphas (bin/test_gauss_smearing.cpp:31-34) and
# other ones, odd and even step counts
SMEAR_CASES = [
    ("s4648", (4, 6, 4, 8), 0.4, -1, 41, 42, 3, 0.5, 5, 4.0),
    ("s6448", (6, 4, 4, 8), 0.3, +1, 51, 52, 4, 0.7, 2, 2.5),
]
# projection of generic (non-unitary) matrices: (tag, L, seed)
PROJ_CASE = ("p4444", (4, 4, 4, 4), 61)


def _gauge_summary(data, key, U, R, sites):
    s = summarize_gauge(U, R, sites)
    data["gnorm2_" + key] = np.array(s["norm2"])
    data["gproj_" + key] = s["proj"]
    data["gsite_" + key] = s["sites"]


def run_smear_case(case, tmpdir):
    tag, L, eps, bt, sg, sp, na, aa, ng, ag = case
    U = q.gen_gauge(L, eps, sg)
    psi = q.gen_gaussian(L, sp)
    fin = os.path.join(tmpdir, "in.bin")
    fout = os.path.join(tmpdir, "out.bin")
    write_input(fin, L, 0.1, 0.0, 1, bt, 1e-10, 1, U, psi, np.zeros_like(psi))
    subprocess.run([HARNESS, "smear", fin, fout, str(na), repr(aa), str(ng), repr(ag)], check=True,
                   stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    R, sites = probes(L)
    GR, gsites = gauge_probes(L)
    data = {"L": np.array(L), "params": np.array([eps, bt, sg, sp, na, aa, ng, ag]),
            "input_hash": np.array(input_hash(U, psi)), "sites": sites, "gsites": gsites}
    data["plaq_in"] = recs["plaq_in"]
    data["plaq_ape"] = recs["plaq_ape"]
    _gauge_summary(data, "ape", recs["ape"].view(np.complex128), GR, gsites)
    for name in ("gauss", "gauss_ape"):
        s = summarize(recs[name].view(np.complex128), R, sites)
        data["norm2_" + name] = np.array(s["norm2"])
        data["proj_" + name] = s["proj"]
        data["site_" + name] = s["sites"]
    np.savez_compressed(os.path.join(OUT, f"smear_{tag}.npz"), **data)
    print(f"smear_{tag}.npz: plaquette {recs['plaq_in'][0]:.12f} -> {recs['plaq_ape'][0]:.12f}")


def run_proj_case(tmpdir):
    tag, L, seed = PROJ_CASE
    M = generic_matrices(L, seed)
    psi = np.zeros((M.shape[0], 4, 3), np.complex128)
    fin = os.path.join(tmpdir, "in.bin")
    fout = os.path.join(tmpdir, "out.bin")
    write_input(fin, L, 0.1, 0.0, 1, 1, 1e-10, 1, M, psi, psi)
    subprocess.run([HARNESS, "proj", fin, fout], check=True, stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    GR, gsites = gauge_probes(L)
    gsites = np.union1d(gsites, [0, 1])
    data = {"L": np.array(L), "seed": np.array(seed), "input_hash": np.array(input_hash(M)), "gsites": gsites}
    _gauge_summary(data, "proj", recs["proj"].view(np.complex128), GR, gsites)
    np.savez_compressed(os.path.join(OUT, f"proj_{tag}.npz"), **data)
    print(f"proj_{tag}.npz")


# propagator assembly and rotations: (tag, L, kappa, psi seed, phi seed)
PROP_CASE = ("r4448", (4, 4, 4, 8), 0.1632, 71, 72)


def prop_columns(psi, phi):
    """The 12 columns the harness builds: column (nu, c2) = (nu + 1) psi + (c2 + 1) i phi,
    as a propagator [V, 4, 4, 3, 3]."""
    M = np.zeros(psi.shape[:1] + (4, 4, 3, 3), np.complex128)
    for nu in range(4):
        for c2 in range(3):
            M[:, :, nu, :, c2] = (nu + 1) * psi + (c2 + 1) * 1j * phi
    return M


def run_prop_case(tmpdir):
    tag, L, kappa, sp, sf = PROP_CASE
    psi, phi = q.gen_gaussian(L, sp), q.gen_gaussian(L, sf)
    fin = os.path.join(tmpdir, "in.bin")
    fout = os.path.join(tmpdir, "out.bin")
    write_input(fin, L, kappa, 0.0, 1, 1, 1e-10, 1, q.gen_gauge(L, 0.3, 1), psi, phi)
    subprocess.run([HARNESS, "prop", fin, fout], check=True, stdout=subprocess.DEVNULL)
    recs = read_records(fout)
    _, sites = probes(L)
    data = {"L": np.array(L), "params": np.array([kappa, sp, sf]), "input_hash": np.array(input_hash(psi, phi)),
            "sites": sites, "prop_norm2": recs["prop_norm2"]}
    for name in ("prop_in", "prop_plus", "prop_minus"):
        M = recs[name].view(np.complex128).reshape(-1, 4, 4, 3, 3)
        data["site_" + name] = M[sites]
        data["norm2_" + name] = np.array(np.vdot(M.ravel(), M.ravel()).real)
    data["site_rescale"] = recs["rescale"].view(np.complex128).reshape(-1, 4, 3)[sites]
    np.savez_compressed(os.path.join(OUT, f"prop_{tag}.npz"), **data)
    print(f"prop_{tag}.npz")


def main():
    if not os.path.exists(HARNESS):
        sys.exit(f"{HARNESS} missing: make -C oracle/ref (needs the reference sources)")
    only = set(sys.argv[1:])   # sections: ops, cfg2, smear, prop (default: all)
    with tempfile.TemporaryDirectory() as d:
        if not only or "ops" in only:
            for c in CASES:
                run_ops_case(c, d)
            run_cg_case(d)
        if not only or "cfg2" in only:
            # the reference CG at 24^3x48 to tol 1e-14: several minutes on one core
            run_cg_case(d, CG_CASE2, nsites=64)
        if not only or "smear" in only:
            for c in SMEAR_CASES:
                run_smear_case(c, d)
            run_proj_case(d)
        if not only or "prop" in only:
            run_prop_case(d)

if __name__ == "__main__":
    main()
