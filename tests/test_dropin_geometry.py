"""The drop-in LatticeGeometry (no GPU): its neighbour tables
(lib/lattice_util.cpp:86-99) checked site by site by the test_geometry driver
-- every pointPlus / pointMinus entry one step away with periodic wrap, the
two tables inverse to each other, one shared set per lattice size -- and its
gammaM against the UKQCD matrices the operator tests use (which the
Clifford-algebra test pins)."""
import os
import re
import subprocess

import numpy as np
import pytest

from test_oracle_math import G5_UKQCD, G_UKQCD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "qkxtm_mpi_amd", "host", "test_geometry")


@pytest.mark.parametrize("L", [(4, 6, 8, 10), (2, 2, 2, 2), (8, 4, 6, 16)])
def test_neighbour_tables(L):
    r = subprocess.run([DRIVER] + [str(v) for v in L], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"geometry ok V={int(np.prod(L))}" in r.stdout


def test_gamma_matrices():
    r = subprocess.run([DRIVER, "4", "4", "4", "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    g = np.zeros((5, 4, 4), complex)
    for m in re.finditer(r"gamma (\d) (\d) (\d) (\S+) (\S+)", r.stdout):
        k, a, b = int(m.group(1)), int(m.group(2)), int(m.group(3))
        g[k, a, b] = float(m.group(4)) + 1j * float(m.group(5))
    for mu in range(4):
        assert np.array_equal(g[mu], G_UKQCD[mu]), mu
    assert np.array_equal(g[4], G5_UKQCD)
