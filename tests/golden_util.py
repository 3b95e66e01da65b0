"""Helpers shared by tests/golden/make_golden.py and the tests that read the
fixtures (hash of the regenerated inputs, probe vectors, summaries)."""
import hashlib

import numpy as np

import qkxtm_mpi_amd as q


def input_hash(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def probes(L, nsites=8):
    V = int(np.prod(L))
    R = np.stack([q.gen_gaussian(L, 900 + k) for k in range(4)])
    rng = np.random.default_rng(7)
    sites = np.sort(rng.choice(V, size=min(nsites, V), replace=False))
    return R, sites


def summarize(y, R, sites):
    y = np.asarray(y).reshape(-1, 4, 3)
    proj = np.array([np.vdot(R[k].ravel(), y.ravel()) for k in range(R.shape[0])])
    return {"norm2": np.vdot(y.ravel(), y.ravel()).real, "proj": proj, "sites": y[sites].copy()}
