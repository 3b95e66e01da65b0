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


def gauge_probes(L, nsites=8):
    """Four seeded gauge-shaped Gaussian probes [V, 4, 3, 3] and 8 seeded sites."""
    R = np.stack([np.stack([q.gen_gaussian(L, 950 + 3 * k + c) for c in range(3)], axis=-1) for k in range(4)])
    rng = np.random.default_rng(11)
    V = int(np.prod(L))
    sites = np.sort(rng.choice(V, size=min(nsites, V), replace=False))
    return R, sites


def summarize_gauge(U, R, sites):
    U = np.asarray(U).reshape(-1, 4, 3, 3)
    proj = np.array([np.vdot(R[k].ravel(), U.ravel()) for k in range(R.shape[0])])
    return {"norm2": np.vdot(U.ravel(), U.ravel()).real, "proj": proj, "sites": U[sites].copy()}


def generic_matrices(L, seed):
    """Gauge-shaped field of generic complex matrices (projection input): three
    seeded Gaussian fields as the columns, plus at site 0 diag(2, 3i, 1) and at
    site 1 2.5 times the identity in every direction (the diagonal shortcut of
    the projection)."""
    M = np.stack([q.gen_gaussian(L, seed + c) for c in range(3)], axis=-1)
    M[0, :] = np.diag([2.0, 3.0j, 1.0])
    M[1, :] = 2.5 * np.eye(3)
    return np.ascontiguousarray(M)
