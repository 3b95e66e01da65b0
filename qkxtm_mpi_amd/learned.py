"""Training data and training loops of the learned preconditioner.

Two kinds of pairs.  Plain-solver pairs: for each gauge field and source b, the
reference CG on D^dag D x = D^dag b is run from x = 0 and its iterate x_k is
kept every few iterations (qkx_cg_record); the error e_k = x* - x_k (x* the
FP64 solution) and the residual r_k = b - D x_k = D e_k form one pair.  Only
e_k is stored (FP32, normalised to |e_k| = 1); r_k = D e_k is recomputed on
the GPU when the pair is used.  The loss of a pair is |M r_k - e_k|^2, so M
learns D^-1 on the error directions plain CG leaves behind.  Every step runs
on the GPU through the library's own kernels (forward, adjoint, weight
gradient, Adam).

Random pairs (train_random): e is device-generated Gaussian noise and r = D e,
the cost |M D e - e|^2 of Lehner and Wettig (Phys. Rev. D 108, 034503 (2023)).
It asks M to invert D on every direction, not only on the slow ones that the
plain-solver pairs sample, which is what a right preconditioner of the normal
equations needs: trained on plain-solver pairs, the network raised the
even-odd CG iterations at 16^3x32, kappa = 0.1625, from 57 to 179; trained on
random pairs, it cut them to 39.  No pair is stored: e is regenerated from its
seed, so fine-tuning at 48^3x96 needs no field per pair.
"""
import time

import numpy as np

from . import (DOUBLE, SINGLE, RECON_18, Gauge, Net, Spinor, X_eq_aX_p_Y, cg_record, gen_gauge, gen_z2, params,
               solve, sync)


class PairSet:
    """Normalised error fields e_k (FP32) per gauge field."""

    def __init__(self):
        self.gauges = []   # (u64, u32, [e fields])
        self.iterations = []

    def __len__(self):
        return sum(len(es) for _, _, es in self.gauges)


def make_pairs(L, kappa, mu, eps, gauge_seeds, sources_per_gauge, every, max_pairs=None, source_seed=1000):
    """Pairs from plain-CG iterates on the given seeded gauge fields."""
    ps = PairSet()
    for gs in gauge_seeds:
        U = gen_gauge(L, eps, gs)
        u64 = Gauge(L, DOUBLE, RECON_18, -1, host=U)
        u32 = Gauge(L, SINGLE, RECON_18, -1)
        u32.convert_from(u64)
        p = params(L, kappa, mu, 1, -1, tol=1e-10, maxIter=20000)
        es = []
        zero32 = Spinor(L, SINGLE)
        for j in range(sources_per_gauge):
            b = Spinor(L, host=gen_z2(L, source_seed * gs + j))
            xs = Spinor(L)
            solve(xs, b, u64, p, eo=True, inner_prec=DOUBLE, tol=1e-12, maxIter=20000)
            bn = Spinor(L)
            bn.tmDag(b, u64, p)
            x = Spinor(L)
            nmax = 20000 // every
            store = []
            st, n = None, 0
            # grow the store until the run fits (CG length is not known in advance)
            k = 64
            while True:
                store = [Spinor(L, SINGLE) for _ in range(min(k, nmax))]
                x.zero()
                st, n = cg_record(x, bn, u64, p, every, store)
                if n < len(store) or len(store) == nmax:
                    break
                k *= 2
            ps.iterations.append(st.iterations)
            xs32 = Spinor(L, SINGLE)
            xs32.copy(xs)
            for f in store[:n]:
                X_eq_aX_p_Y(f, -1.0, xs32)                   # e = x* - x_k
                X_eq_aX_p_Y(f, 1.0 / np.sqrt(f.norm2()), zero32)
                es.append(f)
                if max_pairs is not None and len(ps) + len(es) >= max_pairs:
                    break
            if max_pairs is not None and len(ps) + len(es) >= max_pairs:
                break
        ps.gauges.append((u64, u32, es))
        if max_pairs is not None and len(ps) >= max_pairs:
            break
    return ps


def train(net: Net, pairs: PairSet, kappa, mu, epochs=10, batch=16, lr=2e-3, seed=0, log=None):
    """Adam on the pair loss; returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    index = [(g, k) for g, (_, _, es) in enumerate(pairs.gauges) for k in range(len(es))]
    L = pairs.gauges[0][0].L
    p = params(L, kappa, mu, 1, -1)
    r = Spinor(L, SINGLE)
    history = []
    for ep in range(epochs):
        t0 = time.perf_counter()
        net.zero_grad()
        order = rng.permutation(len(index))
        pending = 0
        for i in order:
            g, k = index[i]
            _, u32, es = pairs.gauges[g]
            r.tm(es[k], u32, p)                 # r = D e
            net.accumulate_grad(r, es[k], u32)
            pending += 1
            if pending == batch:
                net.adam(lr, count=batch)
                pending = 0
        if pending:
            net.adam(lr, count=pending)
        loss = net.read_loss() / len(index)
        history.append(loss)
        if log:
            sync()
            log(f"epoch {ep}: loss {loss:.4f} ({time.perf_counter() - t0:.2f} s)")
    return history


def pair_loss(net: Net, pairs: PairSet, kappa, mu):
    """Mean pair loss of the current weights (no update)."""
    L = pairs.gauges[0][0].L
    p = params(L, kappa, mu, 1, -1)
    r = Spinor(L, SINGLE)
    net.zero_grad()
    n = 0
    for _, u32, es in pairs.gauges:
        for e in es:
            r.tm(e, u32, p)
            net.accumulate_grad(r, e, u32)
            n += 1
    loss = net.read_loss() / n
    net.zero_grad()
    return loss


def train_random(net: Net, gauges, kappa, mu, n_pairs, epochs=4, batch=16, lr=2e-3, seed=0, log=None):
    """Adam on |M D e - e|^2 over n_pairs random unit vectors e per epoch,
    spread over the FP32 gauge fields `gauges` (18-real, the network's
    links); e_k is regenerated on the device from seed + k.  Returns the mean
    loss per epoch."""
    L = gauges[0].L
    p = params(L, kappa, mu, 1, -1)
    e = Spinor(L, SINGLE)
    r = Spinor(L, SINGLE)
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        t0 = time.perf_counter()
        net.zero_grad()
        pending = 0
        for k in rng.permutation(n_pairs):
            u32 = gauges[int(k) % len(gauges)]
            e.gaussian(1_000_003 * (seed + 1) + int(k))
            e.rescale(1.0 / np.sqrt(e.norm2()))
            r.tm(e, u32, p)                      # r = D e
            net.accumulate_grad(r, e, u32)
            pending += 1
            if pending == batch:
                net.adam(lr, count=batch)
                pending = 0
        if pending:
            net.adam(lr, count=pending)
        loss = net.read_loss() / n_pairs
        history.append(loss)
        if log:
            sync()
            log(f"epoch {ep}: loss {loss:.4f} ({time.perf_counter() - t0:.2f} s)")
    return history


def random_loss(net: Net, gauges, kappa, mu, n_pairs, seed=12345):
    """Mean |M D e - e|^2 over n_pairs fresh random unit vectors (no update)."""
    L = gauges[0].L
    p = params(L, kappa, mu, 1, -1)
    e = Spinor(L, SINGLE)
    r = Spinor(L, SINGLE)
    net.zero_grad()
    for k in range(n_pairs):
        u32 = gauges[k % len(gauges)]
        e.gaussian(seed * 7919 + k)
        e.rescale(1.0 / np.sqrt(e.norm2()))
        r.tm(e, u32, p)
        net.accumulate_grad(r, e, u32)
    loss = net.read_loss() / n_pairs
    net.zero_grad()
    return loss
