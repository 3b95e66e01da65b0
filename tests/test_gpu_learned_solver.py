"""GPU: the learned, mixed-precision, multi-RHS solver (qkx_solve_learned).

The network on even-parity fields (N_e = P_e M P_e, forward and adjoint,
batched, every storage precision) against the CPU oracle's layers
(oracle/qkx_oracle_ptc.c, pinned by finite differences and the adjoint
identity in test_oracle_ptc.py); the solver's solutions against the oracle's
FP64 true residual (the reference operator, lib/lattice_util.cpp:1925-1928)
and against the same batch solved without the network (qkx_solve_batch, the
lockstep form of lib/solvers.cpp:16-81); what the trained network gains in
iterations; the plain-CG fallback when the network does not help; and the
transfer of trained weights to a larger lattice with a fine-tune there."""
import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu

REV = [0, 2, 1, 4, 3, 6, 5, 8, 7]


def even_mask(L):
    x = np.indices(L[::-1]).reshape(4, -1)[::-1]   # lexicographic x fastest: coordinates (x, y, z, t)
    return ((x.sum(axis=0) & 1) == 0)


def fp32(w):
    return np.asarray(w, np.complex64).astype(np.complex128)


def rand_weights(ch, seed, scale=0.3):
    r = np.random.default_rng(seed)
    W1 = scale * (r.standard_normal((9, ch, 4)) + 1j * r.standard_normal((9, ch, 4)))
    W2 = scale * (r.standard_normal((9, 4, ch)) + 1j * r.standard_normal((9, 4, ch)))
    return fp32(W1), fp32(W2)


def oracle_eo(psi, Ub, L, W1, W2, adjoint):
    """P_e M P_e psi (or its adjoint) with the oracle's layers."""
    m = even_mask(L)[:, None, None]
    pe = psi * m
    if adjoint:
        h = orc.ptc_adj(pe, Ub, L, W2)        # L2^dag: 4 -> ch
        out = orc.ptc_adj(h, Ub, L, W1)       # L1^dag: ch -> 4
    else:
        out = orc.net_apply(pe, Ub, L, W1, W2)
    return out * m


@pytest.mark.parametrize("ch", [4, 8, 16])
@pytest.mark.parametrize("prec,tol", [("DOUBLE", 2e-6), ("SINGLE", 2e-6), ("HALF", 1e-2), ("BF16", 2e-2)])
@pytest.mark.parametrize("adjoint", [False, True])
def test_net_eo_matches_oracle(q, ch, prec, tol, adjoint):
    L = (4, 6, 4, 8)
    nrhs = 3
    U = q.gen_gauge(L, 0.4, 41)
    Ub = orc.apply_bc(U, L, -1)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(q.Gauge(L, boundaryT=-1, host=U))
    W1, W2 = rand_weights(ch, 50 + ch)
    net = q.Net(ch, W1, W2)
    net.set_bf16(False)
    P = getattr(q, prec)
    srcs = np.stack([q.gen_gaussian(L, 60 + r) for r in range(nrhs)])
    print(f"N_e ch={ch} {prec} adjoint={adjoint}", flush=True)
    x = q.Spinor(L, P, q.EVEN, host=srcs, nrhs=nrhs)
    y = q.Spinor(L, P, q.EVEN, nrhs=nrhs)
    net.apply_eo(y, x, u32, adjoint)
    got = y.download()
    for r in range(nrhs):
        ref = oracle_eo(srcs[r], Ub, L, W1, W2, adjoint)
        assert rel(got[r], ref) <= tol, r
    # in place, single field: the same result
    x1 = q.Spinor(L, P, q.EVEN, host=srcs[1])
    net.apply_eo(x1, x1, u32, adjoint)
    assert rel(x1.download(), got[1]) <= tol


def test_net_eo_rhs_chunks_identical(q):
    """A hidden-buffer bound that forces the batch through in chunks of one
    and of two right-hand sides gives the same bits as one launch pair."""
    L, nrhs = (4, 6, 4, 8), 5
    U = q.gen_gauge(L, 0.4, 45)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(q.Gauge(L, boundaryT=-1, host=U))
    W1, W2 = rand_weights(8, 88)
    net = q.Net(8, W1, W2)
    srcs = np.stack([q.gen_gaussian(L, 70 + r) for r in range(nrhs)])
    per_rhs = 8 * 3 * (np.prod(L) // 2) * 8       # bytes of one right-hand side's hidden channels
    outs = []
    for limit in (0, per_rhs, 2 * per_rhs):
        net.set_hidden_limit(limit)
        for adj in (False, True):
            x = q.Spinor(L, q.HALF, q.EVEN, host=srcs, nrhs=nrhs)
            y = q.Spinor(L, q.HALF, q.EVEN, nrhs=nrhs)
            net.apply_eo(y, x, u32, adj)
            net.apply_eo(x, x, u32, adj)          # in place, chunk by chunk
            outs.append((y.download(), x.download()))
    net.set_hidden_limit(0)
    for k in range(2, len(outs)):
        ref = outs[k % 2]
        assert np.array_equal(outs[k][0], ref[0]) and np.array_equal(outs[k][1], ref[0])


def test_net_eo_bf16_weights(q):
    """Inference with the BF16 copies of the weights: the oracle with the
    weights rounded to BF16."""
    from test_gpu_learned import bf16_round
    L = (4, 4, 4, 8)
    U = q.gen_gauge(L, 0.4, 43)
    Ub = orc.apply_bc(U, L, -1)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(q.Gauge(L, boundaryT=-1, host=U))
    W1, W2 = rand_weights(8, 77)
    net = q.Net(8, W1, W2)
    net.set_bf16(True)
    psi = q.gen_gaussian(L, 5)
    x, y = q.Spinor(L, q.SINGLE, q.EVEN, host=psi), q.Spinor(L, q.SINGLE, q.EVEN)
    for adj in (False, True):
        net.apply_eo(y, x, u32, adj)
        assert rel(y.download(), oracle_eo(psi, Ub, L, bf16_round(W1), bf16_round(W2), adj)) <= 2e-6


def test_gaussian_noise_statistics(q):
    """Device noise: deterministic per seed, different seeds differ, moments
    of a unit normal per real component."""
    L = (16, 16, 16, 16)
    a, b, c = q.Spinor(L), q.Spinor(L), q.Spinor(L)
    a.gaussian(11)
    b.gaussian(11)
    c.gaussian(12)
    A, B, C = a.download(), b.download(), c.download()
    assert np.array_equal(A, B) and not np.array_equal(A, C)
    v = np.concatenate([A.real.ravel(), A.imag.ravel()])
    assert abs(v.mean()) < 0.01 and abs(v.var() - 1.0) < 0.01
    assert abs(np.mean(v ** 4) - 3.0) < 0.05     # Gaussian kurtosis
    assert abs(np.corrcoef(A.real.ravel(), C.real.ravel())[0, 1]) < 0.01


def _system(q, L, seed, kappa, mu, nrhs, eps=0.3):
    U = q.gen_gauge(L, eps, seed)
    u = q.Gauge(L, boundaryT=-1, host=U)
    u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(u)
    srcs = np.stack([q.gen_z2(L, 1000 * seed + r) for r in range(nrhs)])
    return U, u, u32, srcs


def _check_solution(q, x, srcs, U, L, kappa, mu, st):
    Ub = orc.apply_bc(U, L, -1)
    xs = x.download()
    for r in range(len(srcs)):
        res = rel(orc.tm(xs[r], Ub, L, kappa, mu, 1), srcs[r])
        assert st[r].converged and res <= 1e-10, (r, res)
        assert abs(st[r].true_residual - res) <= 1e-3 * res + 1e-15
    return xs


@pytest.fixture(scope="module")
def trained(q):
    """A network of 4 hidden channels trained at 16^3x32, kappa = 0.1625, on
    1024 random pairs from 4 gauge fields (the benchmark's recipe)."""
    from qkxtm_mpi_amd import learned
    L, kappa, mu = (16, 16, 16, 32), 0.1625, 0.004
    gauges = []
    for gs in (101, 102, 103, 104):
        u = q.Gauge(L, boundaryT=-1, host=q.gen_gauge(L, 0.3, gs))
        u32 = q.Gauge(L, q.SINGLE, q.RECON_18, -1)
        u32.convert_from(u)
        gauges.append(u32)
    net = q.Net(4, seed=1)
    loss0 = learned.random_loss(net, gauges, kappa, mu, 64)
    hist = learned.train_random(net, gauges, kappa, mu, 1024, epochs=4, batch=16, lr=2e-3, seed=3)
    loss1 = learned.random_loss(net, gauges, kappa, mu, 64)
    print(f"random-pair loss {loss0:.4f} -> {loss1:.4f} (epochs {hist})", flush=True)
    assert loss1 < 0.5 * loss0
    return net


@pytest.mark.parametrize("prec,delta", [("SINGLE", 0.1), ("HALF", 0.1), ("BF16", 0.1)])
def test_learned_solver_gain_16x16x16x32(q, trained, prec, delta):
    """Four Z2 sources on an unseen gauge field at the learned config: the
    network cuts the batch's CG iterations by at least 20% against the same
    solver without it, and both meet the oracle's true residual."""
    L, kappa, mu = (16, 16, 16, 32), 0.1625, 0.004
    U, u, u32, srcs = _system(q, L, 201, kappa, mu, 4)
    p = q.params(L, kappa, mu, 1, -1)
    P = getattr(q, prec)
    b = q.Spinor(L, host=srcs, nrhs=4)
    x_net, x_plain = q.Spinor(L, nrhs=4), q.Spinor(L, nrhs=4)
    print(f"learned solver {prec} delta={delta}", flush=True)
    s_net = q.solve_learned(x_net, b, u, p, trained, u32, inner_prec=P, delta=delta, tol=1e-10, stall=0.5)
    s_plain = q.solve_learned(x_plain, b, u, p, None, None, inner_prec=P, delta=delta, tol=1e-10)
    print(f"  iterations with the network {s_net[0].iterations} (fallback {s_net[0].fallback_iterations}), "
          f"without {s_plain[0].iterations}", flush=True)
    a = _check_solution(q, x_net, srcs, U, L, kappa, mu, s_net)
    c = _check_solution(q, x_plain, srcs, U, L, kappa, mu, s_plain)
    for r in range(4):
        assert rel(a[r], c[r]) <= 1e-8
    assert s_net[0].fallback_iterations == 0
    assert s_net[0].iterations <= 0.8 * s_plain[0].iterations


def test_learned_solver_falls_back(q):
    """A network that maps every field to zero (all weights zero): the first
    preconditioned cycle cannot lower the residual, the plain CG takes over
    from there and the FP64 true residual is still met."""
    L, kappa, mu = (8, 8, 8, 16), 0.1625, 0.004
    U, u, u32, srcs = _system(q, L, 7, kappa, mu, 2)
    p = q.params(L, kappa, mu, 1, -1)
    net = q.Net(8, np.zeros((9, 8, 4), complex), np.zeros((9, 4, 8), complex))
    b = q.Spinor(L, host=srcs, nrhs=2)
    x = q.Spinor(L, nrhs=2)
    st = q.solve_learned(x, b, u, p, net, u32, inner_prec=q.SINGLE, delta=0.1, tol=1e-10, stall=0.5,
                         maxIter=20000)
    print(f"fallback after {st[0].iterations - st[0].fallback_iterations} preconditioned iterations, "
          f"{st[0].fallback_iterations} plain", flush=True)
    assert st[0].fallback_iterations > 0
    _check_solution(q, x, srcs, U, L, kappa, mu, st)


def test_learned_transfer_and_fine_tune_24x24x24x48(q, trained):
    """Weights trained at 16^3x32 act at 24^3x48, kappa = 0.1632 (the network
    is local, so they apply to any volume); 128 random pairs at the new
    volume lower the loss further, and the fine-tuned network still gains."""
    from qkxtm_mpi_amd import learned
    L, kappa, mu = (24, 24, 24, 48), 0.1632, 0.004
    U, u, u32, srcs = _system(q, L, 1, kappa, mu, 2)
    W1, W2 = trained.get_weights()
    net = q.Net(4, W1, W2)
    before = learned.random_loss(net, [u32], kappa, mu, 16)
    learned.train_random(net, [u32], kappa, mu, 128, epochs=1, batch=16, lr=5e-4, seed=9)
    after = learned.random_loss(net, [u32], kappa, mu, 16)
    print(f"fine-tune at 24^3x48: loss {before:.4f} -> {after:.4f}", flush=True)
    assert after <= before * 1.02
    p = q.params(L, kappa, mu, 1, -1)
    b = q.Spinor(L, host=srcs, nrhs=2)
    x, x0 = q.Spinor(L, nrhs=2), q.Spinor(L, nrhs=2)
    st = q.solve_learned(x, b, u, p, net, u32, inner_prec=q.HALF, delta=0.1, tol=1e-10, stall=0.5)
    s0 = q.solve_learned(x0, b, u, p, None, None, inner_prec=q.HALF, delta=0.1, tol=1e-10)
    print(f"  iterations {st[0].iterations} with the network, {s0[0].iterations} without", flush=True)
    assert all(s.converged and s.true_residual <= 1e-10 for s in st)
    assert st[0].fallback_iterations == 0 and st[0].iterations <= 0.8 * s0[0].iterations
    assert rel(x.download_rhs(0), x0.download_rhs(0)) <= 1e-8
