"""GPU: the learned preconditioner.  The network's forward pass and weight
gradients against the CPU oracle (whose gradients are pinned by finite
differences in test_oracle_ptc.py), the Adam step against its closed form,
training on plain-CG pairs, and flexible GCR with the trained network against
GCR without it (iterations) and against the oracle's true residual."""
import numpy as np
import pytest

import oracle as orc
from refops import rel

pytestmark = pytest.mark.gpu

LS = (4, 4, 4, 8)


def bf16_round(w):
    """Round complex weights to BF16 (round to nearest even), as the device does."""
    f = np.ascontiguousarray(np.asarray(w, np.complex64)).view(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    return (u.astype(np.uint32) << 16).view(np.float32).view(np.complex64).astype(np.complex128).reshape(w.shape)


def fp32(w):
    return np.asarray(w, np.complex64).astype(np.complex128)


@pytest.fixture(scope="module")
def lat(q):
    U = q.gen_gauge(LS, 0.4, 31)
    u64 = q.Gauge(LS, boundaryT=-1, host=U)
    u32 = q.Gauge(LS, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(u64)
    return U, orc.apply_bc(U, LS, -1), u64, u32


def rand_weights(ch, seed, scale=0.3):
    r = np.random.default_rng(seed)
    W1 = scale * (r.standard_normal((9, ch, 4)) + 1j * r.standard_normal((9, ch, 4)))
    W2 = scale * (r.standard_normal((9, 4, ch)) + 1j * r.standard_normal((9, 4, ch)))
    return fp32(W1), fp32(W2)


@pytest.mark.parametrize("ch", [4, 8, 16])
@pytest.mark.parametrize("bf16", [False, True])
def test_net_apply_matches_oracle(q, lat, ch, bf16):
    U, Ub, u64, u32 = lat
    W1, W2 = rand_weights(ch, 7 + ch)
    net = q.Net(ch, W1, W2)
    net.set_bf16(bf16)
    psi = q.gen_gaussian(LS, 3)
    print(f"net apply ch={ch} bf16={bf16}", flush=True)
    out = q.Spinor(LS)
    net.apply(out, q.Spinor(LS, host=psi), u32)
    Wr = (bf16_round(W1), bf16_round(W2)) if bf16 else (W1, W2)
    ref = orc.net_apply(psi, Ub, LS, *Wr)
    assert rel(out.download(), ref) <= 1e-5
    got1, got2 = net.get_weights()
    assert np.array_equal(got1, W1) and np.array_equal(got2, W2)


@pytest.mark.parametrize("ch", [4, 8, 16])
def test_net_gradient_matches_oracle(q, lat, ch):
    U, Ub, u64, u32 = lat
    W1, W2 = rand_weights(ch, 11 + ch)
    net = q.Net(ch, W1, W2)
    net.set_bf16(False)
    r, e = q.gen_gaussian(LS, 5), q.gen_gaussian(LS, 6)
    print(f"net gradient ch={ch}", flush=True)
    net.zero_grad()
    net.accumulate_grad(q.Spinor(LS, host=r), q.Spinor(LS, host=e), u32)
    G1, G2, loss = net.read_grad()
    loss_ref, G1r, G2r = orc.net_loss_grad(r, e, Ub, LS, W1, W2)
    assert abs(loss - loss_ref) <= 1e-5 * loss_ref
    assert rel(G1, G1r) <= 1e-4 and rel(G2, G2r) <= 1e-4
    # accumulation: a second call doubles gradient and loss
    net.accumulate_grad(q.Spinor(LS, host=r), q.Spinor(LS, host=e), u32)
    H1, H2, loss2 = net.read_grad()
    assert rel(H1, 2 * G1) <= 1e-6 and abs(loss2 - 2 * loss) <= 1e-9 * loss


def test_adam_first_step_closed_form(q, lat):
    """First Adam step with bias correction: w -= lr g / (|g| + eps) per real
    component, g = accumulated gradient / count."""
    U, Ub, u64, u32 = lat
    W1, W2 = rand_weights(8, 21)
    net = q.Net(8, W1, W2)
    net.set_bf16(False)
    r, e = q.gen_gaussian(LS, 8), q.gen_gaussian(LS, 9)
    net.zero_grad()
    for _ in range(2):
        net.accumulate_grad(q.Spinor(LS, host=r), q.Spinor(LS, host=e), u32)
    G1, G2, _ = net.read_grad()
    lr = 1e-3
    net.adam(lr, 0.9, 0.999, 1e-8, count=2)
    N1, N2 = net.get_weights()

    def step(w, g):
        g = g / 2
        gr, gi = g.real.astype(np.float32), g.imag.astype(np.float32)
        return w - lr * (gr / (np.abs(gr) + 1e-8) + 1j * gi / (np.abs(gi) + 1e-8))

    assert np.max(np.abs(N1 - step(W1, G1))) <= 2e-6
    assert np.max(np.abs(N2 - step(W2, G2))) <= 2e-6
    Z1, Z2, _ = net.read_grad()
    assert not np.any(Z1) and not np.any(Z2)


def test_gcr_without_net_matches_oracle(q, lat):
    U, Ub, u64, u32 = lat
    kappa, mu = 0.15, 0.05
    src = q.gen_z2(LS, 4)
    p = q.params(LS, kappa, mu, 1, -1)
    x, b = q.Spinor(LS), q.Spinor(LS, host=src)
    st = q.solve_gcr(x, b, u64, p, restart=8, tol=1e-10)
    res = rel(orc.tm(x.download(), Ub, LS, kappa, mu, 1), src)
    print(f"gcr iterations {st.iterations} res {res:.2e}", flush=True)
    assert st.converged and res <= 1e-10 and abs(st.true_residual - res) <= 1e-3 * res + 1e-14


# kappa below the critical value of these eps = 0.3 fields, where restarted
# GCR on D converges without help, so the network's gain is measurable
L8 = (8, 8, 8, 8)
K8, M8, E8 = 0.13, 0.01, 0.3


def test_training_and_gain(q):
    """Train on plain-CG pairs from two gauge fields, then solve on an unseen
    gauge field: the trained network lowers the pair loss and cuts the GCR
    iteration count by at least 20%, and the solution meets 1e-10 against the
    oracle."""
    from qkxtm_mpi_amd import learned
    print("making pairs", flush=True)
    pairs = learned.make_pairs(L8, K8, M8, E8, gauge_seeds=[11, 12], sources_per_gauge=2, every=2, max_pairs=256)
    net = q.Net(16, seed=1)
    loss0 = learned.pair_loss(net, pairs, K8, M8)
    hist = learned.train(net, pairs, K8, M8, epochs=10, batch=8, lr=3e-3, seed=2, log=print)
    print(f"pairs {len(pairs)} loss {loss0:.4f} -> {hist[-1]:.4f}", flush=True)
    assert hist[-1] < 0.7 * loss0

    U = q.gen_gauge(L8, E8, 99)
    u64 = q.Gauge(L8, boundaryT=-1, host=U)
    u32 = q.Gauge(L8, q.SINGLE, q.RECON_18, -1)
    u32.convert_from(u64)
    src = q.gen_z2(L8, 5)
    p = q.params(L8, K8, M8, 1, -1)
    b = q.Spinor(L8, host=src)
    x0, x1 = q.Spinor(L8), q.Spinor(L8)
    st0 = q.solve_gcr(x0, b, u64, p, restart=8, tol=1e-10, maxIter=5000)
    st1 = q.solve_gcr(x1, b, u64, p, net=net, u_net=u32, restart=8, tol=1e-10, maxIter=5000)
    res1 = rel(orc.tm(x1.download(), orc.apply_bc(U, L8, -1), L8, K8, M8, 1), src)
    print(f"GCR iterations without net {st0.iterations}, with net {st1.iterations} (fallback "
          f"{st1.fallback_iterations}); residual {res1:.2e}", flush=True)
    assert st0.converged and st1.converged and res1 <= 1e-10
    assert st1.fallback_iterations == 0 and st1.iterations <= 0.8 * st0.iterations
