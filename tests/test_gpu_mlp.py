"""GPU parity of the per-DOF MLP kernels (forward, gradient, Adam, sampling,
scales, guess) against the CPU oracle oracle/mlp.py, and the fallback rule."""
import numpy as np
import pytest
import torch

from oracle import mlp as om

pytestmark = pytest.mark.gpu

# bf16 activations: the GPU accumulates in fp32, the oracle in fp64, so an
# activation whose fp32 sum sits next to a bf16 rounding boundary may round one
# bf16 ulp (2^-8 relative) apart. These tolerances bound that effect.
Y_RTOL = 2e-3      # ||y_gpu - y_oracle|| / ||y_oracle||
G_RTOL = 2e-2      # per tensor, ||g_gpu - g_oracle|| / ||g_oracle||


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = om.init_params(seed)
    return (p + 0.05 * rng.standard_normal(p.size)).astype(np.float32)


def dev(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device='cuda')


@pytest.mark.parametrize('n', [1, 37, 128, 1000, 4133])
def test_forward_matches_oracle(gpu_ctx, n):
    from opencmp_amd.learned import MLPGuess
    p = rand_params(n)
    net = MLPGuess(gpu_ctx, params=p, max_batch=8192)
    X = np.random.default_rng(n + 1).standard_normal((n, 4)).astype(np.float32)
    y = net.forward(dev(X, torch.float32)).cpu().numpy().astype(np.float64)
    yo = om.forward(p, X, bf16=True)
    assert np.linalg.norm(y - yo) <= Y_RTOL * np.linalg.norm(yo) + 1e-6
    # the exact fp32-weight network is within bf16 storage error
    ye = om.forward(p, X)
    assert np.linalg.norm(y - ye) <= 5e-2 * np.linalg.norm(ye)


def test_forward_is_deterministic(gpu_ctx):
    from opencmp_amd.learned import MLPGuess
    net = MLPGuess(gpu_ctx, params=rand_params(3), max_batch=1024)
    X = dev(np.random.default_rng(0).standard_normal((70000, 4)), torch.float32)
    a = net.forward(X)
    b = net.forward(X)
    assert torch.equal(a, b)


@pytest.mark.parametrize('B', [64, 1000, 20000])
def test_gradient_matches_oracle(gpu_ctx, B):
    from opencmp_amd.learned import MLPGuess
    p = rand_params(B)
    net = MLPGuess(gpu_ctx, params=p, max_batch=B)
    rng = np.random.default_rng(B)
    X = rng.standard_normal((B, 4)).astype(np.float32)
    T = rng.standard_normal(B).astype(np.float32)
    g, loss = net.grad(dev(X, torch.float32), dev(T, torch.float32))
    g = g.cpu().numpy().astype(np.float64)
    lo, go = om.loss_and_grad(p, X, T.astype(np.float64), bf16=True)
    le, ge = om.loss_and_grad(p, X, T.astype(np.float64), bf16=False)
    assert abs(loss - lo) <= 1e-3 * lo
    gd, god, ged = om.unpack(g), om.unpack(go), om.unpack(ge)
    for name, _ in om.SHAPES:
        ref = np.linalg.norm(god[name])
        assert np.linalg.norm(gd[name] - god[name]) <= G_RTOL * ref + 1e-7, name
        # against the exact (fp32-weight, fp64-sum) gradient: bf16 storage error only
        assert np.linalg.norm(gd[name] - ged[name]) <= 1e-1 * np.linalg.norm(ged[name]) + 1e-7, name
    # repeatable bit for bit
    g2, loss2 = net.grad(dev(X, torch.float32), dev(T, torch.float32))
    assert loss2 == loss and np.array_equal(g2.cpu().numpy().astype(np.float64), g)


def test_gradient_rejects_oversized_batch(gpu_ctx):
    from opencmp_amd import _lib
    from opencmp_amd.learned import MLPGuess
    net = MLPGuess(gpu_ctx, max_batch=100)
    X = torch.zeros((101, 4), dtype=torch.float32, device='cuda')
    T = torch.zeros(101, dtype=torch.float32, device='cuda')
    with pytest.raises(_lib.OcmpError, match='max_batch'):
        net.grad(X, T)


def test_adam_step_matches_formula(gpu_ctx):
    from opencmp_amd.learned import MLPGuess
    p = rand_params(11)
    net = MLPGuess(gpu_ctx, params=p, max_batch=4096)
    rng = np.random.default_rng(11)
    X = dev(rng.standard_normal((3000, 4)), torch.float32)
    T = dev(rng.standard_normal(3000), torch.float32)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    q = p.astype(np.float64)
    for k in range(1, 4):
        g, _ = net.grad(X, T)       # gradient at the current parameters
        net.train_step(X, T, 1e-3)
        q, m, v = om.adam(q, g.cpu().numpy().astype(np.float64), m, v, k, lr=1e-3)
        got = net.params.astype(np.float64)
        assert np.max(np.abs(got - q)) < 1e-6, k
        q = got


def test_training_reduces_loss(gpu_ctx):
    """Targets from a smooth function of the features: 300 Adam steps cut the loss 10x."""
    from opencmp_amd.learned import MLPGuess
    net = MLPGuess(gpu_ctx, seed=0, max_batch=16384)
    rng = np.random.default_rng(0)
    Xh = rng.standard_normal((16384, 4)).astype(np.float32)
    Th = (0.8 * Xh[:, 0] - 0.3 * Xh[:, 1] + 0.2 * np.tanh(Xh[:, 3])).astype(np.float32)
    X, T = dev(Xh, torch.float32), dev(Th, torch.float32)
    l0 = net.train_step(X, T, 3e-3, want_loss=True)
    for _ in range(300):
        net.train_step(X, T, 3e-3)
    l1 = net.grad(X, T)[1]
    assert l1 < 0.1 * l0, (l0, l1)


def snapshots(n=300, K=12, seed=0):
    rng = np.random.default_rng(seed)
    om_ = rng.uniform(0.5, 2.0, n)
    ph = rng.uniform(0, 6.28, n)
    amp = rng.uniform(0.5, 2.0, n)
    t = 0.05 * np.arange(K)
    return amp[None, :] * np.sin(om_[None, :] * t[:, None] + ph[None, :]) + 0.1 * np.arange(n)[None, :] / n


def test_scales_match_oracle(gpu_ctx):
    from opencmp_amd.learned import MLPGuess
    snap = snapshots()
    off = [0, 100, 250, 300]
    net = MLPGuess(gpu_ctx, max_batch=128)
    net.fit_scales(dev(snap, torch.float64), off)
    o, s, U = net.scales()
    so, Uo = om.scales(snap, off)
    assert o == off
    np.testing.assert_allclose(s, so, rtol=1e-7)
    np.testing.assert_allclose(U, Uo, rtol=1e-7)


@pytest.mark.parametrize('base', [0, 1])
def test_samples_come_from_the_snapshots(gpu_ctx, base):
    from opencmp_amd.learned import MLPGuess
    snap = snapshots(n=60, K=9)
    off = [0, 60]
    net = MLPGuess(gpu_ctx, max_batch=128, base=base)
    assert net.base == base
    net.fit_scales(dev(snap, torch.float64), off)
    _, s, U = net.scales()
    X, T = net.sample(dev(snap, torch.float64), 60, seed=5, batch=500)
    X, T = X.cpu().numpy(), T.cpu().numpy()
    s32, U32 = np.float32(s[0]), np.float32(U[0])
    cand = {}
    for j in range(3, 8):
        f = om.features(snap[j], snap[j - 1], snap[j - 2], snap[j - 3], float(s32), float(U32))
        t = om.target(snap[j + 1], snap[j], snap[j - 1], float(s32), base)
        for i in range(60):
            cand[(j, i)] = (f[i], t[i])
    keys = list(cand)
    F = np.array([cand[k][0] for k in keys])
    Tt = np.array([cand[k][1] for k in keys])
    seen = set()
    for b in range(500):
        d = np.abs(F - X[b]).max(1) + np.abs(Tt - T[b])
        k = int(np.argmin(d))
        assert d[k] < 1e-5 * (1 + np.abs(F[k]).max()), b
        seen.add(keys[k][0])
    assert seen == {3, 4, 5, 6, 7}     # every admissible snapshot index is drawn


@pytest.mark.parametrize('base', [0, 1])
def test_guess_matches_oracle(gpu_ctx, base):
    from opencmp_amd.learned import MLPGuess
    p = rand_params(21)
    net = MLPGuess(gpu_ctx, params=p, max_batch=128, base=base)
    snap = snapshots(n=1000, K=6, seed=2)
    off = [0, 400, 1000]
    s, U = [0.02, 0.05], [1.3, 0.7]
    net.set_scales(off, s, U)
    hist = [dev(snap[5 - k], torch.float64) for k in range(4)]
    g = net.guess(hist).cpu().numpy()
    go = om.guess(p, [snap[5 - k] for k in range(4)], off, s, U, bf16=True, base=base)
    # the increment s*y carries the network's error; u0 + base (u0 - u1) is exact
    b0 = snap[5] + base * (snap[5] - snap[4])
    assert np.linalg.norm((g - b0) - (go - b0)) <= Y_RTOL * np.linalg.norm(go - b0) + 1e-9
    # a network whose output layer is zero returns the base exactly
    net.params = np.where(np.arange(len(p)) >= len(p) - 129, 0.0, p).astype(np.float32)   # W4, b4 = 0
    g0 = net.guess(hist).cpu().numpy()
    assert np.abs(g0 - b0).max() <= 1e-12 * np.abs(b0).max()


def test_fallback_keeps_the_better_guess(gpu_ctx):
    """choose_guess keeps the plain guess when the learned one has a larger FP64
    residual, and takes the learned one when it is better."""
    from opencmp_amd.learned import choose_guess, residual_norms
    from opencmp_amd.mesh import unit_square
    from opencmp_amd.models import Poisson, const
    from opencmp_amd.solvers import StationarySolver
    hm = unit_square(8)
    m = Poisson(hm, interp_ord=2, f=const(1.0), dirichlet={'u': {b: const(0.0) for b in hm.bnd_names}},
                linear_solver='CG', linear_tolerance=1e-12, ctx=gpu_ctx)
    s = StationarySolver(m)
    exact = s.solve().vec.clone()
    a, b = s.a, s.L.vec
    free = torch.tensor(m.fes.FreeDofs().astype(np.float64), device=exact.device)
    plain = torch.zeros_like(exact)
    bad = exact + 10.0 * free
    choice, rp, rl = choose_guess(a, b, plain, bad)
    assert choice == 'plain' and rl > rp and torch.count_nonzero(plain) == 0
    choice, rp, rl = choose_guess(a, b, plain, exact)
    assert choice == 'learned' and rl < 1e-10 * rp and torch.equal(plain, exact)
    r, bn = residual_norms(a, b, exact)
    assert r < 1e-10 * bn
