"""GPU parity of the U-net kernels against oracle/unet.py, flexible GMRES against
GMRES, and the learned correction inside FGMRES on the diffuse-interface Stokes
model (opencmp/models/stokes_dim.py:32-153) with its fallback."""
import numpy as np
import pytest
import torch

from opencmp_amd import _lib as L
from oracle import unet as ou

pytestmark = pytest.mark.gpu

Y_RTOL = 5e-3     # bf16 activations: fp32 (GPU) vs fp64 (oracle) sums round a few activations one ulp apart


def say(*a):
    print('[gpu-test]', *a, flush=True)


def rand_params(seed):
    rng = np.random.default_rng(seed)
    return (ou.init_params(seed) + 0.05 * rng.standard_normal(ou.NPARAM)).astype(np.float32)


@pytest.mark.parametrize('N', [4, 8])
def test_unet_forward_matches_oracle(gpu_ctx, N):
    from opencmp_amd.learned import UNetPrecond
    say('unet forward N', N)
    p = rand_params(N)
    U = UNetPrecond(N, gpu_ctx, params=p)
    n = N + 1
    x = np.random.default_rng(N).standard_normal((4, n, n, n)).astype(np.float32)
    y = U.forward(torch.tensor(x.reshape(4, -1), device='cuda')).cpu().numpy().reshape(4, n, n, n)
    yo = ou.forward(p, x, bf16=True)
    say('  rel error vs bf16 oracle', np.linalg.norm(y - yo) / np.linalg.norm(yo))
    assert np.linalg.norm(y - yo) <= Y_RTOL * np.linalg.norm(yo)
    ye = ou.forward(p, x)
    assert np.linalg.norm(y - ye) <= 5e-2 * np.linalg.norm(ye)


def test_unet_grad_matches_oracle(gpu_ctx):
    from opencmp_amd.learned import UNetPrecond
    N = 8
    n = N + 1
    p = rand_params(3)
    U = UNetPrecond(N, gpu_ctx, params=p)
    rng = np.random.default_rng(4)
    x = rng.standard_normal((4, n, n, n)).astype(np.float32)
    gy = rng.standard_normal((4, n, n, n)).astype(np.float32)
    g = U.grad(torch.tensor(x.reshape(4, -1), device='cuda'), torch.tensor(gy.reshape(4, -1), device='cuda'))
    g = g.cpu().numpy().astype(np.float64)
    # the backward kernels against the oracle's backward through the GPU's own stored
    # activations (tight), and against the oracle's full bf16 / exact passes (loose: a few
    # activations whose fp32 sums round to a neighbouring bf16 value move the gradients
    # by a few per cent on a grid this small)
    shapes = {-1: (4, n), 0: (8, n), 1: (16, 5), 2: (16, 5), 3: (32, 3), 4: (32, 3), 5: (16, 5), 6: (8, n)}
    acts = {('x' if l < 0 else f'a{l}'): U.activation(l).cpu().numpy().reshape(c, m, m, m)
            for l, (c, m) in shapes.items()}
    _, gt = ou.backward(p, x, gy, bf16=True, acts=acts)
    _, go = ou.backward(p, x, gy, bf16=True)
    _, ge = ou.backward(p, x, gy, bf16=False)
    o = 0
    errs = []
    for name, kind, ci, co, s in ou.LAYERS:
        for part, size in (('W', ci * co * 27), ('b', co)):
            a, t, b, e = g[o:o + size], gt[o:o + size], go[o:o + size], ge[o:o + size]
            rt = np.linalg.norm(a - t) / (np.linalg.norm(t) + 1e-30)
            rb = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
            re = np.linalg.norm(a - e) / (np.linalg.norm(e) + 1e-30)
            say(f'  {name}.{part}: rel vs oracle on GPU activations {rt:.2e}, vs bf16 oracle {rb:.2e}, '
                f'vs exact {re:.2e}')
            errs.append((name, part, rt, rb, re))
            o += size
    # one-hot dL/dy: the OUT.W gradient then lists a6 around the chosen voxel
    _, acts, _ = ou.forward(p, x, bf16=True, keep=True)
    oh = np.zeros((4, n, n, n), dtype=np.float32)
    oh[0, 4, 4, 4] = 1.0
    g1 = U.grad(torch.tensor(x.reshape(4, -1), device='cuda'), torch.tensor(oh.reshape(4, -1), device='cuda'))
    w_out = g1.cpu().numpy()[ou.NPARAM - 4 - 4 * 8 * 27: ou.NPARAM - 4].reshape(4, 8, 27)
    ref = np.array([[acts['a6'][i, 4 + k // 9 - 1, 4 + (k // 3) % 3 - 1, 4 + k % 3 - 1] for k in range(27)]
                    for i in range(8)])
    say('  one-hot OUT.W vs a6: max diff', np.abs(w_out[0] - ref).max(), 'rest', np.abs(w_out[1:]).max())
    say('  gpu', w_out[0, 0, :9], 'ref', ref[0, :9])
    for name, part, rt, rb, re in errs:
        assert rt <= 1e-4 and rb <= 0.15 and re <= 0.2, (name, part, rt, rb, re)
    g2 = U.grad(torch.tensor(x.reshape(4, -1), device='cuda'), torch.tensor(gy.reshape(4, -1), device='cuda'))
    assert np.array_equal(g2.cpu().numpy().astype(np.float64), g)    # bitwise repeatable


def dim_stokes(ctx, N, precision='fp64'):
    from opencmp_amd.mesh import unit_cube, sphere_obstacles, obstacle_phi
    from opencmp_amd.models import StokesDIM, const
    hm = unit_cube(N)
    phi = obstacle_phi(hm.coords, sphere_obstacles(), 1.0 / N)
    inflow = [const(1.0), const(0.0), const(0.0)]
    zero = [const(0.0)] * 3
    walls = {b: zero for b in ('front', 'back', 'bottom', 'top')}
    m = StokesDIM(hm, torch.tensor(phi), phi_order=1, penalty=1e4, kv=1.0, interp_ord=2,
                  dirichlet={'u': {'left': inflow, **walls}}, linear_solver='GMRes', preconditioner='schur',
                  linear_tolerance=1e-8, restart=50, precision=precision, ctx=ctx)
    return m


def assembled(m):
    from opencmp_amd import core as C
    a = C.BilinearForm(m.fes, m.construct_bilinear_time_coefficient(1.0) + m.construct_bilinear_time_ODE(1.0))
    f = C.LinearForm(m.fes, m.construct_linear(1.0))
    a.Assemble()
    f.Assemble()
    pre = m.construct_preconditioners(a)
    pre.Update()
    return a, f, pre


def test_fgmres_without_correction_matches_gmres(gpu_ctx):
    """With a fixed preconditioner flexible GMRES spans the same spaces as right-preconditioned GMRES."""
    from opencmp_amd import core as C
    m = dim_stokes(gpu_ctx, 4)
    a, f, pre = assembled(m)
    x1 = C.GridFunction(m.fes)
    m.apply_dirichlet_bcs_to(x1)
    x2 = C.GridFunction(m.fes, x1.vec.clone())
    r1 = C.solve(a, f.vec, x1.vec, pre, 'GMRes', tol=1e-9, restart=50, maxit=200000)
    r2 = C.solve(a, f.vec, x2.vec, pre, 'FGMRes', tol=1e-9, restart=50, maxit=200000)
    say('  gmres', r1, 'fgmres', r2)
    assert r1.converged and r2.converged and not r2.fallback
    assert abs(r1.iters - r2.iters) <= max(5, 0.05 * r1.iters)
    rel = float(torch.linalg.norm(x1.vec - x2.vec) / torch.linalg.norm(x1.vec))
    assert rel < 1e-5


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_learned_correction_keeps_fgmres_converging(gpu_ctx, precision):
    """Train the U-net on residuals of a plain solve, attach it: FGMRES converges to the same
    solution and is never much slower than the plain path (the fallback bounds the loss to about
    two restart cycles). No reduction of iterations is claimed: the correction trained this way
    does not yet gain over the plain preconditioner."""
    from opencmp_amd import core as C
    from opencmp_amd.learned import UNetPrecond, residual_samples
    N = 8
    m = dim_stokes(gpu_ctx, N, precision)
    a, f, pre = assembled(m)
    x0 = C.GridFunction(m.fes)
    m.apply_dirichlet_bcs_to(x0)
    # plain reference solve
    xp = C.GridFunction(m.fes, x0.vec.clone())
    rp = C.solve(a, f.vec, xp.vec, pre, 'FGMRes', tol=1e-8, restart=50, maxit=200000, precision=precision)
    # training data: normalised residuals along a plain solve
    b0 = f.vec.clone()
    S = residual_samples(a, pre, b0, 16, 20, x0=x0.vec)
    U = UNetPrecond(N, gpu_ctx, seed=0)
    blocks = m.fes.comp_blocks[0] + m.fes.comp_blocks[1]
    off = [m.fes.blocks[k][0] for k in blocks]
    losses = U.train(a, pre, S, off, steps=60, batch=4, lr=2e-3, seed=1)
    say(f'  [{precision}] loss {losses[0]:.4e} -> {np.mean(losses[-10:]):.4e}')
    assert np.isfinite(losses).all() and np.mean(losses[-10:]) < losses[0]
    U.attach(pre)
    xl = C.GridFunction(m.fes, x0.vec.clone())
    rl = C.solve(a, f.vec, xl.vec, pre, 'FGMRes', tol=1e-8, restart=50, maxit=200000, precision=precision)
    U.detach(pre)
    say(f'  [{precision}] plain {rp.iters} iters, learned {rl.iters} iters (fallback {rl.fallback})')
    assert rp.converged and rl.converged and rl.rel_residual <= 1e-8
    assert rl.iters <= rp.iters + 2 * 50 + 10
    rel = float(torch.linalg.norm(xl.vec - xp.vec) / torch.linalg.norm(xp.vec))
    assert rel < 1e-5
