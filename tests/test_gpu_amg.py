"""Smoothed-aggregation AMG on the Schur matrix S = C + B D^-1 B^T (ocmp_precond_set_schur_amg) as the
pressure part of the FGMRES preconditioner for the Taylor-Hood saddle points: the solution equals the
oracle's direct solve and the inner-PCG variant's, in FP64 and mixed precision, on a hierarchy of
several levels (the INS channel) and on a single dense level (a small Stokes problem)."""
import numpy as np
import pytest
import torch

from oracle import fem
from opencmp_amd import _lib as L

pytestmark = pytest.mark.gpu


def say(*a):
    print('[gpu-test]', *a, flush=True)


def channel(ctx, n, precision, **kw):
    from opencmp_amd.mesh import channel_with_holes
    from opencmp_amd.models import INS, const
    hm = channel_with_holes(n)
    inflow = [(L.FN_POLY2, [0, 0, 6.0, 0, 0, -6.0]), const(0.0)]
    zero = [const(0.0), const(0.0)]
    return INS(hm, kv=1e-3, interp_ord=2, dirichlet={'u': {'inlet': inflow, 'wall': zero, 'cyl': zero}},
               stress={'u': {'outlet': zero}}, linear_solver='GMRes', preconditioner='schur',
               linear_tolerance=1e-10, restart=50, precision=precision, ctx=ctx, **kw), hm


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_schur_amg_fgmres_matches_inner_pcg(gpu_ctx, precision):
    from opencmp_amd import core as C
    say('Schur AMG vs inner PCG on the INS channel n=32,', precision)
    m, hm = channel(gpu_ctx, 32, precision)
    dt = 1e-3
    a = C.BilinearForm(m.fes, m.time_derivative_terms() + m.construct_bilinear_time_coefficient(dt) +
                       m.construct_bilinear_time_ODE(dt))
    a.Assemble()
    g = torch.Generator(device='cpu').manual_seed(3)
    b = torch.randn(m.fes.ndof, dtype=torch.float64, generator=g).to(gpu_ctx.torch_device)
    b *= torch.from_numpy(m.fes.FreeDofs().astype(np.float64)).to(b.device)
    out = {}
    for tag, kw in (('pcg', dict(schur_iters=200, schur_tol=1e-1)), ('amg1', dict(schur_amg=1)),
                    ('amg2', dict(schur_amg=2)), ('amg1tri', dict(schur_amg=1))):
        pre = C.Preconditioner(a, 'schur', precision=precision, schur_blocks=m.schur_blocks(), **kw)
        if tag.endswith('tri'):
            pre.set_schur_triangular(True)
        pre.Update()
        x = torch.zeros_like(b)
        r = C.solve(a, b, x, pre, 'FGMRes', tol=1e-10, restart=50, maxit=20000, precision=precision)
        say(f'  {tag}: {r}')
        assert r.converged and r.rel_residual <= 1e-10 and r.precond_iters > 0
        out[tag] = (x, r)
    for tag in ('amg1', 'amg2', 'amg1tri'):
        rel = float(torch.linalg.norm(out[tag][0] - out['pcg'][0]) / torch.linalg.norm(out['pcg'][0]))
        say(f'  {tag} vs pcg: {rel:.2e}')
        assert rel < 1e-6
    # one V-cycle is about as good a Schur solve as the inner PCG to 1e-1 (within 30 % more outer iterations)
    assert out['amg1'][1].iters <= 1.3 * out['pcg'][1].iters + 5
    assert out['amg2'][1].iters <= out['amg1'][1].iters + 5
    # the upper block-triangular velocity rows take fewer outer iterations than the block-diagonal step
    assert out['amg1tri'][1].iters < out['amg1'][1].iters


def test_schur_amg_single_dense_level_matches_oracle(gpu_ctx):
    """A Stokes problem small enough that S is inverted densely (one level): FGMRES with the exact
    Schur solve converges in few iterations to the oracle's direct solution."""
    from opencmp_amd import core as C
    from opencmp_amd.mesh import unit_square
    from opencmp_amd.models import Stokes, const
    say('Schur AMG, one dense level: Stokes 5x5')
    hm = unit_square(5)                       # 36 pressure dofs: at most 48 rows stay a single dense level
    inflow = [(L.FN_POLY2, [0, 0, 4.0, 0, 0, -4.0]), const(0.0)]
    wall = [const(0.0), const(0.0)]
    m = Stokes(hm, kv=1.0, interp_ord=2, dirichlet={'u': {'left': inflow, 'top': wall, 'bottom': wall}},
               linear_solver='GMRes', preconditioner='schur', linear_tolerance=1e-11, schur_amg=1, ctx=gpu_ctx)
    assert m.linear_solver == 'FGMRes'
    a = C.BilinearForm(m.fes, m.construct_bilinear_time_coefficient(1.0) + m.construct_bilinear_time_ODE(1.0))
    a.Assemble()
    f = C.LinearForm(m.fes, m.construct_linear(1.0))
    f.Assemble()
    pre = m.construct_preconditioners(a)
    pre.Update()
    x = C.GridFunction(m.fes)
    m.apply_dirichlet_bcs_to(x)
    x0 = x.numpy().copy()
    r = C.solve(a, f.vec, x.vec, pre, 'FGMRes', tol=1e-11, restart=100, maxit=5000)
    say('  ', r)
    assert r.converged
    K = a.to_scipy().tocsr()
    xo = fem.solve_free(K, f.vec.cpu().numpy(), x0, m.fes.FreeDofs())
    uu = slice(0, 2 * m.fes.blocks[0][1])
    assert np.abs(x.numpy()[uu] - xo[uu]).max() <= 1e-8 * np.abs(xo[uu]).max()


def test_schur_amg_matches_cpu_restatement(gpu_ctx):
    """The GPU's Schur AMG (host-built aggregates, product-list values, smoothing folded into the
    transfers) against the oracle's textbook restatement (oracle/cpu_solver.cpp: sparse products,
    standard V-cycle) on the same INS system: the same hierarchy level by level, FGMRES outer
    iteration counts within rounding of each other, and the same solution."""
    import ctypes as C
    import os
    from oracle import cbuild
    from opencmp_amd import core as Co
    say('Schur AMG: GPU vs the CPU restatement on the INS channel n=32')
    m, hm = channel(gpu_ctx, 32, 'fp64')
    dt = 1e-3
    a = Co.BilinearForm(m.fes, m.time_derivative_terms() + m.construct_bilinear_time_coefficient(dt) +
                        m.construct_bilinear_time_ODE(dt))
    a.Assemble()
    free = m.fes.FreeDofs()
    g = np.random.default_rng(11)
    bh = np.where(free, g.standard_normal(m.fes.ndof), 0.0)
    b = torch.from_numpy(bh).to(gpu_ctx.torch_device)
    pre = Co.Preconditioner(a, 'schur', schur_blocks=m.schur_blocks(), schur_amg=1)
    pre.Update()
    lv_gpu = [r for r, _ in pre.amg_levels()]
    x = torch.zeros_like(b)
    r = Co.solve(a, b, x, pre, 'FGMRes', tol=1e-10, restart=50, maxit=20000)
    say('  gpu', r, 'levels', lv_gpu)
    assert r.converged and len(lv_gpu) >= 3
    if not os.path.exists(cbuild.LIB):
        raise RuntimeError('oracle/libocmp_cpu.so is missing: run __graft_entry__.build()')
    lib = cbuild.load()
    rp, ci = a.pattern()
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    va = np.ascontiguousarray(a.values(), dtype=np.float64)
    off = [int(v) for v in m.fes.block_offsets()]
    pb = m.fes.comp_blocks[1][0]
    s0, s1 = off[pb], off[pb + 1]
    xc = np.zeros(m.fes.ndof)
    fr = np.ascontiguousarray(free.astype(np.uint8))
    lv = np.zeros(10, np.int64)
    rel, inner = C.c_double(), C.c_int64()
    it = lib.ocmp_cpu_fgmres_schur(m.fes.ndof, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, fr.ctypes.data,
                                   s0, s1, bh.ctypes.data, xc.ctypes.data, 50, 1e-10, 20000, 1e-1, 200,
                                   C.byref(rel), C.byref(inner), 1, lv.ctypes.data, 0)
    lv_cpu = [int(v) for v in lv if v]
    say(f'  cpu: {it} outer iterations, rel {rel.value:.2e}, levels {lv_cpu}')
    assert lv_cpu == lv_gpu
    assert rel.value <= 1e-10 and abs(it - r.iters) <= max(2, 0.05 * it)
    xg = x.cpu().numpy()
    uu = slice(0, s0)
    assert np.abs(xg[uu] - xc[uu]).max() <= 1e-7 * np.abs(xc[uu]).max()
    # the upper block-triangular velocity rows (ocmp_precond_set_schur_triangular): the same count again
    pre.set_schur_triangular(True)
    pre.Update()
    x.zero_()
    rt = Co.solve(a, b, x, pre, 'FGMRes', tol=1e-10, restart=50, maxit=20000)
    xc[:] = 0.0
    itt = lib.ocmp_cpu_fgmres_schur(m.fes.ndof, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, fr.ctypes.data,
                                    s0, s1, bh.ctypes.data, xc.ctypes.data, 50, 1e-10, 20000, 1e-1, 200,
                                    C.byref(rel), C.byref(inner), 1, None, 1)
    say(f'  triangular: gpu {rt.iters}, cpu {itt} outer iterations')
    assert rt.converged and rel.value <= 1e-10 and abs(itt - rt.iters) <= max(2, 0.05 * itt) and itt < it
    assert np.abs(x.cpu().numpy()[uu] - xc[uu]).max() <= 1e-7 * np.abs(xc[uu]).max()


def test_jacobi_only_update_keeps_a_lagged_schur_step(gpu_ctx):
    """ocmp_precond_update_jacobi: after the operator changes (a new time step size), refreshing only
    the Jacobi rows keeps the Schur AMG of the old operator; FGMRES still reaches the tolerance and the
    same solution as with a full update. Refused before any full update and for a plain Jacobi."""
    from opencmp_amd import core as C
    say('Jacobi-only update with a lagged Schur AMG step')
    m, hm = channel(gpu_ctx, 32, 'mixed')

    def form(dt):
        a = C.BilinearForm(m.fes, m.time_derivative_terms() + m.construct_bilinear_time_coefficient(dt) +
                           m.construct_bilinear_time_ODE(dt))
        return a

    a = form(1e-3)
    a.Assemble()
    pre = C.Preconditioner(a, 'schur', precision='mixed', schur_blocks=m.schur_blocks(), schur_amg=1)
    with pytest.raises(L.OcmpError, match='full preconditioner update'):
        pre.UpdateJacobi()
    pre.Update()
    g = torch.Generator(device='cpu').manual_seed(5)
    b = torch.randn(m.fes.ndof, dtype=torch.float64, generator=g).to(gpu_ctx.torch_device)
    b *= torch.from_numpy(m.fes.FreeDofs().astype(np.float64)).to(b.device)
    # the same matrix object re-assembled with dt 2e-3: lagged (Jacobi only) vs full update
    a.terms = form(2e-3).terms
    a.Assemble()
    pre.UpdateJacobi()
    xl = torch.zeros_like(b)
    rl = C.solve(a, b, xl, pre, 'FGMRes', tol=1e-10, restart=50, maxit=20000, precision='mixed')
    pre.Update()
    xf = torch.zeros_like(b)
    rf = C.solve(a, b, xf, pre, 'FGMRes', tol=1e-10, restart=50, maxit=20000, precision='mixed')
    say(f'  lagged {rl}\n  full   {rf}')
    assert rl.converged and rf.converged and rl.iters <= 1.5 * rf.iters + 10
    assert float(torch.linalg.norm(xl - xf) / torch.linalg.norm(xf)) < 1e-6
    plain = C.Preconditioner(a, 'local')
    plain.Update()
    with pytest.raises(L.OcmpError, match='Jacobi-only update'):
        plain.UpdateJacobi()
