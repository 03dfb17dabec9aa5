"""MinRes (ngs.solvers.MinRes, opencmp/models/base_model.py:929-932) on the GPU path: preconditioned
MINRES in FP64 for symmetric systems. Pinned against the oracle's direct solve of the same free-dof
system (fem.solve_free) on an SPD Poisson problem and on the symmetric indefinite Taylor-Hood Stokes
system, and against scipy's own MINRES on the Stokes system preconditioned the same way: in exact
arithmetic both minimise the same preconditioned residual over the same Krylov space, so their
iteration counts agree up to the different stopping tests (ours: the FP64 true residual).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from oracle import fem
from opencmp_amd import _lib as L
from opencmp_amd.mesh import unit_square

pytestmark = pytest.mark.gpu


def say(*a):
    print('[gpu-test]', *a, flush=True)


def test_minres_spd_matches_cg_and_oracle(gpu_ctx):
    from opencmp_amd import core as C
    say('MinRes on P2 Poisson 16x16 vs CG and the oracle')
    hm = unit_square(16)
    names = '|'.join(hm.bnd_names)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.H1(2, names)])
    a = C.BilinearForm(gs)
    a += C.Vol('gradgrad', 0, 0)
    a.Assemble()
    f = C.LinearForm(gs)
    f += C.Source(0, (L.FN_CONST, (1.0,)))
    f.Assemble()
    pre = C.Preconditioner(a, 'local')
    pre.Update()
    x1, x2 = C.GridFunction(gs), C.GridFunction(gs)
    r1 = C.solve(a, f.vec, x1.vec, pre, 'MinRes', tol=1e-11, maxit=10000)
    r2 = C.solve(a, f.vec, x2.vec, pre, 'CG', tol=1e-11, maxit=10000)
    say('  minres', r1, 'cg', r2)
    assert r1.converged and r1.rel_residual <= 1e-11
    K = a.to_scipy().tocsr()
    xo = fem.solve_free(K, f.vec.cpu().numpy(), np.zeros(gs.ndof), gs.FreeDofs())
    assert np.linalg.norm(x1.numpy() - xo) <= 1e-9 * np.linalg.norm(xo)
    assert np.linalg.norm(x1.numpy() - x2.numpy()) <= 1e-9 * np.linalg.norm(xo)
    # a few more iterations than CG at most (MINRES minimises the residual, CG the energy error)
    assert r1.iters <= 1.5 * r2.iters + 5


def test_minres_stokes_matches_oracle_and_scipy(gpu_ctx):
    from opencmp_amd import core as C
    from opencmp_amd.models import Stokes, const
    say('MinRes on Taylor-Hood Stokes 8x8 with the positive Schur-diagonal preconditioner')
    hm = unit_square(8)
    inflow = [(L.FN_POLY2, [0, 0, 4.0, 0, 0, -4.0]), const(0.0)]
    wall = [const(0.0), const(0.0)]
    m = Stokes(hm, kv=1.0, interp_ord=2, dirichlet={'u': {'left': inflow, 'top': wall, 'bottom': wall}},
               linear_solver='MinRes', preconditioner='schur', linear_tolerance=1e-10, ctx=gpu_ctx)
    a = C.BilinearForm(m.fes, m.construct_bilinear_time_coefficient(1.0) + m.construct_bilinear_time_ODE(1.0))
    a.Assemble()
    f = C.LinearForm(m.fes, m.construct_linear(1.0))
    f.Assemble()
    pre = m.construct_preconditioners(a)
    pre.Update()
    x = C.GridFunction(m.fes)
    m.apply_dirichlet_bcs_to(x)
    x0 = x.numpy().copy()
    r = C.solve(a, f.vec, x.vec, pre, 'MinRes', tol=1e-10, maxit=20000)
    say('  ', r)
    assert r.converged and r.rel_residual <= 1e-10
    K = a.to_scipy().tocsr()
    free = m.fes.FreeDofs()
    xo = fem.solve_free(K, f.vec.cpu().numpy(), x0, free)
    uu = slice(0, 2 * m.fes.blocks[0][1])
    err = np.abs(x.numpy()[uu] - xo[uu]).max() / np.abs(xo[uu]).max()
    say('  velocity rel error vs oracle', err)
    assert err <= 1e-7
    # scipy's MINRES on the same free-dof system with the same positive diagonal preconditioner
    F = np.nonzero(free)[0]
    Kf = K[F][:, F].tocsr()
    bf = (f.vec.cpu().numpy() - K @ x0)[F]
    d = np.abs(Kf.diagonal())
    nu = np.count_nonzero(F < uu.stop)
    B = Kf[nu:, :nu]
    sd = np.asarray(B.multiply(B) @ (1.0 / d[:nu])).ravel() + np.abs(Kf.diagonal()[nu:])
    dinv = np.concatenate([1.0 / d[:nu], np.where(sd > 1e-10 * sd.max(), 1.0 / sd, 0.0)])
    M = spla.LinearOperator(Kf.shape, matvec=lambda v: dinv * v)
    its = []
    sol, info = spla.minres(Kf, bf, M=M, rtol=1e-10, maxiter=20000, callback=lambda xk: its.append(1))
    say('  scipy minres iterations', len(its), 'info', info, 'ours', r.iters)
    assert info == 0 and r.iters <= 1.5 * len(its) + 20
