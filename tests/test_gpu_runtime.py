"""Runtime behaviour of the library's handles on the GPU: resource lifetime."""
import gc

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dg_solve_cycle(ctx, hm):
    """Create a DG-P2 matrix, its block-Jacobi preconditioner and a GMRES(50) workspace, solve
    a few cycles, and drop every handle."""
    from opencmp_amd import core as C, _lib as L
    gm = C.Mesh(hm, ctx)
    gs = C.FESpace(gm, [C.L2(2)], dgjumps=True)
    w = (np.cos(0.3), np.sin(0.3))
    a = C.BilinearForm(gs)
    a += C.Vol('gradgrad', 0, 0, coef=1e-3)
    a += C.Vol('conv', 0, 0, coef=-1.0, wind=(w[0], w[1], 0.0))
    a += C.DGFacet(0, eps=1e-3, alpha=40.0, wind=(w[0], w[1]), dirichlet=hm.marker_mask(hm.bnd_names))
    a.Assemble()
    f = C.LinearForm(gs)
    f += C.Source(0, (L.FN_CONST, (1.0,)))
    f.Assemble()
    pre = C.Preconditioner(a, 'local')
    pre.Update()
    x = C.GridFunction(gs)
    r = C.solve(a, f.vec, x.vec, pre, 'GMRes', tol=1e-30, restart=50, maxit=150)
    assert r.iters > 50   # at least one full cycle: the whole Krylov basis was allocated
    n = gs.ndof
    del a, f, pre, x, gs, gm
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return n


def test_matrix_destroy_frees_solver_workspace(gpu_ctx):
    """The Krylov workspace belongs to its matrix: destroying the matrix frees it, so building
    and solving many systems one after another leaves device memory flat (each cycle's basis
    is 51 vectors of the 98,304-dof system, ~40 MB)."""
    from opencmp_amd.mesh import unit_square
    print('[gpu-test] workspace lifetime: DG-P2 128x128, 30 create/solve/destroy cycles', flush=True)
    hm = unit_square(128)
    n = _dg_solve_cycle(gpu_ctx, hm)
    _dg_solve_cycle(gpu_ctx, hm)
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(30):
        _dg_solve_cycle(gpu_ctx, hm)
    free1, _ = torch.cuda.mem_get_info()
    basis = 51 * n * 8
    drift = free0 - free1
    print(f'[gpu-test]   basis {basis / 2**20:.1f} MiB per matrix, free-memory drift {drift / 2**20:.1f} MiB',
          flush=True)
    assert drift < 2 * basis, f'device memory fell by {drift / 2**20:.1f} MiB over 30 cycles'
