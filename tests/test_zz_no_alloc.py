"""The allocation-free residual entry point against the allocating one, through the C-ABI.

ocmp_residual_norms_noalloc runs the same SpMV and reduction kernels as
ocmp_residual_norms (it only skips storing the residual vector), with its partial
sums in the library's static per-device scratch array, so both results must agree
bit for bit. Two calls in a row, the second on a smaller matrix, show that what
the first call left in the shared array does not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _system(ctx, n, seed):
    from opencmp_amd import core as CC
    from opencmp_amd.mesh import unit_square
    hm = unit_square(n)
    gm = CC.Mesh(hm, ctx)
    fes = CC.FESpace(gm, [CC.H1(2, '|'.join(hm.bnd_names))])
    a = CC.BilinearForm(fes)
    a += CC.Vol('gradgrad', 0, 0)
    a += CC.Vol('mass', 0, 0, coef=0.5)
    a.Assemble()
    g = torch.Generator().manual_seed(seed)
    x = torch.randn(fes.ndof, generator=g, dtype=torch.float64).to(ctx.torch_device)
    b = torch.randn(fes.ndof, generator=g, dtype=torch.float64).to(ctx.torch_device)
    return gm, fes, a, x, b


def _both(ctx, a, x, b):
    lib = ctx.lib
    host = (C.c_double * 2)()
    rc = lib.ocmp_residual_norms(a.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), host)
    assert rc == 0, lib.ocmp_last_error()
    out = torch.full((2,), float('nan'), dtype=torch.float64, device=ctx.torch_device)
    rc = lib.ocmp_residual_norms_noalloc(a.h, C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()),
                                         C.c_void_p(out.data_ptr()))
    assert rc == 0, lib.ocmp_last_error()
    torch.cuda.synchronize()
    return np.array([host[0], host[1]]), out.cpu().numpy()


def test_noalloc_residual_matches_allocating_entry_point(gpu_ctx):
    ctx = gpu_ctx
    sizes = [(96, 1), (24, 2)]  # the second call has the smaller input
    for n, seed in sizes:
        print(f'noalloc residual: unit square {n}x{n}, H1 P2, seed {seed}', flush=True)
        gm, fes, a, x, b = _system(ctx, n, seed)
        ref, got = _both(ctx, a, x, b)
        assert np.all(np.isfinite(got))
        assert got[0] > 0 and got[1] > 0
        np.testing.assert_array_equal(got, ref)
        # an independent FP64 check of the same quantities (scipy on the host)
        A = a.to_scipy()
        free = fes.FreeDofs().astype(bool)
        xn, bn = x.cpu().numpy(), b.cpu().numpy()
        r = (bn - A @ xn)[free]
        np.testing.assert_allclose(got, [np.linalg.norm(r), np.linalg.norm(bn[free])], rtol=1e-12)
