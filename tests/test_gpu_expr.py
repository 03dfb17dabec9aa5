"""Config expressions (OCMP_FN_EXPR), whole-domain Set, the P3 face lift and vector norms on the GPU.

Pins: closed forms the expression must equal (a POLY2 function of the same
polynomial), exact reproduction of polynomials in the space (P3 Set of a cubic,
DG L2 projection of a polynomial of the block's order), the oracle's lift and
mass-matrix L2 projection, and L1/L2 norms with closed-form values.
"""
import numpy as np
import pytest
import torch

from oracle import fem

pytestmark = pytest.mark.gpu


def say(*a):
    print('[gpu-test]', *a, flush=True)


def test_expression_equals_closed_form(gpu_ctx):
    from opencmp_amd import core as C, _lib as L, expr as E
    from opencmp_amd.mesh import unit_square
    say('expression vs closed form: boundary lift and source, H1 P3 on 6x6')
    hm = unit_square(6)
    names = '|'.join(hm.bnd_names)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.H1(3, names)])
    cc = [1.0, 1.0, -2.0, 1.0, 0.5, -1.0]
    text = '1 + x - 2*y + x^2 + 0.5*x*y - y^2'
    g1, g2 = C.GridFunction(gs), C.GridFunction(gs)
    g1.Set((L.FN_POLY2, cc), names)
    g2.Set(E.Expr(text).bind(gpu_ctx), names)
    assert float((g1.vec - g2.vec).abs().max()) <= 1e-14
    f1, f2 = C.LinearForm(gs), C.LinearForm(gs)
    f1 += C.Source(0, (L.FN_POLY2, cc))
    f2 += C.Source(0, E.Expr(text).bind(gpu_ctx))
    f1.Assemble()
    f2.Assemble()
    assert float((f1.vec - f2.vec).abs().max()) <= 1e-14 * float(f1.vec.abs().max())
    # scale and shift, and a time parameter read at assembly time
    t = E.Parameter(0.0)
    b = E.Expr('x*exp(-2*t)').bind(gpu_ctx, t, scale=3.0, shift=0.5)
    g3 = C.GridFunction(gs)
    t.Set(0.5)
    g3.Set(b, names)
    g4 = C.GridFunction(gs)
    g4.Set((L.FN_POLY2, [0.5, 3.0 * np.exp(-1.0), 0, 0, 0, 0]), names)
    assert float((g3.vec - g4.vec).abs().max()) <= 1e-14


@pytest.mark.parametrize('dim', [2, 3])
def test_whole_domain_set_reproduces_cubics(gpu_ctx, dim):
    from opencmp_amd import core as C, _lib as L, expr as E
    from opencmp_amd.mesh import unit_square, unit_cube
    say('whole-domain Set of a cubic, H1 P3, dim', dim)
    hm = unit_square(5) if dim == 2 else unit_cube(3)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.H1(3)])
    text = 'x^3 - 2*x*y^2 + y + 1' if dim == 2 else 'x*y*z + x^3 - 2*y^2*z + 1'
    e = E.Expr(text).bind(gpu_ctx)
    g = C.GridFunction(gs)
    g.Set(e)
    l2, l1, h1 = vec_norms(gpu_ctx, gs, [0], g.vec, [e])
    say('  L2', l2, 'L1', l1)
    assert l2 < 1e-13 and l1 < 1e-13 and np.isnan(h1)
    # a non-polynomial is interpolated with the O(h^4) error of P3
    g.Set(E.Expr('sin(3*x)*cos(2*y)').bind(gpu_ctx))
    l2s, _, _ = vec_norms(gpu_ctx, gs, [0], g.vec, [E.Expr('sin(3*x)*cos(2*y)').bind(gpu_ctx)])
    assert 1e-7 < l2s < 1e-2


def test_p3_lift_projects_faces_like_the_oracle(gpu_ctx):
    """3D P3 lift of a cubic whose face traces have face-bubble components: equal to the oracle's
    lift, and lift + Galerkin reproduces the cubic."""
    from opencmp_amd import core as C, _lib as L, expr as E
    from opencmp_amd.mesh import unit_cube
    say('3D P3 Dirichlet lift with face projection')
    hm = unit_cube(2)
    names = '|'.join(hm.bnd_names)
    mask = hm.marker_mask(names)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.H1(3, names)])
    u = lambda x, y, z: x * y * z + x ** 3 - 2 * y * y * z + 1
    g = C.GridFunction(gs)
    g.Set(E.Expr('x*y*z + x^3 - 2*y^2*z + 1').bind(gpu_ctx), names)
    t = fem.Topology(hm)
    s = fem.Space(t, 'H1', 3, dirichlet_mask=mask)
    ref = fem.h1_lift(s, u, mask)
    face0 = t.nv + 2 * t.nedge
    assert np.abs(ref[face0:]).max() > 1e-3        # the face dofs matter for this cubic
    assert np.abs(g.numpy() - ref).max() <= 1e-12
    a = C.BilinearForm(gs)
    a += C.Vol('gradgrad', 0, 0)
    a.Assemble()
    f = C.LinearForm(gs)
    f += C.Source(0, E.Expr('-(6*x - 4*z)').bind(gpu_ctx))
    f.Assemble()
    pre = C.Preconditioner(a, 'local')
    pre.Update()
    r = C.solve(a, f.vec, g.vec, pre, 'CG', tol=1e-13)
    assert r.converged
    assert fem.l2_error(s, g.numpy(), u) < 1e-10


def test_dg_set_is_the_l2_projection(gpu_ctx):
    from opencmp_amd import core as C, _lib as L, expr as E
    from opencmp_amd.mesh import unit_square
    say('DG P2 Set = element L2 projection')
    hm = unit_square(8)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.L2(2)], dgjumps=True)
    g = C.GridFunction(gs)
    g.Set(E.Expr('sin(3*x)*cos(2*y) + x').bind(gpu_ctx))
    t = fem.Topology(hm)
    s = fem.Space(t, 'L2', 2)
    cs = fem.Compound([s])
    A = fem.Assembler(cs)
    fem.assemble_volume(A, [dict(kind='mass', trial=0, test=0, coef=1.0)])
    M = A.matrix()
    b = fem.assemble_source(cs, 0, lambda x, y: np.sin(3 * x) * np.cos(2 * y) + x, qdeg=14)
    import scipy.sparse.linalg as spla
    ref = spla.spsolve(M.tocsc(), b)
    assert np.abs(g.numpy() - ref).max() <= 1e-11 * np.abs(ref).max()
    # a quadratic is reproduced exactly
    g.Set(E.Expr('x^2 - x*y + 3').bind(gpu_ctx))
    l2, _, _ = vec_norms(gpu_ctx, gs, [0], g.vec, [E.Expr('x^2 - x*y + 3').bind(gpu_ctx)])
    assert l2 < 1e-13


def vec_norms(ctx, gs, blocks, vec, fns, ref=None):
    import ctypes
    from opencmp_amd import _lib as L
    from opencmp_amd.expr import resolve_fn
    nb = len(blocks)
    bl = (ctypes.c_int32 * nb)(*blocks)
    fid = (ctypes.c_int32 * nb)()
    c = (ctypes.c_double * (10 * nb))()
    for k, f in enumerate(fns or []):
        i, prm = resolve_fn(f)
        fid[k] = i
        for j, v in enumerate(prm):
            c[10 * k + j] = v
    out = (ctypes.c_double * 3)()
    L.check(ctx.lib.ocmp_error_norms_vec(gs.h, nb, bl, ctypes.c_void_p(vec.data_ptr()),
                                         None if ref is None else ctypes.c_void_p(ref.data_ptr()), fid, c, out),
            'error_norms_vec')
    return out[0], out[1], out[2]


def test_vector_l1_l2_norms_closed_forms(gpu_ctx):
    from opencmp_amd import core as C, _lib as L, expr as E
    from opencmp_amd.mesh import unit_square, unit_cube
    say('vector L1/L2 norms against closed forms')
    hm = unit_square(4)
    gm = C.Mesh(hm, gpu_ctx)
    gs = C.FESpace(gm, [C.H1(2, ncomp=2), C.H1(1)])
    ub = gs.comp_blocks[0]
    zero = C.GridFunction(gs)
    # |u_h - g| with u_h = 0 and g = (3, 4): L1 = 5 |Omega|, L2 = 5 sqrt|Omega|
    l2, l1, h1 = vec_norms(gpu_ctx, gs, ub, zero.vec, [(L.FN_CONST, (3.0,)), (L.FN_CONST, (4.0,))])
    assert l1 == pytest.approx(5.0, rel=1e-13) and l2 == pytest.approx(5.0, rel=1e-13) and abs(h1) < 1e-13
    # u_h = (x, 0): int |x| = 1/2, sqrt(int x^2) = 1/sqrt(3); against a reference vector on the space
    g = C.GridFunction(gs)
    g.Set([E.Expr('x').bind(gpu_ctx), E.Expr('0').bind(gpu_ctx)], component=0)
    l2, l1, h1 = vec_norms(gpu_ctx, gs, ub, g.vec, None, ref=zero.vec)
    assert l1 == pytest.approx(0.5, rel=1e-12) and l2 == pytest.approx(3 ** -0.5, rel=1e-12)
    assert h1 == pytest.approx(1.0, rel=1e-12)
    # the pointwise Euclidean norm: u_h = (x, y) has int sqrt(x^2 + y^2) over the unit square
    g.Set([E.Expr('x').bind(gpu_ctx), E.Expr('y').bind(gpu_ctx)], component=0)
    l2, l1, _ = vec_norms(gpu_ctx, gs, ub, g.vec, None, ref=zero.vec)
    exact_l1 = (np.sqrt(2) + np.log(1 + np.sqrt(2))) / 3
    assert l1 == pytest.approx(exact_l1, rel=1e-5)   # the integrand has a kink at the origin
    assert l2 == pytest.approx(np.sqrt(2 / 3), rel=1e-12)
