"""Model/solver drivers of the GPU path against the oracle and the reference's known answers."""
import os
import sys

import numpy as np
import pytest
import torch

from oracle import fem
from opencmp_amd import _lib as L
from opencmp_amd.mesh import unit_square, unit_cube
from opencmp_amd.io import load_mesh

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), 'golden', 'pytests', 'mesh_files')
PI = np.pi


def say(*a):
    print('[gpu-test]', *a, flush=True)


def test_poisson_cg_stationary_config1(gpu_ctx):
    """Config 1 through StationarySolver: errors equal the oracle's for the same discretisation."""
    from opencmp_amd.models import Poisson, const
    from opencmp_amd.solvers import StationarySolver
    say('poisson cg config1 via StationarySolver')
    hm = unit_square(32)
    allb = {b: const(0.0) for b in hm.bnd_names}
    m = Poisson(hm, dc=1.0, f=(L.FN_SINPROD, (2 * PI * PI,)), interp_ord=2, dirichlet={'u': allb},
                linear_solver='CG', linear_tolerance=1e-10, ctx=gpu_ctx)
    s = StationarySolver(m)
    gfu = s.solve()
    assert s.result.converged
    e_l2, e_h1 = m.l2_error(gfu, 'u', (L.FN_SINPROD, (1.0,)))
    t = fem.Topology(hm)
    sp = fem.Space(t, 'H1', 2, dirichlet_mask=hm.marker_mask(hm.bnd_names))
    x = gfu.numpy()
    o_l2, o_h1 = fem.l2_error(sp, x, lambda x, y: np.sin(PI * x) * np.sin(PI * y),
                              grad_exact=[lambda x, y: PI * np.cos(PI * x) * np.sin(PI * y),
                                          lambda x, y: PI * np.sin(PI * x) * np.cos(PI * y)])
    say('  L2', e_l2, o_l2, 'H1', e_h1, o_h1)
    assert abs(e_l2 - o_l2) <= 1e-10 * o_l2 and abs(e_h1 - o_h1) <= 1e-10 * o_h1
    assert 8.0e-6 < e_l2 < 9.2e-6      # oracle value at N=32, P2: 8.6e-6


def test_poisson_dg_reproduces_quadratic(gpu_ctx):
    """DG SIPG Poisson (opencmp/models/poisson.py:81-158) is consistent: exact for a global quadratic."""
    from opencmp_amd.models import Poisson
    from opencmp_amd.solvers import StationarySolver
    say('poisson dg quadratic')
    hm = unit_square(6)
    c = [1.0, 0.5, -1.0, 1.0, 0.5, 2.0]        # u = 1 + x/2 - y + x^2 + xy/2 + 2y^2, -lap u = -6
    g = (L.FN_POLY2, c)
    m = Poisson(hm, dc=1.0, f=(L.FN_CONST, (-6.0,)), interp_ord=2, DG=True,
                dirichlet={'u': {b: g for b in hm.bnd_names}}, linear_solver='GMRes', linear_tolerance=1e-12,
                ctx=gpu_ctx)
    s = StationarySolver(m)
    gfu = s.solve()
    assert s.result.converged
    e_l2, _ = m.l2_error(gfu, 'u', g)
    assert e_l2 < 1e-9, e_l2


def test_reference_transient_poisson_known_answer(gpu_ctx):
    """pytests/full_system/poisson/test_poisson.py:60-78 on the GPU: unit_square_coarse.vol, H1 P3,
    implicit Euler dt=1e-3 to t=1, u0 = x(1-x); printed L2 error 9e-6 at rtol 3. Also equal to the oracle."""
    from opencmp_amd.models import Poisson, const
    from opencmp_amd.solvers import ImplicitEulerSolver
    from opencmp_amd import core as C
    say('transient poisson known answer')
    hm = load_mesh(os.path.join(GOLD, 'unit_square_coarse.vol'))
    m = Poisson(hm, dc=1.0, f=const(0.0), interp_ord=3, dirichlet={'u': {'left': const(0.0), 'right': const(0.0)}},
                linear_solver='CG', linear_tolerance=1e-13, ctx=gpu_ctx)
    # initial condition: L2 projection of x(1-x) (in P3, so exact)
    Mf = C.BilinearForm(m.fes, [C.Vol('mass', 0, 0)])
    Mf.Assemble()
    b = C.LinearForm(m.fes, [C.Source(0, (L.FN_POLY2, [0, 1.0, 0, -1.0, 0, 0]))])
    b.Assemble()
    u0 = C.GridFunction(m.fes)
    pre = C.Preconditioner(Mf, 'jacobi')
    pre.Update()
    r = C.solve(Mf, b.vec, u0.vec, pre, 'CG', tol=1e-14)
    assert r.converged
    s = ImplicitEulerSolver(m, dt=1e-3, t_range=(0.0, 1.0), gfu0=u0)
    gfu = s.solve()
    err, _ = m.l2_error(gfu, 'u', const(0.0))
    say('  L2 norm at t=1:', err)
    assert np.isclose(err, 9e-6, rtol=3, atol=1e-12)
    # the oracle's value for the same discretisation (tests/test_oracle.py restates the loop)
    t = fem.Topology(hm)
    sp = fem.Space(t, 'H1', 3, dirichlet_mask=hm.marker_mask('left|right'))
    assert fem.l2_error(sp, gfu.numpy(), 0.0) == pytest.approx(err, rel=1e-12)


def test_stokes_taylor_hood_matches_oracle(gpu_ctx):
    """Stationary Stokes, Taylor-Hood P2/P1, parabolic inflow (opencmp/models/stokes.py:43-129)."""
    from opencmp_amd.models import Stokes, const
    from opencmp_amd.solvers import StationarySolver
    say('stokes taylor-hood')
    hm = unit_square(6)
    inflow = [(L.FN_POLY2, [0, 0, 4.0, 0, 0, -4.0]), const(0.0)]    # u = (4y(1-y), 0) on left
    wall = [const(0.0), const(0.0)]
    m = Stokes(hm, kv=1.0, interp_ord=2,
               dirichlet={'u': {'left': inflow, 'top': wall, 'bottom': wall}},
               linear_solver='GMRes', preconditioner='None', linear_tolerance=1e-11, restart=200, ctx=gpu_ctx)
    s = StationarySolver(m)
    gfu = s.solve()
    say('  ', s.result)
    assert s.result.converged
    # oracle: same forms, direct solve with the same lift
    t = fem.Topology(hm)
    dm = hm.marker_mask('left|top|bottom')
    su = fem.Space(t, 'H1', 2, ncomp=2, dirichlet_mask=dm)
    sp = fem.Space(t, 'H1', 1)
    cs = fem.Compound([su, sp])
    A = fem.Assembler(cs)
    fem.assemble_volume(A, [dict(kind='grad_grad', trial=0, test=0, coef=1.0),
                            dict(kind='dtrial', trial=0, test=1, coef=-1.0, dir=-1),
                            dict(kind='dtest', trial=1, test=0, coef=-1.0, dir=-1),
                            dict(kind='mass', trial=1, test=1, coef=-1e-10)])
    K = A.matrix()
    assert np.linalg.norm(s.a.values() - K.data) <= 1e-13 * np.linalg.norm(K.data)
    x0 = gfu.numpy().copy()
    # the lift values the GPU set must be the oracle's lift of the same data
    lift = np.zeros(cs.ndof)
    lift[:su.n1] = fem.h1_lift(fem.Space(t, 'H1', 2, dirichlet_mask=hm.marker_mask('left')),
                               lambda x, y: 4 * y * (1 - y), hm.marker_mask('left'))
    bd = ~cs.freedofs
    assert np.abs(x0[bd] - lift[bd]).max() < 1e-13
    xd = fem.solve_free(K, np.zeros(cs.ndof), lift, cs.freedofs)
    uu = slice(0, 2 * su.n1)
    assert np.abs(x0[uu] - xd[uu]).max() < 1e-7 * np.abs(xd[uu]).max()


def test_oseen_picard_converges(gpu_ctx):
    """Stationary INS by Oseen/Picard (opencmp/models/ins.py:323-379) at low Reynolds number:
    the Picard loop converges and the answer is close to the Stokes solution."""
    from opencmp_amd.models import INS, const
    from opencmp_amd.solvers import StationarySolver
    say('oseen picard')
    hm = unit_square(6)
    inflow = [(L.FN_POLY2, [0, 0, 4.0, 0, 0, -4.0]), const(0.0)]
    wall = [const(0.0), const(0.0)]
    m = INS(hm, kv=1.0, interp_ord=2, dirichlet={'u': {'left': inflow, 'top': wall, 'bottom': wall}},
            linear_solver='GMRes', preconditioner='None', linear_tolerance=1e-11, restart=200,
            nonlinear_tolerance=(1e-10, 1e-10), nonlinear_max_iters=20, ctx=gpu_ctx)
    s = StationarySolver(m)
    s.solve()
    say('  picard iterations', m.picard_iterations, 'error', m.picard_error, s.result)
    # inner solves are measured against their own initial residual (no body force), so the
    # criterion that matters is the Picard one: ||W - u||_L2 < abs + rel ||u||_L2
    assert m.picard_converged and 2 <= m.picard_iterations < 20


def test_picard_norms_match_the_separate_norms(gpu_ctx):
    """ocmp_l2_pair_norms (the Picard test's ||W - u|| and ||u|| in one pass, rule of degree 2p)
    equals the two separate L2 norms (rule of degree 2p + 6) on random P2 and P3 vector fields,
    and both equal the closed form for a linear field: ||(x, y)|| on the unit square is sqrt(2/3)."""
    from opencmp_amd import core as C
    from opencmp_amd.models import INS
    say('picard pair norms')
    hm = unit_square(6)
    g = np.random.default_rng(4)
    for order in (2, 3):
        m = INS(hm, kv=1.0, interp_ord=order, ctx=gpu_ctx)
        dev = gpu_ctx.torch_device
        W = C.GridFunction(m.fes, torch.tensor(g.standard_normal(m.fes.ndof), device=dev))
        u = C.GridFunction(m.fes, torch.tensor(g.standard_normal(m.fes.ndof), device=dev))
        err, un = m.picard_norms(W, u)
        zero = C.GridFunction(m.fes, torch.zeros(m.fes.ndof, dtype=torch.float64, device=dev))
        e2, u2 = m.l2_diff(W, u, 'u'), m.l2_diff(u, zero, 'u')
        say(f'  P{order}: pair {err:.15g} {un:.15g}, separate {e2:.15g} {u2:.15g}')
        assert abs(err - e2) <= 1e-13 * e2 and abs(un - u2) <= 1e-13 * u2
    off = m.fes.block_offsets()
    bx, by = m.fes.comp_blocks[0]
    v = np.zeros(m.fes.ndof)
    nv = hm.coords.shape[0]
    v[off[bx]:off[bx] + nv] = hm.coords[:, 0]
    v[off[by]:off[by] + nv] = hm.coords[:, 1]
    lin = C.GridFunction(m.fes, torch.tensor(v, device=dev))
    err, un = m.picard_norms(zero, lin)
    assert abs(err - np.sqrt(2 / 3)) <= 1e-14 and abs(un - np.sqrt(2 / 3)) <= 1e-14


def test_boundary_upwind_term_closed_forms(gpu_ctx):
    """The stress-BC no-backflow term coef * max(w.n, 0) u v (opencmp/models/ins.py:210-216) on the
    unit square, H1 P2: it couples boundary dofs only, and P2 traces interpolate x and y exactly, so
    u^T A u = sum over edges of int max(w.n, 0) u^2 ds in closed form. u = x separates the right edge
    (x = 1) from the left one (x = 0), so a flipped normal fails; a wind field checks the field path."""
    from opencmp_amd import core as C
    from opencmp_amd.models import const
    say('boundary upwind term')
    hm = unit_square(8)
    gm = C.Mesh(hm, gpu_ctx)
    names = '|'.join(hm.bnd_names)
    allm = hm.marker_mask(hm.bnd_names)
    sp = C.FESpace(gm, [C.H1(2, '')])

    def trace(c):
        g = C.GridFunction(sp)
        g.Set((L.FN_POLY2, c), names)
        return g.numpy()

    one, x, y = trace([1, 0, 0, 0, 0, 0]), trace([0, 1, 0, 0, 0, 0]), trace([0, 0, 1, 0, 0, 0])

    def form(**kw):
        a = C.BilinearForm(sp, [C.BndUpwind(0, kw.pop('markers', allm), **kw)])
        a.Assemble()
        return a.to_scipy()

    A = form(coef=1.0, wind=(1.0, 0.0, 0.0))
    vals = [float(v @ (A @ v)) for v in (one, x, y)]
    say('  wind (1,0): u=1, x, y ->', vals)
    assert np.allclose(vals, [1.0, 1.0, 1.0 / 3.0], rtol=0, atol=1e-12)
    A = form(coef=2.0, wind=(0.6, 0.8, 0.0))
    vals = [float(v @ (A @ v)) for v in (one, x)]
    say('  wind (0.6,0.8), coef 2: u=1, x ->', vals)
    assert np.allclose(vals, [2 * 1.4, 2 * (0.6 + 0.8 / 3.0)], rtol=0, atol=1e-12)
    # only the marked boundary: the top edge alone
    A = form(coef=1.0, wind=(0.6, 0.8, 0.0), markers=hm.marker_mask(['top']))
    assert abs(float(one @ (A @ one)) - 0.8) <= 1e-12 and abs(float(x @ (A @ x)) - 0.8 / 3.0) <= 1e-12
    # wind field w = (x, 0) set on the boundary (the trace on an edge depends on its own dofs only)
    ws = C.FESpace(gm, [C.H1(2, '', ncomp=2)])
    W = C.GridFunction(ws)
    W.Set([(L.FN_POLY2, [0, 1, 0, 0, 0, 0]), const(0.0)], names, component=0)
    A = form(coef=1.0, wind_gf=(W, [0, 1]))
    vals = [float(v @ (A @ v)) for v in (one, y)]
    say('  wind field (x,0): u=1, y ->', vals)
    assert np.allclose(vals, [1.0, 1.0 / 3.0], rtol=0, atol=1e-12)
    # symmetric, and zero off the boundary dofs (those a space with Dirichlet data everywhere fixes)
    assert abs(A - A.T).max() <= 1e-14
    bd = ~C.FESpace(gm, [C.H1(2, names)]).FreeDofs()
    assert abs(A[~bd]).max() == 0.0 and abs(A[:, ~bd]).max() == 0.0


def test_divergence_and_linf_norms_closed_forms(gpu_ctx):
    """_divergence and _l_inf of opencmp/helpers/error.py:80-88, :151-167 on linear P2 vector fields
    on the unit square (vertex coefficients = the field, edge coefficients 0 reproduce it exactly):
    sqrt(int (div u)^2) = |div u| for the unit area, and the largest coefficient difference is the
    largest vertex difference."""
    from opencmp_amd import core as C
    from opencmp_amd.models import Stokes
    say('divergence and L-inf norms')
    hm = unit_square(8)
    m = Stokes(hm, kv=1.0, interp_ord=2, ctx=gpu_ctx)
    off = m.fes.block_offsets()
    bx, by = m.fes.comp_blocks[0]
    X, Y = hm.coords[:, 0], hm.coords[:, 1]
    nv = len(X)

    def field(ux, uy):
        v = np.zeros(m.fes.ndof)
        v[off[bx]:off[bx] + nv] = ux
        v[off[by]:off[by] + nv] = uy
        return C.GridFunction(m.fes, torch.tensor(v, device=gpu_ctx.torch_device))

    a = field(X, Y)                       # div 2
    b = field(X + 2 * Y, 3 * X - Y)       # div 0
    c = field(3 * X, -Y)                  # div 2
    d = [m.divergence(g, 'u') for g in (a, b, c)]
    say('  div norms', d)
    assert np.allclose(d, [2.0, 0.0, 2.0], rtol=0, atol=1e-12)
    # max |a - b| over the vertex coefficients: |-2y| <= 2, |y - 3x| <= 3 (at (1, 0))
    li = m.linf_diff(a, b, 'u')
    l0 = m.linf_diff(a, None, 'u')
    say('  linf', li, l0)
    assert abs(li - 3.0) <= 1e-14 and abs(l0 - 1.0) <= 1e-14


def test_neumann_robin_terms_and_exact_solutions(gpu_ctx):
    """Neumann and Robin data of the Poisson model (opencmp/models/poisson.py:94-96, :151-158): the
    boundary source and boundary mass against closed forms on the unit square (b . interpolant of p
    = int g p ds for p = 1, y), and -lap u = 0 with u = x + 1 reproduced to rounding by P2 with
    Dirichlet data on the left and, on the right, Neumann data h = du/dn = 1 or Robin data
    du/dn = r (u - q) with r = -1, q = 3 (top and bottom: the natural condition du/dn = 0)."""
    from opencmp_amd import core as C
    from opencmp_amd.models import Poisson, const
    from opencmp_amd.solvers import StationarySolver
    say('neumann / robin')
    hm = unit_square(8)
    gm = C.Mesh(hm, gpu_ctx)
    sp = C.FESpace(gm, [C.H1(2, '')])
    right = hm.marker_mask(['right'])
    f = C.LinearForm(sp, [C.BndSource(0, right, fn=(L.FN_POLY2, [0, 0, 1, 0, 0, 0]), coef=2.0)])   # 2 y
    f.Assemble()
    b = f.vec.cpu().numpy()
    def trace(c):
        g = C.GridFunction(sp)
        g.Set((L.FN_POLY2, c), '|'.join(hm.bnd_names))
        return g.numpy()

    one, y = trace([1, 0, 0, 0, 0, 0]), trace([0, 0, 1, 0, 0, 0])
    say('  b.1', b @ one, 'b.y', b @ y)
    assert abs(b @ one - 1.0) <= 1e-13 and abs(b @ y - 2.0 / 3.0) <= 1e-13
    a = C.BilinearForm(sp, [C.BndMass(0, right, coef=3.0)])
    a.Assemble()
    A = a.to_scipy()
    assert abs(float(y @ (A @ y)) - 1.0) <= 1e-13   # 3 int_0^1 y^2 dy
    exact = (L.FN_POLY2, [1, 1, 0, 0, 0, 0])
    for kw in ({'neumann': {'u': {'right': const(1.0)}}}, {'robin': {'u': {'right': (-1.0, const(3.0))}}}):
        m = Poisson(hm, dc=1.0, f=const(0.0), interp_ord=2, dirichlet={'u': {'left': const(1.0)}},
                    linear_solver='CG', linear_tolerance=1e-13, ctx=gpu_ctx, **kw)
        s = StationarySolver(m)
        u = s.solve()
        e_l2, e_h1 = m.l2_error(u, 'u', exact)
        say('  ', list(kw)[0], s.result, 'L2', e_l2, 'H1', e_h1)
        assert s.result.converged and e_l2 <= 1e-10 and e_h1 <= 1e-9


def test_boundary_terms_on_tetrahedra(gpu_ctx):
    """The boundary-facet terms on the unit cube (tetrahedral faces): u^T A u = int over the marked
    faces in closed form for the Robin mass and the no-backflow term (u = x picks the x = 1 face
    only, so a flipped normal fails), b . interpolant of p = int g p dA for the boundary source, and
    -lap u = 0 with u = x + 1 reproduced by P2 with Neumann data du/dn = 1 on x = 1."""
    from opencmp_amd import core as C
    from opencmp_amd.models import Poisson, const
    from opencmp_amd.solvers import StationarySolver
    say('boundary terms on tetrahedra')
    hm = unit_cube(3)
    gm = C.Mesh(hm, gpu_ctx)
    sp = C.FESpace(gm, [C.H1(2, '')])
    names = '|'.join(hm.bnd_names)
    right, allm = hm.marker_mask(['right']), hm.marker_mask(hm.bnd_names)

    def trace(c):
        g = C.GridFunction(sp)
        g.Set((L.FN_POLY2, c + [0] * (10 - len(c))), names)
        return g.numpy()

    one, x, y = trace([1]), trace([0, 1]), trace([0, 0, 1])

    def form(t):
        a = C.BilinearForm(sp, [t])
        a.Assemble()
        return a.to_scipy()

    A = form(C.BndMass(0, right, coef=3.0))
    vals = [float(v @ (A @ v)) for v in (one, y)]
    say('  robin mass on x=1: u=1, y ->', vals)
    assert np.allclose(vals, [3.0, 1.0], rtol=0, atol=1e-12)
    A = form(C.BndUpwind(0, allm, coef=1.0, wind=(1.0, 0.0, 0.0)))
    vals = [float(v @ (A @ v)) for v in (one, x, y)]
    say('  no-backflow, wind (1,0,0): u=1, x, y ->', vals)
    assert np.allclose(vals, [1.0, 1.0, 1.0 / 3.0], rtol=0, atol=1e-12)
    # the wind as a P2 field on the faces (field_at_facet on tetrahedra): w = (x, 0, z^2); w.n is x on
    # x = 1, -x (= 0) on x = 0, z^2 on z = 1 and -z^2 (= 0) on z = 0, so for u = 1 the term is
    # int_{x=1} 1 + int_{z=1} 1 = 2 and for u = y it is 1/3 + 1/3
    ws = C.FESpace(gm, [C.H1(2, '', ncomp=3)])
    W = C.GridFunction(ws)
    W.Set([(L.FN_POLY2, [0, 1] + [0] * 8), const(0.0), (L.FN_POLY2, [0, 0, 0, 0, 0, 0, 1, 0, 0, 0])], names,
          component=0)
    A = form(C.BndUpwind(0, allm, coef=1.0, wind_gf=(W, [0, 1, 2])))
    vals = [float(v @ (A @ v)) for v in (one, y)]
    say('  no-backflow, wind field (x, 0, z^2): u=1, y ->', vals)
    assert np.allclose(vals, [2.0, 2.0 / 3.0], rtol=0, atol=1e-12)
    f = C.LinearForm(sp, [C.BndSource(0, right, fn=(L.FN_POLY2, [0, 0, 1, 0, 0, 0, 0, 0, 0, 0]), coef=2.0)])
    f.Assemble()
    b = f.vec.cpu().numpy()
    say('  boundary source 2y on x=1: b.1, b.y ->', b @ one, b @ y)
    assert abs(b @ one - 1.0) <= 1e-12 and abs(b @ y - 2.0 / 3.0) <= 1e-12
    m = Poisson(hm, dc=1.0, f=const(0.0), interp_ord=2, dirichlet={'u': {'left': const(1.0)}},
                neumann={'u': {'right': const(1.0)}}, linear_solver='CG', linear_tolerance=1e-13, ctx=gpu_ctx)
    s = StationarySolver(m)
    u = s.solve()
    e_l2, e_h1 = m.l2_error(u, 'u', (L.FN_POLY2, [1, 1, 0, 0, 0, 0, 0, 0, 0, 0]))
    say('  neumann exact', s.result, 'L2', e_l2, 'H1', e_h1)
    assert s.result.converged and e_l2 <= 1e-10 and e_h1 <= 1e-9


def test_boundary_terms_match_oracle_on_channel(gpu_ctx):
    """The boundary-facet terms against the oracle's restatement on the channel with holes: a straight
    outlet and curved, body-fitted hole boundaries give every facet orientation (matrix and load
    vector within relative Frobenius 1e-12)."""
    from opencmp_amd import core as C
    from opencmp_amd.mesh import channel_with_holes
    say('boundary terms vs oracle on the channel')
    hm = channel_with_holes(10)
    gm = C.Mesh(hm, gpu_ctx)
    sp = C.FESpace(gm, [C.H1(2, '')])
    t = fem.Topology(hm)
    cs = fem.Compound([fem.Space(t, 'H1', 2)])
    mask = hm.marker_mask(['outlet', 'cyl'])
    for kw, okw in (({'wind': (1.0, 0.3, 0.0)}, {'wind': (1.0, 0.3)}), ({}, {})):
        term = C.BndUpwind(0, mask, coef=0.7, **kw) if kw else C.BndMass(0, mask, coef=0.7)
        a = C.BilinearForm(sp, [term])
        a.Assemble()
        A = a.to_scipy()
        asm = fem.Assembler(cs)
        fem.assemble_bnd_facets(asm, 0, 0.7, mask, **okw)
        R = asm.matrix()
        rel = sp_norm(A - R) / sp_norm(R)
        say('  ', 'upwind' if kw else 'mass', 'rel', rel)
        assert rel <= 1e-12
    f = C.LinearForm(sp, [C.BndSource(0, mask, fn=(L.FN_POLY2, [0.5, 1.0, -2.0, 0.3, 0.0, 1.0]), coef=1.3)])
    f.Assemble()
    ref = fem.assemble_bnd_source(cs, 0, lambda X, Y: 0.5 + X - 2 * Y + 0.3 * X * X + Y * Y, mask, coef=1.3)
    rel = np.linalg.norm(f.vec.cpu().numpy() - ref) / np.linalg.norm(ref)
    say('  source rel', rel)
    assert rel <= 1e-12


def sp_norm(M):
    return float(np.sqrt((M.multiply(M)).sum()))


def test_ins_stress_traction_load(gpu_ctx):
    """The load side of the INS stress condition, dt h.v ds (opencmp/models/ins.py:299-305): with
    h = (1, 2y) on the right edge of the unit square and dt = 0.5, the x block of the load against
    the interpolant of 1 gives 0.5 * int 1 ds = 0.5 and the y block against the interpolant of y
    gives 0.5 * int 2 y^2 ds = 1/3; the zero traction adds nothing."""
    from opencmp_amd import core as C
    from opencmp_amd.models import INS, const
    say('ins stress traction')
    hm = unit_square(6)
    h = [const(1.0), (L.FN_POLY2, [0, 0, 2.0, 0, 0, 0])]
    m = INS(hm, kv=1.0, interp_ord=2, stress={'u': {'right': h, 'top': [const(0.0), const(0.0)]}}, ctx=gpu_ctx)
    f = C.LinearForm(m.fes, m.construct_linear(0.5))
    f.Assemble()
    b = f.vec.cpu().numpy()
    off = m.fes.block_offsets()
    bx, by = m.fes.comp_blocks[0]
    X, Y = hm.coords[:, 0], hm.coords[:, 1]
    nv = len(X)
    vx = b[off[bx]:off[bx] + nv].sum()              # vertex dofs only: the interpolant of 1
    vy = b[off[by]:off[by] + nv] @ Y
    say('  x block . 1', vx, 'y block . y', vy)
    assert abs(vx - 0.5) <= 1e-13 and abs(vy - 1.0 / 3.0) <= 1e-13
    assert np.abs(b[off[bx + 2]:]).max() == 0.0     # pressure block untouched


def test_oseen_picard_matches_oracle(gpu_ctx):
    """The Picard loop of opencmp/models/ins.py:323-359 on the GPU against the oracle's restatement
    (oracle/fem.py oseen_picard: the same Oseen form, scipy direct solves): unit square, P2/P1,
    kv = 0.1, parabolic inflow, walls, a zero-traction stress outlet, body force (0, 5 sin(pi x)
    sin(pi y)). The Picard errors ||W - u|| agree iteration by iteration, the loops stop after the
    same number of iterations, and the solutions agree."""
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    from test_oracle import _th_space, th_lift
    from opencmp_amd.models import INS, const
    from opencmp_amd.solvers import StationarySolver
    say('oseen picard vs oracle')
    hm = unit_square(4)
    t, cs, mask = _th_space(hm)
    x0 = th_lift(t, cs, mask, lambda x, y: 4 * y * (1 - y) + 0 * x, lambda x, y: 0 * x)
    b = fem.assemble_source(cs, 0, [lambda x, y: 0 * x, lambda x, y: 5 * np.sin(np.pi * x) * np.sin(np.pi * y)])
    xo, errs_o = fem.oseen_picard(cs, 0.1, x0, b, 1e-10, 1e-10, 40, stress_mask=hm.marker_mask(['right']))
    inflow = [(L.FN_POLY2, [0, 0, 4.0, 0, 0, -4.0]), const(0.0)]
    wall = [const(0.0), const(0.0)]
    m = INS(hm, kv=0.1, interp_ord=2, f=[const(0.0), (L.FN_SINPROD, (5.0,))],
            dirichlet={'u': {'left': inflow, 'top': wall, 'bottom': wall}}, stress={'u': {'right': None}},
            linear_solver='GMRes', preconditioner='None', linear_tolerance=1e-13, restart=200,
            nonlinear_tolerance=(1e-10, 1e-10), nonlinear_max_iters=40, ctx=gpu_ctx)
    s = StationarySolver(m)
    s.solve()
    eg = np.array(m.picard_errors)
    say('  GPU', len(eg), 'iterations, oracle', len(errs_o))
    say('  GPU errors   ', ' '.join('%.6e' % e for e in eg[:6]))
    say('  oracle errors', ' '.join('%.6e' % e for e in errs_o[:6]))
    assert len(eg) == len(errs_o)
    big = np.array(errs_o) > 1e-7
    np.testing.assert_allclose(eg[big], np.array(errs_o)[big], rtol=1e-6)
    x = s.gfu.numpy()
    nu = cs.spaces[0].ndof
    assert np.abs(x[:nu] - xo[:nu]).max() <= 1e-8 * np.abs(xo[:nu]).max()
