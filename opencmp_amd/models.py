"""Weak forms of the OpenCMP models on the GPU path.

Each model mirrors the reference class's ``construct_bilinear_time_ODE``,
``construct_bilinear_time_coefficient`` and ``construct_linear`` term by term,
but returns explicit integrator terms (opencmp_amd.core) instead of NGSolve
symbolic forms. Coefficients are constants, per-element arrays, closed-form
functions ``(FN_*, params)`` or grid functions.

Special functions follow opencmp/helpers/ngsolve_.py:48-70: ``alpha = nu / h``
with ``nu = ipc * p^2`` (opencmp/models/base_model.py:150) and ``h`` the
element size |det J|^(1/dim); ``dt`` multiplies every term as in the reference
(its ``Parameter(1.0)`` for stationary solves).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _lib as L
from . import core as C
from . import expr as E
from .mesh import Mesh as HostMesh

Fn = Tuple[int, Sequence[float]]


def const(v: float) -> Fn:
    return (L.FN_CONST, (float(v),))


def _shifted(fn: Fn, shift: float) -> Fn:
    """fn + shift for a closed form whose constant term is its first parameter."""
    fid, prm = fn
    if fid not in (L.FN_CONST, L.FN_POLY2):
        raise NotImplementedError('a shifted reference needs a constant, a quadratic or an expression')
    prm = list(prm) + [0.0] * (10 - len(prm))
    prm[0] += shift
    return (fid, tuple(prm))


ZERO = const(0.0)


class Model:
    """Common part of the GPU models (opencmp/models/base_model.py:39-987).

    Parameters mirror the reference config: ``interp_ord``, ``DG``, ``ipc``
    (interior penalty coefficient), boundary conditions as dicts
    {marker_name: Fn}, and the [SOLVER] keys (``linear_solver``,
    ``preconditioner``, ``linear_tolerance``, ``linear_max_iterations``) plus the
    GPU-path keys ``precision`` (fp64|mixed), ``restart``, ``s_step`` (5: s-step GMRES
    cycles where the operator allows them, the DG space with element-block Jacobi; 0: standard)
    and ``schur_iters`` / ``schur_tol`` (saddle-point models with preconditioner 'schur': > 0
    solves the pressure rows of every preconditioner step by at most that many PCG iterations on
    C + B diag(A)^{-1} B^T, fewer once the preconditioned residual has fallen by schur_tol; this
    makes GMRes run as FGMRes).
    """

    #: names of the model components in compound-space order
    components: List[str] = ['u']
    nonlinear = False

    def __init__(self, mesh: HostMesh, interp_ord: int = 2, DG: bool = False, ipc: float = 10.0,
                 dirichlet: Optional[Dict[str, Dict[str, Fn]]] = None, ctx: Optional[C.Context] = None,
                 linear_solver: str = 'CG', preconditioner: str = 'default', linear_tolerance: float = 1e-10,
                 linear_max_iterations: int = 100000, precision: str = 'fp64', restart: int = 50,
                 inner_tol: float = 0.0, s_step: int = 5, schur_iters: int = 0, schur_tol: float = 0.0,
                 t_param: Optional[List[E.Parameter]] = None, vel_iters: int = 0, vel_tol: float = 0.0,
                 bf16_precond: bool = False, schur_amg: int = 0, schur_lag: bool = False, schur_tri: bool = False):
        self.ctx = ctx or C.Context.default()
        # t_param[k]: the time of step n+1-k (opencmp/solvers/base_solver.py: t_param); config
        # expressions of t (boundary data, sources) are read at the level a form is built for
        self.t_param = t_param if t_param is not None else [E.Parameter(0.0) for _ in range(3)]
        self.host_mesh = mesh
        self.mesh = C.Mesh(mesh, self.ctx)
        self.interp_ord = interp_ord
        self.DG = DG
        self.ipc = ipc
        self.nu = ipc * interp_ord ** 2           # opencmp/models/base_model.py:150
        self.dirichlet = dirichlet or {}
        if linear_solver == 'default':
            linear_solver = 'direct'
        if linear_solver == 'direct':
            # the GPU path has no sparse direct factorisation; the reference's default
            # 'direct' maps to a preconditioned Krylov solve to a tolerance near machine precision,
            # without the iteration cap of the iterative solvers' config key
            # (DG spaces take the element-block Jacobi preconditioner, which CG does not accept)
            linear_solver = 'GMRes' if (self._indefinite_or_nonsymmetric() or DG) else 'CG'
            linear_tolerance = min(linear_tolerance, 1e-13)
            linear_max_iterations = max(linear_max_iterations, 1000000)
        self.preconditioner = 'local' if preconditioner in ('default', 'local') else preconditioner
        self.schur_iters = schur_iters if self.preconditioner == 'schur' else 0
        self.schur_tol = schur_tol
        # 'schur' with schur_amg > 0: AMG V-cycles on the Schur matrix instead of the inner PCG
        self.schur_amg = schur_amg if self.preconditioner == 'schur' else 0
        if self.schur_amg:
            self.schur_iters = 0
        # Picard iterations after a step's first refresh only the Jacobi rows of the Schur
        # preconditioner (its Schur step lags by at most the step's Picard iterations)
        self.schur_lag = bool(schur_lag) and bool(self.schur_iters or self.schur_amg)
        # the velocity rows as the upper block-triangular step after the Schur step
        self.schur_tri = bool(schur_tri) and bool(self.schur_iters or self.schur_amg)
        if (self.schur_iters or self.schur_amg) and linear_solver == 'GMRes':
            linear_solver = 'FGMRes'   # the inner PCG makes the preconditioner step nonlinear
        self.vel_iters = vel_iters if self.preconditioner == 'block_triangular' else 0
        self.vel_tol = vel_tol
        # mixed precision: keep the inner solves' matrices (Schur S, the velocity block) in BF16. Suits a
        # strongly diagonally dominant block (the diffuse-interface velocity block); a Laplacian-like
        # S (the INS channel's pressure Schur complement at 1.8M dofs) can lose definiteness in BF16
        # and stall its inner PCG
        self.bf16_precond = bool(bf16_precond) and precision == 'mixed'
        if self.preconditioner == 'block_triangular':
            if not self.vel_iters:
                raise ValueError("preconditioner 'block_triangular' needs vel_iters > 0")
            linear_solver = 'FGMRes'
        self.linear_solver = linear_solver
        if self.preconditioner in (None, 'None') and linear_solver in ('CG', 'MinRes'):
            raise ValueError('Preconditioner cannot be None if using CG or MinRes solvers.')
        self.linear_tolerance = linear_tolerance
        self.linear_max_iterations = linear_max_iterations
        self.precision = precision
        self.restart = restart
        self.inner_tol = inner_tol
        self.s_step = s_step
        self.fes = self._construct_fes()

    # ---------------------------------------------------------------- hooks
    def _indefinite_or_nonsymmetric(self) -> bool:
        return False

    def _fn(self, v, time_step: int = 0):
        """A boundary-data / source value as an Fn of the GPU terms: config expressions are bound to
        the time of step n+1-time_step; closed forms pass through; lists map per component."""
        if isinstance(v, list):
            return [self._fn(x, time_step) for x in v]
        if isinstance(v, E.Expr):
            return v.bind(self.ctx, self.t_param[time_step])
        return v

    def _construct_fes(self) -> C.FESpace:
        raise NotImplementedError

    def dirichlet_names(self, comp: str) -> str:
        return '|'.join(self.dirichlet.get(comp, {}).keys())

    def construct_bilinear_time_ODE(self, dt: float = 1.0) -> list:
        return []

    def construct_bilinear_time_coefficient(self, dt: float = 1.0) -> list:
        return []

    def construct_linear(self, dt: float = 1.0, time_step: int = 0) -> list:
        return []

    def time_derivative_terms(self) -> list:
        """U.V mass terms of every time-differentiated component (opencmp/models/base_model.py:418-497)."""
        out = []
        for comp in self.time_derivative_components():
            for b in self.fes.comp_blocks[self.components.index(comp)]:
                out.append(C.Vol('mass', b, b, coef=1.0, weight_gf=self._time_weight()))
        return out

    def time_derivative_components(self) -> List[str]:
        return list(self.components)

    def _time_weight(self):
        return None

    # ---------------------------------------------------------------- boundary conditions, solve
    def apply_dirichlet_bcs_to(self, gfu: C.GridFunction, time_step: int = 0) -> None:
        """gfu.Set(g, definedon=Boundaries(...)) per component (opencmp/models/base_model.py:321-341),
        with time-dependent data read at the time of step n+1-time_step."""
        for comp, bcs in self.dirichlet.items():
            if comp not in self.components:
                continue          # e.g. pressure data, which the reference does not impose either
            ci = self.components.index(comp)
            if not isinstance(self.fes.comps[ci], C.H1):
                continue
            for marker, g in bcs.items():
                gfu.Set(self._fn(g, time_step), marker, component=ci)

    def construct_preconditioners(self, a_assembled: C.BilinearForm) -> Optional[C.Preconditioner]:
        """ngs.Preconditioner(a, 'local') (opencmp/models/base_model.py:365-383)."""
        if self.preconditioner in (None, 'None'):
            return None
        if self.preconditioner == 'schur':
            pre = C.Preconditioner(a_assembled, 'schur', precision=self.precision,
                                   schur_blocks=self.schur_blocks(), schur_iters=self.schur_iters,
                                   schur_tol=self.schur_tol, schur_amg=self.schur_amg)
            if self.bf16_precond:
                pre.set_bf16(True)
            if self.schur_tri:
                pre.set_schur_triangular(True)
            return pre
        if self.preconditioner == 'block_triangular':
            pre = C.Preconditioner(a_assembled, 'schur', precision=self.precision, schur_blocks=self.schur_blocks())
            pre.set_block_triangular(self.schur_diagonal(), self.vel_iters, self.vel_tol)
            if self.bf16_precond:
                pre.set_bf16(True)
            return pre
        return C.Preconditioner(a_assembled, self.preconditioner, precision=self.precision)

    def schur_diagonal(self) -> torch.Tensor:
        """A positive diagonal approximation of the pressure Schur complement over the Schur rows
        (saddle-point models; the block-triangular preconditioner)."""
        raise NotImplementedError(f'{type(self).__name__} has no Schur complement approximation')

    def schur_blocks(self) -> List[int]:
        """Scalar blocks of the constraint (pressure) components of a saddle-point model."""
        return []

    def linear_solve(self, a_assembled: C.BilinearForm, L_assembled: C.LinearForm,
                     precond: Optional[C.Preconditioner], gfu: C.GridFunction) -> C.SolveResult:
        """Model.linear_solve (opencmp/models/base_model.py:886-947): solve on the free dofs starting
        from gfu (which carries the Dirichlet lift). Unlike the reference, returns the solver status."""
        return C.solve(a_assembled, L_assembled.vec, gfu.vec, precond, self.linear_solver,
                       tol=self.linear_tolerance, maxit=self.linear_max_iterations, precision=self.precision,
                       restart=self.restart, inner_tol=self.inner_tol, s_step=self.s_step)

    def solve_single_step(self, a: C.BilinearForm, Lf: C.LinearForm, pre, gfu: C.GridFunction,
                          reassemble=None, assembled: bool = False) -> C.SolveResult:
        return self.linear_solve(a, Lf, pre, gfu)

    # ---------------------------------------------------------------- error norms
    def l2_error(self, gfu: C.GridFunction, comp: str, exact: Fn = ZERO, block: int = 0) -> Tuple[float, float]:
        """(L2 error, H1-seminorm error) of one scalar block against a closed form (calc_error)."""
        import ctypes
        b = self.fes.comp_blocks[self.components.index(comp)][block]
        c = (ctypes.c_double * 10)(*([float(v) for v in exact[1]] + [0.0] * (10 - len(exact[1]))))
        out = (ctypes.c_double * 2)()
        L.check(self.ctx.lib.ocmp_error_norms(self.fes.h, b, ctypes.c_void_p(gfu.vec.data_ptr()), None, exact[0],
                                              c, out), 'error_norms')
        return out[0], out[1]

    def _vec_norms(self, blocks, x: torch.Tensor, fns=None, ref: Optional[torch.Tensor] = None):
        """(L2, L1, H1-seminorm) of the vector of the given scalar blocks against per-block Fns or a
        reference vector on the same space (ocmp_error_norms_vec)."""
        import ctypes
        nb = len(blocks)
        bl = (ctypes.c_int32 * nb)(*blocks)
        fid = (ctypes.c_int32 * nb)()
        c = (ctypes.c_double * (10 * nb))()
        for k, f in enumerate(fns or []):
            i, prm = E.resolve_fn(f)
            fid[k] = i
            for j, v in enumerate(prm):
                c[10 * k + j] = v
        out = (ctypes.c_double * 3)()
        L.check(self.ctx.lib.ocmp_error_norms_vec(self.fes.h, nb, bl, ctypes.c_void_p(x.data_ptr()),
                                                  None if ref is None else ctypes.c_void_p(ref.data_ptr()),
                                                  fid, c, out), 'error_norms_vec')
        return out[0], out[1], out[2]

    def norm(self, norm_type: str, gfu: C.GridFunction, comp: str, ref, average: bool = False) -> float:
        """norm() of opencmp/helpers/error.py:91-128 for one component against its reference (one
        bound expression or closed form per scalar block). l1_norm integrates the pointwise Euclidean
        norm (:70-72), l2_norm is the L2 norm (:75-77), linfinity_norm the largest coefficient of
        the interpolant of the difference (:80-88). average: both sides first offset by their
        ``mean`` (:35-68), which the reference computes as sqrt(int u^2), and Set into the space."""
        ci = self.components.index(comp)
        blocks = self.fes.comp_blocks[ci]
        ref = ref if isinstance(ref, list) else [ref]
        if len(ref) != len(blocks):
            raise ValueError(f'{len(ref)} reference components for {comp!r}, which has {len(blocks)}')
        ref_vec = None
        if average or norm_type == 'linfinity_norm':
            shift = 0.0
            if average:
                if len(blocks) != 1:
                    raise NotImplementedError('error_average for vector components')
                zero = torch.zeros_like(gfu.vec)
                a_h = self._vec_norms(blocks, gfu.vec, [ZERO])[0]
                a_g = self._vec_norms(blocks, zero, ref)[0]
                shift = a_h - a_g
            tmp = C.GridFunction(self.fes)
            tmp.Set([r.scaled(shift=shift) if hasattr(r, 'scaled') else _shifted(r, shift) for r in ref],
                    component=ci)
            ref_vec = tmp.vec
        if norm_type == 'linfinity_norm':
            import ctypes
            out = (ctypes.c_double * 1)()
            m = 0.0
            for blk in blocks:
                L.check(self.ctx.lib.ocmp_linf_norm(self.fes.h, blk, ctypes.c_void_p(gfu.vec.data_ptr()),
                                                    ctypes.c_void_p(ref_vec.data_ptr()), out), 'linf_norm')
                m = max(m, out[0])
            return m
        l2, l1, _ = self._vec_norms(blocks, gfu.vec, None if ref_vec is not None else ref, ref=ref_vec)
        if norm_type == 'l2_norm':
            return l2
        if norm_type == 'l1_norm':
            return l1
        raise ValueError(f'{norm_type} has not been implemented yet.')

    def divergence(self, gfu: C.GridFunction, comp: str) -> float:
        """sqrt(int (div u)^2) of a vector component (opencmp/helpers/error.py:151-167)."""
        import ctypes
        b0 = self.fes.comp_blocks[self.components.index(comp)][0]
        out = (ctypes.c_double * 1)()
        L.check(self.ctx.lib.ocmp_divergence_norm(self.fes.h, b0, ctypes.c_void_p(gfu.vec.data_ptr()), out),
                'divergence_norm')
        return out[0]

    def linf_diff(self, a: C.GridFunction, b: Optional[C.GridFunction], comp: str) -> float:
        """max |a - b| over the component's dof coefficients (opencmp/helpers/error.py:80-88 with a
        grid-function reference on the same space; b None: max |a|)."""
        import ctypes
        out = (ctypes.c_double * 1)()
        m = 0.0
        for blk in self.fes.comp_blocks[self.components.index(comp)]:
            L.check(self.ctx.lib.ocmp_linf_norm(self.fes.h, blk, ctypes.c_void_p(a.vec.data_ptr()),
                                                None if b is None else ctypes.c_void_p(b.vec.data_ptr()), out),
                    'linf_norm')
            m = max(m, out[0])
        return m

    def picard_norms(self, W: C.GridFunction, gfu: C.GridFunction) -> Tuple[float, float]:
        """(norm('l2_norm', W, u), mean(u)) of the Picard test (opencmp/models/ins.py:346-347) in one
        pass (ocmp_l2_pair_norms)."""
        import ctypes
        blocks = self.fes.comp_blocks[self.components.index('u')]
        bl = (ctypes.c_int32 * len(blocks))(*blocks)
        out = (ctypes.c_double * 2)()
        L.check(self.ctx.lib.ocmp_l2_pair_norms(self.fes.h, len(blocks), bl, ctypes.c_void_p(W.vec.data_ptr()),
                                                ctypes.c_void_p(gfu.vec.data_ptr()), out), 'l2_pair_norms')
        return out[0], out[1]

    def l2_diff(self, a: C.GridFunction, b: C.GridFunction, comp: str) -> float:
        """sqrt(sum over the component's blocks of int (a - b)^2) -- norm('l2_norm', W, u) of the Picard loop."""
        import ctypes
        c = (ctypes.c_double * 10)()
        out = (ctypes.c_double * 2)()
        tot = 0.0
        for blk in self.fes.comp_blocks[self.components.index(comp)]:
            L.check(self.ctx.lib.ocmp_error_norms(self.fes.h, blk, ctypes.c_void_p(a.vec.data_ptr()),
                                                  ctypes.c_void_p(b.vec.data_ptr()), L.FN_CONST, c, out),
                    'error_norms')
            tot += out[0] ** 2
        return float(np.sqrt(tot))


class Poisson(Model):
    """opencmp/models/poisson.py:28-168: -div(dc grad u) = f, H1 (CG) or L2 with SIPG (DG).
    ``neumann``: {'u': {marker: h}} adds dt dc h v ds to the load (:151-154); ``robin``:
    {'u': {marker: (r, q)}} adds -dt dc r u v ds to the form (:94-96) and -dt dc r q v ds to the
    load (:156-158); h and q closed forms, r a number; H1 (triangles or tetrahedra)."""
    components = ['u']

    def __init__(self, mesh, dc: float = 1.0, f: Fn = ZERO, neumann=None, robin=None, **kw):
        self.dc = dc
        self.f = f
        self.neumann = neumann or {}
        self.robin = robin or {}
        super().__init__(mesh, **kw)
        if (self.neumann.get('u') or self.robin.get('u')) and self.DG:
            raise ValueError('Neumann/Robin data are implemented for the H1 (CG) Poisson model')

    def _construct_fes(self):
        if self.DG:
            return C.FESpace(self.mesh, [C.L2(self.interp_ord)], dgjumps=True)
        return C.FESpace(self.mesh, [C.H1(self.interp_ord, self.dirichlet_names('u'))])

    def construct_bilinear_time_ODE(self, dt=1.0):
        a = [C.Vol('gradgrad', 0, 0, coef=dt * self.dc)]                      # poisson.py:78
        if self.DG and self.dirichlet.get('u'):
            # Dirichlet Nitsche terms (poisson.py:81-91) are part of the facet term below
            pass
        for marker, (r, q) in self.robin.get('u', {}).items():                # poisson.py:94-96
            a.append(C.BndMass(0, self.mesh.marker_mask(marker), coef=-dt * self.dc * float(r)))
        return a

    def construct_bilinear_time_coefficient(self, dt=1.0):
        if not self.DG:
            return []                                                         # poisson.py:124
        dmask = self.mesh.marker_mask(self.dirichlet_names('u')) if self.dirichlet.get('u') else 0
        # poisson.py:106-121 (interior SIPG) and :81-91 (boundary Nitsche), no convection
        return [C.DGFacet(0, eps=dt * self.dc, alpha=self.nu, dirichlet=dmask, upwind=False)]

    def construct_linear(self, dt=1.0, time_step=0):
        f = self._fn
        Lf = [C.Source(0, f(self.f, time_step), coef=dt)]                     # poisson.py:133
        if self.DG:
            for marker, g in self.dirichlet.get('u', {}).items():             # poisson.py:136-146
                Lf.append(C.DGDirichletData(0, eps=dt * self.dc, alpha=self.nu, wind=(0.0, 0.0),
                                            dirichlet=self.mesh.marker_mask(marker), fn=f(g, time_step)))
        for marker, h in self.neumann.get('u', {}).items():                   # poisson.py:151-154
            Lf.append(C.BndSource(0, self.mesh.marker_mask(marker), fn=f(h, time_step), coef=dt * self.dc))
        for marker, (r, q) in self.robin.get('u', {}).items():                # poisson.py:156-158
            Lf.append(C.BndSource(0, self.mesh.marker_mask(marker), fn=f(q, time_step),
                                  coef=-dt * self.dc * float(r)))
        return Lf


class DGTransport(Model):
    """One species of MultiComponentINS with a fixed velocity, DG (opencmp/models/multi_component_ins.py):

    D grad c . grad r - c w . grad r                             (bulk, :223-224, :258-260)
    skeleton SIPG + upwind                                       (:185-197)
    Dirichlet Nitsche + upwind, upwind on other boundaries       (:262-301)
    L = f r + Dirichlet data                                     (:343-349, :371-386)
    """
    components = ['c']

    def __init__(self, mesh, D: float = 1e-3, wind=(1.0, 0.0), f: Fn = const(1.0), **kw):
        self.D = D
        self.wind = tuple(float(v) for v in wind)
        self.f = f
        kw.setdefault('DG', True)
        kw.setdefault('linear_solver', 'GMRes')
        super().__init__(mesh, **kw)
        if not self.DG:
            raise NotImplementedError('the GPU transport model is implemented with DG (L2) spaces')

    def _indefinite_or_nonsymmetric(self):
        return True

    def _construct_fes(self):
        return C.FESpace(self.mesh, [C.L2(self.interp_ord)], dgjumps=True)

    def construct_bilinear_time_ODE(self, dt=1.0):
        w = self.wind
        return [C.Vol('gradgrad', 0, 0, coef=dt * self.D),
                C.Vol('conv', 0, 0, coef=-dt, wind=(w[0], w[1], 0.0))]

    def construct_bilinear_time_coefficient(self, dt=1.0):
        dmask = self.mesh.marker_mask(self.dirichlet_names('c')) if self.dirichlet.get('c') else 0
        return [C.DGFacet(0, eps=self.D, alpha=self.nu, wind=self.wind, dirichlet=dmask, coef=dt)]

    def construct_linear(self, dt=1.0, time_step=0):
        Lf = [C.Source(0, self._fn(self.f, time_step), coef=dt)]
        for marker, g in self.dirichlet.get('c', {}).items():
            Lf.append(C.DGDirichletData(0, eps=self.D, alpha=self.nu, wind=self.wind,
                                        dirichlet=self.mesh.marker_mask(marker), fn=self._fn(g, time_step),
                                        coef=dt))
        return Lf


class Stokes(Model):
    """opencmp/models/stokes.py:28-147 (CG Taylor-Hood): kv grad u : grad v - div(u) q - div(v) p - 1e-10 p q.
    ``stress``: {'u': {marker: traction}} adds dt h.v ds of each traction h (one Fn per velocity
    component, or one for all; None: the do-nothing condition), stokes.py:131-137."""
    components = ['u', 'p']

    def __init__(self, mesh, kv: float = 1.0, f=None, stress: Optional[Dict[str, Dict[str, Fn]]] = None, **kw):
        self.kv = kv
        self.f = f
        self.stress = stress or {}
        kw.setdefault('linear_solver', 'GMRes')
        super().__init__(mesh, **kw)

    def _indefinite_or_nonsymmetric(self):
        return True

    def schur_blocks(self):
        return list(self.fes.comp_blocks[1])

    def _construct_fes(self):
        d = self.host_mesh.dim
        return C.FESpace(self.mesh, [C.H1(self.interp_ord, self.dirichlet_names('u'), ncomp=d),
                                     C.H1(self.interp_ord - 1, self.dirichlet_names('p'))])

    def time_derivative_components(self):
        return ['u']

    def _schur_weight(self):
        return None

    def schur_diagonal(self) -> torch.Tensor:
        """The lumped pressure mass over kv, int psi_i / kv (weighted by the model's pressure weight):
        the Schur complement B A^{-1} B^T of the viscous operator is spectrally equivalent to the
        pressure mass over the viscosity, and the lumped mass (the P1 functions sum to one) is
        equivalent to the mass matrix."""
        pb = self.fes.comp_blocks[1][0]
        if len(self.fes.comp_blocks[1]) != 1 or self.interp_ord != 2:
            raise NotImplementedError('the lumped pressure mass needs one P1 pressure block')
        f = C.LinearForm(self.fes, [C.Source(pb, (L.FN_CONST, (1.0 / self.kv,)), weight_gf=self._schur_weight())])
        f.Assemble()
        sl = self.fes.block_slice(pb)
        return f.vec[sl].clone()

    def _velocity_terms(self, dt):
        ub = self.fes.comp_blocks[0]
        return [C.Vol('gradgrad', b, b, coef=dt * self.kv) for b in ub]

    def _pressure_terms(self, dt, weight=None):
        ub = self.fes.comp_blocks[0]
        pb = self.fes.comp_blocks[1][0]
        t = []
        for k, b in enumerate(ub):
            t.append(C.Vol('dtrial', b, pb, coef=-dt, dir=k, weight_gf=weight))       # -div(u) q
            t.append(C.Vol('dtest', pb, b, coef=-dt, dir=k, weight_gf=weight))        # -div(v) p
        t.append(C.Vol('mass', pb, pb, coef=-1e-10 * dt, weight_gf=weight))          # -1e-10 p q
        return t

    def construct_bilinear_time_ODE(self, dt=1.0):
        return self._velocity_terms(dt)

    def construct_bilinear_time_coefficient(self, dt=1.0):
        return self._pressure_terms(dt)

    def construct_linear(self, dt=1.0, time_step=0):
        ub = self.fes.comp_blocks[0]
        Lf = []
        if self.f is not None:
            fs = self.f if isinstance(self.f, list) else [self.f] * len(ub)
            if len(fs) != len(ub):
                raise ValueError(f'source has {len(fs)} components, the velocity has {len(ub)}')
            Lf += [C.Source(b, self._fn(fk, time_step), coef=dt) for b, fk in zip(ub, fs)]
        for marker, h in self.stress.get('u', {}).items():                 # stokes.py:131-137, ins.py:299-305
            if h is None:
                continue
            fns = h if isinstance(h, list) else [h] * len(ub)             # one Fn or one per component
            if len(fns) != len(ub):
                raise ValueError(f'traction on {marker!r} has {len(fns)} components, the velocity has {len(ub)}')
            for b, fn in zip(ub, fns):
                if isinstance(fn, E.Expr) or hasattr(fn, 'fn_args') or any(float(c) != 0.0 for c in fn[1]):
                    Lf.append(C.BndSource(b, self.host_mesh.marker_mask(marker), fn=self._fn(fn, time_step),
                                          coef=dt))
        return Lf


class INS(Stokes):
    """opencmp/models/ins.py (CG, Oseen linearisation): adds -(u (x) w) : grad v with the wind w the
    previous iterate (:191-195); solved by the Picard loop of solve_single_step (:323-359).
    ``stress``: {'u': {marker: traction}} boundaries with the stress condition: each gets the
    no-backflow term dt max(w.n, 0) u.v (:210-216) and the load dt h.v ds of its traction h (one
    closed form per velocity component, or one for all; None or zero: the do-nothing outlet,
    :299-305)."""
    nonlinear = True

    def __init__(self, mesh, kv: float = 1e-3, nonlinear_tolerance=(1e-8, 1e-8), nonlinear_max_iters: int = 10,
                 stress: Optional[Dict[str, Dict[str, Fn]]] = None, **kw):
        self.abs_tol, self.rel_tol = nonlinear_tolerance
        self.nonlinear_max_iters = nonlinear_max_iters
        super().__init__(mesh, kv=kv, stress=stress, **kw)
        self.W = C.GridFunction(self.fes)    # linearisation point (opencmp/models/ins.py:109-116)

    def set_linearisation_point(self, gfu: C.GridFunction) -> None:
        """W = the initial condition's velocity (opencmp/models/ins.py:125-128)."""
        self.W.vec.copy_(gfu.vec)

    def construct_bilinear_time_ODE(self, dt=1.0):
        ub = self.fes.comp_blocks[0]
        terms = self._velocity_terms(dt)
        for b in ub:
            # -(u (x) w) : grad v = -sum_i u_i (w . grad v_i)
            terms.append(C.Vol('conv', b, b, coef=-dt, wind_gf=(self.W, ub)))
        names = list(self.stress.get('u', {}).keys())
        if names:
            mask = self.host_mesh.marker_mask(names)
            for b in ub:
                terms.append(C.BndUpwind(b, mask, coef=dt, wind_gf=(self.W, ub)))
        return terms

    def solve_single_step(self, a, Lf, pre, gfu, reassemble=None, assembled=False):
        """Picard iteration (opencmp/models/ins.py:323-359): re-assemble with the current wind,
        solve, stop when ||W - u||_L2 < abs + rel * ||u||_L2. assembled: `a` and `pre` already
        hold the first iteration's operator (the caller assembled it with the current wind)."""
        it = 1
        res = None
        self.linear_iterations = 0
        self.precond_iterations = 0   # inner PCG iterations of a Schur inner solve
        self.picard_log = []      # (iterations, refinement steps, initial rel residual, Picard error) per iteration
        self.picard_errors = []   # ||W - u||_L2 per iteration, full precision
        while True:
            # the Dirichlet values are those of this step at every Picard iteration, and the solves
            # leave the Dirichlet dofs as they are (their Krylov rows are zero), so one lift suffices
            if it == 1:
                self.apply_dirichlet_bcs_to(gfu)
            if reassemble is not None and not (assembled and it == 1):
                reassemble(full=it == 1 or not getattr(self, 'schur_lag', False))
            res = self.linear_solve(a, Lf, pre, gfu)
            self.linear_iterations += res.iters
            self.precond_iterations += res.precond_iters
            err, unorm = self.picard_norms(self.W, gfu)
            self.picard_errors.append(err)
            self.picard_log.append((res.iters, res.outer, float('%.3g' % res.initial_rel_residual), float('%.3g' % err)))
            it += 1
            self.W.vec.copy_(gfu.vec)
            self.picard_converged = err < self.abs_tol + self.rel_tol * unorm
            self.picard_error = err
            if self.picard_converged or it > self.nonlinear_max_iters:
                break
        self.picard_iterations = it - 1
        return res


class StokesDIM(Stokes):
    """opencmp/models/stokes_dim.py:28-153 with phi = 1 in the fluid, 0 in the solid:

    kv grad u : grad v phi + alpha u.v (1 - phi)
    (-div(u) q - div(v) p - 1e-10 p q) phi - p div(v) (1 - phi)
    L = f.v phi
    """

    def __init__(self, mesh, phi: torch.Tensor, phi_order: int = 1, penalty: float = 1e4, **kw):
        self.phi_order = phi_order
        self.penalty = penalty
        self._phi_vec = phi
        super().__init__(mesh, **kw)
        self.phi_space = C.FESpace(self.mesh, [C.H1(phi_order)])
        self.phi = C.GridFunction(self.phi_space, phi.to(self.ctx.torch_device, torch.float64))

    def construct_bilinear_time_ODE(self, dt=1.0):
        ub = self.fes.comp_blocks[0]
        wphi = (self.phi, 0, L.WEIGHT_FIELD)
        w1m = (self.phi, 0, L.WEIGHT_ONE_MINUS_FIELD)
        t = [C.Vol('gradgrad', b, b, coef=dt * self.kv, weight_gf=wphi) for b in ub]
        t += [C.Vol('mass', b, b, coef=dt * self.penalty, weight_gf=w1m) for b in ub]
        return t

    def construct_bilinear_time_coefficient(self, dt=1.0):
        ub = self.fes.comp_blocks[0]
        pb = self.fes.comp_blocks[1][0]
        t = self._pressure_terms(dt, weight=(self.phi, 0, L.WEIGHT_FIELD))
        for k, b in enumerate(ub):
            t.append(C.Vol('dtest', pb, b, coef=-dt, dir=k, weight_gf=(self.phi, 0, L.WEIGHT_ONE_MINUS_FIELD)))
        return t

    def construct_linear(self, dt=1.0, time_step=0):
        if self.f is None:
            return []
        ub = self.fes.comp_blocks[0]
        return [C.Source(b, self._fn(fk, time_step), coef=dt, weight_gf=(self.phi, 0, L.WEIGHT_FIELD))
                for b, fk in zip(ub, self.f)]

    def _time_weight(self):
        return (self.phi, 0, L.WEIGHT_FIELD)                 # opencmp/solvers/time_integration_schemes.py:678-681

    def _schur_weight(self):
        return (self.phi, 0, L.WEIGHT_FIELD)


class MultiComponentINS(Model):
    """opencmp/models/multi_component_ins.py with ``velocity_fixed = True``: species carried by a fixed
    velocity w (the initial condition's; a constant vector here). Per species c with test function
    r, every term times dt:

        D grad c . grad r  -  c w . grad r                                 (:223-224, :258-260)
        r c max(w.n, 0) on total-flux boundaries (neither Dirichlet nor Neumann)   (:262-275)
        r c (w.n) on Neumann boundaries                                    (:277-285)
        DG: skeleton SIPG (D != 0) + upwind [r](c max(w.n,0) + c' min(w.n,0))    (:171-197),
            Dirichlet facets D (alpha r c - r grad c.n - c grad r.n) + r c max(w.n, 0)   (:287-301)
        L: f r;  -h r on total-flux and Neumann boundaries (:357-367);
           DG Dirichlet data D g (alpha r - grad r.n) - r g min(w.n, 0)  (:371-386)

    H1 species take their Dirichlet data as a lift (strongly), L2 species through the facet terms.
    """
    nonlinear = False

    def __init__(self, mesh, components: Sequence[str], kinds: Dict[str, str], D: Dict[str, float],
                 f: Dict[str, Fn], wind=(0.0, 0.0), dirichlet=None, neumann=None, total_flux=None, **kw):
        self.components = list(components)
        self.kinds = dict(kinds)
        for c in self.components:
            if self.kinds[c] not in ('H1', 'L2'):
                raise NotImplementedError(f'species {c!r} in a {self.kinds[c]} space')
        self.D = dict(D)
        self.f = dict(f)
        self.wind = tuple(float(v) for v in wind) + (0.0,) * (3 - len(wind))
        self.neumann = neumann or {}
        self.total_flux = total_flux or {}
        kw.setdefault('linear_solver', 'GMRes')
        super().__init__(mesh, dirichlet=dirichlet, **kw)
        for c in self.components:
            if self.kinds[c] == 'L2' and self.neumann.get(c):
                raise NotImplementedError('Neumann boundaries of a DG species (its facet term upwinds every '
                                          'boundary facet)')

    def _indefinite_or_nonsymmetric(self):
        return True

    def _construct_fes(self):
        comps = []
        for c in self.components:
            if self.kinds[c] == 'H1':
                comps.append(C.H1(self.interp_ord, self.dirichlet_names(c)))
            else:
                comps.append(C.L2(self.interp_ord))
        return C.FESpace(self.mesh, comps, dgjumps=self.DG)

    def _markers(self, c: str):
        names = list(self.host_mesh.bnd_names)
        d = set(self.dirichlet.get(c, {}).keys())
        n = set(self.neumann.get(c, {}).keys())
        return [m for m in names if m not in d and m not in n], sorted(n), sorted(d)

    def construct_bilinear_time_ODE(self, dt=1.0):
        w = self.wind
        terms = []
        for k, c in enumerate(self.components):
            b = self.fes.comp_blocks[k][0]
            if self.D[c] > 0:
                terms.append(C.Vol('gradgrad', b, b, coef=dt * self.D[c]))
            terms.append(C.Vol('conv', b, b, coef=-dt, wind=w))
            if self.kinds[c] == 'H1':
                tf, nm, _ = self._markers(c)
                if tf:
                    terms.append(C.BndUpwind(b, self.host_mesh.marker_mask(tf), coef=dt, wind=w))
                if nm:
                    if self.D[c] == 0:
                        raise ValueError('Trying to apply a neuman boundary condition for purely advective flow '
                                         '(diffusion coefficient is 0).')
                    terms.append(C.BndUpwind(b, self.host_mesh.marker_mask(nm), coef=dt, wind=w, signed=True))
        return terms

    def construct_bilinear_time_coefficient(self, dt=1.0):
        terms = []
        for k, c in enumerate(self.components):
            if self.kinds[c] != 'L2':
                continue
            b = self.fes.comp_blocks[k][0]
            dm = self.host_mesh.marker_mask(list(self.dirichlet.get(c, {}).keys())) if self.dirichlet.get(c) else 0
            terms.append(C.DGFacet(b, eps=self.D[c], alpha=self.nu, wind=self.wind[:2], dirichlet=dm,
                                   sipg=self.D[c] != 0, upwind=True, coef=dt))
        return terms

    def construct_linear(self, dt=1.0, time_step=0):
        Lf = []
        for k, c in enumerate(self.components):
            b = self.fes.comp_blocks[k][0]
            Lf.append(C.Source(b, self._fn(self.f.get(c, ZERO), time_step), coef=dt))
            for marker, h in list(self.total_flux.get(c, {}).items()) + list(self.neumann.get(c, {}).items()):
                if not isinstance(h, E.Expr) and not hasattr(h, 'fn_args') and \
                        all(float(v) == 0.0 for v in h[1]):
                    continue
                if self.kinds[c] == 'L2':
                    raise NotImplementedError('non-zero flux data on a DG species')
                Lf.append(C.BndSource(b, self.host_mesh.marker_mask(marker), fn=self._fn(h, time_step), coef=-dt))
            if self.kinds[c] == 'L2':
                for marker, g in self.dirichlet.get(c, {}).items():
                    Lf.append(C.DGDirichletData(b, eps=self.D[c], alpha=self.nu, wind=self.wind[:2],
                                                dirichlet=self.host_mesh.marker_mask(marker),
                                                fn=self._fn(g, time_step), coef=dt))
        return Lf
