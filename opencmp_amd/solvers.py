"""Solver drivers of the GPU path: the caller side of the assemble+solve boundary.

Mirrors opencmp/solvers/: ``Solver._assemble`` (base_solver.py:368-377),
``StationarySolver`` (stationary.py:29-85) and the implicit-Euler step of
``TransientMultiStepSolver`` (transient_multistep.py:32-169) built by
``implicit_euler`` (time_integration_schemes.py:79-132, :663-685):

    A_step = M + dt (A_ode + A_coef),   L_step = dt L + M u_n

Differences from the reference, all on the performance side: the stationary
path assembles once (the reference assembles at :812 and again at :620), and
a transient linear model with time-independent coefficients keeps its
assembled step matrix instead of re-assembling every step; the mass matrix M is
assembled once and applied as a matrix-vector product for M u_n.
"""
from __future__ import annotations

import ctypes
import time
from typing import Optional

import torch

from . import _lib as L
from . import core as C
from .models import Model


class StationarySolver:
    def __init__(self, model: Model, gfu0: Optional[C.GridFunction] = None):
        """gfu0: the initial condition, the solve's starting point (base_solver.py:813)."""
        self.model = model
        self.gfu = C.GridFunction(model.fes)
        if gfu0 is not None:
            self.gfu.vec.copy_(gfu0.vec)
        self.a = None
        self.L = None
        self.pre = None
        self.result = None
        self.timings = {}

    def _create_linear_and_bilinear_forms(self):
        m = self.model
        self.a = C.BilinearForm(m.fes, m.construct_bilinear_time_coefficient(1.0) + m.construct_bilinear_time_ODE(1.0))
        self.L = C.LinearForm(m.fes, m.construct_linear(1.0))

    def _create_preconditioners(self):
        self.pre = self.model.construct_preconditioners(self.a)

    def _assemble(self, full: bool = True):
        """Solver._assemble: a.Assemble(), L.Assemble(), preconditioner.Update() (full=False: the
        Jacobi rows only, keeping a lagged Schur step)."""
        self.a.Assemble()
        self.L.Assemble()
        if self.pre is not None:
            self.pre.Update() if full else self.pre.UpdateJacobi()

    def solve(self) -> C.GridFunction:
        t0 = time.perf_counter()
        self._create_linear_and_bilinear_forms()
        self._create_preconditioners()
        t1 = time.perf_counter()
        self._assemble()
        self.model.apply_dirichlet_bcs_to(self.gfu)
        reassemble = self._assemble if self.model.nonlinear else None
        self.result = self.model.solve_single_step(self.a, self.L, self.pre, self.gfu, reassemble=reassemble)
        torch.cuda.synchronize()
        self.timings = {'setup_s': t1 - t0, 'assemble_solve_s': time.perf_counter() - t1}
        return self.gfu


class ImplicitEulerSolver:
    """Fixed-step implicit Euler (time_integration_schemes.py:79-132)."""

    def __init__(self, model: Model, dt: float, t_range=(0.0, 1.0), gfu0: Optional[C.GridFunction] = None,
                 reassemble_each_step: Optional[bool] = None, guess=None, linearize_at_guess: bool = False):
        """guess: an optional :class:`opencmp_amd.learned.GuessPolicy` that records the
        solution history (and snapshots for training) and proposes initial-guess candidates
        for each step (one is kept only if it lowers the FP64 residual of the step's first
        system). For a nonlinear model that first system is linearised about the previous
        solution, as in the reference, and the kept guess is only the Krylov start vector;
        linearize_at_guess makes the kept guess also the first linearisation point (the
        Picard iterate can then start closer to the fixed point, at the risk of a harder
        first system)."""
        self.model = model
        self.guess = guess
        self.linearize_at_guess = bool(linearize_at_guess)
        self.dt = dt
        self.t0, self.t1 = t_range
        self.gfu = C.GridFunction(model.fes)
        self.gfu_prev = C.GridFunction(model.fes)
        if gfu0 is not None:
            self.gfu_prev.vec.copy_(gfu0.vec)
        fes = model.fes
        mass = model.time_derivative_terms()
        self.a = C.BilinearForm(fes, mass + model.construct_bilinear_time_coefficient(dt) +
                                model.construct_bilinear_time_ODE(dt))
        self.M = C.BilinearForm(fes, mass)
        self.L = C.LinearForm(fes, model.construct_linear(dt))
        self.pre = model.construct_preconditioners(self.a)
        self.tmp = model.ctx.zeros(fes.ndof)
        self.reassemble_each_step = model.nonlinear if reassemble_each_step is None else reassemble_each_step
        self.M.Assemble()
        self._assemble_a()
        self.step_results = []
        self.guess_log = []
        self._fixed = None  # device mask of the space's fixed (non-free) dofs, for the guess candidates
        if self.guess is not None:
            self.guess.push(self.gfu_prev.vec)

    def _assemble_a(self, full: bool = True):
        self.a.Assemble()
        if self.pre is not None:
            self.pre.Update() if full else self.pre.UpdateJacobi()

    def _rhs(self):
        """L_step = dt L + M u_n (time_integration_schemes.py:663-685)."""
        self.L.Assemble()
        self.M.mult(self.gfu_prev.vec, self.tmp)
        n = self.model.fes.ndof
        L.check(self.model.ctx.lib.ocmp_axpby(self.model.ctx.h, n, 1.0, ctypes.c_void_p(self.tmp.data_ptr()), 1.0,
                                              ctypes.c_void_p(self.L.vec.data_ptr())), 'axpby')

    def step(self):
        self._rhs()
        # initial guess: previous solution (what the reference's gfu holds), then the Dirichlet lift
        self.gfu.vec.copy_(self.gfu_prev.vec)
        if self.reassemble_each_step and not self.model.nonlinear:
            self._assemble_a()
        self.model.apply_dirichlet_bcs_to(self.gfu)
        assembled = False
        pending = None
        if self.guess is not None and self.guess.ready():
            if self.model.nonlinear:
                # the step's first Picard system, linearised about the previous solution (W holds it)
                self._assemble_a()
                assembled = True
            if self.model.nonlinear and self.linearize_at_guess:
                # the choice decides the linearisation point: read it now
                if self._learned_guess() != 'plain':
                    self.model.W.vec.copy_(self.gfu.vec)
                    assembled = False
            else:
                # the choice only sets the Krylov start vector: made on the device, read after the solve
                pending = self._learned_guess(wait=False)
        reassemble = self._assemble_a if self.model.nonlinear else None
        res = self.model.solve_single_step(self.a, self.L, self.pre, self.gfu, reassemble=reassemble,
                                           assembled=assembled)
        if pending is not None:
            self.guess_log.append(pending())
        self.step_results.append(res)
        self.gfu_prev.vec.copy_(self.gfu.vec)
        if self.guess is not None:
            self.guess.push(self.gfu.vec)
        return res

    def _learned_guess(self, wait: bool = True):
        """Replace the plain initial guess (previous solution + Dirichlet lift) by the
        policy's candidate with the lowest FP64 residual against the current operator,
        if that beats the plain one; returns the choice (wait=True), or with wait=False a
        function that returns (choice, plain residual, best candidate's residual) once the
        device has made the choice (no synchronisation before the solve is enqueued)."""
        from .learned import choose_best, choose_best_async
        if self._fixed is None:
            self._fixed = torch.from_numpy(~self.model.fes.FreeDofs()).to(self.gfu.vec.device)
        cands = []
        for name, v in self.guess.candidates():
            # the plain guess already holds this step's Dirichlet lift on the fixed dofs: the candidates
            # take those values (what apply_dirichlet_bcs_to would set) and keep their own on the free ones
            torch.where(self._fixed, self.gfu.vec, v, out=v)
            cands.append((name, v))
        if not wait:
            return choose_best_async(self.a, self.L.vec, self.gfu.vec, cands)
        choice, rp, rb = choose_best(self.a, self.L.vec, self.gfu.vec, cands)
        self.guess_log.append((choice, rp, rb))
        return choice

    def solve(self, nsteps: Optional[int] = None) -> C.GridFunction:
        if nsteps is None:
            nsteps = int(round((self.t1 - self.t0) / self.dt))
        for _ in range(nsteps):
            self.step()
        return self.gfu_prev


SCHEMES = ('implicit euler', 'explicit euler', 'crank nicolson')


class TransientMultiStepSolver:
    """Fixed-step time integration (opencmp/solvers/transient_multistep.py:32-169) with the one-step
    schemes of opencmp/solvers/time_integration_schemes.py, each a combination of device-resident
    operators: the mass matrix M of the time-differentiated components, the time-coefficient form
    A_c and the ODE form A_o (both built for the step's dt), and the linear forms L(t) at the new and
    the previous time:

        implicit euler (:79-132):  (M + A_c + A_o) u = L(t_n+1) + M u_n
        explicit euler (:25-76):   (M + A_c) u = L(t_n) - A_o u_n + M u_n
        crank nicolson (:135-190): (M + A_c + A_o/2) u = (L(t_n+1) + L(t_n))/2 - A_o u_n/2 + M u_n

    The step size is clipped so the last step lands on the end time (_dt_for_next_time_to_hit);
    a step of another size rebuilds the operators. Boundary data are applied at t_n+1. A linear model
    keeps its assembled operators between steps; a nonlinear (Oseen) model re-assembles in its Picard
    loop and supports implicit euler only, as a wind that changes within the step has no explicit part
    here. Differences from the reference: a DG model's boundary facet terms belong to its facet term,
    which the explicit schemes treat as part of A_c (implicitly).
    """

    def __init__(self, model: Model, scheme: str, dt: float, t_range=(0.0, 1.0), t_param=None,
                 gfu0: Optional[C.GridFunction] = None):
        if scheme not in SCHEMES:
            raise NotImplementedError(f'time scheme {scheme!r} (the GPU path has {", ".join(SCHEMES)})')
        if model.nonlinear and scheme != 'implicit euler':
            raise NotImplementedError(f'{scheme} for a nonlinear (Oseen) model')
        self.model = model
        self.scheme = scheme
        self.dt0 = float(dt)
        self.t0, self.t1 = float(t_range[0]), float(t_range[1])
        from .expr import Parameter
        self.t_param = t_param if t_param is not None else model.t_param
        self.t_param[0].Set(self.t0)
        self.t_param[1].Set(self.t0 - self.dt0)
        self.gfu = C.GridFunction(model.fes)
        self.gfu_prev = C.GridFunction(model.fes)
        if gfu0 is not None:
            self.gfu_prev.vec.copy_(gfu0.vec)
        self.gfu.vec.copy_(self.gfu_prev.vec)
        self.tmp = model.ctx.zeros(model.fes.ndof)
        self.steps = 0
        self.linear_iterations = 0
        self.step_results = []
        self.M = C.BilinearForm(model.fes, model.time_derivative_terms())
        self.M.Assemble()
        self.dt = None
        self._build(self.dt0)

    def _build(self, dt: float):
        m = self.model
        fes = m.fes
        mass = m.time_derivative_terms()
        coef = m.construct_bilinear_time_coefficient(dt)
        if self.scheme == 'implicit euler':
            a_terms = mass + coef + m.construct_bilinear_time_ODE(dt)
            self.L_new, self.L_old, self.Aode = C.LinearForm(fes, m.construct_linear(dt, 0)), None, None
        elif self.scheme == 'explicit euler':
            a_terms = mass + coef
            self.L_new, self.L_old = None, C.LinearForm(fes, m.construct_linear(dt, 1))
            self.Aode = C.BilinearForm(fes, m.construct_bilinear_time_ODE(dt))
        else:
            a_terms = mass + coef + m.construct_bilinear_time_ODE(0.5 * dt)
            self.L_new = C.LinearForm(fes, m.construct_linear(0.5 * dt, 0))
            self.L_old = C.LinearForm(fes, m.construct_linear(0.5 * dt, 1))
            self.Aode = C.BilinearForm(fes, m.construct_bilinear_time_ODE(0.5 * dt))
        self.a = C.BilinearForm(fes, a_terms)
        self.pre = m.construct_preconditioners(self.a)
        self.a.Assemble()
        if self.pre is not None:
            self.pre.Update()
        if self.Aode is not None:
            self.Aode.Assemble()
        self.dt = dt

    def _assemble_a(self, full: bool = True):
        self.a.Assemble()
        if self.pre is not None:
            self.pre.Update() if full else self.pre.UpdateJacobi()

    def _axpy(self, a: float, x, y):
        L.check(self.model.ctx.lib.ocmp_axpby(self.model.ctx.h, self.model.fes.ndof, a,
                                              ctypes.c_void_p(x.data_ptr()), 1.0, ctypes.c_void_p(y.data_ptr())),
                'axpby')

    def _rhs(self) -> torch.Tensor:
        """The scheme's right-hand side in a fresh device vector."""
        rhs = self.model.ctx.zeros(self.model.fes.ndof)
        if self.L_new is not None:
            self.L_new.Assemble()
            self._axpy(1.0, self.L_new.vec, rhs)
        if self.L_old is not None:
            self.L_old.Assemble()
            self._axpy(1.0, self.L_old.vec, rhs)
        if self.Aode is not None:
            self.Aode.mult(self.gfu_prev.vec, self.tmp)
            self._axpy(-1.0, self.tmp, rhs)
        self.M.mult(self.gfu_prev.vec, self.tmp)
        self._axpy(1.0, self.tmp, rhs)
        return rhs

    def step(self, dt: float):
        if self.dt is None or abs(dt - self.dt) > 1e-12 * self.dt0:
            self._build(dt)
        tn = self.t_param[0].value
        self.t_param[1].Set(tn)
        self.t_param[0].Set(tn + dt)
        rhs = self._rhs()
        self.gfu.vec.copy_(self.gfu_prev.vec)
        self.model.apply_dirichlet_bcs_to(self.gfu)
        Lwrap = _VecForm(rhs)
        reassemble = self._assemble_a if self.model.nonlinear else None
        res = self.model.solve_single_step(self.a, Lwrap, self.pre, self.gfu, reassemble=reassemble)
        self.step_results.append(res)
        self.linear_iterations += getattr(self.model, 'linear_iterations', None) or res.iters
        self.gfu_prev.vec.copy_(self.gfu.vec)
        self.steps += 1
        return res

    def solve(self, on_step=None) -> C.GridFunction:
        """Steps to the end time; `on_step(self)` after every step (solution saving)."""
        import numpy as np
        # opencmp/solvers/base_solver.py:815: step while t < t_end and t is not close to t_end
        while self.t_param[0].value < self.t1 and not np.isclose(self.t_param[0].value, self.t1):
            self.step(min(self.dt0, self.t1 - self.t_param[0].value))
            if on_step is not None:
                on_step(self)
        return self.gfu


class _VecForm:
    """A linear form already assembled into a device vector (what solve_single_step reads)."""

    def __init__(self, vec: torch.Tensor):
        self.vec = vec

    def Assemble(self):
        return self
