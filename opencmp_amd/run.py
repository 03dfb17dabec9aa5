"""Run an OpenCMP config file on the GPU path.

``run(config_file_path)`` mirrors opencmp/run.py:30-69: read the config
(:mod:`opencmp_amd.config`), build the model named in ``[OTHER] model`` from the
run directory's boundary-condition, initial-condition, model and
reference-solution sub-configs, solve (stationary, or time-stepped with the
``[TRANSIENT] scheme``), and print the error metrics of ``[METRICS]`` in the
reference's format (opencmp/helpers/error.py:238-301), which its test harness
parses (opencmp/helpers/testing.py:58-86).

Supported: the models Poisson (CG, DG), Stokes (CG), INS (CG, Oseen) and
MultiComponentINS with ``velocity_fixed = True`` and a constant velocity (CG,
DG); the schemes implicit euler, explicit euler and crank nicolson; the metrics
l1_norm, l2_norm, linfinity_norm and divergence. Anything else raises
NotImplementedError naming what is missing.
"""
from __future__ import annotations

import logging
import math
import os
import sys
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import core as C
from . import expr as E
from . import models as M
from .config import ConfigParser, SubConfig, parse_dict, parse_list, run_paths


@dataclass
class RunResult:
    gfu: C.GridFunction
    model: M.Model
    errors: List[float] = field(default_factory=list)
    lines: List[str] = field(default_factory=list)
    steps: int = 0
    t: float = 0.0
    linear_iterations: int = 0


# ------------------------------------------------------------------ config values

def _exprs(text: str) -> List[E.Expr]:
    return E.vector(text)


def _const(text: str, what: str) -> float:
    e = E.Expr(text)
    if not e.is_constant():
        raise NotImplementedError(f'{what} = {text!r}: only constant model parameters are supported')
    return e.value()


def _fn_of(exprs: List[E.Expr]):
    """One scalar Fn per component: a constant closed form where possible, else the expression."""
    out = []
    for e in exprs:
        out.append(M.const(e.value()) if e.is_constant() else e)
    return out if len(out) > 1 else out[0]


def _bc_dict(bc: SubConfig, section: str) -> Dict[str, Dict[str, str]]:
    return bc.dict_section(section)


def _all_markers(mesh) -> List[str]:
    return list(mesh.bnd_names)


# ------------------------------------------------------------------ checkpoints (.sol files)

# previous solutions each scheme needs to restart (opencmp/solvers/base_solver.py:56-69)
SCHEME_HISTORY_ORDER = {'explicit euler': 1, 'implicit euler': 1, 'crank nicolson': 1, 'adaptive two step': 1,
                        'adaptive three step': 1, 'euler IMEX': 1, 'CNLF': 2, 'SBDF': 3, 'adaptive IMEX': 3,
                        'RK 222': 1, 'RK 232': 1}


def model_file_name(model_name: str) -> str:
    """Model.name (opencmp/models/base_model.py:648-662): the class name in lower case, without a
    trailing 'dim'."""
    n = model_name.lower()
    return n[:-3] if n.endswith('dim') else n


def checkpoint_time(checkpoint: Path, model_name: str) -> float:
    """opencmp/solvers/base_solver.py:91-113: the time in a '<model>_<time>.sol' file name."""
    prefix = model_name + '_'
    if not checkpoint.name.startswith(prefix) or checkpoint.suffix != '.sol':
        raise ValueError(f'Restart checkpoint "{checkpoint}" must be named "{prefix}<time>.sol" for the active model.')
    text = checkpoint.name[len(prefix):-len(checkpoint.suffix)]
    try:
        t = float(text)
    except ValueError:
        raise ValueError(f'Could not determine the restart time from checkpoint "{checkpoint}". Expected the '
                         f'filename format "{prefix}<time>.sol".')
    if not math.isfinite(t):
        raise ValueError(f'Restart checkpoint "{checkpoint}" does not contain a finite time.')
    return t


def select_restart_checkpoint(restart_from: str, run_dir: Path, model_name: str) -> Tuple[Path, float]:
    """opencmp/solvers/base_solver.py:115-146: an explicit path (relative to run_dir) or LATEST, the
    latest valid checkpoint of the model under run_dir/output/<model>_sol."""
    if restart_from.upper() == 'LATEST':
        out_dir = run_dir / 'output' / (model_name + '_sol')
        found = []
        if out_dir.exists():
            for c in out_dir.rglob(model_name + '_*.sol'):
                try:
                    found.append((checkpoint_time(c, model_name), c.resolve()))
                except ValueError:
                    continue
        if not found:
            raise FileNotFoundError(f'No valid .sol checkpoint for model "{model_name}" was found in "{out_dir}". '
                                    'Set resume_from_previous to False to start a new simulation.')
        t, c = max(found, key=lambda item: item[0])
        return c, t
    c = Path(restart_from).expanduser()
    if not c.is_absolute():
        c = run_dir / c
    c = c.resolve()
    if not c.is_file():
        raise FileNotFoundError(f'The restart checkpoint "{c}" does not exist.')
    return c, checkpoint_time(c, model_name)


def resume_point(cfg: ConfigParser, run_dir: Path) -> Optional[Tuple[Path, float]]:
    """The resume guard of opencmp/solvers/base_solver.py:158-206, run before anything is built:
    None for a fresh start, else the checkpoint and its time. Multi-step schemes are refused (their
    history cannot be rebuilt from saved files); a run that is already complete exits."""
    resume = cfg.get_item(['OTHER', 'resume_from_previous'], bool)
    restart_from = cfg.get_item(['OTHER', 'restart_from'], str).strip()
    if restart_from and not resume:
        raise ValueError(f'restart_from is set to "{restart_from}", but resume_from_previous is False. Enable '
                         'resume_from_previous or remove restart_from.')
    if not resume:
        return None
    scheme = cfg.get_item(['TRANSIENT', 'scheme'], str)
    order = SCHEME_HISTORY_ORDER.get(scheme)
    if order is None:
        raise NotImplementedError(f'time scheme {scheme!r}')
    if order > 1:
        raise NotImplementedError(f'resume_from_previous only supports single-step time discretization schemes. '
                                  f'The "{scheme}" scheme needs {order} previous time steps, which cannot be '
                                  'recovered from the saved .sol files since save_frequency may skip time steps.')
    if not restart_from:
        restart_from = 'LATEST'
        logging.warning('resume_from_previous without restart_from is deprecated and currently means '
                        'restart_from = LATEST. Set restart_from explicitly in the [OTHER] section.')
    name = model_file_name(cfg.get_item(['OTHER', 'model'], str))
    c, t = select_restart_checkpoint(restart_from, run_dir, name)
    t_end = cfg.get_list(['TRANSIENT', 'time_range'], float)[-1]
    if t >= t_end or math.isclose(t, t_end, rel_tol=1e-9):
        print(f'Simulation is already complete at t = {t}, the end of time_range. Nothing to resume.')
        sys.exit(0)
    print(f'Resuming from checkpoint: {c}')
    return c, t


class SolSaver:
    """SolutionFileSaver for .sol files (opencmp/helpers/saving.py:28-93) with the save_frequency
    rule of opencmp/solvers/base_solver.py:560-583: run_dir/output/<model>_sol/<model>_<t>.sol."""

    def __init__(self, cfg: ConfigParser, run_dir: Path, name: str, t_start: float):
        self.on = cfg.get_item(['VISUALIZATION', 'save_to_file'], bool)
        if not self.on:
            logging.warning('save_to_file is set to False — no solution files will be written for this run.')
            return
        kind = cfg.get_item(['VISUALIZATION', 'save_type'], str)
        if kind != '.sol':
            raise NotImplementedError(f'save_type {kind!r} (the GPU path writes .sol files)')
        freq = cfg.get_list(['VISUALIZATION', 'save_frequency'], str)
        self.freq = (float(freq[0]), freq[1])
        if self.freq[1] not in ('time', 'numit'):
            raise ValueError('Save frequency must be specified either per amount of time ("time") or per number '
                             'of iterations ("numit").')
        self.dir = run_dir / 'output' / (name + '_sol')
        self.dir.mkdir(parents=True, exist_ok=True)
        self.base = name + '_'
        self.t_start = t_start

    def save(self, vec, t: float) -> None:
        from .io import save_sol
        save_sol(str(self.dir / f'{self.base}{t}.sol'), vec)

    def after_step(self, vec, t: float, num_iters: int) -> None:
        if not self.on:
            return
        if self.freq[1] == 'numit':
            if num_iters % self.freq[0] == 0:
                self.save(vec, t)
        else:
            # a save time is a multiple of the period after the start (up to round-off)
            k = (t - self.t_start) / self.freq[0]
            if math.isclose(k, round(k), abs_tol=1e-8):
                self.save(vec, t)


# ------------------------------------------------------------------ model construction

def _solver_keys(cfg: ConfigParser) -> dict:
    """[SOLVER] keys of the reference (opencmp/models/base_model.py:152-180) plus the GPU-path ones."""
    backend = cfg.get_item(['SOLVER', 'backend'], str)
    if backend != 'mi355x':
        raise ValueError(f'[SOLVER] backend = {backend!r}: this package runs the mi355x backend only')
    precision = cfg.get_item(['SOLVER', 'precision'], str)
    if precision not in ('fp64', 'mixed'):
        raise ValueError(f'[SOLVER] precision must be fp64 or mixed, not {precision!r}')
    return dict(linear_solver=cfg.get_item(['SOLVER', 'linear_solver'], str),
                preconditioner=cfg.get_item(['SOLVER', 'preconditioner'], str),
                linear_tolerance=cfg.get_item(['SOLVER', 'linear_tolerance'], float),
                linear_max_iterations=cfg.get_item(['SOLVER', 'linear_max_iterations'], int),
                precision=precision)


def _elements(cfg: ConfigParser) -> Dict[str, str]:
    return parse_dict(cfg.raw('FINITE ELEMENT SPACE', 'elements'))


def build_model(cfg: ConfigParser, mesh, bc: SubConfig, mcfg: SubConfig, t_param, ctx=None,
                ic: Optional[SubConfig] = None) -> M.Model:
    name = cfg.get_item(['OTHER', 'model'], str)
    if cfg.get_item(['DIM', 'diffuse_interface_method'], bool):
        raise NotImplementedError('diffuse-interface runs from config files are not supported')
    order = cfg.get_item(['FINITE ELEMENT SPACE', 'interpolant_order'], int)
    DG = cfg.get_item(['DG', 'DG'], bool)
    ipc = cfg.get_item(['DG', 'interior_penalty_coefficient'], float)
    elements = _elements(cfg)
    params = {k: parse_dict(v) for k, v in mcfg.section('PARAMETERS').items()}
    funcs = {k: parse_dict(v) for k, v in mcfg.section('FUNCTIONS').items()}
    dirichlet = _bc_dict(bc, 'DIRICHLET')
    common = dict(interp_ord=order, DG=DG, ipc=ipc, ctx=ctx, t_param=t_param, **_solver_keys(cfg))

    def vec_bcs(section, drop_zero=False):
        out = {}
        for var, d in _bc_dict(bc, section).items():
            dd = {}
            for m, v in d.items():
                ex = _exprs(v)
                if drop_zero and all(e.is_constant() and e.value() == 0.0 for e in ex):
                    continue
                dd[m] = _fn_of(ex)
            out[var] = dd
        return out

    if name == 'Poisson':
        if elements.get('u') not in ('H1', 'L2'):
            raise NotImplementedError(f"Poisson with u -> {elements.get('u')}")
        robin = _bc_dict(bc, 'ROBIN')
        if any(robin.values()):
            raise NotImplementedError('Robin data from config files')
        model = M.Poisson(mesh, dc=_const(params['diffusion_coefficient']['all'], 'diffusion_coefficient'),
                          f=_fn_of(_exprs(funcs['source']['all'])),
                          dirichlet=vec_bcs('DIRICHLET'), neumann=vec_bcs('NEUMANN', drop_zero=True), **common)
        return model
    if name in ('Stokes', 'INS'):
        if DG or elements.get('u') != 'VectorH1' or elements.get('p') != 'H1':
            raise NotImplementedError(f'{name} on the GPU path is Taylor-Hood (VectorH1/H1, CG)')
        kv = _const(params['kinematic_viscosity']['all'], 'kinematic_viscosity')
        src = funcs.get('source', {}).get('u')
        f = None if src is None else [(_fn_of([e])) for e in _exprs(src)]
        stress = vec_bcs('STRESS')
        for var, d in stress.items():
            for m, h in d.items():
                if not isinstance(h, list):
                    d[m] = [h] * mesh.dim
        dirichlet_u = vec_bcs('DIRICHLET')
        for var, d in dirichlet_u.items():
            for m, g in d.items():
                if not isinstance(g, list):
                    d[m] = [g] * mesh.dim
        common.setdefault('preconditioner', 'default')
        if common['preconditioner'] in ('default', 'direct', 'local'):
            common['preconditioner'] = 'schur'
            step = cfg.get_item(['SOLVER', 'schur_step'], str)
            if step == 'amg':
                common['schur_amg'] = 1
            elif step == 'pcg':
                common['schur_iters'] = 200
                common['schur_tol'] = 1e-3
            else:
                raise ValueError(f'[SOLVER] schur_step must be amg or pcg, not {step!r}')
        if name == 'Stokes':
            return M.Stokes(mesh, kv=kv, f=f, dirichlet=dirichlet_u, stress=stress, **common)
        tol = {k: float(v) for k, v in cfg.get_dict(['SOLVER', 'nonlinear_tolerance']).items()}
        if cfg.get_item(['SOLVER', 'linearization_method'], str) != 'Oseen':
            raise NotImplementedError('INS with IMEX linearisation')
        return M.INS(mesh, kv=kv, f=f, dirichlet=dirichlet_u, stress=stress,
                     nonlinear_tolerance=(tol.get('absolute', 0.0), tol['relative']),
                     nonlinear_max_iters=cfg.get_item(['SOLVER', 'nonlinear_max_iterations'], int), **common)
    if name == 'MultiComponentINS':
        if not cfg.get_item(['OTHER', 'velocity_fixed'], bool):
            raise NotImplementedError('MultiComponentINS with a solved velocity')
        comps = cfg.get_list(['OTHER', 'component_names'], str)
        Ds = params.get('diffusion_coefficients', {})
        srcs = funcs.get('source', {})
        kinds = {c: elements.get(c, 'L2' if DG else 'H1') for c in comps}
        # the fixed velocity is the initial condition's (opencmp/models/multi_component_ins.py:129-133)
        sec = ic.section(name) if ic is not None else {}
        u_ic = parse_dict(sec.get('u', 'all -> [0.0, 0.0]')).get('all', '[0.0, 0.0]')
        wind = []
        for e in _exprs(u_ic):
            if not e.is_constant():
                raise NotImplementedError(f'a fixed velocity that varies in space or time ({u_ic!r})')
            wind.append(e.value())
        return M.MultiComponentINS(
            mesh, components=comps, kinds=kinds, wind=wind,
            D={c: _const(Ds.get(c, '0'), f'diffusion_coefficients {c}') for c in comps},
            f={c: _fn_of(_exprs(srcs.get(c, '0'))) for c in comps},
            dirichlet=vec_bcs('DIRICHLET'), neumann=vec_bcs('NEUMANN'), total_flux=vec_bcs('TOTAL_FLUX', drop_zero=True),
            **common)
    raise NotImplementedError(f'model {name!r} is not on the GPU path')


def _initial_condition(model: M.Model, ic: SubConfig, name: str, cfg: ConfigParser, gfu: C.GridFunction):
    """gfu.Set(IC) per component (opencmp/models/base_model.py:216-221); 'None' leaves zero."""
    sec = ic.section(name)
    for comp, txt in sec.items():
        d = parse_dict(txt)
        if comp == 'all' or comp not in model.components:
            continue
        val = d.get('all')
        if val is None or val.strip() == 'None':
            continue
        ci = model.components.index(comp)
        exprs = _exprs(val)
        fns = [e.bind(model.ctx, model.t_param[0]) for e in exprs]
        gfu.Set(fns if len(fns) > 1 else fns[0], component=ci)


# ------------------------------------------------------------------ error metrics

def calc_error(cfg: ConfigParser, model: M.Model, gfu: C.GridFunction, ref: SubConfig, t: E.Parameter,
               out=print) -> List[float]:
    """The [METRICS] of the reference-solution config, printed as opencmp/helpers/error.py:238-301
    prints them."""
    average = set(parse_list(cfg.raw('ERROR ANALYSIS', 'error_average')))
    refs = {k: v for k, v in ref.section('REFERENCE SOLUTIONS').items()}
    errors = []
    for metric, vars_txt in ref.section('METRICS').items():
        m = metric.lower()
        for var in parse_list(vars_txt):
            if m in ('l1_norm', 'l2_norm', 'linfinity_norm'):
                err = model.norm(m, gfu, var, [e.bind(model.ctx, t) for e in _exprs(refs[var])],
                                 average=var in average)
                out('{0} in {1}: {2}'.format(metric.replace('_', ' '), var, err))
            elif m == 'divergence':
                err = model.divergence(gfu, var)
                out('divergence of {0}: {1}'.format(var, err))
            else:
                raise NotImplementedError(f'metric {metric!r}')
            errors.append(err)
    return errors


# ------------------------------------------------------------------ driver

def run(config_file_path: str, config_parser: Optional[ConfigParser] = None, root: Optional[str] = None,
        out=print, ctx=None) -> RunResult:
    """opencmp/run.py:30-69 on the GPU path. ``root``: the directory the config's relative paths
    (mesh file, run_dir) are taken from; the reference takes them from the working directory."""
    from .io import load_mesh
    from . import solvers as S
    from .io import load_sol
    cfg = config_parser if config_parser is not None else ConfigParser(config_file_path)
    paths = run_paths(cfg, root)
    run_dir = Path(paths['run_dir']).resolve()
    restart = resume_point(cfg, run_dir)          # before anything is built, like the reference's solver
    mesh = load_mesh(paths['mesh'])
    bc, ic, mcfg, ref = (SubConfig(paths[k]) for k in ('bc', 'ic', 'model', 'ref_sol'))
    transient = cfg.get_item(['TRANSIENT', 'transient'], bool)
    t0, t1 = [float(v) for v in cfg.get_list(['TRANSIENT', 'time_range'], float)]
    if restart is not None:
        t0 = restart[1]
    dt = cfg.get_item(['TRANSIENT', 'dt'], float)
    t_param = [E.Parameter(t0 - k * dt) for k in range(3)]
    model = build_model(cfg, mesh, bc, mcfg, t_param, ctx=ctx, ic=ic)
    name = cfg.get_item(['OTHER', 'model'], str)
    gfu0 = C.GridFunction(model.fes)
    if restart is not None:
        # the checkpoint replaces the configured initial condition (opencmp/models/base_model.py:216-221)
        gfu0.vec.copy_(torch.from_numpy(load_sol(str(restart[0]), model.fes.ndof)).to(gfu0.vec.device))
    else:
        _initial_condition(model, ic, name, cfg, gfu0)
    saver = SolSaver(cfg, run_dir, model_file_name(name), t0)
    if saver.on:
        saver.save(gfu0.vec, t0)
    if hasattr(model, 'set_linearisation_point'):
        model.set_linearisation_point(gfu0)
    lines: List[str] = []

    def emit(s):
        lines.append(s)
        out(s)

    if not transient:
        solver = S.StationarySolver(model, gfu0=gfu0)
        gfu = solver.solve()
        steps, its = 1, getattr(solver.result, 'iters', 0)
    else:
        scheme = cfg.get_item(['TRANSIENT', 'scheme'], str)
        solver = S.TransientMultiStepSolver(model, scheme, dt, (t0, t1), t_param, gfu0)
        gfu = solver.solve(on_step=lambda s_: saver.after_step(s_.gfu.vec, s_.t_param[0].value, s_.steps))
        steps, its = solver.steps, solver.linear_iterations
    errors = []
    if cfg.get_item(['ERROR ANALYSIS', 'check_error'], bool):
        errors = calc_error(cfg, model, gfu, ref, t_param[0], out=emit)
    return RunResult(gfu, model, errors, lines, steps, t_param[0].value, its)
