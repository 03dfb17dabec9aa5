"""Checkpoints and restart (pytests/full_system/restart/test_restart.py:76-220 of the reference) on the GPU
path's run(): the resume guard runs before anything is built, so its refusals are checked on the CPU; the
resumed solve itself is checked against the uninterrupted one on the GPU. The case directory is the
reference's restart/transient_poisson, copied unmodified under tests/golden.
"""
import os
import shutil
from pathlib import Path

import numpy as np
import pytest

from opencmp_amd import run as R
from opencmp_amd.io import load_sol, save_sol

ROOT = Path(__file__).parent / 'golden'
TEMPLATE = ROOT / 'pytests' / 'full_system' / 'restart' / 'transient_poisson'
MULTISTEP_SCHEMES = ['CNLF', 'SBDF', 'adaptive IMEX']
RESTARTABLE_SCHEMES = ['implicit euler', 'adaptive three step', 'RK 222', 'RK 232']
T_RESTART, T_FINAL = 0.05, 0.1


def make_case(tmp_path: Path) -> Path:
    case = tmp_path / 'transient_poisson'
    shutil.copytree(TEMPLATE, case)
    cfg = case / 'config'
    cfg.write_text(cfg.read_text().replace('run_dir = .', 'run_dir = ' + str(case)))
    return case


def sol_time(p: Path) -> float:
    return float(p.stem.split('_')[1])


def sol_files(d: Path):
    return sorted(d.glob('poisson_*.sol'), key=sol_time)


def edit(case: Path, *pairs, append=''):
    cfg = case / 'config'
    t = cfg.read_text()
    for a, b in pairs:
        t = t.replace(a, b)
    cfg.write_text(t + append)
    return str(cfg)


def test_sol_format_matches_ngsolve_file():
    """The reference's phi.sol (written by NGSolve's GridFunction.Save, H1 order 3 on its 40x40 quad mesh:
    1,600 vertices, 3,120 edges, 1,521 cells) is a headerless float64 dump of exactly
    1600 + 2 * 3120 + 4 * 1521 = 13,924 coefficients, which is the layout save_sol writes."""
    p = ROOT / 'pytests' / 'full_system' / 'dim' / 'dim_poisson_2' / 'dim_dir' / 'phi.sol'
    a = load_sol(str(p), 1600 + 2 * 3120 + 4 * 1521)
    assert np.isfinite(a).all() and 0.0 <= a[:1600].min() and a[:1600].max() <= 1.0 + 1e-12
    with pytest.raises(ValueError):
        load_sol(str(p), 13923)


def test_sol_round_trip(tmp_path):
    v = np.random.default_rng(0).standard_normal(1234)
    save_sol(str(tmp_path / 'poisson_0.5.sol'), v)
    assert os.path.getsize(tmp_path / 'poisson_0.5.sol') == 8 * 1234
    assert np.array_equal(load_sol(str(tmp_path / 'poisson_0.5.sol'), 1234), v)


@pytest.mark.parametrize('scheme', MULTISTEP_SCHEMES)
def test_resume_rejects_multistep_schemes(tmp_path, scheme):
    case = make_case(tmp_path)
    cfg = edit(case, ('scheme = implicit euler', 'scheme = ' + scheme),
               ('resume_from_previous = False', 'resume_from_previous = True'))
    with pytest.raises(NotImplementedError, match='single-step'):
        R.run(cfg, root=str(ROOT))


@pytest.mark.parametrize('scheme', RESTARTABLE_SCHEMES)
def test_resume_allowed_for_single_step_schemes(tmp_path, scheme):
    """The guard accepts the single-step schemes; with nothing saved yet the missing checkpoint is the error."""
    case = make_case(tmp_path)
    assert R.SCHEME_HISTORY_ORDER[scheme] == 1
    cfg = edit(case, ('scheme = implicit euler', 'scheme = ' + scheme),
               ('resume_from_previous = False', 'resume_from_previous = True'))
    with pytest.raises(FileNotFoundError, match='No valid .sol checkpoint'):
        R.run(cfg, root=str(ROOT))


def test_resume_without_any_checkpoint_fails(tmp_path):
    case = make_case(tmp_path)
    cfg = edit(case, ('resume_from_previous = False', 'resume_from_previous = True'))
    assert not (case / 'output').exists()
    with pytest.raises(FileNotFoundError, match='No valid .sol checkpoint'):
        R.run(cfg, root=str(ROOT))


def test_explicit_restart_checkpoint_must_exist(tmp_path):
    case = make_case(tmp_path)
    cfg = edit(case, ('resume_from_previous = False', 'resume_from_previous = True'),
               append='\nrestart_from = output/poisson_sol/poisson_0.05.sol\n')
    with pytest.raises(FileNotFoundError, match='does not exist'):
        R.run(cfg, root=str(ROOT))


def test_restart_from_needs_resume(tmp_path):
    case = make_case(tmp_path)
    cfg = edit(case, append='\nrestart_from = LATEST\n')
    with pytest.raises(ValueError, match='resume_from_previous is False'):
        R.run(cfg, root=str(ROOT))


def test_restart_checkpoint_must_match_active_model(tmp_path):
    case = make_case(tmp_path)
    d = case / 'output' / 'poisson_sol'
    d.mkdir(parents=True)
    save_sol(str(d / 'ins_0.05.sol'), np.zeros(10))
    cfg = edit(case, ('resume_from_previous = False', 'resume_from_previous = True'),
               append='\nrestart_from = ' + str(d / 'ins_0.05.sol') + '\n')
    with pytest.raises(ValueError, match='must be named'):
        R.run(cfg, root=str(ROOT))


def test_latest_ignores_other_models_and_bad_names(tmp_path):
    d = tmp_path / 'output' / 'poisson_sol'
    d.mkdir(parents=True)
    for name in ('poisson_0.01.sol', 'poisson_0.09999999999999999.sol', 'poisson_x.sol', 'mcpoisson_999.0.sol'):
        save_sol(str(d / name), np.zeros(3))
    c, t = R.select_restart_checkpoint('LATEST', tmp_path, 'poisson')
    assert c.name == 'poisson_0.09999999999999999.sol' and t == pytest.approx(0.1)
    assert R.model_file_name('StokesDIM') == 'stokes' and R.model_file_name('INS') == 'ins'


def test_complete_run_is_not_resumed(tmp_path):
    case = make_case(tmp_path)
    d = case / 'output' / 'poisson_sol'
    d.mkdir(parents=True)
    save_sol(str(d / 'poisson_0.09999999999999999.sol'), np.zeros(3))
    cfg = edit(case, ('resume_from_previous = False', 'resume_from_previous = True'),
               append='\nrestart_from = LATEST\n')
    with pytest.raises(SystemExit) as e:
        R.run(cfg, root=str(ROOT))
    assert e.value.code == 0


@pytest.mark.gpu
@pytest.mark.parametrize('selector', ['relative', 'absolute', 'LATEST', 'legacy'])
def test_resume_matches_uninterrupted_solve(tmp_path, selector, caplog, gpu_ctx):
    """Solve 0 -> 0.1 in one go, drop the files after 0.05, resume: the final solutions agree to the
    reference's bound (relative L2 difference < 1e-10; here on the coefficients, which bound it)."""
    print(f'[gpu-test] restart: resume via {selector}', flush=True)
    case = make_case(tmp_path)
    d = case / 'output' / 'poisson_sol'
    R.run(str(case / 'config'), root=str(ROOT), ctx=gpu_ctx, out=lambda s: None)
    saved = sol_files(d)
    assert sol_time(saved[0]) == 0.0 and sol_time(saved[-1]) == pytest.approx(T_FINAL)
    assert len(saved) == 11                       # the initial condition and every step (save_frequency 1, numit)
    ref = load_sol(str(saved[-1]), os.path.getsize(saved[-1]) // 8)
    for f in saved:
        if sol_time(f) > T_RESTART + 1e-12:
            f.unlink()
    restart_from = sol_files(d)[-1]
    assert sol_time(restart_from) == pytest.approx(T_RESTART)
    if selector == 'LATEST':
        shutil.copy(restart_from, d / 'mcpoisson_999.0.sol')
    append = {'relative': '\nrestart_from = output/poisson_sol/' + restart_from.name + '\n',
              'absolute': '\nrestart_from = ' + str(restart_from.resolve()) + '\n',
              'LATEST': '\nrestart_from = LATEST\n', 'legacy': ''}[selector]
    cfg = edit(case, ('resume_from_previous = False', 'resume_from_previous = True'), append=append)
    R.run(cfg, root=str(ROOT), ctx=gpu_ctx, out=lambda s: None)
    if selector == 'legacy':
        assert 'without restart_from is deprecated' in caplog.text
    final = sol_files(d)[-1]
    assert sol_time(final) == pytest.approx(T_FINAL)
    got = load_sol(str(final), ref.size)
    assert np.linalg.norm(ref) > 0.0
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-10
