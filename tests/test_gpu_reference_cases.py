"""The reference's own full-system cases, run from their unmodified config directories through
opencmp_amd.run.run on the GPU, checked the way the reference checks them:
numpy.isclose(error, expected, rtol=3, atol=1e-12) (opencmp/helpers/testing.py:58-86).

Expected values are the reference's (file:line of the test that states them). These cases do not
depend on NGSolve's basis or numbering: each compares a discrete solution with a closed-form one.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.join(os.path.dirname(__file__), 'golden')


def cfg(case):
    from opencmp_amd.config import ConfigParser
    return ConfigParser(os.path.join(ROOT, 'pytests', 'full_system', case, 'config'))


def check(case, c, expected):
    from opencmp_amd.run import run
    print(f'[gpu-test] reference case {case}', flush=True)
    r = run('', c, root=ROOT, out=lambda s: print('   ', s, flush=True))
    print(f'[gpu-test]   {r.steps} steps to t = {r.t:.6g}, {r.linear_iterations} linear iterations', flush=True)
    assert len(r.errors) == len(expected)
    for got, exp in zip(r.errors, expected):
        assert np.isclose(got, exp, rtol=3, atol=1e-12), (got, exp)
    return r


@pytest.mark.parametrize('scheme', ['implicit euler', 'crank nicolson'])
def test_mcins_purely_convective_dg(scheme):
    """DG-P3 upwind transport, fixed velocity, to t = 20:
    pytests/full_system/mcins/test_mcins.py:678-690 expects [1e-8] for both schemes."""
    c = cfg('mcins/purely_convective')
    c['TRANSIENT']['scheme'] = scheme
    r = check(f'mcins/purely_convective {scheme}', c, [1e-8])
    assert r.steps == 400


def test_stokes_stationary_pipe_cg():
    """Taylor-Hood P3/P2 Poiseuille flow with stress tractions: the seven metrics of
    pytests/full_system/stokes/test_stokes.py:34-37."""
    check('stokes/stationary_pipe', cfg('stokes/stationary_pipe'),
          [1e-10, 6e-12, 3e-11, 2e-12, 6e-10, 2e-11, 3e-10])


def test_ins_sinusoidal_transient_cg():
    """Oseen-Picard INS, Taylor-Hood P3/P2, implicit euler to t = 0.1:
    pytests/full_system/ins/test_ins.py:74-77 expects [1e-4, 2e-3]."""
    r = check('ins/sinusoidal_transient', cfg('ins/sinusoidal_transient'), [1e-4, 2e-3])
    assert r.steps == 100


@pytest.mark.parametrize('scheme', ['explicit euler', 'implicit euler', 'crank nicolson'])
@pytest.mark.parametrize('dg', [False, True])
def test_poisson_transient_schemes(scheme, dg):
    """H1/L2 P3 heat equation on the 6-triangle mesh to t = 1 with each one-step scheme:
    pytests/full_system/poisson/test_poisson.py:57-90 expects [9e-6]."""
    c = cfg('poisson/transient_coarse')
    c['TRANSIENT']['scheme'] = scheme
    if dg:
        c['DG']['DG'] = 'True'
        c['FINITE ELEMENT SPACE']['elements'] = 'u -> L2'
    r = check(f'poisson/transient_coarse {scheme} {"DG" if dg else "CG"}', c, [9e-6])
    assert r.steps == 1000


def test_mcins_diffusion_convection_cg():
    """Two species, CG P3, fixed velocity, Neumann and ramped Dirichlet data, implicit euler to t = 10:
    pytests/full_system/mcins/test_mcins.py:223-226 expects [1e-6, 4e-8]."""
    check('mcins/diffusion_convection', cfg('mcins/diffusion_convection'), [1e-6, 4e-8])
