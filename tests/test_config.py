"""Reading the reference's config files (no GPU): the fixtures under tests/golden/pytests are the
reference's own case directories, unmodified."""
import os

import pytest

from opencmp_amd import config as K
from opencmp_amd import expr as E

ROOT = os.path.join(os.path.dirname(__file__), 'golden')
CASES = ['mcins/purely_convective', 'mcins/diffusion_convection', 'stokes/stationary_pipe',
         'ins/sinusoidal_transient', 'ins/pressure_flow_in_pipe_velocity', 'poisson/transient_coarse']


def cfg_path(case):
    return os.path.join(ROOT, 'pytests', 'full_system', case, 'config')


def test_main_config_items_lists_dicts_and_defaults():
    c = K.ConfigParser(cfg_path('stokes/stationary_pipe'))
    assert c.get_item(['OTHER', 'model'], str) == 'Stokes'
    assert c.get_item(['FINITE ELEMENT SPACE', 'interpolant_order'], int) == 3
    assert K.parse_dict(c.raw('FINITE ELEMENT SPACE', 'elements')) == {'u': 'VectorH1', 'p': 'H1'}
    assert c.get_item(['DG', 'DG'], bool) is False
    assert c.get_list(['ERROR ANALYSIS', 'error_average'], str) == ['p']
    # defaults of opencmp/config_functions/expanded_config_parser.py:26-90 and the GPU-path keys
    assert c.get_item(['SOLVER', 'linear_tolerance'], float) == 1e-15
    assert c.get_item(['SOLVER', 'linear_max_iterations'], int) == 100
    assert c.get_item(['SOLVER', 'backend'], str) == 'mi355x'
    assert c.get_item(['SOLVER', 'precision'], str) == 'fp64'
    assert c.get_item(['SOLVER', 'schur_step'], str) == 'pcg'
    t = K.ConfigParser(cfg_path('mcins/purely_convective'))
    assert t.get_list(['TRANSIENT', 'time_range'], float) == [0.0, 20.0]
    assert t.get_dict(['SOLVER', 'nonlinear_tolerance']) == {'relative': '1e-4', 'absolute': '1e-8'}
    assert t.get_item(['OTHER', 'velocity_fixed'], bool) is True
    # keys can be overridden before a run, as the reference's tests do
    t['DG']['DG'] = 'False'
    assert t.get_item(['DG', 'DG'], bool) is False
    with pytest.raises(ValueError):
        t.get_item(['MESH', 'nonexistent_required'], str)


def test_sub_configs():
    c = K.ConfigParser(cfg_path('stokes/stationary_pipe'))
    p = K.run_paths(c, ROOT)
    assert os.path.isfile(p['mesh']) and p['mesh'].endswith('channel_3bcs.vol')
    bc = K.SubConfig(p['bc'])
    assert bc.dict_section('DIRICHLET') == {'u': {'wall': '[0.0, 0.0]'}}
    assert bc.dict_section('STRESS') == {'u': {'inlet': '[10.0, 0.0]', 'outlet': '[-5.0, 0.0]'}}
    ref = K.SubConfig(p['ref_sol'])
    # configparser lower-cases keys; the order of the metrics is the order of the printed errors
    assert list(ref.section('METRICS').keys()) == ['l2_norm', 'l1_norm', 'linfinity_norm', 'divergence']
    assert ref.section('REFERENCE SOLUTIONS')['p'] == '5.0*(1.0 - x) + 5.0'
    ic = K.SubConfig(K.run_paths(K.ConfigParser(cfg_path('mcins/purely_convective')), ROOT)['ic'])
    assert K.parse_dict(ic.section('MultiComponentINS')['u']) == {'all': '[0.1, 0.0]'}


@pytest.mark.parametrize('case', CASES)
def test_every_fixture_expression_parses(case):
    c = K.ConfigParser(cfg_path(case))
    p = K.run_paths(c, ROOT)
    assert os.path.isfile(p['mesh'])
    n = 0
    for key in ('bc', 'ic', 'model'):
        sub = K.SubConfig(p[key])
        for sec in sub.sections():
            for var, txt in sub.section(sec).items():
                for marker, val in K.parse_dict(txt).items():
                    if val.strip() != 'None':
                        assert E.vector(val)
                        n += 1
    ref = K.SubConfig(p['ref_sol'])
    for var, val in ref.section('REFERENCE SOLUTIONS').items():
        assert E.vector(val)
        n += 1
    assert n >= 3
