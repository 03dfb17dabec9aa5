"""OpenCMP config files for the GPU path.

Reads the reference's INI layout (a main ``config`` plus the ``bc_dir``,
``ic_dir``, ``model_dir`` and ``ref_sol_dir`` sub-configs of the run directory)
with the defaults of opencmp/config_functions/expanded_config_parser.py:26-90.
Values keep the reference syntax: comma-separated lists, ``key -> value`` lines
for dictionaries, arithmetic strings (parsed by :mod:`opencmp_amd.expr`).

Three keys select and tune the GPU path; all default so that an unmodified
reference config runs on it:

    [SOLVER]
    backend    = mi355x         # the only backend of this package
    precision  = fp64           # fp64 | mixed (FP32 inner solve in FP64 refinement)
    schur_step = pcg            # Stokes/INS: pcg (inner Jacobi-PCG on the Schur matrix) | amg (one
                                # AMG V-cycle: faster where the velocity block is mass-dominated,
                                # transient INS; steady Stokes P3/P2 stalls with it)
"""
from __future__ import annotations

import configparser
import os
from typing import Any, Dict, List, Optional

CONFIG_DEFAULTS: Dict[str, Dict[str, Any]] = {
    # opencmp/config_functions/expanded_config_parser.py:26-90 (the sections this path reads)
    'MESH': {'filename': 'REQUIRED', 'curved_elements': 'False'},
    'FINITE ELEMENT SPACE': {'elements': 'REQUIRED', 'interpolant_order': 'REQUIRED',
                             'no_constrained_dofs': 'False'},
    'DG': {'DG': 'False', 'interior_penalty_coefficient': '10.0'},
    'SOLVER': {'linear_solver': 'default', 'preconditioner': 'default', 'linear_tolerance': '1e-15',
               'linear_max_iterations': '100', 'linearization_method': 'Oseen', 'nonlinear_solver': 'default',
               'nonlinear_tolerance': 'absolute -> 0.0', 'nonlinear_max_iterations': '10',
               'backend': 'mi355x', 'precision': 'fp64', 'schur_step': 'pcg'},
    'TRANSIENT': {'transient': 'False', 'scheme': 'implicit euler', 'time_range': '0.0, 5.0', 'dt': '0.001',
                  'dt_tolerance': 'absolute -> 0.0', 'dt_range': '1e-6, 0.1', 'maximum_rejected_solves': '1000'},
    'ERROR ANALYSIS': {'check_error': 'False', 'check_error_every_timestep': 'False',
                       'save_error_every_timestep': 'False', 'error_average': '', 'num_refinements': '4'},
    'OTHER': {'num_threads': '1', 'messaging_level': '0', 'model': 'REQUIRED', 'component_names': '',
              'parameter_names': '', 'velocity_fixed': 'False', 'run_dir': 'REQUIRED',
              'resume_from_previous': 'False', 'restart_from': ''},
    'VISUALIZATION': {'save_to_file': 'False', 'save_type': '.sol', 'save_frequency': '1, numit',
                      'subdivision': '-1', 'split_components': 'False'},
    'DIM': {'diffuse_interface_method': 'False'},
}


def _read(path: str) -> configparser.ConfigParser:
    if not os.path.isfile(path):
        raise FileNotFoundError(f'The given config file "{path}" does not exist.')
    cp = configparser.ConfigParser()     # the reference's parser: lower-cased keys, '#'/';' comments
    cp.read(path)
    return cp


def parse_dict(text: str) -> Dict[str, str]:
    """``a -> v`` lines (the first may follow the '=' sign) into an ordered dict of strings."""
    out: Dict[str, str] = {}
    for line in str(text).split('\n'):
        line = line.strip()
        if not line:
            continue
        if '->' not in line:
            raise ValueError(f'expected "key -> value", got {line!r}')
        k, v = line.split('->', 1)
        out[k.strip()] = v.strip()
    return out


def parse_list(text: str) -> List[str]:
    t = str(text).strip()
    if not t:
        return []
    return [v.strip() for v in t.replace('\n', ',').split(',') if v.strip()]


def to_bool(v: str) -> bool:
    s = str(v).strip()
    if s in ('True', 'true', '1'):
        return True
    if s in ('False', 'false', '0'):
        return False
    raise ValueError(f'expected True or False, got {v!r}')


class ConfigParser:
    """The main config file (mirrors ConfigParser.get_item/get_list/get_dict of
    opencmp/config_functions/expanded_config_parser.py:390-499). Sections and keys can be set
    like a dict before ``run`` (the reference's tests do ``config['DG']['DG'] = 'True'``)."""

    def __init__(self, config_file_path: str):
        self.path = os.path.abspath(config_file_path)
        self.cp = _read(config_file_path)

    def __getitem__(self, section: str):
        if not self.cp.has_section(section):
            self.cp.add_section(section)
        return self.cp[section]

    def raw(self, section: str, key: str, quiet: bool = True) -> str:
        if self.cp.has_option(section, key):
            return self.cp.get(section, key)
        default = CONFIG_DEFAULTS.get(section, {}).get(key)
        if default is None or default == 'REQUIRED':
            raise ValueError(f'No default available for {section}, {key}. Please specify a value in the config '
                             'file.')
        return default

    def get_item(self, keys: List[str], typ=str):
        v = self.raw(*keys).strip()
        if typ is bool:
            return to_bool(v)
        return typ(v)

    def get_list(self, keys: List[str], typ=str) -> list:
        vals = parse_list(self.raw(*keys))
        return [to_bool(v) if typ is bool else typ(v) for v in vals]

    def get_dict(self, keys: List[str]) -> Dict[str, str]:
        d = parse_dict(self.raw(*keys))
        default = CONFIG_DEFAULTS.get(keys[0], {}).get(keys[1])
        if default and '->' in default:
            for k, v in parse_dict(default).items():
                d.setdefault(k, v)
        return d


class SubConfig:
    """One of the run directory's sub-configs (bc_config, ic_config, model_config, ref_sol_config):
    section -> variable -> raw value (a ``marker -> value`` dict or a plain string)."""

    def __init__(self, path: str):
        self.path = path
        self.cp = _read(path) if os.path.isfile(path) else None

    def sections(self) -> List[str]:
        return self.cp.sections() if self.cp is not None else []

    def section(self, name: str) -> Dict[str, str]:
        if self.cp is None:
            return {}
        for s in self.cp.sections():
            if s.lower() == name.lower():
                return {k: v for k, v in self.cp[s].items()}
        return {}

    def dict_section(self, name: str) -> Dict[str, Dict[str, str]]:
        """variable -> {marker: value} for the ``u = marker -> value`` layout."""
        return {var: parse_dict(v) for var, v in self.section(name).items() if v.strip()}


def run_paths(config: ConfigParser, root: Optional[str] = None) -> Dict[str, str]:
    """The mesh file and the run directory's sub-configs. Relative paths resolve against ``root``
    (the reference resolves them against the working directory)."""
    root = os.getcwd() if root is None else root
    rd = config.get_item(['OTHER', 'run_dir'], str)
    run_dir = rd if os.path.isabs(rd) else os.path.join(root, rd)
    mesh = config.get_item(['MESH', 'filename'], str)
    mesh = mesh if os.path.isabs(mesh) else os.path.join(root, mesh)
    return {'run_dir': run_dir, 'mesh': mesh,
            'bc': os.path.join(run_dir, 'bc_dir', 'bc_config'),
            'ic': os.path.join(run_dir, 'ic_dir', 'ic_config'),
            'model': os.path.join(run_dir, 'model_dir', 'model_config'),
            'ref_sol': os.path.join(run_dir, 'ref_sol_dir', 'ref_sol_config')}
