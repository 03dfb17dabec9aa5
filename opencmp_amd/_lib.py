"""ctypes bindings of libocmp_mi355x.so (declarations in include/ocmp.h).

The library is built in-tree by ``opencmp_amd.build``. Loading fails loudly if
it is missing: there is no CPU fallback for anything on the assemble+solve path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libocmp_mi355x.so')

OCMP_OK = 0
OCMP_ERR_ARG, OCMP_ERR_HIP, OCMP_ERR_STATE, OCMP_ERR_UNSUPPORTED, OCMP_ERR_INTERNAL, OCMP_ERR_NUMERIC = -1, -2, -3, -4, -5, -6
OCMP_H1, OCMP_L2 = 0, 1
BFI_GRADGRAD, BFI_MASS, BFI_CONV, BFI_DTRIAL, BFI_DTEST, BFI_BND_UPWIND, BFI_BND_MASS, BFI_BND_FLUX = 0, 1, 2, 3, 4, 5, 6, 7
BFI_DG_FACET = 16
WEIGHT_NONE, WEIGHT_FIELD, WEIGHT_ONE_MINUS_FIELD = 0, 1, 2
LFI_SOURCE, LFI_DG_DIRICHLET, LFI_BND_SOURCE = 0, 1, 2
FN_CONST, FN_SINPROD, FN_POLY2, FN_FIELD, FN_EXPR = 0, 1, 2, 3, 4
PRE_NONE, PRE_JACOBI, PRE_BLOCK_JACOBI = 0, 1, 2
FP64, MIXED = 0, 1
SOLVER_CG, SOLVER_GMRES, SOLVER_RICHARDSON, SOLVER_FGMRES, SOLVER_MINRES = 0, 1, 2, 3, 4
UNET_NPARAM = 70980
MLP_HIDDEN, MLP_NPARAM, MLP_MAXSEG = 128, 33793, 8


class FeDesc(C.Structure):
    _fields_ = [('kind', C.c_int32), ('order', C.c_int32), ('ncomp', C.c_int32), ('reserved', C.c_int32),
                ('dirichlet_mask', C.c_uint64)]


class Field(C.Structure):
    _fields_ = [('space', C.c_void_p), ('block', C.c_int32), ('reserved', C.c_int32), ('vec', C.c_void_p)]


class Term(C.Structure):
    _fields_ = [('kind', C.c_int32), ('trial', C.c_int32), ('test', C.c_int32), ('dir', C.c_int32),
                ('coef', C.c_double), ('elem_coef', C.c_void_p), ('wind', C.c_double * 3),
                ('wind_field', C.c_int32), ('weight_mode', C.c_int32), ('weight_field', C.c_int32),
                ('sipg', C.c_int32), ('upwind', C.c_int32), ('reserved', C.c_int32),
                ('eps', C.c_double), ('alpha', C.c_double), ('dirichlet_mask', C.c_uint64)]


class LTerm(C.Structure):
    _fields_ = [('kind', C.c_int32), ('test', C.c_int32), ('fn', C.c_int32), ('field', C.c_int32),
                ('c', C.c_double * 10), ('coef', C.c_double), ('weight_mode', C.c_int32),
                ('weight_field', C.c_int32), ('eps', C.c_double), ('alpha', C.c_double),
                ('wind', C.c_double * 3), ('dirichlet_mask', C.c_uint64)]


class SolveOpts(C.Structure):
    _fields_ = [('solver', C.c_int32), ('precision', C.c_int32), ('restart', C.c_int32), ('maxit', C.c_int32),
                ('tol', C.c_double), ('inner_tol', C.c_double), ('check_every', C.c_int32),
                ('max_outer', C.c_int32), ('s_step', C.c_int32), ('reserved', C.c_int32)]


class SolveStats(C.Structure):
    _fields_ = [('iters', C.c_int32), ('outer', C.c_int32), ('converged', C.c_int32), ('fallback', C.c_int32),
                ('rel_residual', C.c_double), ('initial_rel_residual', C.c_double), ('precond_iters', C.c_int64)]


# (name, argtypes) for every entry point declared in include/ocmp.h
P = C.c_void_p
I64P = C.POINTER(C.c_int64)
SIGNATURES = {
    'ocmp_ctx_create': [C.c_int, P, C.POINTER(P)],
    'ocmp_ctx_destroy': [P],
    'ocmp_ctx_sync': [P],
    'ocmp_ctx_stream': [P],
    'ocmp_ctx_timing': [P, C.c_int],
    'ocmp_ctx_timing_report': [P, C.c_char_p, C.c_int64],
    'ocmp_last_error': [],
    'ocmp_version': [],
    'ocmp_mesh_create': [P, C.c_int, C.c_int64, P, C.c_int64, P, C.c_int64, P, P, C.POINTER(P)],
    'ocmp_mesh_info': [P, I64P],
    'ocmp_mesh_destroy': [P],
    'ocmp_space_create': [P, C.c_int, C.POINTER(FeDesc), C.c_int, C.POINTER(P)],
    'ocmp_space_info': [P, I64P],
    'ocmp_space_blocks': [P, I64P],
    'ocmp_space_freedofs': [P, P],
    'ocmp_space_element_dofs': [P, P],
    'ocmp_space_destroy': [P],
    'ocmp_matrix_create': [P, C.POINTER(P)],
    'ocmp_matrix_info': [P, I64P],
    'ocmp_matrix_pattern': [P, P, P],
    'ocmp_matrix_values': [P, P],
    'ocmp_matrix_destroy': [P],
    'ocmp_assemble_bilinear': [P, C.c_int, C.POINTER(Term), C.c_int, C.POINTER(Field), C.c_int],
    'ocmp_assemble_linear': [P, P, C.c_int, C.POINTER(LTerm), C.c_int, C.POINTER(Field), C.c_int],
    'ocmp_dirichlet_set': [P, C.c_int, P, C.c_int, C.POINTER(C.c_double), C.c_uint64],
    'ocmp_matvec': [P, P, P],
    'ocmp_axpby': [P, C.c_int64, C.c_double, P, C.c_double, P],
    'ocmp_error_norms': [P, C.c_int, P, P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_residual_norms': [P, P, P, C.POINTER(C.c_double)],
    'ocmp_residual_norms_noalloc': [P, P, P, P],
    'ocmp_error_norms_vec': [P, C.c_int, C.POINTER(C.c_int32), P, P, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                             C.POINTER(C.c_double)],
    'ocmp_set': [P, C.c_int, P, C.c_int, C.POINTER(C.c_double)],
    'ocmp_expr_create': [P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)],
    'ocmp_divergence_norm': [P, C.c_int, P, C.POINTER(C.c_double)],
    'ocmp_l2_pair_norms': [P, C.c_int, P, P, P, C.POINTER(C.c_double)],
    'ocmp_linf_norm': [P, C.c_int, P, P, C.POINTER(C.c_double)],
    'ocmp_residual': [P, P, P, P, C.POINTER(C.c_double)],
    'ocmp_precond_create': [P, C.c_int, C.c_int, C.POINTER(P)],
    'ocmp_precond_create_schur': [P, C.c_int, C.c_uint64, C.POINTER(P)],
    'ocmp_precond_set_schur_solve': [P, C.c_int, C.c_double],
    'ocmp_precond_set_block_triangular': [P, P, C.c_int, C.c_double],
    'ocmp_precond_set_bf16': [P, C.c_int],
    'ocmp_precond_set_schur_amg': [P, C.c_int],
    'ocmp_precond_schur_amg_levels': [P, P, P, C.c_int, C.POINTER(C.c_int)],
    'ocmp_precond_set_schur_triangular': [P, C.c_int],
    'ocmp_precond_update': [P],
    'ocmp_precond_update_jacobi': [P],
    'ocmp_precond_destroy': [P],
    'ocmp_solve': [P, P, P, P, C.POINTER(SolveOpts), C.POINTER(SolveStats)],
    'ocmp_hessenberg_eigvals': [C.c_int32, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)],
    'ocmp_newton_shifts': [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                           C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_mlp_create': [P, P, C.c_int64, C.POINTER(P)],
    'ocmp_mlp_destroy': [P],
    'ocmp_mlp_get_params': [P, P],
    'ocmp_mlp_set_params': [P, P],
    'ocmp_mlp_set_scales': [P, C.c_int, I64P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_mlp_get_scales': [P, C.POINTER(C.c_int), I64P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_mlp_fit_scales': [P, P, C.c_int, C.c_int64, C.c_int, I64P],
    'ocmp_mlp_sample': [P, P, C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.c_int64, P, P],
    'ocmp_mlp_forward': [P, C.c_int64, P, P],
    'ocmp_mlp_grad': [P, C.c_int64, P, P, P, C.POINTER(C.c_double)],
    'ocmp_mlp_train_step': [P, C.c_int64, P, P, C.c_float, C.POINTER(C.c_double)],
    'ocmp_mlp_guess': [P, C.c_int64, P, P, P, P, P],
    'ocmp_mlp_set_base': [P, C.c_int],
    'ocmp_mlp_get_base': [P, C.POINTER(C.c_int)],
    'ocmp_unet_create': [P, C.c_int, P, C.POINTER(P)],
    'ocmp_unet_destroy': [P],
    'ocmp_unet_get_params': [P, P],
    'ocmp_unet_set_params': [P, P],
    'ocmp_unet_set_channels': [P, I64P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_unet_forward': [P, P, P],
    'ocmp_unet_grad': [P, P, P, P],
    'ocmp_unet_train_step': [P, P, P, P, C.c_int, C.c_int64, P, C.c_int, C.c_float, C.POINTER(C.c_double)],
    'ocmp_unet_fit_channels': [P, P, P, C.c_int, C.c_int64, I64P],
    'ocmp_unet_get_channels': [P, I64P, C.POINTER(C.c_double), C.POINTER(C.c_double)],
    'ocmp_unet_activation': [P, C.c_int, P, I64P],
    'ocmp_precond_attach_unet': [P, P],
}
RESTYPES = {'ocmp_last_error': C.c_char_p, 'ocmp_ctx_stream': C.c_void_p}

_lib = None


class OcmpError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OcmpError(f'{LIB_PATH} is missing: build it with `python -m opencmp_amd.build` '
                        '(the GPU path has no fallback)')
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(rc: int, what: str = '') -> None:
    if rc != OCMP_OK:
        msg = load().ocmp_last_error().decode(errors='replace')
        raise OcmpError(f'{what} failed ({rc}): {msg}')
