"""Assemble+solve seconds per step on one MI355X: mixed precision vs plain FP64.

Workload (the single-GPU configuration the headline metric is quoted on):
steady convection-diffusion -eps Lap c + w.grad c = 1 on the unit square,
512x512 squares split into 524,288 triangles, DG P2 (3,145,728 DOFs),
symmetric interior penalty (sigma = 10 p^2 / h) + upwind flux, eps = 1e-3,
w = (cos 0.3, sin 0.3), c = 0 on the boundary (Nitsche), GMRES(50) with
element-block Jacobi to ||b - A x|| / ||b|| <= 1e-8.

A step = bilinear + linear assembly, preconditioner setup, and the Krylov
solve from a zero initial guess; inputs (mesh, pattern, schedules) are resident
in HBM before the timed region. Paths:
  fp64  : FP64 assembly, FP64 GMRES on the block-Jacobi-scaled operator
  mixed : FP64 assembly and outer residual, FP32 inner GMRES cycles (iterative refinement)

    python bench.py [--steps K] [--warmup W] [--n 512] [--tol 1e-8]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'assemble+solve s/step at 1e-8 rel residual: mixed+learned vs plain FP64 GPU'
PEAK_HBM_GBS = 8000.0


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


def build_problem(n, ctx):
    from opencmp_amd import core as C, _lib as L
    from opencmp_amd.mesh import unit_square
    hm = unit_square(n)
    gm = C.Mesh(hm, ctx)
    gs = C.FESpace(gm, [C.L2(2)], dgjumps=True)
    eps, w, alpha = 1e-3, (np.cos(0.3), np.sin(0.3)), 10.0 * 2 * 2
    allm = hm.marker_mask(hm.bnd_names)
    a = C.BilinearForm(gs)
    a += C.Vol('gradgrad', 0, 0, coef=eps)
    a += C.Vol('conv', 0, 0, coef=-1.0, wind=(w[0], w[1], 0.0))
    a += C.DGFacet(0, eps=eps, alpha=alpha, wind=(w[0], w[1]), dirichlet=allm)
    f = C.LinearForm(gs)
    f += C.Source(0, (L.FN_CONST, (1.0,)))
    return hm, gm, gs, a, f


def run_path(C, ctx, a, f, gs, precision, tol, restart, steps, warmup, timing=False):
    import torch
    pre = C.Preconditioner(a, 'local', precision=precision)
    x = C.GridFunction(gs)
    res = None

    def step():
        a.Assemble()
        f.Assemble()
        pre.Update()
        x.vec.zero_()
        return C.solve(a, f.vec, x.vec, pre, 'GMRes', tol=tol, restart=restart, precision=precision,
                       maxit=200000, inner_tol=1e-5 if precision == 'mixed' else 0.0)

    for i in range(warmup):
        res = step()
        log(f'{precision} warmup {i}: iters {res.iters} outer {res.outer} rel_res {res.rel_residual:.3e}')
    torch.cuda.synchronize()
    if timing:
        ctx.timing(True)
        ctx.timing_report()  # reset
    t0 = time.perf_counter()
    its = []
    for i in range(steps):
        res = step()
        its.append(res.iters)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    rep = None
    if timing:
        rep = ctx.timing_report()
        ctx.timing(False)
    log(f'{precision}: {steps} steps in {t1 - t0:.3f}s, iters {its}, rel_res {res.rel_residual:.3e}')
    return (t1 - t0) / steps, res, its, x.vec.clone(), rep


def cpu_baseline(gpu_iters_fp64, n_full):
    """Oracle (numpy/scipy FP64, 1 thread) timed on a bounded sample of the same workload.

    Sample: the 128x128 instance of the same discretisation; assembly is timed
    in full and GMRES(50) + block Jacobi for 100 iterations; both scale linearly
    in the number of elements/DOFs, so s/step at the full size is
    t_asm * (n_full/128)^2 + t_iter * (n_full/128)^2 * (iterations of the FP64 GPU path)."""
    from oracle import fem
    from opencmp_amd.mesh import unit_square
    ns = 128
    eps, w = 1e-3, np.array([np.cos(0.3), np.sin(0.3)])
    m = unit_square(ns)
    t0 = time.perf_counter()
    t = fem.Topology(m)
    cs = fem.Compound([fem.Space(t, 'L2', 2)], dgjumps=True)
    A_ = fem.Assembler(cs)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    fem.assemble_volume(A_, [dict(kind='grad_grad', trial=0, test=0, coef=eps),
                             dict(kind='conv', trial=0, test=0, coef=-1.0, wind=w)])
    fem.assemble_dg_facets(A_, 0, eps, 40.0, w, m.marker_mask(m.bnd_names), 0)
    A = A_.matrix()
    b = fem.assemble_source(cs, 0, 1.0)
    Bi = fem.block_jacobi_inverse(A, 6)
    t_asm = time.perf_counter() - t0
    M = lambda v: np.einsum('eij,ej->ei', Bi, v.reshape(-1, 6)).ravel()
    nit = 100
    t0 = time.perf_counter()
    fem.gmres_right(A, b, np.zeros(cs.ndof), M, 1e-30, m=50, maxit=nit)
    t_it = (time.perf_counter() - t0) / nit
    scale = (n_full / ns) ** 2
    s_step = t_asm * scale + t_it * scale * gpu_iters_fp64
    return {'value': s_step, 'unit': 's/step', 'cores': 1, 'kind': 'port',
            'sample': f'oracle/fem.py (numpy+scipy, FP64, 1 thread) on the {ns}x{ns} instance: full assembly '
                      f'+ block-Jacobi setup ({t_asm:.2f}s) and {nit} GMRES(50) iterations ({t_it * 1e3:.2f} ms/iter), '
                      f'scaled by DOFs ({scale:.0f}x) to {n_full}x{n_full} and by the FP64 GPU path\'s '
                      f'{gpu_iters_fp64} iterations'}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=512)
    ap.add_argument('--tol', type=float, default=1e-8)
    ap.add_argument('--restart', type=int, default=50)
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU only: this assemble+solve path is not sharded across GPUs '
              '(see the sharding design); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    from opencmp_amd import core as C
    ctx = C.Context.default(0)
    t0 = time.perf_counter()
    hm, gm, gs, a, f = build_problem(args.n, ctx)
    log(f'setup {time.perf_counter() - t0:.1f}s: ndof {gs.ndof} nnz {a.nnz}')
    s64, r64, it64, x64, rep64 = run_path(C, ctx, a, f, gs, 'fp64', args.tol, args.restart, args.steps,
                                          args.warmup, timing=False)
    smx, rmx, itmx, xmx, repmx = run_path(C, ctx, a, f, gs, 'mixed', args.tol, args.restart, args.steps,
                                          args.warmup, timing=False)
    # instrumented step (HIP events around every launch) for the per-kernel roofline
    _, _, _, _, rep = run_path(C, ctx, a, f, gs, 'mixed', args.tol, args.restart, 1, 0, timing=True)
    _, _, _, _, rep_fp = run_path(C, ctx, a, f, gs, 'fp64', args.tol, args.restart, 1, 0, timing=True)
    err = float(torch.linalg.norm(xmx - x64) / torch.linalg.norm(x64))
    kernels = {}
    for k, (cnt, ms, byt) in rep.items():
        kernels[k] = {'launches': cnt, 'ms': round(ms, 3), 'avg_us': round(1e3 * ms / max(cnt, 1), 2),
                      'GBps': round(byt / (ms * 1e-3) / 1e9, 1) if byt > 0 and ms > 0 else None}
    dom = max((k for k in rep if rep[k][2] > 0), key=lambda k: rep[k][1])
    cnt, ms, byt = rep[dom]
    achieved = byt / (ms * 1e-3) / 1e9
    roof = {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
            'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': None,
            'alg_bytes_per_launch': byt / cnt, 'avg_launch_us': 1e3 * ms / cnt}
    out = {
        'metric': METRIC, 'value': round(smx, 5), 'unit': 's/step', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(smx * 1e3, 3), 'higher_is_better': False, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64+f32', 'data': 'synthetic',
        'config': {'workload': f'DG-P2 convection-diffusion (SIPG+upwind), {args.n}x{args.n} split unit square, '
                               f'{hm.nel} triangles, GMRES({args.restart}) + block Jacobi to rel residual {args.tol:g}',
                   'ndof': gs.ndof, 'nnz': a.nnz, 'eps': 1e-3, 'tol': args.tol},
        'plain_fp64': {'s_per_step': round(s64, 5), 'iters': it64[-1], 'rel_residual': r64.rel_residual},
        'mixed': {'s_per_step': round(smx, 5), 'iters': itmx[-1], 'outer': rmx.outer,
                  'rel_residual': rmx.rel_residual, 'rel_error_vs_fp64': err},
        'speedup_mixed_vs_fp64': round(s64 / smx, 3),
        'dofs_per_s': round(gs.ndof / smx, 1),
        'roofline': roof,
        'kernels_mixed': kernels,
        'kernels_fp64': {k: {'launches': c, 'ms': round(m_, 3),
                             'GBps': round(b_ / (m_ * 1e-3) / 1e9, 1) if b_ > 0 and m_ > 0 else None}
                         for k, (c, m_, b_) in rep_fp.items()},
    }
    if not args.no_cpu:
        log('cpu baseline ...')
        cb = cpu_baseline(it64[-1], args.n)
        cb['speedup_gpu_mixed'] = round(cb['value'] / smx, 1)
        out['cpu_baseline'] = cb
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
