"""Assemble+solve seconds per step on one MI355X: mixed precision + learned guess vs FP64.

Headline workload (the default, config 4 of the benchmark set): transient incompressible
Navier-Stokes, Taylor-Hood P2/P1, implicit Euler dt 1e-3 with the Oseen-Picard linearisation
(<= 3 iterations, tolerance 1e-6), nu 1e-3, on the channel [0,4]x[0,1] with 8 seeded circular
holes (n=228: 396,953 triangles, 1,794,910 DOFs), parabolic inflow Umax 1.5, do-nothing (stress)
outlet with the no-backflow term; FGMRES(20) with the Schur preconditioner (velocity Jacobi,
one smoothed-aggregation AMG V-cycle on C + B D^-1 B^T for the pressure, then the velocity rows
as the upper block-triangular step; the Schur step is rebuilt at each step's first Picard
iteration, the Jacobi rows at every one) to
||b - A x|| / ||b|| <= 1e-8.
A plain FP64 run records 40 steps, the per-DOF MLP is trained on them, then from the same
state both legs time the same steps:
  fp64           : FP64 assembly, FP64 FGMRES + FP64 AMG (the best FP64 path: the same
                   FGMRES/Schur algorithm as the mixed leg)
  mixed+learned  : FP64 assembly and outer residual, FP32 FGMRES + FP32 AMG in iterative
                   refinement, MLP initial guess with the residual fallback
A step = the right-hand side, every Picard iteration's bilinear assembly, preconditioner
setup and solve; inputs (mesh, pattern, schedules) are resident in HBM before the timed
region. The line also carries config 2 (DG-P2 512^2 convection-diffusion, GMRES(50) +
block Jacobi; mixed s-step vs FP64 s-step vs FP64 standard) as extra keys, the measured
device-copy bandwidth, and a CPU baseline timed on the host cores.

    python bench.py [--steps K] [--warmup W] [--config 1..5] [--no-extra] [--no-cpu]
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


def build_problem(config, n, ctx):
    """Returns (host mesh, gpu mesh, space, bilinear form, linear form, solver name, description)."""
    import torch
    from opencmp_amd import core as C, _lib as L
    from opencmp_amd.mesh import unit_square, unit_cube
    if config == 2:
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
        desc = (f'DG-P2 convection-diffusion (SIPG+upwind, eps 1e-3), {n}x{n} split unit square, '
                f'{hm.nel} triangles, GMRES(50) + element-block Jacobi')
        return hm, gm, gs, a, f, 'GMRes', desc
    if config == 3:
        hm = unit_cube(n)
        gm = C.Mesh(hm, ctx)
        gs = C.FESpace(gm, [C.H1(3, '|'.join(hm.bnd_names))])
        # kappa = exp(g), g ~ N(0,1) i.i.d. per element in generator order, seed 0
        kappa = np.exp(np.random.default_rng(0).standard_normal(hm.nel))
        a = C.BilinearForm(gs)
        a += C.Vol('gradgrad', 0, 0, elem_coef=torch.tensor(kappa))
        f = C.LinearForm(gs)
        f += C.Source(0, (L.FN_CONST, (1.0,)))
        desc = (f'3D diffusion -div(kappa grad u) = 1, kappa = exp(N(0,1)) per element, {n}^3 cubes x 6 tets '
                f'({hm.nel} tets), H1 P3, u = 0 on the boundary, Jacobi-PCG')
        return hm, gm, gs, a, f, 'CG', desc
    if config == 1:
        hm = unit_square(n)
        gm = C.Mesh(hm, ctx)
        gs = C.FESpace(gm, [C.H1(2, '|'.join(hm.bnd_names))])
        a = C.BilinearForm(gs)
        a += C.Vol('gradgrad', 0, 0)
        f = C.LinearForm(gs)
        f += C.Source(0, (L.FN_SINPROD, (2 * np.pi ** 2,)))
        desc = f'Poisson H1 P2, {n}x{n} split unit square ({hm.nel} triangles), u = sin(pi x) sin(pi y), Jacobi-PCG'
        return hm, gm, gs, a, f, 'CG', desc
    raise SystemExit(f'unknown config {config}')


def run_path(C, ctx, a, f, gs, precision, tol, restart, steps, warmup, timing=False, solver='GMRes', inner_tol=None,
             maxit=500000, s_step=0):
    import torch
    pre = C.Preconditioner(a, 'local', precision=precision)
    x = C.GridFunction(gs)
    res = None
    inner = (inner_tol or 1e-5) if precision == 'mixed' else 0.0

    def step():
        a.Assemble()
        f.Assemble()
        pre.Update()
        x.vec.zero_()
        return C.solve(a, f.vec, x.vec, pre, solver, tol=tol, restart=restart, precision=precision,
                       maxit=maxit, inner_tol=inner, s_step=s_step)

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
    log(f'{precision}{" s-step" if s_step else ""}: {steps} steps in {t1 - t0:.3f}s, iters {its}, '
        f'rel_res {res.rel_residual:.3e}')
    return (t1 - t0) / steps, res, its, x.vec.clone(), rep


def cpu_baseline(config, a, f, gs, gpu_iters_fp64, n_full):
    """Host-CPU baseline, timed on a bounded sample of the same workload.

    Solve: the oracle's C++/OpenMP restatement (oracle/cpu_solver.cpp, FP64, all host
    cores) runs a fixed number of Krylov iterations of the same algorithm on the same
    full-size CSR matrix (copied from the GPU); its time per iteration times the FP64 GPU
    path's iteration count is the solve time per step. Assembly: the numpy oracle
    (1 thread) on a smaller instance of the same discretisation, scaled linearly in
    elements. Both parts are stated in `sample`."""
    import ctypes as C
    from oracle import fem, cbuild
    from opencmp_amd.mesh import unit_square, unit_cube
    lib = cbuild.load()
    cores = len(os.sched_getaffinity(0))
    lib.ocmp_cpu_threads(cores)
    rp, ci = a.pattern()
    va = np.ascontiguousarray(a.values(), dtype=np.float64)
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    b = f.vec.detach().cpu().numpy().astype(np.float64)
    n = b.size
    x = np.zeros(n)
    rel = C.c_double()
    if config == 2:
        nit = 100
        t0 = time.perf_counter()
        it = lib.ocmp_cpu_gmres_bj(n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, 6, b.ctypes.data,
                                   x.ctypes.data, 50, 1e-30, nit, C.byref(rel))
        t_it = (time.perf_counter() - t0) / max(it, 1)
        ns = 128
        eps, w = 1e-3, np.array([np.cos(0.3), np.sin(0.3)])
        m = unit_square(ns)
        t = fem.Topology(m)
        cs = fem.Compound([fem.Space(t, 'L2', 2)], dgjumps=True)
        A_ = fem.Assembler(cs)
        t0 = time.perf_counter()
        fem.assemble_volume(A_, [dict(kind='grad_grad', trial=0, test=0, coef=eps),
                                 dict(kind='conv', trial=0, test=0, coef=-1.0, wind=w)])
        fem.assemble_dg_facets(A_, 0, eps, 40.0, w, m.marker_mask(m.bnd_names), 0)
        A_.matrix()
        fem.assemble_source(cs, 0, 1.0)
        t_asm = time.perf_counter() - t0
        scale = (n_full / ns) ** 2
        what = f'GMRES(50) + block Jacobi, {it} iterations'
    else:
        nit = 100 if config == 3 else 200
        free = np.ascontiguousarray(gs.FreeDofs().astype(np.uint8))
        t0 = time.perf_counter()
        it = lib.ocmp_cpu_pcg_jacobi(n, rp.ctypes.data, ci.ctypes.data, va.ctypes.data, free.ctypes.data,
                                     b.ctypes.data, x.ctypes.data, 1e-30, nit, C.byref(rel))
        t_it = (time.perf_counter() - t0) / max(it, 1)
        if config == 3:
            ns = 8
            m = unit_cube(ns)
            p, coef, fsrc = 3, np.exp(np.random.default_rng(0).standard_normal(m.nel)), 1.0
        else:
            ns = n_full
            m = unit_square(ns)
            p, coef = 2, 1.0
            fsrc = lambda x_, y_: 2 * np.pi ** 2 * np.sin(np.pi * x_) * np.sin(np.pi * y_)
        t = fem.Topology(m)
        cs = fem.Compound([fem.Space(t, 'H1', p, dirichlet_mask=m.marker_mask(m.bnd_names))])
        A_ = fem.Assembler(cs)
        t0 = time.perf_counter()
        fem.assemble_volume(A_, [dict(kind='grad_grad', trial=0, test=0, coef=coef)])
        A_.matrix()
        fem.assemble_source(cs, 0, fsrc)
        t_asm = time.perf_counter() - t0
        scale = (n_full / ns) ** m.dim
        what = f'Jacobi-PCG, {it} iterations'
    s_solve = t_it * gpu_iters_fp64
    s_asm = t_asm * scale
    return {'value': s_asm + s_solve, 'unit': 's/step', 'cores': cores, 'kind': 'port',
            'solve_s': round(s_solve, 3), 'assembly_s': round(s_asm, 3),
            'sample': f'solve: oracle/cpu_solver.cpp (C++ -O3 OpenMP, FP64, {cores} threads) on the full-size '
                      f'CSR ({n} dofs, {va.size} nnz): {what} at {t_it * 1e3:.2f} ms/iteration, times the FP64 '
                      f'GPU path\'s {gpu_iters_fp64} iterations; assembly: oracle/fem.py (numpy, 1 thread) on '
                      f'the {ns}^{m.dim} instance ({t_asm:.2f} s) scaled by elements ({scale:.0f}x). '
                      f'Host-CPU restatement, not NGSolve.'}


def copy_bandwidth(mib: int = 512, reps: int = 10) -> float:
    """Measured HBM bandwidth of a plain device copy (bytes read + written per second, GB/s)."""
    import torch
    nb = mib * (1 << 20) // 8
    src = torch.ones(nb, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    gbs = 2.0 * nb * 8 * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    return gbs


def cpu_baseline_ins(r, gpu_outer_fp64, amg=0, restart=50, tri=0, n_iter=30):
    """Config 4 on the host cores, timed on the full-size problem: the oracle's C++/OpenMP
    restatement (oracle/cpu_solver.cpp, FP64) assembles the Oseen operator with the wind of the
    last timed step (textbook element loop: quadrature per element, atomic scatter) and runs the
    same FGMRES(restart) + Schur preconditioner (velocity Jacobi; for the pressure `amg`
    smoothed-aggregation V-cycles on C + B D^-1 B^T, or Jacobi-PCG on it to 1e-1 when amg is 0)
    on the FP64 path's last system: its preconditioner setup (S built by rows, the AMG hierarchy
    built) alone, then `n_iter` outer iterations. Per step: (assembly + setup) x the Picard
    iterations of the FP64 leg + time per outer iteration x the FP64 leg's outer iterations."""
    import ctypes as C
    from oracle import cbuild
    from oracle.cpu_baseline import OseenTables
    lib = cbuild.load()
    cores = len(os.sched_getaffinity(0))
    lib.ocmp_cpu_threads(cores)
    m, s = r['fp64_model'], r['fp64_solver']
    rp, ci = s.a.pattern()
    rp = np.ascontiguousarray(rp, dtype=np.int32)
    ci = np.ascontiguousarray(ci, dtype=np.int32)
    va = np.ascontiguousarray(s.a.values(), dtype=np.float64)
    b = s.L.vec.detach().cpu().numpy().astype(np.float64)
    free = np.ascontiguousarray(m.fes.FreeDofs().astype(np.uint8))
    off = [int(v) for v in m.fes.block_offsets()]
    pb = m.fes.comp_blocks[1][0]
    s0, s1 = off[pb], off[pb + 1]
    W = m.W.vec.detach().cpu().numpy().astype(np.float64)
    n = b.size
    T = OseenTables(r['hm'], 2)
    if T.ndof != n or T.poff != s0:
        raise RuntimeError('oracle and GPU compound layouts differ')
    T.assemble(lib, W, r['dt'], r['nu'], rp, ci)            # first touch of the host arrays
    t0 = time.perf_counter()
    T.assemble(lib, W, r['dt'], r['nu'], rp, ci)
    t_asm = time.perf_counter() - t0
    rel, inner = C.c_double(), C.c_int64()
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    x = np.zeros(n)
    t0 = time.perf_counter()
    lib.ocmp_cpu_fgmres_schur(n, P(rp), P(ci), P(va), P(free), s0, s1, P(b), P(x), restart, 1e-30, 0, 1e-1, 200,
                              C.byref(rel), C.byref(inner), amg, None, tri)
    t_setup = time.perf_counter() - t0
    x[:] = 0.0
    t0 = time.perf_counter()
    it = lib.ocmp_cpu_fgmres_schur(n, P(rp), P(ci), P(va), P(free), s0, s1, P(b), P(x), restart, 1e-30, n_iter,
                                   1e-1, 200, C.byref(rel), C.byref(inner), amg, None, tri)
    t_it = (time.perf_counter() - t0 - t_setup) / max(it, 1)
    picard = float(np.mean([row[1] for row in r['rows64']]))
    outer = float(np.mean(gpu_outer_fp64))
    s_asm = (t_asm + t_setup) * picard
    s_solve = t_it * outer
    return {'value': round(s_asm + s_solve, 3), 'unit': 's/step', 'cores': cores, 'kind': 'port',
            'assembly_and_setup_s': round(s_asm, 3), 'solve_s': round(s_solve, 3),
            'sample': f'oracle/cpu_solver.cpp (C++ -O3 OpenMP, FP64, {cores} threads) on the full-size problem '
                      f'({n} dofs, {va.size} nnz): Oseen assembly {t_asm:.3f} s and Schur setup {t_setup:.3f} s, '
                      f'times {picard:.1f} Picard iterations per step; FGMRES({restart}) + Schur ('
                      + (f'{amg} smoothed-aggregation AMG V-cycle(s) per step, the hierarchy rebuilt at each '
                         f'setup' if amg else 'inner Jacobi-PCG to 1e-1') +
                      (', upper block-triangular velocity rows' if tri else '') +
                      f') {it} outer iterations ({inner.value} inner) at {t_it * 1e3:.1f} ms per outer '
                      f'iteration, times the FP64 GPU leg\'s {outer:.0f} outer iterations per step. The outlet\'s '
                      f'boundary term and the right-hand side are not in the CPU assembly. Host-CPU '
                      f'restatement, not NGSolve.'}


def bench_learned(args, ctx):
    """Config 4: transient INS (Taylor-Hood P2/P1, implicit Euler, Oseen Picard) on the
    channel [0,4]x[0,1] with 8 seeded holes. A plain FP64 run records `--record`
    steps; the per-DOF MLP is trained on them with the library's own kernels;
    then, from the same state, `--warmup` + `--steps` steps run on the plain FP64
    path and on the mixed+learned path (FP32 inner GMRES in iterative
    refinement + learned initial guess with the residual fallback)."""
    import torch
    from opencmp_amd import _lib as L
    from opencmp_amd.learned import GuessPolicy, MLPGuess
    from opencmp_amd.mesh import channel_with_holes
    from opencmp_amd.models import INS, const
    from opencmp_amd.solvers import ImplicitEulerSolver
    n = args.n or 228
    if args.schur_iters is None:
        args.schur_iters = 200   # a cap: the tolerance ends the inner solves
    dt = 1e-3
    t0 = time.perf_counter()
    hm = channel_with_holes(n)
    inflow = [(L.FN_POLY2, [0, 0, 6.0, 0, 0, -6.0]), const(0.0)]
    zero = [const(0.0), const(0.0)]

    def model(precision, s_step=5):
        # the plain FP64 path is standard GMRES (s_step 0), as in the other configs
        # do-nothing outlet: the stress condition with zero traction, which carries the no-backflow term
        return INS(hm, kv=1e-3, interp_ord=2, dirichlet={'u': {'inlet': inflow, 'wall': zero, 'cyl': zero}},
                   stress={'u': {'outlet': zero}},
                   linear_solver='GMRes', preconditioner='schur', linear_tolerance=args.tol, restart=args.restart,
                   linear_max_iterations=args.maxit,
                   nonlinear_tolerance=(1e-6, 1e-6), nonlinear_max_iters=3, precision=precision,
                   inner_tol=(args.inner_tol or 1e-4) if precision == 'mixed' else 0.0, s_step=s_step,
                   schur_iters=args.schur_iters, schur_tol=args.schur_tol, schur_amg=args.schur_amg,
                   schur_lag=args.schur_lag, schur_tri=args.schur_tri, ctx=ctx)

    m0 = model('fp64')
    ndof = m0.fes.ndof
    log(f'config 4 setup {time.perf_counter() - t0:.1f}s: {hm.nel} triangles, ndof {ndof}')
    pol = GuessPolicy(ctx, ndof, max_snapshots=args.record + 1)
    s0 = ImplicitEulerSolver(m0, dt=dt, guess=pol)
    t0 = time.perf_counter()
    rec_its = []
    for k in range(args.record):
        s0.step()
        rec_its.append((m0.linear_iterations, m0.picard_iterations))
        log(f'record step {k}: picard {m0.picard_iterations} lin iters {m0.linear_iterations} '
            f'(inner {m0.precond_iterations}) ({time.perf_counter() - t0:.1f}s)')
    torch.cuda.synchronize()
    t_record = time.perf_counter() - t0
    # the network predicts the velocity blocks' correction to the linear extrapolation 2 u^n - u^{n-1};
    # the pressure keeps the plain guess
    net = MLPGuess(ctx, seed=0, max_batch=65536, base=1)
    t0 = time.perf_counter()
    seg = m0.fes.block_offsets()[:len(m0.fes.comp_blocks[0]) + 1]
    losses = net.train(pol.snapshots(), seg, steps=args.train_steps, batch=65536, lr=5e-4,
                       seed=1, log_every=max(1, args.train_steps // 10))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    log(f'training {args.train_steps} Adam steps x 65536 samples: {t_train:.2f}s, loss {losses[0]:.3e} -> '
        f'{losses[-1]:.3e}')
    state = s0.gfu_prev.vec.clone()
    ring = [h.clone() for h in pol.history()]
    wind = m0.W.vec.clone()
    del s0, m0, pol

    def branch(precision, use_net, timing=False):
        m = model(precision, s_step=0 if precision == 'fp64' else 5)
        p = GuessPolicy(ctx, ndof, net=net if use_net else None, n_learned=seg[-1])
        for h in reversed(ring[1:]):
            p.push(h)
        s = ImplicitEulerSolver(m, dt=dt, gfu0=type(m.W)(m.fes, state.clone()), guess=p)
        m.W.vec.copy_(wind)
        rows = []
        for k in range(args.warmup):
            s.step()
            if args.verbose_steps:
                log(f'  {precision} warmup {k}: lin iters {m.linear_iterations}')
        torch.cuda.synchronize()
        if timing:
            ctx.timing(True)
            ctx.timing_report()
        t1 = time.perf_counter()
        per = []
        for k in range(args.steps):
            tk = time.perf_counter()
            r = s.step()
            torch.cuda.synchronize()
            per.append(time.perf_counter() - tk)
            rows.append((m.linear_iterations, m.picard_iterations, r.rel_residual, m.picard_log))
            if args.verbose_steps:
                log(f'  {precision} step {k}: lin iters {m.linear_iterations} (inner {m.precond_iterations}) '
                    f'({time.perf_counter() - t1:.1f}s)')
        torch.cuda.synchronize()
        dt_ = (time.perf_counter() - t1) / args.steps
        rep = None
        if timing:
            rep = ctx.timing_report()
            ctx.timing(False)
        log(f'{precision}{"+learned" if use_net else ""}: {dt_:.4f} s/step, (iters, picard, res) {rows}, '
            f'choices {[c for c, _, _ in s.guess_log]}')
        return dt_, rows, s.gfu_prev.vec.clone(), s.guess_log, rep, float(np.median(per)), (m, s)

    s64, rows64, u64, _, _, med64, ms64 = branch('fp64', False)
    smp, rowsmp, _, _, _, medmp, _ = branch('mixed', False)
    sml, rowsml, uml, glog, _, medml, _ = branch('mixed', True)
    _, _, _, _, rep, _, _ = branch('mixed', True, timing=True)
    err = float(torch.linalg.norm(uml - u64) / torch.linalg.norm(u64))
    return dict(desc=(f'transient INS Taylor-Hood P2/P1, implicit Euler dt 1e-3 + Oseen Picard (<=3, tol 1e-6), '
                      f'channel [0,4]x[0,1] with 8 holes, n={n} ({hm.nel} triangles), do-nothing (stress) outlet '
                      f'with the no-backflow term, ' +
                      (f'FGMRES({args.restart}) + Schur preconditioner with {args.schur_amg} smoothed-aggregation '
                       f'AMG V-cycle(s) on C + B diag(A)^-1 B^T' if args.schur_amg else
                       f'FGMRES({args.restart}) + Schur preconditioner with inner Jacobi-PCG on '
                       f'C + B diag(A)^-1 B^T (to {args.schur_tol:g}, at most {args.schur_iters} iterations)'
                       if args.schur_iters else
                       f'GMRES({args.restart}) + Schur-diagonal preconditioner') +
                      (', upper block-triangular velocity rows' if args.schur_tri and (args.schur_amg or args.schur_iters)
                       else '') +
                      (' (Schur step rebuilt at the first Picard iteration of each step, Jacobi rows at every one)'
                       if args.schur_lag and (args.schur_amg or args.schur_iters) else '')),
                ndof=ndof, s64=s64, sml=sml, rows64=rows64, rowsml=rowsml, err=err, rep=rep, med64=med64,
                medml=medml, fp64_model=ms64[0], fp64_solver=ms64[1], hm=hm, dt=dt, nu=1e-3, n=n,
                smp=smp, medmp=medmp, rowsmp=rowsmp,
                learned=dict(record_steps=args.record, record_s=round(t_record, 2), train_steps=args.train_steps,
                             train_batch=65536, train_s=round(t_train, 3), loss_first=losses[0],
                             loss_last=losses[-1], choices=[c for c, _, _ in glog],
                             residual_plain_vs_learned=[(rp, rl) for _, rp, rl in glog]))


def bench_unet(args, ctx):
    """Config 5: diffuse-interface Stokes on the unit cube, n^3 cubes x 6 tets, Taylor-Hood
    P2/P1, 20 seeded spheres (phi = 0.5 (1 + tanh(d / 2h)), d > 0 in the fluid), penalty
    1e4 (1 - phi), unit inflow in x at x = 0, no-slip on y/z walls, do-nothing at x = 1.
    Plain path: FP64 FGMRES(m) with the Schur-diagonal preconditioner. Learned path: FP32
    inner FGMRES cycles in FP64 iterative refinement with the U-net correction (trained on
    residuals of a plain solve by the library's kernels) and its fallback."""
    import torch
    from opencmp_amd import core as C
    from opencmp_amd.learned import UNetPrecond, residual_samples, residual_norms
    from opencmp_amd.mesh import unit_cube, sphere_obstacles, obstacle_phi
    from opencmp_amd.models import StokesDIM, const
    n = args.n or 64
    t0 = time.perf_counter()
    hm = unit_cube(n)
    phi = obstacle_phi(hm.coords, sphere_obstacles(), 1.0 / n)
    inflow = [const(1.0), const(0.0), const(0.0)]
    zero = [const(0.0)] * 3
    walls = {b: zero for b in ('front', 'back', 'bottom', 'top')}

    def model(precision):
        return StokesDIM(hm, torch.tensor(phi), phi_order=1, penalty=1e4, kv=1.0, interp_ord=2,
                         dirichlet={'u': {'left': inflow, **walls}}, linear_solver='FGMRes', preconditioner='schur',
                         linear_tolerance=args.tol, restart=args.restart, precision=precision,
                         inner_tol=(args.inner_tol or 1e-4) if precision == 'mixed' else 0.0, ctx=ctx)

    m64 = model('fp64')
    a = C.BilinearForm(m64.fes, m64.construct_bilinear_time_coefficient(1.0) + m64.construct_bilinear_time_ODE(1.0))
    f = C.LinearForm(m64.fes, m64.construct_linear(1.0))
    pre64 = m64.construct_preconditioners(a)
    pre32 = C.Preconditioner(a, 'schur', precision='mixed', schur_blocks=m64.schur_blocks())
    if args.vel_iters:
        # block-triangular step: lumped phi-weighted pressure mass for the Schur complement, inner
        # Jacobi-PCG on the velocity block (BF16-stored in the mixed leg with --bf16)
        d = m64.schur_diagonal()
        for p_ in (pre64, pre32):
            p_.set_block_triangular(d, args.vel_iters, args.vel_tol)
        if not args.no_bf16:
            pre32.set_bf16(True)
    ndof = m64.fes.ndof
    log(f'config 5 setup {time.perf_counter() - t0:.1f}s: {hm.nel} tets, ndof {ndof}, nnz {a.nnz}')
    x = C.GridFunction(m64.fes)

    def step(pre, precision):
        """One assemble+solve step. The solve is issued in slices of 20 restart cycles (each
        cycle restarts from the FP64 residual, so slicing at cycle boundaries changes nothing
        for the plain preconditioner) and prints its progress after each slice."""
        a.Assemble()
        f.Assemble()
        pre.Update()
        x.vec.zero_()
        m64.apply_dirichlet_bcs_to(x)
        inner = (args.inner_tol or 1e-4) if precision == 'mixed' else 0.0
        learned = getattr(pre, 'unet', None) is not None
        chunk = args.maxit if learned else 20 * args.restart
        # the solve measures ||b - Ax|| against ||b|| on the free dofs, or against its own initial
        # residual when b vanishes there (flow driven by the inflow data): keep that reference fixed
        r0, bn = residual_norms(a, f.vec, x.vec)
        bref = bn if bn > 0 else r0
        total, outer, r, rc = 0, 0, None, r0
        while total < args.maxit:
            tol = args.tol if bn > 0 else args.tol * r0 / rc
            r = C.solve(a, f.vec, x.vec, pre, 'FGMRes', tol=tol, restart=args.restart,
                        maxit=min(chunk, args.maxit - total), precision=precision, inner_tol=inner)
            total += r.iters
            outer += r.outer
            rc = residual_norms(a, f.vec, x.vec)[0]
            if rc <= args.tol * bref or r.iters == 0:
                break
            log(f'  ... {total} iterations, rel residual {rc / bref:.2e}')
        r.iters, r.outer, r.rel_residual = total, outer, rc / bref
        r.converged = rc <= args.tol * bref
        return r

    def timed(pre, precision, label, timing=False):
        for _ in range(args.warmup):
            step(pre, precision)
        torch.cuda.synchronize()
        if timing:
            ctx.timing(True)
            ctx.timing_report()
        t1 = time.perf_counter()
        rs = [step(pre, precision) for _ in range(args.steps)]
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t1) / args.steps
        rep = None
        if timing:
            rep = ctx.timing_report()
            ctx.timing(False)
        log(f'{label}: {dt:.3f} s/step, iters {[r.iters for r in rs]}, res {rs[-1].rel_residual:.2e}, '
            f'fallback {[r.fallback for r in rs]}')
        return dt, rs[-1], x.vec.clone(), rep

    s64, r64, u64, _ = timed(pre64, 'fp64', 'fp64 plain')
    # training data: normalised residuals along a plain solve from the Dirichlet lift
    a.Assemble()
    f.Assemble()
    pre64.Update()
    x0 = C.GridFunction(m64.fes)
    m64.apply_dirichlet_bcs_to(x0)
    t1 = time.perf_counter()
    S = residual_samples(a, pre64, f.vec, args.unet_samples, 20, x0=x0.vec)
    torch.cuda.synchronize()
    t_samples = time.perf_counter() - t1
    U = UNetPrecond(n, ctx, seed=0)
    blocks = m64.fes.comp_blocks[0] + m64.fes.comp_blocks[1]
    off = [m64.fes.blocks[k][0] for k in blocks]
    t1 = time.perf_counter()
    losses = U.train(a, pre64, S, off, steps=args.train_steps, batch=4, lr=2e-3, seed=1)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    log(f'U-net: {args.unet_samples} samples in {t_samples:.1f}s, {args.train_steps} Adam steps x 4 in '
        f'{t_train:.1f}s, loss {losses[0]:.3e} -> {np.mean(losses[-10:]):.3e}')
    del S
    U.attach(pre32)
    sml, rml, uml, rep = timed(pre32, 'mixed', 'mixed+learned', timing=True)
    U.detach(pre32)
    smx, rmx, umx, _ = timed(pre32, 'mixed', 'mixed (no correction)')
    err = float(torch.linalg.norm(uml - u64) / torch.linalg.norm(u64))
    pdesc = (f'block-triangular preconditioner (lumped phi-weighted pressure mass, inner Jacobi-PCG on the '
             f'velocity block to {args.vel_tol:g}, at most {args.vel_iters} iterations'
             f'{", BF16-stored in the mixed leg" if not args.no_bf16 else ""})' if args.vel_iters else
             'Schur-diagonal preconditioner')
    return dict(desc=(f'diffuse-interface Stokes, unit cube {n}^3 x 6 tets ({hm.nel} tets), Taylor-Hood P2/P1, '
                      f'20 spheres, penalty 1e4, FGMRES({args.restart}) + {pdesc}'),
                ndof=ndof, nnz=a.nnz, s64=s64, sml=sml, smx=smx, r64=r64, rml=rml, rmx=rmx, err=err, rep=rep,
                learned=dict(samples=args.unet_samples, sample_s=round(t_samples, 2), train_steps=args.train_steps,
                             train_s=round(t_train, 2), loss_first=losses[0], loss_last10=float(np.mean(losses[-10:])),
                             fallback=rml.fallback))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--config', type=int, default=4, choices=(1, 2, 3, 4, 5))
    ap.add_argument('--no-extra', action='store_true', help='config 4: skip the config-2 extra keys')
    ap.add_argument('--vel-iters', type=int, default=2000,
                    help='config 5: block-triangular preconditioner with at most this many velocity PCG iterations '
                         '(0: the Schur-diagonal preconditioner, which stalls from 64^3 on)')
    ap.add_argument('--vel-tol', type=float, default=1e-3, help='config 5: velocity PCG reduction per step')
    ap.add_argument('--no-bf16', action='store_true', help='config 5: FP32 (not BF16) velocity block in the mixed leg')
    ap.add_argument('--unet-samples', type=int, default=24, help='config 5: residual samples for training')
    ap.add_argument('--record', type=int, default=40, help='config 4: plain steps recorded for training')
    ap.add_argument('--verbose-steps', action='store_true', help='config 4: log every step of the timed legs')
    ap.add_argument('--schur-tol', type=float, default=1e-1,
                    help='config 4: the inner PCG stops once its preconditioned residual has fallen by this')
    ap.add_argument('--schur-iters', type=int, default=None,
                    help='config 4: at most this many inner PCG iterations on the pressure Schur complement '
                         'per FGMRES step (0: the Schur diagonal under GMRES)')
    ap.add_argument('--train-steps', type=int, default=1000, help='config 4: Adam steps')
    ap.add_argument('--schur-amg', type=int, default=1,
                    help='config 4: smoothed-aggregation AMG V-cycles on the Schur matrix per FGMRES step '
                         '(0: the inner Jacobi-PCG of --schur-tol / --schur-iters instead)')
    ap.add_argument('--schur-lag', type=int, default=1,
                    help='config 4: 1 = Picard iterations after the first refresh only the Jacobi rows of the '
                         'Schur preconditioner (the Schur step lags within a step; 0: full update each time)')
    ap.add_argument('--schur-tri', type=int, default=1,
                    help='config 4: 1 = upper block-triangular velocity rows after the Schur step (0: the '
                         'block-diagonal step)')
    ap.add_argument('--n', type=int, default=None,
                    help='mesh cells per direction (defaults: 32 / 512 / 64 / 228 / 64; config 5 at 64^3 needs the '
                         'block-triangular preconditioner, the default)')
    ap.add_argument('--tol', type=float, default=1e-8)
    ap.add_argument('--restart', type=int, default=None,
                    help='GMRES restart (defaults: 50; 20 for config 4; 200 for config 5 with the Schur diagonal, '
                         'whose diffuse-interface saddle point stagnates under GMRES(50) and GMRES(100) from 32^3 on)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--inner-tol', type=float, default=None, help='mixed path: relative reduction per inner solve')
    ap.add_argument('--maxit', type=int, default=500000, help='Krylov iteration cap (short profiling runs)')
    ap.add_argument('--s-step', type=int, default=5, help='GMRES on the DG operator: 5 = s-step cycles in the '
                    'mixed path (0: standard); the plain FP64 path is always standard GMRES')
    ap.add_argument('--calib-copy', type=int, default=0, metavar='MB',
                    help='first copy MB megabytes with a plain device copy (HBM counter calibration)')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU only: this assemble+solve path is not sharded across GPUs '
              '(see the sharding design); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    n = args.n or {1: 32, 2: 512, 3: 64, 4: 228, 5: 64}[args.config]
    if args.restart is None:
        # config 4: FGMRES(20) (each Picard solve takes 20-45 iterations; a sweep over 10-50 at
        # n=228 put 20 fastest: the orthogonalisation shrinks, the iteration count barely grows)
        args.restart = 200 if (args.config == 5 and not args.vel_iters) else (20 if args.config == 4 else 50)
    import torch
    from opencmp_amd import core as C
    ctx = C.Context.default(0)
    if args.calib_copy:
        # a plain copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch calibrate the HBM counters
        nb = args.calib_copy * (1 << 20) // 8
        src = torch.ones(nb, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f'calibration copy: {args.calib_copy} MiB read + {args.calib_copy} MiB written per dispatch')
        del src, dst
    if args.config == 5:
        r = bench_unet(args, ctx)
        rep = r['rep']
        dom = max((k for k in rep if rep[k][2] > 0), key=lambda k: rep[k][1])
        cnt, ms, byt = rep[dom]
        achieved = byt / (ms * 1e-3) / 1e9
        out = {
            'metric': METRIC, 'value': round(r['sml'], 5), 'unit': 's/step', 'n_gpus': 1, 'steps': args.steps,
            'warmup': args.warmup, 'ms_per_step': round(r['sml'] * 1e3, 3), 'higher_is_better': False,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64+f32+bf16', 'data': 'synthetic',
            'config': {'workload': f"{r['desc']} to rel residual {args.tol:g}", 'config_index': 5,
                       'ndof': r['ndof'], 'nnz': r['nnz'], 'tol': args.tol},
            'plain_fp64': {'s_per_step': round(r['s64'], 5), 'iters': r['r64'].iters,
                           'rel_residual': r['r64'].rel_residual},
            'mixed_learned': {'s_per_step': round(r['sml'], 5), 'iters': r['rml'].iters, 'outer': r['rml'].outer,
                              'fallback': r['rml'].fallback, 'rel_residual': r['rml'].rel_residual,
                              'rel_error_vs_fp64': r['err']},
            'mixed_plain': {'s_per_step': round(r['smx'], 5), 'iters': r['rmx'].iters},
            'speedup_mixed_learned_vs_fp64': round(r['s64'] / r['sml'], 3),
            'dofs_per_s': round(r['ndof'] / r['sml'], 1),
            'learned': r['learned'],
            'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
                         'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': None,
                         'alg_bytes_per_launch': round(byt / cnt), 'avg_launch_us': round(1e3 * ms / cnt, 2)},
            'kernels_mixed_learned': {k: {'launches': c, 'ms': round(m_, 3), 'avg_us': round(1e3 * m_ / max(c, 1), 2),
                                          'alg_bytes_per_launch': round(b_ / max(c, 1))}
                                      for k, (c, m_, b_) in rep.items()},
        }
        print(json.dumps(out), flush=True)
        return
    if args.config == 4:
        r = bench_learned(args, ctx)
        rep = r['rep']
        dom = max((k for k in rep if rep[k][2] > 0), key=lambda k: rep[k][1])
        cnt, ms, byt = rep[dom]
        achieved = byt / (ms * 1e-3) / 1e9
        copy_gbs = copy_bandwidth()
        it64 = [row[0] for row in r['rows64']]
        itml = [row[0] for row in r['rowsml']]
        out = {
            'metric': METRIC, 'value': round(r['sml'], 5), 'unit': 's/step', 'n_gpus': 1, 'steps': args.steps,
            'warmup': args.warmup, 'ms_per_step': round(r['sml'] * 1e3, 3), 'higher_is_better': False,
            'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64+f32+bf16', 'data': 'synthetic',
            'config': {'workload': f"{r['desc']} to rel residual {args.tol:g}", 'config_index': 4,
                       'ndof': r['ndof'], 'tol': args.tol},
            'plain_fp64': {'s_per_step': round(r['s64'], 5), 'median_s_per_step': round(r['med64'], 5),
                           'algorithm': (f'FP64 FGMRES({args.restart}) + Schur (' + (f'FP64 AMG, {args.schur_amg} V-cycle(s)'
                                                                         if args.schur_amg else 'FP64 inner PCG') +
                                         '): the same algorithm as the mixed leg, the best FP64 path of this build'),
                           'outer_iters_per_step': it64, 'iters_picard_res': r['rows64']},
            'mixed_learned': {'s_per_step': round(r['sml'], 5), 'median_s_per_step': round(r['medml'], 5),
                              'outer_iters_per_step': itml, 'iters_picard_res': r['rowsml'],
                              'rel_error_vs_fp64': r['err']},
            'mixed_plain': {'s_per_step': round(r['smp'], 5), 'median_s_per_step': round(r['medmp'], 5),
                            'outer_iters_per_step': [row[0] for row in r['rowsmp']],
                            'first_picard_outer_iters': [row[3][0][0] for row in r['rowsmp']]},
            'speedup_mixed_learned_vs_fp64': round(r['s64'] / r['sml'], 3),
            'learned_gain': {'s_per_step_mixed_plain_over_learned': round(r['smp'] / r['sml'], 3),
                             'outer_iters_plain': sum(row[0] for row in r['rowsmp']),
                             'outer_iters_learned': sum(row[0] for row in r['rowsml']),
                             'first_picard_outer_iters_learned': [row[3][0][0] for row in r['rowsml']]},
            'dofs_per_s': round(r['ndof'] / r['sml'], 1),
            'learned': r['learned'],
            'schur_amg_levels': (r['fp64_solver'].pre.amg_levels() if args.schur_amg else []),
            'roofline': {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
                         'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': None,
                         'alg_bytes_per_launch': round(byt / cnt), 'avg_launch_us': round(1e3 * ms / cnt, 2),
                         'measured_copy_GBps': round(copy_gbs, 1),
                         'frac_of_measured_copy': round(achieved / copy_gbs, 4)},
            'kernels_mixed_learned': {k: {'launches': c, 'ms': round(m_, 3), 'avg_us': round(1e3 * m_ / max(c, 1), 2),
                                          'alg_bytes_per_launch': round(b_ / max(c, 1))}
                                      for k, (c, m_, b_) in rep.items()},
        }
        # HBM bytes per launch: the calibrated FETCH_SIZE + WRITE_SIZE ratio to the algorithmic bytes,
        # measured with rocprofv3 --pmc on this workload, times this run's algorithmic bytes per launch
        tf = os.path.join(ROOT, 'profiles', 'hbm_traffic_ins_channel_n228.json')
        if (args.n or 228) == 228 and os.path.exists(tf):
            ratios = json.load(open(tf))['hbm_bytes_over_algorithmic']
            if dom in ratios:
                out['roofline']['traffic'] = round(ratios[dom] * byt / cnt)
                out['roofline']['traffic_source'] = 'profiles/ins_channel_n228_fgmres_schur_hbm_traffic.md'
        if not args.no_cpu:
            log('cpu baseline (config 4) ...')
            cb = cpu_baseline_ins(r, it64, amg=args.schur_amg, restart=args.restart,
                                  tri=args.schur_tri if (args.schur_amg or args.schur_iters) else 0)
            cb['speedup_gpu_mixed_learned'] = round(cb['value'] / r['sml'], 1)
            cb['speedup_gpu_fp64'] = round(cb['value'] / r['s64'], 1)
            out['cpu_baseline'] = cb
        del r
        if not args.no_extra:
            log('config 2 (extra keys) ...')
            c2 = bench_plain(args, ctx, 2, 512, 2, 1, False)
            out['config2'] = {k: c2[k] for k in ('config', 'plain_fp64', 'mixed', 'fp64_s_step', 'roofline')}
            out['config2']['speedup_mixed_vs_fp64_standard'] = c2['speedup_mixed_vs_fp64']
            if c2.get('fp64_s_step'):
                out['config2']['speedup_mixed_vs_fp64_s_step'] = round(
                    c2['fp64_s_step']['s_per_step'] / c2['mixed']['s_per_step'], 3)
        print(json.dumps(out), flush=True)
        return
    print(json.dumps(bench_plain(args, ctx, args.config, n, args.steps, args.warmup, not args.no_cpu)), flush=True)


def bench_plain(args, ctx, config, n, steps, warmup, with_cpu):
    """Configs 1-3 (and config 2 as the headline line's extra keys): the plain FP64 path, the
    mixed path, the FP64 s-step path where it applies, per-kernel timings and the CPU baseline."""
    import torch
    from opencmp_amd import core as C
    t0 = time.perf_counter()
    hm, gm, gs, a, f, solver, desc = build_problem(config, n, ctx)
    log(f'setup {time.perf_counter() - t0:.1f}s: ndof {gs.ndof} nnz {a.nnz}')
    kw = dict(solver=solver, inner_tol=args.inner_tol, maxit=args.maxit)
    ss = args.s_step if solver == 'GMRes' else 0
    # configs 1-3 keep GMRES(50) when they ride as extra keys on another config's line
    restart = args.restart if args.config == config else 50
    s64, r64, it64, x64, _ = run_path(C, ctx, a, f, gs, 'fp64', args.tol, restart, steps, warmup, **kw)
    smx, rmx, itmx, xmx, _ = run_path(C, ctx, a, f, gs, 'mixed', args.tol, restart, steps, warmup,
                                      s_step=ss, **kw)
    # the same s-step algorithm in FP64, for reference (not the plain path)
    s64s = None
    if ss:
        s64s, r64s, it64s, _, _ = run_path(C, ctx, a, f, gs, 'fp64', args.tol, restart, 1, 0, s_step=ss, **kw)
    # one more step of each path with HIP events around every launch: the per-kernel roofline
    _, _, _, _, rep = run_path(C, ctx, a, f, gs, 'mixed', args.tol, restart, 1, 0, timing=True, s_step=ss, **kw)
    _, _, _, _, rep_fp = run_path(C, ctx, a, f, gs, 'fp64', args.tol, restart, 1, 0, timing=True, **kw)
    err = float(torch.linalg.norm(xmx - x64) / torch.linalg.norm(x64))

    def table(r):
        return {k: {'launches': c, 'ms': round(m_, 3), 'avg_us': round(1e3 * m_ / max(c, 1), 2),
                    'GBps': round(b_ / (m_ * 1e-3) / 1e9, 1) if b_ > 0 and m_ > 0 else None}
                for k, (c, m_, b_) in r.items()}
    dom = max((k for k in rep if rep[k][2] > 0), key=lambda k: rep[k][1])
    cnt, ms, byt = rep[dom]
    achieved = byt / (ms * 1e-3) / 1e9
    roof = {'bound': 'hbm', 'kernel': dom, 'achieved': round(achieved, 1), 'peak': PEAK_HBM_GBS,
            'unit': 'GB/s', 'frac': round(achieved / PEAK_HBM_GBS, 4), 'traffic': None,
            'alg_bytes_per_launch': round(byt / cnt), 'avg_launch_us': round(1e3 * ms / cnt, 2)}
    # HBM bytes per launch: the calibrated FETCH_SIZE + WRITE_SIZE ratio to the algorithmic bytes,
    # measured with rocprofv3 --pmc on this workload, times this run's algorithmic bytes per launch
    tf = os.path.join(ROOT, 'profiles', 'hbm_traffic_dg_p2_512.json')
    if config == 2 and n == 512 and os.path.exists(tf):
        ratios = json.load(open(tf))['hbm_bytes_over_algorithmic']
        if dom in ratios:
            roof['traffic'] = round(ratios[dom] * byt / cnt)
            roof['traffic_source'] = 'profiles/dg_p2_512_gmres_hbm_traffic.md'
    out = {
        'metric': METRIC, 'value': round(smx, 5), 'unit': 's/step', 'n_gpus': 1, 'steps': steps,
        'warmup': warmup, 'ms_per_step': round(smx * 1e3, 3), 'higher_is_better': False, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64+f32', 'data': 'synthetic',
        'config': {'workload': f'{desc} to rel residual {args.tol:g}', 'config_index': config,
                   'ndof': gs.ndof, 'nnz': a.nnz, 'tol': args.tol},
        'plain_fp64': {'s_per_step': round(s64, 5), 'iters': it64[-1], 'rel_residual': r64.rel_residual},
        'mixed': {'s_per_step': round(smx, 5), 'iters': itmx[-1], 'outer': rmx.outer, 's_step': ss,
                  'rel_residual': rmx.rel_residual, 'rel_error_vs_fp64': err},
        'fp64_s_step': ({'s_per_step': round(s64s, 5), 'iters': it64s[-1], 'rel_residual': r64s.rel_residual}
                        if s64s is not None else None),
        'speedup_mixed_vs_fp64': round(s64 / smx, 3),
        'dofs_per_s': round(gs.ndof / smx, 1),
        'roofline': roof,
        'kernels_mixed': table(rep),
        'kernels_fp64': table(rep_fp),
    }
    if with_cpu:
        log('cpu baseline ...')
        cb = cpu_baseline(config, a, f, gs, it64[-1], n)
        cb['speedup_gpu_mixed'] = round(cb['value'] / smx, 1)
        cb['speedup_gpu_fp64'] = round(cb['value'] / s64, 1)
        out['cpu_baseline'] = cb
    return out




if __name__ == '__main__':
    main()
