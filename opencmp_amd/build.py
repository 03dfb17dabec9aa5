"""Builds libocmp_mi355x.so (HIP for gfx950) in-tree with hipcc.

    python -m opencmp_amd.build            # or opencmp_amd.build.build()

The library is one shared object with a C ABI (include/ocmp.h); the Python
side loads it with ctypes. Nothing is JIT-compiled at import time.
"""
from __future__ import annotations

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libocmp_mi355x.so')

SOURCES = ['api.cpp', 'tables.cpp', 'ritz.cpp', 'assemble.hip', 'linalg.hip', 'mlp.hip', 'unet.hip']

# Packed-FP32 instructions are disabled for all device code: hipcc would
# otherwise form v_pk_* FP32 ops from scalar float code.
HIPFLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-fopenmp',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-I' + os.path.join(ROOT, 'include'), '-Wno-unused-result']


def _hipcc():
    for p in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if p and (os.path.isabs(p) and os.path.exists(p) or not os.path.isabs(p)):
            return p
    raise RuntimeError('hipcc not found')


def build(verbose: bool = True, jobs: int = 8) -> str:
    hipcc = _hipcc()
    objdir = os.path.join(HERE, '_build')
    os.makedirs(objdir, exist_ok=True)
    t0 = time.time()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + '.o')
        objs.append(obj)
        lang = ['-x', 'hip'] if src.endswith('.hip') else ['-x', 'hip']
        cmd = [hipcc] + HIPFLAGS + lang + ['-c', os.path.join(CSRC, src), '-o', obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            time.sleep(0.05)
    failed = []
    for src, p in procs:
        out = p.communicate()[0].decode(errors='replace')
        # hipcc's host pass does not know the packed-fp32 target feature; that warning is expected
        lines = [ln for ln in out.splitlines() if 'packed-fp32-ops' not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            sys.stderr.write(f'--- {src} ---\n' + '\n'.join(lines) + '\n')
        elif verbose and lines:
            sys.stderr.write(f'--- {src} (warnings) ---\n' + '\n'.join(lines[:40]) + '\n')
    if failed:
        raise RuntimeError('HIP build failed: ' + ', '.join(failed))
    tmp = LIB + '.tmp'
    cmd = [hipcc, '--offload-arch=gfx950', '-shared', '-fopenmp', '-o', tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors='replace'))
        raise RuntimeError('link failed')
    os.replace(tmp, LIB)
    if verbose:
        print(f'built {LIB} in {time.time() - t0:.1f}s')
    return LIB


if __name__ == '__main__':
    build()
