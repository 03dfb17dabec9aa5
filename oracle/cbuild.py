"""Builds the oracle's C++ restatement of the assembly and the Krylov solves
(oracle/cpu_solver.cpp) into oracle/libocmp_cpu.so -- test infrastructure for the CPU
baseline, never the product."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'cpu_solver.cpp')
LIB = os.path.join(HERE, 'libocmp_cpu.so')


def build(verbose: bool = False) -> str:
    # x86-64-v3 (AVX2/FMA): runs on the Intel host that builds it and on the AMD EPYC hosts that load it
    cmd = ['g++', '-O3', '-march=x86-64-v3', '-fopenmp', '-fPIC', '-shared', '-std=c++17', SRC, '-o', LIB]
    if verbose:
        print(' '.join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    return LIB


def load():
    import ctypes as C
    if not os.path.exists(LIB):
        raise RuntimeError(f'{LIB} is missing: run oracle.cbuild.build()')
    lib = C.CDLL(LIB)
    P = C.c_void_p
    lib.ocmp_cpu_threads.argtypes = [C.c_int]
    lib.ocmp_cpu_threads.restype = C.c_int
    lib.ocmp_cpu_gmres_bj.argtypes = [C.c_int64, P, P, P, C.c_int, P, P, C.c_int, C.c_double, C.c_int,
                                      C.POINTER(C.c_double)]
    lib.ocmp_cpu_gmres_bj.restype = C.c_int
    lib.ocmp_cpu_pcg_jacobi.argtypes = [C.c_int64, P, P, P, P, P, P, C.c_double, C.c_int, C.POINTER(C.c_double)]
    lib.ocmp_cpu_pcg_jacobi.restype = C.c_int
    I64, I32, D = C.c_int64, C.c_int, C.c_double
    lib.ocmp_cpu_assemble_oseen.argtypes = [I32, I64, P, I32, P, I32, P, I64, I64, I32, P, P, P, P, P, P, D, D, D,
                                            P, P, P]
    lib.ocmp_cpu_assemble_oseen.restype = C.c_int
    lib.ocmp_cpu_fgmres_schur.argtypes = [I64, P, P, P, P, I64, I64, P, P, I32, D, I32, D, I32,
                                          C.POINTER(C.c_double), C.POINTER(C.c_int64), I32, P, I32]
    lib.ocmp_cpu_pcg_amg.argtypes = [I64, P, P, P, P, P, P, D, I32, I32, P, C.POINTER(C.c_double)]
    lib.ocmp_cpu_pcg_amg.restype = C.c_int
    lib.ocmp_cpu_fgmres_schur.restype = C.c_int
    return lib


if __name__ == '__main__':
    print(build(verbose=True))
