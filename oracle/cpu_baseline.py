"""Inputs for the C++ restatement's Oseen assembly (oracle/cpu_solver.cpp) -- TEST
INFRASTRUCTURE for the CPU baseline and its tests, never the product.

The element tables come from the oracle's own topology and spaces (oracle/fem.py), so the
C++ element loop assembles exactly the discretisation of the oracle's assemble_volume: the
Taylor-Hood pair (H1 order p velocity, order p-1 pressure), dof numbering of the compound
space [velocity components..., pressure], reference basis values and barycentric derivatives
tabulated at a quadrature rule exact for every term (degree 3p - 1: the convection integrand
phi_j (w . grad phi_i) with a P_p wind).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from oracle import fem


class OseenTables:
    def __init__(self, host_mesh, p: int = 2):
        self.topo = t = fem.Topology(host_mesh)
        self.dim = d = t.dim
        su = fem.Space(t, 'H1', p, ncomp=d)
        sp_ = fem.Space(t, 'H1', p - 1)
        self.cs = fem.Compound([su, sp_])
        self.su, self.sp = su, sp_
        self.nu1 = su.n1
        self.poff = int(self.cs.offsets[1])
        self.ndof = self.cs.ndof
        lam, w = fem.simplex_rule(d, 3 * p - 1)
        self.nq = len(w)
        self.qw = np.ascontiguousarray(w, dtype=np.float64)
        pu, du = fem.h1_eval(d, p, lam)
        pp, dp = fem.h1_eval(d, p - 1, lam)
        self.phiu, self.dlu = np.ascontiguousarray(pu), np.ascontiguousarray(du)
        self.phip, self.dlp = np.ascontiguousarray(pp), np.ascontiguousarray(dp)
        self.udofs = np.ascontiguousarray(su.edofs1, dtype=np.int32)
        self.pdofs = np.ascontiguousarray(sp_.edofs1, dtype=np.int32)
        X = t.coords[t.elems]                                    # (nel, d+1, d), sorted local vertices
        self.X = np.ascontiguousarray(X, dtype=np.float64)

    def assemble(self, lib, W, dt, nu, rp, col, peps=1e-10):
        """CSR values of the implicit-Euler Oseen operator on the pattern (rp, col)."""
        rp = np.ascontiguousarray(rp, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        W = np.ascontiguousarray(W, dtype=np.float64)
        val = np.zeros(col.size)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        bad = lib.ocmp_cpu_assemble_oseen(self.dim, self.topo.nel, P(self.X), self.udofs.shape[1], P(self.udofs),
                                          self.pdofs.shape[1], P(self.pdofs), self.nu1, self.poff, self.nq,
                                          P(self.qw), P(self.phiu), P(self.dlu), P(self.phip), P(self.dlp), P(W),
                                          float(dt), float(nu), float(peps), P(rp), P(col), P(val))
        if bad:
            raise ValueError(f'{bad} element entries outside the CSR pattern')
        return val
