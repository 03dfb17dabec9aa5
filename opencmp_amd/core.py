"""Host objects over the C-ABI: context, mesh, spaces, grid functions, forms, preconditioners.

Names mirror the NGSolve objects OpenCMP's models and solvers use
(``FESpace``, ``GridFunction``, ``BilinearForm``, ``LinearForm``,
``Preconditioner``), with the weak forms written as explicit integrator terms
instead of symbolic coefficient-function trees (NGSolve is not part of this
path). Device vectors are torch tensors on ``cuda:N``; torch is used for
device memory and the stream only, all arithmetic runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib as L
from .expr import resolve_fn
from .mesh import Mesh as HostMesh


def _ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


class Context:
    """One HIP stream on one device (replaces the TaskManager scope of opencmp/run.py:57-61)."""

    _default = {}

    def __init__(self, device: int = 0):
        self.lib = L.load()
        if not torch.cuda.is_available():
            raise L.OcmpError('no HIP device visible: the assemble+solve path runs on the GPU only')
        self.device = device
        self.torch_device = torch.device('cuda', device)
        with torch.cuda.device(device):
            self.stream = torch.cuda.current_stream(device)
        h = C.c_void_p()
        L.check(self.lib.ocmp_ctx_create(device, C.c_void_p(self.stream.cuda_stream), C.byref(h)), 'ctx_create')
        self.h = h

    @classmethod
    def default(cls, device: int = 0) -> 'Context':
        if device not in cls._default:
            cls._default[device] = cls(device)
        return cls._default[device]

    def sync(self):
        L.check(self.lib.ocmp_ctx_sync(self.h), 'ctx_sync')

    def timing(self, on: bool):
        L.check(self.lib.ocmp_ctx_timing(self.h, 1 if on else 0), 'ctx_timing')

    def timing_report(self) -> dict:
        import json
        buf = C.create_string_buffer(1 << 16)
        L.check(self.lib.ocmp_ctx_timing_report(self.h, buf, len(buf)), 'timing_report')
        return json.loads(buf.value.decode())

    def empty(self, n, dtype=torch.float64):
        return torch.empty(n, dtype=dtype, device=self.torch_device)

    def zeros(self, n, dtype=torch.float64):
        return torch.zeros(n, dtype=dtype, device=self.torch_device)


class Mesh:
    """Device-side mesh (replaces ngs.Mesh from opencmp/helpers/io.py:69-110)."""

    def __init__(self, host: HostMesh, ctx: Optional[Context] = None):
        host.validate()
        self.ctx = ctx or Context.default()
        self.host = host
        self.dim = host.dim
        coords = np.ascontiguousarray(host.coords, dtype=np.float64)
        elems = np.ascontiguousarray(host.elems, dtype=np.int32)
        bf = np.ascontiguousarray(host.bnd_facets, dtype=np.int32)
        bm = np.ascontiguousarray(host.bnd_markers, dtype=np.int32)
        h = C.c_void_p()
        L.check(self.ctx.lib.ocmp_mesh_create(self.ctx.h, self.dim, coords.shape[0], coords.ctypes.data,
                                              elems.shape[0], elems.ctypes.data, bf.shape[0],
                                              bf.ctypes.data if bf.size else None,
                                              bm.ctypes.data if bm.size else None, C.byref(h)), 'mesh_create')
        self.h = h
        info = (C.c_int64 * 8)()
        L.check(self.ctx.lib.ocmp_mesh_info(self.h, info), 'mesh_info')
        self.nv, self.nedge, self.nface, self.ne, self.nfacet, self.ninterior, self.nbnd, _ = list(info)

    def marker_mask(self, names) -> int:
        return self.host.marker_mask(names)

    def GetBoundaries(self):
        return list(self.host.bnd_names)

    def __del__(self):
        try:
            self.ctx.lib.ocmp_mesh_destroy(self.h)
        except Exception:
            pass


@dataclass
class H1:
    order: int
    dirichlet: str = ''
    ncomp: int = 1


@dataclass
class L2:
    order: int


def VectorH1(order: int, dirichlet: str = '', dim: int = 2) -> H1:
    return H1(order, dirichlet, ncomp=dim)


class FESpace:
    """Compound space (ngs.FESpace([...], dgjumps=...)); components blocked in order.

    ``comp_blocks[c]`` lists the scalar block indices of component c (a vector
    component expands into ``dim`` blocks).
    """

    def __init__(self, mesh: Mesh, comps: Sequence[Union[H1, L2]], dgjumps: bool = False):
        self.mesh = mesh
        self.ctx = mesh.ctx
        self.comps = list(comps)
        self.dgjumps = dgjumps
        descs = (L.FeDesc * len(comps))()
        for i, c in enumerate(comps):
            if isinstance(c, H1):
                mask = mesh.marker_mask(c.dirichlet) if c.dirichlet else 0
                descs[i] = L.FeDesc(L.OCMP_H1, c.order, c.ncomp, 0, mask)
            else:
                descs[i] = L.FeDesc(L.OCMP_L2, c.order, 1, 0, 0)
        h = C.c_void_p()
        L.check(self.ctx.lib.ocmp_space_create(mesh.h, len(comps), descs, 1 if dgjumps else 0, C.byref(h)),
                'space_create')
        self.h = h
        info = (C.c_int64 * 4)()
        L.check(self.ctx.lib.ocmp_space_info(self.h, info), 'space_info')
        self.ndof, self.nloc, self.nblocks, self.nfree = list(info)
        b = (C.c_int64 * (4 * self.nblocks))()
        L.check(self.ctx.lib.ocmp_space_blocks(self.h, b), 'space_blocks')
        self.blocks = [tuple(b[4 * i:4 * i + 4]) for i in range(self.nblocks)]   # (dof_off, n, loc_off, nloc)
        self.comp_blocks = []
        k = 0
        for c in comps:
            nc = c.ncomp if isinstance(c, H1) else 1
            self.comp_blocks.append(list(range(k, k + nc)))
            k += nc
        self._free = None
        self._free_dev = None

    def FreeDofs(self) -> np.ndarray:
        if self._free is None:
            m = np.empty(self.ndof, dtype=np.uint8)
            L.check(self.ctx.lib.ocmp_space_freedofs(self.h, m.ctypes.data), 'space_freedofs')
            self._free = m.astype(bool)
        return self._free

    def element_dofs(self) -> np.ndarray:
        out = np.empty((self.mesh.ne, self.nloc), dtype=np.int32)
        L.check(self.ctx.lib.ocmp_space_element_dofs(self.h, out.ctypes.data), 'space_element_dofs')
        return out

    def block_offsets(self) -> List[int]:
        """[off_0, off_1, ..., ndof]: the dof ranges of the scalar blocks."""
        return [b[0] for b in self.blocks] + [self.ndof]

    def block_slice(self, blk: int) -> slice:
        off, n, _, _ = self.blocks[blk]
        return slice(off, off + n)

    def __del__(self):
        try:
            self.ctx.lib.ocmp_space_destroy(self.h)
        except Exception:
            pass


class GridFunction:
    """A device dof vector on a space (ngs.GridFunction)."""

    def __init__(self, fes: FESpace, vec: Optional[torch.Tensor] = None):
        self.space = fes
        self.vec = vec if vec is not None else fes.ctx.zeros(fes.ndof)

    def field(self, block: int) -> L.Field:
        return L.Field(self.space.h, block, 0, _ptr(self.vec))

    def Set(self, fn, boundaries: Optional[str] = None, component: int = 0):
        """gfu.Set(g, definedon=mesh.Boundaries(...)) for a closed-form or expression g on each block of a
        component; without ``boundaries``, gfu.Set(g) over the whole domain (projection-based
        interpolation on H1 blocks, element-wise L2 projection on L2 blocks).

        fn = (FN_*, params) or a bound expression; for a vector component pass a list, one per block.
        """
        blocks = self.space.comp_blocks[component]
        fns = fn if isinstance(fn, list) else [fn] * len(blocks)
        if len(fns) != len(blocks):
            raise ValueError(f'{len(fns)} functions for a component of {len(blocks)} scalar blocks')
        mask = None if boundaries is None else self.space.mesh.marker_mask(boundaries)
        for blk, f in zip(blocks, fns):
            fid, params = resolve_fn(f)
            c = (C.c_double * 10)(*([float(v) for v in params] + [0.0] * (10 - len(params))))
            if mask is None:
                L.check(self.space.ctx.lib.ocmp_set(self.space.h, blk, C.c_void_p(_ptr(self.vec)), fid, c), 'set')
            else:
                L.check(self.space.ctx.lib.ocmp_dirichlet_set(self.space.h, blk, C.c_void_p(_ptr(self.vec)), fid,
                                                              c, mask), 'dirichlet_set')

    def numpy(self) -> np.ndarray:
        return self.vec.detach().cpu().numpy()


# ------------------------------------------------------------------ integrator terms


@dataclass
class Vol:
    """Volume term on scalar blocks: kind in gradgrad|mass|conv|dtrial|dtest."""
    kind: str
    trial: int
    test: int
    coef: float = 1.0
    elem_coef: Optional[torch.Tensor] = None
    wind: Tuple[float, ...] = (0.0, 0.0, 0.0)
    wind_gf: Optional[Tuple['GridFunction', Sequence[int]]] = None   # (gf, blocks per dimension)
    weight_gf: Optional[Tuple['GridFunction', int, int]] = None      # (gf, block, WEIGHT_* mode)
    dir: int = 0


@dataclass
class BndUpwind:
    """coef * max(w.n, 0) * u v on the boundary facets whose marker is in `markers` (bit mask), one
    H1 block (trial == test): the no-backflow part of the stress boundary condition. signed=True
    gives coef * (w.n) * u v instead: the convective flux of a Neumann boundary."""
    block: int
    markers: int
    coef: float = 1.0
    wind: Tuple[float, ...] = (0.0, 0.0, 0.0)
    wind_gf: Optional[Tuple['GridFunction', Sequence[int]]] = None   # (gf, blocks per dimension)
    signed: bool = False


@dataclass
class BndMass:
    """coef * u v on the boundary facets whose marker is in `markers` (bit mask), one H1 block:
    the Robin term."""
    block: int
    markers: int
    coef: float = 1.0


@dataclass
class BndSource:
    """coef * g * v on the boundary facets whose marker is in `markers`, g = (FN_*, params):
    Neumann and Robin data."""
    test: int
    markers: int
    fn: Tuple[int, Sequence[float]] = (L.FN_CONST, (0.0,))
    coef: float = 1.0


@dataclass
class DGFacet:
    """Interior-penalty + upwind skeleton terms of an L2 block."""
    block: int
    eps: float
    alpha: float
    wind: Tuple[float, float] = (0.0, 0.0)
    dirichlet: int = 0          # marker mask
    sipg: bool = True
    upwind: bool = True
    coef: float = 1.0


@dataclass
class Source:
    """coef * f * v with f = (FN_*, params) or a field (gf, block)."""
    test: int
    fn: Tuple[int, Sequence[float]] = (L.FN_CONST, (1.0,))
    coef: float = 1.0
    field_gf: Optional[Tuple['GridFunction', int]] = None
    weight_gf: Optional[Tuple['GridFunction', int, int]] = None


@dataclass
class DGDirichletData:
    """eps*g*(alpha/h v - grad v.n) - v g min(w.n,0) on Dirichlet facets."""
    test: int
    eps: float
    alpha: float
    wind: Tuple[float, float]
    dirichlet: int
    fn: Tuple[int, Sequence[float]] = (L.FN_CONST, (0.0,))
    coef: float = 1.0


_KIND = {'gradgrad': L.BFI_GRADGRAD, 'mass': L.BFI_MASS, 'conv': L.BFI_CONV, 'dtrial': L.BFI_DTRIAL,
         'dtest': L.BFI_DTEST}


class _FieldTable:
    def __init__(self):
        self.items = []
        self.keys = {}

    def index(self, gf: GridFunction, block: int) -> int:
        key = (id(gf), block)
        if key not in self.keys:
            self.keys[key] = len(self.items)
            self.items.append(gf.field(block))
        return self.keys[key]

    def array(self):
        arr = (L.Field * max(1, len(self.items)))()
        for i, f in enumerate(self.items):
            arr[i] = f
        return arr, len(self.items)


class BilinearForm:
    """A list of terms assembled into the device CSR matrix (BilinearForm + Assemble())."""

    def __init__(self, fes: FESpace, terms: Optional[list] = None):
        self.space = fes
        self.terms = list(terms or [])
        self.ctx = fes.ctx
        h = C.c_void_p()
        L.check(self.ctx.lib.ocmp_matrix_create(fes.h, C.byref(h)), 'matrix_create')
        self.h = h
        info = (C.c_int64 * 4)()
        L.check(self.ctx.lib.ocmp_matrix_info(self.h, info), 'matrix_info')
        self.n, self.nnz, self.ncolours, self.is_dg = list(info)
        self._keep = []

    def __iadd__(self, term):
        self.terms.append(term)
        return self

    def _ctypes_terms(self, terms, ft: _FieldTable):
        arr = (L.Term * len(terms))()
        keep = []
        for i, t in enumerate(terms):
            tt = L.Term()
            tt.wind_field = -1
            tt.weight_field = -1
            if isinstance(t, BndMass):
                tt.kind = L.BFI_BND_MASS
                tt.trial = tt.test = t.block
                tt.coef = t.coef
                tt.dirichlet_mask = t.markers
            elif isinstance(t, BndUpwind):
                tt.kind = L.BFI_BND_FLUX if t.signed else L.BFI_BND_UPWIND
                tt.trial = tt.test = t.block
                tt.coef = t.coef
                tt.dirichlet_mask = t.markers
                for c in range(min(3, len(t.wind))):
                    tt.wind[c] = t.wind[c]
                if t.wind_gf is not None:
                    gf, blocks = t.wind_gf
                    idx = [ft.index(gf, b) for b in blocks]
                    if idx != list(range(idx[0], idx[0] + len(idx))):
                        raise ValueError('wind components must be consecutive fields')
                    tt.wind_field = idx[0]
            elif isinstance(t, DGFacet):
                tt.kind = L.BFI_DG_FACET
                tt.trial = tt.test = t.block
                tt.coef = t.coef
                tt.eps = t.eps
                tt.alpha = t.alpha
                tt.wind[0], tt.wind[1] = t.wind[0], t.wind[1]
                tt.dirichlet_mask = t.dirichlet
                tt.sipg = 1 if t.sipg else 0
                tt.upwind = 1 if t.upwind else 0
            else:
                tt.kind = _KIND[t.kind]
                tt.trial, tt.test, tt.dir = t.trial, t.test, t.dir
                tt.coef = t.coef
                if t.elem_coef is not None:
                    ec = t.elem_coef.to(device=self.ctx.torch_device, dtype=torch.float64).contiguous()
                    if ec.numel() != self.space.mesh.ne:
                        raise ValueError(f'elem_coef has {ec.numel()} entries, the mesh {self.space.mesh.ne} elements')
                    keep.append(ec)
                    tt.elem_coef = _ptr(ec)
                for c in range(min(3, len(t.wind))):
                    tt.wind[c] = t.wind[c]
                if t.wind_gf is not None:
                    gf, blocks = t.wind_gf
                    idx = [ft.index(gf, b) for b in blocks]
                    if idx != list(range(idx[0], idx[0] + len(idx))):
                        raise ValueError('wind components must be consecutive fields')
                    tt.wind_field = idx[0]
                if t.weight_gf is not None:
                    gf, blk, mode = t.weight_gf
                    tt.weight_field = ft.index(gf, blk)
                    tt.weight_mode = mode
            arr[i] = tt
        return arr, keep

    def Assemble(self):
        ft = _FieldTable()
        # register wind fields first so each wind gets consecutive indices
        for t in self.terms:
            if isinstance(t, (Vol, BndUpwind)) and t.wind_gf is not None:
                gf, blocks = t.wind_gf
                for b in blocks:
                    ft.index(gf, b)
        # H1: one call (the library splits what its element scatter cannot take at once, and the row
        # gather forms all volume terms in one pass); DG: at most 8 terms per call
        per = 8 if self.is_dg else 32
        chunks = [self.terms[i:i + per] for i in range(0, max(1, len(self.terms)), per)]
        for ci, chunk in enumerate(chunks):
            arr, keep = self._ctypes_terms(chunk, ft)
            farr, nf = ft.array()
            L.check(self.ctx.lib.ocmp_assemble_bilinear(self.h, len(chunk), arr, nf, farr, 1 if ci else 0),
                    'assemble_bilinear')
            self._keep = keep   # device coefficient tensors must outlive the enqueued kernels
        return self

    def mult(self, x: torch.Tensor, y: torch.Tensor):
        L.check(self.ctx.lib.ocmp_matvec(self.h, C.c_void_p(_ptr(x)), C.c_void_p(_ptr(y))), 'matvec')

    def pattern(self):
        rp = np.empty(self.n + 1, dtype=np.int32)
        col = np.empty(self.nnz, dtype=np.int32)
        L.check(self.ctx.lib.ocmp_matrix_pattern(self.h, rp.ctypes.data, col.ctypes.data), 'matrix_pattern')
        return rp, col

    def values(self) -> np.ndarray:
        v = np.empty(self.nnz, dtype=np.float64)
        L.check(self.ctx.lib.ocmp_matrix_values(self.h, v.ctypes.data), 'matrix_values')
        return v

    def to_scipy(self):
        import scipy.sparse as sp
        rp, col = self.pattern()
        return sp.csr_matrix((self.values(), col, rp), shape=(self.n, self.n))

    def __del__(self):
        try:
            self.ctx.lib.ocmp_matrix_destroy(self.h)
        except Exception:
            pass


class LinearForm:
    def __init__(self, fes: FESpace, terms: Optional[list] = None):
        self.space = fes
        self.terms = list(terms or [])
        self.ctx = fes.ctx
        self.vec = self.ctx.zeros(fes.ndof)

    def __iadd__(self, term):
        self.terms.append(term)
        return self

    def Assemble(self):
        ft = _FieldTable()
        chunks = [self.terms[i:i + 8] for i in range(0, max(1, len(self.terms)), 8)]
        for ci, chunk in enumerate(chunks):
            arr = (L.LTerm * max(1, len(chunk)))()
            for i, t in enumerate(chunk):
                lt = L.LTerm()
                lt.field = -1
                lt.weight_field = -1
                if isinstance(t, Source):
                    lt.kind = L.LFI_SOURCE
                    lt.test = t.test
                    lt.coef = t.coef
                    if t.field_gf is not None:
                        gf, blk = t.field_gf
                        lt.fn = L.FN_FIELD
                        lt.field = ft.index(gf, blk)
                        lt.c[0] = 1.0
                    else:
                        fid, prm = resolve_fn(t.fn)
                        lt.fn = fid
                        for k, v in enumerate(prm):
                            lt.c[k] = v
                    if t.weight_gf is not None:
                        gf, blk, mode = t.weight_gf
                        lt.weight_field = ft.index(gf, blk)
                        lt.weight_mode = mode
                elif isinstance(t, BndSource):
                    lt.kind = L.LFI_BND_SOURCE
                    lt.test = t.test
                    lt.coef = t.coef
                    lt.dirichlet_mask = t.markers
                    fid, prm = resolve_fn(t.fn)
                    lt.fn = fid
                    for k, v in enumerate(prm):
                        lt.c[k] = v
                else:
                    lt.kind = L.LFI_DG_DIRICHLET
                    lt.test = t.test
                    lt.coef = t.coef
                    lt.eps = t.eps
                    lt.alpha = t.alpha
                    lt.wind[0], lt.wind[1] = t.wind[0], t.wind[1]
                    lt.dirichlet_mask = t.dirichlet
                    fid, prm = resolve_fn(t.fn)
                    lt.fn = fid
                    for k, v in enumerate(prm):
                        lt.c[k] = v
                arr[i] = lt
            farr, nf = ft.array()
            L.check(self.ctx.lib.ocmp_assemble_linear(self.space.h, C.c_void_p(_ptr(self.vec)), len(chunk), arr,
                                                      nf, farr, 1 if ci else 0), 'assemble_linear')
        return self


class Preconditioner:
    """ngs.Preconditioner(a, 'local') on the GPU: point Jacobi (H1) or element-block Jacobi (L2)."""

    KINDS = {'local': None, 'jacobi': L.PRE_JACOBI, 'block_jacobi': L.PRE_BLOCK_JACOBI, 'none': L.PRE_NONE}

    def __init__(self, a: BilinearForm, kind: str = 'local', precision: str = 'fp64',
                 schur_blocks: Sequence[int] = (), schur_iters: int = 0, schur_tol: float = 0.0,
                 schur_amg: int = 0):
        """kind 'schur': point Jacobi on all rows except those of `schur_blocks` (the
        pressure blocks of a saddle-point system), which get the diagonal of the
        Schur complement (ocmp_precond_create_schur); with `schur_iters` > 0 FGMRES
        solves those rows by at most that many PCG iterations on C + B diag(A)^{-1} B^T
        instead, fewer once the preconditioned residual has fallen by `schur_tol`
        (ocmp_precond_set_schur_solve); with `schur_amg` > 0 FGMRES applies that many smoothed-
        aggregation AMG V-cycles on the same S instead (ocmp_precond_set_schur_amg)."""
        self.a = a
        self.precision = L.MIXED if precision == 'mixed' else L.FP64
        h = C.c_void_p()
        if kind == 'schur':
            mask = 0
            for b in schur_blocks:
                mask |= 1 << int(b)
            self.kind = L.PRE_JACOBI
            L.check(a.ctx.lib.ocmp_precond_create_schur(a.h, self.precision, C.c_uint64(mask), C.byref(h)),
                    'precond_create_schur')
            if schur_iters:
                L.check(a.ctx.lib.ocmp_precond_set_schur_solve(h, int(schur_iters), float(schur_tol)),
                        'precond_set_schur_solve')
            if schur_amg:
                L.check(a.ctx.lib.ocmp_precond_set_schur_amg(h, int(schur_amg)), 'precond_set_schur_amg')
        else:
            k = self.KINDS[kind]
            if k is None:
                k = L.PRE_BLOCK_JACOBI if a.is_dg else L.PRE_JACOBI
            self.kind = k
            L.check(a.ctx.lib.ocmp_precond_create(a.h, k, self.precision, C.byref(h)), 'precond_create')
        self.h = h

    def set_block_triangular(self, schur_diag: torch.Tensor, vel_iters: int, vel_tol: float) -> None:
        """Block-triangular step for FGMRES (ocmp_precond_set_block_triangular; 'schur' kind with the
        Schur blocks last): z_p = -v_p / schur_diag on the Schur rows, then z_u from at most
        `vel_iters` Jacobi-PCG iterations on the velocity block for v_u - A_up z_p, stopped once the
        preconditioned residual has fallen by `vel_tol`. `schur_diag`: a device tensor over the Schur
        rows, a positive diagonal approximation of the Schur complement (0 iters: off)."""
        d = schur_diag.to(self.a.ctx.torch_device, torch.float64).contiguous()
        L.check(self.a.ctx.lib.ocmp_precond_set_block_triangular(self.h, C.c_void_p(_ptr(d)), int(vel_iters),
                                                                 float(vel_tol)), 'precond_set_block_triangular')

    def set_schur_triangular(self, on: bool = True) -> None:
        """Upper block-triangular velocity rows after the flexible Schur step,
        z_u = diag(A)^-1 (v_u - A_up z_p) (ocmp_precond_set_schur_triangular); the next Update gathers A_up."""
        L.check(self.a.ctx.lib.ocmp_precond_set_schur_triangular(self.h, 1 if on else 0), 'precond_set_schur_triangular')

    def set_bf16(self, on: bool = True) -> None:
        """Keep the inner solves' matrices (Schur S, the block-triangular velocity block) in BF16 for the
        mixed path (ocmp_precond_set_bf16); the next Update makes the copies."""
        L.check(self.a.ctx.lib.ocmp_precond_set_bf16(self.h, 1 if on else 0), 'precond_set_bf16')

    def Update(self):
        L.check(self.a.ctx.lib.ocmp_precond_update(self.h), 'precond_update')

    def UpdateJacobi(self):
        """The point-Jacobi rows only; the Schur step keeps its last full Update (ocmp_precond_update_jacobi)."""
        L.check(self.a.ctx.lib.ocmp_precond_update_jacobi(self.h), 'precond_update_jacobi')

    def amg_levels(self) -> list:
        """(rows, nonzeros) of each level of the Schur AMG hierarchy the last Update built (empty if none)."""
        rows, nnz, nl = (C.c_int64 * 16)(), (C.c_int64 * 16)(), C.c_int()
        L.check(self.a.ctx.lib.ocmp_precond_schur_amg_levels(self.h, rows, nnz, 16, C.byref(nl)),
                'precond_schur_amg_levels')
        return [(int(rows[l]), int(nnz[l])) for l in range(min(nl.value, 16))]

    def __del__(self):
        try:
            self.a.ctx.lib.ocmp_precond_destroy(self.h)
        except Exception:
            pass


@dataclass
class SolveResult:
    iters: int
    outer: int
    converged: bool
    rel_residual: float
    initial_rel_residual: float
    fallback: bool = False        # FGMRES: the learned correction was dropped for the plain preconditioner
    precond_iters: int = 0        # FGMRES with a Schur inner solve: its PCG iterations over all steps


def solve(a: BilinearForm, b: torch.Tensor, x: torch.Tensor, pre: Optional[Preconditioner], solver: str = 'CG',
          tol: float = 1e-10, maxit: int = 100000, precision: str = 'fp64', restart: int = 50,
          inner_tol: float = 0.0, check_every: int = 0, max_outer: int = 0, s_step: int = 0) -> SolveResult:
    o = L.SolveOpts()
    o.solver = {'CG': L.SOLVER_CG, 'GMRes': L.SOLVER_GMRES, 'Richardson': L.SOLVER_RICHARDSON,
                'FGMRes': L.SOLVER_FGMRES, 'MinRes': L.SOLVER_MINRES}[solver]
    o.precision = L.MIXED if precision == 'mixed' else L.FP64
    o.restart = restart
    o.maxit = maxit
    o.tol = tol
    o.inner_tol = inner_tol
    o.check_every = check_every
    o.max_outer = max_outer
    o.s_step = s_step
    s = L.SolveStats()
    L.check(a.ctx.lib.ocmp_solve(a.h, pre.h if pre is not None else None, C.c_void_p(_ptr(b)),
                                 C.c_void_p(_ptr(x)), C.byref(o), C.byref(s)), 'solve')
    return SolveResult(s.iters, s.outer, bool(s.converged), s.rel_residual, s.initial_rel_residual,
                       bool(s.fallback), int(s.precond_iters))
