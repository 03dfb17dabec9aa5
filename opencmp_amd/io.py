"""Mesh file readers: Netgen ``.vol`` and Gmsh v2.2 ``.msh`` (simplex meshes).

Replaces the mesh half of ``load_mesh`` (opencmp/helpers/io.py:69-110), which
hands the file to Netgen (``ngmsh.Mesh().Load``) or ``ReadGmsh``. Boundary
names come from the file (``bcnames`` / ``$PhysicalNames``); boundary facet
markers index into that list.
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh


def _sections(lines):
    """Netgen .vol: map section keyword -> line index of the keyword."""
    out = {}
    for i, ln in enumerate(lines):
        s = ln.strip()
        if s and not s.startswith('#') and s.replace('_', '').isalnum() and s[0].isalpha():
            out.setdefault(s, i)
    return out


def read_vol(path: str) -> Mesh:
    with open(path) as f:
        lines = f.read().splitlines()
    sec = _sections(lines)
    dim = int(lines[sec['dimension'] + 1].split()[0])

    def block(name):
        if name not in sec:
            return []
        i = sec[name] + 1
        while not lines[i].strip() or lines[i].strip().startswith('#'):
            i += 1
        n = int(lines[i].split()[0])
        rows = []
        i += 1
        while len(rows) < n:
            s = lines[i].strip()
            if s and not s.startswith('#'):
                rows.append(s.split())
            i += 1
        return rows

    pts = np.array([[float(v) for v in r[:3]] for r in block('points')], dtype=np.float64)
    names = {}
    for r in block('bcnames'):
        names[int(r[0])] = r[1]
    if dim == 2:
        surf = block('surfaceelements')
        elems = []
        for r in surf:
            npn = int(r[4])
            if npn != 3:
                raise NotImplementedError('only triangular 2D .vol meshes are supported')
            elems.append([int(v) - 1 for v in r[5:8]])
        segs = block('edgesegmentsgi2') or block('edgesegmentsgi') or block('edgesegments')
        bnd, mk = [], []
        for r in segs:
            bnd.append([int(r[2]) - 1, int(r[3]) - 1])
            mk.append(int(r[0]))
        coords = pts[:, :2]
    else:
        vol = block('volumeelements')
        elems = []
        for r in vol:
            npn = int(r[1])
            if npn != 4:
                raise NotImplementedError('only tetrahedral 3D .vol meshes are supported')
            elems.append([int(v) - 1 for v in r[2:6]])
        bnd, mk = [], []
        for r in block('surfaceelements'):
            if int(r[4]) != 3:
                raise NotImplementedError('only triangular boundary faces are supported')
            bnd.append([int(v) - 1 for v in r[5:8]])
            mk.append(int(r[1]))
        coords = pts
    keys = sorted(set(mk) | set(names))
    # several bc numbers may carry one name: they form one marker
    bnames = []
    kidx = {}
    for k in keys:
        nm = names.get(k, f'default{k}')
        if nm not in bnames:
            bnames.append(nm)
        kidx[k] = bnames.index(nm)
    return Mesh(np.ascontiguousarray(coords), np.array(elems, dtype=np.int32),
                np.array(bnd, dtype=np.int32).reshape(-1, dim), np.array([kidx[k] for k in mk], dtype=np.int32),
                bnames)


def read_msh(path: str) -> Mesh:
    """Gmsh 2.2 ASCII: triangles (type 2) or tetrahedra (type 4); boundary lines (1) / triangles (2)."""
    with open(path) as f:
        lines = [ln.strip() for ln in f.read().splitlines()]
    i = lines.index('$PhysicalNames')
    nphys = int(lines[i + 1])
    phys = {}
    for r in lines[i + 2:i + 2 + nphys]:
        d, tag, name = r.split(maxsplit=2)
        phys[int(tag)] = (int(d), name.strip('"'))
    i = lines.index('$Nodes')
    nn = int(lines[i + 1])
    ids, pts = [], []
    for r in lines[i + 2:i + 2 + nn]:
        v = r.split()
        ids.append(int(v[0]))
        pts.append([float(x) for x in v[1:4]])
    idmap = {k: j for j, k in enumerate(ids)}
    pts = np.array(pts)
    i = lines.index('$Elements')
    ne = int(lines[i + 1])
    tris, tets, lines2, ltag, ttag = [], [], [], [], []
    for r in lines[i + 2:i + 2 + ne]:
        v = [int(x) for x in r.split()]
        typ, ntag = v[1], v[2]
        tags = v[3:3 + ntag]
        nodes = [idmap[x] for x in v[3 + ntag:]]
        if typ == 1:
            lines2.append(nodes)
            ltag.append(tags[0])
        elif typ == 2:
            tris.append(nodes)
            ttag.append(tags[0])
        elif typ == 4:
            tets.append(nodes)
    if tets:
        dim = 3
        elems = tets
        bnd, mk = tris, ttag
        coords = pts
    else:
        dim = 2
        elems = tris
        bnd, mk = lines2, ltag
        coords = pts[:, :2]
    keys = sorted({t for t in mk})
    names = [phys.get(k, (dim - 1, f'default{k}'))[1] for k in keys]
    kidx = {k: j for j, k in enumerate(keys)}
    # drop unreferenced nodes (gmsh keeps geometry points)
    used = np.unique(np.array(elems).ravel())
    remap = -np.ones(len(coords), dtype=np.int64)
    remap[used] = np.arange(len(used))
    elems = remap[np.array(elems)]
    bnd = remap[np.array(bnd)].reshape(-1, dim)
    return Mesh(np.ascontiguousarray(coords[used]), elems.astype(np.int32), bnd.astype(np.int32),
                np.array([kidx[t] for t in mk], dtype=np.int32), names)


def load_mesh(path: str) -> Mesh:
    if path.endswith('.vol'):
        return read_vol(path)
    if path.endswith('.msh'):
        return read_msh(path)
    raise TypeError('Only .vol (Netgen) and .msh (GMSH) meshes can be used.')


# ------------------------------------------------------------------ solution files

def save_sol(path: str, vec) -> None:
    """GridFunction.Save (opencmp/helpers/saving.py:74-93): the coefficient vector as raw native
    float64 values, no header, in dof order (vertex, edge, face/cell dofs per component block; NGSolve's
    node order, which the dof numbering follows). `vec`: a device tensor or an array."""
    a = vec.detach().cpu().numpy() if hasattr(vec, 'detach') else np.asarray(vec)
    np.ascontiguousarray(a, dtype=np.float64).tofile(path)


def load_sol(path: str, ndof: int) -> np.ndarray:
    """GridFunction.Load: the raw float64 coefficient vector of a space with `ndof` dofs."""
    a = np.fromfile(path, dtype=np.float64)
    if a.size != ndof:
        raise ValueError(f'"{path}" holds {a.size} values; the space has {ndof} dofs')
    return a
