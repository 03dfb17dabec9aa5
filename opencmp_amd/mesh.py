"""Simplex meshes for the assemble+solve path.

A mesh here is plain host arrays: vertex coordinates, element->vertex table,
boundary facets with an integer marker and the marker names. This is what the
C-ABI's ``ocmp_mesh_create`` takes; the library derives edges, faces, facets
and their element adjacency itself.

The structured generators produce the meshes the workloads are defined on
(the reference has no generator for them: ``get_Netgen_nonconformal``
quarters squares around a centre point, opencmp/diffuse_interface/
mesh_helpers.py:547-565, and emits pyramids in 3D, :609-633). Boundary names
follow the reference's structured meshes: ``bottom/right/top/left`` in 2D
(opencmp/diffuse_interface/mesh_helpers.py:533-537) and additionally
``front/back`` in 3D (:681-683).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List

import numpy as np


@dataclass
class Mesh:
    coords: np.ndarray                  # (nv, dim) float64
    elems: np.ndarray                   # (nel, dim+1) int32, vertex ids
    bnd_facets: np.ndarray              # (nbnd, dim) int32, vertex ids
    bnd_markers: np.ndarray             # (nbnd,) int32, index into bnd_names
    bnd_names: List[str] = field(default_factory=list)

    @property
    def dim(self) -> int:
        return int(self.coords.shape[1])

    @property
    def nv(self) -> int:
        return int(self.coords.shape[0])

    @property
    def nel(self) -> int:
        return int(self.elems.shape[0])

    def marker_mask(self, names) -> int:
        """Bit mask over marker indices for a list (or '|'-joined string) of boundary names."""
        if isinstance(names, str):
            names = [s for s in names.replace(',', '|').split('|') if s]
        mask = 0
        for nm in names:
            if nm not in self.bnd_names:
                raise ValueError(f'unknown boundary name {nm!r}; mesh has {self.bnd_names}')
            mask |= 1 << self.bnd_names.index(nm)
        return mask

    def validate(self) -> None:
        d = self.dim
        if self.elems.ndim != 2 or self.elems.shape[1] != d + 1:
            raise ValueError('elems must be (nel, dim+1)')
        if self.elems.min() < 0 or self.elems.max() >= self.nv:
            raise ValueError('element vertex index out of range')
        if self.bnd_facets.shape[0] and (self.bnd_facets.min() < 0 or self.bnd_facets.max() >= self.nv):
            raise ValueError('boundary facet vertex index out of range')
        if self.bnd_markers.shape[0] != self.bnd_facets.shape[0]:
            raise ValueError('one marker per boundary facet')
        if self.bnd_markers.shape[0] and self.bnd_markers.max() >= max(1, len(self.bnd_names)):
            raise ValueError('boundary marker without a name')
        if len(self.bnd_names) > 64:
            raise ValueError('at most 64 boundary markers')


def unit_square(n: int) -> Mesh:
    """n x n squares on [0,1]^2, each split along its (0,0)->(1,1) diagonal into 2 triangles.

    Vertex (i, j) (x index i, y index j) has id j*(n+1)+i, the row-major order of
    opencmp/diffuse_interface/mesh_helpers.py:520-528.
    """
    h = 1.0 / n
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing='xy')
    coords = np.stack([i.ravel() * h, j.ravel() * h], axis=1).astype(np.float64)
    # exact end points
    coords[np.isclose(coords, 1.0)] = 1.0
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing='xy')
    ci = ci.ravel(); cj = cj.ravel()
    p1 = cj * (n + 1) + ci
    p2 = p1 + 1
    p3 = p1 + n + 2
    p4 = p1 + n + 1
    t1 = np.stack([p1, p2, p3], 1)
    t2 = np.stack([p1, p3, p4], 1)
    elems = np.empty((2 * n * n, 3), dtype=np.int32)
    elems[0::2] = t1
    elems[1::2] = t2
    k = np.arange(n)
    bottom = np.stack([k, k + 1], 1)
    right = np.stack([k * (n + 1) + n, (k + 1) * (n + 1) + n], 1)
    top = np.stack([n * (n + 1) + k, n * (n + 1) + k + 1], 1)
    left = np.stack([k * (n + 1), (k + 1) * (n + 1)], 1)
    bnd = np.concatenate([bottom, right, top, left]).astype(np.int32)
    markers = np.repeat(np.arange(4, dtype=np.int32), n)
    return Mesh(coords, elems, bnd, markers, ['bottom', 'right', 'top', 'left'])


# The 6 Kuhn tetrahedra of the unit cube: paths 000 -> 111 along the axis
# permutations. Corner c = (bx, by, bz) has local index bx + 2*by + 4*bz.
_KUHN = []
for _perm in ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)):
    _c = [0, 0, 0]
    _path = [0]
    for _ax in _perm:
        _c[_ax] = 1
        _path.append(_c[0] + 2 * _c[1] + 4 * _c[2])
    _KUHN.append(_path)
_KUHN = np.array(_KUHN, dtype=np.int64)


def unit_cube(n: int) -> Mesh:
    """n^3 cubes on [0,1]^3, each split into the 6 Kuhn tetrahedra around its 000-111 diagonal.

    Vertex (i, j, k) has id i + (n+1)*(j + (n+1)*k). Boundary names: left/right
    are x=0/1, front/back are y=0/1, bottom/top are z=0/1.
    """
    m = n + 1
    h = 1.0 / n
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing='ij')
    coords = np.stack([i.ravel() * h, j.ravel() * h, k.ravel() * h], axis=1).astype(np.float64)
    coords[np.isclose(coords, 1.0)] = 1.0
    ck, cj, ci = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij')
    base = (ci + m * (cj + m * ck)).ravel()
    corner_off = np.array([bx + m * (by + m * bz) for bz in (0, 1) for by in (0, 1) for bx in (0, 1)],
                          dtype=np.int64)
    # corner_off index = bx + 2*by + 4*bz matches _KUHN's corner numbering
    elems = (base[:, None, None] + corner_off[_KUHN][None, :, :]).reshape(-1, 4).astype(np.int32)

    a, b = np.meshgrid(np.arange(n), np.arange(n), indexing='ij')
    a = a.ravel(); b = b.ravel()

    def quad_tris(v00, v10, v01, v11):
        # split each boundary square along the diagonal v00-v11, consistent with the Kuhn split
        return np.concatenate([np.stack([v00, v10, v11], 1), np.stack([v00, v01, v11], 1)])

    def vid(ii, jj, kk):
        return ii + m * (jj + m * kk)

    faces = []
    marks = []
    names = ['left', 'right', 'front', 'back', 'bottom', 'top']
    for mk, (axis, val) in enumerate([(0, 0), (0, n), (1, 0), (1, n), (2, 0), (2, n)]):
        if axis == 0:
            f = quad_tris(vid(val, a, b), vid(val, a + 1, b), vid(val, a, b + 1), vid(val, a + 1, b + 1))
        elif axis == 1:
            f = quad_tris(vid(a, val, b), vid(a + 1, val, b), vid(a, val, b + 1), vid(a + 1, val, b + 1))
        else:
            f = quad_tris(vid(a, b, val), vid(a + 1, b, val), vid(a, b + 1, val), vid(a + 1, b + 1, val))
        faces.append(f)
        marks.append(np.full(f.shape[0], mk, dtype=np.int32))
    bnd = np.concatenate(faces).astype(np.int32)
    return Mesh(coords, elems, bnd, np.concatenate(marks), names)
