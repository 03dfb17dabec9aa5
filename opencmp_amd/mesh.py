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


def _hole_centres(n_holes, r_range, h, seed, length, height, x_margin=0.25, max_tries=100000):
    """Radii r ~ U[r_range] and centres rejection-sampled from default_rng(seed):
    clear of the walls by 2h and of each other by a gap of at least 2h."""
    rng = np.random.default_rng(seed)
    holes = []
    tries = 0
    while len(holes) < n_holes:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('could not place the holes; lower n_holes or the radii')
        r = rng.uniform(*r_range)
        cx = rng.uniform(x_margin + r, length - x_margin - r)
        cy = rng.uniform(r + 2 * h, height - r - 2 * h)
        if all(np.hypot(cx - x, cy - y) >= r + q + 2 * h for x, y, q in holes):
            holes.append((cx, cy, r))
    return holes


def channel_with_holes(n: int = 228, n_holes: int = 8, r_range=(0.05, 0.12), seed: int = 0,
                       length: float = 4.0, height: float = 1.0) -> Mesh:
    """Channel [0,length] x [0,height] with circular holes, body-fitted by snapping.

    A structured grid of (length*n) x (height*n) squares split along their
    (0,0)->(1,1) diagonals; triangles whose centroid lies in a hole are removed,
    and every vertex on a hole's staircase boundary is moved radially onto the
    circle. Boundary names: ``inlet`` (x=0), ``outlet`` (x=length), ``wall``
    (y=0 and y=height) and ``cyl`` (the holes), the names of the channel meshes
    of the reference's INS examples. n=228 gives about 4e5 triangles.
    """
    h = 1.0 / n
    nx, ny = int(round(length * n)), int(round(height * n))
    holes = _hole_centres(n_holes, r_range, h, seed, length, height)
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing='xy')
    coords = np.stack([i.ravel() * (length / nx), j.ravel() * (height / ny)], axis=1).astype(np.float64)
    coords[:, 0][np.isclose(coords[:, 0], length)] = length
    coords[:, 1][np.isclose(coords[:, 1], height)] = height
    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing='xy')
    ci = ci.ravel(); cj = cj.ravel()
    p1 = cj * (nx + 1) + ci
    t1 = np.stack([p1, p1 + 1, p1 + nx + 2], 1)
    t2 = np.stack([p1, p1 + nx + 2, p1 + nx + 1], 1)
    elems = np.empty((2 * nx * ny, 3), dtype=np.int64)
    elems[0::2] = t1
    elems[1::2] = t2
    cen = coords[elems].mean(1)
    keep = np.ones(len(elems), dtype=bool)
    for cx, cy, r in holes:
        keep &= np.hypot(cen[:, 0] - cx, cen[:, 1] - cy) >= r
    elems = elems[keep]
    # boundary edges of what is left, and the hole ones among them
    e = np.concatenate([elems[:, [0, 1]], elems[:, [1, 2]], elems[:, [2, 0]]])
    es = np.sort(e, 1)
    key = es[:, 0] * (nx + 1) * (ny + 1) + es[:, 1]
    uk, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    bmask = cnt[inv] == 1
    bedges = e[bmask]                      # oriented as in their element (counter-clockwise)
    mid = coords[bedges].mean(1)
    outer = (np.isclose(mid[:, 0], 0.0) | np.isclose(mid[:, 0], length) | np.isclose(mid[:, 1], 0.0) |
             np.isclose(mid[:, 1], height))
    # Fit the holes: hole boundary vertices move radially onto their circle and
    # the vertices near a hole are relaxed towards the mean of their neighbours.
    # Every move is accepted only if no triangle's area drops below a quarter of
    # the grid's, so no triangle inverts.
    hv = np.unique(bedges[~outer])
    deg = np.bincount(bedges[~outer].ravel(), minlength=len(coords))
    if np.any(deg[hv] != 2):
        raise RuntimeError('a hole boundary touches itself at a vertex; change n or the seed')
    owner = np.zeros(len(coords), dtype=np.int64)
    best = np.full(len(coords), np.inf)
    for k, (cx, cy, r) in enumerate(holes):
        dd = np.abs(np.hypot(coords[:, 0] - cx, coords[:, 1] - cy) - r)
        upd = dd < best
        best[upd] = dd[upd]
        owner[upd] = k
    hc = np.array([[cx, cy] for cx, cy, _ in holes])
    hr = np.array([r for _, _, r in holes])
    near = best < 6 * h
    near[np.unique(bedges[outer])] = False
    used = np.zeros(len(coords), dtype=bool)
    used[elems.ravel()] = True
    inner = np.nonzero(near & used)[0]
    inner = np.setdiff1d(inner, hv)
    ue = np.unique(np.sort(e, 1), axis=0)
    ue = ue[near[ue[:, 0]] | near[ue[:, 1]]]
    loc = np.nonzero(np.isin(elems, np.nonzero(near)[0]).any(1))[0]
    min_area = 0.25 * 0.5 * (length / nx) * (height / ny)

    def tri_area(c, el):
        P = c[el]
        return 0.5 * ((P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) -
                      (P[:, 2, 0] - P[:, 0, 0]) * (P[:, 1, 1] - P[:, 0, 1]))

    def safe_move(c, verts, target):
        new = c.copy()
        new[verts] = target
        moved = np.zeros(len(c), dtype=bool)
        moved[verts] = True
        el = elems[loc]
        while True:
            bad = tri_area(new, el) < min_area
            if not bad.any():
                return new
            rv = np.unique(el[bad])
            rv = rv[moved[rv]]
            new[rv] = c[rv]
            moved[rv] = False

    for _ in range(40):
        d = coords[hv] - hc[owner[hv]]
        coords = safe_move(coords, hv, hc[owner[hv]] + d * (hr[owner[hv]] / np.hypot(d[:, 0], d[:, 1]))[:, None])
        acc = np.zeros_like(coords)
        cnt2 = np.zeros(len(coords))
        np.add.at(acc, ue[:, 0], coords[ue[:, 1]])
        np.add.at(acc, ue[:, 1], coords[ue[:, 0]])
        np.add.at(cnt2, ue[:, 0], 1.0)
        np.add.at(cnt2, ue[:, 1], 1.0)
        coords = safe_move(coords, inner, acc[inner] / cnt2[inner][:, None])
    d = coords[hv] - hc[owner[hv]]
    fit_err = np.abs(np.hypot(d[:, 0], d[:, 1]) - hr[owner[hv]]).max() if len(hv) else 0.0
    # renumber the vertices still in use
    used = np.unique(elems)
    remap = -np.ones(len(coords), dtype=np.int64)
    remap[used] = np.arange(len(used))
    coords = coords[used]
    elems = remap[elems].astype(np.int32)
    bedges = remap[bedges].astype(np.int32)
    names = ['inlet', 'outlet', 'wall', 'cyl']
    markers = np.full(len(bedges), 3, dtype=np.int32)
    markers[outer & np.isclose(mid[:, 0], 0.0)] = 0
    markers[outer & np.isclose(mid[:, 0], length)] = 1
    markers[outer & (np.isclose(mid[:, 1], 0.0) | np.isclose(mid[:, 1], height))] = 2
    m = Mesh(coords, elems, bedges, markers, names)
    m.holes = holes
    m.hole_fit_error = float(fit_err)
    return m


def sphere_obstacles(n_spheres: int = 20, r_range=(0.03, 0.1), seed: int = 0, lo: float = 0.1, hi: float = 0.9):
    """Spheres with r ~ U[r_range] and centres ~ U[lo, hi]^3 drawn from default_rng(seed)
    (radius first, then the centre, sphere by sphere)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_spheres):
        r = rng.uniform(*r_range)
        c = rng.uniform(lo, hi, 3)
        out.append((c, r))
    return out


def obstacle_phi(coords: np.ndarray, spheres, h: float) -> np.ndarray:
    """Diffuse-interface indicator phi = 0.5 (1 + tanh(d / (2h))) with d the signed distance
    to the union of the spheres, positive in the fluid: phi -> 1 in the fluid and 0 inside
    the obstacles (the convention of opencmp/models/stokes_dim.py:42-48, where the
    penalty acts on 1 - phi)."""
    d = np.full(coords.shape[0], np.inf)
    for c, r in spheres:
        d = np.minimum(d, np.linalg.norm(coords - c[None, :], axis=1) - r)
    return 0.5 * (1.0 + np.tanh(d / (2.0 * h)))
