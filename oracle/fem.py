"""FP64 CPU oracle for the OpenCMP assemble+solve path.

TEST INFRASTRUCTURE. This module is the checker for the HIP path: only the
test-suite, the smoke check and the CPU-baseline leg of the benchmark may
import it. It is never the thing measured or shipped, and the product package
never calls it.

It restates, in plain numpy/scipy, what NGSolve computes for the forms the
OpenCMP models build (the reference delegates all arithmetic to NGSolve, an
unpinned third-party C++ library absent from this image):

* simplex topology, H1 (hierarchical, vertex-first numbering) and L2
  (orthonormal Dubiner, element-blocked) spaces, compound spaces blocked per
  component -- the layout of ``ngs.FESpace([...])`` used by
  opencmp/models/poisson.py:60-64 and opencmp/models/ins.py:95-128;
* the full-element-clique CSR pattern (plus facet neighbours for
  ``dgjumps=True``, opencmp/models/multi_component_ins.py:79-102);
* element integrals by plain Gauss quadrature on every element (the HIP path
  uses precomputed reference tensors instead, so the two are independent
  computations);
* DG interior-penalty and upwind facet terms of
  opencmp/models/multi_component_ins.py:173-197 and :287-301,
  opencmp/models/poisson.py:81-121, with alpha = ipc*p^2/h
  (opencmp/models/base_model.py:150, opencmp/helpers/ngsolve_.py:65-67);
* free-DOF linear solves (opencmp/models/base_model.py:886-947) and L2 error
  norms (opencmp/helpers/error.py:54-128).

Parity status per function is listed in ``PARITY`` at the bottom.
"""
from __future__ import annotations

from math import factorial

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# =============================================================== quadrature


def _gl01(n):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def line_rule(degree):
    """Gauss-Legendre on [0,1], exact to ``degree``; weights sum to 1."""
    return _gl01(max(1, (degree + 2) // 2))


def simplex_rule(dim, degree):
    """Collapsed (Duffy) Gauss rule on the reference simplex, exact to ``degree``.

    Returns barycentric points (nq, dim+1) and weights summing to 1/dim!.
    """
    if dim == 1:
        t, w = line_rule(degree)
        return np.stack([1 - t, t], 1), w
    if dim == 2:
        n = max(1, (degree + 3) // 2)
        u, wu = _gl01(n)
        U, V = np.meshgrid(u, u, indexing='ij')
        WU, WV = np.meshgrid(wu, wu, indexing='ij')
        x = U.ravel()
        y = ((1 - U) * V).ravel()
        w = (WU * WV * (1 - U)).ravel()
        return np.stack([1 - x - y, x, y], 1), w
    if dim == 3:
        n = max(1, (degree + 4) // 2)
        u, wu = _gl01(n)
        U, V, W = np.meshgrid(u, u, u, indexing='ij')
        WU, WV, WW = np.meshgrid(wu, wu, wu, indexing='ij')
        x = U.ravel()
        y = ((1 - U) * V).ravel()
        z = ((1 - U) * (1 - V) * W).ravel()
        w = (WU * WV * WW * (1 - U) ** 2 * (1 - V)).ravel()
        return np.stack([1 - x - y - z, x, y, z], 1), w
    raise ValueError(dim)


def simplex_monomial_integral(exps):
    """Closed form  int_ref prod lambda_i^a_i = prod a_i! * d! / (sum a_i + d)!  * (1/d!)."""
    d = len(exps) - 1
    num = 1
    for a in exps:
        num *= factorial(a)
    return num / factorial(sum(exps) + d)


# =============================================================== local layout

TRI_EDGES = [(0, 1), (0, 2), (1, 2)]
TET_EDGES = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
TET_FACES = [(1, 2, 3), (0, 2, 3), (0, 1, 3), (0, 1, 2)]      # face k opposite vertex k
TRI_FACETS = [(1, 2), (0, 2), (0, 1)]                         # facet k opposite vertex k


def local_edges(dim):
    return TRI_EDGES if dim == 2 else TET_EDGES


def local_facets(dim):
    return TRI_FACETS if dim == 2 else TET_FACES


def h1_ndof_local(dim, p):
    nv = dim + 1
    ne = 3 if dim == 2 else 6
    n = nv + ne * (p - 1)
    if p >= 3:
        n += 1 if dim == 2 else 4
    if p >= 4:
        raise NotImplementedError('H1 order <= 3')
    return n


def h1_eval(dim, p, lam):
    """H1 hierarchical basis on the reference simplex (local vertices sorted by global id).

    vertex i: lambda_i; edge (a,b), degree k: lambda_a lambda_b (lambda_b-lambda_a)^(k-2);
    face (a,b,c) (order 3): lambda_a lambda_b lambda_c.
    Returns phi (nq, nloc) and dphi/dlambda (nq, nloc, dim+1).
    """
    nq = lam.shape[0]
    nl = h1_ndof_local(dim, p)
    phi = np.zeros((nq, nl))
    dl = np.zeros((nq, nl, dim + 1))
    c = 0
    for i in range(dim + 1):
        phi[:, c] = lam[:, i]
        dl[:, c, i] = 1.0
        c += 1
    for (a, b) in local_edges(dim):
        la, lb = lam[:, a], lam[:, b]
        for k in range(2, p + 1):
            if k == 2:
                phi[:, c] = la * lb
                dl[:, c, a] = lb
                dl[:, c, b] = la
            else:  # k == 3
                phi[:, c] = la * lb * (lb - la)
                dl[:, c, a] = lb * lb - 2 * la * lb
                dl[:, c, b] = 2 * la * lb - la * la
            c += 1
    if p >= 3:
        faces = [(0, 1, 2)] if dim == 2 else TET_FACES
        for (a, b, cc) in faces:
            la, lb, lc = lam[:, a], lam[:, b], lam[:, cc]
            phi[:, c] = la * lb * lc
            dl[:, c, a] = lb * lc
            dl[:, c, b] = la * lc
            dl[:, c, cc] = la * lb
            c += 1
    assert c == nl
    return phi, dl


# ---- orthonormal Dubiner basis on the reference triangle, as 2D polynomials

def _p2mul(A, B):
    out = np.zeros((A.shape[0] + B.shape[0] - 1, A.shape[1] + B.shape[1] - 1))
    for i in range(A.shape[0]):
        for j in range(A.shape[1]):
            if A[i, j] != 0.0:
                out[i:i + B.shape[0], j:j + B.shape[1]] += A[i, j] * B
    return out


def _p2add(A, B):
    s = (max(A.shape[0], B.shape[0]), max(A.shape[1], B.shape[1]))
    out = np.zeros(s)
    out[:A.shape[0], :A.shape[1]] += A
    out[:B.shape[0], :B.shape[1]] += B
    return out


def _jacobi_coeffs_in_y(j, alpha):
    """P_j^{(alpha,0)}(2y-1) as a 2D poly in (x, y) via the three-term recurrence."""
    z = np.array([[-1.0, 2.0]])                       # 2y - 1, shape (1, 2): x^0 y^k
    P = [np.array([[1.0]])]
    if j >= 1:
        # P1 = ((alpha+2) z + alpha)/2
        P.append(_p2add(0.5 * (alpha + 2) * z, np.array([[0.5 * alpha]])))
    for n in range(1, j):
        a = alpha
        c1 = 2 * (n + 1) * (n + a + 1) * (2 * n + a)
        c2 = (2 * n + a + 1) * (a * a)
        c3 = (2 * n + a) * (2 * n + a + 1) * (2 * n + a + 2)
        c4 = 2 * (n + a) * n * (2 * n + a + 2)
        nxt = _p2add(_p2mul(c3 * z, P[n]) + 0 * 0, c2 * P[n])
        nxt = _p2add(nxt, -c4 * P[n - 1])
        P.append(nxt / c1)
    return P[j]


def dubiner_polys(p):
    """Orthonormal (int_ref psi_i psi_j = delta_ij) Dubiner basis of degree <= p.

    Order: by total degree, then by decreasing x-degree: (i, j) = (0,0),(1,0),(0,1),(2,0),(1,1),(0,2),...
    psi_ij = Q_i(2x+y-1, 1-y) * P_j^{(2i+1,0)}(2y-1), Q_i the scaled Legendre polynomial.
    Returns a list of coefficient arrays C[a, b] for x^a y^b.
    """
    t = np.array([[-1.0, 1.0], [2.0, 0.0]])            # 2x + y - 1
    s = np.array([[1.0, -1.0]])                         # 1 - y
    Q = [np.array([[1.0]]), t]
    for n in range(1, p):
        Q.append(_p2add((2 * n + 1) * _p2mul(t, Q[n]), -n * _p2mul(_p2mul(s, s), Q[n - 1])) / (n + 1))
    polys = []
    lam, w = simplex_rule(2, 2 * p + 2)
    for deg in range(p + 1):
        for i in range(deg, -1, -1):
            j = deg - i
            C = _p2mul(Q[i], _jacobi_coeffs_in_y(j, 2 * i + 1))
            v = poly2_eval(C, lam[:, 1], lam[:, 2])
            C = C / np.sqrt(np.sum(w * v * v))
            polys.append(C)
    return polys


def poly2_eval(C, x, y):
    out = np.zeros_like(x)
    for a in range(C.shape[0]):
        for b in range(C.shape[1]):
            if C[a, b] != 0.0:
                out = out + C[a, b] * x ** a * y ** b
    return out


def poly2_grad(C):
    Cx = np.zeros((max(1, C.shape[0] - 1), C.shape[1]))
    Cy = np.zeros((C.shape[0], max(1, C.shape[1] - 1)))
    for a in range(1, C.shape[0]):
        Cx[a - 1, :] = a * C[a, :]
    for b in range(1, C.shape[1]):
        Cy[:, b - 1] = b * C[:, b]
    return Cx, Cy


def l2_ndof_local(dim, p):
    if dim != 2:
        raise NotImplementedError('L2 spaces on triangles only')
    return (p + 1) * (p + 2) // 2


def l2_eval(dim, p, lam):
    """Dubiner basis values and reference gradients d/d(x^, y^) at barycentric points."""
    polys = dubiner_polys(p)
    x, y = lam[:, 1], lam[:, 2]
    phi = np.stack([poly2_eval(C, x, y) for C in polys], 1)
    g = []
    for C in polys:
        Cx, Cy = poly2_grad(C)
        g.append(np.stack([poly2_eval(Cx, x, y), poly2_eval(Cy, x, y)], 1))
    return phi, np.stack(g, 1)        # (nq, nl), (nq, nl, 2)


def h1_ref_grad(dim, p, lam):
    phi, dl = h1_eval(dim, p, lam)
    return phi, dl[:, :, 1:] - dl[:, :, :1]


def ref_basis(kind, dim, p, lam):
    """values (nq, nl) and reference gradients (nq, nl, dim)."""
    if kind == 'H1':
        return h1_ref_grad(dim, p, lam)
    return l2_eval(dim, p, lam)


# =============================================================== topology


class Topology:
    """Edges, faces and facets of a simplex mesh, with local vertices sorted by global id."""

    def __init__(self, mesh):
        self.mesh = mesh
        self.dim = d = mesh.dim
        self.coords = np.asarray(mesh.coords, dtype=np.float64)
        self.elems = np.sort(np.asarray(mesh.elems, dtype=np.int64), axis=1)
        self.nel = self.elems.shape[0]
        self.nv = self.coords.shape[0]
        E = self.elems
        # edges
        le = local_edges(d)
        pairs = np.concatenate([E[:, [a, b]] for (a, b) in le])             # (nedge_loc*nel, 2)
        key = pairs[:, 0] * self.nv + pairs[:, 1]
        ukey, inv = np.unique(key, return_inverse=True)
        self.edges = np.stack([ukey // self.nv, ukey % self.nv], 1)
        self.elem_edges = inv.reshape(len(le), self.nel).T                  # (nel, nle)
        self.nedge = self.edges.shape[0]
        # faces (3D)
        if d == 3:
            tr = np.concatenate([E[:, list(f)] for f in TET_FACES])
            key = (tr[:, 0] * self.nv + tr[:, 1]) * self.nv + tr[:, 2]
            ukey, inv = np.unique(key, return_inverse=True)
            self.faces = np.stack([ukey // (self.nv * self.nv), (ukey // self.nv) % self.nv, ukey % self.nv], 1)
            self.elem_faces = inv.reshape(4, self.nel).T
            self.nface = self.faces.shape[0]
            self.elem_facets = self.elem_faces
            self.facets = self.faces
        else:
            # facet k opposite vertex k: TRI_FACETS[k] = edge index in TRI_EDGES order (1,2)->2, (0,2)->1, (0,1)->0
            self.elem_facets = self.elem_edges[:, [2, 1, 0]]
            self.facets = self.edges
        nf = self.facets.shape[0]
        self.nfacet = nf
        # facet -> (elem, local facet) for both sides
        el = np.repeat(np.arange(self.nel), d + 1)
        lk = np.tile(np.arange(d + 1), self.nel)
        fid = self.elem_facets.ravel()
        order = np.lexsort((el, fid))
        fid_s, el_s, lk_s = fid[order], el[order], lk[order]
        first = np.ones(len(fid_s), bool)
        first[1:] = fid_s[1:] != fid_s[:-1]
        self.f_e0 = np.full(nf, -1)
        self.f_k0 = np.full(nf, -1)
        self.f_e1 = np.full(nf, -1)
        self.f_k1 = np.full(nf, -1)
        self.f_e0[fid_s[first]] = el_s[first]
        self.f_k0[fid_s[first]] = lk_s[first]
        sec = ~first
        self.f_e1[fid_s[sec]] = el_s[sec]
        self.f_k1[fid_s[sec]] = lk_s[sec]
        # boundary markers per facet
        self.f_marker = np.full(nf, -1)
        bf = np.sort(np.asarray(mesh.bnd_facets, dtype=np.int64), axis=1)
        if d == 2:
            bkey = bf[:, 0] * self.nv + bf[:, 1]
            fkey = self.facets[:, 0] * self.nv + self.facets[:, 1]
        else:
            bkey = (bf[:, 0] * self.nv + bf[:, 1]) * self.nv + bf[:, 2]
            fkey = (self.facets[:, 0] * self.nv + self.facets[:, 1]) * self.nv + self.facets[:, 2]
        pos = np.searchsorted(fkey, bkey)
        if np.any(pos >= nf) or np.any(fkey[np.minimum(pos, nf - 1)] != bkey):
            raise ValueError('boundary facet not in mesh')
        self.f_marker[pos] = np.asarray(mesh.bnd_markers)
        # geometry
        X = self.coords[E]                                                  # (nel, d+1, d)
        J = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))              # columns v_i - v_0
        self.J = J
        self.detJ = np.linalg.det(J)
        self.absdet = np.abs(self.detJ)
        self.Jinv = np.linalg.inv(J)
        self.h = self.absdet ** (1.0 / d)                                    # NGSolve-style element size
        # gradients of barycentrics: grad lambda_i (i>=1) = rows of J^{-1}; lambda_0 = -sum
        gl = self.Jinv                                                      # (nel, d, d): row i = grad lambda_{i+1}
        self.grad_lam = np.concatenate([-gl.sum(1, keepdims=True), gl], 1)  # (nel, d+1, d)

    def map_points(self, lam):
        """physical points (nel, nq, d) for barycentric points (nq, d+1)."""
        X = self.coords[self.elems]
        return np.einsum('qi,eid->eqd', lam, X)


# =============================================================== spaces


class Space:
    """A single-component FE space: kind 'H1' or 'L2', order p, ncomp copies (VectorH1 = ncomp=dim).

    Global numbering (ndof of one copy = n1): H1 vertex dofs (vertex id), then
    (p-1) dofs per edge in edge order, then face dofs (3D faces, or triangles in 2D);
    L2: (p+1)(p+2)/2 consecutive dofs per element. A vector space repeats the
    scalar block ncomp times.
    """

    def __init__(self, topo, kind, order, ncomp=1, dirichlet_mask=0):
        self.topo = topo
        self.kind = kind
        self.order = p = order
        self.ncomp = ncomp
        d = topo.dim
        if kind == 'H1':
            self.nloc1 = h1_ndof_local(d, p)
            nv, ne = topo.nv, topo.nedge
            ldofs = [topo.elems]
            if p >= 2:
                ed = nv + (p - 1) * topo.elem_edges[:, :, None] + np.arange(p - 1)[None, None, :]
                ldofs.append(ed.reshape(topo.nel, -1))
            n1 = nv + (p - 1) * ne
            if p >= 3:
                if d == 2:
                    ldofs.append((n1 + np.arange(topo.nel))[:, None])
                    n1 += topo.nel
                else:
                    ldofs.append(n1 + topo.elem_faces)
                    n1 += topo.nface
            self.edofs1 = np.concatenate(ldofs, 1)
            self.n1 = n1
            # dirichlet dofs
            bd = np.zeros(n1, bool)
            sel = (topo.f_marker >= 0) & (((dirichlet_mask >> np.maximum(topo.f_marker, 0)) & 1) == 1)
            fverts = topo.facets[sel]
            bd[fverts.ravel()] = True
            if p >= 2:
                if d == 2:
                    eids = np.nonzero(sel)[0]
                else:
                    # edges of selected faces
                    fe = topo.elem_edges  # map faces to edges through elements
                    eids = _face_edges(topo, np.nonzero(sel)[0])
                for k in range(p - 1):
                    bd[nv + (p - 1) * eids + k] = True
            if p >= 3 and d == 3:
                bd[nv + (p - 1) * ne + np.nonzero(sel)[0]] = True
            self.dirichlet1 = bd
        else:
            self.nloc1 = l2_ndof_local(d, p)
            self.n1 = topo.nel * self.nloc1
            self.edofs1 = np.arange(self.n1).reshape(topo.nel, self.nloc1)
            self.dirichlet1 = np.zeros(self.n1, bool)
        self.ndof = self.n1 * ncomp
        self.nloc = self.nloc1 * ncomp
        self.edofs = np.concatenate([self.edofs1 + c * self.n1 for c in range(ncomp)], 1)
        self.dirichlet = np.tile(self.dirichlet1, ncomp)


def _face_edges(topo, face_ids):
    """edge ids of the given faces (3D)."""
    F = topo.faces[face_ids]
    nv = topo.nv
    ekey = topo.edges[:, 0] * nv + topo.edges[:, 1]
    out = []
    for (a, b) in [(0, 1), (0, 2), (1, 2)]:
        k = F[:, a] * nv + F[:, b]
        out.append(np.searchsorted(ekey, k))
    return np.unique(np.concatenate(out))


class Compound:
    """Compound space: components blocked one after another (ngs.FESpace([...]))."""

    def __init__(self, spaces, dgjumps=False):
        self.spaces = spaces
        self.topo = spaces[0].topo
        self.offsets = np.cumsum([0] + [s.ndof for s in spaces])
        self.ndof = int(self.offsets[-1])
        self.dgjumps = dgjumps
        self.edofs = np.concatenate([s.edofs + o for s, o in zip(spaces, self.offsets[:-1])], 1)
        self.freedofs = ~np.concatenate([s.dirichlet for s in spaces])

    def comp_slice(self, c):
        return slice(int(self.offsets[c]), int(self.offsets[c + 1]))


# =============================================================== pattern


def build_pattern(cs):
    """Full element-clique CSR pattern (+ facet-neighbour cliques when dgjumps)."""
    topo = cs.topo
    ed = cs.edofs
    nl = ed.shape[1]
    rows = [np.repeat(ed, nl, axis=1).ravel()]
    cols = [np.tile(ed, (1, nl)).ravel()]
    if cs.dgjumps:
        inner = topo.f_e1 >= 0
        a = ed[topo.f_e0[inner]]
        b = ed[topo.f_e1[inner]]
        rows += [np.repeat(a, nl, 1).ravel(), np.repeat(b, nl, 1).ravel()]
        cols += [np.tile(b, (1, nl)).ravel(), np.tile(a, (1, nl)).ravel()]
    r = np.concatenate(rows).astype(np.int64)
    c = np.concatenate(cols).astype(np.int64)
    key = np.unique(r * cs.ndof + c)
    rr = key // cs.ndof
    cc = key % cs.ndof
    indptr = np.zeros(cs.ndof + 1, np.int64)
    np.add.at(indptr, rr + 1, 1)
    indptr = np.cumsum(indptr)
    return indptr, cc, key


class Assembler:
    """Accumulates (row, col, value) triples into a fixed CSR pattern."""

    def __init__(self, cs):
        self.cs = cs
        self.indptr, self.indices, self.keys = build_pattern(cs)
        self.nnz = len(self.indices)
        self.parts = []

    def add(self, rows, cols, vals):
        k = rows.astype(np.int64).ravel() * self.cs.ndof + cols.astype(np.int64).ravel()
        slot = np.searchsorted(self.keys, k)
        if np.any(self.keys[slot] != k):
            raise AssertionError('entry outside pattern')
        self.parts.append((slot, vals.ravel()))

    def matrix(self):
        v = np.zeros(self.nnz)
        for slot, vals in self.parts:
            v += np.bincount(slot, weights=vals, minlength=self.nnz)
        n = self.cs.ndof
        return sp.csr_matrix((v, self.indices.copy(), self.indptr.copy()), shape=(n, n))


# =============================================================== coefficient helpers


def _coef_at(coef, topo, X):
    """coef -> array (nel, nq). scalar | (nel,) per-element | callable(x, y[, z]) -> values."""
    nel, nq = X.shape[0], X.shape[1]
    if callable(coef):
        return np.broadcast_to(coef(*[X[..., i] for i in range(X.shape[2])]), (nel, nq)).astype(float)
    coef = np.asarray(coef, dtype=float)
    if coef.ndim == 0:
        return np.full((nel, nq), float(coef))
    if coef.shape == (nel,):
        return np.repeat(coef[:, None], nq, 1)
    if coef.shape == (nel, nq):
        return coef
    raise ValueError('bad coefficient shape')


class GridFunctionEval:
    """Evaluates an H1 (possibly vector) field given by its dof vector on a Space."""

    def __init__(self, space, vec):
        self.space = space
        self.vec = np.asarray(vec, dtype=float)

    def at(self, lam):
        """values (nel, nq, ncomp) and gradients (nel, nq, ncomp, d) at barycentric points."""
        s = self.space
        topo = s.topo
        phi, gref = ref_basis(s.kind, topo.dim, s.order, lam)
        gphys = np.einsum('eab,qla->eqlb', topo.Jinv, gref)     # J^{-T} grad_ref
        vals, grads = [], []
        for c in range(s.ncomp):
            u = self.vec[s.edofs1 + c * s.n1]                    # (nel, nl)
            vals.append(np.einsum('ql,el->eq', phi, u))
            grads.append(np.einsum('eqld,el->eqd', gphys, u))
        return np.stack(vals, 2), np.stack(grads, 2)


def _wind_at(wind, topo, lam, X):
    if isinstance(wind, GridFunctionEval):
        return wind.at(lam)[0]
    if callable(wind):
        w = wind(*[X[..., i] for i in range(X.shape[2])])
        return np.stack([np.broadcast_to(wi, X.shape[:2]) for wi in w], 2)
    w = np.asarray(wind, dtype=float)
    return np.broadcast_to(w, X.shape[:2] + (w.shape[0],))


# =============================================================== volume integrators


def _vol_tables(space, qdeg):
    topo = space.topo
    lam, w = simplex_rule(topo.dim, qdeg)
    phi, gref = ref_basis(space.kind, topo.dim, space.order, lam)
    g = np.einsum('eab,qla->eqlb', topo.Jinv, gref)               # (nel, nq, nl1, d)
    return lam, w, phi, g


def assemble_volume(asm, terms, qdeg=None):
    """Volume bilinear terms. Each term is a dict with
       kind: 'grad_grad' (coef grad u . grad v), 'mass' (coef u v),
             'conv' (coef u (w . grad v)), 'dtrial' (coef d_{dir} u v), 'dtest' (coef u d_{dir} v)
       trial, test: component indices in the compound space; coef; wind; dir.
    """
    cs = asm.cs
    topo = cs.topo
    for t in terms:
        su, sv = cs.spaces[t['trial']], cs.spaces[t['test']]
        pmax = max(su.order, sv.order)
        q = qdeg if qdeg is not None else 2 * pmax + 2
        lam, w = simplex_rule(topo.dim, q)
        X = topo.map_points(lam)
        fu, gu = ref_basis(su.kind, topo.dim, su.order, lam)
        fv, gv = ref_basis(sv.kind, topo.dim, sv.order, lam)
        Gu = np.einsum('eab,qla->eqlb', topo.Jinv, gu)
        Gv = np.einsum('eab,qla->eqlb', topo.Jinv, gv)
        cf = t.get('coef', 1.0)
        if isinstance(cf, tuple) and cf[0] == 'field':
            # ('field', GridFunctionEval, mode, scalar): scalar * phi (mode 1) or scalar * (1 - phi) (mode 2)
            ph = cf[1].at(lam)[0][..., 0]
            cv = cf[3] * (ph if cf[2] == 1 else 1.0 - ph)
        else:
            cv = _coef_at(cf, topo, X)
        c = cv * (w[None, :] * topo.absdet[:, None])   # (nel, nq)
        k = t['kind']
        for cu in range(su.ncomp):
            for cv in range(sv.ncomp):
                if k in ('grad_grad', 'mass', 'conv') and su.ncomp > 1 and cu != cv:
                    continue
                if k == 'grad_grad':
                    E = np.einsum('eq,eqid,eqjd->eij', c, Gv, Gu)
                elif k == 'mass':
                    E = np.einsum('eq,qi,qj->eij', c, fv, fu)
                elif k == 'conv':
                    W = _wind_at(t['wind'], topo, lam, X)                                # (nel, nq, d)
                    E = np.einsum('eq,eqd,eqid,qj->eij', c, W, Gv, fu)
                elif k == 'dtrial':
                    # d_{dir} of trial (component cu of a vector trial is used with dir=cu when 'div')
                    dd = t['dir'] if t['dir'] >= 0 else cu
                    if t['dir'] < 0 and su.ncomp == 1:
                        raise ValueError
                    if su.ncomp > 1 and t['dir'] >= 0:
                        raise ValueError('vector trial needs dir=-1 (divergence)')
                    E = np.einsum('eq,qi,eqj->eij', c, fv, Gu[..., dd])
                elif k == 'dtest':
                    dd = t['dir'] if t['dir'] >= 0 else cv
                    if sv.ncomp > 1 and t['dir'] >= 0:
                        raise ValueError('vector test needs dir=-1 (divergence)')
                    E = np.einsum('eq,eqi,qj->eij', c, Gv[..., dd], fu)
                else:
                    raise ValueError(k)
                ro = cs.offsets[t['test']] + cv * sv.n1
                co = cs.offsets[t['trial']] + cu * su.n1
                R = (sv.edofs1 + ro)
                C = (su.edofs1 + co)
                asm.add(np.repeat(R[:, :, None], C.shape[1], 2), np.repeat(C[:, None, :], R.shape[1], 1), E)


def assemble_source(cs, comp, f, qdeg=None, vcomp=None):
    """int f v dx for component ``comp`` (scalar f; or vector f with one entry per vector component)."""
    s = cs.spaces[comp]
    topo = cs.topo
    q = qdeg if qdeg is not None else 2 * s.order + 4
    lam, w = simplex_rule(topo.dim, q)
    X = topo.map_points(lam)
    phi, _ = ref_basis(s.kind, topo.dim, s.order, lam)
    b = np.zeros(cs.ndof)
    fl = f if isinstance(f, (list, tuple)) else [f]
    for c in range(s.ncomp):
        if c >= len(fl):
            break
        fv = _coef_at(fl[c], topo, X) * (w[None, :] * topo.absdet[:, None])
        loc = np.einsum('eq,qi->ei', fv, phi)
        np.add.at(b, s.edofs1 + cs.offsets[comp] + c * s.n1, loc)
    return b


# =============================================================== facet integrators (DG, triangles)


def _facet_lam(dim, k, t):
    """barycentric coords of facet points: facet k of local vertices sorted; t (nq,) on [0,1] (2D)."""
    a, b = TRI_FACETS[k]
    lam = np.zeros((t.shape[0], 3))
    lam[:, a] = 1 - t
    lam[:, b] = t
    return lam


def assemble_dg_facets(asm, comp, eps, alpha_coef, wind, dirichlet_mask, outflow_mask, qdeg=None,
                       sipg=True, upwind=True):
    """DG interior-penalty + upwind facet terms of the MCINS DG transport form (one species):

    interior facets  eps*(alpha[c][r] - {grad c}.n[r] - {grad r}.n[c]) + [r](c max(w.n,0) + c' min(w.n,0))
        (opencmp/models/multi_component_ins.py:185-197)
    Dirichlet facets eps*(alpha c r - r grad c.n - c grad r.n) + r c max(w.n,0)   (:287-301)
    other boundary facets (total-flux / unlabelled)  r c max(w.n,0)               (:262-275)
    with alpha = alpha_coef / h, h = |det J|^(1/d) (face value: mean of the two sides).
    """
    cs = asm.cs
    s = cs.spaces[comp]
    topo = cs.topo
    if topo.dim != 2:
        raise NotImplementedError
    q = qdeg if qdeg is not None else 2 * s.order + 2
    t, wq = line_rule(q)
    off = cs.offsets[comp]
    tabs = []
    for k in range(3):
        lam = _facet_lam(2, k, t)
        phi, gref = ref_basis(s.kind, 2, s.order, lam)
        tabs.append((lam, phi, gref))
    X = topo.coords
    for f_side in ('interior', 'boundary'):
        if f_side == 'interior':
            sel = np.nonzero(topo.f_e1 >= 0)[0]
        else:
            sel = np.nonzero(topo.f_e1 < 0)[0]
        if len(sel) == 0:
            continue
        e0, k0 = topo.f_e0[sel], topo.f_k0[sel]
        fv = topo.facets[sel]
        L = np.linalg.norm(X[fv[:, 1]] - X[fv[:, 0]], axis=1)
        # points: param t from lower to higher global vertex
        P = X[fv[:, 0]][:, None, :] * (1 - t)[None, :, None] + X[fv[:, 1]][:, None, :] * t[None, :, None]
        # outward normal of e0
        gl = topo.grad_lam[e0, k0]                                   # (nf, d)
        n = -gl / np.linalg.norm(gl, axis=1, keepdims=True)
        if isinstance(wind, GridFunctionEval):
            raise NotImplementedError
        W = _wind_at(wind, topo, None, P) if not isinstance(wind, (list, tuple, np.ndarray)) else \
            np.broadcast_to(np.asarray(wind, float), P.shape)
        wn = np.einsum('fqd,fd->fq', W, n)
        ww = wq[None, :] * L[:, None]                                # (nf, nq)

        def side(e, k):
            phi = np.stack([tabs[kk][1] for kk in range(3)])[k]          # (nf, nq, nl)
            gref = np.stack([tabs[kk][2] for kk in range(3)])[k]         # (nf, nq, nl, d)
            g = np.einsum('fab,fqla->fqlb', topo.Jinv[e], gref)
            return phi, g

        f0, g0 = side(e0, k0)
        dn0 = np.einsum('fqld,fd->fql', g0, n)
        R0 = cs.edofs[e0][:, s.edofs1.shape[1] * 0:][:, :s.nloc1] + 0
        R0 = s.edofs1[e0] + off
        if f_side == 'interior':
            e1, k1 = topo.f_e1[sel], topo.f_k1[sel]
            f1, g1 = side(e1, k1)
            dn1 = np.einsum('fqld,fd->fql', g1, n)
            R1 = s.edofs1[e1] + off
            h = 0.5 * (topo.h[e0] + topo.h[e1])
            a = alpha_coef / h
            # blocks (test side, trial side); jump = this - other, n outward of e0
            # test e0: [r]=f0 ; test e1: [r]=-f1 ; trial e0: [c]=f0,{dc}=dn0/2 ; trial e1: [c]=-f1,{dc}=dn1/2
            blocks = []
            for (tr, tf, tdn, Rt) in ((+1, f0, dn0, R0), (-1, f1, dn1, R1)):
                for (sc, cf, cdn, Rc, is_this) in ((+1, f0, dn0, R0, True), (-1, f1, dn1, R1, False)):
                    E = np.zeros((len(sel), s.nloc1, s.nloc1))
                    if sipg:
                        E += eps * (np.einsum('fq,fqi,fqj->fij', a[:, None] * ww, tr * tf, sc * cf)
                                    - np.einsum('fq,fqi,fqj->fij', ww, tr * tf, 0.5 * cdn)
                                    - np.einsum('fq,fqi,fqj->fij', ww, 0.5 * tdn, sc * cf))
                    if upwind:
                        up = np.maximum(wn, 0.0) if is_this else np.minimum(wn, 0.0)
                        E += np.einsum('fq,fqi,fqj->fij', ww * up, tr * tf, cf)
                    blocks.append((Rt, Rc, E))
        else:
            mk = topo.f_marker[sel]
            isd = ((dirichlet_mask >> np.maximum(mk, 0)) & 1) == 1
            a = alpha_coef / topo.h[e0]
            E = np.zeros((len(sel), s.nloc1, s.nloc1))
            if sipg:
                Ed = eps * (np.einsum('fq,fqi,fqj->fij', a[:, None] * ww, f0, f0)
                            - np.einsum('fq,fqi,fqj->fij', ww, f0, dn0)
                            - np.einsum('fq,fqi,fqj->fij', ww, dn0, f0))
                E += Ed * isd[:, None, None]
            if upwind:
                E += np.einsum('fq,fqi,fqj->fij', ww * np.maximum(wn, 0.0), f0, f0)
            blocks = [(R0, R0, E)]
        for (Rt, Rc, E) in blocks:
            asm.add(np.repeat(Rt[:, :, None], Rc.shape[1], 2), np.repeat(Rc[:, None, :], Rt.shape[1], 1), E)


def assemble_dg_dirichlet_rhs(cs, comp, eps, alpha_coef, wind, g, dirichlet_mask, qdeg=None):
    """RHS Dirichlet terms eps*g*(alpha r - grad r.n) - r g min(w.n,0) (opencmp/models/multi_component_ins.py:371-386)."""
    s = cs.spaces[comp]
    topo = cs.topo
    q = qdeg if qdeg is not None else 2 * s.order + 4
    t, wq = line_rule(q)
    b = np.zeros(cs.ndof)
    sel = np.nonzero((topo.f_e1 < 0) & (topo.f_marker >= 0))[0]
    sel = sel[((dirichlet_mask >> topo.f_marker[sel]) & 1) == 1]
    if len(sel) == 0:
        return b
    e0, k0 = topo.f_e0[sel], topo.f_k0[sel]
    fv = topo.facets[sel]
    X = topo.coords
    L = np.linalg.norm(X[fv[:, 1]] - X[fv[:, 0]], axis=1)
    P = X[fv[:, 0]][:, None, :] * (1 - t)[None, :, None] + X[fv[:, 1]][:, None, :] * t[None, :, None]
    gl = topo.grad_lam[e0, k0]
    n = -gl / np.linalg.norm(gl, axis=1, keepdims=True)
    W = np.broadcast_to(np.asarray(wind, float), P.shape) if not callable(wind) else _wind_at(wind, topo, None, P)
    wn = np.einsum('fqd,fd->fq', W, n)
    gv = g(P[..., 0], P[..., 1]) if callable(g) else np.full(P.shape[:2], float(g))
    ww = wq[None, :] * L[:, None]
    phis, grs = [], []
    for k in range(3):
        lam = _facet_lam(2, k, t)
        ph, gr = ref_basis(s.kind, 2, s.order, lam)
        phis.append(ph)
        grs.append(gr)
    f0 = np.stack(phis)[k0]
    g0 = np.einsum('fab,fqla->fqlb', topo.Jinv[e0], np.stack(grs)[k0])
    dn0 = np.einsum('fqld,fd->fql', g0, n)
    a = alpha_coef / topo.h[e0]
    loc = eps * (np.einsum('fq,fqi->fi', a[:, None] * ww * gv, f0) - np.einsum('fq,fqi->fi', ww * gv, dn0)) \
        - np.einsum('fq,fqi->fi', ww * gv * np.minimum(wn, 0.0), f0)
    np.add.at(b, s.edofs1[e0] + cs.offsets[comp], loc)
    return b


# =============================================================== Dirichlet lift and solves


def _bnd_facets(cs, comp, marker_mask, qdeg):
    """Marked boundary edges of an H1 component (triangles): owner element, basis values at the edge
    rule, rule weights times edge length, points and outward normals."""
    s = cs.spaces[comp]
    topo = cs.topo
    if topo.dim != 2:
        raise NotImplementedError
    t, wq = line_rule(qdeg)
    mk = topo.f_marker
    sel = np.nonzero((topo.f_e1 < 0) & (mk >= 0) & (((marker_mask >> np.maximum(mk, 0)) & 1) == 1))[0]
    e0, k0 = topo.f_e0[sel], topo.f_k0[sel]
    X = topo.coords
    fv = topo.facets[sel]
    L = np.linalg.norm(X[fv[:, 1]] - X[fv[:, 0]], axis=1)
    P = X[fv[:, 0]][:, None, :] * (1 - t)[None, :, None] + X[fv[:, 1]][:, None, :] * t[None, :, None]
    gl = topo.grad_lam[e0, k0]
    n = -gl / np.linalg.norm(gl, axis=1, keepdims=True)
    phis = np.stack([ref_basis(s.kind, 2, s.order, _facet_lam(2, k, t))[0] for k in range(3)])
    return s, e0, phis[k0], wq[None, :] * L[:, None], P, n


def assemble_bnd_facets(asm, comp, coef, marker_mask, wind=None, qdeg=None):
    """coef * max(w.n, 0) u v (wind given: the stress-BC no-backflow term,
    opencmp/models/ins.py:210-216) or coef * u v (wind None: the Robin term,
    opencmp/models/poisson.py:94-96) on the boundary edges whose marker is in marker_mask, every
    copy of H1 component ``comp`` on its own block (triangles)."""
    s = asm.cs.spaces[comp]
    q = qdeg if qdeg is not None else 2 * s.order + 2
    s, e0, phi, ww, P, n = _bnd_facets(asm.cs, comp, marker_mask, q)
    if len(e0) == 0:
        return
    if wind is not None:
        if isinstance(wind, GridFunctionEval):
            # a field on this component's space: its trace from the facet basis values
            ws = wind.space
            if ws.order != s.order or ws.kind != s.kind:
                raise ValueError('a field wind must live on the component space')
            W = np.stack([np.einsum('fql,fl->fq', phi, wind.vec[ws.edofs1[e0] + c * ws.n1])
                          for c in range(ws.ncomp)], 2)
        else:
            W = _wind_at(wind, asm.cs.topo, None, P)
        ww = ww * np.maximum(np.einsum('fqd,fd->fq', W, n), 0.0)
    Ke = coef * np.einsum('fq,fqi,fqj->fij', ww, phi, phi)
    for c in range(s.ncomp):
        d = s.edofs1[e0] + asm.cs.offsets[comp] + c * s.n1
        nl = d.shape[1]
        asm.add(np.repeat(d, nl, axis=1), np.tile(d, (1, nl)), Ke.reshape(len(e0), -1))


def assemble_bnd_source(cs, comp, g, marker_mask, coef=1.0, qdeg=None):
    """coef * int g v ds over the boundary edges whose marker is in marker_mask (Neumann and Robin
    data, opencmp/models/poisson.py:151-158), g a number or g(x, y), first copy of H1 component
    ``comp`` (triangles)."""
    s = cs.spaces[comp]
    q = qdeg if qdeg is not None else 2 * s.order + 4
    s, e0, phi, ww, P, n = _bnd_facets(cs, comp, marker_mask, q)
    b = np.zeros(cs.ndof)
    if len(e0) == 0:
        return b
    gv = g(P[..., 0], P[..., 1]) if callable(g) else np.full(P.shape[:2], float(g))
    loc = coef * np.einsum('fq,fqi->fi', ww * gv, phi)
    np.add.at(b, s.edofs1[e0] + cs.offsets[comp], loc)
    return b


def h1_interpolate_vertices(space, g):
    """Values of g at the vertex dofs (higher-order dofs set to 0): exact for P1-representable data."""
    topo = space.topo
    out = np.zeros(space.ndof)
    gv = g(*[topo.coords[:, i] for i in range(topo.dim)])
    out[:topo.nv] = gv
    return out


def h1_lift(space, g, dirichlet_mask):
    """Boundary values of an H1 space on the facets with the given markers:
    vertex dofs take g at the vertex; edge dofs the 1D L2 projection of g minus its
    linear interpolant onto the edge bubbles; face dofs (3D, order 3) the L2 projection over
    the face of g minus its vertex and edge parts onto the face bubble lambda_a lambda_b
    lambda_c. Stands in for GridFunction.Set(g, definedon=Boundaries)
    (opencmp/models/base_model.py:321-341); NGSolve's own projection is parity unpinned, the
    tests pin exactness on the trace space (a cubic's lift reproduces it on every face)."""
    topo = space.topo
    d = topo.dim
    out = np.zeros(space.ndof)
    sel = np.nonzero((topo.f_marker >= 0) & (((dirichlet_mask >> np.maximum(topo.f_marker, 0)) & 1) == 1))[0]
    verts = np.unique(topo.facets[sel].ravel())
    X = topo.coords
    gv = lambda P: g(*[P[..., i] for i in range(d)])
    out[verts] = gv(X[verts])
    p = space.order
    if p >= 2:
        eids = sel if d == 2 else _face_edges(topo, sel)
        a, b = topo.edges[eids, 0], topo.edges[eids, 1]
        t, w = line_rule(12)
        P = X[a][:, None, :] * (1 - t)[None, :, None] + X[b][:, None, :] * t[None, :, None]
        res = gv(P) - (out[a][:, None] * (1 - t) + out[b][:, None] * t)
        b2 = (1 - t) * t
        b3 = (1 - t) * t * (2 * t - 1)
        out[topo.nv + (p - 1) * eids] = (res * b2) @ w / (b2 * b2 @ w)
        if p >= 3:
            out[topo.nv + (p - 1) * eids + 1] = (res * b3) @ w / (b3 * b3 @ w)
    if p >= 3 and d == 3 and len(sel):
        F = topo.faces[sel]                                   # ascending vertex ids a < b < c
        nv = topo.nv
        ekey = topo.edges[:, 0] * nv + topo.edges[:, 1]
        E = [np.searchsorted(ekey, F[:, i] * nv + F[:, j]) for (i, j) in [(0, 1), (0, 2), (1, 2)]]
        lam, w = simplex_rule(2, 12)                          # barycentrics (a, b, c) of each point
        P = np.einsum('qk,fkd->fqd', lam, X[F])
        la, lb, lc = lam[:, 0], lam[:, 1], lam[:, 2]
        part = out[F] @ lam.T                                 # vertex part, (faces, points)
        for k, (l1, l2) in enumerate([(la, lb), (la, lc), (lb, lc)]):
            e2 = out[topo.nv + 2 * E[k]]
            e3 = out[topo.nv + 2 * E[k] + 1]
            part += np.outer(e2, l1 * l2) + np.outer(e3, l1 * l2 * (l2 - l1))
        res = gv(P) - part
        B = la * lb * lc
        out[topo.nv + 2 * topo.nedge + sel] = (res * B) @ w / ((B * B) @ w)
    return out


def solve_free(A, b, x0, free):
    """x_F = A_FF^{-1} (b_F - A_FD x_D); x_D = x0_D. Matches the direct branch of
    opencmp/models/base_model.py:908-922 (r = L - A gfu; gfu += inv r on freedofs)."""
    x = x0.copy()
    r = b - A @ x
    F = np.nonzero(free)[0]
    AFF = A[F][:, F].tocsc()
    x[F] += spla.spsolve(AFF, r[F])
    return x


def pcg_jacobi(A, b, x0, free, tol, maxit=100000):
    """Jacobi-PCG on the free dofs, stopping on ||b-Ax||/||b|| (free dofs) <= tol."""
    x = x0.copy()
    m = free.astype(float)
    d = A.diagonal()
    dinv = np.where(free, 1.0 / np.where(d == 0, 1, d), 0.0)
    r = (b - A @ x) * m
    bn = np.linalg.norm(b * m)
    z = dinv * r
    p = z.copy()
    rz = r @ z
    it = 0
    while np.linalg.norm(r) > tol * bn and it < maxit:
        q = (A @ p) * m
        al = rz / (p @ q)
        x += al * p
        r -= al * q
        z = dinv * r
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
        it += 1
    return x, it, np.linalg.norm(r) / bn


def block_jacobi_inverse(A, nb):
    """Inverse of the nb x nb diagonal blocks (element blocks of an L2 space)."""
    n = A.shape[0]
    ne = n // nb
    Ad = A.tocsr()
    blocks = np.zeros((ne, nb, nb))
    rows = np.arange(n)
    for j in range(nb):
        cols = (rows // nb) * nb + j
        blocks[:, :, j] = np.asarray(Ad[rows, cols]).reshape(ne, nb)
    return np.linalg.inv(blocks)


def gmres_right(A, b, x0, Minv_apply, tol, m=50, maxit=100000, free=None):
    """Restarted right-preconditioned GMRES(m), CGS2 orthogonalisation, stopping on the true relative residual."""
    mask = np.ones_like(b) if free is None else free.astype(float)
    x = x0.copy()
    bn = np.linalg.norm(b * mask)
    total = 0
    while True:
        r = (b - A @ x) * mask
        beta = np.linalg.norm(r)
        if beta <= tol * bn or total >= maxit:
            return x, total, beta / bn
        V = np.zeros((m + 1, len(b)))
        Z = np.zeros((m, len(b)))
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs_, sn_ = np.zeros(m), np.zeros(m)
        j = 0
        for j in range(m):
            Z[j] = Minv_apply(V[j]) * mask
            w = (A @ Z[j]) * mask
            h = V[:j + 1] @ w
            w -= h @ V[:j + 1]
            h2 = V[:j + 1] @ w
            w -= h2 @ V[:j + 1]
            h += h2
            H[:j + 1, j] = h
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t0 = cs_[i] * H[i, j] + sn_[i] * H[i + 1, j]
                H[i + 1, j] = -sn_[i] * H[i, j] + cs_[i] * H[i + 1, j]
                H[i, j] = t0
            den = np.hypot(H[j, j], H[j + 1, j])
            cs_[j], sn_[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn_[j] * g[j]
            g[j] = cs_[j] * g[j]
            total += 1
            if abs(g[j + 1]) <= tol * bn * 0.5 or total >= maxit:
                break
        k = j + 1
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x += y @ Z[:k]


# =============================================================== error norms


def l2_error(space, vec, exact, comp=0, qdeg=None, grad_exact=None):
    """sqrt(int (u_h - u)^2) (opencmp/helpers/error.py:75-77) and, if grad_exact given, the H1-seminorm error."""
    topo = space.topo
    q = qdeg if qdeg is not None else 2 * space.order + 6
    lam, w = simplex_rule(topo.dim, q)
    X = topo.map_points(lam)
    ge = GridFunctionEval(space, vec)
    vals, grads = ge.at(lam)
    wq = w[None, :] * topo.absdet[:, None]
    ex = _coef_at(exact, topo, X)
    e2 = np.sum(wq * (vals[..., comp] - ex) ** 2)
    if grad_exact is None:
        return float(np.sqrt(e2))
    gx = np.stack([_coef_at(gf, topo, X) for gf in grad_exact], 2)
    h2 = np.sum(wq[..., None] * (grads[..., comp, :] - gx) ** 2)
    return float(np.sqrt(e2)), float(np.sqrt(h2))


# =============================================================== parity status

PARITY = {
    'oseen_picard': 'pinned: Poiseuille flow (an exact Navier-Stokes solution in the P2/P1 space) is '
                    'reproduced to round-off, the Picard errors fall monotonically at low Reynolds number',
    'h1_lift (3D P3 faces)': 'pinned: a cubic lift reproduces the cubic exactly (a zero face dof does not)',
    'simplex_rule': 'pinned: closed-form simplex monomial integrals',
    'h1_eval': 'pinned: partition of unity of the vertex part, polynomial reproduction (P2/P3 exact solves)',
    'dubiner_polys': 'pinned: orthonormality, reproduction of P2 polynomials',
    'Topology': 'pinned: Euler characteristic, reference .vol topology counts',
    'Space numbering': 'NGSolve vertex-first layout pinned by .sol fixture sizes; edge/face order and '
                       'basis polynomials are parity unpinned against NGSolve itself (not importable)',
    'build_pattern': 'pinned: brute-force clique enumeration on tiny meshes',
    'assemble_volume': 'pinned: P1 closed-form stiffness, K.1=0, sum(M)=|T|, exact P2/P3 solves',
    'assemble_dg_facets': 'pinned: consistency (exact DG solution for global quadratics), symmetry of SIPG part',
    'solve_free': 'pinned: dense numpy solve on tiny meshes',
    'gmres_right/pcg_jacobi': 'pinned: true residual check against solve_free',
    'l2_error': 'pinned: manufactured solution rates (P2: L2 3, H1 2) and the reference transient Poisson value',
}


def oseen_picard(cs, kv, lift, b, abs_tol, rel_tol, max_iter, W0=None, qdeg=None, stress_mask=0):
    """Stationary Oseen/Picard iteration of opencmp/models/ins.py:323-359 (linearized_solve :361-379)
    for Taylor-Hood spaces cs = Compound([vector H1 u, H1 p]):

        repeat: assemble kv grad u:grad v - (u (x) W):grad v - div(u) q - div(v) p - 1e-10 p q
                (ins.py:191-195, :240-244) + max(W.n, 0) u.v on the stress boundaries in
                stress_mask (ins.py:210-216), solve on the free dofs from the Dirichlet lift ``lift``
                (scipy direct), err = ||W - u||_L2 (velocity), W <- u,
        until err < abs_tol + rel_tol ||u||_L2 or max_iter iterations.

    b: the load vector (sources, tractions). Returns (x, [err per iteration])."""
    su = cs.spaces[0]
    W = np.zeros(cs.ndof) if W0 is None else np.asarray(W0, dtype=float).copy()
    errs = []
    x = lift.copy()
    for _ in range(max_iter):
        A = Assembler(cs)
        wind = GridFunctionEval(su, W[:su.ndof])
        assemble_volume(A, [dict(kind='grad_grad', trial=0, test=0, coef=kv),
                            dict(kind='conv', trial=0, test=0, coef=-1.0, wind=wind),
                            dict(kind='dtrial', trial=0, test=1, coef=-1.0, dir=-1),
                            dict(kind='dtest', trial=1, test=0, coef=-1.0, dir=-1),
                            dict(kind='mass', trial=1, test=1, coef=-1e-10)], qdeg=qdeg)
        if stress_mask:
            assemble_bnd_facets(A, 0, 1.0, stress_mask, wind=wind)
        K = A.matrix()
        x = solve_free(K, b, lift, cs.freedofs)
        du = x[:su.ndof] - W[:su.ndof]
        err = np.sqrt(sum(l2_error(su, du, lambda *X: 0 * X[0], comp=c) ** 2 for c in range(su.ncomp)))
        un = np.sqrt(sum(l2_error(su, x[:su.ndof], lambda *X: 0 * X[0], comp=c) ** 2 for c in range(su.ncomp)))
        errs.append(err)
        W = x.copy()
        if err < abs_tol + rel_tol * un:
            break
    return x, errs
