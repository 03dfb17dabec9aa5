// C ABI entry points, context, mesh topology, spaces, CSR pattern and the
// deterministic assembly schedule (host side of the library).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <array>
#include <cstring>
#include <numeric>

#include "internal.h"

namespace ocmp {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

template <class F>
static int guard(F &&f) {
  try {
    f();
    return OCMP_OK;
  } catch (const Error &e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return OCMP_ERR_INTERNAL;
  } catch (const std::exception &e) {
    set_error(e.what());
    return OCMP_ERR_INTERNAL;
  }
}

// ------------------------------------------------------------------ context timing

int Ctx::cls(const std::string &name) {
  auto it = class_ids.find(name);
  if (it != class_ids.end()) return it->second;
  int id = (int)class_names.size();
  class_names.push_back(name);
  class_ids[name] = id;
  return id;
}

hipEvent_t Ctx::get_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  OCMP_HIP(hipEventCreate(&e));
  return e;
}

void Ctx::tic(int c, hipEvent_t *a) {
  *a = get_event();
  OCMP_HIP(hipEventRecord(*a, stream));
  (void)c;
}

void Ctx::toc(int c, hipEvent_t a, double bytes) {
  hipEvent_t b = get_event();
  OCMP_HIP(hipEventRecord(b, stream));
  pending.push_back({c, a, b, bytes});
  if (pending.size() > 4096) flush_timing();
}

void Ctx::flush_timing() {
  if (pending.empty()) return;
  OCMP_HIP(hipEventSynchronize(pending.back().b));
  for (auto &r : pending) {
    float ms = 0.f;
    OCMP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    auto &slot = acc[r.cls];
    slot.first += 1;
    slot.second += ms;
    acc_bytes[r.cls] += r.bytes;
    event_pool.push_back(r.a);
    event_pool.push_back(r.b);
  }
  pending.clear();
}

// ------------------------------------------------------------------ topology

static const int TRI_EDGES[3][2] = {{0, 1}, {0, 2}, {1, 2}};
static const int TET_EDGES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int TET_FACES[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
static const int TRI_FACET_EDGE[3] = {2, 1, 0};  // facet k (opposite vertex k) -> local edge

static void unique_keys(const std::vector<uint64_t> &keys, std::vector<uint64_t> &uniq, std::vector<int32_t> &inv) {
  std::vector<int64_t> idx(keys.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
  inv.assign(keys.size(), -1);
  uniq.clear();
  for (size_t t = 0; t < idx.size(); ++t) {
    uint64_t k = keys[idx[t]];
    if (uniq.empty() || uniq.back() != k) uniq.push_back(k);
    inv[idx[t]] = (int32_t)(uniq.size() - 1);
  }
}

static void build_topology(Mesh &m, int64_t nbnd, const int32_t *bnd, const int32_t *bmark) {
  const int d = m.dim, nvl = d + 1;
  const uint64_t nv = (uint64_t)m.nv;
  // edges
  const int nle = d == 2 ? 3 : 6;
  std::vector<uint64_t> keys((size_t)m.nel * nle), uniq;
  std::vector<int32_t> inv;
  for (int64_t e = 0; e < m.nel; ++e)
    for (int k = 0; k < nle; ++k) {
      int a = d == 2 ? TRI_EDGES[k][0] : TET_EDGES[k][0];
      int b = d == 2 ? TRI_EDGES[k][1] : TET_EDGES[k][1];
      keys[e * nle + k] = (uint64_t)m.elems[e * nvl + a] * nv + (uint64_t)m.elems[e * nvl + b];
    }
  unique_keys(keys, uniq, inv);
  m.nedge = (int64_t)uniq.size();
  m.edges.resize(2 * m.nedge);
  for (int64_t i = 0; i < m.nedge; ++i) {
    m.edges[2 * i] = (int32_t)(uniq[i] / nv);
    m.edges[2 * i + 1] = (int32_t)(uniq[i] % nv);
  }
  m.elem_edges = inv;
  std::vector<uint64_t> facet_keys;
  if (d == 3) {
    keys.assign((size_t)m.nel * 4, 0);
    for (int64_t e = 0; e < m.nel; ++e)
      for (int k = 0; k < 4; ++k) {
        const int32_t *v = &m.elems[e * 4];
        keys[e * 4 + k] = ((uint64_t)v[TET_FACES[k][0]] * nv + (uint64_t)v[TET_FACES[k][1]]) * nv +
                          (uint64_t)v[TET_FACES[k][2]];
      }
    unique_keys(keys, uniq, inv);
    m.nface = (int64_t)uniq.size();
    m.faces.resize(3 * m.nface);
    for (int64_t i = 0; i < m.nface; ++i) {
      m.faces[3 * i] = (int32_t)(uniq[i] / (nv * nv));
      m.faces[3 * i + 1] = (int32_t)((uniq[i] / nv) % nv);
      m.faces[3 * i + 2] = (int32_t)(uniq[i] % nv);
    }
    m.elem_faces = inv;
    m.elem_facets = inv;
    m.nfacet = m.nface;
    m.facet_verts = m.faces;
    facet_keys = uniq;
  } else {
    m.elem_facets.resize((size_t)m.nel * 3);
    for (int64_t e = 0; e < m.nel; ++e)
      for (int k = 0; k < 3; ++k) m.elem_facets[e * 3 + k] = m.elem_edges[e * 3 + TRI_FACET_EDGE[k]];
    m.nfacet = m.nedge;
    m.facet_verts = m.edges;
    facet_keys.resize(m.nedge);
    for (int64_t i = 0; i < m.nedge; ++i)
      facet_keys[i] = (uint64_t)m.edges[2 * i] * nv + (uint64_t)m.edges[2 * i + 1];
  }
  const int64_t nf = m.nfacet;
  m.f_e0.assign(nf, -1);
  m.f_k0.assign(nf, -1);
  m.f_e1.assign(nf, -1);
  m.f_k1.assign(nf, -1);
  m.f_marker.assign(nf, -1);
  for (int64_t e = 0; e < m.nel; ++e)
    for (int k = 0; k < nvl; ++k) {
      int32_t f = m.elem_facets[e * nvl + k];
      if (m.f_e0[f] < 0) {
        m.f_e0[f] = (int32_t)e;
        m.f_k0[f] = k;
      } else if (m.f_e1[f] < 0) {
        m.f_e1[f] = (int32_t)e;
        m.f_k1[f] = k;
      } else {
        throw Error(OCMP_ERR_ARG, "non-manifold mesh: facet shared by more than two elements");
      }
    }
  m.ninterior = 0;
  for (int64_t f = 0; f < nf; ++f) m.ninterior += m.f_e1[f] >= 0;
  m.nbnd = nbnd;
  for (int64_t i = 0; i < nbnd; ++i) {
    int32_t v[3];
    for (int a = 0; a < d; ++a) v[a] = bnd[i * d + a];
    std::sort(v, v + d);
    for (int a = 0; a < d; ++a) OCMP_REQUIRE(v[a] >= 0 && v[a] < m.nv, "boundary facet vertex out of range");
    uint64_t key = d == 2 ? (uint64_t)v[0] * nv + (uint64_t)v[1]
                          : ((uint64_t)v[0] * nv + (uint64_t)v[1]) * nv + (uint64_t)v[2];
    auto it = std::lower_bound(facet_keys.begin(), facet_keys.end(), key);
    OCMP_REQUIRE(it != facet_keys.end() && *it == key, "boundary facet is not a facet of the mesh");
    int64_t f = it - facet_keys.begin();
    OCMP_REQUIRE(m.f_e1[f] < 0, "boundary facet is interior");
    OCMP_REQUIRE(bmark[i] >= 0 && bmark[i] < 64, "boundary marker must be in [0,64)");
    m.f_marker[f] = bmark[i];
  }
  m.elem_nb.assign((size_t)m.nel * nvl, 0);
  m.elem_nbk.assign((size_t)m.nel * nvl, -1);
  for (int64_t e = 0; e < m.nel; ++e)
    for (int k = 0; k < nvl; ++k) {
      int32_t f = m.elem_facets[e * nvl + k];
      int32_t other = m.f_e0[f] == e ? m.f_e1[f] : m.f_e0[f];
      int32_t ok = m.f_e0[f] == e ? m.f_k1[f] : m.f_k0[f];
      if (other >= 0) {
        m.elem_nb[e * nvl + k] = other;
        m.elem_nbk[e * nvl + k] = (int8_t)ok;
      } else {
        int mk = m.f_marker[f] >= 0 ? m.f_marker[f] : 64;
        m.elem_nb[e * nvl + k] = -1 - mk;
      }
    }
}

// ------------------------------------------------------------------ spaces

static int64_t find_edge(const Mesh &m, int32_t a, int32_t b) {
  if (a > b) std::swap(a, b);
  // edges are sorted by (a, b)
  int64_t lo = 0, hi = m.nedge;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    int32_t ma = m.edges[2 * mid], mb = m.edges[2 * mid + 1];
    if (ma < a || (ma == a && mb < b))
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= m.nedge || m.edges[2 * lo] != a || m.edges[2 * lo + 1] != b)
    throw Error(OCMP_ERR_INTERNAL, "edge lookup failed");
  return lo;
}

static void build_space(Space &s, int ncomp, const ocmp_fe_desc *comps) {
  Mesh &m = *s.mesh;
  const int d = m.dim, nvl = d + 1;
  int loc = 0;
  int64_t off = 0;
  for (int c = 0; c < ncomp; ++c) {
    const ocmp_fe_desc &fd = comps[c];
    OCMP_REQUIRE(fd.ncomp == 1 || fd.ncomp == d, "ncomp must be 1 or dim");
    OCMP_REQUIRE(fd.kind == OCMP_H1 || fd.dirichlet_mask == 0, "Dirichlet dofs exist only for H1");
    for (int k = 0; k < fd.ncomp; ++k) {
      Block b;
      b.kind = fd.kind;
      b.order = fd.order;
      b.dim = d;
      b.nloc = nloc1(fd.kind, d, fd.order);
      if (fd.kind == OCMP_H1) {
        b.n1 = m.nv + (int64_t)(fd.order - 1) * m.nedge;
        if (fd.order >= 3) b.n1 += d == 2 ? m.nel : m.nface;
      } else {
        b.n1 = m.nel * (int64_t)b.nloc;
      }
      b.dof_off = off;
      b.loc_off = loc;
      b.dirichlet_mask = fd.dirichlet_mask;
      b.comp = c;
      off += b.n1;
      loc += b.nloc;
      s.blocks.push_back(b);
    }
  }
  OCMP_REQUIRE(off < (int64_t)INT32_MAX, "too many dofs for 32-bit indices");
  s.ndof = off;
  s.nloc = loc;
  s.edofs.assign((size_t)m.nel * loc, 0);
  const int nle = d == 2 ? 3 : 6;
  for (const Block &b : s.blocks) {
    for (int64_t e = 0; e < m.nel; ++e) {
      int32_t *ed = &s.edofs[e * loc + b.loc_off];
      if (b.kind == OCMP_L2) {
        for (int i = 0; i < b.nloc; ++i) ed[i] = (int32_t)(b.dof_off + e * b.nloc + i);
        continue;
      }
      int c = 0;
      for (int i = 0; i < nvl; ++i) ed[c++] = (int32_t)(b.dof_off + m.elems[e * nvl + i]);
      for (int k = 0; k < nle; ++k)
        for (int q = 0; q < b.order - 1; ++q)
          ed[c++] = (int32_t)(b.dof_off + m.nv + (int64_t)(b.order - 1) * m.elem_edges[e * nle + k] + q);
      if (b.order >= 3) {
        int64_t base = b.dof_off + m.nv + 2 * m.nedge;
        if (d == 2)
          ed[c++] = (int32_t)(base + e);
        else
          for (int f = 0; f < 4; ++f) ed[c++] = (int32_t)(base + m.elem_faces[e * 4 + f]);
      }
    }
  }
  // free dofs
  s.free.assign(s.ndof, 1);
  for (const Block &b : s.blocks) {
    if (b.kind != OCMP_H1 || !b.dirichlet_mask) continue;
    for (int64_t f = 0; f < m.nfacet; ++f) {
      int mk = m.f_marker[f];
      if (m.f_e1[f] >= 0 || mk < 0 || !((b.dirichlet_mask >> mk) & 1)) continue;
      const int32_t *fv = &m.facet_verts[f * d];
      for (int a = 0; a < d; ++a) s.free[b.dof_off + fv[a]] = 0;
      if (b.order >= 2) {
        if (d == 2) {
          for (int q = 0; q < b.order - 1; ++q) s.free[b.dof_off + m.nv + (b.order - 1) * f + q] = 0;
        } else {
          const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
          for (auto &p : pr) {
            int64_t ei = find_edge(m, fv[p[0]], fv[p[1]]);
            for (int q = 0; q < b.order - 1; ++q) s.free[b.dof_off + m.nv + (b.order - 1) * ei + q] = 0;
          }
        }
      }
      if (b.order >= 3 && d == 3) s.free[b.dof_off + m.nv + 2 * m.nedge + f] = 0;
    }
  }
}

// ------------------------------------------------------------------ pattern + schedule

void build_pattern(Matrix *A) {
  Space &s = *A->space;
  Mesh &m = *s.mesh;
  const int d = m.dim, nvl = d + 1;
  bool all_l2 = true, all_h1 = true;
  for (auto &b : s.blocks) {
    all_l2 &= b.kind == OCMP_L2;
    all_h1 &= b.kind == OCMP_H1;
  }
  OCMP_REQUIRE(all_l2 || all_h1, "compound spaces mixing H1 and L2 blocks are not supported");
  A->n = s.ndof;
  if (all_l2) {
    OCMP_REQUIRE(s.blocks.size() == 1, "DG path supports a single L2 block");
    const int nb = s.blocks[0].nloc;
    A->dg = true;
    A->dg_nb = nb;
    std::vector<int8_t> pos((size_t)m.nel * (nvl + 1), -1);
    std::vector<int32_t> bcol((size_t)m.nel * 4, -1);
    A->row_ptr.assign(A->n + 1, 0);
    for (int64_t e = 0; e < m.nel; ++e) {
      int32_t blist[4];
      int nbk = 0;
      blist[nbk++] = (int32_t)e;
      if (s.dgjumps)
        for (int k = 0; k < nvl; ++k)
          if (m.elem_nb[e * nvl + k] >= 0) blist[nbk++] = m.elem_nb[e * nvl + k];
      std::sort(blist, blist + nbk);
      for (int t = 0; t < nbk; ++t) {
        bcol[e * 4 + t] = blist[t];
        if (blist[t] == e) pos[e * (nvl + 1)] = (int8_t)t;
        for (int k = 0; k < nvl; ++k)
          if (s.dgjumps && m.elem_nb[e * nvl + k] == blist[t]) pos[e * (nvl + 1) + 1 + k] = (int8_t)t;
      }
      for (int i = 0; i < nb; ++i) A->row_ptr[e * nb + i + 1] = nbk * nb;
    }
    for (int64_t r = 0; r < A->n; ++r) A->row_ptr[r + 1] += A->row_ptr[r];
    A->nnz = A->row_ptr[A->n];
    A->col.resize(A->nnz);
    for (int64_t e = 0; e < m.nel; ++e)
      for (int i = 0; i < nb; ++i) {
        int64_t p = A->row_ptr[e * nb + i];
        for (int t = 0; t < 4 && bcol[e * 4 + t] >= 0; ++t)
          for (int j = 0; j < nb; ++j) A->col[p++] = bcol[e * 4 + t] * nb + j;
      }
    A->d_dg_pos.upload(pos, A->ctx->stream);
    A->d_dg_bcol.upload(bcol, A->ctx->stream);
  } else {
    A->dg = false;
    const int nl = s.nloc;
    // dof -> elements incidence
    std::vector<int64_t> inc_ptr(s.ndof + 1, 0);
    for (int64_t t = 0; t < m.nel * (int64_t)nl; ++t) inc_ptr[s.edofs[t] + 1]++;
    for (int64_t r = 0; r < s.ndof; ++r) inc_ptr[r + 1] += inc_ptr[r];
    std::vector<int32_t> inc(inc_ptr[s.ndof]);
    {
      std::vector<int64_t> fill(inc_ptr.begin(), inc_ptr.end() - 1);
      for (int64_t e = 0; e < m.nel; ++e)
        for (int i = 0; i < nl; ++i) inc[fill[s.edofs[e * nl + i]]++] = (int32_t)e;
    }
    A->row_ptr.assign(s.ndof + 1, 0);
    auto row_cols = [&](int64_t r, std::vector<int32_t> &buf) {
      buf.clear();
      for (int64_t t = inc_ptr[r]; t < inc_ptr[r + 1]; ++t) {
        const int32_t *ed = &s.edofs[(int64_t)inc[t] * nl];
        buf.insert(buf.end(), ed, ed + nl);
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
    };
#pragma omp parallel
    {
      std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 4096)
      for (int64_t r = 0; r < s.ndof; ++r) {
        row_cols(r, buf);
        A->row_ptr[r + 1] = (int32_t)buf.size();
      }
    }
    std::vector<int64_t> rp(s.ndof + 1, 0);
    for (int64_t r = 0; r < s.ndof; ++r) rp[r + 1] = rp[r] + A->row_ptr[r + 1];
    OCMP_REQUIRE(rp[s.ndof] < (int64_t)INT32_MAX, "nnz exceeds 32-bit CSR offsets");
    for (int64_t r = 0; r <= s.ndof; ++r) A->row_ptr[r] = (int32_t)rp[r];
    A->nnz = rp[s.ndof];
    A->col.resize(A->nnz);
#pragma omp parallel
    {
      std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 4096)
      for (int64_t r = 0; r < s.ndof; ++r) {
        row_cols(r, buf);
        std::copy(buf.begin(), buf.end(), A->col.begin() + A->row_ptr[r]);
      }
    }
  }
  hipStream_t st = A->ctx->stream;
  A->d_row_ptr.upload(A->row_ptr, st);
  A->d_col.upload(A->col, st);
  A->d_val.alloc(A->nnz);
  OCMP_HIP(hipMemsetAsync(A->d_val.p, 0, A->nnz * sizeof(double), st));
  build_row_blocks(A, A->rb64, SPMV_NNZB64);
}

void build_row_blocks(Matrix *A, RowBlocks &rb, int nnzb) {
  // greedy: extend a block while its nonzeros fit in nnzb; a row longer than that is a block of
  // its own (the kernel then loops over it in chunks and keeps global columns)
  rb.rblk.clear();
  rb.rblk.push_back(0);
  int64_t start_nz = 0;
  for (int64_t r = 0; r < A->n; ++r) {
    const int64_t end_nz = A->row_ptr[r + 1];
    if (end_nz - start_nz > nnzb && rb.rblk.back() != r) {
      rb.rblk.push_back((int32_t)r);
      start_nz = A->row_ptr[r];
    }
  }
  rb.rblk.push_back((int32_t)A->n);
  rb.nrblk = (int64_t)rb.rblk.size() - 1;
  rb.d_rblk.upload(rb.rblk, A->ctx->stream);
  // per block: its distinct columns, sorted, and for every nonzero the 16-bit position of its
  // column in that list; the SpMV gathers x once per distinct column instead of once per nonzero
  const int64_t nb = rb.nrblk;
  std::vector<int32_t> nu(nb + 1, 0);
  std::vector<uint16_t> lcol(A->nnz);
  std::vector<std::vector<int32_t>> ub(nb);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 256)
    for (int64_t b = 0; b < nb; ++b) {
      const int64_t j0 = A->row_ptr[rb.rblk[b]], j1 = A->row_ptr[rb.rblk[b + 1]];
      if (j1 - j0 > nnzb) continue;
      buf.assign(A->col.begin() + j0, A->col.begin() + j1);
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      for (int64_t j = j0; j < j1; ++j)
        lcol[j] = (uint16_t)(std::lower_bound(buf.begin(), buf.end(), A->col[j]) - buf.begin());
      nu[b + 1] = (int32_t)buf.size();
      ub[b] = buf;
    }
  }
  for (int64_t b = 0; b < nb; ++b) nu[b + 1] += nu[b];
  std::vector<int32_t> ucol(nu[nb]);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nb; ++b) std::copy(ub[b].begin(), ub[b].end(), ucol.begin() + nu[b]);
  rb.nucol = nu[nb];
  rb.d_ubase.upload(nu, A->ctx->stream);
  rb.d_ucol.upload(ucol, A->ctx->stream);
  rb.d_lcol.upload(lcol, A->ctx->stream);
  rb.nnzb = nnzb;
}

// The Krylov operator of A: the entries whose (row block, column block) pair is active, with their
// slots in A's values (the values are gathered by the solver when stale). Its rows are A's rows in
// the Z-order of their dofs' positions (the mean centroid of the elements holding the dof), so a
// row block of consecutive operator rows is a compact patch of the mesh whose rows share most of
// their columns; columns keep A's numbering and each row's result goes to its row of A.
void build_solver_operator(Matrix *A) {
  const Space &s = *A->space;
  const Mesh &m = *s.mesh;
  const int nb = (int)s.blocks.size();
  std::vector<int8_t> blk(A->n, 0);
  for (int k = 0; k < nb; ++k)
    for (int64_t i = 0; i < s.blocks[k].n1; ++i) blk[s.blocks[k].dof_off + i] = (int8_t)k;
  // dof positions and their Morton keys
  const int d = m.dim, nvl = d + 1, nl = s.nloc;
  std::vector<double> pos((size_t)A->n * d, 0.0);
  std::vector<int32_t> cnt(A->n, 0);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t e = 0; e < m.nel; ++e) {
    double c[3] = {0, 0, 0};
    for (int k = 0; k < nvl; ++k)
      for (int a = 0; a < d; ++a) c[a] += m.coords[(size_t)m.elems[e * nvl + k] * d + a] / nvl;
    for (int i = 0; i < nl; ++i) {
      const int32_t r = s.edofs[e * nl + i];
      for (int a = 0; a < d; ++a) pos[(size_t)r * d + a] += c[a];
      ++cnt[r];
    }
  }
  for (int64_t r = 0; r < A->n; ++r)
    for (int a = 0; a < d; ++a) {
      double &p = pos[(size_t)r * d + a];
      p = cnt[r] ? p / cnt[r] : 0.0;
      lo[a] = std::min(lo[a], p);
      hi[a] = std::max(hi[a], p);
    }
  const int bits = d == 2 ? 31 : 21;
  std::vector<uint64_t> key(A->n, 0);
  for (int64_t r = 0; r < A->n; ++r) {
    uint64_t q[3] = {0, 0, 0};
    for (int a = 0; a < d; ++a) {
      const double t = hi[a] > lo[a] ? (pos[(size_t)r * d + a] - lo[a]) / (hi[a] - lo[a]) : 0.0;
      q[a] = (uint64_t)std::min(t * (double)((1ull << bits) - 1), (double)((1ull << bits) - 1));
    }
    uint64_t z = 0;
    for (int bit = bits - 1; bit >= 0; --bit)
      for (int a = 0; a < d; ++a) z = (z << 1) | ((q[a] >> bit) & 1);
    key[r] = z;
  }
  std::vector<int32_t> order(A->n);
  for (int64_t r = 0; r < A->n; ++r) order[r] = (int32_t)r;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  auto *M = new Matrix();
  M->space = A->space;
  M->ctx = A->ctx;
  M->n = A->n;
  M->row_ptr.assign(A->n + 1, 0);
  std::vector<int32_t> src;
  for (int64_t i = 0; i < A->n; ++i) {
    const int32_t r = order[i];
    for (int32_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; ++k)
      if (A->active_pairs >> (blk[r] * nb + blk[A->col[k]]) & 1) {
        M->col.push_back(A->col[k]);
        src.push_back(k);
      }
    M->row_ptr[i + 1] = (int32_t)M->col.size();
  }
  M->nnz = (int64_t)M->col.size();
  hipStream_t st = A->ctx->stream;
  M->d_row_ptr.upload(M->row_ptr, st);
  M->d_col.upload(M->col.empty() ? std::vector<int32_t>(1, 0) : M->col, st);
  M->d_src.upload(src.empty() ? std::vector<int32_t>(1, 0) : src, st);
  M->d_rowmap.upload(order, st);
  M->d_val.alloc(std::max<int64_t>(M->nnz, 1));
  build_row_blocks(M, M->rb64, SPMV_NNZB64);
  M->assembled = true;
  OCMP_HIP(hipStreamSynchronize(st));
  delete A->op;
  A->op = M;
  A->op_pairs = A->active_pairs;
  A->op_version = A->op_version32 = ~0ull;
}

void ensure_diag_slots(Matrix *A) {
  if (A->d_dslot.p) return;
  std::vector<int32_t> d(A->n, -1);
  const int32_t *rp = A->row_ptr.data(), *col = A->col.data();
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < A->n; ++r) {
    const int32_t *q = std::lower_bound(col + rp[r], col + rp[r + 1], (int32_t)r);
    if (q != col + rp[r + 1] && *q == r) d[r] = (int32_t)(q - col);
  }
  A->d_dslot.upload(d, A->ctx->stream);
  OCMP_HIP(hipStreamSynchronize(A->ctx->stream));
}

void ensure_row_blocks32(Matrix *A) {
  if (!A->rb32.nnzb) build_row_blocks(A, A->rb32, SPMV_NNZB32);
}

// the dof range [s0, s1) of the Schur blocks, which must be adjacent
static void schur_range(const Precond *P, int64_t &s0, int64_t &s1) {
  const auto &bl = P->A->space->blocks;
  s0 = -1;
  s1 = -1;
  for (size_t k = 0; k < bl.size(); ++k) {
    if (!(P->schur_blocks >> k & 1)) continue;
    if (s0 < 0) {
      s0 = bl[k].dof_off;
    } else {
      OCMP_REQUIRE(bl[k].dof_off == s1, "the Schur inner solve needs the Schur blocks to be adjacent");
    }
    s1 = bl[k].dof_off + bl[k].n1;
  }
  OCMP_REQUIRE(s0 >= 0 && s1 > s0, "no Schur rows");
}

void build_schur_matrix(Precond *P) {
  Matrix *A = P->A;
  int64_t s0, s1;
  schur_range(P, s0, s1);
  const int64_t n = A->n, ns = s1 - s0;
  const int32_t *rp = A->row_ptr.data(), *col = A->col.data();
  const uint8_t *freed = A->space->free.data();
  // per row: the CSR slots of its Schur columns (sorted columns, so one range)
  std::vector<int32_t> seg(2 * n);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    seg[2 * r] = (int32_t)(std::lower_bound(col + rp[r], col + rp[r + 1], (int32_t)s0) - col);
    seg[2 * r + 1] = (int32_t)(std::lower_bound(col + rp[r], col + rp[r + 1], (int32_t)s1) - col);
  }
  // row p of S: the Schur columns of row p (C) and of every row u that p couples to through a
  // free non-Schur column (B diag(A)^{-1} B^T)
  std::vector<std::vector<int32_t>> rows(ns);
#pragma omp parallel
  {
    std::vector<int32_t> buf;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t i = 0; i < ns; ++i) {
      const int64_t p = s0 + i;
      buf.clear();
      for (int32_t j = seg[2 * p]; j < seg[2 * p + 1]; ++j) buf.push_back(col[j] - (int32_t)s0);
      for (int pass = 0; pass < 2; ++pass) {
        const int32_t a = pass ? seg[2 * p + 1] : rp[p], e = pass ? rp[p + 1] : seg[2 * p];
        for (int32_t j = a; j < e; ++j) {
          const int32_t u = col[j];
          if (!freed[u]) continue;
          for (int32_t t = seg[2 * u]; t < seg[2 * u + 1]; ++t) buf.push_back(col[t] - (int32_t)s0);
        }
      }
      std::sort(buf.begin(), buf.end());
      buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
      rows[i] = buf;
    }
  }
  auto *S = new Matrix();
  S->space = A->space;
  S->ctx = A->ctx;
  S->n = ns;
  S->row_ptr.assign(ns + 1, 0);
  int64_t nnz = 0;
  for (int64_t i = 0; i < ns; ++i) {
    nnz += (int64_t)rows[i].size();
    OCMP_REQUIRE(nnz < (int64_t)INT32_MAX, "Schur matrix nnz exceeds 32-bit CSR offsets");
    S->row_ptr[i + 1] = (int32_t)nnz;
  }
  S->nnz = nnz;
  S->col.resize(nnz);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < ns; ++i) std::copy(rows[i].begin(), rows[i].end(), S->col.begin() + S->row_ptr[i]);
  hipStream_t st = A->ctx->stream;
  S->d_row_ptr.upload(S->row_ptr, st);
  S->d_col.upload(S->col, st);
  S->d_val.alloc(nnz);
  build_row_blocks(S, S->rb64, SPMV_NNZB64);
  build_row_blocks(S, P->srb, SCG_NNZB);
  S->assembled = true;
  P->d_sseg.upload(seg, st);
  // the product lists of S's values: for entry (p, q), the velocity columns u of row p in slot
  // order, each with the slot of column q in row u (one per u, since u's Schur columns are sorted)
  {
    std::vector<int64_t> rcnt(ns + 1, 0);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < ns; ++i) {
      const int64_t p = s0 + i;
      int64_t c = 0;
      for (int32_t j = rp[p]; j < rp[p + 1]; ++j) {
        const int32_t u = col[j];
        if (u >= s0 && u < s1) continue;
        if (freed[u]) c += seg[2 * u + 1] - seg[2 * u];
      }
      rcnt[i + 1] = c;
    }
    for (int64_t i = 0; i < ns; ++i) rcnt[i + 1] += rcnt[i];
    if (rcnt[ns] <= SCHUR_MAX_PAIRS) {
      std::vector<int32_t> cslot(nnz, -1), pptr(nnz + 1, 0), pj(rcnt[ns]), pt(rcnt[ns]);
#pragma omp parallel
      {
        std::vector<int32_t> cnt;
#pragma omp for schedule(dynamic, 1024)
        for (int64_t i = 0; i < ns; ++i) {
          const int64_t p = s0 + i;
          const int32_t k0 = S->row_ptr[i], len = S->row_ptr[i + 1] - k0;
          const int32_t *sc = S->col.data() + k0;
          auto pos = [&](int32_t q) { return (int32_t)(std::lower_bound(sc, sc + len, q) - sc); };
          for (int32_t j = seg[2 * p]; j < seg[2 * p + 1]; ++j) cslot[k0 + pos(col[j] - (int32_t)s0)] = j;
          cnt.assign(len, 0);
          auto each = [&](auto &&fn) {
            for (int pass = 0; pass < 2; ++pass) {
              const int32_t a = pass ? seg[2 * p + 1] : rp[p], e = pass ? rp[p + 1] : seg[2 * p];
              for (int32_t j = a; j < e; ++j) {
                const int32_t u = col[j];
                if (!freed[u]) continue;
                for (int32_t t = seg[2 * u]; t < seg[2 * u + 1]; ++t) fn(j, t, pos(col[t] - (int32_t)s0));
              }
            }
          };
          each([&](int32_t, int32_t, int32_t k) { ++cnt[k]; });
          // the row's lists start at the row's first pair; entry k's list after entries < k
          int64_t off = rcnt[i];
          std::vector<int64_t> at(len);
          for (int32_t k = 0; k < len; ++k) {
            at[k] = off;
            pptr[k0 + k] = (int32_t)off;
            off += cnt[k];
          }
          each([&](int32_t j, int32_t t, int32_t k) {
            pj[at[k]] = j;
            pt[at[k]] = t;
            ++at[k];
          });
        }
      }
      pptr[nnz] = (int32_t)rcnt[ns];
      P->d_scslot.upload(cslot, st);
      P->d_spptr.upload(pptr, st);
      P->d_spj.upload(pj, st);
      P->d_spt.upload(pt, st);
      OCMP_HIP(hipStreamSynchronize(st));  // the host vectors end with this scope
    }
  }
  OCMP_HIP(hipStreamSynchronize(st));
  P->s0 = s0;
  P->s1 = s1;
  P->S = S;
}

// The velocity block A_uu (rows and columns [0, s0)) and the coupling A_up (rows [0, s0), columns
// [s0, n) shifted to 0) of a system whose Schur rows are the last ones: each row's columns are
// sorted, so its A_uu part is a prefix and its A_up part the rest. Values are gathered through
// the slot arrays at every update.
// ------------------------------------------------------------------ AMG hierarchy on S (host)
//
// Smoothed aggregation (Vanek, Mandel and Brezina, Computing 56 (1996) 179-196): strength
// |a_ij| >= theta sqrt(|a_ii a_jj|); aggregates by the three greedy passes (roots whose strong
// neighbours are all free take them; the rest join a neighbouring first-pass aggregate; leftovers
// form their own); tentative prolongation = aggregate indicator, smoothed once by damped Jacobi,
// P = (I - omega D^-1 A) P_t with omega = 4 / (3 lambda_max(D^-1 A)); Galerkin A_c = P^T A P.
// Built once with the first values of S; the device recomputes every value from the lists at each
// update, so the aggregates and patterns stay while the values follow the operator.

namespace {

constexpr double AMG_THETA = 0.08;
constexpr int64_t AMG_COARSE = 48;    // dense inverse (in LDS) on the last level at or below this many rows
constexpr int64_t AMG_DENSE_LIMIT = 1200;  // the largest last level accepted when coarsening stalls
constexpr int AMG_MAX_LEVELS = 10;

struct HCsr {
  int64_t n = 0;
  std::vector<int32_t> rp, col;
  std::vector<double> val;
};

Matrix *amg_matrix(Matrix *like, const HCsr &h) {
  auto *M = new Matrix();
  M->space = like->space;
  M->ctx = like->ctx;
  M->n = h.n;
  M->row_ptr = h.rp;
  M->col = h.col;
  M->nnz = (int64_t)h.col.size();
  hipStream_t st = like->ctx->stream;
  M->d_row_ptr.upload(M->row_ptr, st);
  M->d_col.upload(M->col, st);
  M->d_val.alloc(std::max<int64_t>(M->nnz, 1));
  build_row_blocks(M, M->rb64, SPMV_NNZB64);
  M->assembled = true;
  return M;
}

// largest eigenvalue of D^-1 A (A symmetric, D its diagonal) on the usable rows, by power iteration
double amg_lambda_max(const HCsr &A, const std::vector<double> &dg, const std::vector<uint8_t> &use) {
  const int64_t n = A.n;
  std::vector<double> x(n), y(n);
  for (int64_t i = 0; i < n; ++i) x[i] = use[i] ? 1.0 + 0.5 * std::sin(0.7 * (double)i) : 0.0;
  double lam = 1.0;
  for (int it = 0; it < 30; ++it) {
    double nn = 0.0, dot = 0.0;
#pragma omp parallel for reduction(+ : nn, dot) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      if (use[i] && dg[i] != 0.0)
        for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
          if (use[A.col[k]]) s += A.val[k] * x[A.col[k]];
      y[i] = use[i] && dg[i] != 0.0 ? s / dg[i] : 0.0;
      nn += y[i] * y[i];
      dot += y[i] * x[i];
    }
    double xx = 0.0;
    for (int64_t i = 0; i < n; ++i) xx += x[i] * x[i];
    lam = xx > 0.0 ? dot / xx : 1.0;
    const double inv = nn > 0.0 ? 1.0 / std::sqrt(nn) : 0.0;
    for (int64_t i = 0; i < n; ++i) x[i] = y[i] * inv;
  }
  return std::max(lam, 1e-300);
}

}  // namespace

Amg::~Amg() {
  for (auto &L : lv) {
    if (L->own_A) delete L->A;
    delete L->P;
    delete L->W;
    delete L->Wt;
  }
}

Precond::~Precond() {
  delete S;
  delete Auu;
  delete Aup;
  delete amg;
}

void build_schur_amg(Precond *Pc, const std::vector<double> &sval, const std::vector<double> &sdinv) {
  Matrix *S = Pc->S;
  hipStream_t st = S->ctx->stream;
  auto *amg = new Amg();
  HCsr A;
  A.n = S->n;
  A.rp = S->row_ptr;
  A.col = S->col;
  A.val = sval;
  std::vector<uint8_t> use(A.n);
  for (int64_t i = 0; i < A.n; ++i) use[i] = sdinv[i] != 0.0;
  Matrix *Adev = S;
  for (int l = 0;; ++l) {
    auto Lp = std::make_unique<AmgLevel>();
    AmgLevel &L = *Lp;
    L.A = Adev;
    L.own_A = l > 0;
    const int64_t n = A.n;
    if (n <= AMG_COARSE || l == AMG_MAX_LEVELS - 1) {
      OCMP_REQUIRE(n <= AMG_DENSE_LIMIT, "AMG: the coarsest level is too large for its dense inverse");
      amg->lv.push_back(std::move(Lp));
      amg->ncoarse = n;
      break;
    }
    std::vector<double> dg(n, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      const int32_t *c0 = A.col.data() + A.rp[i], *c1 = A.col.data() + A.rp[i + 1];
      const int32_t *q = std::lower_bound(c0, c1, (int32_t)i);
      if (q != c1 && *q == i) dg[i] = A.val[q - A.col.data()];
    }
    auto strong = [&](int64_t i, int32_t k) {
      const int32_t j = A.col[k];
      return j != i && use[j] && std::fabs(A.val[k]) >= AMG_THETA * std::sqrt(std::fabs(dg[i] * dg[j]));
    };
    std::vector<int32_t> agg(n, -1);
    int32_t na = 0;
    for (int64_t i = 0; i < n; ++i) {  // pass 1: roots with all strong neighbours free
      if (!use[i] || agg[i] >= 0) continue;
      bool ok = true;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1] && ok; ++k)
        if (strong(i, k) && agg[A.col[k]] >= 0) ok = false;
      if (!ok) continue;
      agg[i] = na;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (strong(i, k)) agg[A.col[k]] = na;
      ++na;
    }
    const std::vector<int32_t> agg1 = agg;
    for (int64_t i = 0; i < n; ++i) {  // pass 2: join a strongly connected first-pass aggregate
      if (!use[i] || agg[i] >= 0) continue;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (strong(i, k) && agg1[A.col[k]] >= 0) {
          agg[i] = agg1[A.col[k]];
          break;
        }
    }
    for (int64_t i = 0; i < n; ++i) {  // pass 3: the rest form aggregates of their own
      if (!use[i] || agg[i] >= 0) continue;
      agg[i] = na;
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
        if (strong(i, k) && agg[A.col[k]] < 0) agg[A.col[k]] = na;
      ++na;
    }
    if (na == 0 || na > (int64_t)(0.9 * (double)n)) {
      OCMP_REQUIRE(n <= AMG_DENSE_LIMIT, "AMG: coarsening stalled above the dense-inverse size");
      amg->lv.push_back(std::move(Lp));
      amg->ncoarse = n;
      break;
    }
    const double omega = 4.0 / (3.0 * amg_lambda_max(A, dg, use));
    L.omega = omega;
    // P: row i (usable) has its own aggregate and its usable neighbours' aggregates; each entry
    // sums the A entries of row i whose column lies in that aggregate
    HCsr Ph;
    Ph.n = n;
    Ph.rp.assign(n + 1, 0);
    std::vector<std::vector<int32_t>> pc(n), pl(n), pll(n);  // columns, list lengths, A slots
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      if (!use[i]) continue;
      std::vector<std::pair<int32_t, int32_t>> t;  // (aggregate, A slot)
      t.emplace_back(agg[i], -1);
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.col[k];
        if (use[j] && agg[j] >= 0) t.emplace_back(agg[j], k);
      }
      std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
      for (size_t a = 0; a < t.size();) {
        size_t b = a;
        int32_t cnt = 0;
        while (b < t.size() && t[b].first == t[a].first) {
          if (t[b].second >= 0) {
            pll[i].push_back(t[b].second);
            ++cnt;
          }
          ++b;
        }
        pc[i].push_back(t[a].first);
        pl[i].push_back(cnt);
        a = b;
      }
    }
    for (int64_t i = 0; i < n; ++i) Ph.rp[i + 1] = Ph.rp[i] + (int32_t)pc[i].size();
    const int64_t np = Ph.rp[n];
    Ph.col.resize(np);
    Ph.val.resize(np);
    std::vector<int32_t> prow(np), plptr(np + 1, 0), plist;
    std::vector<uint8_t> pident(np);
    for (int64_t i = 0; i < n; ++i)
      for (size_t e = 0; e < pc[i].size(); ++e) plptr[Ph.rp[i] + e + 1] = pl[i][e];
    for (int64_t e = 0; e < np; ++e) plptr[e + 1] += plptr[e];
    plist.resize(plptr[np]);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      size_t off = 0;
      for (size_t e = 0; e < pc[i].size(); ++e) {
        const int64_t pe = Ph.rp[i] + e;
        Ph.col[pe] = pc[i][e];
        prow[pe] = (int32_t)i;
        pident[pe] = pc[i][e] == agg[i] ? 1 : 0;
        double sum = 0.0;
        for (int32_t t = 0; t < pl[i][e]; ++t) {
          const int32_t k = pll[i][off + t];
          plist[plptr[pe] + t] = k;
          sum += A.val[k];
        }
        off += pl[i][e];
        Ph.val[pe] = (double)pident[pe] - (dg[i] != 0.0 ? omega / dg[i] : 0.0) * sum;
      }
    }
    std::vector<std::vector<int32_t>>().swap(pc);
    std::vector<std::vector<int32_t>>().swap(pl);
    std::vector<std::vector<int32_t>>().swap(pll);
    // Q = A P: entry (i, J) sums A[i,j] P[j,J] over the usable columns j of row i
    HCsr Qh;
    Qh.n = n;
    Qh.rp.assign(n + 1, 0);
    std::vector<std::vector<int32_t>> qc(n), qn(n), qa(n), qp(n);
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t i = 0; i < n; ++i) {
      std::vector<std::array<int32_t, 3>> t;  // (J, A slot, P slot)
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const int32_t j = A.col[k];
        for (int32_t e = Ph.rp[j]; e < Ph.rp[j + 1]; ++e) t.push_back({Ph.col[e], k, e});
      }
      std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a[0] < b[0]; });
      for (size_t a = 0; a < t.size();) {
        size_t b = a;
        while (b < t.size() && t[b][0] == t[a][0]) {
          qa[i].push_back(t[b][1]);
          qp[i].push_back(t[b][2]);
          ++b;
        }
        qc[i].push_back(t[a][0]);
        qn[i].push_back((int32_t)(b - a));
        a = b;
      }
    }
    for (int64_t i = 0; i < n; ++i) Qh.rp[i + 1] = Qh.rp[i] + (int32_t)qc[i].size();
    const int64_t nq = Qh.rp[n];
    Qh.col.resize(nq);
    Qh.val.assign(nq, 0.0);
    std::vector<int64_t> qptr(nq + 1, 0);
    for (int64_t i = 0; i < n; ++i)
      for (size_t e = 0; e < qc[i].size(); ++e) qptr[Qh.rp[i] + e + 1] = qn[i][e];
    for (int64_t e = 0; e < nq; ++e) qptr[e + 1] += qptr[e];
    OCMP_REQUIRE(qptr[nq] < (int64_t)INT32_MAX, "AMG: Q pair lists exceed 32-bit offsets");
    std::vector<int32_t> qptr32(nq + 1), qa_all(qptr[nq]), qp_all(qptr[nq]);
    for (int64_t e = 0; e <= nq; ++e) qptr32[e] = (int32_t)qptr[e];
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < n; ++i) {
      size_t off = 0;
      for (size_t e = 0; e < qc[i].size(); ++e) {
        const int64_t qe = Qh.rp[i] + e;
        Qh.col[qe] = qc[i][e];
        double sum = 0.0;
        for (int32_t t = 0; t < qn[i][e]; ++t) {
          qa_all[qptr[qe] + t] = qa[i][off + t];
          qp_all[qptr[qe] + t] = qp[i][off + t];
          sum += A.val[qa[i][off + t]] * Ph.val[qp[i][off + t]];
        }
        off += qn[i][e];
        Qh.val[qe] = sum;
      }
    }
    std::vector<std::vector<int32_t>>().swap(qc);
    std::vector<std::vector<int32_t>>().swap(qn);
    std::vector<std::vector<int32_t>>().swap(qa);
    std::vector<std::vector<int32_t>>().swap(qp);
    // P's entries by column (counting sort): the fine rows of each aggregate with their P slots
    std::vector<int32_t> tptr(na + 1, 0), trow(np), tslot(np);
    for (int64_t e = 0; e < np; ++e) tptr[Ph.col[e] + 1]++;
    for (int32_t I = 0; I < na; ++I) tptr[I + 1] += tptr[I];
    {
      std::vector<int32_t> fill(tptr.begin(), tptr.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        for (int32_t e = Ph.rp[i]; e < Ph.rp[i + 1]; ++e) {
          const int32_t k = fill[Ph.col[e]]++;
          trow[k] = (int32_t)i;
          tslot[k] = e;
        }
    }
    // A_c = P^T Q: entry (I, J) sums P[i,I] Q[i,J] over the fine rows i of aggregate I
    HCsr C;
    C.n = na;
    C.rp.assign(na + 1, 0);
    std::vector<std::vector<int32_t>> cc(na), cn(na), cp(na), cq(na);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t I = 0; I < na; ++I) {
      std::vector<std::array<int32_t, 3>> t;  // (J, P slot, Q slot)
      for (int32_t k = tptr[I]; k < tptr[I + 1]; ++k) {
        const int32_t i = trow[k], e = tslot[k];
        for (int32_t q = Qh.rp[i]; q < Qh.rp[i + 1]; ++q) t.push_back({Qh.col[q], e, q});
      }
      std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return a[0] < b[0]; });
      for (size_t a = 0; a < t.size();) {
        size_t b = a;
        while (b < t.size() && t[b][0] == t[a][0]) {
          cp[I].push_back(t[b][1]);
          cq[I].push_back(t[b][2]);
          ++b;
        }
        cc[I].push_back(t[a][0]);
        cn[I].push_back((int32_t)(b - a));
        a = b;
      }
    }
    for (int64_t I = 0; I < na; ++I) C.rp[I + 1] = C.rp[I] + (int32_t)cc[I].size();
    const int64_t nc = C.rp[na];
    C.col.resize(nc);
    C.val.assign(nc, 0.0);
    std::vector<int64_t> cptr(nc + 1, 0);
    for (int64_t I = 0; I < na; ++I)
      for (size_t e = 0; e < cc[I].size(); ++e) cptr[C.rp[I] + e + 1] = cn[I][e];
    for (int64_t e = 0; e < nc; ++e) cptr[e + 1] += cptr[e];
    OCMP_REQUIRE(cptr[nc] < (int64_t)INT32_MAX, "AMG: Galerkin pair lists exceed 32-bit offsets");
    std::vector<int32_t> cptr32(nc + 1), cp_all(cptr[nc]), cq_all(cptr[nc]);
    for (int64_t e = 0; e <= nc; ++e) cptr32[e] = (int32_t)cptr[e];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t I = 0; I < na; ++I) {
      size_t off = 0;
      for (size_t e = 0; e < cc[I].size(); ++e) {
        const int64_t ce = C.rp[I] + e;
        C.col[ce] = cc[I][e];
        double sum = 0.0;
        for (int32_t t = 0; t < cn[I][e]; ++t) {
          cp_all[cptr[ce] + t] = cp[I][off + t];
          cq_all[cptr[ce] + t] = cq[I][off + t];
          sum += Ph.val[cp[I][off + t]] * Qh.val[cq[I][off + t]];
        }
        off += cn[I][e];
        C.val[ce] = sum;
      }
    }
    // W = P - w D^-1 Q on the union of P's and Q's patterns (both rows sorted by column), and W^T
    HCsr Wh;
    Wh.n = n;
    Wh.rp.assign(n + 1, 0);
    std::vector<int32_t> wrow, wps, wqs;
    for (int64_t i = 0; i < n; ++i) {
      int32_t a = Ph.rp[i], b = Qh.rp[i];
      while (a < Ph.rp[i + 1] || b < Qh.rp[i + 1]) {
        const int32_t ca = a < Ph.rp[i + 1] ? Ph.col[a] : INT32_MAX, cb = b < Qh.rp[i + 1] ? Qh.col[b] : INT32_MAX;
        const int32_t c = std::min(ca, cb);
        Wh.col.push_back(c);
        wrow.push_back((int32_t)i);
        wps.push_back(ca == c ? a++ : -1);
        wqs.push_back(cb == c ? b++ : -1);
      }
      Wh.rp[i + 1] = (int32_t)Wh.col.size();
    }
    const int64_t nw = (int64_t)Wh.col.size();
    Wh.val.assign(nw, 0.0);
    HCsr Wth;
    Wth.n = na;
    Wth.rp.assign(na + 1, 0);
    for (int64_t e = 0; e < nw; ++e) Wth.rp[Wh.col[e] + 1]++;
    for (int32_t I = 0; I < na; ++I) Wth.rp[I + 1] += Wth.rp[I];
    Wth.col.resize(nw);
    std::vector<int32_t> wtslot(nw);
    {
      std::vector<int32_t> fill(Wth.rp.begin(), Wth.rp.end() - 1);
      for (int64_t i = 0; i < n; ++i)
        for (int32_t e = Wh.rp[i]; e < Wh.rp[i + 1]; ++e) {
          const int32_t k = fill[Wh.col[e]]++;
          Wth.col[k] = (int32_t)i;
          wtslot[k] = e;
        }
    }
    // device copies of the level's structure
    L.P = amg_matrix(S, Ph);
    L.W = amg_matrix(S, Wh);
    L.Wt = amg_matrix(S, Wth);
    L.d_wrow.upload(wrow.empty() ? std::vector<int32_t>(1, 0) : wrow, st);
    L.d_wps.upload(wps.empty() ? std::vector<int32_t>(1, 0) : wps, st);
    L.d_wqs.upload(wqs.empty() ? std::vector<int32_t>(1, 0) : wqs, st);
    L.d_wtslot.upload(wtslot.empty() ? std::vector<int32_t>(1, 0) : wtslot, st);
    L.nq = nq;
    L.d_prow.upload(prow, st);
    L.d_plptr.upload(plptr, st);
    L.d_plist.upload(plist.empty() ? std::vector<int32_t>(1, 0) : plist, st);
    L.d_pident.upload(pident, st);
    L.d_qptr.upload(qptr32, st);
    L.d_qa.upload(qa_all.empty() ? std::vector<int32_t>(1, 0) : qa_all, st);
    L.d_qp.upload(qp_all.empty() ? std::vector<int32_t>(1, 0) : qp_all, st);
    L.d_cptr.upload(cptr32, st);
    L.d_cp.upload(cp_all.empty() ? std::vector<int32_t>(1, 0) : cp_all, st);
    L.d_cq.upload(cq_all.empty() ? std::vector<int32_t>(1, 0) : cq_all, st);
    L.q.alloc(std::max<int64_t>(nq, 1));
    OCMP_HIP(hipStreamSynchronize(st));  // the host vectors end with this iteration
    amg->lv.push_back(std::move(Lp));
    Adev = amg_matrix(S, C);
    A = std::move(C);
    use.assign(A.n, 1);
  }
  Pc->amg = amg;
}

void build_velocity_blocks(Precond *P) {
  Matrix *A = P->A;
  int64_t s0, s1;
  schur_range(P, s0, s1);
  const int64_t n = A->n;
  OCMP_REQUIRE(s1 == n, "the block-triangular step needs the Schur blocks last");
  P->s0 = s0;
  P->s1 = s1;
  const int32_t *rp = A->row_ptr.data(), *col = A->col.data();
  std::vector<int32_t> cut(s0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < s0; ++r)
    cut[r] = (int32_t)(std::lower_bound(col + rp[r], col + rp[r + 1], (int32_t)s0) - col);
  // entries of block pairs no term of the last assembly touched are structural zeros: left out
  const Space &sp = *A->space;
  const int nb = (int)sp.blocks.size();
  const bool known = nb <= 8 && A->active_pairs != 0;
  std::vector<int8_t> blk(n, 0);
  for (int k = 0; k < nb; ++k)
    for (int64_t i = 0; i < sp.blocks[k].n1; ++i) blk[sp.blocks[k].dof_off + i] = (int8_t)k;
  auto live = [&](int64_t r, int32_t c) { return !known || (A->active_pairs >> (blk[r] * nb + blk[c]) & 1); };
  hipStream_t st = A->ctx->stream;
  auto make = [&](bool uu, std::vector<int32_t> &slot) {
    auto *M = new Matrix();
    M->space = A->space;
    M->ctx = A->ctx;
    M->n = s0;
    M->row_ptr.assign(s0 + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s0; ++r) {
      int32_t c = 0;
      for (int32_t j = uu ? rp[r] : cut[r]; j < (uu ? cut[r] : rp[r + 1]); ++j) c += live(r, col[j]) ? 1 : 0;
      M->row_ptr[r + 1] = c;
    }
    for (int64_t r = 0; r < s0; ++r) M->row_ptr[r + 1] += M->row_ptr[r];
    M->nnz = M->row_ptr[s0];
    M->col.resize(M->nnz);
    slot.resize(M->nnz);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < s0; ++r) {
      const int32_t a = uu ? rp[r] : cut[r], e = uu ? cut[r] : rp[r + 1];
      int32_t k = M->row_ptr[r];
      for (int32_t j = a; j < e; ++j) {
        if (!live(r, col[j])) continue;
        M->col[k] = uu ? col[j] : col[j] - (int32_t)s0;
        slot[k] = j;
        ++k;
      }
    }
    M->d_row_ptr.upload(M->row_ptr, st);
    M->d_col.upload(M->col, st);
    M->d_val.alloc(M->nnz);
    build_row_blocks(M, M->rb64, SPMV_NNZB64);
    M->assembled = true;
    return M;
  };
  std::vector<int32_t> uus, ups;
  Matrix *Auu = make(true, uus);
  Matrix *Aup = make(false, ups);
  build_row_blocks(Auu, P->vrb, SCG_NNZB);
  P->d_uuslot.upload(uus, st);
  P->d_upslot.upload(ups, st);
  OCMP_HIP(hipStreamSynchronize(st));
  P->Auu = Auu;
  P->Aup = Aup;
  P->vel_pairs = A->active_pairs;
}

void build_colouring(Space &s) {
  Mesh &m = *s.mesh;
  const int d = m.dim, nvl = d + 1;
  // greedy colouring: elements sharing a vertex get different colours
  std::vector<int64_t> vptr(m.nv + 1, 0);
  for (int64_t t = 0; t < m.nel * nvl; ++t) vptr[m.elems[t] + 1]++;
  for (int64_t v = 0; v < m.nv; ++v) vptr[v + 1] += vptr[v];
  std::vector<int32_t> vel(vptr[m.nv]);
  {
    std::vector<int64_t> fill(vptr.begin(), vptr.end() - 1);
    for (int64_t e = 0; e < m.nel; ++e)
      for (int k = 0; k < nvl; ++k) vel[fill[m.elems[e * nvl + k]]++] = (int32_t)e;
  }
  std::vector<int16_t> colour(m.nel, -1);
  int ncol = 0;
  for (int64_t e = 0; e < m.nel; ++e) {
    uint64_t used[4] = {0, 0, 0, 0};
    for (int k = 0; k < nvl; ++k) {
      int32_t v = m.elems[e * nvl + k];
      for (int64_t t = vptr[v]; t < vptr[v + 1]; ++t) {
        int c = colour[vel[t]];
        if (c >= 0) used[c >> 6] |= 1ull << (c & 63);
      }
    }
    int c = 0;
    while (c < 256 && ((used[c >> 6] >> (c & 63)) & 1)) ++c;
    OCMP_REQUIRE(c < 256, "colouring needs more than 256 colours");
    colour[e] = (int16_t)c;
    ncol = std::max(ncol, c + 1);
  }
  s.ncolours = ncol;
  s.colour_off.assign(ncol + 1, 0);
  for (int64_t e = 0; e < m.nel; ++e) s.colour_off[colour[e] + 1]++;
  for (int c = 0; c < ncol; ++c) s.colour_off[c + 1] += s.colour_off[c];
  std::vector<int32_t> order(m.nel);
  {
    std::vector<int64_t> fill(s.colour_off.begin(), s.colour_off.end() - 1);
    for (int64_t e = 0; e < m.nel; ++e) order[fill[colour[e]]++] = (int32_t)e;
  }
  s.d_colour_elems.upload(order, m.ctx->stream);
  // the elements of each colour with a boundary facet, same order (boundary-facet terms)
  const int nf = m.dim + 1;
  std::vector<int32_t> bnd;
  s.bnd_colour_off.assign(ncol + 1, 0);
  for (int c = 0; c < ncol; ++c) {
    for (int64_t i = s.colour_off[c]; i < s.colour_off[c + 1]; ++i) {
      const int32_t e = order[i];
      bool b = false;
      for (int k = 0; k < nf; ++k) b |= m.elem_nb[(size_t)e * nf + k] < 0;
      if (b) bnd.push_back(e);
    }
    s.bnd_colour_off[c + 1] = (int64_t)bnd.size();
  }
  s.d_bnd_colour_elems.upload(bnd, m.ctx->stream);
  OCMP_HIP(hipStreamSynchronize(m.ctx->stream));
}

// Row-owner gather structures of an H1 matrix: for every row r the entries (element e, local row
// i) with edofs[e][i] == r, in increasing e, and for each entry the position within row r of every
// column edofs[e][j]. Built once per pattern (OpenMP over rows for the positions).
void build_h1_gather(Matrix *A) {
  Space &s = *A->space;
  Mesh &m = *s.mesh;
  const int nl = s.nloc;
  OCMP_REQUIRE(nl <= 64, "the row gather takes at most 64 local dofs per element");
  const int64_t n = A->n;
  // row order: every row's entries (element e, local row i) with edofs[e][i] == r, increasing e
  std::vector<int32_t> pptr(n + 1, 0);
  for (int64_t t = 0; t < m.nel * nl; ++t) pptr[s.edofs[t] + 1]++;
  for (int64_t r = 0; r < n; ++r) pptr[r + 1] += pptr[r];
  const int64_t np = pptr[n];
  OCMP_REQUIRE(np == m.nel * nl && np < INT32_MAX, "patch count mismatch");
  std::vector<int32_t> pe(np);
  std::vector<uint8_t> pli(np);
  {
    std::vector<int32_t> fill(pptr.begin(), pptr.end() - 1);
    for (int64_t e = 0; e < m.nel; ++e)
      for (int i = 0; i < nl; ++i) {
        const int32_t r = s.edofs[e * nl + i];
        pe[fill[r]] = (int32_t)e;
        pli[fill[r]] = (uint8_t)i;
        fill[r]++;
      }
  }
  int maxrow = 0, maxent = 0;
  for (int64_t r = 0; r < n; ++r) {
    maxrow = std::max(maxrow, A->row_ptr[r + 1] - A->row_ptr[r]);
    maxent = std::max(maxent, pptr[r + 1] - pptr[r]);
  }
  OCMP_REQUIRE(maxrow <= 65536, "a row with more than 65536 nonzeros");
  const bool narrow = maxrow <= 256;
  const int sb = narrow ? 1 : 2;
  // row blocks: staged values (entries x nl doubles + positions) of about 28 KB and at most 2048
  // nonzeros per block, greedily over consecutive rows; a row that alone exceeds that gets a block
  // of its own and the buffers grow to it, up to 48 KB of LDS in all
  const int emax = std::max(std::max(1, 12288 / (nl * (8 + sb) + 2)), maxent);
  const int accmax = std::max(1024, maxrow);
  A->gblk.clear();
  A->maxrow = maxrow;
  A->gemax = emax;
  A->gaccmax = accmax;
  A->npatch = np;
  if ((size_t)emax * (nl * (8 + sb) + 2) + (size_t)accmax * 8 > 49152 || emax > 65535) return;  // colour path
  A->gblk.push_back(0);
  for (int64_t r = 0; r < n;) {
    int64_t r1 = r + 1;
    while (r1 < n && pptr[r1 + 1] - pptr[r] <= emax && A->row_ptr[r1 + 1] - A->row_ptr[r] <= accmax) ++r1;
    A->gblk.push_back((int32_t)r1);
    r = r1;
  }
  const int64_t nblk = (int64_t)A->gblk.size() - 1;
  // per block: entries sorted by local row (stable), their positions, and each entry's column
  // positions within its row
  std::vector<int32_t> spe(np);
  std::vector<uint8_t> spli(np);
  std::vector<uint16_t> gpos(np), gloff((size_t)nblk * (nl + 1));
  std::vector<uint8_t> s8(narrow ? (size_t)np * nl : 0);
  std::vector<uint16_t> s16(narrow ? 0 : (size_t)np * nl);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t b = 0; b < nblk; ++b) {
    const int32_t r0 = A->gblk[b], r1 = A->gblk[b + 1];
    const int32_t e0 = pptr[r0], ne = pptr[r1] - e0;
    std::vector<int> cnt(nl + 1, 0);
    for (int32_t k = 0; k < ne; ++k) cnt[pli[e0 + k] + 1]++;
    for (int L = 0; L < nl; ++L) cnt[L + 1] += cnt[L];
    for (int L = 0; L <= nl; ++L) gloff[(size_t)b * (nl + 1) + L] = (uint16_t)cnt[L];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int32_t r = r0; r < r1; ++r) {
      const int32_t *c0 = A->col.data() + A->row_ptr[r];
      const int32_t *c1 = A->col.data() + A->row_ptr[r + 1];
      for (int32_t p = pptr[r]; p < pptr[r + 1]; ++p) {
        const int q = fill[pli[p]]++;             // sorted position within the block
        gpos[p] = (uint16_t)q;
        const int64_t sp = e0 + q;
        spe[sp] = pe[p];
        spli[sp] = pli[p];
        const int32_t *ed = &s.edofs[(int64_t)pe[p] * nl];
        for (int j = 0; j < nl; ++j) {
          const int pos = (int)(std::lower_bound(c0, c1, ed[j]) - c0);
          if (narrow)
            s8[(size_t)sp * nl + j] = (uint8_t)pos;
          else
            s16[(size_t)sp * nl + j] = (uint16_t)pos;
        }
      }
    }
  }
  hipStream_t st = A->ctx->stream;
  A->d_pptr.upload(pptr, st);
  A->d_pe.upload(spe, st);
  A->d_pli.upload(spli, st);
  A->d_gpos.upload(gpos, st);
  A->d_gloff.upload(gloff, st);
  A->d_gblk.upload(A->gblk, st);
  if (narrow)
    A->d_pslot8.upload(s8, st);
  else
    A->d_pslot16.upload(s16, st);
  OCMP_HIP(hipStreamSynchronize(st));
}

void build_h1_schedule(Matrix *A) {
  Space &s = *A->space;
  Mesh &m = *s.mesh;
  const int nl = s.nloc;
  // element -> CSR slot map
  std::vector<int32_t> slot((size_t)m.nel * nl * nl);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m.nel; ++e) {
    const int32_t *ed = &s.edofs[e * nl];
    for (int i = 0; i < nl; ++i) {
      int32_t r = ed[i];
      const int32_t *c0 = A->col.data() + A->row_ptr[r];
      const int32_t *c1 = A->col.data() + A->row_ptr[r + 1];
      for (int j = 0; j < nl; ++j) {
        const int32_t *p = std::lower_bound(c0, c1, ed[j]);
        slot[((size_t)e * nl + i) * nl + j] = (int32_t)(p - A->col.data());
      }
    }
  }
  A->d_slot.upload(slot, A->ctx->stream);
  OCMP_HIP(hipStreamSynchronize(A->ctx->stream));
}

}  // namespace ocmp

using namespace ocmp;

// ====================================================================== C ABI

extern "C" {

const char *ocmp_last_error(void) { return g_err.c_str(); }
int ocmp_version(void) { return 1; }

int ocmp_ctx_create(int device, void *hip_stream, ocmp_ctx **out) {
  return guard([&] {
    OCMP_REQUIRE(out, "out is NULL");
    int n = 0;
    OCMP_HIP(hipGetDeviceCount(&n));
    OCMP_REQUIRE(device >= 0 && device < n, "device index out of range");
    OCMP_HIP(hipSetDevice(device));
    // this device's copy of the allocation-free entry points' scratch (looked up once per device)
    double *noalloc = noalloc_scratch_init(device);
    auto *c = new ocmp_ctx();
    c->c.device = device;
    c->c.noalloc_scratch = noalloc;
    // NULL selects the device's null stream, which is also PyTorch's default stream:
    // work enqueued by the library and by torch is then ordered without extra events.
    c->c.stream = (hipStream_t)hip_stream;
    c->c.own_stream = false;
    c->c.scal.alloc(256);
    c->c.red.alloc(1 << 20);
    OCMP_HIP(hipHostMalloc((void **)&c->c.host_scal, 256 * sizeof(double), hipHostMallocDefault));
    *out = c;
  });
}

int ocmp_ctx_destroy(ocmp_ctx *ctx) {
  return guard([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    (void)hipStreamSynchronize(ctx->c.stream);
    for (auto &r : ctx->c.pending) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (auto e : ctx->c.event_pool) (void)hipEventDestroy(e);
    if (ctx->c.host_scal) (void)hipHostFree(ctx->c.host_scal);
    if (ctx->c.own_stream) (void)hipStreamDestroy(ctx->c.stream);
    delete ctx;
  });
}

int ocmp_ctx_sync(ocmp_ctx *ctx) {
  return guard([&] {
    OCMP_REQUIRE(ctx, "ctx is NULL");
    OCMP_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

void *ocmp_ctx_stream(ocmp_ctx *ctx) { return ctx ? (void *)ctx->c.stream : nullptr; }

int ocmp_ctx_timing(ocmp_ctx *ctx, int enable) {
  return guard([&] {
    OCMP_REQUIRE(ctx, "ctx is NULL");
    ctx->c.timing = enable != 0;
  });
}

int ocmp_ctx_timing_report(ocmp_ctx *ctx, char *buf, int64_t cap) {
  return guard([&] {
    OCMP_REQUIRE(ctx && buf && cap > 2, "bad arguments");
    ctx->c.flush_timing();
    std::string s = "{";
    bool first = true;
    for (auto &kv : ctx->c.acc) {
      char tmp[256];
      snprintf(tmp, sizeof tmp, "%s\"%s\": [%ld, %.6f, %.6e]", first ? "" : ", ",
               ctx->c.class_names[kv.first].c_str(), kv.second.first, kv.second.second, ctx->c.acc_bytes[kv.first]);
      s += tmp;
      first = false;
    }
    s += "}";
    OCMP_REQUIRE((int64_t)s.size() + 1 <= cap, "timing report buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    ctx->c.acc.clear();
    ctx->c.acc_bytes.clear();
  });
}

int ocmp_mesh_create(ocmp_ctx *ctx, int dim, int64_t nv, const double *coords, int64_t nel, const int32_t *elems,
                     int64_t nbnd, const int32_t *bnd_facets, const int32_t *bnd_markers, ocmp_mesh **out) {
  return guard([&] {
    OCMP_REQUIRE(ctx && out && coords && elems, "NULL argument");
    OCMP_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    OCMP_REQUIRE(nv > 0 && nel > 0 && nbnd >= 0, "empty mesh");
    OCMP_REQUIRE(nbnd == 0 || (bnd_facets && bnd_markers), "NULL boundary arrays");
    OCMP_REQUIRE(nv < INT32_MAX && nel < INT32_MAX, "mesh too large for 32-bit ids");
    OCMP_HIP(hipSetDevice(ctx->c.device));
    auto *mm = new ocmp_mesh();
    Mesh &m = mm->m;
    try {
      m.ctx = &ctx->c;
      m.dim = dim;
      m.nv = nv;
      m.nel = nel;
      m.coords.assign(coords, coords + nv * dim);
      m.elems.assign(elems, elems + nel * (dim + 1));
      for (int64_t e = 0; e < nel; ++e) {
        int32_t *v = &m.elems[e * (dim + 1)];
        for (int a = 0; a <= dim; ++a) OCMP_REQUIRE(v[a] >= 0 && v[a] < nv, "element vertex out of range");
        std::sort(v, v + dim + 1);
        for (int a = 0; a < dim; ++a) OCMP_REQUIRE(v[a] != v[a + 1], "degenerate element (repeated vertex)");
      }
      build_topology(m, nbnd, bnd_facets, bnd_markers);
      hipStream_t st = ctx->c.stream;
      m.d_coords.upload(m.coords, st);
      m.d_elems.upload(m.elems, st);
      m.d_elem_nb.upload(m.elem_nb, st);
      m.d_elem_nbk.upload(m.elem_nbk, st);
      OCMP_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete mm;
      throw;
    }
    *out = mm;
  });
}

int ocmp_mesh_info(const ocmp_mesh *mesh, int64_t *out) {
  return guard([&] {
    OCMP_REQUIRE(mesh && out, "NULL argument");
    const Mesh &m = mesh->m;
    int64_t v[8] = {m.nv, m.nedge, m.nface, m.nel, m.nfacet, m.ninterior, m.nbnd, m.dim};
    std::memcpy(out, v, sizeof v);
  });
}

int ocmp_mesh_destroy(ocmp_mesh *mesh) {
  return guard([&] { delete mesh; });
}

int ocmp_space_create(ocmp_mesh *mesh, int ncomp, const ocmp_fe_desc *comps, int dgjumps, ocmp_space **out) {
  return guard([&] {
    OCMP_REQUIRE(mesh && comps && out && ncomp > 0, "bad arguments");
    auto *sp = new ocmp_space();
    try {
      sp->s.mesh = &mesh->m;
      sp->s.dgjumps = dgjumps != 0;
      build_space(sp->s, ncomp, comps);
      hipStream_t st = mesh->m.ctx->stream;
      sp->s.d_edofs.upload(sp->s.edofs, st);
      sp->s.d_free.upload(sp->s.free, st);
      sp->s.d_edges.upload(mesh->m.edges, st);
      bool all_h1 = true;
      for (auto &b : sp->s.blocks) all_h1 &= b.kind == OCMP_H1;
      if (all_h1) build_colouring(sp->s);
      OCMP_HIP(hipStreamSynchronize(st));
    } catch (...) {
      delete sp;
      throw;
    }
    *out = sp;
  });
}

int ocmp_space_info(const ocmp_space *s, int64_t *out) {
  return guard([&] {
    OCMP_REQUIRE(s && out, "NULL argument");
    int64_t nfree = 0;
    for (auto f : s->s.free) nfree += f;
    out[0] = s->s.ndof;
    out[1] = s->s.nloc;
    out[2] = (int64_t)s->s.blocks.size();
    out[3] = nfree;
  });
}

int ocmp_space_blocks(const ocmp_space *s, int64_t *out) {
  return guard([&] {
    OCMP_REQUIRE(s && out, "NULL argument");
    for (size_t i = 0; i < s->s.blocks.size(); ++i) {
      const Block &b = s->s.blocks[i];
      out[4 * i + 0] = b.dof_off;
      out[4 * i + 1] = b.n1;
      out[4 * i + 2] = b.loc_off;
      out[4 * i + 3] = b.nloc;
    }
  });
}

int ocmp_space_freedofs(const ocmp_space *s, uint8_t *mask) {
  return guard([&] {
    OCMP_REQUIRE(s && mask, "NULL argument");
    std::memcpy(mask, s->s.free.data(), s->s.free.size());
  });
}

int ocmp_space_element_dofs(const ocmp_space *s, int32_t *out) {
  return guard([&] {
    OCMP_REQUIRE(s && out, "NULL argument");
    std::memcpy(out, s->s.edofs.data(), s->s.edofs.size() * sizeof(int32_t));
  });
}

int ocmp_space_destroy(ocmp_space *s) {
  return guard([&] { delete s; });
}

int ocmp_matrix_create(ocmp_space *s, ocmp_matrix **out) {
  return guard([&] {
    OCMP_REQUIRE(s && out, "NULL argument");
    auto *A = new ocmp_matrix();
    try {
      A->A.space = &s->s;
      A->A.ctx = s->s.mesh->ctx;
      OCMP_HIP(hipSetDevice(A->A.ctx->device));
      build_pattern(&A->A);
      if (!A->A.dg) build_h1_gather(&A->A);
      OCMP_HIP(hipStreamSynchronize(A->A.ctx->stream));
    } catch (...) {
      delete A;
      throw;
    }
    *out = A;
  });
}

int ocmp_matrix_info(const ocmp_matrix *A, int64_t *out) {
  return guard([&] {
    OCMP_REQUIRE(A && out, "NULL argument");
    out[0] = A->A.n;
    out[1] = A->A.nnz;
    out[2] = A->A.space->ncolours;
    out[3] = A->A.dg ? 1 : 0;
  });
}

int ocmp_matrix_pattern(const ocmp_matrix *A, int32_t *row_ptr, int32_t *col) {
  return guard([&] {
    OCMP_REQUIRE(A && row_ptr && col, "NULL argument");
    std::memcpy(row_ptr, A->A.row_ptr.data(), A->A.row_ptr.size() * sizeof(int32_t));
    std::memcpy(col, A->A.col.data(), A->A.col.size() * sizeof(int32_t));
  });
}

int ocmp_matrix_values(const ocmp_matrix *A, double *vals) {
  return guard([&] {
    OCMP_REQUIRE(A && vals, "NULL argument");
    OCMP_HIP(hipMemcpyAsync(vals, A->A.d_val.p, A->A.nnz * sizeof(double), hipMemcpyDeviceToHost,
                            A->A.ctx->stream));
    OCMP_HIP(hipStreamSynchronize(A->A.ctx->stream));
  });
}

int ocmp_matrix_destroy(ocmp_matrix *A) {
  return guard([&] {
    if (A) (void)hipStreamSynchronize(A->A.ctx->stream);
    delete A;
  });
}

int ocmp_assemble_bilinear(ocmp_matrix *A, int nterms, const ocmp_term *terms, int nfields,
                           const ocmp_field *fields, int accumulate) {
  return guard([&] {
    OCMP_REQUIRE(A && (nterms == 0 || terms) && (nfields == 0 || fields), "NULL argument");
    std::vector<ocmp_term> t(terms, terms + nterms);
    std::vector<ocmp_field> f(fields, fields + nfields);
    const Space &s = *A->A.space;
    for (auto &x : t) {
      OCMP_REQUIRE(x.trial >= 0 && x.trial < (int)s.blocks.size() && x.test >= 0 && x.test < (int)s.blocks.size(),
                   "term block index out of range");
      OCMP_REQUIRE(x.kind == OCMP_BFI_DG_FACET || (x.kind >= 0 && x.kind <= OCMP_BFI_BND_FLUX),
                   "unknown term kind");
      OCMP_REQUIRE((x.kind != OCMP_BFI_BND_UPWIND && x.kind != OCMP_BFI_BND_MASS && x.kind != OCMP_BFI_BND_FLUX) ||
                       (!A->A.dg && x.trial == x.test),
                   "the boundary terms are for one H1 block (trial == test)");
      OCMP_REQUIRE(x.weight_mode == OCMP_WEIGHT_NONE || (x.weight_field >= 0 && x.weight_field < nfields),
                   "weight field index out of range");
      OCMP_REQUIRE(x.wind_field < 0 || x.wind_field + s.mesh->dim <= nfields, "wind field index out of range");
      OCMP_REQUIRE(x.kind != OCMP_BFI_DTRIAL && x.kind != OCMP_BFI_DTEST ||
                       (x.dir >= 0 && x.dir < s.mesh->dim),
                   "derivative direction out of range");
    }
    for (auto &x : f) {
      OCMP_REQUIRE(x.space && x.vec, "field without space or vector");
      OCMP_REQUIRE(x.space->s.mesh == s.mesh, "field lives on another mesh");
      OCMP_REQUIRE(x.block >= 0 && x.block < (int)x.space->s.blocks.size(), "field block out of range");
      OCMP_REQUIRE(x.space->s.blocks[x.block].kind == OCMP_H1, "fields must be H1 blocks");
    }
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    if (A->A.dg)
      launch_assemble_dg(&A->A, t, accumulate != 0);
    else
      launch_assemble_h1(&A->A, t, f, accumulate != 0);
    A->A.assembled = true;
    const int nb = (int)s.blocks.size();
    uint64_t pairs = 0;
    for (auto &x : t) pairs |= nb <= 8 ? 1ull << (x.test * nb + x.trial) : ~0ull;
    A->A.active_pairs = (accumulate ? A->A.active_pairs : 0) | pairs;
    ++A->A.version;
  });
}

int ocmp_assemble_linear(ocmp_space *s, double *b, int nterms, const ocmp_lterm *terms, int nfields,
                         const ocmp_field *fields, int accumulate) {
  return guard([&] {
    OCMP_REQUIRE(s && b && (nterms == 0 || terms) && (nfields == 0 || fields), "NULL argument");
    std::vector<ocmp_lterm> t(terms, terms + nterms);
    std::vector<ocmp_field> f(fields, fields + nfields);
    for (auto &x : t) {
      OCMP_REQUIRE(x.test >= 0 && x.test < (int)s->s.blocks.size(), "term block index out of range");
      OCMP_REQUIRE(x.fn != OCMP_FN_FIELD || (x.field >= 0 && x.field < nfields), "field index out of range");
      OCMP_REQUIRE(x.weight_mode == OCMP_WEIGHT_NONE || (x.weight_field >= 0 && x.weight_field < nfields),
                   "weight field index out of range");
      OCMP_REQUIRE(x.kind >= OCMP_LFI_SOURCE && x.kind <= OCMP_LFI_BND_SOURCE, "unknown linear term kind");
      OCMP_REQUIRE(x.kind != OCMP_LFI_DG_DIRICHLET || s->s.blocks[x.test].kind == OCMP_L2,
                   "DG Dirichlet data needs an L2 block");
      OCMP_REQUIRE(x.kind != OCMP_LFI_BND_SOURCE || s->s.blocks[x.test].kind == OCMP_H1,
                   "boundary sources need an H1 block");
    }
    for (auto &x : f) {
      OCMP_REQUIRE(x.space && x.vec, "field without space or vector");
      OCMP_REQUIRE(x.space->s.mesh == s->s.mesh, "field lives on another mesh");
      OCMP_REQUIRE(x.block >= 0 && x.block < (int)x.space->s.blocks.size(), "field block out of range");
      OCMP_REQUIRE(x.space->s.blocks[x.block].kind == OCMP_H1, "fields must be H1 blocks");
    }
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    launch_assemble_linear(&s->s, b, t, f, accumulate != 0);
  });
}

int ocmp_dirichlet_set(ocmp_space *s, int block, double *x, int fn, const double *c, uint64_t marker_mask) {
  return guard([&] {
    OCMP_REQUIRE(s && x && c, "NULL argument");
    OCMP_REQUIRE(block >= 0 && block < (int)s->s.blocks.size(), "block out of range");
    OCMP_REQUIRE(s->s.blocks[block].kind == OCMP_H1, "Dirichlet values apply to H1 blocks");
    OCMP_REQUIRE(fn != OCMP_FN_FIELD, "Dirichlet data must be a closed form or an expression");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    launch_dirichlet_set(&s->s, block, x, fn, c, marker_mask);
  });
}

int ocmp_set(ocmp_space *s, int block, double *x, int fn, const double *c) {
  return guard([&] {
    OCMP_REQUIRE(s && x && c, "NULL argument");
    OCMP_REQUIRE(block >= 0 && block < (int)s->s.blocks.size(), "block out of range");
    OCMP_REQUIRE(fn != OCMP_FN_FIELD, "Set takes a closed form or an expression");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    launch_set(&s->s, block, x, fn, c);
  });
}

int ocmp_expr_create(ocmp_ctx *ctx, int32_t n, const double *code, int32_t *id) {
  return guard([&] {
    OCMP_REQUIRE(ctx && code && id, "NULL argument");
    OCMP_HIP(hipSetDevice(ctx->c.device));
    *id = expr_register(&ctx->c, n, code);
  });
}

int ocmp_error_norms(ocmp_space *s, int block, const double *x, const double *ref, int fn, const double *c,
                     double *out) {
  return guard([&] {
    OCMP_REQUIRE(s && x && c && out, "NULL argument");
    OCMP_REQUIRE(block >= 0 && block < (int)s->s.blocks.size(), "block out of range");
    OCMP_REQUIRE(ref || fn != OCMP_FN_FIELD, "a field reference is passed as ref");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    error_norms(&s->s, block, x, ref, fn, c, out);
  });
}

int ocmp_error_norms_vec(ocmp_space *s, int nblocks, const int32_t *blocks, const double *x, const double *ref,
                         const int32_t *fns, const double *c, double *out) {
  return guard([&] {
    OCMP_REQUIRE(s && blocks && x && out && (ref || (fns && c)), "NULL argument");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    error_norms_vec(&s->s, nblocks, blocks, x, ref, fns, c, out);
  });
}

int ocmp_l2_pair_norms(ocmp_space *s, int nblocks, const int32_t *blocks, const double *a, const double *b,
                       double *out) {
  return guard([&] {
    OCMP_REQUIRE(s && blocks && a && b && out, "NULL argument");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    l2_pair_norms(&s->s, nblocks, blocks, a, b, out);
  });
}

int ocmp_divergence_norm(ocmp_space *s, int block0, const double *x, double *out) {
  return guard([&] {
    OCMP_REQUIRE(s && x && out, "NULL argument");
    OCMP_REQUIRE(block0 >= 0 && block0 + s->s.mesh->dim <= (int)s->s.blocks.size(), "block out of range");
    OCMP_REQUIRE(s->s.blocks[block0].kind == OCMP_H1, "the divergence norm is for H1 blocks");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    divergence_norm(&s->s, block0, x, out);
  });
}

int ocmp_linf_norm(ocmp_space *s, int block, const double *x, const double *ref, double *out) {
  return guard([&] {
    OCMP_REQUIRE(s && x && out, "NULL argument");
    OCMP_REQUIRE(block >= 0 && block < (int)s->s.blocks.size(), "block out of range");
    OCMP_HIP(hipSetDevice(s->s.mesh->ctx->device));
    linf_norm(&s->s, block, x, ref, out);
  });
}

int ocmp_axpby(ocmp_ctx *ctx, int64_t n, double a, const double *x, double b, double *y) {
  return guard([&] {
    OCMP_REQUIRE(ctx && x && y && n >= 0, "bad arguments");
    OCMP_HIP(hipSetDevice(ctx->c.device));
    if (n) axpby(&ctx->c, n, a, x, b, y);
  });
}

int ocmp_matvec(ocmp_matrix *A, const double *x, double *y) {
  return guard([&] {
    OCMP_REQUIRE(A && x && y, "NULL argument");
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    spmv_f64(&A->A, x, y);
  });
}

int ocmp_residual_norms(ocmp_matrix *A, const double *b, const double *x, double *out) {
  return guard([&] {
    OCMP_REQUIRE(A && b && x && out, "NULL argument");
    OCMP_REQUIRE(A->A.assembled, "matrix not assembled");
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    residual_norms(&A->A, b, x, out);
  });
}

int ocmp_residual_norms_noalloc(ocmp_matrix *A, const double *b, const double *x, double *out) {
  return guard([&] {
    OCMP_REQUIRE(A && b && x && out, "NULL argument");
    OCMP_REQUIRE(A->A.assembled, "matrix not assembled");
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    residual_norms_noalloc(&A->A, b, x, out);
  });
}

int ocmp_hessenberg_eigvals(int32_t n, const double *H, int32_t ldh, double *wr, double *wi) {
  return guard([&] {
    OCMP_REQUIRE(H && wr && wi, "NULL argument");
    OCMP_REQUIRE(n >= 1 && ldh >= n, "need n >= 1 and ldh >= n");
    if (!hessenberg_eigvals(n, H, ldh, wr, wi)) throw Error(OCMP_ERR_NUMERIC, "Hessenberg QR did not converge");
  });
}

int ocmp_newton_shifts(int32_t n, const double *wr, const double *wi, int32_t s, double *d, double *c) {
  return guard([&] {
    OCMP_REQUIRE(wr && wi && d && c, "NULL argument");
    OCMP_REQUIRE(n >= 0 && s >= 1, "need n >= 0 and s >= 1");
    newton_shifts(n, wr, wi, s, d, c);
  });
}

int ocmp_residual(ocmp_matrix *A, const double *b, const double *x, double *r, double *norm) {
  return guard([&] {
    OCMP_REQUIRE(A && b && x && r && norm, "NULL argument");
    OCMP_REQUIRE(A->A.assembled, "matrix not assembled");
    OCMP_REQUIRE(r != b && r != x, "r must not alias b or x");
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    *norm = residual_vector(&A->A, b, x, r);
  });
}

int ocmp_precond_create(ocmp_matrix *A, int kind, int precision, ocmp_precond **out) {
  return guard([&] {
    OCMP_REQUIRE(A && out, "NULL argument");
    OCMP_REQUIRE(kind == OCMP_PRE_NONE || kind == OCMP_PRE_JACOBI || kind == OCMP_PRE_BLOCK_JACOBI,
                 "unknown preconditioner");
    OCMP_REQUIRE(precision == OCMP_FP64 || precision == OCMP_MIXED, "unknown precision");
    OCMP_REQUIRE(kind != OCMP_PRE_BLOCK_JACOBI || A->A.dg, "block Jacobi needs an L2 (DG) matrix");
    auto *P = new ocmp_precond();
    P->P.A = &A->A;
    P->P.kind = kind;
    P->P.precision = precision;
    *out = P;
  });
}

int ocmp_precond_create_schur(ocmp_matrix *A, int precision, uint64_t schur_blocks, ocmp_precond **out) {
  return guard([&] {
    OCMP_REQUIRE(A && out, "NULL argument");
    OCMP_REQUIRE(precision == OCMP_FP64 || precision == OCMP_MIXED, "unknown precision");
    const auto &bl = A->A.space->blocks;
    OCMP_REQUIRE(schur_blocks != 0, "empty Schur block mask");
    OCMP_REQUIRE(bl.size() >= 64 || (schur_blocks >> bl.size()) == 0, "Schur block mask names a missing block");
    OCMP_REQUIRE(!A->A.dg, "the Schur-diagonal preconditioner is for H1 (Taylor-Hood) systems");
    auto *P = new ocmp_precond();
    P->P.A = &A->A;
    P->P.kind = OCMP_PRE_JACOBI;
    P->P.precision = precision;
    P->P.schur_blocks = schur_blocks;
    *out = P;
  });
}

int ocmp_precond_set_schur_solve(ocmp_precond *P, int max_iters, double tol) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.schur_blocks != 0, "the Schur inner solve needs a Schur preconditioner");
    OCMP_REQUIRE(max_iters >= 0 && max_iters <= 1000, "Schur inner iterations must be in [0, 1000]");
    OCMP_REQUIRE(tol >= 0.0 && tol < 1.0, "Schur inner tolerance must be in [0, 1)");
    P->P.schur_iters = max_iters;
    P->P.schur_tol = tol;
    if (max_iters > 0 && !P->P.S) P->P.ready = false;  // S is built by the next update
  });
}

int ocmp_precond_set_block_triangular(ocmp_precond *P, const double *schur_diag, int vel_iters, double vel_tol) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.schur_blocks != 0, "the block-triangular step needs a Schur preconditioner");
    OCMP_REQUIRE(vel_iters >= 0 && vel_iters <= 10000, "velocity inner iterations must be in [0, 10000]");
    OCMP_REQUIRE(vel_tol >= 0.0 && vel_tol < 1.0, "velocity inner tolerance must be in [0, 1)");
    if (vel_iters == 0) {
      P->P.btri = false;
      return;
    }
    OCMP_REQUIRE(schur_diag, "NULL Schur diagonal");
    OCMP_REQUIRE(P->P.schur_iters == 0, "the block-triangular step replaces the Schur inner solve: turn that off");
    Matrix *A = P->P.A;
    int64_t s0, s1;
    schur_range(&P->P, s0, s1);
    OCMP_REQUIRE(s0 > 0 && s1 == A->n, "the block-triangular step needs the Schur blocks last");
    OCMP_HIP(hipSetDevice(A->ctx->device));
    P->P.pdiag.alloc(A->n - s0);
    OCMP_HIP(hipMemcpyAsync(P->P.pdiag.p, schur_diag, (A->n - s0) * sizeof(double), hipMemcpyDeviceToDevice,
                            A->ctx->stream));
    OCMP_HIP(hipStreamSynchronize(A->ctx->stream));
    P->P.btri = true;
    P->P.vel_iters = vel_iters;
    P->P.vel_tol = vel_tol;
    P->P.vel_cap = 0;
    P->P.ready = false;
  });
}

int ocmp_precond_set_bf16(ocmp_precond *P, int on) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.precision == OCMP_MIXED || !on, "BF16 storage is for preconditioners built with OCMP_MIXED");
    if ((on != 0) != P->P.bf16) P->P.ready = false;  // the copies are made by the next update
    P->P.bf16 = on != 0;
  });
}

int ocmp_precond_set_schur_amg(ocmp_precond *P, int vcycles) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.schur_blocks != 0, "the Schur AMG step needs a Schur preconditioner");
    OCMP_REQUIRE(vcycles >= 0 && vcycles <= 16, "V-cycles per step must be in [0, 16]");
    OCMP_REQUIRE(vcycles == 0 || (P->P.schur_iters == 0 && !P->P.btri),
                 "the Schur AMG step replaces the Schur inner solve and the block-triangular step");
    if (vcycles > 0 && !P->P.amg) P->P.ready = false;  // the hierarchy is built by the next update
    P->P.schur_amg = vcycles;
  });
}

int ocmp_precond_set_schur_triangular(ocmp_precond *P, int on) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.schur_blocks != 0, "the triangular velocity rows need a Schur preconditioner");
    OCMP_REQUIRE(!on || !P->P.btri, "the block-triangular step already solves the velocity rows");
    if (on && !P->P.schur_tri) P->P.ready = false;  // A_up is gathered by the next update
    P->P.schur_tri = on != 0;
  });
}

int ocmp_precond_schur_amg_levels(const ocmp_precond *P, int64_t *rows, int64_t *nnz, int max_levels,
                                  int *levels) {
  return guard([&] {
    OCMP_REQUIRE(P && levels, "NULL argument");
    OCMP_REQUIRE(max_levels >= 0 && (max_levels == 0 || (rows && nnz)), "bad level arrays");
    const Amg *M = P->P.amg;
    *levels = M ? (int)M->lv.size() : 0;
    for (int l = 0; l < max_levels; ++l) {
      const bool on = l < *levels;
      rows[l] = on ? M->lv[l]->A->n : 0;
      nnz[l] = on ? M->lv[l]->A->nnz : 0;
    }
  });
}

int ocmp_precond_update(ocmp_precond *P) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.A->assembled, "matrix not assembled");
    OCMP_HIP(hipSetDevice(P->P.A->ctx->device));
    precond_update(&P->P);
  });
}

int ocmp_precond_update_jacobi(ocmp_precond *P) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.A->assembled, "matrix not assembled");
    OCMP_HIP(hipSetDevice(P->P.A->ctx->device));
    precond_update_jacobi(&P->P);
  });
}

int ocmp_precond_destroy(ocmp_precond *P) {
  return guard([&] { delete P; });
}

int ocmp_solve(ocmp_matrix *A, ocmp_precond *P, const double *b, double *x, const ocmp_solve_opts *opts,
               ocmp_solve_stats *stats) {
  return guard([&] {
    OCMP_REQUIRE(A && b && x && opts && stats, "NULL argument");
    OCMP_REQUIRE(A->A.assembled, "matrix not assembled");
    OCMP_REQUIRE(!P || P->P.A == &A->A, "preconditioner built for another matrix");
    OCMP_REQUIRE(opts->tol > 0 && opts->maxit > 0, "tol and maxit must be positive");
    OCMP_REQUIRE((opts->solver != OCMP_SOLVER_GMRES && opts->solver != OCMP_SOLVER_FGMRES) ||
                     (opts->restart >= 1 && opts->restart <= 200),
                 "GMRES restart must be in [1, 200]");
    OCMP_REQUIRE(opts->s_step == 0 || opts->s_step == 5, "s_step must be 0 (standard) or 5");
    OCMP_HIP(hipSetDevice(A->A.ctx->device));
    if (P && !P->P.ready) precond_update(&P->P);
    std::memset(stats, 0, sizeof(*stats));
    solve(&A->A, P ? &P->P : nullptr, b, x, *opts, *stats);
  });
}

// ------------------------------------------------------------------ learned initial guess

#define MLP_ENTER(m)                                   \
  OCMP_REQUIRE((m) && (m)->M, "NULL mlp handle");     \
  OCMP_HIP(hipSetDevice((m)->ctx->device))

int ocmp_mlp_create(ocmp_ctx *ctx, const float *params, int64_t max_batch, ocmp_mlp **out) {
  return guard([&] {
    OCMP_REQUIRE(ctx && params && out, "NULL argument");
    OCMP_HIP(hipSetDevice(ctx->c.device));
    auto *h = new ocmp_mlp();
    h->ctx = &ctx->c;
    try {
      h->M = ocmp::mlp_create(&ctx->c, params, max_batch);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int ocmp_mlp_destroy(ocmp_mlp *mlp) {
  return guard([&] {
    if (!mlp) return;
    ocmp::mlp_destroy(mlp->M);
    delete mlp;
  });
}

int ocmp_mlp_get_params(ocmp_mlp *mlp, float *params) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_get_params(mlp->M, params);
  });
}

int ocmp_mlp_set_params(ocmp_mlp *mlp, const float *params) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_set_params(mlp->M, params);
  });
}

int ocmp_mlp_set_scales(ocmp_mlp *mlp, int nseg, const int64_t *seg_off, const double *s, const double *U) {
  return guard([&] {
    MLP_ENTER(mlp);
    OCMP_REQUIRE(s && U, "NULL argument");
    ocmp::mlp_set_scales(mlp->M, nseg, seg_off, s, U);
  });
}

int ocmp_mlp_get_scales(ocmp_mlp *mlp, int *nseg, int64_t *seg_off, double *s, double *U) {
  return guard([&] {
    MLP_ENTER(mlp);
    OCMP_REQUIRE(nseg && seg_off && s && U, "NULL argument");
    ocmp::mlp_get_scales(mlp->M, nseg, seg_off, s, U);
  });
}

int ocmp_mlp_fit_scales(ocmp_mlp *mlp, const double *snap, int nsnap, int64_t ld, int nseg, const int64_t *seg_off) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_fit_scales(mlp->M, snap, nsnap, ld, nseg, seg_off);
  });
}

int ocmp_mlp_sample(ocmp_mlp *mlp, const double *snap, int nsnap, int64_t ld, int64_t n, uint64_t seed,
                    int64_t batch, float *X, float *T) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_sample(mlp->M, snap, nsnap, ld, n, seed, batch, X, T);
  });
}

int ocmp_mlp_forward(ocmp_mlp *mlp, int64_t n, const float *X, float *Y) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_forward(mlp->M, n, X, Y);
  });
}

int ocmp_mlp_grad(ocmp_mlp *mlp, int64_t n, const float *X, const float *T, float *grad, double *loss) {
  return guard([&] {
    MLP_ENTER(mlp);
    OCMP_REQUIRE(grad, "NULL gradient buffer");
    ocmp::mlp_grad(mlp->M, n, X, T, grad, loss);
  });
}

int ocmp_mlp_train_step(ocmp_mlp *mlp, int64_t n, const float *X, const float *T, float lr, double *loss) {
  return guard([&] {
    MLP_ENTER(mlp);
    ocmp::mlp_train_step(mlp->M, n, X, T, lr, loss);
  });
}

int ocmp_mlp_set_base(ocmp_mlp *mlp, int base) {
  return guard([&] {
    MLP_ENTER(mlp);
    mlp_set_base(mlp->M, base);
  });
}

int ocmp_mlp_get_base(ocmp_mlp *mlp, int *base) {
  return guard([&] {
    MLP_ENTER(mlp);
    OCMP_REQUIRE(base, "NULL argument");
    *base = mlp_get_base(mlp->M);
  });
}

int ocmp_mlp_guess(ocmp_mlp *mlp, int64_t n, const double *u0, const double *u1, const double *u2, const double *u3,
                   double *out) {
  return guard([&] {
    MLP_ENTER(mlp);
    OCMP_REQUIRE(out != u0 && out != u1 && out != u2 && out != u3, "out must not alias the history");
    ocmp::mlp_guess(mlp->M, n, u0, u1, u2, u3, out);
  });
}

// ------------------------------------------------------------------ learned preconditioner correction

#define UNET_ENTER(u)                                  \
  OCMP_REQUIRE((u) && (u)->U, "NULL unet handle");    \
  OCMP_HIP(hipSetDevice((u)->ctx->device))

int ocmp_unet_create(ocmp_ctx *ctx, int N, const float *params, ocmp_unet **out) {
  return guard([&] {
    OCMP_REQUIRE(ctx && params && out, "NULL argument");
    OCMP_HIP(hipSetDevice(ctx->c.device));
    auto *h = new ocmp_unet();
    h->ctx = &ctx->c;
    try {
      h->U = ocmp::unet_create(&ctx->c, N, params);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

int ocmp_unet_destroy(ocmp_unet *u) {
  return guard([&] {
    if (!u) return;
    ocmp::unet_destroy(u->U);
    delete u;
  });
}

int ocmp_unet_get_params(ocmp_unet *u, float *params) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(params, "NULL argument");
    ocmp::unet_get_params(u->U, params);
  });
}

int ocmp_unet_set_params(ocmp_unet *u, const float *params) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(params, "NULL argument");
    ocmp::unet_set_params(u->U, params);
  });
}

int ocmp_unet_set_channels(ocmp_unet *u, const int64_t *dof_off, const double *in_scale, const double *out_scale) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(dof_off && in_scale && out_scale, "NULL argument");
    ocmp::unet_set_channels(u->U, dof_off, in_scale, out_scale);
  });
}

int ocmp_unet_forward(ocmp_unet *u, const float *x, float *y) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(x && y, "NULL argument");
    ocmp::unet_forward_grid(u->U, x, y);
  });
}

int ocmp_unet_grad(ocmp_unet *u, const float *x, const float *gy, float *grad) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(x && gy && grad, "NULL argument");
    ocmp::unet_grad_grid(u->U, x, gy, grad);
  });
}

int ocmp_unet_train_step(ocmp_unet *u, ocmp_matrix *A, ocmp_precond *P, const double *samples, int nsamples,
                         int64_t ld, const int32_t *idx, int nbatch, float lr, double *loss) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(A && P && samples && idx && loss, "NULL argument");
    OCMP_REQUIRE(P->P.A == &A->A, "preconditioner built for another matrix");
    OCMP_REQUIRE(A->A.assembled, "matrix not assembled");
    OCMP_REQUIRE(ld >= A->A.n && nsamples >= 1, "bad sample layout");
    OCMP_REQUIRE(lr > 0.f, "learning rate must be > 0");
    for (int b = 0; b < nbatch; ++b) OCMP_REQUIRE(idx[b] >= 0 && idx[b] < nsamples, "sample index out of range");
    OCMP_REQUIRE(ocmp::unet_max_index(u->U) <= A->A.n, "U-net channels reach beyond the system's dofs");
    if (!P->P.ready) ocmp::precond_update(&P->P);
    *loss = ocmp::unet_train_step(u->U, &A->A, &P->P, samples, ld, idx, nbatch, lr);
  });
}

int ocmp_unet_fit_channels(ocmp_unet *u, ocmp_precond *P, const double *samples, int nsamples, int64_t ld,
                           const int64_t *dof_off) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(P && samples && dof_off && nsamples >= 1, "bad arguments");
    OCMP_REQUIRE(ld >= P->P.A->n, "sample stride shorter than the system");
    if (!P->P.ready) ocmp::precond_update(&P->P);
    ocmp::unet_fit_channels(u->U, &P->P, samples, nsamples, ld, dof_off);
  });
}

int ocmp_unet_get_channels(ocmp_unet *u, int64_t *dof_off, double *in_scale, double *out_scale) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(dof_off && in_scale && out_scale, "NULL argument");
    ocmp::unet_get_channels(u->U, dof_off, in_scale, out_scale);
  });
}

int ocmp_unet_activation(ocmp_unet *u, int layer, float *out, int64_t *count) {
  return guard([&] {
    UNET_ENTER(u);
    OCMP_REQUIRE(layer >= -1 && layer <= 7, "layer out of range [-1, 7]");
    OCMP_REQUIRE(count, "NULL argument");
    *count = ocmp::unet_activation(u->U, layer, out);
  });
}

int ocmp_precond_attach_unet(ocmp_precond *P, ocmp_unet *u) {
  return guard([&] {
    OCMP_REQUIRE(P, "NULL argument");
    OCMP_REQUIRE(P->P.kind == OCMP_PRE_JACOBI, "the learned correction extends a point-Jacobi preconditioner");
    if (u) OCMP_REQUIRE(ocmp::unet_max_index(u->U) <= P->P.A->n, "U-net channels reach beyond the system's dofs");
    P->P.unet = u ? u->U : nullptr;
  });
}

}  // extern "C"
