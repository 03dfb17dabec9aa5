// Assembly kernels: element matrices of the OpenCMP weak forms on affine
// simplices, scattered deterministically into CSR.
//
//  * H1 (continuous) spaces: one thread per (element, local test row).
//    Constant-coefficient terms contract precomputed reference tensors with
//    the element metric (K = |det J| sum_k g_k(e) T_k); terms with field
//    coefficients (diffuse-interface weight phi, Oseen wind) integrate by
//    quadrature. Elements are launched colour by colour (no two elements of a
//    colour share a vertex), so the read-modify-write into CSR through the
//    element->slot map is race-free and bitwise reproducible.
//  * L2 (DG) spaces: owner-computes. The thread of (element e, row i) forms the
//    whole CSR row: the volume block plus, for every facet of e, its own side of
//    the facet integral (diagonal block) and the coupling to the neighbour's
//    dofs. Each value is written exactly once; no colouring, no atomics.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <type_traits>

#include "internal.h"

namespace ocmp {

constexpr int MAXT = 8;   // terms per launch of the element-scatter kernels
constexpr int MAXTG = 32; // volume terms per launch of the row gather (they share factor sets)
constexpr int MAXF = 8;   // fields per launch
constexpr int MAXB = 8;   // scalar blocks per space
constexpr int MAXL = 40;  // compound local dofs
constexpr int MAXNB = 10; // DG block size (P3)

struct KTerm {
  int kind, trial, test, dir;
  double coef;
  const double *elem_coef;
  double wind[3];
  int wind_field, weight_mode, weight_field, quad;
  int tab;  // contraction: offset of the tensor set (ntens tensors of nl_test x nl_trial)
  int ntens;
  int sipg, upwind;
  double eps, alpha;
  unsigned long long dmask;
};

struct KField {
  const double *vec;
  const int32_t *edofs;  // element dof table of the field's space
  int ld;                // its nloc
  int loc_off;           // block offset within the element dof list
  int nl;                // block local dofs
  int tab_phi;           // offset of phi table (nq x nl) at the volume rule
  int tab_fphi;          // boundary terms: phi at the facet rule ((dim+1) x nqf x nl)
};

struct KBlocks {
  int n;
  int loc_off[MAXB];
  int nloc[MAXB];
  int tab_phi[MAXB];   // volume rule: phi (nq x nl)
  int tab_grad[MAXB];  // volume rule: reference gradients (nq x nl x dim)
  int tab_fphi[MAXB];  // boundary terms: phi at the facet rule of each local facet ((dim+1) x nqf x nl)
};

struct KArgs {
  int nterms, nfields, nq, nloc;
  KTerm t[MAXT];
  KField f[MAXF];
  KBlocks b;
  int tab_w;   // quadrature weights of the volume rule
  int tab_lam; // barycentric points (nq x (dim+1))
  const double *tab;
  // boundary terms: facet rule on [0, 1] (nqf points, weights at tab_fw) and the element-facet
  // neighbours (-(marker+1) on the boundary)
  int nqf, tab_fw;
  const int32_t *elem_nb;
};

// ------------------------------------------------------------------ geometry

template <int D>
struct Geo {
  double Jinv[D][D];  // Jinv[a][c] = d xhat_a / d x_c
  double adet;
  double x0[D];
  double J[D][D];     // J[c][a] = (v_{a+1} - v_0)_c
};

template <int D>
__device__ __forceinline__ void geometry(const double *__restrict__ coords, const int32_t *__restrict__ v, Geo<D> &g) {
  double X[D + 1][D];
#pragma unroll
  for (int i = 0; i <= D; ++i)
#pragma unroll
    for (int c = 0; c < D; ++c) X[i][c] = coords[(int64_t)v[i] * D + c];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    g.x0[c] = X[0][c];
#pragma unroll
    for (int a = 0; a < D; ++a) g.J[c][a] = X[a + 1][c] - X[0][c];
  }
  if constexpr (D == 2) {
    double det = g.J[0][0] * g.J[1][1] - g.J[0][1] * g.J[1][0];
    double id = 1.0 / det;
    g.Jinv[0][0] = g.J[1][1] * id;
    g.Jinv[0][1] = -g.J[0][1] * id;
    g.Jinv[1][0] = -g.J[1][0] * id;
    g.Jinv[1][1] = g.J[0][0] * id;
    g.adet = fabs(det);
  } else {
    const double(*J)[3] = g.J;
    double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
    double id = 1.0 / det;
    g.Jinv[0][0] = c00 * id;
    g.Jinv[1][0] = c01 * id;
    g.Jinv[2][0] = c02 * id;
    g.Jinv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    g.Jinv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    g.Jinv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    g.Jinv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    g.Jinv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    g.Jinv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    g.adet = fabs(det);
  }
}

// ------------------------------------------------------------------ expressions (OCMP_FN_EXPR)
//
// Coefficient, boundary-data and reference-solution expressions of the OpenCMP config language
// (opencmp/config_functions/parse_arithmetic.py:37-268) arrive compiled to postfix programs
// (opcodes OCMP_OP_* with inline constants, see ocmp.h). Programs live in one table per device
// and are evaluated per point by a stack machine whose 8 stack slots are named registers: a push
// or pop shifts them, so the stack never goes to scratch memory.

constexpr int EXPR_CAP = 1 << 16;  // doubles of program text per device
constexpr int EXPR_DEPTH = 8;
static __device__ double expr_code[EXPR_CAP];

// sigmoid and hyperbolic tangent as the reference writes them (opencmp/helpers/math.py:23-59)
__device__ __forceinline__ double ref_sig(double x) { return 1.0 / (1.0 + exp(-x)); }
__device__ __forceinline__ double ref_tanh(double x) {
  const double xa = x - 350.0 > 0.0 ? 350.0 : x;
  const double e2 = exp(2.0 * xa);
  return (e2 - 1.0) / (e2 + 1.0);
}

__device__ double eval_expr(int pc, double t, const double *x, int dim) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0, s5 = 0.0, s6 = 0.0, s7 = 0.0;
  for (;;) {
    const int op = (int)expr_code[pc++];
    if (op == OCMP_OP_END) return s0;
    if (op <= OCMP_OP_T) {
      double v;
      if (op == OCMP_OP_CONST) v = expr_code[pc++];
      else if (op == OCMP_OP_X) v = x[0];
      else if (op == OCMP_OP_Y) v = dim > 1 ? x[1] : 0.0;
      else if (op == OCMP_OP_Z) v = dim > 2 ? x[2] : 0.0;
      else v = t;
      s7 = s6; s6 = s5; s5 = s4; s4 = s3; s3 = s2; s2 = s1; s1 = s0; s0 = v;
    } else if (op >= OCMP_OP_ADD && op <= OCMP_OP_POW) {
      const double a = s1, b = s0;
      double r;
      if (op == OCMP_OP_ADD) r = a + b;
      else if (op == OCMP_OP_SUB) r = a - b;
      else if (op == OCMP_OP_MUL) r = a * b;
      else if (op == OCMP_OP_DIV) r = a / b;
      else r = pow(a, b);
      s0 = r; s1 = s2; s2 = s3; s3 = s4; s4 = s5; s5 = s6; s6 = s7;
    } else if (op == OCMP_OP_NEG || (op >= OCMP_OP_SIN && op <= OCMP_OP_ROUND) || op == OCMP_OP_HS ||
               op == OCMP_OP_RAMP) {
      const double a = s0;
      double r;
      switch (op) {
        case OCMP_OP_NEG: r = -a; break;
        case OCMP_OP_SIN: r = sin(a); break;
        case OCMP_OP_COS: r = cos(a); break;
        case OCMP_OP_TAN: r = tan(a); break;
        case OCMP_OP_EXP: r = exp(a); break;
        case OCMP_OP_SQRT: r = sqrt(a); break;
        case OCMP_OP_TANH: r = ref_tanh(a); break;
        case OCMP_OP_SIG: r = ref_sig(a); break;
        case OCMP_OP_ABS: r = fabs(a); break;
        case OCMP_OP_SGN: r = a < -1e-12 ? -1.0 : (a > 1e-12 ? 1.0 : 0.0); break;
        case OCMP_OP_TRUNC: r = trunc(a); break;
        case OCMP_OP_ROUND: r = rint(a); break;
        case OCMP_OP_HS: r = ref_sig(2.0 * 9.21025 / 0.1 * a - 9.21025); break;  // H_s(t), w = 0.1
        default: r = 1.0 - a > 0.0 ? 0.5 * (0.0 - 1.0) * cos(a * M_PI) + 0.5 : 1.0;  // ramp_cos(t)
      }
      s0 = r;
    } else if (op == OCMP_OP_HS2) {  // H_s(t, w)
      const double r = ref_sig(2.0 * 9.21025 / s0 * s1 - 9.21025);
      s0 = r; s1 = s2; s2 = s3; s3 = s4; s4 = s5; s5 = s6; s6 = s7;
    } else {  // OCMP_OP_RAMP4: ramp_cos(t, p1, p2, tr), opencmp/helpers/math.py:116-131
      const double tt = s3, p1 = s2, p2 = s1, tr = s0;
      const double r = tr - tt > 0.0 ? 0.5 * (p1 - p2) * cos(tt * M_PI / tr) + 0.5 * (p1 + p2) : p2;
      s0 = r; s1 = s4; s2 = s5; s3 = s6; s4 = s7;
    }
  }
}

__device__ __forceinline__ double eval_fn(int fn, const double *c, const double *x, int dim) {
  if (fn == OCMP_FN_CONST) return c[0];
  if (fn == OCMP_FN_SINPROD) {
    double s = c[0];
    for (int d = 0; d < dim; ++d) s *= sin(M_PI * x[d]);
    return s;
  }
  if (fn == OCMP_FN_EXPR) return eval_expr((int)c[0], c[1], x, dim) * c[2] + c[3];
  // POLY2
  if (dim == 2) {
    double X = x[0], Y = x[1];
    return c[0] + c[1] * X + c[2] * Y + c[3] * X * X + c[4] * X * Y + c[5] * Y * Y;
  }
  double X = x[0], Y = x[1], Z = x[2];
  return c[0] + c[1] * X + c[2] * Y + c[3] * Z + c[4] * X * X + c[5] * Y * Y + c[6] * Z * Z + c[7] * X * Y +
         c[8] * X * Z + c[9] * Y * Z;
}

// value of a field (H1 block) at volume quadrature point q of element e
__device__ __forceinline__ double field_at(const KField &f, const double *__restrict__ tab, int64_t e, int q) {
  const double *phi = tab + f.tab_phi + q * f.nl;
  const int32_t *ed = f.edofs + e * f.ld + f.loc_off;
  double s = 0.0;
  for (int l = 0; l < f.nl; ++l) s += phi[l] * f.vec[ed[l]];
  return s;
}

// value of a field at facet-rule point q of local facet k of element e
__device__ __forceinline__ double field_at_facet(const KField &f, const double *__restrict__ tab, int64_t e, int k,
                                                 int q, int nqf) {
  const double *phi = tab + f.tab_fphi + (k * nqf + q) * f.nl;
  const int32_t *ed = f.edofs + e * f.ld + f.loc_off;
  double s = 0.0;
  for (int l = 0; l < f.nl; ++l) s += phi[l] * f.vec[ed[l]];
  return s;
}

// Row il (block bi) of coef * max(w.n, 0) u v (OCMP_BFI_BND_UPWIND: the no-backflow part of the
// stress boundary condition), coef * (w.n) u v (OCMP_BFI_BND_FLUX: the convective flux of a
// Neumann boundary) or of coef * u v (OCMP_BFI_BND_MASS: the Robin term) over the
// element's boundary facets whose marker is in T.dmask. Facet k is the edge (face) opposite vertex
// k, n its outward unit normal; in 2-D the rule maps [0, 1] onto the edge.
__host__ __device__ inline bool is_bnd_kind(int k) {
  return k == OCMP_BFI_BND_UPWIND || k == OCMP_BFI_BND_MASS || k == OCMP_BFI_BND_FLUX;
}

template <int D>
__device__ void bnd_facet_row(const KArgs &A, const KTerm &T, const Geo<D> &g, int64_t e, int bi, int il, int bj,
                               double ce, double *r) {
  if constexpr (D == 2) {
    const int nlt = A.b.nloc[bi], nlj = A.b.nloc[bj];
    double P[3][2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      P[0][c] = g.x0[c];
      P[1][c] = g.x0[c] + g.J[c][0];
      P[2][c] = g.x0[c] + g.J[c][1];
    }
    const double *fw = A.tab + A.tab_fw;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const int32_t nb = A.elem_nb[e * 3 + k];
      if (nb >= 0) continue;
      const int marker = -nb - 1;
      if (marker >= 64 || !((T.dmask >> marker) & 1ull)) continue;
      // edge k runs from vertex fa to vertex fb: (1, 2), (0, 2), (0, 1)
      const int fa = k == 0 ? 1 : 0, fb = k == 2 ? 1 : 2;
      const double tx = P[fb][0] - P[fa][0], ty = P[fb][1] - P[fa][1];
      const double len = sqrt(tx * tx + ty * ty);
      double nx = ty / len, ny = -tx / len;
      if ((P[k][0] - P[fa][0]) * nx + (P[k][1] - P[fa][1]) * ny > 0.0) {
        nx = -nx;
        ny = -ny;
      }
      const double *phit = A.tab + A.b.tab_fphi[bi] + (size_t)k * A.nqf * nlt;
      const double *phij = A.tab + A.b.tab_fphi[bj] + (size_t)k * A.nqf * nlj;
      for (int q = 0; q < A.nqf; ++q) {
        double wn = 1.0;
        if (T.kind != OCMP_BFI_BND_MASS) {
          double wx = T.wind[0], wy = T.wind[1];
          if (T.wind_field >= 0) {
            wx = field_at_facet(A.f[T.wind_field], A.tab, e, k, q, A.nqf);
            wy = field_at_facet(A.f[T.wind_field + 1], A.tab, e, k, q, A.nqf);
          }
          wn = wx * nx + wy * ny;
          if (T.kind == OCMP_BFI_BND_UPWIND && !(wn > 0.0)) continue;
        }
        const double c = ce * fw[q] * len * wn * phit[q * nlt + il];
        for (int j = 0; j < nlj; ++j) r[j] += c * phij[q * nlj + j];
      }
    }
  } else {
    // tetrahedra: face k has the vertices other than k, (a, b, c) in increasing order; the face rule's
    // weights sum to 1, so a face integral is its area times the weighted sum
    const int nlt = A.b.nloc[bi], nlj = A.b.nloc[bj];
    double P[4][3];
#pragma unroll
    for (int cc = 0; cc < 3; ++cc) {
      P[0][cc] = g.x0[cc];
#pragma unroll
      for (int i = 0; i < 3; ++i) P[i + 1][cc] = g.x0[cc] + g.J[cc][i];
    }
    const double *fw = A.tab + A.tab_fw;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int32_t nb = A.elem_nb[e * 4 + k];
      if (nb >= 0) continue;
      const int marker = -nb - 1;
      if (marker >= 64 || !((T.dmask >> marker) & 1ull)) continue;
      const int a = k == 0 ? 1 : 0, b = k <= 1 ? 2 : 1, c = k == 3 ? 2 : 3;
      double u[3], v[3], nrm[3];
#pragma unroll
      for (int cc = 0; cc < 3; ++cc) {
        u[cc] = P[b][cc] - P[a][cc];
        v[cc] = P[c][cc] - P[a][cc];
      }
      nrm[0] = u[1] * v[2] - u[2] * v[1];
      nrm[1] = u[2] * v[0] - u[0] * v[2];
      nrm[2] = u[0] * v[1] - u[1] * v[0];
      const double a2 = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      double n3[3] = {nrm[0] / a2, nrm[1] / a2, nrm[2] / a2};
      if ((P[k][0] - P[a][0]) * n3[0] + (P[k][1] - P[a][1]) * n3[1] + (P[k][2] - P[a][2]) * n3[2] > 0.0) {
        n3[0] = -n3[0];
        n3[1] = -n3[1];
        n3[2] = -n3[2];
      }
      const double area = 0.5 * a2;
      const double *phit = A.tab + A.b.tab_fphi[bi] + (size_t)k * A.nqf * nlt;
      const double *phij = A.tab + A.b.tab_fphi[bj] + (size_t)k * A.nqf * nlj;
      for (int q = 0; q < A.nqf; ++q) {
        double wn = 1.0;
        if (T.kind != OCMP_BFI_BND_MASS) {
          double w3[3] = {T.wind[0], T.wind[1], T.wind[2]};
          if (T.wind_field >= 0)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) w3[cc] = field_at_facet(A.f[T.wind_field + cc], A.tab, e, k, q, A.nqf);
          wn = w3[0] * n3[0] + w3[1] * n3[1] + w3[2] * n3[2];
          if (T.kind == OCMP_BFI_BND_UPWIND && !(wn > 0.0)) continue;
        }
        const double cq = ce * fw[q] * area * wn * phit[q * nlt + il];
        for (int j = 0; j < nlj; ++j) r[j] += cq * phij[q * nlj + j];
      }
    }
  }
}

// ------------------------------------------------------------------ H1 row kernel

template <int D>
__device__ void h1_row(const KArgs &A, const double *__restrict__ coords, const int32_t *__restrict__ elems, int64_t e,
                       int i, double *row) {
  Geo<D> g;
  geometry<D>(coords, elems + e * (D + 1), g);
  // which block is row i in
  int bi = 0;
  while (bi + 1 < A.b.n && i >= A.b.loc_off[bi + 1]) ++bi;
  const int il = i - A.b.loc_off[bi];
  const int nlt = A.b.nloc[bi];
  for (int j = 0; j < A.nloc; ++j) row[j] = 0.0;
  for (int t = 0; t < A.nterms; ++t) {
    const KTerm &T = A.t[t];
    if (T.test != bi) continue;
    const int bj = T.trial;
    const int nlj = A.b.nloc[bj];
    double *r = row + A.b.loc_off[bj];
    double ce = T.coef * (T.elem_coef ? T.elem_coef[e] : 1.0);
    if (is_bnd_kind(T.kind)) continue;  // k_h1_bnd_assemble
    if (!T.quad) {
      // reference-tensor contraction
      double fac[9];
      int nf = 0;
      if (T.kind == OCMP_BFI_GRADGRAD) {
#pragma unroll
        for (int a = 0; a < D; ++a)
#pragma unroll
          for (int b = 0; b < D; ++b) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < D; ++c) s += g.Jinv[a][c] * g.Jinv[b][c];
            fac[nf++] = s * g.adet;
          }
      } else if (T.kind == OCMP_BFI_MASS) {
        fac[nf++] = g.adet;
      } else if (T.kind == OCMP_BFI_CONV) {
#pragma unroll
        for (int a = 0; a < D; ++a) {
          double s = 0.0;
#pragma unroll
          for (int c = 0; c < D; ++c) s += g.Jinv[a][c] * T.wind[c];
          fac[nf++] = s * g.adet;
        }
      } else {  // DTRIAL / DTEST
#pragma unroll
        for (int a = 0; a < D; ++a) fac[nf++] = g.Jinv[a][T.dir] * g.adet;
      }
      const double *ten = A.tab + T.tab;
      const int tsz = nlt * nlj;
      for (int j = 0; j < nlj; ++j) {
        double s = 0.0;
        for (int k = 0; k < nf; ++k) s += fac[k] * ten[k * tsz + il * nlj + j];
        r[j] += ce * s;
      }
    } else {
      // quadrature with field coefficients
      const double *w = A.tab + A.tab_w;
      const double *phit = A.tab + A.b.tab_phi[bi];
      const double *grt = A.tab + A.b.tab_grad[bi];
      const double *phij = A.tab + A.b.tab_phi[bj];
      const double *grj = A.tab + A.b.tab_grad[bj];
      for (int q = 0; q < A.nq; ++q) {
        double c = ce * w[q] * g.adet;
        if (T.weight_mode != OCMP_WEIGHT_NONE) {
          double ph = field_at(A.f[T.weight_field], A.tab, e, q);
          c *= T.weight_mode == OCMP_WEIGHT_FIELD ? ph : (1.0 - ph);
        }
        // physical gradient of the test function
        double gi[D];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < D; ++a) s += g.Jinv[a][cc] * grt[(q * nlt + il) * D + a];
          gi[cc] = s;
        }
        const double pi = phit[q * nlt + il];
        if (T.kind == OCMP_BFI_GRADGRAD) {
          for (int j = 0; j < nlj; ++j) {
            double s = 0.0;
#pragma unroll
            for (int cc = 0; cc < D; ++cc) {
              double gj = 0.0;
#pragma unroll
              for (int a = 0; a < D; ++a) gj += g.Jinv[a][cc] * grj[(q * nlj + j) * D + a];
              s += gi[cc] * gj;
            }
            r[j] += c * s;
          }
        } else if (T.kind == OCMP_BFI_MASS) {
          for (int j = 0; j < nlj; ++j) r[j] += c * pi * phij[q * nlj + j];
        } else if (T.kind == OCMP_BFI_CONV) {
          double wv[D];
#pragma unroll
          for (int cc = 0; cc < D; ++cc) wv[cc] = T.wind_field >= 0 ? field_at(A.f[T.wind_field + cc], A.tab, e, q)
                                                                    : T.wind[cc];
          double wg = 0.0;
#pragma unroll
          for (int cc = 0; cc < D; ++cc) wg += wv[cc] * gi[cc];
          for (int j = 0; j < nlj; ++j) r[j] += c * wg * phij[q * nlj + j];
        } else if (T.kind == OCMP_BFI_DTEST) {
          for (int j = 0; j < nlj; ++j) r[j] += c * gi[T.dir] * phij[q * nlj + j];
        } else {  // DTRIAL
          for (int j = 0; j < nlj; ++j) {
            double gj = 0.0;
#pragma unroll
            for (int a = 0; a < D; ++a) gj += g.Jinv[a][T.dir] * grj[(q * nlj + j) * D + a];
            r[j] += c * pi * gj;
          }
        }
      }
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_h1_assemble(KArgs A, const double *__restrict__ coords,
                                                     const int32_t *__restrict__ elems,
                                                     const int32_t *__restrict__ colour_elems, int64_t cbeg,
                                                     int64_t ccount, const int32_t *__restrict__ slot,
                                                     double *__restrict__ val) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ccount * A.nloc) return;
  const int64_t e = colour_elems[cbeg + t / A.nloc];
  const int i = (int)(t % A.nloc);
  double row[MAXL];
  h1_row<D>(A, coords, elems, e, i, row);
  const int32_t *sl = slot + (e * A.nloc + i) * A.nloc;
  for (int j = 0; j < A.nloc; ++j) val[sl[j]] += row[j];
}

// the boundary-facet terms, over the elements of one colour that have a boundary facet: thread =
// (element, local row), which owns that row of the element's slots within the launch
template <int D>
__global__ void __launch_bounds__(256) k_h1_bnd_assemble(KArgs A, const double *__restrict__ coords,
                                                         const int32_t *__restrict__ elems,
                                                         const int32_t *__restrict__ bnd_elems, int64_t cbeg,
                                                         int64_t ccount, const int32_t *__restrict__ slot,
                                                         double *__restrict__ val) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ccount * A.nloc) return;
  const int64_t e = bnd_elems[cbeg + t / A.nloc];
  const int i = (int)(t % A.nloc);
  Geo<D> g;
  geometry<D>(coords, elems + e * (D + 1), g);
  int bi = 0;
  while (bi + 1 < A.b.n && i >= A.b.loc_off[bi + 1]) ++bi;
  const int il = i - A.b.loc_off[bi];
  double row[MAXL];
  for (int j = 0; j < A.nloc; ++j) row[j] = 0.0;
  bool any = false;
  for (int tt = 0; tt < A.nterms; ++tt) {
    const KTerm &T = A.t[tt];
    if (T.test != bi || !is_bnd_kind(T.kind)) continue;
    const double ce = T.coef * (T.elem_coef ? T.elem_coef[e] : 1.0);
    bnd_facet_row<D>(A, T, g, e, bi, il, T.trial, ce, row + A.b.loc_off[T.trial]);
    any = true;
  }
  if (!any) return;
  const int32_t *sl = slot + (e * A.nloc + i) * A.nloc;
  for (int j = 0; j < A.nloc; ++j)
    if (row[j] != 0.0) val[sl[j]] += row[j];
}

__global__ void k_fill(double *__restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ H1 volume terms: element factors + row gather
//
// Every volume term on affine simplices is a contraction K_e = sum_f F_e[f] T_f of per-element
// factors with reference tables (nl_test x nl_trial each):
//   grad-grad   F = c |det J| (J^-1 J^-T)_ab, a <= b        T = int (dphi_i/dxh_a dphi_j/dxh_b
//                                                                  + [a < b] dphi_i/dxh_b dphi_j/dxh_a)
//   mass        F = c |det J|                               T = int phi_i phi_j
//   convection / d_dir test / d_dir trial (constant vector v: the wind, or e_dir)
//               F = c |det J| (J^-1 v)_a                    T = int dphi_i/dxh_a phi_j  (test)
//                                                           T = int phi_i dphi_j/dxh_a  (trial)
//   a weight field w = sum_m w_m psi_m (diffuse-interface phi, or 1 - phi) multiplies F by w_m and
//   T by psi_m; a wind field W = sum_m W_m psi_m gives F = c |det J| (J^-1 W_m)_a,
//   T = int psi_m dphi_i/dxh_a phi_j.
// The tables are integrated exactly on the reference element, so field coefficients carry no
// quadrature error beyond the field's own representation.
//
// k_h1_factors: one thread per element computes every factor set once and writes it into the
// element's factor row, where each (test block, trial block) pair finds its factors contiguous.
// k_h1_gather: one workgroup per row block; its element entries are stored sorted by local row
// index L. Phase A, for each L: the entries' element-matrix rows V[q][j] = sum_f F_e[f] TD[L][f][j]
// (TD: every block pair's tables at that pair's factor positions, zero elsewhere) as FP64 MFMA
// tiles of 16 entries x 16 columns, staged in LDS with each value's column position. Phase B: a
// group of nloc lanes per row walks the row's entries in element order, adding each lane's value
// at its column's position in an LDS copy of the row (the columns of one entry are distinct). The
// block's rows, one contiguous CSR range, are then written once. No colouring, no
// read-modify-write of CSR values, a fixed summation order (bitwise reproducible).

constexpr int MAXFS = 16;  // factor sets per launch
enum { FS_GG = 0, FS_MASS = 1, FS_VEC = 2, FS_WIND = 3 };

struct KFactorSet {
  int kind, nf, ndst;
  int dst[4];        // offsets in the factor row where this set's nf values go
  int weight_mode;   // OCMP_WEIGHT_*
  int wfield;        // weight field (weight modes) or first wind field (FS_WIND)
  double coef;
  const double *elem_coef;
  double vec[3];
};

struct KFactors {
  int nsets, NF;
  KFactorSet fs[MAXFS];
  KField f[MAXF];
};

struct KPairG {
  int nfp, fbase, toff, nli, nlj;
};

struct KGather {
  int NL, NF, nb, accumulate, emax, accmax;
  int NFP, NLP;     // NF rounded up to 4 (the MFMA's K step), NL rounded up to 16 (its N)
  int td;           // offset in tab of TD[L][f][j] (NL x NFP x NLP, zero where factor f is not
                    // one of the (block(L), block(j)) pair's)
  int loc_off[MAXB + 1];
  KPairG pg[MAXB * MAXB];
  const double *tab;
};

typedef double f64x4 __attribute__((ext_vector_type(4)));

// block id with consecutive ids on one XCD (gridDim.x is a multiple of 8; blocks b, b+8, ... of a
// launch run on one XCD), so neighbouring row blocks share an L2
__device__ __forceinline__ int64_t xcd_block_id() {
  const int64_t per = gridDim.x / 8;
  return (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// local coefficient m of a field on element e
__device__ __forceinline__ double field_coef(const KField &f, int64_t e, int m) {
  return f.vec[f.edofs[e * f.ld + f.loc_off + m]];
}

template <int D>
__global__ void __launch_bounds__(256) k_h1_factors(const KFactors *__restrict__ Kp, int64_t nel,
                                                    const double *__restrict__ coords,
                                                    const int32_t *__restrict__ elems, double *__restrict__ F) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  const KFactors &K = *Kp;
  Geo<D> g;
  geometry<D>(coords, elems + e * (D + 1), g);
  double *Fe = F + e * K.NF;
  constexpr int NS = D * (D + 1) / 2;
  for (int si = 0; si < K.nsets; ++si) {
    const KFactorSet &S = K.fs[si];
    const double c = S.coef * (S.elem_coef ? S.elem_coef[e] : 1.0) * g.adet;
    if (S.kind == FS_WIND) {  // F[m D + a] = c (J^-1 W_m)_a
      const KField &f0 = K.f[S.wfield];
      for (int m = 0; m < f0.nl; ++m) {
        double w[D];
#pragma unroll
        for (int cc = 0; cc < D; ++cc) w[cc] = field_coef(K.f[S.wfield + cc], e, m);
#pragma unroll
        for (int a = 0; a < D; ++a) {
          double sacc = 0.0;
#pragma unroll
          for (int cc = 0; cc < D; ++cc) sacc += g.Jinv[a][cc] * w[cc];
          for (int dd = 0; dd < S.ndst; ++dd) Fe[S.dst[dd] + m * D + a] = c * sacc;
        }
      }
      continue;
    }
    double base[NS];
    int nb;
    if (S.kind == FS_GG) {
      nb = NS;
      int k = 0;
#pragma unroll
      for (int a = 0; a < D; ++a)
#pragma unroll
        for (int b = a; b < D; ++b) {
          double mm = 0.0;
#pragma unroll
          for (int cc = 0; cc < D; ++cc) mm += g.Jinv[a][cc] * g.Jinv[b][cc];
          base[k++] = c * mm;
        }
    } else if (S.kind == FS_MASS) {
      nb = 1;
#pragma unroll
      for (int k = 0; k < NS; ++k) base[k] = k == 0 ? c : 0.0;
    } else {  // FS_VEC
      nb = D;
#pragma unroll
      for (int a = 0; a < NS; ++a) {
        double mm = 0.0;
        if (a < D)
#pragma unroll
          for (int cc = 0; cc < D; ++cc) mm += g.Jinv[a][cc] * S.vec[cc];
        base[a] = c * mm;
      }
    }
    const int nw = S.weight_mode == OCMP_WEIGHT_NONE ? 1 : K.f[S.wfield].nl;
    for (int m = 0; m < nw; ++m) {
      double wm = 1.0;
      if (S.weight_mode != OCMP_WEIGHT_NONE) {
        wm = field_coef(K.f[S.wfield], e, m);
        // 1 - w: the constant 1 is the sum of the vertex functions (the first D+1 of an H1 basis)
        if (S.weight_mode == OCMP_WEIGHT_ONE_MINUS_FIELD) wm = (m <= D ? 1.0 : 0.0) - wm;
      }
      for (int dd = 0; dd < S.ndst; ++dd)
#pragma unroll
        for (int k = 0; k < NS; ++k)
          if (k < nb) Fe[S.dst[dd] + m * nb + k] = wm * base[k];
    }
  }
}

template <typename SlotT>
__global__ void __launch_bounds__(256) k_h1_gather(const KGather *__restrict__ Kp, int64_t nblk,
                                                   const int32_t *__restrict__ gblk, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ pptr, const int32_t *__restrict__ pe,
                                                   const uint16_t *__restrict__ gloff,
                                                   const uint16_t *__restrict__ gpos,
                                                   const SlotT *__restrict__ pslot, const double *__restrict__ F,
                                                   double *__restrict__ val) {
  extern __shared__ double lds_rows[];
  const KGather &K = *Kp;
  const int64_t b = xcd_block_id();
  if (b >= nblk) return;
  const int NL = K.NL;
  const int r0 = gblk[b], r1 = gblk[b + 1];
  const int e0 = pptr[r0], ne = pptr[r1] - e0;
  const int z0 = rp[r0], nz = rp[r1] - z0;
  double *V = lds_rows;                                     // ne * NL staged values (sorted entries)
  double *acc = lds_rows + (size_t)K.emax * NL;             // nz row values
  SlotT *S = reinterpret_cast<SlotT *>(acc + K.accmax);     // ne * NL positions within the row
  uint16_t *Gp = reinterpret_cast<uint16_t *>(S + (size_t)K.emax * NL);  // ne sorted positions
  for (int t = threadIdx.x; t < nz; t += blockDim.x) acc[t] = 0.0;
  for (int t = threadIdx.x; t < ne; t += blockDim.x) Gp[t] = gpos[e0 + t];
  // phase A: the block's element-matrix rows as FP64 MFMA tiles. The entries with local row L
  // (sorted: [go[L], go[L+1])) give V[q][j] = sum_f F[e_q][f] TD[L][f][j], a (entries x NF) by
  // (NF x NL) product; a 16 x 16 tile takes NFP / 4 v_mfma_f64_16x16x4 steps. Lane l holds the
  // A operand (entry l % 16, factor k0 + l / 16), the B operand (factor k0 + l / 16, column
  // l % 16) and, after the steps, the results of entries 4 r + l / 16 (r = 0..3) at column l % 16.
  // The tiles of all L are dealt round-robin to the block's waves.
  {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const uint16_t *go = gloff + b * (NL + 1);
    const int ntn = K.NLP >> 4;
    int tbase = 0;
    for (int L = 0; L < NL; ++L) {
      const int g0 = go[L], cnt = go[L + 1] - g0;
      const int nt = ((cnt + 15) >> 4) * ntn;
      const double *__restrict__ TL = K.tab + K.td + (size_t)L * K.NFP * K.NLP;
      for (int t = ((wv - tbase) % nw + nw) % nw; t < nt; t += nw) {
        const int mt = t / ntn, n0 = (t - mt * ntn) << 4;
        const bool va = mt * 16 + li < cnt;
        const double *__restrict__ Fe = F + (va ? (int64_t)pe[e0 + g0 + mt * 16 + li] : 0) * K.NF;
        f64x4 d = {0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < K.NFP; k0 += 4) {
          const int k = k0 + lk;
          const double av = (va && k < K.NF) ? Fe[k] : 0.0;
          const double bv = TL[k * K.NLP + n0 + li];
          d = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, d, 0, 0, 0);
        }
        const int col = n0 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = mt * 16 + 4 * r + lk;
          if (row < cnt && col < NL) {
            const int tq = (g0 + row) * NL + col;
            V[tq] = d[r];
            S[tq] = pslot[(int64_t)e0 * NL + tq];
          }
        }
      }
      tbase += nt;
    }
  }
  __syncthreads();
  // phase B: groups of NL lanes within a wave, one row each
  {
    const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Gw = 64 / NL;
    const int grp = lane / NL, lj = lane - grp * NL;
    const int nw = blockDim.x >> 6;
    if (grp < Gw)
      for (int r = r0 + wl * Gw + grp; r < r1; r += nw * Gw) {
        double *ar = acc + (rp[r] - z0);
        const int p1 = pptr[r + 1] - e0;
        for (int p = pptr[r] - e0; p < p1; ++p) {
          const int t = Gp[p] * NL + lj;
          ar[S[t]] += V[t];
        }
      }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nz; t += blockDim.x) {
    double x = acc[t];
    if (K.accumulate) x += val[z0 + t];
    val[z0 + t] = x;
  }
}

// ------------------------------------------------------------------ DG owner-computes kernel (triangles)

struct KDG {
  int nb;          // dofs per element
  int nq;          // facet quadrature points
  int tab_fphi;    // facet phi: [k][q][l]
  int tab_fgrad;   // facet reference gradients: [k][q][l][2]
  int tab_fw;      // facet weights (sum 1)
  int has_facet;
  double eps, alpha;
  double wind[2];
  unsigned long long dmask;
  int sipg, upwind;
};

__global__ void __launch_bounds__(256) k_dg_assemble(KArgs A, KDG F, int64_t nel, const double *__restrict__ coords,
                                                     const int32_t *__restrict__ elems,
                                                     const int32_t *__restrict__ elem_nb,
                                                     const int8_t *__restrict__ elem_nbk,
                                                     const int8_t *__restrict__ dg_pos,
                                                     const int32_t *__restrict__ row_ptr, double *__restrict__ val,
                                                     int accumulate) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nb = F.nb;
  if (t >= nel * nb) return;
  const int64_t e = t / nb;
  const int i = (int)(t % nb);
  double diag[MAXNB];
  h1_row<2>(A, coords, elems, e, i, diag);  // volume terms (contraction tables of the L2 block)
  double off[3][MAXNB];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    for (int j = 0; j < nb; ++j) off[k][j] = 0.0;
  if (F.has_facet) {
    Geo<2> g;
    geometry<2>(coords, elems + e * 3, g);
    const double h_e = sqrt(g.adet);
    const double *fphi = A.tab + F.tab_fphi;
    const double *fgr = A.tab + F.tab_fgrad;
    const double *fw = A.tab + F.tab_fw;
    for (int k = 0; k < 3; ++k) {
      // outward normal: -grad lambda_k / |grad lambda_k|
      double gl[2];
      if (k == 0) {
        gl[0] = -(g.Jinv[0][0] + g.Jinv[1][0]);
        gl[1] = -(g.Jinv[0][1] + g.Jinv[1][1]);
      } else {
        gl[0] = g.Jinv[k - 1][0];
        gl[1] = g.Jinv[k - 1][1];
      }
      const double gn = sqrt(gl[0] * gl[0] + gl[1] * gl[1]);
      const double n0 = -gl[0] / gn, n1 = -gl[1] / gn;
      // facet length = 2 |T| |grad lambda_k| ... |F| = |det J| * |grad lambda_k| (2D)
      const double flen = g.adet * gn;
      const double wn = F.wind[0] * n0 + F.wind[1] * n1;
      const int32_t nbv = elem_nb[e * 3 + k];
      if (nbv >= 0) {
        const int kk = elem_nbk[e * 3 + k];
        Geo<2> g2;
        geometry<2>(coords, elems + (int64_t)nbv * 3, g2);
        const double hF = 0.5 * (h_e + sqrt(g2.adet));
        const double al = F.alpha / hF;
        for (int q = 0; q < F.nq; ++q) {
          const double ww = fw[q] * flen;
          const double *ph = fphi + (k * F.nq + q) * nb;
          const double *gr = fgr + (k * F.nq + q) * nb * 2;
          const double *ph2 = fphi + (kk * F.nq + q) * nb;
          const double *gr2 = fgr + (kk * F.nq + q) * nb * 2;
          const double pi = ph[i];
          const double dni = (g.Jinv[0][0] * gr[2 * i] + g.Jinv[1][0] * gr[2 * i + 1]) * n0 +
                             (g.Jinv[0][1] * gr[2 * i] + g.Jinv[1][1] * gr[2 * i + 1]) * n1;
          const double up = wn > 0.0 ? wn : 0.0, dn = wn < 0.0 ? wn : 0.0;
          for (int j = 0; j < nb; ++j) {
            const double pj = ph[j];
            const double dnj = (g.Jinv[0][0] * gr[2 * j] + g.Jinv[1][0] * gr[2 * j + 1]) * n0 +
                               (g.Jinv[0][1] * gr[2 * j] + g.Jinv[1][1] * gr[2 * j + 1]) * n1;
            const double qj = ph2[j];
            const double dqj = (g2.Jinv[0][0] * gr2[2 * j] + g2.Jinv[1][0] * gr2[2 * j + 1]) * n0 +
                               (g2.Jinv[0][1] * gr2[2 * j] + g2.Jinv[1][1] * gr2[2 * j + 1]) * n1;
            double vd = 0.0, vo = 0.0;
            if (F.sipg) {
              vd += F.eps * (al * pi * pj - 0.5 * dnj * pi - 0.5 * dni * pj);
              vo += F.eps * (-al * pi * qj - 0.5 * dqj * pi + 0.5 * dni * qj);
            }
            if (F.upwind) {
              vd += up * pi * pj;
              vo += dn * pi * qj;
            }
            diag[j] += ww * vd;
            off[k][j] += ww * vo;
          }
        }
      } else {
        const int mk = -1 - nbv;
        const bool dir = mk < 64 && ((F.dmask >> mk) & 1ull);
        const double al = F.alpha / h_e;
        const double up = wn > 0.0 ? wn : 0.0;
        for (int q = 0; q < F.nq; ++q) {
          const double ww = fw[q] * flen;
          const double *ph = fphi + (k * F.nq + q) * nb;
          const double *gr = fgr + (k * F.nq + q) * nb * 2;
          const double pi = ph[i];
          const double dni = (g.Jinv[0][0] * gr[2 * i] + g.Jinv[1][0] * gr[2 * i + 1]) * n0 +
                             (g.Jinv[0][1] * gr[2 * i] + g.Jinv[1][1] * gr[2 * i + 1]) * n1;
          for (int j = 0; j < nb; ++j) {
            const double pj = ph[j];
            double v = 0.0;
            if (F.sipg && dir) {
              const double dnj = (g.Jinv[0][0] * gr[2 * j] + g.Jinv[1][0] * gr[2 * j + 1]) * n0 +
                                 (g.Jinv[0][1] * gr[2 * j] + g.Jinv[1][1] * gr[2 * j + 1]) * n1;
              v += F.eps * (al * pi * pj - pi * dnj - pj * dni);
            }
            if (F.upwind) v += up * pi * pj;
            diag[j] += ww * v;
          }
        }
      }
    }
  }
  // write the owned CSR row
  const int64_t r = e * nb + i;
  double *rv = val + row_ptr[r];
  const int8_t *pos = dg_pos + e * 4;
  if (accumulate) {
    for (int j = 0; j < nb; ++j) rv[pos[0] * nb + j] += diag[j];
    for (int k = 0; k < 3; ++k)
      if (pos[1 + k] >= 0)
        for (int j = 0; j < nb; ++j) rv[pos[1 + k] * nb + j] += off[k][j];
  } else {
    for (int j = 0; j < nb; ++j) rv[pos[0] * nb + j] = diag[j];
    for (int k = 0; k < 3; ++k)
      if (pos[1 + k] >= 0)
        for (int j = 0; j < nb; ++j) rv[pos[1 + k] * nb + j] = off[k][j];
  }
}

// ------------------------------------------------------------------ linear forms

struct KLTerm {
  int kind, test, fn, field;
  double c[10];
  double coef;
  int weight_mode, weight_field;
  double eps, alpha;
  double wind[3];
  unsigned long long dmask;
};

struct KLArgs {
  int nterms, nfields, nq, nloc;
  KLTerm t[MAXT];
  KField f[MAXF];
  KBlocks b;
  int tab_w, tab_lam;
  // DG facet data (L2 spaces)
  int fnq, tab_fphi, tab_fgrad, tab_fw, tab_ft;
  // H1 boundary sources: edge rule (bnq points, weights at tab_bfw, points at tab_bft) with the
  // blocks' phi at it in KBlocks::tab_fphi
  int bnq, tab_bfw, tab_bft;
  const double *tab;
};

template <int D>
__device__ void lin_local(const KLArgs &A, const double *__restrict__ coords, const int32_t *__restrict__ elems,
                          const int32_t *__restrict__ elem_nb, int64_t e, double *loc) {
  Geo<D> g;
  geometry<D>(coords, elems + e * (D + 1), g);
  for (int j = 0; j < A.nloc; ++j) loc[j] = 0.0;
  const double *w = A.tab + A.tab_w;
  const double *lam = A.tab + A.tab_lam;
  for (int t = 0; t < A.nterms; ++t) {
    const KLTerm &T = A.t[t];
    const int bi = T.test;
    const int nl = A.b.nloc[bi];
    double *r = loc + A.b.loc_off[bi];
    if (T.kind == OCMP_LFI_SOURCE) {
      const double *phi = A.tab + A.b.tab_phi[bi];
      for (int q = 0; q < A.nq; ++q) {
        double x[D];
#pragma unroll
        for (int c = 0; c < D; ++c) {
          double s = g.x0[c];
#pragma unroll
          for (int a = 0; a < D; ++a) s += g.J[c][a] * lam[q * (D + 1) + a + 1];
          x[c] = s;
        }
        double fv = T.fn == OCMP_FN_FIELD ? T.c[0] * field_at(A.f[T.field], A.tab, e, q) : eval_fn(T.fn, T.c, x, D);
        double c = T.coef * fv * w[q] * g.adet;
        if (T.weight_mode != OCMP_WEIGHT_NONE) {
          double ph = field_at(A.f[T.weight_field], A.tab, e, q);
          c *= T.weight_mode == OCMP_WEIGHT_FIELD ? ph : (1.0 - ph);
        }
        for (int j = 0; j < nl; ++j) r[j] += c * phi[q * nl + j];
      }
    } else if (T.kind == OCMP_LFI_BND_SOURCE) {
      // coef * g * v on the element's marked boundary edges (edge k opposite vertex k, from vertex
      // fa to vertex fb as in the facet tables)
      if constexpr (D == 2) {
        const double *fw = A.tab + A.tab_bfw;
        const double *ft = A.tab + A.tab_bft;
        const double *phi0 = A.tab + A.b.tab_fphi[bi];
        const int32_t *v = elems + e * 3;
        for (int k = 0; k < 3; ++k) {
          const int32_t nbv = elem_nb[e * 3 + k];
          if (nbv >= 0) continue;
          const int mk = -1 - nbv;
          if (mk >= 64 || !((T.dmask >> mk) & 1ull)) continue;
          const int fa = k == 0 ? 1 : 0, fb = k == 2 ? 1 : 2;
          const double ax = coords[(int64_t)v[fa] * 2], ay = coords[(int64_t)v[fa] * 2 + 1];
          const double bx = coords[(int64_t)v[fb] * 2], by = coords[(int64_t)v[fb] * 2 + 1];
          const double len = sqrt((bx - ax) * (bx - ax) + (by - ay) * (by - ay));
          for (int q = 0; q < A.bnq; ++q) {
            const double x[2] = {(1.0 - ft[q]) * ax + ft[q] * bx, (1.0 - ft[q]) * ay + ft[q] * by};
            const double c = T.coef * eval_fn(T.fn, T.c, x, 2) * fw[q] * len;
            const double *ph = phi0 + (k * A.bnq + q) * nl;
            for (int j = 0; j < nl; ++j) r[j] += c * ph[j];
          }
        }
      } else {
        // tetrahedra: face k has the vertices other than k in increasing order; tab_bft holds the
        // face rule's barycentrics (3 per point), the weights sum to 1
        const double *fw = A.tab + A.tab_bfw;
        const double *fl = A.tab + A.tab_bft;
        const double *phi0 = A.tab + A.b.tab_fphi[bi];
        const int32_t *v = elems + e * 4;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int32_t nbv = elem_nb[e * 4 + k];
          if (nbv >= 0) continue;
          const int mk = -1 - nbv;
          if (mk >= 64 || !((T.dmask >> mk) & 1ull)) continue;
          const int a = k == 0 ? 1 : 0, b = k <= 1 ? 2 : 1, c = k == 3 ? 2 : 3;
          double Pa[3], Pb[3], Pc[3];
#pragma unroll
          for (int cc = 0; cc < 3; ++cc) {
            Pa[cc] = coords[(int64_t)v[a] * 3 + cc];
            Pb[cc] = coords[(int64_t)v[b] * 3 + cc];
            Pc[cc] = coords[(int64_t)v[c] * 3 + cc];
          }
          const double u0 = Pb[0] - Pa[0], u1 = Pb[1] - Pa[1], u2 = Pb[2] - Pa[2];
          const double w0 = Pc[0] - Pa[0], w1 = Pc[1] - Pa[1], w2 = Pc[2] - Pa[2];
          const double n0 = u1 * w2 - u2 * w1, n1 = u2 * w0 - u0 * w2, n2 = u0 * w1 - u1 * w0;
          const double area = 0.5 * sqrt(n0 * n0 + n1 * n1 + n2 * n2);
          for (int q = 0; q < A.bnq; ++q) {
            double x[3];
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) x[cc] = fl[3 * q] * Pa[cc] + fl[3 * q + 1] * Pb[cc] + fl[3 * q + 2] * Pc[cc];
            const double cq = T.coef * eval_fn(T.fn, T.c, x, 3) * fw[q] * area;
            const double *ph = phi0 + (k * A.bnq + q) * nl;
            for (int j = 0; j < nl; ++j) r[j] += cq * ph[j];
          }
        }
      }
    } else if constexpr (D == 2) {
      // DG Dirichlet data on boundary facets of this element
      const double *fphi = A.tab + A.tab_fphi;
      const double *fgr = A.tab + A.tab_fgrad;
      const double *fw = A.tab + A.tab_fw;
      const double *ft = A.tab + A.tab_ft;
      const int32_t *v = elems + e * 3;
      for (int k = 0; k < 3; ++k) {
        const int32_t nbv = elem_nb[e * 3 + k];
        if (nbv >= 0) continue;
        const int mk = -1 - nbv;
        if (mk >= 64 || !((T.dmask >> mk) & 1ull)) continue;
        double gl[2];
        if (k == 0) {
          gl[0] = -(g.Jinv[0][0] + g.Jinv[1][0]);
          gl[1] = -(g.Jinv[0][1] + g.Jinv[1][1]);
        } else {
          gl[0] = g.Jinv[k - 1][0];
          gl[1] = g.Jinv[k - 1][1];
        }
        const double gn = sqrt(gl[0] * gl[0] + gl[1] * gl[1]);
        const double n0 = -gl[0] / gn, n1 = -gl[1] / gn;
        const double flen = g.adet * gn;
        const double al = T.alpha / sqrt(g.adet);
        const double wn = T.wind[0] * n0 + T.wind[1] * n1;
        const double dn = wn < 0.0 ? wn : 0.0;
        // facet k runs from local vertex a to b (sorted) = lower to higher global id
        const int a = k == 0 ? 1 : 0, b = k == 2 ? 1 : 2;
        for (int q = 0; q < A.fnq; ++q) {
          const double tq = ft[q];
          double x[2];
          for (int c = 0; c < 2; ++c)
            x[c] = (1.0 - tq) * coords[(int64_t)v[a] * 2 + c] + tq * coords[(int64_t)v[b] * 2 + c];
          const double gv = eval_fn(T.fn, T.c, x, 2) * T.coef;
          const double ww = fw[q] * flen;
          const double *ph = fphi + (k * A.fnq + q) * nl;
          const double *gr = fgr + (k * A.fnq + q) * nl * 2;
          for (int j = 0; j < nl; ++j) {
            const double dnj = (g.Jinv[0][0] * gr[2 * j] + g.Jinv[1][0] * gr[2 * j + 1]) * n0 +
                               (g.Jinv[0][1] * gr[2 * j] + g.Jinv[1][1] * gr[2 * j + 1]) * n1;
            r[j] += ww * gv * (T.eps * (al * ph[j] - dnj) - dn * ph[j]);
          }
        }
      }
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256) k_linear(KLArgs A, const double *__restrict__ coords,
                                                const int32_t *__restrict__ elems,
                                                const int32_t *__restrict__ elem_nb,
                                                const int32_t *__restrict__ elem_list, int64_t ebeg, int64_t ecount,
                                                const int32_t *__restrict__ edofs, double *__restrict__ b,
                                                int accumulate_owned) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ecount) return;
  const int64_t e = elem_list ? elem_list[ebeg + t] : ebeg + t;
  double loc[MAXL];
  lin_local<D>(A, coords, elems, elem_nb, e, loc);
  const int32_t *ed = edofs + e * A.nloc;
  if (elem_list || accumulate_owned) {
    for (int j = 0; j < A.nloc; ++j) b[ed[j]] += loc[j];
  } else {
    for (int j = 0; j < A.nloc; ++j) b[ed[j]] = loc[j];
  }
}

// ------------------------------------------------------------------ Dirichlet lift

// 6-point Gauss-Legendre rule on [0, 1]
__constant__ double GL6_T[6] = {0.033765242898423986, 0.16939530676686776, 0.38069040695840156,
                                0.61930959304159844, 0.83060469323313224, 0.96623475710157603};
__constant__ double GL6_W[6] = {0.085662246189585172, 0.18038078652406930, 0.23395696728634552,
                                0.23395696728634552, 0.18038078652406930, 0.085662246189585172};

__global__ void k_dir_verts(const int32_t *__restrict__ verts, int64_t nv, const double *__restrict__ coords, int dim,
                            int fn, const double *cc, int64_t off, double *__restrict__ x) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nv) return;
  double c[10];
  for (int i = 0; i < 10; ++i) c[i] = cc[i];
  const int32_t v = verts ? verts[t] : (int32_t)t;
  x[off + v] = eval_fn(fn, c, coords + (int64_t)v * dim, dim);
}

// 1D L2 projection onto the edge bubbles of (g - linear interpolant) along edge a->b (a < b)
__global__ void k_dir_edges(const int32_t *__restrict__ eids, int64_t ne, const int32_t *__restrict__ edges,
                            const double *__restrict__ coords, int dim, int fn, const double *cc, int order,
                            int64_t off, int64_t nvtx, double *__restrict__ x) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ne) return;
  double c[10];
  for (int i = 0; i < 10; ++i) c[i] = cc[i];
  const int32_t ei = eids ? eids[t] : (int32_t)t;
  const int32_t a = edges[2 * ei], b = edges[2 * ei + 1];
  double xa[3] = {0, 0, 0}, xb[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    xa[d] = coords[(int64_t)a * dim + d];
    xb[d] = coords[(int64_t)b * dim + d];
  }
  const double ga = eval_fn(fn, c, xa, dim), gb = eval_fn(fn, c, xb, dim);
  // 6-point Gauss-Legendre on [0,1]
  const double *gt = GL6_T, *gw = GL6_W;
  double r2 = 0.0, r3 = 0.0, m2 = 0.0, m3 = 0.0;
  for (int q = 0; q < 6; ++q) {
    const double s = gt[q];
    double xq[3];
    for (int d = 0; d < dim; ++d) xq[d] = (1.0 - s) * xa[d] + s * xb[d];
    const double res = eval_fn(fn, c, xq, dim) - ((1.0 - s) * ga + s * gb);
    const double b2 = (1.0 - s) * s;             // lambda_a lambda_b
    const double b3 = (1.0 - s) * s * (2.0 * s - 1.0);  // lambda_a lambda_b (lambda_b - lambda_a)
    r2 += gw[q] * res * b2;
    r3 += gw[q] * res * b3;
    m2 += gw[q] * b2 * b2;
    m3 += gw[q] * b3 * b3;
  }
  // b2 and b3 are orthogonal on [0,1]
  const int64_t base = off + nvtx + (int64_t)(order - 1) * ei;
  x[base] = r2 / m2;
  if (order >= 3) x[base + 1] = r3 / m3;
}

// The face-bubble dof (lambda_a lambda_b lambda_c, global vertex ids a < b < c) of triangles whose
// vertex and edge dofs are already set: the L2 projection over the triangle of g minus its vertex
// and edge parts onto the bubble. The 3D faces of a P3 lift or set, and the cell bubbles of a 2D P3
// set. fids: the triangles (nullptr: 0..nf-1), fverts: their vertex triples, fedges[3 t + k]: the
// edge ids of (a,b), (a,c), (b,c) of the t-th listed triangle. Collapsed 6x6 Gauss rule (exact to
// degree 10); the constant Jacobian of the triangle cancels in the quotient.
__global__ void k_set_faces(const int32_t *__restrict__ fids, int64_t nf, const int32_t *__restrict__ fverts,
                            const int32_t *__restrict__ fedges, const double *__restrict__ coords, int dim, int fn,
                            const double *cc, int64_t off, int64_t nvtx, int64_t face_base, double *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf) return;
  double c[10];
  for (int i = 0; i < 10; ++i) c[i] = cc[i];
  const int64_t f = fids ? fids[t] : t;
  const int32_t va = fverts[3 * f], vb = fverts[3 * f + 1], vc = fverts[3 * f + 2];
  double X[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  for (int d = 0; d < dim; ++d) {
    X[0][d] = coords[(int64_t)va * dim + d];
    X[1][d] = coords[(int64_t)vb * dim + d];
    X[2][d] = coords[(int64_t)vc * dim + d];
  }
  const double ga = x[off + va], gb = x[off + vb], gc = x[off + vc];
  double e2[3], e3[3];
  for (int k = 0; k < 3; ++k) {
    const int64_t eb = off + nvtx + 2 * (int64_t)fedges[3 * t + k];
    e2[k] = x[eb];
    e3[k] = x[eb + 1];
  }
  const double *gt = GL6_T, *gw = GL6_W;
  double num = 0.0, den = 0.0;
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double lb = gt[i], lc = gt[j] * (1.0 - gt[i]), la = 1.0 - lb - lc;
      const double w = gw[i] * gw[j] * (1.0 - gt[i]);
      double xq[3];
      for (int d = 0; d < 3; ++d) xq[d] = la * X[0][d] + lb * X[1][d] + lc * X[2][d];
      double r = eval_fn(fn, c, xq, dim) - (ga * la + gb * lb + gc * lc);
      r -= la * lb * (e2[0] + e3[0] * (lb - la));
      r -= la * lc * (e2[1] + e3[1] * (lc - la));
      r -= lb * lc * (e2[2] + e3[2] * (lc - lb));
      const double B = la * lb * lc;
      num += w * r * B;
      den += w * B * B;
    }
  x[face_base + f] = num / den;
}

// Element-wise L2 projection of g onto an L2 block (GridFunction.Set on a DG space):
// coefficients M_ref^{-1} sum_q w_q g(x_q) phi_j(q) per element (the element's |det J| cancels)
struct KL2Set {
  int nq, nl, ld, loc_off;
  int tab_phi, tab_w, tab_lam, tab_minv;
  int fn;
  double c[10];
  const double *tab;
};

template <int D>
__global__ void __launch_bounds__(256) k_l2_set(KL2Set K, int64_t nel, const double *__restrict__ coords,
                                                const int32_t *__restrict__ elems, const int32_t *__restrict__ edofs,
                                                double *__restrict__ x) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  Geo<D> g;
  geometry<D>(coords, elems + e * (D + 1), g);
  const double *phi = K.tab + K.tab_phi;
  const double *w = K.tab + K.tab_w;
  const double *lam = K.tab + K.tab_lam;
  const double *minv = K.tab + K.tab_minv;
  double r[MAXNB];
#pragma unroll
  for (int j = 0; j < MAXNB; ++j) r[j] = 0.0;
  for (int q = 0; q < K.nq; ++q) {
    double xq[D];
    for (int cc = 0; cc < D; ++cc) {
      double sx = g.x0[cc];
      for (int a = 0; a < D; ++a) sx += g.J[cc][a] * lam[q * (D + 1) + a + 1];
      xq[cc] = sx;
    }
    const double gv = w[q] * eval_fn(K.fn, K.c, xq, D);
#pragma unroll
    for (int j = 0; j < MAXNB; ++j)
      if (j < K.nl) r[j] += gv * phi[q * K.nl + j];
  }
  const int32_t *ed = edofs + e * K.ld + K.loc_off;
  for (int i = 0; i < K.nl; ++i) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < MAXNB; ++j)
      if (j < K.nl) v += minv[i * K.nl + j] * r[j];
    x[ed[i]] = v;
  }
}

// ------------------------------------------------------------------ error norms

__device__ __forceinline__ void eval_fn_grad(int fn, const double *c, const double *x, int dim, double *g) {
  for (int d = 0; d < dim; ++d) g[d] = 0.0;
  if (fn == OCMP_FN_EXPR) {
    // programs carry no derivative: the H1 seminorm against an expression is undefined
    for (int d = 0; d < dim; ++d) g[d] = __builtin_nan("");
  } else if (fn == OCMP_FN_SINPROD) {
    for (int d = 0; d < dim; ++d) {
      double s = c[0] * M_PI * cos(M_PI * x[d]);
      for (int e = 0; e < dim; ++e)
        if (e != d) s *= sin(M_PI * x[e]);
      g[d] = s;
    }
  } else if (fn == OCMP_FN_POLY2) {
    if (dim == 2) {
      g[0] = c[1] + 2.0 * c[3] * x[0] + c[4] * x[1];
      g[1] = c[2] + c[4] * x[0] + 2.0 * c[5] * x[1];
    } else {
      g[0] = c[1] + 2.0 * c[4] * x[0] + c[7] * x[1] + c[8] * x[2];
      g[1] = c[2] + 2.0 * c[5] * x[1] + c[7] * x[0] + c[9] * x[2];
      g[2] = c[3] + 2.0 * c[6] * x[2] + c[8] * x[0] + c[9] * x[1];
    }
  }
}

struct KNorm {
  int nq, nl, ld, loc_off;
  int tab_phi, tab_grad, tab_w, tab_lam;
  int fn;
  double c[10];
  const double *tab;
};

// per-block partial sums of int (u_h - g)^2 and int |grad(u_h - g)|^2
template <int D>
__global__ void __launch_bounds__(256) k_error_norms(KNorm K, int64_t nel, const double *__restrict__ coords,
                                                     const int32_t *__restrict__ elems,
                                                     const int32_t *__restrict__ edofs, const double *__restrict__ x,
                                                     const double *__restrict__ ref, double *__restrict__ part) {
  __shared__ double sh[2][4];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0, h2 = 0.0;
  if (e < nel) {
    Geo<D> g;
    geometry<D>(coords, elems + e * (D + 1), g);
    const int32_t *ed = edofs + e * K.ld + K.loc_off;
    const double *phi = K.tab + K.tab_phi;
    const double *gr = K.tab + K.tab_grad;
    const double *w = K.tab + K.tab_w;
    const double *lam = K.tab + K.tab_lam;
    for (int q = 0; q < K.nq; ++q) {
      double u = 0.0, gref[D];
      for (int a = 0; a < D; ++a) gref[a] = 0.0;
      for (int l = 0; l < K.nl; ++l) {
        const double d = x[ed[l]] - (ref ? ref[ed[l]] : 0.0);
        u += phi[q * K.nl + l] * d;
        for (int a = 0; a < D; ++a) gref[a] += gr[(q * K.nl + l) * D + a] * d;
      }
      double gu[D];
      for (int c = 0; c < D; ++c) {
        double s = 0.0;
        for (int a = 0; a < D; ++a) s += g.Jinv[a][c] * gref[a];
        gu[c] = s;
      }
      if (!ref) {
        double xq[D];
        for (int c = 0; c < D; ++c) {
          double s = g.x0[c];
          for (int a = 0; a < D; ++a) s += g.J[c][a] * lam[q * (D + 1) + a + 1];
          xq[c] = s;
        }
        u -= eval_fn(K.fn, K.c, xq, D);
        double gg[3];
        eval_fn_grad(K.fn, K.c, xq, D, gg);
        for (int c = 0; c < D; ++c) gu[c] -= gg[c];
      }
      const double wq = w[q] * g.adet;
      e2 += wq * u * u;
      for (int c = 0; c < D; ++c) h2 += wq * gu[c] * gu[c];
    }
  }
  // block reduction (fixed order)
  for (int o = 32; o > 0; o >>= 1) {
    e2 += __shfl_down(e2, o, 64);
    h2 += __shfl_down(h2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = e2;
    sh[1][threadIdx.x >> 6] = h2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// A vector component (ncomp <= 3 scalar blocks of one kind and order) against closed forms or a
// reference vector: per-block partial sums of int |u_h - g|^2, int |u_h - g| (the pointwise
// Euclidean norm: opencmp/helpers/error.py:70-72 takes sqrt((u - g).(u - g))) and
// int |grad(u_h - g)|^2
struct KNormV {
  int nq, nl, ld, ncomp;
  int loc_off[3];
  int tab_phi, tab_grad, tab_w, tab_lam;
  int fn[3];
  double c[3][10];
  const double *tab;
};

template <int D>
__global__ void __launch_bounds__(256) k_error_norms_vec(KNormV K, int64_t nel, const double *__restrict__ coords,
                                                         const int32_t *__restrict__ elems,
                                                         const int32_t *__restrict__ edofs,
                                                         const double *__restrict__ x, const double *__restrict__ ref,
                                                         double *__restrict__ part) {
  __shared__ double sh[3][4];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0, l1 = 0.0, h2 = 0.0;
  if (e < nel) {
    Geo<D> g;
    geometry<D>(coords, elems + e * (D + 1), g);
    const double *phi = K.tab + K.tab_phi;
    const double *gr = K.tab + K.tab_grad;
    const double *w = K.tab + K.tab_w;
    const double *lam = K.tab + K.tab_lam;
    for (int q = 0; q < K.nq; ++q) {
      double xq[D];
      for (int c = 0; c < D; ++c) {
        double sx = g.x0[c];
        for (int a = 0; a < D; ++a) sx += g.J[c][a] * lam[q * (D + 1) + a + 1];
        xq[c] = sx;
      }
      double sq = 0.0, hq = 0.0;
      for (int k = 0; k < K.ncomp; ++k) {
        const int32_t *ed = edofs + e * K.ld + K.loc_off[k];
        double u = 0.0, gref[D];
        for (int a = 0; a < D; ++a) gref[a] = 0.0;
        for (int l = 0; l < K.nl; ++l) {
          const double dv = x[ed[l]] - (ref ? ref[ed[l]] : 0.0);
          u += phi[q * K.nl + l] * dv;
          for (int a = 0; a < D; ++a) gref[a] += gr[(q * K.nl + l) * D + a] * dv;
        }
        double gu[D];
        for (int c = 0; c < D; ++c) {
          double sg = 0.0;
          for (int a = 0; a < D; ++a) sg += g.Jinv[a][c] * gref[a];
          gu[c] = sg;
        }
        if (!ref) {
          u -= eval_fn(K.fn[k], K.c[k], xq, D);
          double gg[3];
          eval_fn_grad(K.fn[k], K.c[k], xq, D, gg);
          for (int c = 0; c < D; ++c) gu[c] -= gg[c];
        }
        sq += u * u;
        for (int c = 0; c < D; ++c) hq += gu[c] * gu[c];
      }
      const double wq = w[q] * g.adet;
      e2 += wq * sq;
      l1 += wq * sqrt(sq);
      h2 += wq * hq;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    e2 += __shfl_down(e2, o, 64);
    l1 += __shfl_down(l1, o, 64);
    h2 += __shfl_down(h2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = e2;
    sh[1][threadIdx.x >> 6] = l1;
    sh[2][threadIdx.x >> 6] = h2;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = sh[k][0] + sh[k][1] + sh[k][2] + sh[k][3];
}

// The Picard stop test's two norms in one pass over the elements: per-block partial sums of
// int |a - b|^2 and int |b|^2 over the vector component's ncomp scalar blocks (a, b finite element
// vectors on the same space; the rule integrates their degree-2p products exactly). The element's
// dof values are held in registers (NLM >= nl) and the tables read with wave-uniform indices.
struct KPair {
  int nq, nl, ld, ncomp;
  int loc_off[3];
  int tab_phi, tab_w;
  const double *tab;
};

template <int D, int NLM>
__global__ void __launch_bounds__(256) k_l2_pair(KPair K, int64_t nel, const double *__restrict__ coords,
                                                 const int32_t *__restrict__ elems,
                                                 const int32_t *__restrict__ edofs, const double *__restrict__ a,
                                                 const double *__restrict__ b, double *__restrict__ part) {
  __shared__ double sh[2][4];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s_d = 0.0, s_b = 0.0;
  if (e < nel) {
    const double *phi = K.tab + K.tab_phi;
    const double *w = K.tab + K.tab_w;
    // int |a - b|^2 = sum over the scalar blocks of int (a_k - b_k)^2: one block at a time
    for (int k = 0; k < K.ncomp; ++k) {
      const int32_t *ed = edofs + e * K.ld + K.loc_off[k];
      double dv[NLM], bv[NLM];
#pragma unroll
      for (int l = 0; l < NLM; ++l) {
        dv[l] = 0.0;
        bv[l] = 0.0;
        if (l < K.nl) {
          const int32_t i = ed[l];
          bv[l] = b[i];
          dv[l] = a[i] - bv[l];
        }
      }
      for (int q = 0; q < K.nq; ++q) {
        double ud = 0.0, ub = 0.0;
#pragma unroll
        for (int l = 0; l < NLM; ++l) {
          const double f = l < K.nl ? phi[q * K.nl + l] : 0.0;
          ud += f * dv[l];
          ub += f * bv[l];
        }
        s_d += w[q] * ud * ud;
        s_b += w[q] * ub * ub;
      }
    }
    Geo<D> g;
    geometry<D>(coords, elems + e * (D + 1), g);
    s_d *= g.adet;
    s_b *= g.adet;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s_d += __shfl_down(s_d, o, 64);
    s_b += __shfl_down(s_b, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = s_d;
    sh[1][threadIdx.x >> 6] = s_b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[2 * blockIdx.x + 1] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// out = (sqrt(sum part[3i]), sum part[3i+1], sqrt(sum part[3i+2])), one block, fixed order
__global__ void k_sum_triples(const double *__restrict__ part, int64_t np, double *__restrict__ out) {
  __shared__ double sh[3][256];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[3 * i];
    b += part[3 * i + 1];
    c += part[3 * i + 2];
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  sh[2][threadIdx.x] = c;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sqrt(sh[0][0]);
    out[1] = sh[1][0];
    out[2] = sqrt(sh[2][0]);
  }
}

// per-block partial sums of int (div u_h)^2 for the vector field in the D consecutive scalar blocks
// starting at loc_off (same kind and order each, nl local dofs); the pair's second entry is 0
template <int D>
__global__ void __launch_bounds__(256) k_div_norm(KNorm K, int64_t nel, const double *__restrict__ coords,
                                                  const int32_t *__restrict__ elems,
                                                  const int32_t *__restrict__ edofs, const double *__restrict__ x,
                                                  double *__restrict__ part) {
  __shared__ double sh[4];
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s2 = 0.0;
  if (e < nel) {
    Geo<D> g;
    geometry<D>(coords, elems + e * (D + 1), g);
    const double *gr = K.tab + K.tab_grad;
    const double *w = K.tab + K.tab_w;
    for (int q = 0; q < K.nq; ++q) {
      double dv = 0.0;
      for (int c = 0; c < D; ++c) {
        const int32_t *ed = edofs + e * K.ld + K.loc_off + c * K.nl;
        // d u_c / d x_c = sum_a Jinv[a][c] d u_c / d xhat_a
        double gref[D];
        for (int a = 0; a < D; ++a) gref[a] = 0.0;
        for (int l = 0; l < K.nl; ++l) {
          const double xv = x[ed[l]];
          for (int a = 0; a < D; ++a) gref[a] += gr[(q * K.nl + l) * D + a] * xv;
        }
        for (int a = 0; a < D; ++a) dv += g.Jinv[a][c] * gref[a];
      }
      s2 += w[q] * g.adet * dv * dv;
    }
  }
  for (int o = 32; o > 0; o >>= 1) s2 += __shfl_down(s2, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s2;
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    part[2 * blockIdx.x + 1] = 0.0;
  }
}

// per-block maxima of |x_i - ref_i| over i in [0, n) (ref may be NULL)
__global__ void __launch_bounds__(256) k_absmax_part(int64_t n, const double *__restrict__ x,
                                                     const double *__restrict__ ref, double *__restrict__ part) {
  __shared__ double sh[4];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmax(m, fabs(x[i] - (ref ? ref[i] : 0.0)));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_down(m, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__global__ void k_max_of(const double *__restrict__ part, int np, double *__restrict__ out) {
  __shared__ double sh[256];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) m = fmax(m, part[i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

__global__ void k_sum_pairs(const double *__restrict__ part, int64_t np, double *__restrict__ out) {
  __shared__ double sh[2][256];
  double a = 0.0, b = 0.0;
  for (int64_t i = threadIdx.x; i < np; i += blockDim.x) {
    a += part[2 * i];
    b += part[2 * i + 1];
  }
  sh[0][threadIdx.x] = a;
  sh[1][threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sqrt(sh[0][0]);
    out[1] = sqrt(sh[1][0]);
  }
}

// ================================================================== host side

static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Tables for a set of blocks: volume rule, per-block phi/grad, contraction tensors.
struct TableBuilder {
  std::vector<double> v;
  int add(const std::vector<double> &x) {
    int o = (int)v.size();
    v.insert(v.end(), x.begin(), x.end());
    return o;
  }
};

static void fill_block_tables(TableBuilder &tb, const Space &s, const Rule &rule, KBlocks &kb) {
  const int d = s.mesh->dim;
  kb.n = (int)s.blocks.size();
  OCMP_REQUIRE(kb.n <= MAXB, "too many scalar blocks");
  for (int bi = 0; bi < kb.n; ++bi) {
    const Block &b = s.blocks[bi];
    kb.loc_off[bi] = b.loc_off;
    kb.nloc[bi] = b.nloc;
    std::vector<double> phi((size_t)rule.nq * b.nloc), gr((size_t)rule.nq * b.nloc * d);
    for (int q = 0; q < rule.nq; ++q)
      eval_basis(b.kind, d, b.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b.nloc], &gr[(size_t)q * b.nloc * d]);
    kb.tab_phi[bi] = tb.add(phi);
    kb.tab_grad[bi] = tb.add(gr);
  }
}

// reference tensors of a constant-coefficient term
static int add_tensors(TableBuilder &tb, const Space &s, const ocmp_term &t, int &ntens) {
  const int d = s.mesh->dim;
  const Block &bt = s.blocks[t.test], &bj = s.blocks[t.trial];
  Rule rule = simplex_rule(d, bt.order + bj.order + (bt.kind == OCMP_L2 || bj.kind == OCMP_L2 ? 2 : 0));
  const int nt = bt.nloc, nj = bj.nloc;
  std::vector<double> pt(nt), gt(nt * d), pj(nj), gj(nj * d);
  int nk = 0;
  switch (t.kind) {
    case OCMP_BFI_GRADGRAD: nk = d * d; break;
    case OCMP_BFI_MASS: nk = 1; break;
    default: nk = d; break;
  }
  std::vector<double> T((size_t)nk * nt * nj, 0.0);
  for (int q = 0; q < rule.nq; ++q) {
    eval_basis(bt.kind, d, bt.order, &rule.lam[q * (d + 1)], pt.data(), gt.data());
    eval_basis(bj.kind, d, bj.order, &rule.lam[q * (d + 1)], pj.data(), gj.data());
    const double w = rule.w[q];
    for (int i = 0; i < nt; ++i)
      for (int j = 0; j < nj; ++j) {
        double *o = &T[(size_t)i * nj + j];
        const size_t tsz = (size_t)nt * nj;
        if (t.kind == OCMP_BFI_GRADGRAD) {
          for (int a = 0; a < d; ++a)
            for (int b = 0; b < d; ++b) o[(a * d + b) * tsz] += w * gt[i * d + a] * gj[j * d + b];
        } else if (t.kind == OCMP_BFI_MASS) {
          o[0] += w * pt[i] * pj[j];
        } else if (t.kind == OCMP_BFI_CONV || t.kind == OCMP_BFI_DTEST) {
          for (int a = 0; a < d; ++a) o[a * tsz] += w * gt[i * d + a] * pj[j];
        } else {  // DTRIAL
          for (int a = 0; a < d; ++a) o[a * tsz] += w * pt[i] * gj[j * d + a];
        }
      }
  }
  ntens = nk;
  return tb.add(T);
}

static int max_order(const Space &s) {
  int p = 0;
  for (auto &b : s.blocks) p = std::max(p, b.order);
  return p;
}

static DBuf<double> *cached_table(Space &s, const std::string &key, const TableBuilder &tb, hipStream_t st) {
  auto it = s.tables.find(key);
  if (it != s.tables.end()) return it->second;
  auto *buf = new DBuf<double>();
  buf->upload(tb.v, st);
  OCMP_HIP(hipStreamSynchronize(st));  // host vector goes out of scope
  s.tables[key] = buf;
  return buf;
}

static std::string term_key(const std::vector<ocmp_term> &terms, const std::vector<ocmp_field> &fields) {
  std::string k = "B";
  for (auto &t : terms) {
    bool quad = t.weight_mode != OCMP_WEIGHT_NONE || t.wind_field >= 0;
    k += "|" + std::to_string(t.kind) + "," + std::to_string(t.trial) + "," + std::to_string(t.test) + "," +
         std::to_string(quad);
  }
  for (auto &f : fields) k += "|f" + std::to_string((uintptr_t)f.space) + "," + std::to_string(f.block);
  return k;
}

// phi of a (kind, order) block at the facet rules of all local facets (facet k opposite vertex k):
// triangles take the edge rule t (lam = 1 - t, t on the edge's two vertices in increasing order),
// tetrahedra the face rule's barycentrics on the face's three vertices in increasing order
static int add_facet_phi(TableBuilder &tb, int d, int kind, int order, int nl, int nq, const std::vector<double> &ft,
                         const std::vector<double> &flam) {
  const int nf = d + 1;
  std::vector<double> phi((size_t)nf * nq * nl), gr((size_t)nf * nq * nl * d);
  for (int k = 0; k < nf; ++k)
    for (int q = 0; q < nq; ++q) {
      double lam[4] = {0, 0, 0, 0};
      if (d == 2) {
        const int a = k == 0 ? 1 : 0, b = k == 2 ? 1 : 2;
        lam[a] = 1.0 - ft[q];
        lam[b] = ft[q];
      } else {
        const int a = k == 0 ? 1 : 0, b = k <= 1 ? 2 : 1, c = k == 3 ? 2 : 3;
        lam[a] = flam[3 * q];
        lam[b] = flam[3 * q + 1];
        lam[c] = flam[3 * q + 2];
      }
      eval_basis(kind, d, order, lam, &phi[((size_t)k * nq + q) * nl], &gr[((size_t)k * nq + q) * nl * d]);
    }
  return tb.add(phi);
}

static void build_kargs(Space &s, const std::vector<ocmp_term> &terms, const std::vector<ocmp_field> &fields,
                        KArgs &A, TableBuilder &tb, hipStream_t st) {
  const int d = s.mesh->dim;
  OCMP_REQUIRE((int)terms.size() <= MAXT, "too many terms in one assembly call (max 8)");
  OCMP_REQUIRE((int)fields.size() <= MAXF, "too many fields (max 8)");
  OCMP_REQUIRE(s.nloc <= MAXL, "compound element has too many dofs");
  std::memset(&A, 0, sizeof A);
  A.nterms = (int)terms.size();
  A.nfields = (int)fields.size();
  A.nloc = s.nloc;
  int pf = 0;
  for (auto &f : fields) pf = std::max(pf, f.space->s.blocks[f.block].order);
  Rule rule = simplex_rule(d, 2 * max_order(s) + pf + 1);
  A.nq = rule.nq;
  fill_block_tables(tb, s, rule, A.b);
  A.tab_w = tb.add(rule.w);
  A.tab_lam = tb.add(rule.lam);
  bool bnd = false;
  for (auto &t : terms) bnd |= is_bnd_kind(t.kind);
  // facet rule for boundary terms: exact for the product of three functions of the highest order
  std::vector<double> ft, fwt, flam;
  if (bnd) {
    const int deg = 2 * max_order(s) + std::max(pf, 1);
    if (d == 2) {
      line_rule(deg, ft, fwt);
      A.nqf = (int)ft.size();
    } else {
      Rule fr = simplex_rule(2, deg);
      A.nqf = fr.nq;
      flam = fr.lam;
      fwt = fr.w;
      for (auto &w : fwt) w *= 2.0;  // weights summing to 1: a face integral is area * sum
    }
    A.tab_fw = tb.add(fwt);
    A.elem_nb = s.mesh->d_elem_nb.p;
  }
  // phi of a (kind, order) block at the facet rule of each local facet (facet k opposite vertex k)
  auto facet_phi = [&](int kind, int order, int nl) { return add_facet_phi(tb, d, kind, order, nl, A.nqf, ft, flam); };
  if (bnd)
    for (int b = 0; b < A.b.n; ++b) A.b.tab_fphi[b] = facet_phi(s.blocks[b].kind, s.blocks[b].order, s.blocks[b].nloc);
  for (size_t i = 0; i < fields.size(); ++i) {
    const Space &fs = fields[i].space->s;
    const Block &fb = fs.blocks[fields[i].block];
    KField &kf = A.f[i];
    kf.vec = fields[i].vec;
    kf.edofs = fs.d_edofs.p;
    kf.ld = fs.nloc;
    kf.loc_off = fb.loc_off;
    kf.nl = fb.nloc;
    std::vector<double> phi((size_t)rule.nq * fb.nloc), gr((size_t)rule.nq * fb.nloc * d);
    for (int q = 0; q < rule.nq; ++q)
      eval_basis(fb.kind, d, fb.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * fb.nloc],
                 &gr[(size_t)q * fb.nloc * d]);
    kf.tab_phi = tb.add(phi);
    if (bnd) kf.tab_fphi = facet_phi(fb.kind, fb.order, fb.nloc);
  }
  for (size_t i = 0; i < terms.size(); ++i) {
    const ocmp_term &t = terms[i];
    KTerm &k = A.t[i];
    k.kind = t.kind;
    k.trial = t.trial;
    k.test = t.test;
    k.dir = t.dir;
    k.coef = t.coef;
    k.elem_coef = t.elem_coef;
    for (int c = 0; c < 3; ++c) k.wind[c] = t.wind[c];
    k.wind_field = t.wind_field;
    k.weight_mode = t.weight_mode;
    k.weight_field = t.weight_field;
    k.quad = (t.weight_mode != OCMP_WEIGHT_NONE || t.wind_field >= 0) ? 1 : 0;
    k.sipg = t.sipg;
    k.upwind = t.upwind;
    k.eps = t.eps;
    k.alpha = t.alpha;
    k.dmask = t.dirichlet_mask;
    if (t.kind == OCMP_BFI_DG_FACET) {
      k.test = -1;  // never matches a row block in the volume pass
      continue;
    }
    if (is_bnd_kind(t.kind)) {
      OCMP_REQUIRE(t.kind == OCMP_BFI_BND_MASS || t.wind_field < 0 || t.wind_field + 1 < (int)fields.size(),
                   "wind field out of range");
      continue;
    }
    if (!k.quad) k.tab = add_tensors(tb, s, t, k.ntens);
  }
}

// reference tables of one volume term in the contraction form (nf x nl_test x nl_trial), exact;
// grad-grad in the symmetric form (a <= b)
static std::vector<double> gather_tables(const Space &s, const ocmp_term &t, const std::vector<ocmp_field> &fields,
                                         int &nf) {
  const int d = s.mesh->dim;
  const Block &bt = s.blocks[t.test], &bj = s.blocks[t.trial];
  const bool wind_field = t.kind == OCMP_BFI_CONV && t.wind_field >= 0;
  const bool weighted = t.weight_mode != OCMP_WEIGHT_NONE;
  int fo = 0, nw = 1, fkind = 0;
  if (wind_field || weighted) {
    const ocmp_field &f = fields[wind_field ? t.wind_field : t.weight_field];
    const Block &fb = f.space->s.blocks[f.block];
    fo = fb.order;
    nw = fb.nloc;
    fkind = fb.kind;
  }
  const int ns = d * (d + 1) / 2;
  int nb = 0;
  switch (t.kind) {
    case OCMP_BFI_GRADGRAD: nb = ns; break;
    case OCMP_BFI_MASS: nb = 1; break;
    default: nb = d; break;
  }
  nf = nw * nb;
  const int ni = bt.nloc, nj = bj.nloc;
  std::vector<double> T((size_t)nf * ni * nj, 0.0);
  Rule rule = simplex_rule(d, bt.order + bj.order + fo);
  std::vector<double> pt(ni), gt((size_t)ni * d), pj(nj), gj((size_t)nj * d), pw(nw), gw((size_t)nw * d);
  const size_t tsz = (size_t)ni * nj;
  for (int q = 0; q < rule.nq; ++q) {
    const double *lam = &rule.lam[q * (d + 1)];
    eval_basis(bt.kind, d, bt.order, lam, pt.data(), gt.data());
    eval_basis(bj.kind, d, bj.order, lam, pj.data(), gj.data());
    if (wind_field || weighted)
      eval_basis(fkind, d, fo, lam, pw.data(), gw.data());
    else
      pw[0] = 1.0;
    const double w = rule.w[q];
    for (int m = 0; m < nw; ++m) {
      const double wm = w * pw[m];
      for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j) {
          double *o = &T[(size_t)m * nb * tsz + (size_t)i * nj + j];
          if (t.kind == OCMP_BFI_GRADGRAD) {
            int k = 0;
            for (int a = 0; a < d; ++a)
              for (int bb = a; bb < d; ++bb, ++k) {
                double v = gt[i * d + a] * gj[j * d + bb];
                if (bb != a) v += gt[i * d + bb] * gj[j * d + a];
                o[k * tsz] += wm * v;
              }
          } else if (t.kind == OCMP_BFI_MASS) {
            o[0] += wm * pt[i] * pj[j];
          } else if (t.kind == OCMP_BFI_DTRIAL) {
            for (int a = 0; a < d; ++a) o[a * tsz] += wm * pt[i] * gj[j * d + a];
          } else {  // CONV, DTEST: the test function carries the derivative
            for (int a = 0; a < d; ++a) o[a * tsz] += wm * gt[i * d + a] * pj[j];
          }
        }
    }
  }
  return T;
}

// can the row gather assemble these volume terms (field weights need H1 fields; a convection
// term takes either a wind field or a weight, not both)
static bool gather_supported(const Matrix *M, const std::vector<ocmp_term> &terms,
                             const std::vector<ocmp_field> &fields) {
  if (!M->d_pptr.p || M->gblk.empty() || M->space->nloc > 64) return false;
  for (auto &t : terms) {
    if (is_bnd_kind(t.kind)) continue;
    if (t.kind == OCMP_BFI_DG_FACET) return false;
    if (t.kind == OCMP_BFI_CONV && t.wind_field >= 0 && t.weight_mode != OCMP_WEIGHT_NONE) return false;
    if (t.weight_mode != OCMP_WEIGHT_NONE &&
        fields[t.weight_field].space->s.blocks[fields[t.weight_field].block].kind != OCMP_H1)
      return false;
  }
  return true;
}

// returns false (nothing launched) when some block pair needs more than 64 factors
static bool launch_h1_gather(Matrix *M, const std::vector<ocmp_term> &vol, const std::vector<ocmp_field> &fields,
                             bool accumulate) {
  Space &s = *M->space;
  Ctx *ctx = M->ctx;
  hipStream_t st = ctx->stream;
  const Mesh &m = *s.mesh;
  const int d = m.dim;
  OCMP_REQUIRE((int)vol.size() <= MAXTG && (int)fields.size() <= MAXF, "too many terms or fields in one call");
  KFactors KF;
  KGather KG;
  std::memset(&KF, 0, sizeof KF);
  std::memset(&KG, 0, sizeof KG);
  for (size_t i = 0; i < fields.size(); ++i) {
    const Space &fs = fields[i].space->s;
    const Block &fb = fs.blocks[fields[i].block];
    KField &kf = KF.f[i];
    kf.vec = fields[i].vec;
    kf.edofs = fs.d_edofs.p;
    kf.ld = fs.nloc;
    kf.loc_off = fb.loc_off;
    kf.nl = fb.nloc;
  }
  // factor sets (terms with equal factors share one) and each term's tables
  std::vector<std::string> set_keys;
  std::vector<int> term_set(vol.size());
  std::vector<std::vector<double>> term_tab(vol.size());
  for (size_t ti = 0; ti < vol.size(); ++ti) {
    const ocmp_term &t = vol[ti];
    int nf = 0;
    term_tab[ti] = gather_tables(s, t, fields, nf);
    KFactorSet S;
    std::memset(&S, 0, sizeof S);
    S.coef = t.coef;
    S.elem_coef = t.elem_coef;
    S.weight_mode = t.weight_mode;
    S.wfield = t.weight_mode != OCMP_WEIGHT_NONE ? t.weight_field : -1;
    if (t.kind == OCMP_BFI_GRADGRAD) {
      S.kind = FS_GG;
    } else if (t.kind == OCMP_BFI_MASS) {
      S.kind = FS_MASS;
    } else if (t.kind == OCMP_BFI_CONV && t.wind_field >= 0) {
      S.kind = FS_WIND;
      S.wfield = t.wind_field;
      S.weight_mode = OCMP_WEIGHT_NONE;
    } else {
      S.kind = FS_VEC;
      if (t.kind == OCMP_BFI_CONV)
        for (int c = 0; c < 3; ++c) S.vec[c] = t.wind[c];
      else
        S.vec[t.dir] = 1.0;
    }
    S.nf = nf;
    char key[256];
    std::snprintf(key, sizeof key, "%d|%d|%d|%a|%p|%a|%a|%a|%d", S.kind, S.weight_mode, S.wfield, S.coef,
                  (const void *)S.elem_coef, S.vec[0], S.vec[1], S.vec[2], nf);
    auto it = std::find(set_keys.begin(), set_keys.end(), std::string(key));
    if (it == set_keys.end()) {
      OCMP_REQUIRE(KF.nsets < MAXFS, "too many factor sets");
      KF.fs[KF.nsets++] = S;
      set_keys.emplace_back(key);
      term_set[ti] = KF.nsets - 1;
    } else {
      term_set[ti] = (int)(it - set_keys.begin());
    }
  }
  // per (test block, trial block): the distinct sets its terms use (tables of terms sharing a set
  // add up); pairs with the same set list share one contiguous range of the factor row
  const int nb = (int)s.blocks.size();
  std::vector<std::vector<int>> plist(nb * nb);
  std::vector<std::vector<double>> ptab(nb * nb);
  for (int bi = 0; bi < nb; ++bi)
    for (int bj = 0; bj < nb; ++bj) {
      const int pi = bi * nb + bj;
      const size_t tsz = (size_t)s.blocks[bi].nloc * s.blocks[bj].nloc;
      std::vector<std::vector<double>> chunk;
      for (size_t ti = 0; ti < vol.size(); ++ti) {
        if (vol[ti].test != bi || vol[ti].trial != bj) continue;
        auto f = std::find(plist[pi].begin(), plist[pi].end(), term_set[ti]);
        if (f == plist[pi].end()) {
          plist[pi].push_back(term_set[ti]);
          chunk.push_back(term_tab[ti]);
        } else {
          auto &c = chunk[f - plist[pi].begin()];
          for (size_t k = 0; k < c.size(); ++k) c[k] += term_tab[ti][k];
        }
      }
      for (auto &c : chunk) ptab[pi].insert(ptab[pi].end(), c.begin(), c.end());
      (void)tsz;
    }
  std::vector<std::vector<int>> lists;
  std::vector<int> list_base;
  int NF = 0, maxpf = 0;
  TableBuilder tb;
  for (int pi = 0; pi < nb * nb; ++pi) {
    KPairG &P = KG.pg[pi];
    P.nli = s.blocks[pi / nb].nloc;
    P.nlj = s.blocks[pi % nb].nloc;
    if (plist[pi].empty()) continue;
    int nfp = 0;
    for (int si : plist[pi]) nfp += KF.fs[si].nf;
    maxpf = std::max(maxpf, nfp);
    auto it = std::find(lists.begin(), lists.end(), plist[pi]);
    int base;
    if (it == lists.end()) {
      base = NF;
      int o = NF;
      for (int si : plist[pi]) {
        KFactorSet &S = KF.fs[si];
        OCMP_REQUIRE(S.ndst < 4, "a factor set used by more than four factor lists");
        S.dst[S.ndst++] = o;
        o += S.nf;
      }
      NF = o;
      lists.push_back(plist[pi]);
      list_base.push_back(base);
    } else {
      base = list_base[it - lists.begin()];
    }
    P.nfp = nfp;
    P.fbase = base;
    P.toff = tb.add(ptab[pi]);
  }
  if (maxpf > 64) return false;
  KF.NF = NF;
  // TD[L][f][j]: for each compound local row L and column j, the (block(L), block(j)) pair's
  // tables at the pair's factor positions, zero elsewhere and in the padding
  const int NL = s.nloc, NFP = (NF + 3) / 4 * 4, NLP = (NL + 15) / 16 * 16;
  {
    std::vector<double> td((size_t)NL * NFP * NLP, 0.0);
    auto blk = [&](int l) {
      int bb = 0;
      while (bb + 1 < nb && l >= s.blocks[bb + 1].loc_off) ++bb;
      return bb;
    };
    for (int L = 0; L < NL; ++L)
      for (int j = 0; j < NL; ++j) {
        const int bi = blk(L), bj = blk(j), pi = bi * nb + bj;
        const KPairG &P = KG.pg[pi];
        if (plist[pi].empty()) continue;
        const int li = L - s.blocks[bi].loc_off, lj = j - s.blocks[bj].loc_off;
        for (int k = 0; k < P.nfp; ++k)
          td[((size_t)L * NFP + P.fbase + k) * NLP + j] = ptab[pi][(size_t)k * P.nli * P.nlj + (size_t)li * P.nlj + lj];
      }
    KG.td = tb.add(td);
  }
  DBuf<double> *tab = cached_table(s, "G" + term_key(vol, fields), tb, st);
  KG.NL = s.nloc;
  KG.NF = NF;
  KG.NFP = NFP;
  KG.NLP = NLP;
  KG.nb = nb;
  KG.emax = M->gemax;
  KG.accmax = M->gaccmax;
  KG.accumulate = accumulate ? 1 : 0;
  for (int bb = 0; bb < nb; ++bb) KG.loc_off[bb] = s.blocks[bb].loc_off;
  KG.loc_off[nb] = s.nloc;
  KG.tab = tab->p;
  // descriptors go to device memory (the kernels index them at run time)
  std::vector<char> desc(sizeof KF + sizeof KG);
  std::memcpy(desc.data(), &KF, sizeof KF);
  std::memcpy(desc.data() + sizeof KF, &KG, sizeof KG);
  M->d_gdesc.alloc(desc.size());
  OCMP_HIP(hipMemcpyAsync(M->d_gdesc.p, desc.data(), desc.size(), hipMemcpyHostToDevice, st));
  const KFactors *dKF = reinterpret_cast<const KFactors *>(M->d_gdesc.p);
  const KGather *dKG = reinterpret_cast<const KGather *>(M->d_gdesc.p + sizeof KF);
  M->d_factors.alloc((size_t)m.nel * std::max(NF, 1));
  {
    KTimer tm(ctx, ctx->cls("assemble_h1_factors"), (double)m.nel * (8.0 * NF + 4.0 * (d + 1)) + 8.0 * m.nv * d);
    if (d == 2)
      k_h1_factors<2><<<nblk(m.nel), 256, 0, st>>>(dKF, m.nel, m.d_coords.p, m.d_elems.p, M->d_factors.p);
    else
      k_h1_factors<3><<<nblk(m.nel), 256, 0, st>>>(dKF, m.nel, m.d_coords.p, m.d_elems.p, M->d_factors.p);
  }
  {
    const int sb = M->d_pslot8.p ? 1 : 2;
    // algorithmic: every CSR value written once (read too when accumulating), the entries' elements,
    // sorted positions and column positions, every element's factor row once, the row pointers
    KTimer tm(ctx, ctx->cls("assemble_h1_gather"),
              8.0 * M->nnz * (accumulate ? 2 : 1) + (double)M->npatch * (6.0 + s.nloc * sb) +
                  8.0 * m.nel * NF + 8.0 * (M->n + 1));
    const int64_t nb_rows = (int64_t)M->gblk.size() - 1;
    const unsigned grid = (unsigned)((nb_rows + 7) / 8 * 8);
    const size_t shm = (size_t)M->gemax * s.nloc * (sizeof(double) + sb) + (size_t)M->gaccmax * sizeof(double) +
                       (size_t)M->gemax * sizeof(uint16_t);
    if (sb == 1)
      k_h1_gather<uint8_t><<<grid, 256, shm, st>>>(dKG, nb_rows, M->d_gblk.p, M->d_row_ptr.p, M->d_pptr.p,
                                                   M->d_pe.p, M->d_gloff.p, M->d_gpos.p, M->d_pslot8.p,
                                                   M->d_factors.p, M->d_val.p);
    else
      k_h1_gather<uint16_t><<<grid, 256, shm, st>>>(dKG, nb_rows, M->d_gblk.p, M->d_row_ptr.p, M->d_pptr.p,
                                                    M->d_pe.p, M->d_gloff.p, M->d_gpos.p, M->d_pslot16.p,
                                                    M->d_factors.p, M->d_val.p);
  }
  // the descriptor copy reads host memory that goes out of scope on return
  OCMP_HIP(hipStreamSynchronize(st));
  OCMP_HIP(hipGetLastError());
  return true;
}

// The elements with a facet on the markers of `mask`, greedily coloured among themselves (built
// once per marker set): a boundary-only assembly scatters over these few elements in a few colours
// instead of over every boundary element in the mesh's colours.
static Space::BndSet &boundary_set(Space &s, uint64_t mask) {
  auto it = s.bnd_sets.find(mask);
  if (it != s.bnd_sets.end()) return *it->second;
  const Mesh &m = *s.mesh;
  const int nf = m.dim + 1;
  std::vector<int32_t> el;
  for (int64_t e = 0; e < m.nel; ++e) {
    bool on = false;
    for (int k = 0; k < nf; ++k) {
      const int32_t nb = m.elem_nb[(size_t)e * nf + k];
      if (nb < 0 && -(nb + 1) < 64 && (mask >> (-(nb + 1)) & 1)) on = true;
    }
    if (on) el.push_back((int32_t)e);
  }
  std::map<int32_t, std::vector<int>> vcol;  // vertex -> colours of the set's elements on it
  std::vector<int> col(el.size());
  int ncol = 0;
  for (size_t i = 0; i < el.size(); ++i) {
    uint64_t used = 0;
    for (int k = 0; k < nf; ++k)
      for (int c : vcol[m.elems[(size_t)el[i] * nf + k]]) used |= 1ull << c;
    int c = 0;
    while (c < 63 && (used >> c & 1)) ++c;
    col[i] = c;
    ncol = std::max(ncol, c + 1);
    for (int k = 0; k < nf; ++k) vcol[m.elems[(size_t)el[i] * nf + k]].push_back(c);
  }
  auto B = std::make_unique<Space::BndSet>();
  B->off.assign(ncol + 1, 0);
  for (int c : col) B->off[c + 1]++;
  for (int c = 0; c < ncol; ++c) B->off[c + 1] += B->off[c];
  std::vector<int32_t> order(el.size());
  std::vector<int64_t> fill(B->off.begin(), B->off.end() - 1);
  for (size_t i = 0; i < el.size(); ++i) order[fill[col[i]]++] = el[i];
  B->elems.upload(order.empty() ? std::vector<int32_t>(1, 0) : order, m.ctx->stream);
  OCMP_HIP(hipStreamSynchronize(m.ctx->stream));
  auto &ref = *B;
  s.bnd_sets.emplace(mask, std::move(B));
  return ref;
}

// the colour-scheduled element scatter: boundary-facet terms, and volume terms the gather does
// not cover
static void launch_h1_colour(Matrix *M, const std::vector<ocmp_term> &terms, const std::vector<ocmp_field> &fields,
                             bool accumulate, bool volume_terms) {
  Space &s = *M->space;
  Ctx *ctx = M->ctx;
  hipStream_t st = ctx->stream;
  if (!M->d_slot.p) build_h1_schedule(M);
  KArgs A;
  TableBuilder tb;
  build_kargs(s, terms, fields, A, tb, st);
  DBuf<double> *tab = cached_table(s, term_key(terms, fields), tb, st);
  A.tab = tab->p;
  // algorithmic bytes for the volume scatter (every value, the elements, the coordinates); the
  // boundary-only scatter touches a few rows and is timed without a byte count
  KTimer tm(ctx, ctx->cls(volume_terms ? "assemble_h1" : "assemble_h1_bnd"),
            volume_terms ? 8.0 * M->nnz + 4.0 * s.mesh->nel * (s.mesh->dim + 1) + 8.0 * s.mesh->nv * s.mesh->dim
                         : 0.0);
  if (!accumulate) k_fill<<<nblk(M->nnz), 256, 0, st>>>(M->d_val.p, M->nnz, 0.0);
  const Mesh &m = *s.mesh;
  bool has_bnd = false;
  uint64_t bmask = 0;
  for (auto &t : terms)
    if (is_bnd_kind(t.kind)) {
      has_bnd = true;
      bmask |= t.dirichlet_mask;
    }
  OCMP_REQUIRE(!has_bnd || (int64_t)s.bnd_colour_off.size() == s.ncolours + 1, "boundary element lists missing");
  if (has_bnd && !volume_terms) {
    Space::BndSet &B = boundary_set(s, bmask);
    for (size_t c = 0; c + 1 < B.off.size(); ++c) {
      const int64_t bb = B.off[c], bn = B.off[c + 1] - bb;
      if (!bn) continue;
      if (m.dim == 2)
        k_h1_bnd_assemble<2><<<nblk(bn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, B.elems.p, bb, bn,
                                                                 M->d_slot.p, M->d_val.p);
      else
        k_h1_bnd_assemble<3><<<nblk(bn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, B.elems.p, bb, bn,
                                                                 M->d_slot.p, M->d_val.p);
    }
    OCMP_HIP(hipGetLastError());
    return;
  }
  for (int c = 0; c < s.ncolours; ++c) {
    int64_t cb = s.colour_off[c], cn = s.colour_off[c + 1] - cb;
    if (!cn) continue;
    if (volume_terms) {
      if (m.dim == 2)
        k_h1_assemble<2><<<nblk(cn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, s.d_colour_elems.p, cb,
                                                             cn, M->d_slot.p, M->d_val.p);
      else
        k_h1_assemble<3><<<nblk(cn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, s.d_colour_elems.p, cb,
                                                             cn, M->d_slot.p, M->d_val.p);
    }
    if (has_bnd) {
      const int64_t bb = s.bnd_colour_off[c], bn = s.bnd_colour_off[c + 1] - bb;
      if (!bn) continue;
      if (m.dim == 2)
        k_h1_bnd_assemble<2><<<nblk(bn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, s.d_bnd_colour_elems.p,
                                                                 bb, bn, M->d_slot.p, M->d_val.p);
      else
        k_h1_bnd_assemble<3><<<nblk(bn * s.nloc), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, s.d_bnd_colour_elems.p,
                                                                 bb, bn, M->d_slot.p, M->d_val.p);
    }
  }
  OCMP_HIP(hipGetLastError());
}

void launch_assemble_h1(Matrix *M, const std::vector<ocmp_term> &terms, const std::vector<ocmp_field> &fields,
                        bool accumulate) {
  for (auto &t : terms) OCMP_REQUIRE(t.kind != OCMP_BFI_DG_FACET, "facet terms need an L2 (DG) space");
  std::vector<ocmp_term> vol, bnd;
  for (auto &t : terms) (is_bnd_kind(t.kind) ? bnd : vol).push_back(t);
  // the element scatter takes MAXT terms per launch: later launches add to the first
  auto colour = [&](const std::vector<ocmp_term> &ts, bool acc) {
    if (ts.empty()) {
      launch_h1_colour(M, ts, fields, acc, false);
      return;
    }
    for (size_t i = 0; i < ts.size(); i += MAXT) {
      std::vector<ocmp_term> c(ts.begin() + i, ts.begin() + std::min(ts.size(), i + MAXT));
      bool has_vol = false;
      for (auto &t : c) has_vol |= !is_bnd_kind(t.kind);
      launch_h1_colour(M, c, fields, acc || i > 0, has_vol);
    }
  };
  const bool gather = !vol.empty() && (int)vol.size() <= MAXTG && gather_supported(M, vol, fields);
  if (!gather || !launch_h1_gather(M, vol, fields, accumulate)) {
    colour(terms, accumulate);
    return;
  }
  if (!bnd.empty()) colour(bnd, true);
}

void launch_assemble_dg(Matrix *M, const std::vector<ocmp_term> &terms, bool accumulate) {
  Space &s = *M->space;
  Ctx *ctx = M->ctx;
  hipStream_t st = ctx->stream;
  const Mesh &m = *s.mesh;
  OCMP_REQUIRE(m.dim == 2, "DG assembly is implemented on triangles");
  const Block &b = s.blocks[0];
  OCMP_REQUIRE(b.nloc <= MAXNB, "DG block too large");
  std::vector<ocmp_term> vol;
  const ocmp_term *fac = nullptr;
  for (auto &t : terms) {
    OCMP_REQUIRE(t.weight_mode == OCMP_WEIGHT_NONE && t.wind_field < 0,
                 "DG terms take constant coefficients (per-element factors allowed)");
    if (t.kind == OCMP_BFI_DG_FACET) {
      OCMP_REQUIRE(!fac, "one facet term per assembly call");
      OCMP_REQUIRE(t.elem_coef == nullptr, "facet term takes scalar coefficients");
      fac = &t;
    } else {
      vol.push_back(t);
    }
  }
  KArgs A;
  TableBuilder tb;
  std::vector<ocmp_field> nofields;
  build_kargs(s, vol, nofields, A, tb, st);
  KDG F;
  std::memset(&F, 0, sizeof F);
  F.nb = b.nloc;
  std::vector<double> ft, fw;
  line_rule(2 * b.order + 2, ft, fw);
  F.nq = (int)ft.size();
  {
    std::vector<double> phi((size_t)3 * F.nq * b.nloc), gr((size_t)3 * F.nq * b.nloc * 2);
    const int fv[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < F.nq; ++q) {
        double lam[3] = {0, 0, 0};
        lam[fv[k][0]] = 1.0 - ft[q];
        lam[fv[k][1]] = ft[q];
        eval_basis(b.kind, 2, b.order, lam, &phi[((size_t)k * F.nq + q) * b.nloc],
                   &gr[((size_t)k * F.nq + q) * b.nloc * 2]);
      }
    F.tab_fphi = tb.add(phi);
    F.tab_fgrad = tb.add(gr);
    F.tab_fw = tb.add(fw);
  }
  if (fac) {
    F.has_facet = 1;
    F.eps = fac->eps * fac->coef;
    F.alpha = fac->alpha;
    F.wind[0] = fac->wind[0];
    F.wind[1] = fac->wind[1];
    F.dmask = fac->dirichlet_mask;
    F.sipg = fac->sipg;
    F.upwind = fac->upwind;
    if (fac->coef != 1.0 && fac->upwind) {
      // scale the upwind flux with coef as well
      F.wind[0] *= fac->coef;
      F.wind[1] *= fac->coef;
    }
  }
  std::string key = "DG" + term_key(terms, nofields);
  DBuf<double> *tab = cached_table(s, key, tb, st);
  A.tab = tab->p;
  KTimer tm(ctx, ctx->cls("assemble_dg"),
            8.0 * M->nnz + 4.0 * m.nel * 3 * 3 + 4.0 * m.nel * 4 + 8.0 * m.nv * 2 + 4.0 * (M->n + 1));
  k_dg_assemble<<<nblk(m.nel * b.nloc), 256, 0, st>>>(A, F, m.nel, m.d_coords.p, m.d_elems.p, m.d_elem_nb.p,
                                                      m.d_elem_nbk.p, M->d_dg_pos.p, M->d_row_ptr.p, M->d_val.p,
                                                      accumulate ? 1 : 0);
  OCMP_HIP(hipGetLastError());
}

void launch_assemble_linear(Space *sp, double *bvec, const std::vector<ocmp_lterm> &terms,
                            const std::vector<ocmp_field> &fields, bool accumulate) {
  Space &s = *sp;
  Ctx *ctx = s.mesh->ctx;
  hipStream_t st = ctx->stream;
  const Mesh &m = *s.mesh;
  const int d = m.dim;
  OCMP_REQUIRE((int)terms.size() <= MAXT, "too many terms (max 8)");
  OCMP_REQUIRE((int)fields.size() <= MAXF, "too many fields (max 8)");
  OCMP_REQUIRE(s.nloc <= MAXL, "compound element has too many dofs");
  bool all_l2 = true;
  for (auto &b : s.blocks) all_l2 &= b.kind == OCMP_L2;
  KLArgs A;
  std::memset(&A, 0, sizeof A);
  TableBuilder tb;
  A.nterms = (int)terms.size();
  A.nfields = (int)fields.size();
  A.nloc = s.nloc;
  // quadrature degree: exact for polynomial data (constant: p, quadratic: p+2, grid function of
  // order q: p+q, times a weight field), 2p+4 for the transcendental sin-product source
  int qdeg = 0;
  {
    const int p = max_order(s);
    for (auto &t : terms) {
      int dg = 2 * p + 4;
      if (t.kind == OCMP_LFI_SOURCE) {
        if (t.fn == OCMP_FN_CONST) dg = p;
        else if (t.fn == OCMP_FN_POLY2) dg = p + 2;
        else if (t.fn == OCMP_FN_FIELD) dg = p + fields[t.field].space->s.blocks[fields[t.field].block].order;
        if (t.weight_mode != OCMP_WEIGHT_NONE)
          dg += fields[t.weight_field].space->s.blocks[fields[t.weight_field].block].order;
      }
      qdeg = std::max(qdeg, dg);
    }
  }
  Rule rule = simplex_rule(d, qdeg);
  A.nq = rule.nq;
  fill_block_tables(tb, s, rule, A.b);
  A.tab_w = tb.add(rule.w);
  A.tab_lam = tb.add(rule.lam);
  for (size_t i = 0; i < fields.size(); ++i) {
    const Space &fs = fields[i].space->s;
    const Block &fb = fs.blocks[fields[i].block];
    KField &kf = A.f[i];
    kf.vec = fields[i].vec;
    kf.edofs = fs.d_edofs.p;
    kf.ld = fs.nloc;
    kf.loc_off = fb.loc_off;
    kf.nl = fb.nloc;
    std::vector<double> phi((size_t)rule.nq * fb.nloc), gr((size_t)rule.nq * fb.nloc * d);
    for (int q = 0; q < rule.nq; ++q)
      eval_basis(fb.kind, d, fb.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * fb.nloc],
                 &gr[(size_t)q * fb.nloc * d]);
    kf.tab_phi = tb.add(phi);
  }
  bool has_dg = false, has_bnd = false;
  for (size_t i = 0; i < terms.size(); ++i) {
    const ocmp_lterm &t = terms[i];
    KLTerm &k = A.t[i];
    k.kind = t.kind;
    k.test = t.test;
    k.fn = t.fn;
    k.field = t.field;
    for (int c = 0; c < 10; ++c) k.c[c] = t.c[c];
    k.coef = t.coef;
    k.weight_mode = t.weight_mode;
    k.weight_field = t.weight_field;
    k.eps = t.eps;
    k.alpha = t.alpha;
    for (int c = 0; c < 3; ++c) k.wind[c] = t.wind[c];
    k.dmask = t.dirichlet_mask;
    check_fn(s.mesh->ctx, t.fn, t.c);
    if (t.kind == OCMP_LFI_DG_DIRICHLET) {
      OCMP_REQUIRE(d == 2, "DG Dirichlet data on triangles only");
      OCMP_REQUIRE(t.fn != OCMP_FN_FIELD, "DG Dirichlet data must be a closed-form function");
      has_dg = true;
    }
    if (t.kind == OCMP_LFI_BND_SOURCE) {
      OCMP_REQUIRE(t.fn != OCMP_FN_FIELD, "boundary source data must be a closed-form function");
      has_bnd = true;
    }
  }
  if (has_bnd) {
    std::vector<double> ft, fw, flam;
    if (d == 2) {
      line_rule(2 * max_order(s) + 4, ft, fw);
      A.bnq = (int)ft.size();
      A.tab_bft = tb.add(ft);
    } else {
      Rule fr = simplex_rule(2, 2 * max_order(s) + 4);
      A.bnq = fr.nq;
      flam = fr.lam;
      fw = fr.w;
      for (auto &w : fw) w *= 2.0;  // weights summing to 1: a face integral is area * sum
      A.tab_bft = tb.add(flam);
    }
    A.tab_bfw = tb.add(fw);
    for (int bb = 0; bb < A.b.n; ++bb)
      A.b.tab_fphi[bb] = add_facet_phi(tb, d, s.blocks[bb].kind, s.blocks[bb].order, s.blocks[bb].nloc, A.bnq, ft,
                                       flam);
  }
  if (has_dg) {
    const Block &b = s.blocks[0];
    std::vector<double> ft, fw;
    line_rule(2 * b.order + 4, ft, fw);
    A.fnq = (int)ft.size();
    std::vector<double> phi((size_t)3 * A.fnq * b.nloc), gr((size_t)3 * A.fnq * b.nloc * 2);
    const int fv[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    for (int k = 0; k < 3; ++k)
      for (int q = 0; q < A.fnq; ++q) {
        double lam[3] = {0, 0, 0};
        lam[fv[k][0]] = 1.0 - ft[q];
        lam[fv[k][1]] = ft[q];
        eval_basis(b.kind, 2, b.order, lam, &phi[((size_t)k * A.fnq + q) * b.nloc],
                   &gr[((size_t)k * A.fnq + q) * b.nloc * 2]);
      }
    A.tab_fphi = tb.add(phi);
    A.tab_fgrad = tb.add(gr);
    A.tab_fw = tb.add(fw);
    A.tab_ft = tb.add(ft);
  }
  std::string key = "L" + std::to_string(qdeg);
  for (auto &t : terms) key += "|" + std::to_string(t.kind) + "," + std::to_string(t.test) + "," +
                               std::to_string(t.fn) + "," + std::to_string(t.fn == OCMP_FN_FIELD ? t.field : -1);
  for (auto &f : fields) key += "|f" + std::to_string((uintptr_t)f.space) + "," + std::to_string(f.block);
  DBuf<double> *tab = cached_table(s, key, tb, st);
  A.tab = tab->p;
  KTimer tm(ctx, ctx->cls("assemble_linear"));
  if (all_l2) {
    // owner-computes: every dof belongs to exactly one element
    if (d == 2)
      k_linear<2><<<nblk(m.nel), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, m.d_elem_nb.p, nullptr, 0, m.nel,
                                               s.d_edofs.p, bvec, accumulate ? 1 : 0);
    else
      k_linear<3><<<nblk(m.nel), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, m.d_elem_nb.p, nullptr, 0, m.nel,
                                               s.d_edofs.p, bvec, accumulate ? 1 : 0);
  } else {
    OCMP_REQUIRE(s.ncolours > 0, "space has no colouring (mixed H1/L2 compound)");
    if (!accumulate) k_fill<<<nblk(s.ndof), 256, 0, st>>>(bvec, s.ndof, 0.0);
    for (int c = 0; c < s.ncolours; ++c) {
      int64_t cb = s.colour_off[c], cn = s.colour_off[c + 1] - cb;
      if (!cn) continue;
      if (d == 2)
        k_linear<2><<<nblk(cn), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, m.d_elem_nb.p, s.d_colour_elems.p, cb,
                                              cn, s.d_edofs.p, bvec, 1);
      else
        k_linear<3><<<nblk(cn), 256, 0, st>>>(A, m.d_coords.p, m.d_elems.p, m.d_elem_nb.p, s.d_colour_elems.p, cb,
                                              cn, s.d_edofs.p, bvec, 1);
    }
  }
  OCMP_HIP(hipGetLastError());
}

// edge id of the vertex pair (a, b) (edges are sorted pairs)
static int64_t find_edge(const Mesh &m, int32_t a, int32_t b) {
  if (a > b) std::swap(a, b);
  int64_t lo = 0, hi = m.nedge;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    if (m.edges[2 * mid] < a || (m.edges[2 * mid] == a && m.edges[2 * mid + 1] < b))
      lo = mid + 1;
    else
      hi = mid;
  }
  OCMP_REQUIRE(lo < m.nedge && m.edges[2 * lo] == a && m.edges[2 * lo + 1] == b, "edge not in the mesh");
  return lo;
}

// edge ids of (a,b), (a,c), (b,c) of the triangles tri[3 t ..] (ascending vertex ids), for the
// listed triangles (ids) or all ntri of them (ids empty)
static std::vector<int32_t> triangle_edges(const Mesh &m, const int32_t *tri, int64_t ntri,
                                           const std::vector<int32_t> &ids) {
  const int64_t n = ids.empty() ? ntri : (int64_t)ids.size();
  std::vector<int32_t> out((size_t)n * 3);
  for (int64_t t = 0; t < n; ++t) {
    const int32_t *v = tri + 3 * (int64_t)(ids.empty() ? t : ids[t]);
    OCMP_REQUIRE(v[0] < v[1] && v[1] < v[2], "triangle vertices must be ascending");
    out[3 * t] = (int32_t)find_edge(m, v[0], v[1]);
    out[3 * t + 1] = (int32_t)find_edge(m, v[0], v[2]);
    out[3 * t + 2] = (int32_t)find_edge(m, v[1], v[2]);
  }
  return out;
}

static void ensure_tri_tables(Space &s) {
  if (s.d_tri_verts.p) return;
  const Mesh &m = *s.mesh;
  hipStream_t st = m.ctx->stream;
  const std::vector<int32_t> &tri = m.dim == 3 ? m.faces : m.elems;
  const int64_t ntri = m.dim == 3 ? m.nface : m.nel;
  s.d_tri_verts.upload(tri, st);
  s.d_tri_edges.upload(triangle_edges(m, tri.data(), ntri, {}), st);
  s.ntri = ntri;
  OCMP_HIP(hipStreamSynchronize(st));
}

void launch_dirichlet_set(Space *sp, int block, double *x, int fn, const double *c, uint64_t mask) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const Block &b = s.blocks[block];
  check_fn(ctx, fn, c);
  auto key = std::make_pair(block, mask);
  Space::Lift *L;
  auto it = s.lifts.find(key);
  if (it == s.lifts.end()) {
    L = new Space::Lift();
    std::vector<uint8_t> vflag(m.nv, 0), eflag(m.nedge, 0);
    std::vector<int32_t> fl;
    const int d = m.dim;
    for (int64_t f = 0; f < m.nfacet; ++f) {
      int mk = m.f_marker[f];
      if (m.f_e1[f] >= 0 || mk < 0 || !((mask >> mk) & 1)) continue;
      const int32_t *fv = &m.facet_verts[f * d];
      for (int a = 0; a < d; ++a) vflag[fv[a]] = 1;
      if (d == 2)
        eflag[f] = 1;
      else {
        eflag[find_edge(m, fv[0], fv[1])] = 1;
        eflag[find_edge(m, fv[0], fv[2])] = 1;
        eflag[find_edge(m, fv[1], fv[2])] = 1;
        fl.push_back((int32_t)f);
      }
    }
    std::vector<int32_t> vl, el;
    for (int64_t v = 0; v < m.nv; ++v)
      if (vflag[v]) vl.push_back((int32_t)v);
    for (int64_t e = 0; e < m.nedge; ++e)
      if (eflag[e]) el.push_back((int32_t)e);
    L->verts.upload(vl, st);
    L->edges.upload(el, st);
    L->faces.upload(fl, st);
    if (d == 3 && !fl.empty()) L->fedges.upload(triangle_edges(m, m.faces.data(), m.nface, fl), st);
    L->nverts = (int64_t)vl.size();
    L->nedges = (int64_t)el.size();
    L->nfaces = (int64_t)fl.size();
    OCMP_HIP(hipStreamSynchronize(st));
    s.lifts[key] = L;
  } else {
    L = it->second;
  }
  if (b.order >= 3 && m.dim == 3 && L->nfaces) ensure_tri_tables(s);
  // function parameters live in the context's device scalar area (slots 200..209)
  double *dc = ctx->scal.p + 200;
  OCMP_HIP(hipMemcpyAsync(dc, c, 10 * sizeof(double), hipMemcpyHostToDevice, st));
  KTimer tm(ctx, ctx->cls("dirichlet_set"));
  if (L->nverts)
    k_dir_verts<<<nblk(L->nverts), 256, 0, st>>>(L->verts.p, L->nverts, m.d_coords.p, m.dim, fn, dc, b.dof_off, x);
  if (b.order >= 2 && L->nedges)
    k_dir_edges<<<nblk(L->nedges), 256, 0, st>>>(L->edges.p, L->nedges, s.d_edges.p, m.d_coords.p, m.dim, fn, dc,
                                                 b.order, b.dof_off, m.nv, x);
  if (b.order >= 3 && m.dim == 3 && L->nfaces)
    k_set_faces<<<nblk(L->nfaces), 256, 0, st>>>(L->faces.p, L->nfaces, s.d_tri_verts.p, L->fedges.p, m.d_coords.p,
                                                 m.dim, fn, dc, b.dof_off, m.nv, b.dof_off + m.nv + 2 * m.nedge, x);
  OCMP_HIP(hipGetLastError());
  // the parameter copy reads host memory that the caller may free on return
  OCMP_HIP(hipStreamSynchronize(st));
}

void launch_set(Space *sp, int block, double *x, int fn, const double *c) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const Block &b = s.blocks[block];
  const int d = m.dim;
  check_fn(ctx, fn, c);
  OCMP_REQUIRE(fn != OCMP_FN_FIELD, "Set takes a closed-form function or an expression");
  KTimer tm(ctx, ctx->cls("set"));
  if (b.kind == OCMP_H1) {
    OCMP_REQUIRE(b.order >= 1 && b.order <= 3, "Set on H1 blocks of order 1-3");
    double *dc = ctx->scal.p + 200;
    OCMP_HIP(hipMemcpyAsync(dc, c, 10 * sizeof(double), hipMemcpyHostToDevice, st));
    k_dir_verts<<<nblk(m.nv), 256, 0, st>>>(nullptr, m.nv, m.d_coords.p, d, fn, dc, b.dof_off, x);
    if (b.order >= 2)
      k_dir_edges<<<nblk(m.nedge), 256, 0, st>>>(nullptr, m.nedge, s.d_edges.p, m.d_coords.p, d, fn, dc, b.order,
                                                 b.dof_off, m.nv, x);
    if (b.order >= 3) {
      ensure_tri_tables(s);
      k_set_faces<<<nblk(s.ntri), 256, 0, st>>>(nullptr, s.ntri, s.d_tri_verts.p, s.d_tri_edges.p, m.d_coords.p, d,
                                                fn, dc, b.dof_off, m.nv, b.dof_off + m.nv + 2 * m.nedge, x);
    }
    OCMP_HIP(hipGetLastError());
    OCMP_HIP(hipStreamSynchronize(st));
    return;
  }
  OCMP_REQUIRE(b.kind == OCMP_L2 && d == 2 && b.nloc <= MAXNB, "Set on L2 blocks: triangles up to P3");
  KL2Set K;
  std::memset(&K, 0, sizeof K);
  TableBuilder tb;
  const int qdeg = fn == OCMP_FN_CONST ? 2 * b.order : (fn == OCMP_FN_POLY2 ? b.order + 2 : 2 * b.order + 6);
  Rule rule = simplex_rule(d, qdeg);
  K.nq = rule.nq;
  K.nl = b.nloc;
  K.ld = s.nloc;
  K.loc_off = b.loc_off;
  std::vector<double> phi((size_t)rule.nq * b.nloc), gr((size_t)rule.nq * b.nloc * d);
  for (int q = 0; q < rule.nq; ++q)
    eval_basis(b.kind, d, b.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b.nloc], &gr[(size_t)q * b.nloc * d]);
  // reference mass matrix on an exact rule, inverted by Gauss-Jordan (nl <= 10)
  Rule mr = simplex_rule(d, 2 * b.order);
  const int nl = b.nloc;
  std::vector<double> M((size_t)nl * nl, 0.0), Mi((size_t)nl * nl, 0.0), pm((size_t)nl), gm((size_t)nl * d);
  for (int q = 0; q < mr.nq; ++q) {
    eval_basis(b.kind, d, b.order, &mr.lam[q * (d + 1)], pm.data(), gm.data());
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j) M[i * nl + j] += mr.w[q] * pm[i] * pm[j];
  }
  for (int i = 0; i < nl; ++i) Mi[i * nl + i] = 1.0;
  for (int col = 0; col < nl; ++col) {
    int piv = col;
    for (int r = col + 1; r < nl; ++r)
      if (std::fabs(M[r * nl + col]) > std::fabs(M[piv * nl + col])) piv = r;
    for (int j = 0; j < nl; ++j) {
      std::swap(M[col * nl + j], M[piv * nl + j]);
      std::swap(Mi[col * nl + j], Mi[piv * nl + j]);
    }
    const double dv = M[col * nl + col];
    OCMP_REQUIRE(std::fabs(dv) > 0.0, "singular reference mass matrix");
    for (int j = 0; j < nl; ++j) {
      M[col * nl + j] /= dv;
      Mi[col * nl + j] /= dv;
    }
    for (int r = 0; r < nl; ++r) {
      if (r == col) continue;
      const double fct = M[r * nl + col];
      for (int j = 0; j < nl; ++j) {
        M[r * nl + j] -= fct * M[col * nl + j];
        Mi[r * nl + j] -= fct * Mi[col * nl + j];
      }
    }
  }
  K.tab_phi = tb.add(phi);
  K.tab_w = tb.add(rule.w);
  K.tab_lam = tb.add(rule.lam);
  K.tab_minv = tb.add(Mi);
  K.fn = fn;
  for (int i = 0; i < 10; ++i) K.c[i] = c[i];
  DBuf<double> *tab = cached_table(s, "S" + std::to_string(block) + "," + std::to_string(qdeg), tb, st);
  K.tab = tab->p;
  k_l2_set<2><<<nblk(m.nel), 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipStreamSynchronize(st));
}

void error_norms(Space *sp, int block, const double *x, const double *ref, int fn, const double *c, double *out) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const Block &b = s.blocks[block];
  const int d = m.dim;
  if (!ref) check_fn(ctx, fn, c);
  KNorm K;
  std::memset(&K, 0, sizeof K);
  TableBuilder tb;
  Rule rule = simplex_rule(d, 2 * b.order + 6);
  K.nq = rule.nq;
  K.nl = b.nloc;
  K.ld = s.nloc;
  K.loc_off = b.loc_off;
  std::vector<double> phi((size_t)rule.nq * b.nloc), gr((size_t)rule.nq * b.nloc * d);
  for (int q = 0; q < rule.nq; ++q)
    eval_basis(b.kind, d, b.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b.nloc], &gr[(size_t)q * b.nloc * d]);
  K.tab_phi = tb.add(phi);
  K.tab_grad = tb.add(gr);
  K.tab_w = tb.add(rule.w);
  K.tab_lam = tb.add(rule.lam);
  K.fn = fn;
  for (int i = 0; i < 10; ++i) K.c[i] = c[i];
  DBuf<double> *tab = cached_table(s, "N" + std::to_string(block), tb, st);
  K.tab = tab->p;
  const unsigned nb = nblk(m.nel);
  double *part = ctx->red.p;
  OCMP_REQUIRE(ctx->red.n >= 2 * (size_t)nb + 2, "reduction scratch too small for this mesh");
  KTimer tm(ctx, ctx->cls("error_norms"));
  if (d == 2)
    k_error_norms<2><<<nb, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, ref, part);
  else
    k_error_norms<3><<<nb, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, ref, part);
  double *res = ctx->scal.p + 110;
  k_sum_pairs<<<1, 256, 0, st>>>(part, nb, res);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
}

void error_norms_vec(Space *sp, int nb, const int *blocks, const double *x, const double *ref, const int *fns,
                     const double *cs, double *out) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const int d = m.dim;
  OCMP_REQUIRE(nb >= 1 && nb <= 3, "a vector component has 1 to 3 scalar blocks");
  const Block &b = s.blocks[blocks[0]];
  KNormV K;
  std::memset(&K, 0, sizeof K);
  for (int k = 0; k < nb; ++k) {
    OCMP_REQUIRE(blocks[k] >= 0 && blocks[k] < (int)s.blocks.size(), "block out of range");
    const Block &bk = s.blocks[blocks[k]];
    OCMP_REQUIRE(bk.kind == b.kind && bk.order == b.order, "the blocks of a vector norm must share kind and order");
    K.loc_off[k] = bk.loc_off;
    if (!ref) {
      check_fn(ctx, fns[k], cs + 10 * k);
      OCMP_REQUIRE(fns[k] != OCMP_FN_FIELD, "a field reference is passed as ref, not as a function");
    }
    K.fn[k] = ref ? OCMP_FN_CONST : fns[k];
    for (int i = 0; i < 10; ++i) K.c[k][i] = ref ? 0.0 : cs[10 * k + i];
  }
  TableBuilder tb;
  Rule rule = simplex_rule(d, 2 * b.order + 6);
  K.nq = rule.nq;
  K.nl = b.nloc;
  K.ld = s.nloc;
  K.ncomp = nb;
  std::vector<double> phi((size_t)rule.nq * b.nloc), gr((size_t)rule.nq * b.nloc * d);
  for (int q = 0; q < rule.nq; ++q)
    eval_basis(b.kind, d, b.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b.nloc], &gr[(size_t)q * b.nloc * d]);
  K.tab_phi = tb.add(phi);
  K.tab_grad = tb.add(gr);
  K.tab_w = tb.add(rule.w);
  K.tab_lam = tb.add(rule.lam);
  DBuf<double> *tab = cached_table(s, "N" + std::to_string(blocks[0]), tb, st);
  K.tab = tab->p;
  const unsigned nblkk = nblk(m.nel);
  double *part = ctx->red.p;
  OCMP_REQUIRE(ctx->red.n >= 3 * (size_t)nblkk + 3, "reduction scratch too small for this mesh");
  KTimer tm(ctx, ctx->cls("error_norms"));
  if (d == 2)
    k_error_norms_vec<2><<<nblkk, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, ref, part);
  else
    k_error_norms_vec<3><<<nblkk, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, ref, part);
  double *res = ctx->scal.p + 110;
  k_sum_triples<<<1, 256, 0, st>>>(part, nblkk, res);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipMemcpyAsync(out, res, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
}

void l2_pair_norms(Space *sp, int nb, const int *blocks, const double *a, const double *b, double *out) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const int d = m.dim;
  OCMP_REQUIRE(nb >= 1 && nb <= 3, "a vector component has 1 to 3 scalar blocks");
  const Block &b0 = s.blocks[blocks[0]];
  KPair K;
  std::memset(&K, 0, sizeof K);
  for (int k = 0; k < nb; ++k) {
    OCMP_REQUIRE(blocks[k] >= 0 && blocks[k] < (int)s.blocks.size(), "block out of range");
    const Block &bk = s.blocks[blocks[k]];
    OCMP_REQUIRE(bk.kind == b0.kind && bk.order == b0.order, "the blocks of a vector norm must share kind and order");
    K.loc_off[k] = bk.loc_off;
  }
  OCMP_REQUIRE(b0.nloc <= 20, "at most 20 local dofs per block");
  TableBuilder tb;
  Rule rule = simplex_rule(d, 2 * b0.order);
  K.nq = rule.nq;
  K.nl = b0.nloc;
  K.ld = s.nloc;
  K.ncomp = nb;
  std::vector<double> phi((size_t)rule.nq * b0.nloc), gr((size_t)rule.nq * b0.nloc * d);
  for (int q = 0; q < rule.nq; ++q)
    eval_basis(b0.kind, d, b0.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b0.nloc], &gr[(size_t)q * b0.nloc * d]);
  K.tab_phi = tb.add(phi);
  K.tab_w = tb.add(rule.w);
  DBuf<double> *tab = cached_table(s, "P" + std::to_string(blocks[0]), tb, st);
  K.tab = tab->p;
  const unsigned nbk = nblk(m.nel);
  double *part = ctx->red.p;
  OCMP_REQUIRE(ctx->red.n >= 2 * (size_t)nbk + 2, "reduction scratch too small for this mesh");
  KTimer tm(ctx, ctx->cls("picard_norms"));
  auto launch = [&](auto nlm) {
    constexpr int NLM = decltype(nlm)::value;
    if (d == 2)
      k_l2_pair<2, NLM><<<nbk, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, a, b, part);
    else
      k_l2_pair<3, NLM><<<nbk, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, a, b, part);
  };
  if (b0.nloc <= 4)
    launch(std::integral_constant<int, 4>());
  else if (b0.nloc <= 10)
    launch(std::integral_constant<int, 10>());
  else
    launch(std::integral_constant<int, 20>());
  double *res = ctx->scal.p + 110;
  k_sum_pairs<<<1, 256, 0, st>>>(part, nbk, res);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipMemcpyAsync(out, res, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
}

}  // namespace ocmp

namespace ocmp {

void divergence_norm(Space *sp, int block0, const double *x, double *out) {
  Space &s = *sp;
  const Mesh &m = *s.mesh;
  Ctx *ctx = m.ctx;
  hipStream_t st = ctx->stream;
  const int d = m.dim;
  const Block &b = s.blocks[block0];
  for (int c = 1; c < d; ++c) {
    const Block &bc = s.blocks[block0 + c];
    OCMP_REQUIRE(bc.kind == b.kind && bc.order == b.order && bc.loc_off == b.loc_off + c * b.nloc,
                 "the divergence needs dim consecutive blocks of one vector component");
  }
  KNorm K;
  std::memset(&K, 0, sizeof K);
  TableBuilder tb;
  Rule rule = simplex_rule(d, 2 * b.order);
  K.nq = rule.nq;
  K.nl = b.nloc;
  K.ld = s.nloc;
  K.loc_off = b.loc_off;
  std::vector<double> phi((size_t)rule.nq * b.nloc), gr((size_t)rule.nq * b.nloc * d);
  for (int q = 0; q < rule.nq; ++q)
    eval_basis(b.kind, d, b.order, &rule.lam[q * (d + 1)], &phi[(size_t)q * b.nloc], &gr[(size_t)q * b.nloc * d]);
  K.tab_grad = tb.add(gr);
  K.tab_w = tb.add(rule.w);
  DBuf<double> *tab = cached_table(s, "D" + std::to_string(block0), tb, st);
  K.tab = tab->p;
  const unsigned nb = nblk(m.nel);
  double *part = ctx->red.p;
  OCMP_REQUIRE(ctx->red.n >= 2 * (size_t)nb + 2, "reduction scratch too small for this mesh");
  KTimer tm(ctx, ctx->cls("error_norms"));
  if (d == 2)
    k_div_norm<2><<<nb, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, part);
  else
    k_div_norm<3><<<nb, 256, 0, st>>>(K, m.nel, m.d_coords.p, m.d_elems.p, s.d_edofs.p, x, part);
  double *res = ctx->scal.p + 110;
  k_sum_pairs<<<1, 256, 0, st>>>(part, nb, res);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
}

void linf_norm(Space *sp, int block, const double *x, const double *ref, double *out) {
  Space &s = *sp;
  Ctx *ctx = s.mesh->ctx;
  hipStream_t st = ctx->stream;
  const Block &b = s.blocks[block];
  const unsigned nb = (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, nblk(b.n1)));
  double *part = ctx->red.p;
  OCMP_REQUIRE(ctx->red.n >= (size_t)nb + 2, "reduction scratch too small");
  k_absmax_part<<<nb, 256, 0, st>>>(b.n1, x + b.dof_off, ref ? ref + b.dof_off : nullptr, part);
  double *res = ctx->scal.p + 110;
  k_max_of<<<1, 256, 0, st>>>(part, (int)nb, res);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
}


// ------------------------------------------------------------------ expression programs (host)

namespace {
std::mutex g_expr_mu;
struct ExprTable {
  int next = 0;
  std::set<int> starts;
};
std::map<int, ExprTable> g_expr_tables;  // device -> its program table's bookkeeping
}  // namespace

// stack effect of an opcode: (pops, pushes, inline constants); pops < 0: invalid opcode
static void expr_effect(int op, int &pops, int &pushes, int &consts) {
  pops = -1;
  pushes = 1;
  consts = 0;
  if (op == OCMP_OP_CONST) { pops = 0; consts = 1; }
  else if (op >= OCMP_OP_X && op <= OCMP_OP_T) pops = 0;
  else if (op == OCMP_OP_NEG || (op >= OCMP_OP_SIN && op <= OCMP_OP_ROUND) || op == OCMP_OP_HS || op == OCMP_OP_RAMP)
    pops = 1;
  else if ((op >= OCMP_OP_ADD && op <= OCMP_OP_POW) || op == OCMP_OP_HS2) pops = 2;
  else if (op == OCMP_OP_RAMP4) pops = 4;
}

int expr_register(Ctx *ctx, int n, const double *code) {
  OCMP_REQUIRE(n >= 2 && n <= EXPR_CAP, "expression program length out of range");
  // validate: every opcode known and integral, the stack never under- or overflows, exactly one
  // value left at the final OCMP_OP_END, and no END before it (so every evaluation terminates
  // inside the program)
  int depth = 0;
  for (int pc = 0; pc < n;) {
    const double od = code[pc];
    OCMP_REQUIRE(od == std::floor(od) && od >= 0 && od < 64, "expression program: bad opcode");
    const int op = (int)od;
    if (op == OCMP_OP_END) {
      OCMP_REQUIRE(pc == n - 1, "expression program: END before the end");
      OCMP_REQUIRE(depth == 1, "expression program must leave exactly one value");
      break;
    }
    OCMP_REQUIRE(pc < n - 1, "expression program must end with OCMP_OP_END");
    int pops, pushes, consts;
    expr_effect(op, pops, pushes, consts);
    OCMP_REQUIRE(pops >= 0, "expression program: unknown opcode " + std::to_string(op));
    OCMP_REQUIRE(depth >= pops, "expression program: stack underflow");
    depth += pushes - pops;
    OCMP_REQUIRE(depth <= EXPR_DEPTH, "expression program needs more than 8 stack slots");
    pc += 1 + consts;
    OCMP_REQUIRE(pc < n, "expression program: constant past the end");
  }
  std::lock_guard<std::mutex> lock(g_expr_mu);
  ExprTable &T = g_expr_tables[ctx->device];
  OCMP_REQUIRE(T.next + n <= EXPR_CAP, "expression program table full (64K doubles per device)");
  const int start = T.next;
  OCMP_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(expr_code), code, (size_t)n * sizeof(double),
                                  (size_t)start * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  OCMP_HIP(hipStreamSynchronize(ctx->stream));
  T.starts.insert(start);
  T.next += n;
  return start;
}

void check_fn(const Ctx *ctx, int fn, const double *c) {
  OCMP_REQUIRE(fn >= OCMP_FN_CONST && fn <= OCMP_FN_EXPR, "unknown function id");
  if (fn != OCMP_FN_EXPR) return;
  std::lock_guard<std::mutex> lock(g_expr_mu);
  auto it = g_expr_tables.find(ctx->device);
  const double id = c[0];
  OCMP_REQUIRE(it != g_expr_tables.end() && id == std::floor(id) && id >= 0 && id < EXPR_CAP &&
                   it->second.starts.count((int)id),
               "OCMP_FN_EXPR: c[0] is not a program registered on this device");
}

}  // namespace ocmp
