// Reference-element tables: simplex quadrature and the H1 / L2 bases.
//
// H1 (hierarchical, local vertices sorted by global id so that edge and face
// orientation is global): vertex i -> lambda_i; edge (a,b), degree k in 2..3 ->
// lambda_a lambda_b (lambda_b - lambda_a)^(k-2); face (a,b,c), order 3 ->
// lambda_a lambda_b lambda_c. Local order: vertices, edges (degree inner),
// faces. This is the vertex-first layout of NGSolve's H1HighOrderFESpace that
// OpenCMP's spaces use (opencmp/models/poisson.py:60-64).
//
// L2 on triangles: orthonormal Dubiner basis psi_ij = Q_i(2x+y-1, 1-y) *
// P_j^{(2i+1,0)}(2y-1), ordered by total degree then decreasing i.
#include <cmath>
#include <mutex>

#include "internal.h"

namespace ocmp {

static void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      double dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) {
        // final derivative at converged z
        double q0 = 1.0, q1 = 0.0;
        for (int k = 1; k <= n; ++k) {
          double q2 = q1;
          q1 = q0;
          q0 = ((2.0 * k - 1.0) * z * q1 - (k - 1.0) * q2) / k;
        }
        dp = n * (z * q0 - q1) / (z * z - 1.0);
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
        break;
      }
      w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] *= 0.5;
  }
}

void line_rule(int degree, std::vector<double> &t, std::vector<double> &w) {
  int n = std::max(1, (degree + 2) / 2);
  gauss_legendre01(n, t, w);
}

Rule simplex_rule(int dim, int degree) {
  Rule r;
  r.dim = dim;
  std::vector<double> u, wu;
  if (dim == 2) {
    int n = std::max(1, (degree + 3) / 2);
    gauss_legendre01(n, u, wu);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b) {
        double x = u[a], y = (1 - u[a]) * u[b];
        r.lam.insert(r.lam.end(), {1 - x - y, x, y});
        r.w.push_back(wu[a] * wu[b] * (1 - u[a]));
      }
  } else if (dim == 3) {
    int n = std::max(1, (degree + 4) / 2);
    gauss_legendre01(n, u, wu);
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c) {
          double x = u[a], y = (1 - u[a]) * u[b], z = (1 - u[a]) * (1 - u[b]) * u[c];
          r.lam.insert(r.lam.end(), {1 - x - y - z, x, y, z});
          r.w.push_back(wu[a] * wu[b] * wu[c] * (1 - u[a]) * (1 - u[a]) * (1 - u[b]));
        }
  } else {
    throw Error(OCMP_ERR_ARG, "simplex_rule: dim must be 2 or 3");
  }
  r.nq = (int)r.w.size();
  return r;
}

static const int TRI_EDGES[3][2] = {{0, 1}, {0, 2}, {1, 2}};
static const int TET_EDGES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
static const int TET_FACES[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

int nloc1(int kind, int dim, int order) {
  if (kind == OCMP_H1) {
    if (order < 1 || order > 3) throw Error(OCMP_ERR_UNSUPPORTED, "H1 order must be 1..3");
    int ne = dim == 2 ? 3 : 6;
    int n = dim + 1 + ne * (order - 1);
    if (order == 3) n += dim == 2 ? 1 : 4;
    return n;
  }
  if (kind == OCMP_L2) {
    if (dim != 2) throw Error(OCMP_ERR_UNSUPPORTED, "L2 spaces are implemented on triangles");
    if (order < 0 || order > 3) throw Error(OCMP_ERR_UNSUPPORTED, "L2 order must be 0..3");
    return (order + 1) * (order + 2) / 2;
  }
  throw Error(OCMP_ERR_ARG, "unknown space kind");
}

// ---- 2D polynomials for the Dubiner basis: coefficient c[a][b] of x^a y^b, degree <= 3
struct P2 {
  double c[8][8] = {};
};
static P2 pmul(const P2 &A, const P2 &B) {
  P2 o;
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b)
      if (A.c[a][b] != 0.0)
        for (int i = 0; a + i < 8; ++i)
          for (int j = 0; b + j < 8; ++j) o.c[a + i][b + j] += A.c[a][b] * B.c[i][j];
  return o;
}
static P2 padd(const P2 &A, const P2 &B, double sa = 1.0, double sb = 1.0) {
  P2 o;
  for (int a = 0; a < 8; ++a)
    for (int b = 0; b < 8; ++b) o.c[a][b] = sa * A.c[a][b] + sb * B.c[a][b];
  return o;
}
static double peval(const P2 &A, double x, double y) {
  double s = 0.0;
  for (int a = 7; a >= 0; --a) {
    double row = 0.0;
    for (int b = 7; b >= 0; --b) row = row * y + A.c[a][b];
    s = s * x + row;
  }
  return s;
}
static P2 pdx(const P2 &A) {
  P2 o;
  for (int a = 1; a < 8; ++a)
    for (int b = 0; b < 8; ++b) o.c[a - 1][b] = a * A.c[a][b];
  return o;
}
static P2 pdy(const P2 &A) {
  P2 o;
  for (int a = 0; a < 8; ++a)
    for (int b = 1; b < 8; ++b) o.c[a][b - 1] = b * A.c[a][b];
  return o;
}

static std::vector<P2> dubiner(int p) {
  P2 one;
  one.c[0][0] = 1.0;
  P2 t;  // 2x + y - 1
  t.c[0][0] = -1.0;
  t.c[1][0] = 2.0;
  t.c[0][1] = 1.0;
  P2 s;  // 1 - y
  s.c[0][0] = 1.0;
  s.c[0][1] = -1.0;
  P2 z;  // 2y - 1
  z.c[0][0] = -1.0;
  z.c[0][1] = 2.0;
  std::vector<P2> Q = {one, t};
  for (int n = 1; n < p; ++n) {
    P2 a = pmul(t, Q[n]);
    P2 b = pmul(pmul(s, s), Q[n - 1]);
    Q.push_back(padd(a, b, (2.0 * n + 1.0) / (n + 1.0), -double(n) / (n + 1.0)));
  }
  auto jac = [&](int j, double al) {
    std::vector<P2> P = {one};
    if (j >= 1) {
      P2 p1 = padd(z, one, 0.5 * (al + 2.0), 0.5 * al);
      P.push_back(p1);
    }
    for (int n = 1; n < j; ++n) {
      double c1 = 2.0 * (n + 1) * (n + al + 1) * (2 * n + al);
      double c2 = (2 * n + al + 1) * al * al;
      double c3 = (2 * n + al) * (2 * n + al + 1) * (2 * n + al + 2);
      double c4 = 2.0 * (n + al) * n * (2 * n + al + 2);
      P2 a = pmul(z, P[n]);
      P2 r = padd(a, P[n], c3 / c1, c2 / c1);
      r = padd(r, P[n - 1], 1.0, -c4 / c1);
      P.push_back(r);
    }
    return P[j];
  };
  Rule rule = simplex_rule(2, 2 * p + 2);
  std::vector<P2> out;
  for (int deg = 0; deg <= p; ++deg)
    for (int i = deg; i >= 0; --i) {
      int j = deg - i;
      P2 C = pmul(Q[i], jac(j, 2.0 * i + 1.0));
      double nrm = 0.0;
      for (int q = 0; q < rule.nq; ++q) {
        double v = peval(C, rule.lam[3 * q + 1], rule.lam[3 * q + 2]);
        nrm += rule.w[q] * v * v;
      }
      nrm = std::sqrt(nrm);
      for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) C.c[a][b] /= nrm;
      out.push_back(C);
    }
  return out;
}

struct DubinerCache {
  std::mutex mu;
  std::map<int, std::vector<std::array<P2, 3>>> polys;  // value, d/dx, d/dy
  const std::vector<std::array<P2, 3>> &get(int p) {
    std::lock_guard<std::mutex> g(mu);
    auto it = polys.find(p);
    if (it != polys.end()) return it->second;
    std::vector<std::array<P2, 3>> v;
    for (auto &C : dubiner(p)) v.push_back({C, pdx(C), pdy(C)});
    return polys[p] = v;
  }
};
static DubinerCache g_dub;

void eval_basis(int kind, int dim, int order, const double *lam, double *phi, double *g) {
  if (kind == OCMP_L2) {
    const auto &P = g_dub.get(order);
    double x = lam[1], y = lam[2];
    for (size_t i = 0; i < P.size(); ++i) {
      phi[i] = peval(P[i][0], x, y);
      g[2 * i + 0] = peval(P[i][1], x, y);
      g[2 * i + 1] = peval(P[i][2], x, y);
    }
    return;
  }
  // H1: derivatives with respect to barycentrics, then reference gradient d/dx_a = d/dl_{a+1} - d/dl_0
  const int nl = nloc1(kind, dim, order);
  std::vector<double> dl((size_t)nl * (dim + 1), 0.0);
  int c = 0;
  for (int i = 0; i <= dim; ++i) {
    phi[c] = lam[i];
    dl[c * (dim + 1) + i] = 1.0;
    ++c;
  }
  const int ne = dim == 2 ? 3 : 6;
  for (int e = 0; e < ne; ++e) {
    int a = dim == 2 ? TRI_EDGES[e][0] : TET_EDGES[e][0];
    int b = dim == 2 ? TRI_EDGES[e][1] : TET_EDGES[e][1];
    double la = lam[a], lb = lam[b];
    for (int k = 2; k <= order; ++k) {
      if (k == 2) {
        phi[c] = la * lb;
        dl[c * (dim + 1) + a] = lb;
        dl[c * (dim + 1) + b] = la;
      } else {
        phi[c] = la * lb * (lb - la);
        dl[c * (dim + 1) + a] = lb * lb - 2 * la * lb;
        dl[c * (dim + 1) + b] = 2 * la * lb - la * la;
      }
      ++c;
    }
  }
  if (order >= 3) {
    int nf = dim == 2 ? 1 : 4;
    for (int f = 0; f < nf; ++f) {
      int a = dim == 2 ? 0 : TET_FACES[f][0];
      int b = dim == 2 ? 1 : TET_FACES[f][1];
      int cc = dim == 2 ? 2 : TET_FACES[f][2];
      phi[c] = lam[a] * lam[b] * lam[cc];
      dl[c * (dim + 1) + a] = lam[b] * lam[cc];
      dl[c * (dim + 1) + b] = lam[a] * lam[cc];
      dl[c * (dim + 1) + cc] = lam[a] * lam[b];
      ++c;
    }
  }
  for (int i = 0; i < nl; ++i)
    for (int a = 0; a < dim; ++a) g[i * dim + a] = dl[i * (dim + 1) + a + 1] - dl[i * (dim + 1) + 0];
}

}  // namespace ocmp
