// Host-CPU restatement of the assembly and the Krylov solves, for the CPU baseline -- TEST
// INFRASTRUCTURE.
//
// Only bench.py's CPU-baseline leg (and the tests) load this library. It runs the
// same algorithms as the GPU path on the same CSR matrix, in FP64 with OpenMP over
// all host cores: GMRES(m) right-preconditioned with element-block Jacobi (the DG
// workload, opencmp/models/base_model.py:886-947 with Preconditioner 'local'),
// Jacobi-PCG (the H1 workloads), and flexible GMRES with the Schur preconditioner
// (velocity Jacobi, Jacobi-PCG on C + B D^-1 B^T for the pressure) for the saddle
// points. Classical Gram-Schmidt, Givens rotations, the FP64 true residual
// ||b - Ax|| / ||b|| on the free dofs as the stopping test. The Oseen assembly is the
// textbook element loop: quadrature per element, then a scatter with atomic adds.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

namespace {

void spmv(int64_t n, const int32_t *rp, const int32_t *col, const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int32_t j = rp[r]; j < rp[r + 1]; ++j) s += val[j] * x[col[j]];
    y[r] = s;
  }
}

double dot(int64_t n, const double *a, const double *b) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// inverse of each nb x nb diagonal block (Gauss-Jordan with partial pivoting)
void block_inverse(int64_t n, int nb, const int32_t *rp, const int32_t *col, const double *val,
                   std::vector<double> &binv) {
  const int64_t nblk = n / nb;
  binv.assign((size_t)nblk * nb * nb, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < nblk; ++e) {
    std::vector<double> a((size_t)nb * 2 * nb, 0.0);
    for (int i = 0; i < nb; ++i) {
      const int64_t r = e * nb + i;
      for (int32_t j = rp[r]; j < rp[r + 1]; ++j) {
        const int64_t c = col[j];
        if (c >= e * nb && c < (e + 1) * nb) a[(size_t)i * 2 * nb + (c - e * nb)] = val[j];
      }
      a[(size_t)i * 2 * nb + nb + i] = 1.0;
    }
    for (int k = 0; k < nb; ++k) {
      int p = k;
      for (int i = k + 1; i < nb; ++i)
        if (std::fabs(a[(size_t)i * 2 * nb + k]) > std::fabs(a[(size_t)p * 2 * nb + k])) p = i;
      if (p != k)
        for (int j = 0; j < 2 * nb; ++j) std::swap(a[(size_t)k * 2 * nb + j], a[(size_t)p * 2 * nb + j]);
      const double d = a[(size_t)k * 2 * nb + k];
      for (int j = 0; j < 2 * nb; ++j) a[(size_t)k * 2 * nb + j] /= d;
      for (int i = 0; i < nb; ++i) {
        if (i == k) continue;
        const double f = a[(size_t)i * 2 * nb + k];
        if (f != 0.0)
          for (int j = 0; j < 2 * nb; ++j) a[(size_t)i * 2 * nb + j] -= f * a[(size_t)k * 2 * nb + j];
      }
    }
    for (int i = 0; i < nb; ++i)
      for (int j = 0; j < nb; ++j) binv[((size_t)e * nb + i) * nb + j] = a[(size_t)i * 2 * nb + nb + j];
  }
}

void block_apply(int64_t n, int nb, const std::vector<double> &binv, const double *v, double *z) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n / nb; ++e)
    for (int i = 0; i < nb; ++i) {
      double s = 0.0;
      for (int j = 0; j < nb; ++j) s += binv[((size_t)e * nb + i) * nb + j] * v[e * nb + j];
      z[e * nb + i] = s;
    }
}

// ---- smoothed-aggregation AMG, in the textbook form (sparse products P = (I - w D^-1 A) P_t,
// A_c = P^T A P, a V(1,1) cycle with pre-smoothing, restriction, coarse correction, prolongation,
// post-smoothing). Aggregation: strength |a_ij| >= theta sqrt(|a_ii a_jj|); pass 1 roots whose
// strong neighbours are all free, pass 2 joins a strong first-pass aggregate, pass 3 groups the
// rest; prolongation damping 4 / (3 lambda_max(D^-1 A)) from 30 power steps; damped Jacobi (2/3)
// smoothing; a dense inverse on the last level (at most 48 rows, up to 1200 where coarsening
// stalls). Rows with a zero Jacobi entry on the first level are left out of every operator.

constexpr double AMG_THETA = 0.08, AMG_SMOOTH = 2.0 / 3.0;
constexpr int64_t AMG_COARSE = 48, AMG_DENSE_LIMIT = 1200;
constexpr int AMG_MAX_LEVELS = 10;

struct Csr {
  int64_t n = 0, m = 0;  // rows, columns
  std::vector<int32_t> rp, col;
  std::vector<double> val;
};

void csr_mv(const Csr &A, const double *x, double *y) {
  spmv(A.n, A.rp.data(), A.col.data(), A.val.data(), x, y);
}

Csr csr_transpose(const Csr &A) {
  Csr T;
  T.n = A.m;
  T.m = A.n;
  T.rp.assign(T.n + 1, 0);
  for (int32_t c : A.col) T.rp[c + 1]++;
  for (int64_t i = 0; i < T.n; ++i) T.rp[i + 1] += T.rp[i];
  T.col.resize(A.col.size());
  T.val.resize(A.col.size());
  std::vector<int32_t> fill(T.rp.begin(), T.rp.end() - 1);
  for (int64_t i = 0; i < A.n; ++i)
    for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
      const int32_t t = fill[A.col[k]]++;
      T.col[t] = (int32_t)i;
      T.val[t] = A.val[k];
    }
  return T;
}

// C = A B, row by row with a dense accumulator per thread (columns sorted)
Csr csr_product(const Csr &A, const Csr &B) {
  std::vector<std::vector<int32_t>> cc(A.n);
  std::vector<std::vector<double>> cv(A.n);
#pragma omp parallel
  {
    std::vector<double> acc(B.m, 0.0);
    std::vector<char> on(B.m, 0);
    std::vector<int32_t> t;
#pragma omp for schedule(dynamic, 256)
    for (int64_t i = 0; i < A.n; ++i) {
      t.clear();
      for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k) {
        const double a = A.val[k];
        const int32_t j = A.col[k];
        for (int32_t e = B.rp[j]; e < B.rp[j + 1]; ++e) {
          const int32_t c = B.col[e];
          if (!on[c]) {
            on[c] = 1;
            t.push_back(c);
          }
          acc[c] += a * B.val[e];
        }
      }
      std::sort(t.begin(), t.end());
      cc[i] = t;
      cv[i].resize(t.size());
      for (size_t q = 0; q < t.size(); ++q) {
        cv[i][q] = acc[t[q]];
        acc[t[q]] = 0.0;
        on[t[q]] = 0;
      }
    }
  }
  Csr C;
  C.n = A.n;
  C.m = B.m;
  C.rp.assign(A.n + 1, 0);
  for (int64_t i = 0; i < A.n; ++i) C.rp[i + 1] = C.rp[i] + (int32_t)cc[i].size();
  C.col.resize(C.rp[A.n]);
  C.val.resize(C.rp[A.n]);
  for (int64_t i = 0; i < A.n; ++i) {
    std::copy(cc[i].begin(), cc[i].end(), C.col.begin() + C.rp[i]);
    std::copy(cv[i].begin(), cv[i].end(), C.val.begin() + C.rp[i]);
  }
  return C;
}

double csr_diag(const Csr &A, int64_t i) {
  for (int32_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
    if (A.col[k] == i) return A.val[k];
  return 0.0;
}

struct AmgLevelCpu {
  Csr A, P, R;
  std::vector<double> dinv, x, r, bc;
};

struct AmgCpu {
  std::vector<AmgLevelCpu> lv;
  std::vector<double> inv;  // dense inverse of the last level (row-major)
  int64_t applications = 0;
};

// dense inverse of the last level; rows with dinv 0 are left out (zero rows and columns)
void amg_dense_inverse(const Csr &A, const std::vector<double> &dinv, std::vector<double> &inv) {
  const int64_t n = A.n;
  std::vector<int64_t> id;
  for (int64_t i = 0; i < n; ++i)
    if (dinv[i] != 0.0) id.push_back(i);
  const int64_t k = (int64_t)id.size();
  std::vector<int64_t> pos(n, -1);
  for (int64_t a = 0; a < k; ++a) pos[id[a]] = a;
  std::vector<double> M((size_t)k * 2 * k, 0.0);
  for (int64_t a = 0; a < k; ++a) {
    for (int32_t q = A.rp[id[a]]; q < A.rp[id[a] + 1]; ++q)
      if (pos[A.col[q]] >= 0) M[(size_t)a * 2 * k + pos[A.col[q]]] = A.val[q];
    M[(size_t)a * 2 * k + k + a] = 1.0;
  }
  for (int64_t c = 0; c < k; ++c) {  // Gauss-Jordan with partial pivoting
    int64_t p = c;
    for (int64_t a = c + 1; a < k; ++a)
      if (std::fabs(M[(size_t)a * 2 * k + c]) > std::fabs(M[(size_t)p * 2 * k + c])) p = a;
    if (p != c)
      for (int64_t j = 0; j < 2 * k; ++j) std::swap(M[(size_t)c * 2 * k + j], M[(size_t)p * 2 * k + j]);
    const double ip = 1.0 / M[(size_t)c * 2 * k + c];
    for (int64_t j = 0; j < 2 * k; ++j) M[(size_t)c * 2 * k + j] *= ip;
    for (int64_t a = 0; a < k; ++a) {
      const double f = M[(size_t)a * 2 * k + c];
      if (a == c || f == 0.0) continue;
      for (int64_t j = 0; j < 2 * k; ++j) M[(size_t)a * 2 * k + j] -= f * M[(size_t)c * 2 * k + j];
    }
  }
  inv.assign((size_t)n * n, 0.0);
  for (int64_t a = 0; a < k; ++a)
    for (int64_t c = 0; c < k; ++c) inv[(size_t)id[a] * n + id[c]] = M[(size_t)a * 2 * k + k + c];
}

// the hierarchy of A (dinv0: its Jacobi, 0 on rows left out)
void amg_setup(AmgCpu &M, Csr A, std::vector<double> dinv) {
  std::vector<uint8_t> use(A.n);
  for (int64_t i = 0; i < A.n; ++i) use[i] = dinv[i] != 0.0;
  for (int l = 0;; ++l) {
    const int64_t n = A.n;
    AmgLevelCpu L;
    auto finish = [&]() {
      if (n > AMG_DENSE_LIMIT) throw std::runtime_error("AMG: the last level is too large for a dense inverse");
      amg_dense_inverse(A, dinv, M.inv);
      L.A = std::move(A);
      L.dinv = std::move(dinv);
      L.x.resize(n);
      M.lv.push_back(std::move(L));
    };
    if (n <= AMG_COARSE || l == AMG_MAX_LEVELS - 1) {
      finish();
      return;
    }
    std::vector<double> dg(n);
    for (int64_t i = 0; i < n; ++i) dg[i] = csr_diag(A, i);
    auto strong = [&](int64_t i, int32_t q) {
      const int32_t j = A.col[q];
      return j != i && use[j] && std::fabs(A.val[q]) >= AMG_THETA * std::sqrt(std::fabs(dg[i] * dg[j]));
    };
    std::vector<int32_t> agg(n, -1);
    int32_t na = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (!use[i] || agg[i] >= 0) continue;
      bool free_nb = true;
      for (int32_t q = A.rp[i]; q < A.rp[i + 1]; ++q) free_nb = free_nb && !(strong(i, q) && agg[A.col[q]] >= 0);
      if (!free_nb) continue;
      agg[i] = na;
      for (int32_t q = A.rp[i]; q < A.rp[i + 1]; ++q)
        if (strong(i, q)) agg[A.col[q]] = na;
      ++na;
    }
    const std::vector<int32_t> first = agg;
    for (int64_t i = 0; i < n; ++i) {
      if (!use[i] || agg[i] >= 0) continue;
      for (int32_t q = A.rp[i]; q < A.rp[i + 1]; ++q)
        if (strong(i, q) && first[A.col[q]] >= 0) {
          agg[i] = first[A.col[q]];
          break;
        }
    }
    for (int64_t i = 0; i < n; ++i) {
      if (!use[i] || agg[i] >= 0) continue;
      agg[i] = na;
      for (int32_t q = A.rp[i]; q < A.rp[i + 1]; ++q)
        if (strong(i, q) && agg[A.col[q]] < 0) agg[A.col[q]] = na;
      ++na;
    }
    if (na == 0 || na > (int64_t)(0.9 * (double)n)) {
      finish();
      return;
    }
    // lambda_max(D^-1 A) on the usable rows by power steps from a fixed start
    std::vector<double> v(n), y(n);
    for (int64_t i = 0; i < n; ++i) v[i] = use[i] ? 1.0 + 0.5 * std::sin(0.7 * (double)i) : 0.0;
    double lam = 1.0;
    for (int it = 0; it < 30; ++it) {
      double yy = 0.0, yv = 0.0, vv = 0.0;
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        if (use[i] && dg[i] != 0.0) {
          for (int32_t q = A.rp[i]; q < A.rp[i + 1]; ++q)
            if (use[A.col[q]]) s += A.val[q] * v[A.col[q]];
          s /= dg[i];
        }
        y[i] = s;
        yy += s * s;
        yv += s * v[i];
        vv += v[i] * v[i];
      }
      lam = vv > 0.0 ? yv / vv : 1.0;
      for (int64_t i = 0; i < n; ++i) v[i] = yy > 0.0 ? y[i] / std::sqrt(yy) : 0.0;
    }
    const double omega = 4.0 / (3.0 * std::max(lam, 1e-300));
    // tentative prolongation, then P = P_t - omega D^-1 A P_t on the usable rows
    Csr Pt;
    Pt.n = n;
    Pt.m = na;
    Pt.rp.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (use[i]) {
        Pt.col.push_back(agg[i]);
        Pt.val.push_back(1.0);
      }
      Pt.rp[i + 1] = (int32_t)Pt.col.size();
    }
    const Csr AP = csr_product(A, Pt);
    Csr P;
    P.n = n;
    P.m = na;
    P.rp.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) {
      if (use[i]) {
        const double f = dg[i] != 0.0 ? omega / dg[i] : 0.0;
        for (int32_t q = AP.rp[i]; q < AP.rp[i + 1]; ++q) {
          P.col.push_back(AP.col[q]);
          P.val.push_back((AP.col[q] == agg[i] ? 1.0 : 0.0) - f * AP.val[q]);
        }
      }
      P.rp[i + 1] = (int32_t)P.col.size();
    }
    Csr R = csr_transpose(P);
    Csr Ac = csr_product(R, csr_product(A, P));
    L.A = std::move(A);
    L.P = std::move(P);
    L.R = std::move(R);
    L.dinv = std::move(dinv);
    L.x.resize(n);
    L.r.resize(n);
    L.bc.resize(na);
    M.lv.push_back(std::move(L));
    A = std::move(Ac);
    dinv.assign(A.n, 0.0);
    for (int64_t i = 0; i < A.n; ++i) {
      const double d = csr_diag(A, i);
      dinv[i] = d != 0.0 ? 1.0 / d : 0.0;
    }
    use.assign(A.n, 1);
  }
}

// x = V(1,1)-cycle approximation of A_l^{-1} b
void amg_vcycle(AmgCpu &M, size_t l, const double *b, double *x) {
  AmgLevelCpu &L = M.lv[l];
  const int64_t n = L.A.n;
  if (l + 1 == M.lv.size()) {
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t j = 0; j < n; ++j) s += M.inv[(size_t)i * n + j] * b[j];
      x[i] = s;
    }
    return;
  }
  AmgLevelCpu &Cl = M.lv[l + 1];
  double *r = L.r.data();
  const double *dv = L.dinv.data();
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) x[i] = AMG_SMOOTH * dv[i] * b[i];
  csr_mv(L.A, x, r);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
  csr_mv(L.R, r, L.bc.data());
  double *xc = Cl.x.data();
  amg_vcycle(M, l + 1, L.bc.data(), xc);
  csr_mv(L.P, xc, r);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) x[i] += r[i];
  csr_mv(L.A, x, r);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) x[i] += AMG_SMOOTH * dv[i] * (b[i] - r[i]);
}

// x = (cycles V-cycles) b: the first from zero, each further one on the current residual
void amg_apply(AmgCpu &M, int cycles, const double *b, double *x) {
  const AmgLevelCpu &L0 = M.lv[0];
  const int64_t n = L0.A.n;
  amg_vcycle(M, 0, b, x);
  std::vector<double> rr(cycles > 1 ? n : 0), ee(cycles > 1 ? n : 0);
  for (int c = 1; c < cycles; ++c) {
    csr_mv(L0.A, x, rr.data());
    for (int64_t i = 0; i < n; ++i) rr[i] = b[i] - rr[i];
    amg_vcycle(M, 0, rr.data(), ee.data());
    for (int64_t i = 0; i < n; ++i) x[i] += ee[i];
  }
  ++M.applications;
}

void amg_levels(const AmgCpu &M, int64_t *levels) {
  if (!levels) return;
  for (int l = 0; l < AMG_MAX_LEVELS; ++l) levels[l] = l < (int)M.lv.size() ? M.lv[l].A.n : 0;
}

}  // namespace

extern "C" {

int ocmp_cpu_threads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
}

// GMRES(m) right-preconditioned with element-block Jacobi (blocks of nb consecutive dofs).
// Runs until ||b - Ax|| <= tol ||b|| or maxit iterations; returns the iterations done and
// writes the final relative residual to *rel.
int ocmp_cpu_gmres_bj(int64_t n, const int32_t *rp, const int32_t *col, const double *val, int nb,
                      const double *b, double *x, int m, double tol, int maxit, double *rel) {
  std::vector<double> binv;
  block_inverse(n, nb, rp, col, val, binv);
  std::vector<double> V((size_t)(m + 1) * n), z(n), w(n), r(n);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  const double bn = std::sqrt(dot(n, b, b));
  int it = 0;
  double rn = 0.0;
  for (;;) {
    spmv(n, rp, col, val, x, r.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    rn = std::sqrt(dot(n, r.data(), r.data()));
    if (rn <= tol * bn || it >= maxit) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) V[i] = r[i] / rn;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rn;
    int k = 0;
    for (; k < m && it < maxit; ++k, ++it) {
      block_apply(n, nb, binv, &V[(size_t)k * n], z.data());
      spmv(n, rp, col, val, z.data(), w.data());
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = dot(n, &V[(size_t)j * n], w.data());  // CGS
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        double s = w[i];
        for (int j = 0; j <= k; ++j) s -= H[(size_t)j * m + k] * V[(size_t)j * n + i];
        w[i] = s;
      }
      const double hn = std::sqrt(dot(n, w.data(), w.data()));
      H[(size_t)(k + 1) * m + k] = hn;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) V[(size_t)(k + 1) * n + i] = w[i] / hn;
      for (int j = 0; j < k; ++j) {
        const double a = H[(size_t)j * m + k], c = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * c;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * c;
      }
      const double a = H[(size_t)k * m + k], c = H[(size_t)(k + 1) * m + k];
      const double d = std::hypot(a, c);
      cs[k] = a / d;
      sn[k] = c / d;
      H[(size_t)k * m + k] = d;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) <= tol * bn) {
        ++k;
        ++it;
        break;
      }
    }
    for (int i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
      y[i] = s / H[(size_t)i * m + i];
    }
    std::fill(w.begin(), w.end(), 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += y[j] * V[(size_t)j * n + i];
      w[i] = s;
    }
    block_apply(n, nb, binv, w.data(), z.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] += z[i];
  }
  *rel = bn > 0 ? rn / bn : rn;
  return it;
}

// Jacobi-PCG on the free dofs (free[i] != 0); Dirichlet entries of x are left as given.
int ocmp_cpu_pcg_jacobi(int64_t n, const int32_t *rp, const int32_t *col, const double *val, const uint8_t *freem,
                        const double *b, double *x, double tol, int maxit, double *rel) {
  std::vector<double> dinv(n), r(n), z(n), p(n), q(n);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double d = 0.0;
    for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
      if (col[j] == i) d = val[j];
    dinv[i] = freem[i] && d != 0.0 ? 1.0 / d : 0.0;
  }
  spmv(n, rp, col, val, x, q.data());
  double bn2 = 0.0;
#pragma omp parallel for reduction(+ : bn2) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = freem[i] ? b[i] - q[i] : 0.0;
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
    bn2 += freem[i] ? b[i] * b[i] : 0.0;
  }
  const double bn = std::sqrt(bn2);
  double rz = dot(n, r.data(), z.data());
  int it = 0;
  double rn = std::sqrt(dot(n, r.data(), r.data()));
  while (rn > tol * bn && it < maxit) {
    spmv(n, rp, col, val, p.data(), q.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      if (!freem[i]) q[i] = 0.0;
    const double alpha = rz / dot(n, p.data(), q.data());
    double rr = 0.0, rz2 = 0.0;
#pragma omp parallel for reduction(+ : rr, rz2) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
      rr += r[i] * r[i];
      rz2 += r[i] * z[i];
    }
    const double beta = rz2 / rz;
    rz = rz2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    rn = std::sqrt(rr);
    ++it;
  }
  *rel = bn > 0 ? rn / bn : rn;
  return it;
}

// CG on the free dofs preconditioned with `cycles` AMG V-cycles of the matrix (its Jacobi on the free
// rows, the other rows left out); Dirichlet entries of x are left as given. levels (optional, 10
// entries) receives the rows of each level, 0 past the last.
int ocmp_cpu_pcg_amg(int64_t n, const int32_t *rp, const int32_t *col, const double *val, const uint8_t *freem,
                     const double *b, double *x, double tol, int maxit, int cycles, int64_t *levels,
                     double *rel) {
  Csr A;
  A.n = A.m = n;
  A.rp.assign(rp, rp + n + 1);
  A.col.assign(col, col + rp[n]);
  A.val.assign(val, val + rp[n]);
  std::vector<double> dinv(n, 0.0);
  for (int64_t i = 0; i < n; ++i) {
    const double d = csr_diag(A, i);
    dinv[i] = freem[i] && d != 0.0 ? 1.0 / d : 0.0;
  }
  AmgCpu M;
  amg_setup(M, A, dinv);
  amg_levels(M, levels);
  std::vector<double> r(n), z(n), p(n), q(n);
  spmv(n, rp, col, val, x, r.data());
  for (int64_t i = 0; i < n; ++i) r[i] = freem[i] ? b[i] - r[i] : 0.0;
  double bn2 = 0.0;
  for (int64_t i = 0; i < n; ++i) bn2 += freem[i] ? b[i] * b[i] : 0.0;
  const double bn = std::sqrt(bn2);
  amg_apply(M, cycles, r.data(), z.data());
  p = z;
  double g = dot(n, r.data(), z.data());
  int it = 0;
  double rn = std::sqrt(dot(n, r.data(), r.data()));
  while (rn > tol * bn && it < maxit) {
    spmv(n, rp, col, val, p.data(), q.data());
    for (int64_t i = 0; i < n; ++i)
      if (!freem[i]) q[i] = 0.0;
    const double a = g / dot(n, p.data(), q.data());
    for (int64_t i = 0; i < n; ++i) {
      x[i] += a * p[i];
      r[i] -= a * q[i];
    }
    amg_apply(M, cycles, r.data(), z.data());
    const double g2 = dot(n, r.data(), z.data());
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + (g2 / g) * p[i];
    g = g2;
    rn = std::sqrt(dot(n, r.data(), r.data()));
    ++it;
  }
  *rel = bn > 0.0 ? rn / bn : rn;
  return it;
}

// Implicit-Euler Oseen operator of a Taylor-Hood pair on simplices (opencmp/models/ins.py:
// 174-216 without the outlet's boundary term, with the time derivative of
// opencmp/solvers/time_integration_schemes.py:135-170): per velocity component the block
//   M + dt nu K - dt C(w),  C(w)_ij = int phi_j (w . grad phi_i),
// and the couplings -dt int psi_k d_c phi_j (pressure rows) and -dt int d_c phi_i psi_k
// (velocity rows), plus -peps dt int psi_k psi_l. Reference basis values and their
// barycentric derivatives come tabulated at the quadrature points; the wind is the
// velocity part of W. Element matrices are added to the CSR values with atomic adds.
int ocmp_cpu_assemble_oseen(int dim, int64_t nel, const double *X, int nlu, const int32_t *udofs, int nlp,
                            const int32_t *pdofs, int64_t nu1, int64_t poff, int nq, const double *qw,
                            const double *phiu, const double *dlu, const double *phip, const double *dlp,
                            const double *W, double dt, double nu, double peps, const int32_t *rp,
                            const int32_t *col, double *val) {
  const int nv = dim + 1;
  const int nloc = dim * nlu + nlp;
  int bad = 0;
#pragma omp parallel
  {
    std::vector<double> gu((size_t)nq * nlu * 3), gp((size_t)nq * nlp * 3), K((size_t)nloc * nloc);
    std::vector<int64_t> g(nloc);
#pragma omp for schedule(static) reduction(+ : bad)
    for (int64_t e = 0; e < nel; ++e) {
      const double *xe = X + e * nv * dim;
      // J columns x_k - x_0; grad lambda_k (k >= 1) = rows of J^{-1}; grad lambda_0 = -sum
      double J[3][3] = {{0}}, Ji[3][3] = {{0}}, det;
      for (int a = 0; a < dim; ++a)
        for (int k = 0; k < dim; ++k) J[a][k] = xe[(k + 1) * dim + a] - xe[a];
      if (dim == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0] = J[1][1] / det;
        Ji[0][1] = -J[0][1] / det;
        Ji[1][0] = -J[1][0] / det;
        Ji[1][1] = J[0][0] / det;
      } else {
        det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const int a1 = (b + 1) % 3, a2 = (b + 2) % 3, b1 = (a + 1) % 3, b2 = (a + 2) % 3;
            Ji[a][b] = (J[a1][b1] * J[a2][b2] - J[a1][b2] * J[a2][b1]) / det;
          }
      }
      double gl[4][3] = {{0}};
      for (int k = 1; k <= dim; ++k)
        for (int a = 0; a < dim; ++a) gl[k][a] = Ji[k - 1][a];
      for (int a = 0; a < dim; ++a) {
        double s = 0.0;
        for (int k = 1; k <= dim; ++k) s -= gl[k][a];
        gl[0][a] = s;
      }
      const double vol = std::fabs(det);
      for (int q = 0; q < nq; ++q) {
        for (int i = 0; i < nlu; ++i)
          for (int a = 0; a < dim; ++a) {
            double s = 0.0;
            for (int k = 0; k < nv; ++k) s += dlu[((size_t)q * nlu + i) * nv + k] * gl[k][a];
            gu[((size_t)q * nlu + i) * 3 + a] = s;
          }
        for (int i = 0; i < nlp; ++i)
          for (int a = 0; a < dim; ++a) {
            double s = 0.0;
            for (int k = 0; k < nv; ++k) s += dlp[((size_t)q * nlp + i) * nv + k] * gl[k][a];
            gp[((size_t)q * nlp + i) * 3 + a] = s;
          }
      }
      std::fill(K.begin(), K.end(), 0.0);
      for (int q = 0; q < nq; ++q) {
        const double wq = qw[q] * vol;
        const double *pu = phiu + (size_t)q * nlu, *pp = phip + (size_t)q * nlp;
        double w[3] = {0, 0, 0};
        for (int c = 0; c < dim; ++c)
          for (int j = 0; j < nlu; ++j) w[c] += W[c * nu1 + udofs[e * nlu + j]] * pu[j];
        for (int i = 0; i < nlu; ++i) {
          const double *gi = &gu[((size_t)q * nlu + i) * 3];
          double wgi = 0.0;
          for (int a = 0; a < dim; ++a) wgi += w[a] * gi[a];
          for (int j = 0; j < nlu; ++j) {
            const double *gj = &gu[((size_t)q * nlu + j) * 3];
            double gg = 0.0;
            for (int a = 0; a < dim; ++a) gg += gi[a] * gj[a];
            const double v = wq * (pu[i] * pu[j] + dt * nu * gg - dt * pu[j] * wgi);
            for (int c = 0; c < dim; ++c) K[(size_t)(c * nlu + i) * nloc + c * nlu + j] += v;
          }
          for (int k = 0; k < nlp; ++k)
            for (int c = 0; c < dim; ++c) {
              const double v = -dt * wq * gi[c] * pp[k];
              K[(size_t)(c * nlu + i) * nloc + dim * nlu + k] += v;  // velocity row, pressure column
              K[(size_t)(dim * nlu + k) * nloc + c * nlu + i] += v;  // pressure row, velocity column
            }
        }
        for (int k = 0; k < nlp; ++k)
          for (int l = 0; l < nlp; ++l) K[(size_t)(dim * nlu + k) * nloc + dim * nlu + l] -= peps * dt * wq * pp[k] * pp[l];
      }
      for (int c = 0; c < dim; ++c)
        for (int i = 0; i < nlu; ++i) g[c * nlu + i] = c * nu1 + udofs[e * nlu + i];
      for (int k = 0; k < nlp; ++k) g[dim * nlu + k] = poff + pdofs[e * nlp + k];
      for (int i = 0; i < nloc; ++i) {
        const int32_t *c0 = col + rp[g[i]], *c1 = col + rp[g[i] + 1];
        for (int j = 0; j < nloc; ++j) {
          const int32_t *at = std::lower_bound(c0, c1, (int32_t)g[j]);
          if (at == c1 || *at != g[j]) {
            ++bad;
            continue;
          }
#pragma omp atomic
          val[at - col] += K[(size_t)i * nloc + j];
        }
      }
    }
  }
  return bad;
}

// Flexible GMRES(m) with the Schur preconditioner of the GPU path: on the rows outside [s0, s1)
// z = v / a_ii; on [s0, s1) z = -S^{-1} v by Jacobi-PCG from zero until (r, Mr) <= stol^2 (r0, Mr0)
// or smax iterations, S = C + B D^{-1} B^T with D_uu = max(|a_uu|, 1e-2 max_j |a_uj|) (columns
// outside [s0, s1)) over the free velocity rows, and S_ii below 1e-10 max S_ii left out. With amg > 0
// the pressure step is amg smoothed-aggregation V-cycles on S instead of the PCG (levels, optional,
// gets the hierarchy's rows per level). With tri, the velocity rows follow the Schur rows as the
// upper block-triangular step: z_u = (v_u - A_up z_p) / a_uu. Returns the outer iterations; *inner
// gets the inner PCG iterations (or the V-cycle applications) summed.
int ocmp_cpu_fgmres_schur(int64_t n, const int32_t *rp, const int32_t *col, const double *val, const uint8_t *freem,
                          int64_t s0, int64_t s1, const double *b, double *x, int m, double tol, int maxit,
                          double stol, int smax, double *rel, int64_t *inner, int amg, int64_t *levels, int tri) {
  const int64_t ns = s1 - s0;
  std::vector<double> dinv(n, 0.0), dpos(n, 0.0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    if (!freem[r] || (r >= s0 && r < s1)) continue;
    double d = 0.0, mx = 0.0;
    for (int32_t j = rp[r]; j < rp[r + 1]; ++j) {
      if (col[j] >= s0 && col[j] < s1) continue;
      const double a = std::fabs(val[j]);
      if (col[j] == r) d = val[j];
      mx = std::max(mx, a);
    }
    dinv[r] = d != 0.0 ? 1.0 / d : 0.0;
    const double dd = std::max(std::fabs(d), 1e-2 * mx);
    dpos[r] = dd > 0.0 ? 1.0 / dd : 0.0;
  }
  // S by rows: a dense accumulator per thread over the ns pressure columns
  std::vector<std::vector<int32_t>> scol(ns);
  std::vector<std::vector<double>> sval(ns);
#pragma omp parallel
  {
    std::vector<double> acc(ns, 0.0);
    std::vector<char> on(ns, 0);
    std::vector<int32_t> touched;
#pragma omp for schedule(dynamic, 512)
    for (int64_t i = 0; i < ns; ++i) {
      const int64_t p = s0 + i;
      touched.clear();
      auto add = [&](int64_t q, double v) {
        if (!on[q]) {
          on[q] = 1;
          touched.push_back((int32_t)q);
        }
        acc[q] += v;
      };
      for (int32_t j = rp[p]; j < rp[p + 1]; ++j) {
        const int64_t u = col[j];
        if (u >= s0 && u < s1) {
          add(u - s0, -val[j]);
          continue;
        }
        const double f = val[j] * dpos[u];
        if (f == 0.0) continue;
        for (int32_t t = rp[u]; t < rp[u + 1]; ++t)
          if (col[t] >= s0 && col[t] < s1) add(col[t] - s0, f * val[t]);
      }
      std::sort(touched.begin(), touched.end());
      scol[i] = touched;
      sval[i].resize(touched.size());
      for (size_t k = 0; k < touched.size(); ++k) {
        sval[i][k] = acc[touched[k]];
        acc[touched[k]] = 0.0;
        on[touched[k]] = 0;
      }
    }
  }
  std::vector<double> sd(ns, 0.0);
  double smx = 0.0;
  for (int64_t i = 0; i < ns; ++i) {
    for (size_t k = 0; k < scol[i].size(); ++k)
      if (scol[i][k] == i) sd[i] = sval[i][k];
    if (!freem[s0 + i] || !(sd[i] > 0.0)) sd[i] = 0.0;
    smx = std::max(smx, sd[i]);
  }
  for (int64_t i = 0; i < ns; ++i) sd[i] = sd[i] > 1e-10 * smx ? 1.0 / sd[i] : 0.0;
  std::vector<double> pr(ns), pu(ns), pp(ns), pq(ns);
  int64_t inner_total = 0;
  AmgCpu amgS;
  if (amg > 0) {
    Csr S;
    S.n = S.m = ns;
    S.rp.assign(ns + 1, 0);
    for (int64_t i = 0; i < ns; ++i) S.rp[i + 1] = S.rp[i] + (int32_t)scol[i].size();
    for (int64_t i = 0; i < ns; ++i) {
      S.col.insert(S.col.end(), scol[i].begin(), scol[i].end());
      S.val.insert(S.val.end(), sval[i].begin(), sval[i].end());
    }
    amg_setup(amgS, std::move(S), sd);
    amg_levels(amgS, levels);
  }
  // z[s0:s1] = -S^{-1} v[s0:s1]
  auto schur_solve = [&](const double *v, double *z) {
    double *xz = z + s0;
    if (amg > 0) {
      for (int64_t i = 0; i < ns; ++i) pr[i] = -v[s0 + i];
      amg_apply(amgS, amg, pr.data(), xz);
      inner_total = amgS.applications;
      return;
    }
    double g = 0.0;
#pragma omp parallel for reduction(+ : g) schedule(static)
    for (int64_t i = 0; i < ns; ++i) {
      xz[i] = 0.0;
      pr[i] = -v[s0 + i];
      pu[i] = sd[i] * pr[i];
      pp[i] = pu[i];
      g += pr[i] * pu[i];
    }
    const double g0 = g;
    for (int it = 0; it < smax && g > stol * stol * g0; ++it) {
      double d = 0.0;
#pragma omp parallel for reduction(+ : d) schedule(static)
      for (int64_t i = 0; i < ns; ++i) {
        double s = 0.0;
        for (size_t k = 0; k < scol[i].size(); ++k) s += sval[i][k] * pp[scol[i][k]];
        pq[i] = s;
        d += pp[i] * s;
      }
      const double a = d != 0.0 ? g / d : 0.0;
      double g2 = 0.0;
#pragma omp parallel for reduction(+ : g2) schedule(static)
      for (int64_t i = 0; i < ns; ++i) {
        xz[i] += a * pp[i];
        pr[i] -= a * pq[i];
        pu[i] = sd[i] * pr[i];
        g2 += pr[i] * pu[i];
      }
      const double be = g != 0.0 ? g2 / g : 0.0;
      g = g2;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < ns; ++i) pp[i] = pu[i] + be * pp[i];
      ++inner_total;
    }
  };
  auto masked_spmv = [&](const double *xv, double *y) {
    spmv(n, rp, col, val, xv, y);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      if (!freem[i]) y[i] = 0.0;
  };
  std::vector<double> V((size_t)(m + 1) * n), Z((size_t)m * n), w(n), r(n);
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  double bn2 = 0.0;
#pragma omp parallel for reduction(+ : bn2) schedule(static)
  for (int64_t i = 0; i < n; ++i) bn2 += freem[i] ? b[i] * b[i] : 0.0;
  const double bn = std::sqrt(bn2);
  int it = 0;
  double rn = 0.0;
  for (;;) {
    spmv(n, rp, col, val, x, r.data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = freem[i] ? b[i] - r[i] : 0.0;
    rn = std::sqrt(dot(n, r.data(), r.data()));
    if (rn <= tol * bn || it >= maxit) break;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) V[i] = r[i] / rn;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = rn;
    int k = 0;
    for (; k < m && it < maxit; ++k, ++it) {
      const double *vk = &V[(size_t)k * n];
      double *zk = &Z[(size_t)k * n];
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) zk[i] = dinv[i] * vk[i];
      schur_solve(vk, zk);
      if (tri) {
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; ++r) {
          if ((r >= s0 && r < s1) || dinv[r] == 0.0) continue;
          double sc = 0.0;
          for (int32_t j = rp[r]; j < rp[r + 1]; ++j)
            if (col[j] >= s0 && col[j] < s1) sc += val[j] * zk[col[j]];
          zk[r] -= dinv[r] * sc;
        }
      }
      masked_spmv(zk, w.data());
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = dot(n, &V[(size_t)j * n], w.data());
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) {
        double s = w[i];
        for (int j = 0; j <= k; ++j) s -= H[(size_t)j * m + k] * V[(size_t)j * n + i];
        w[i] = s;
      }
      const double hn = std::sqrt(dot(n, w.data(), w.data()));
      H[(size_t)(k + 1) * m + k] = hn;
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < n; ++i) V[(size_t)(k + 1) * n + i] = hn > 0.0 ? w[i] / hn : 0.0;
      for (int j = 0; j < k; ++j) {
        const double a = H[(size_t)j * m + k], c = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * c;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * c;
      }
      const double a = H[(size_t)k * m + k], c = H[(size_t)(k + 1) * m + k];
      const double d = std::hypot(a, c);
      cs[k] = a / d;
      sn[k] = c / d;
      H[(size_t)k * m + k] = d;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (std::fabs(g[k + 1]) <= tol * bn) {
        ++k;
        ++it;
        break;
      }
    }
    for (int i = k - 1; i >= 0; --i) {
      double s = g[i];
      for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
      y[i] = s / H[(size_t)i * m + i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = 0; j < k; ++j) s += y[j] * Z[(size_t)j * n + i];
      x[i] += s;
    }
  }
  *rel = bn > 0 ? rn / bn : rn;
  *inner = inner_total;
  return it;
}

}  // extern "C"
