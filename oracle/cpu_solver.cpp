// Host-CPU restatement of the Krylov solves, for the CPU baseline -- TEST INFRASTRUCTURE.
//
// Only bench.py's CPU-baseline leg (and the tests) load this library. It runs the
// same algorithms as the GPU path on the same CSR matrix, in FP64 with OpenMP over
// all host cores: GMRES(m) right-preconditioned with element-block Jacobi (the DG
// workload, opencmp/models/base_model.py:886-947 with Preconditioner 'local'), and
// Jacobi-PCG (the H1 workloads). Classical Gram-Schmidt, Givens rotations, the
// FP64 true residual ||b - Ax|| / ||b|| on the free dofs as the stopping test.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>
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

}  // extern "C"
