// Ritz values of a GMRES cycle and the Newton-basis shifts of the s-step GMRES.
//
// The s-step cycle builds its Krylov block as v_{t+1} = (A - d_t I) v_t + c_t v_{t-1}. With the
// shifts at Leja-ordered Ritz values of an earlier cycle the block is far better conditioned than
// the monomial one, which matters when the block is stored in FP32. Host code only.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "internal.h"

namespace ocmp {

// Eigenvalues of the n x n upper Hessenberg matrix H (column-major, leading dimension ldh) by the
// Francis double-shift QR iteration on a working copy: deflate at negligible subdiagonal entries,
// read 1x1 and 2x2 diagonal blocks off as they split, exceptional shifts after 10 and 20 sweeps
// without deflation. Returns false if an eigenvalue takes more than 60 sweeps.
bool hessenberg_eigvals(int n, const double *H, int ldh, double *wr, double *wi) {
  std::vector<double> A((size_t)n * n, 0.0);
  auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
  double anorm = 0.0;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i <= std::min(j + 1, n - 1); ++i) {
      a(i, j) = H[(size_t)j * ldh + i];
      anorm += std::fabs(a(i, j));
    }
  int hi = n - 1;       // active window is rows/columns lo..hi
  double shift = 0.0;   // accumulated exceptional shifts
  int sweeps = 0;
  while (hi >= 0) {
    // lowest row lo of the unreduced block ending at hi
    int lo = hi;
    while (lo > 0) {
      double s = std::fabs(a(lo - 1, lo - 1)) + std::fabs(a(lo, lo));
      if (s == 0.0) s = anorm;
      if (std::fabs(a(lo, lo - 1)) + s == s) {
        a(lo, lo - 1) = 0.0;
        break;
      }
      --lo;
    }
    double x = a(hi, hi);
    if (lo == hi) {  // 1x1 block
      wr[hi] = x + shift;
      wi[hi] = 0.0;
      --hi;
      sweeps = 0;
      continue;
    }
    double y = a(hi - 1, hi - 1), w = a(hi, hi - 1) * a(hi - 1, hi);
    if (lo == hi - 1) {  // 2x2 block: its two eigenvalues
      const double p = 0.5 * (y - x), q = p * p + w, z = std::sqrt(std::fabs(q));
      x += shift;
      if (q >= 0.0) {
        const double zz = p + std::copysign(z, p);
        wr[hi - 1] = wr[hi] = x + zz;
        if (zz != 0.0) wr[hi] = x - w / zz;
        wi[hi - 1] = wi[hi] = 0.0;
      } else {
        wr[hi - 1] = wr[hi] = x + p;
        wi[hi - 1] = z;
        wi[hi] = -z;
      }
      hi -= 2;
      sweeps = 0;
      continue;
    }
    if (sweeps == 60) return false;
    if (sweeps == 10 || sweeps == 20) {  // exceptional shift
      shift += x;
      for (int i = 0; i <= hi; ++i) a(i, i) -= x;
      const double s = std::fabs(a(hi, hi - 1)) + std::fabs(a(hi - 1, hi - 2));
      x = y = 0.75 * s;
      w = -0.4375 * s * s;
    }
    ++sweeps;
    // start row m of the double-shift bulge: where two consecutive subdiagonal entries are small
    // enough that the first column of (H - s1)(H - s2) decouples from the rows above
    int m = hi - 2;
    double p = 0.0, q = 0.0, r = 0.0;
    for (;; --m) {
      const double z = a(m, m), rr = x - z, ss = y - z;
      p = (rr * ss - w) / a(m + 1, m) + a(m, m + 1);
      q = a(m + 1, m + 1) - z - rr - ss;
      r = a(m + 2, m + 1);
      const double s = std::fabs(p) + std::fabs(q) + std::fabs(r);
      p /= s;
      q /= s;
      r /= s;
      if (m == lo) break;
      const double u = std::fabs(a(m, m - 1)) * (std::fabs(q) + std::fabs(r));
      const double v = std::fabs(p) * (std::fabs(a(m - 1, m - 1)) + std::fabs(z) + std::fabs(a(m + 1, m + 1)));
      if (u + v == v) break;
    }
    for (int i = m + 2; i <= hi; ++i) {
      a(i, i - 2) = 0.0;
      if (i != m + 2) a(i, i - 3) = 0.0;
    }
    // chase the bulge with 3x3 Householder reflections
    for (int k = m; k <= hi - 1; ++k) {
      const bool last = (k == hi - 1);
      if (k != m) {
        p = a(k, k - 1);
        q = a(k + 1, k - 1);
        r = last ? 0.0 : a(k + 2, k - 1);
        x = std::fabs(p) + std::fabs(q) + std::fabs(r);
        if (x == 0.0) continue;
        p /= x;
        q /= x;
        r /= x;
      }
      const double s = std::copysign(std::sqrt(p * p + q * q + r * r), p);
      if (s == 0.0) continue;
      if (k == m) {
        if (lo != m) a(k, k - 1) = -a(k, k - 1);
      } else {
        a(k, k - 1) = -s * x;
      }
      p += s;
      x = p / s;
      y = q / s;
      const double z = r / s;
      q /= p;
      r /= p;
      for (int j = k; j < n; ++j) {  // rows k..k+2
        double t = a(k, j) + q * a(k + 1, j);
        if (!last) {
          t += r * a(k + 2, j);
          a(k + 2, j) -= t * z;
        }
        a(k + 1, j) -= t * y;
        a(k, j) -= t * x;
      }
      const int imax = std::min(hi, k + 3);
      for (int i = 0; i <= imax; ++i) {  // columns k..k+2
        double t = x * a(i, k) + y * a(i, k + 1);
        if (!last) {
          t += z * a(i, k + 2);
          a(i, k + 2) -= t * r;
        }
        a(i, k + 1) -= t * q;
        a(i, k) -= t;
      }
    }
  }
  return true;
}

// s shifts in modified Leja order from the n Ritz values (wr, wi): the first is the value of
// largest modulus, each next one maximises the product of distances to the points chosen so far
// (sums of logarithms); a complex value is taken with its conjugate right after it, as the pair
// (d, 0), (d, (Im z)^2) of the real three-term step. A complex value that would be the last
// shift enters by its real part.
void newton_shifts(int n, const double *wr, const double *wi, int s, double *d, double *c) {
  struct Cand {
    double re, im;
    bool used;
  };
  std::vector<Cand> cand;
  for (int j = 0; j < n; ++j)
    if (wi[j] >= 0.0 && std::isfinite(wr[j]) && std::isfinite(wi[j])) cand.push_back({wr[j], wi[j], false});
  std::vector<std::pair<double, double>> pts;
  int t = 0;
  while (t < s) {
    int best = -1;
    double bscore = -INFINITY;
    for (size_t j = 0; j < cand.size(); ++j) {
      if (cand[j].used) continue;
      double score;
      if (pts.empty()) {
        score = std::hypot(cand[j].re, cand[j].im);
      } else {
        score = 0.0;
        for (auto &u : pts) score += std::log(std::hypot(cand[j].re - u.first, cand[j].im - u.second));
      }
      if (best < 0 || score > bscore) {
        best = (int)j;
        bscore = score;
      }
    }
    if (best < 0) {  // fewer Ritz values than shifts: repeat the last shift (a monomial tail)
      d[t] = t > 0 ? d[t - 1] : 1.0;
      c[t] = 0.0;
      ++t;
      continue;
    }
    Cand &z = cand[best];
    z.used = true;
    if (z.im > 0.0 && t + 1 < s) {
      d[t] = z.re;
      c[t] = 0.0;
      d[t + 1] = z.re;
      c[t + 1] = z.im * z.im;
      pts.push_back({z.re, z.im});
      pts.push_back({z.re, -z.im});
      t += 2;
    } else {
      d[t] = z.re;
      c[t] = 0.0;
      pts.push_back({z.re, 0.0});
      ++t;
    }
  }
}

}  // namespace ocmp
