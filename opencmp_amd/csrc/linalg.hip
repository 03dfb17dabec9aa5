// Sparse linear algebra on gfx950: CSR SpMV, block-Jacobi, Jacobi-PCG and
// restarted GMRES(m), in FP64 or as FP32 inner solves inside FP64 iterative
// refinement. All reductions are two-stage with a fixed order (partials per
// workgroup, then one workgroup per value), so results are bitwise
// reproducible run to run.
//
// Launch geometry: reduction-carrying kernels use a capped grid whose blocks
// own contiguous row ranges; block ids are remapped so that each of the 8
// XCDs (which receive workgroups round-robin) works on one contiguous slice
// of the rows, keeping the x-gathers of neighbouring rows in one L2.
#include <cmath>
#include <cstring>

#include "internal.h"

namespace ocmp {

constexpr int NTHR = 256;
constexpr int NXCD = 8;
constexpr int MAX_RED_BLOCKS = 2048;

static inline unsigned nblk(int64_t n, int b = NTHR) { return (unsigned)((n + b - 1) / b); }

__device__ __forceinline__ int xcd_block() {
  // gridDim.x is a multiple of NXCD; blocks b, b+8, b+16.. land on one XCD
  const int per = gridDim.x / NXCD;
  return (blockIdx.x % NXCD) * per + blockIdx.x / NXCD;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// sum over the block; valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += sh[i];
  }
  __syncthreads();
  return s;
}

static inline int red_grid(int64_t work_items) {
  int64_t g = (work_items + NTHR - 1) / NTHR;
  g = std::min<int64_t>(g, MAX_RED_BLOCKS);
  g = std::max<int64_t>(g, NXCD);
  g = (g + NXCD - 1) / NXCD * NXCD;
  return (int)g;
}

// out[v] = sum_b part[v*nparts + b], one block per value, fixed order
__global__ void k_reduce(const double *__restrict__ part, int nparts, double *__restrict__ out, int accumulate_sqrt) {
  __shared__ double sh[NTHR / 64];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += part[(int64_t)v * nparts + b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[v] = accumulate_sqrt ? sqrt(s) : s;
}

// ------------------------------------------------------------------ CSR SpMV

// mode 0: y = A x
// mode 1: y = mask*(b - A x), partial sum y^2
// mode 2: y = mask*(A x), partial sum x.y                (CG: q = A p, p.q)
// mode 3: y = mask*(A (cs*sx*x))   cs: column scaling (right Jacobi), sx: device scalar  (GMRES on H1)
template <int G, typename VT, typename XT>
__global__ void __launch_bounds__(NTHR) k_spmv_csr(int64_t n, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col, const VT *__restrict__ val,
                                                   const XT *__restrict__ x, XT *__restrict__ y,
                                                   const double *__restrict__ b, const uint8_t *__restrict__ mask,
                                                   const XT *__restrict__ cs, const double *__restrict__ sx,
                                                   int mode, double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  constexpr int RPB = NTHR / G;  // rows per block-iteration
  const int lane = threadIdx.x % G;
  const int64_t ngroups = (n + RPB - 1) / RPB;
  const int lb = xcd_block();
  const int64_t chunk = (ngroups + gridDim.x - 1) / gridDim.x;
  const int64_t g0 = lb * chunk, g1 = min(ngroups, g0 + chunk);
  const XT scale = (mode == 3 && sx) ? (XT)sx[0] : (XT)1;
  double acc = 0.0;
  for (int64_t gr = g0; gr < g1; ++gr) {
    const int64_t r = gr * RPB + threadIdx.x / G;
    XT s = 0;
    if (r < n) {
      const int32_t j0 = rp[r], j1 = rp[r + 1];
      if (mode == 3 && cs) {
        for (int32_t j = j0 + lane; j < j1; j += G) {
          const int32_t c = col[j];
          s += (XT)val[j] * (cs[c] * x[c]);
        }
      } else {
        for (int32_t j = j0 + lane; j < j1; j += G) s += (XT)val[j] * x[col[j]];
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (r < n && lane == 0) {
      if (mode == 0) {
        y[r] = s;
      } else if (mode == 1) {
        const double v = (mask && !mask[r]) ? 0.0 : b[r] - (double)s;
        y[r] = (XT)v;
        acc += v * v;
      } else if (mode == 2) {
        const XT v = (mask && !mask[r]) ? (XT)0 : s;
        y[r] = v;
        acc += (double)x[r] * (double)v;
      } else {
        y[r] = (mask && !mask[r]) ? (XT)0 : s * scale;
      }
    }
  }
  if (part) {
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) part[lb] = acc;
  }
}

template <typename VT, typename XT>
static void spmv_dispatch(Matrix *A, const VT *val, const XT *x, XT *y, const double *b, const uint8_t *mask,
                          const XT *cs, const double *sx, int mode, double *part, int grid, hipStream_t st) {
  const int64_t n = A->n;
#define OCMP_SPMV_CASE(GG)                                                                                     \
  case GG:                                                                                                     \
    k_spmv_csr<GG, VT, XT><<<grid, NTHR, 0, st>>>(n, A->d_row_ptr.p, A->d_col.p, val, x, y, b, mask, cs, sx, \
                                                  mode, part);                                                 \
    break;
  switch (A->spmv_group) {
    OCMP_SPMV_CASE(4)
    OCMP_SPMV_CASE(8)
    OCMP_SPMV_CASE(16)
    OCMP_SPMV_CASE(32)
    OCMP_SPMV_CASE(64)
    default:
      throw Error(OCMP_ERR_INTERNAL, "bad SpMV group size");
  }
#undef OCMP_SPMV_CASE
}

// algorithmic bytes of one CSR SpMV: values + column indices + row pointers + x once + y once
static double csr_bytes(const Matrix *A, int wv, int wx) {
  return (double)A->nnz * (wv + 4) + 4.0 * (A->n + 1) + 2.0 * A->n * wx;
}

static int spmv_grid(Matrix *A) {
  const int rpb = NTHR / A->spmv_group;
  return red_grid((A->n + rpb - 1) / rpb * NTHR);
}

void spmv_f64(Matrix *A, const double *x, double *y) {
  Ctx *ctx = A->ctx;
  KTimer tm(ctx, ctx->cls("spmv_csr_f64"), csr_bytes(A, 8, 8));
  spmv_dispatch<double, double>(A, A->d_val.p, x, y, nullptr, nullptr, nullptr, nullptr, 0, nullptr, spmv_grid(A),
                                ctx->stream);
  OCMP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ DG right-preconditioned operator (ELL layout)
//
// Ahat = A * blockdiag(A)^{-1}. Its diagonal blocks are the identity, so only
// the three facet-neighbour blocks are stored, column-major "ELL":
// ahat[(k*nb + j) * n + r] = Ahat(r, nbr_k*nb + j), r = e*nb + i, k = local facet.
// Boundary facets store zero blocks and point at the element itself.
// nbr[k*nel + e] is the neighbour element across facet k.

template <typename T, int NB>
__global__ void __launch_bounds__(NTHR) k_dg_ahat_spmv(int64_t nel, const int32_t *__restrict__ nbr,
                                                       const T *__restrict__ ahat, const T *__restrict__ x,
                                                       const double *__restrict__ sx, T *__restrict__ y) {
  const int64_t n = nel * NB;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = r / NB;
  const T scale = sx ? (T)sx[0] : (T)1;
  T s = x[r];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int64_t c0 = (int64_t)nbr[k * nel + e] * NB;
    const T *a = ahat + (int64_t)(k * NB) * n + r;
#pragma unroll
    for (int j = 0; j < NB; ++j) s += a[(int64_t)j * n] * x[c0 + j];
  }
  y[r] = s * scale;
}

// invert the diagonal blocks (Gauss-Jordan, partial pivoting), one thread per element
template <int NB>
__global__ void k_dg_block_inverse(int64_t nel, const int32_t *__restrict__ rp, const int8_t *__restrict__ dg_pos,
                                   const double *__restrict__ val, double *__restrict__ binv,
                                   float *__restrict__ binv32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double a[NB][NB], v[NB][NB];
  const int p0 = dg_pos[e * 4];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const double *row = val + rp[e * NB + i] + p0 * NB;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      a[i][j] = row[j];
      v[i][j] = i == j ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    int piv = c;
    double best = fabs(a[c][c]);
#pragma unroll
    for (int r = c + 1; r < NB; ++r)
      if (fabs(a[r][c]) > best) {
        best = fabs(a[r][c]);
        piv = r;
      }
    if (piv != c) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double t = a[c][j];
        a[c][j] = a[piv][j];
        a[piv][j] = t;
        t = v[c][j];
        v[c][j] = v[piv][j];
        v[piv][j] = t;
      }
    }
    const double inv = 1.0 / a[c][c];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      a[c][j] *= inv;
      v[c][j] *= inv;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r)
      if (r != c) {
        const double f = a[r][c];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          a[r][j] -= f * a[c][j];
          v[r][j] -= f * v[c][j];
        }
      }
  }
  double *o = binv + e * NB * NB;
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      o[i * NB + j] = v[i][j];
      if (binv32) binv32[e * NB * NB + i * NB + j] = (float)v[i][j];
    }
}

// ahat(r, block k) = A(r, block nbr_k) * inv(D_nbr_k); one thread per row
template <int NB>
__global__ void k_dg_build_ahat(int64_t nel, const int32_t *__restrict__ rp, const int8_t *__restrict__ dg_pos,
                                const int32_t *__restrict__ elem_nb, const double *__restrict__ val,
                                const double *__restrict__ binv, double *__restrict__ ahat,
                                float *__restrict__ ahat32, int32_t *__restrict__ nbr) {
  const int64_t n = nel * NB;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = r / NB;
  const int i = (int)(r % NB);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t nb = elem_nb[e * 3 + k];
    const int pos = dg_pos[e * 4 + 1 + k];
    double out[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) out[j] = 0.0;
    if (nb >= 0 && pos >= 0) {
      const double *arow = val + rp[r] + pos * NB;
      const double *bi = binv + (int64_t)nb * NB * NB;
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        const double al = arow[l];
#pragma unroll
        for (int j = 0; j < NB; ++j) out[j] += al * bi[l * NB + j];
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (ahat) ahat[(int64_t)(k * NB + j) * n + r] = out[j];
      if (ahat32) ahat32[(int64_t)(k * NB + j) * n + r] = (float)out[j];
    }
    if (i == 0) nbr[k * nel + e] = (nb >= 0 && pos >= 0) ? nb : (int32_t)e;
  }
}

// x[e] += inv(D_e) u[e]
template <int NB>
__global__ void k_dg_block_apply_add(int64_t nel, const double *__restrict__ binv, const double *__restrict__ u,
                                     double *__restrict__ x) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nel * NB) return;
  const int64_t e = r / NB;
  const int i = (int)(r % NB);
  const double *bi = binv + e * NB * NB + i * NB;
  double s = 0.0;
#pragma unroll
  for (int j = 0; j < NB; ++j) s += bi[j] * u[e * NB + j];
  x[r] += s;
}

// ------------------------------------------------------------------ Jacobi

__global__ void k_jacobi_diag(int64_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                              const double *__restrict__ val, const uint8_t *__restrict__ mask,
                              double *__restrict__ dinv, float *__restrict__ dinv32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int32_t j = rp[r]; j < rp[r + 1]; ++j)
    if (col[j] == r) d = val[j];
  const double v = (mask && !mask[r]) || d == 0.0 ? 0.0 : 1.0 / d;
  dinv[r] = v;
  if (dinv32) dinv32[r] = (float)v;
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// ------------------------------------------------------------------ vector kernels

template <typename T>
__global__ void k_fill_t(T *__restrict__ p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// partial sums of a.b and (optionally) c.c
template <typename T>
__global__ void __launch_bounds__(NTHR) k_dot(int64_t n, const T *__restrict__ a, const T *__restrict__ b,
                                              const uint8_t *__restrict__ mask, double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int lb = xcd_block();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(n, i0 + chunk);
  double s = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
    if (!mask || mask[i]) s += (double)a[i] * (double)b[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[lb] = s;
}

// CG update with device scalars: alpha = rz[it&1] / pq
// x += alpha p; r -= alpha q; z = dinv r; partials of r.z and r.r
template <typename T>
__global__ void __launch_bounds__(NTHR) k_cg_update(int64_t n, T *__restrict__ x, T *__restrict__ r,
                                                    T *__restrict__ z, const T *__restrict__ p,
                                                    const T *__restrict__ q, const T *__restrict__ dinv,
                                                    const double *__restrict__ sc, int it, double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const double alpha_d = sc[it & 1] / sc[2];
  const T alpha = (T)alpha_d;
  const int lb = xcd_block();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(n, i0 + chunk);
  double rz = 0.0, rr = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = dinv[i] * ri;
    z[i] = zi;
    rz += (double)ri * (double)zi;
    rr += (double)ri * (double)ri;
  }
  rz = block_sum(rz, sh);
  rr = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part[lb] = rz;
    part[gridDim.x + lb] = rr;
  }
}

// p = z + beta p, beta = rz[(it+1)&1] / rz[it&1]
template <typename T>
__global__ void k_cg_p(int64_t n, T *__restrict__ p, const T *__restrict__ z, const double *__restrict__ sc, int it) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T beta = (T)(sc[(it + 1) & 1] / sc[it & 1]);
  p[i] = z[i] + beta * p[i];
}

// z = dinv * r (with optional scale), p = z
template <typename T, typename S>
__global__ void k_jacobi_apply(int64_t n, const T *__restrict__ dinv, const S *__restrict__ r, double scale,
                               T *__restrict__ z, T *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = dinv[i] * (T)((double)r[i] * scale);
  z[i] = v;
  if (p) p[i] = v;
}

// dst = (T) (src * scale)
template <typename S, typename T>
__global__ void k_scale_copy(int64_t n, const S *__restrict__ src, double scale, T *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (T)((double)src[i] * scale);
}

// x64 += scale * d
template <typename T>
__global__ void k_axpy64(int64_t n, double scale, const T *__restrict__ d, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += scale * (double)d[i];
}

// ------------------------------------------------------------------ GMRES kernels
//
// The Krylov basis is stored unnormalised: V[k] with device scale vs[k], so
// v_k = vs[k] * V[k]. This saves the separate normalisation pass.

constexpr int KC = 8;  // dot products per thread chunk

// partials of V[k].w for k in [0, nk) and of w.w (as index nk)
template <typename T>
__global__ void __launch_bounds__(NTHR) k_mdot(int64_t n, const T *__restrict__ V, int64_t ldv,
                                               const T *__restrict__ w, int nk, double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int lb = xcd_block();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(n, i0 + chunk);
  const int k0 = blockIdx.y * KC;
  double acc[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) acc[c] = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double wi = (double)w[i];
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      const int k = k0 + c;
      if (k < nk)
        acc[c] += (double)V[(int64_t)k * ldv + i] * wi;
      else if (k == nk)
        acc[c] += wi * wi;
    }
  }
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    const int k = k0 + c;
    const double s = block_sum(acc[c], sh);
    if (threadIdx.x == 0 && k <= nk) part[(int64_t)k * gridDim.x + lb] = s;
  }
}

// w -= sum_k c_k V[k] with c_k = vs[k]^2 * dot[k]; partial of |w|^2
template <typename T>
__global__ void __launch_bounds__(NTHR) k_gs_update(int64_t n, const T *__restrict__ V, int64_t ldv,
                                                    T *__restrict__ w, int nk, const double *__restrict__ dot,
                                                    const double *__restrict__ vs, double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  __shared__ double cks[256];
  for (int k = threadIdx.x; k < nk; k += blockDim.x) cks[k] = vs[k] * vs[k] * dot[k];
  __syncthreads();
  const int lb = xcd_block();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(n, i0 + chunk);
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    double s = (double)w[i];
    for (int k = 0; k < nk; ++k) s -= cks[k] * (double)V[(int64_t)k * ldv + i];
    const T o = (T)s;
    w[i] = o;
    acc += (double)o * (double)o;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[lb] = acc;
}

// Givens step for column j (single thread).
// dot[0..j] = V[k].w, wn2 = |w_after|^2 (from the update partials, reduced into dot[j+1]).
// State: R (column-major, ldr = m+1), cs/sn, g; vs[j+1] = 1/h_{j+1,j}; res[it] = |g_{j+1}| * gscale.
__global__ void k_givens(int j, int m, const double *__restrict__ dot, double *__restrict__ vs,
                         double *__restrict__ R, double *__restrict__ csn, double *__restrict__ g,
                         double *__restrict__ res, int it, double gscale) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ld = m + 1;
  double *col = R + (int64_t)j * ld;
  for (int k = 0; k <= j; ++k) col[k] = vs[k] * dot[k];
  const double hn = sqrt(dot[j + 1]);
  col[j + 1] = hn;
  vs[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
  for (int k = 0; k < j; ++k) {
    const double c = csn[2 * k], s = csn[2 * k + 1];
    const double a = col[k], b = col[k + 1];
    col[k] = c * a + s * b;
    col[k + 1] = -s * a + c * b;
  }
  const double a = col[j], b = col[j + 1];
  const double den = hypot(a, b);
  const double c = den > 0.0 ? a / den : 1.0, s = den > 0.0 ? b / den : 0.0;
  csn[2 * j] = c;
  csn[2 * j + 1] = s;
  col[j] = den;
  col[j + 1] = 0.0;
  const double gj = g[j];
  g[j] = c * gj;
  g[j + 1] = -s * gj;
  res[it] = fabs(g[j + 1]) * gscale;
}

// back substitution R y = g (k columns), then coef[l] = y_l * vs[l]
__global__ void k_gmres_solve(int k, int m, const double *__restrict__ R, const double *__restrict__ g,
                              const double *__restrict__ vs, double *__restrict__ coef) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ld = m + 1;
  for (int i = k - 1; i >= 0; --i) {
    double s = g[i];
    for (int l = i + 1; l < k; ++l) s -= R[(int64_t)l * ld + i] * coef[l];
    coef[i] = s / R[(int64_t)i * ld + i];
  }
  for (int i = 0; i < k; ++i) coef[i] *= vs[i];
}

// u = scale * sum_l coef[l] V[l]   (in FP64)
template <typename T>
__global__ void k_combine(int64_t n, const T *__restrict__ V, int64_t ldv, int k, const double *__restrict__ coef,
                          double scale, double *__restrict__ u) {
  __shared__ double cf[256];
  for (int l = threadIdx.x; l < k; l += blockDim.x) cf[l] = coef[l] * scale;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int l = 0; l < k; ++l) s += cf[l] * (double)V[(int64_t)l * ldv + i];
  u[i] = s;
}

__global__ void k_gmres_init(int m, double beta_scaled, double *__restrict__ g, double *__restrict__ vs,
                             double inv_beta) {
  const int i = threadIdx.x;
  for (int k = i; k <= m; k += blockDim.x) g[k] = 0.0;
  if (i == 0) {
    g[0] = beta_scaled;
    vs[0] = inv_beta;
  }
}

// x += dinv .* u  (H1 right Jacobi)  or x += u (no preconditioner)
__global__ void k_diag_apply_add(int64_t n, const double *__restrict__ dinv, const double *__restrict__ u,
                                 double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (dinv ? dinv[i] : 1.0) * u[i];
}

// ================================================================== host drivers

struct Work {
  // scratch owned by the matrix' solver state; grown on demand
  DBuf<double> r64, u64, t64;
  DBuf<double> vd;   // FP64 work vectors (n * k)
  DBuf<float> vf;    // FP32 work vectors
  DBuf<double> part; // partial sums
  DBuf<double> small;// device scalars / Hessenberg etc.
  double *hbuf = nullptr;  // pinned host
  size_t hcap = 0;
  DBuf<int32_t> nbr;       // DG neighbour table for the ELL operator
};

static std::map<Matrix *, Work *> g_work;

static Work &work(Matrix *A) {
  auto it = g_work.find(A);
  if (it != g_work.end()) return *it->second;
  auto *w = new Work();
  g_work[A] = w;
  return *w;
}

static double *pinned(Work &w, size_t n) {
  if (w.hcap < n) {
    if (w.hbuf) (void)hipHostFree(w.hbuf);
    OCMP_HIP(hipHostMalloc((void **)&w.hbuf, n * sizeof(double), hipHostMallocDefault));
    w.hcap = n;
  }
  return w.hbuf;
}

// reduce nvals partial rows of `grid` entries into dst (device)
static void reduce_to(const double *part, int grid, int nvals, double *dst, hipStream_t st, bool do_sqrt = false) {
  k_reduce<<<nvals, NTHR, 0, st>>>(part, grid, dst, do_sqrt ? 1 : 0);
}

// FP64 residual r = mask*(b - A x); returns ||r|| (syncs)
static double residual64(Matrix *A, const double *b, const double *x, double *r, Work &w) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int grid = spmv_grid(A);
  w.part.alloc(std::max<size_t>(w.part.n, (size_t)4 * MAX_RED_BLOCKS + 64 * MAX_RED_BLOCKS));
  {
    KTimer tm(ctx, ctx->cls("spmv_csr_f64_residual"), csr_bytes(A, 8, 8) + 8.0 * A->n);
    spmv_dispatch<double, double>(A, A->d_val.p, x, r, b, A->space->d_free.p, (const double *)nullptr, nullptr, 1,
                                  w.part.p, grid, st);
  }
  double *sc = ctx->scal.p;
  reduce_to(w.part.p, grid, 1, sc + 100, st);
  double *h = pinned(w, 8);
  OCMP_HIP(hipMemcpyAsync(h, sc + 100, sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  return std::sqrt(h[0]);
}

static double norm_masked(Matrix *A, const double *v, Work &w) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int grid = red_grid(A->n);
  k_dot<double><<<grid, NTHR, 0, st>>>(A->n, v, v, A->space->d_free.p, w.part.p);
  reduce_to(w.part.p, grid, 1, ctx->scal.p + 101, st);
  double *h = pinned(w, 8);
  OCMP_HIP(hipMemcpyAsync(h, ctx->scal.p + 101, sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  return std::sqrt(h[0]);
}

void precond_update(Precond *P) {
  Matrix *A = P->A;
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const bool mixed = P->precision == OCMP_MIXED;
  KTimer tm(ctx, ctx->cls("precond_setup"));
  if (P->kind == OCMP_PRE_JACOBI || P->kind == OCMP_PRE_NONE) {
    P->dinv.alloc(A->n);
    if (mixed) P->dinv32.alloc(A->n);
    if (P->kind == OCMP_PRE_JACOBI)
      k_jacobi_diag<<<nblk(A->n), NTHR, 0, st>>>(A->n, A->d_row_ptr.p, A->d_col.p, A->d_val.p, A->space->d_free.p,
                                                 P->dinv.p, mixed ? P->dinv32.p : nullptr);
    if (mixed) {
      A->d_val32.alloc(A->nnz);
      k_to_f32<<<nblk(A->nnz), NTHR, 0, st>>>(A->nnz, A->d_val.p, A->d_val32.p);
    }
  } else {
    const Mesh &m = *A->space->mesh;
    const int nb = A->dg_nb;
    P->binv.alloc((size_t)m.nel * nb * nb);
    P->ahat.alloc(mixed ? 0 : (size_t)A->n * 3 * nb);
    if (mixed) P->ahat32.alloc((size_t)A->n * 3 * nb);
    Work &w = work(A);
    w.nbr.alloc((size_t)3 * m.nel);
#define OCMP_DG_CASE(NBV)                                                                                          \
  case NBV:                                                                                                        \
    k_dg_block_inverse<NBV><<<nblk(m.nel, 128), 128, 0, st>>>(m.nel, A->d_row_ptr.p, A->d_dg_pos.p, A->d_val.p,       \
                                                         P->binv.p, nullptr);                                      \
    k_dg_build_ahat<NBV><<<nblk(A->n), NTHR, 0, st>>>(m.nel, A->d_row_ptr.p, A->d_dg_pos.p, m.d_elem_nb.p,       \
                                                      A->d_val.p, P->binv.p, mixed ? nullptr : P->ahat.p,          \
                                                      mixed ? P->ahat32.p : nullptr, w.nbr.p);                     \
    break;
    switch (nb) {
      OCMP_DG_CASE(1)
      OCMP_DG_CASE(3)
      OCMP_DG_CASE(6)
      OCMP_DG_CASE(10)
      default:
        throw Error(OCMP_ERR_UNSUPPORTED, "block Jacobi block size");
    }
#undef OCMP_DG_CASE
  }
  OCMP_HIP(hipGetLastError());
  P->ready = true;
}

// ------------------------------------------------------------------ CG

template <typename VT, typename T>
static int cg_inner(Matrix *A, const VT *val, const T *dinv, T *x, T *r, T *z, T *p, T *q, double tol_rr,
                    int maxit, int check_every, Work &w, int &iters_done, double *last_rr) {
  // x, r, z, p set by the caller: r = residual, z = dinv r, p = z, sc[0] = r.z
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  const int sgrid = spmv_grid(A);
  const int vgrid = red_grid(n);
  double *sc = ctx->scal.p;  // [0],[1]: rz ring, [2]: pq, [3]: rr
  DBuf<double> &hist = w.small;
  const int H = std::max(check_every, 1);
  hist.alloc(std::max<size_t>(hist.n, 4096));
  double *hh = pinned(w, H + 8);
  const int c_spmv = ctx->cls(sizeof(VT) == 8 ? "spmv_csr_f64_cg" : "spmv_csr_f32_cg");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "cg_update_f64" : "cg_update_f32");
  int it = 0;
  while (it < maxit) {
    const int nb = std::min(H, maxit - it);
    for (int s = 0; s < nb; ++s, ++it) {
      {
        KTimer tm(ctx, c_spmv, csr_bytes(A, sizeof(VT), sizeof(T)));
        spmv_dispatch<VT, T>(A, val, p, q, nullptr, A->space->d_free.p, (const T *)nullptr, nullptr, 2, w.part.p,
                             sgrid, st);
      }
      reduce_to(w.part.p, sgrid, 1, sc + 2, st);
      {
        KTimer tm(ctx, c_upd, 8.0 * n * sizeof(T));
        k_cg_update<T><<<vgrid, NTHR, 0, st>>>(n, x, r, z, p, q, dinv, sc, it, w.part.p);
      }
      // rz_new -> sc[(it+1)&1], rr -> hist[s]
      reduce_to(w.part.p, vgrid, 1, sc + ((it + 1) & 1), st);
      reduce_to(w.part.p + vgrid, vgrid, 1, hist.p + s, st);
      k_cg_p<T><<<nblk(n), NTHR, 0, st>>>(n, p, z, sc, it);
    }
    OCMP_HIP(hipMemcpyAsync(hh, hist.p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    OCMP_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < nb; ++s) {
      *last_rr = hh[s];
      if (!(hh[s] > tol_rr)) {  // also stops on NaN
        iters_done = it - nb + s + 1;
        return 1;
      }
    }
  }
  iters_done = it;
  return 0;
}

static void solve_cg(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                     ocmp_solve_stats &st_) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  Work &w = work(A);
  const int64_t n = A->n;
  OCMP_REQUIRE(P && P->kind == OCMP_PRE_JACOBI, "CG needs a Jacobi preconditioner (as NGSolve's CG needs one)");
  w.part.alloc(std::max<size_t>(w.part.n, (size_t)4 * MAX_RED_BLOCKS + 64 * MAX_RED_BLOCKS));
  w.r64.alloc(n);
  const double bn = norm_masked(A, b, w);
  if (bn == 0.0) {
    // b = 0 on free dofs: the solution is x_F = -A_FF^{-1} A_FD x_D; fall through with tol relative to ||A x0||
  }
  const double bref = bn > 0.0 ? bn : 1.0;
  double rn = residual64(A, b, x, w.r64.p, w);
  st_.initial_rel_residual = rn / bref;
  st_.rel_residual = rn / bref;
  const int check = o.check_every > 0 ? o.check_every : 10;
  if (rn <= o.tol * bref) {
    st_.converged = 1;
    st_.outer = 0;
    return;
  }
  const bool mixed = o.precision == OCMP_MIXED;
  int total = 0;
  if (!mixed) {
    w.vd.alloc((size_t)4 * n);
    double *z = w.vd.p, *p = z + n, *q = p + n;
    double *r = w.r64.p;
    // z = dinv r, p = z, rz
    k_jacobi_apply<double, double><<<nblk(n), NTHR, 0, st>>>(n, P->dinv.p, r, 1.0, z, p);
    const int vgrid = red_grid(n);
    k_dot<double><<<vgrid, NTHR, 0, st>>>(n, r, z, nullptr, w.part.p);
    reduce_to(w.part.p, vgrid, 1, ctx->scal.p + 0, st);
    int outer = 0;
    while (true) {
      ++outer;
      int done = 0;
      double last = 0.0;
      const int conv = cg_inner<double, double>(A, A->d_val.p, P->dinv.p, x, r, z, p, q,
                                                o.tol * o.tol * bref * bref, o.maxit - total, check, w, done, &last);
      total += done;
      rn = residual64(A, b, x, w.r64.p, w);  // true residual (also restarts r)
      st_.rel_residual = rn / bref;
      if (rn <= o.tol * bref || total >= o.maxit || !std::isfinite(rn)) break;
      (void)conv;
      // recurrence drifted: restart from the true residual
      k_jacobi_apply<double, double><<<nblk(n), NTHR, 0, st>>>(n, P->dinv.p, r, 1.0, z, p);
      k_dot<double><<<vgrid, NTHR, 0, st>>>(n, r, z, nullptr, w.part.p);
      reduce_to(w.part.p, vgrid, 1, ctx->scal.p + 0, st);
    }
    st_.outer = outer;
  } else {
    OCMP_REQUIRE(P->dinv32.p && A->d_val32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
    w.vf.alloc((size_t)5 * n);
    float *d = w.vf.p, *r = d + n, *z = r + n, *p = z + n, *q = p + n;
    const int vgrid = red_grid(n);
    const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-4;
    int outer = 0;
    while (true) {
      ++outer;
      // inner: A d = r/|r|, d0 = 0
      k_fill_t<float><<<nblk(n), NTHR, 0, st>>>(d, n, 0.0f);
      k_scale_copy<double, float><<<nblk(n), NTHR, 0, st>>>(n, w.r64.p, 1.0 / rn, r);
      k_jacobi_apply<float, float><<<nblk(n), NTHR, 0, st>>>(n, P->dinv32.p, r, 1.0, z, p);
      k_dot<float><<<vgrid, NTHR, 0, st>>>(n, r, z, nullptr, w.part.p);
      reduce_to(w.part.p, vgrid, 1, ctx->scal.p + 0, st);
      // ask the inner solve for the larger of inner_tol and what the outer still needs
      const double need = std::max(inner, 0.5 * o.tol * bref / rn);
      int done = 0;
      double last = 0.0;
      cg_inner<float, float>(A, A->d_val32.p, P->dinv32.p, d, r, z, p, q, need * need, o.maxit - total, check, w,
                             done, &last);
      total += done;
      k_axpy64<float><<<nblk(n), NTHR, 0, st>>>(n, rn, d, x);
      const double rn_old = rn;
      rn = residual64(A, b, x, w.r64.p, w);
      st_.rel_residual = rn / bref;
      if (rn <= o.tol * bref || total >= o.maxit || !std::isfinite(rn)) break;
      if (outer >= (o.max_outer > 0 ? o.max_outer : 50)) break;
      if (rn > 0.9 * rn_old && done == 0) break;  // stagnation
    }
    st_.outer = outer;
  }
  st_.iters = total;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
}

// ------------------------------------------------------------------ GMRES

// one restart cycle of right-preconditioned GMRES(m) in precision T on rhs r0 (FP64, masked),
// accumulating the update into x (FP64). Returns iterations used.
template <typename T>
static int gmres_cycle(Matrix *A, Precond *P, const double *r0, double beta, double *x, int m, double stop_abs,
                       int maxit, int check, Work &w) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  const bool dg = A->dg && P && P->kind == OCMP_PRE_BLOCK_JACOBI;
  const uint8_t *mask = A->dg ? nullptr : A->space->d_free.p;
  DBuf<T> *Vbuf;
  if constexpr (sizeof(T) == 8) {
    w.vd.alloc(std::max<size_t>(w.vd.n, (size_t)(m + 1) * n));
    Vbuf = &w.vd;
  } else {
    w.vf.alloc(std::max<size_t>(w.vf.n, (size_t)(m + 1) * n));
    Vbuf = &w.vf;
  }
  T *V = Vbuf->p;
  const int64_t ldv = n;
  // small device state: dot[m+2], vs[m+2], R[(m+1)*m], csn[2m], g[m+1], res[maxit], coef[m]
  const size_t ns = (size_t)(m + 2) * 2 + (size_t)(m + 1) * m + 2 * m + (m + 1) + 4096 + m;
  w.small.alloc(std::max(w.small.n, ns));
  double *dot = w.small.p, *vs = dot + (m + 2), *R = vs + (m + 2), *csn = R + (size_t)(m + 1) * m,
         *g = csn + 2 * m, *res = g + (m + 1), *coef = res + 4096;
  const int grid = red_grid(n);
  const int nchunk = (m + 1 + KC) / KC;
  w.part.alloc(std::max<size_t>(w.part.n, (size_t)(m + 2 + KC) * grid));
  // v0 = r0 (scale 1/beta, stored unnormalised)
  k_scale_copy<double, T><<<nblk(n), NTHR, 0, st>>>(n, r0, 1.0, V);
  k_gmres_init<<<1, 64, 0, st>>>(m, 1.0, g, vs, 1.0 / beta);
  double *hres = pinned(w, check + 8);
  const int c_op = ctx->cls(dg ? (sizeof(T) == 8 ? "dg_ahat_spmv_f64" : "dg_ahat_spmv_f32")
                               : (sizeof(T) == 8 ? "spmv_csr_f64_gmres" : "spmv_csr_f32_gmres"));
  const int c_dot = ctx->cls(sizeof(T) == 8 ? "gmres_mdot_f64" : "gmres_mdot_f32");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "gmres_update_f64" : "gmres_update_f32");
  int j = 0, k_used = 0;
  bool stop = false;
  const double gscale = beta;  // g is relative to beta
  while (!stop) {
    const int j_end = std::min(m, std::min(j + check, maxit));
    for (int jj = j; jj < j_end; ++jj) {
      T *vj = V + (int64_t)jj * ldv;
      T *wv = V + (int64_t)(jj + 1) * ldv;
      {
        const double opb = dg ? (double)n * 3 * A->dg_nb * sizeof(T) + 12.0 * A->space->mesh->nel + 2.0 * n * sizeof(T)
                              : csr_bytes(A, sizeof(T), sizeof(T));
        KTimer tm(ctx, c_op, opb);
        if (dg) {
          const T *ah;
          if constexpr (sizeof(T) == 8)
            ah = P->ahat.p;
          else
            ah = P->ahat32.p;
          const int nbk = A->dg_nb;
          const int64_t nel = A->space->mesh->nel;
          switch (nbk) {
            case 1: k_dg_ahat_spmv<T, 1><<<nblk(n), NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 3: k_dg_ahat_spmv<T, 3><<<nblk(n), NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 6: k_dg_ahat_spmv<T, 6><<<nblk(n), NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 10: k_dg_ahat_spmv<T, 10><<<nblk(n), NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            default: throw Error(OCMP_ERR_UNSUPPORTED, "DG block size");
          }
        } else {
          const T *cs = nullptr;
          const T *vals;
          if constexpr (sizeof(T) == 8) {
            vals = A->d_val.p;
            cs = (P && P->kind == OCMP_PRE_JACOBI) ? P->dinv.p : nullptr;
          } else {
            vals = A->d_val32.p;
            cs = (P && P->kind == OCMP_PRE_JACOBI) ? P->dinv32.p : nullptr;
          }
          spmv_dispatch<T, T>(A, vals, vj, wv, nullptr, mask, cs, vs + jj, 3, nullptr, spmv_grid(A), st);
        }
      }
      {
        KTimer tm(ctx, c_dot, (double)(jj + 2) * n * sizeof(T));
        dim3 gd(grid, (jj + 2 + KC - 1) / KC);
        k_mdot<T><<<gd, NTHR, 0, st>>>(n, V, ldv, wv, jj + 1, w.part.p);
      }
      reduce_to(w.part.p, grid, jj + 1, dot, st);
      {
        KTimer tm(ctx, c_upd, (double)(jj + 3) * n * sizeof(T));
        k_gs_update<T><<<grid, NTHR, 0, st>>>(n, V, ldv, wv, jj + 1, dot, vs, w.part.p);
      }
      reduce_to(w.part.p, grid, 1, dot + jj + 1, st);
      k_givens<<<1, 64, 0, st>>>(jj, m, dot, vs, R, csn, g, res, jj, gscale);
    }
    (void)nchunk;
    const int nb = j_end - j;
    OCMP_HIP(hipMemcpyAsync(hres, res + j, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    OCMP_HIP(hipStreamSynchronize(st));
    int hit = -1;
    for (int s = 0; s < nb; ++s)
      if (!(hres[s] > stop_abs)) {
        hit = s;
        break;
      }
    if (hit >= 0) {
      k_used = j + hit + 1;
      stop = true;
    } else {
      j = j_end;
      k_used = j;
      if (j >= m || j >= maxit) stop = true;
    }
  }
  // x += M^{-1} sum_l y_l v_l
  k_gmres_solve<<<1, 64, 0, st>>>(k_used, m, R, g, vs, coef);
  w.u64.alloc(n);
  k_combine<T><<<nblk(n), NTHR, 0, st>>>(n, V, ldv, k_used, coef, beta, w.u64.p);
  if (dg) {
    const int64_t nel = A->space->mesh->nel;
    switch (A->dg_nb) {
      case 1: k_dg_block_apply_add<1><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 3: k_dg_block_apply_add<3><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 6: k_dg_block_apply_add<6><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 10: k_dg_block_apply_add<10><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
    }
  } else {
    const double *dinv = (P && P->kind == OCMP_PRE_JACOBI) ? P->dinv.p : nullptr;
    k_diag_apply_add<<<nblk(n), NTHR, 0, st>>>(n, dinv, w.u64.p, x);
  }
  OCMP_HIP(hipGetLastError());
  return k_used;
}

static void solve_gmres(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                        ocmp_solve_stats &st_) {
  Work &w = work(A);
  const int64_t n = A->n;
  w.part.alloc(std::max<size_t>(w.part.n, (size_t)4 * MAX_RED_BLOCKS + 64 * MAX_RED_BLOCKS));
  w.r64.alloc(n);
  const bool mixed = o.precision == OCMP_MIXED;
  if (mixed && P && P->kind == OCMP_PRE_JACOBI)
    OCMP_REQUIRE(P->dinv32.p && A->d_val32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
  if (mixed && P && P->kind == OCMP_PRE_BLOCK_JACOBI)
    OCMP_REQUIRE(P->ahat32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
  if (!mixed && P && P->kind == OCMP_PRE_BLOCK_JACOBI)
    OCMP_REQUIRE(P->ahat.p, "FP64 GMRES needs a preconditioner built with OCMP_FP64");
  if (mixed && !P) {
    // plain GMRES in FP32 needs an FP32 copy of A
    A->d_val32.alloc(A->nnz);
    k_to_f32<<<nblk(A->nnz), NTHR, 0, A->ctx->stream>>>(A->nnz, A->d_val.p, A->d_val32.p);
  }
  const double bn = norm_masked(A, b, w);
  const double bref = bn > 0.0 ? bn : 1.0;
  double rn = residual64(A, b, x, w.r64.p, w);
  st_.initial_rel_residual = rn / bref;
  st_.rel_residual = rn / bref;
  const int check = o.check_every > 0 ? o.check_every : 5;
  const int m = o.restart;
  int total = 0, outer = 0;
  const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-5;
  while (rn > o.tol * bref && total < o.maxit && std::isfinite(rn)) {
    ++outer;
    const double stop_abs = mixed ? std::max(o.tol * bref, inner * rn) : o.tol * bref;
    int used;
    if (mixed)
      used = gmres_cycle<float>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w);
    else
      used = gmres_cycle<double>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w);
    total += used;
    const double rn_old = rn;
    rn = residual64(A, b, x, w.r64.p, w);
    st_.rel_residual = rn / bref;
    if (used == 0 || (mixed && o.max_outer > 0 && outer >= o.max_outer)) break;
    if (!(rn < rn_old) && used < m) break;  // no progress and the cycle thought it had converged
  }
  st_.iters = total;
  st_.outer = outer;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
}

static void solve_richardson(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                             ocmp_solve_stats &st_) {
  // x += M^{-1} (b - A x), M = Jacobi or block Jacobi
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  Work &w = work(A);
  const int64_t n = A->n;
  OCMP_REQUIRE(P && (P->kind == OCMP_PRE_JACOBI || P->kind == OCMP_PRE_BLOCK_JACOBI),
               "Richardson needs a preconditioner");
  w.part.alloc(std::max<size_t>(w.part.n, (size_t)4 * MAX_RED_BLOCKS + 64 * MAX_RED_BLOCKS));
  w.r64.alloc(n);
  const double bn = norm_masked(A, b, w);
  const double bref = bn > 0.0 ? bn : 1.0;
  double rn = residual64(A, b, x, w.r64.p, w);
  st_.initial_rel_residual = rn / bref;
  int it = 0;
  while (rn > o.tol * bref && it < o.maxit && std::isfinite(rn)) {
    if (P->kind == OCMP_PRE_JACOBI) {
      k_diag_apply_add<<<nblk(n), NTHR, 0, st>>>(n, P->dinv.p, w.r64.p, x);
    } else {
      const int64_t nel = A->space->mesh->nel;
      switch (A->dg_nb) {
        case 1: k_dg_block_apply_add<1><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.r64.p, x); break;
        case 3: k_dg_block_apply_add<3><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.r64.p, x); break;
        case 6: k_dg_block_apply_add<6><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.r64.p, x); break;
        case 10: k_dg_block_apply_add<10><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.r64.p, x); break;
      }
    }
    ++it;
    rn = residual64(A, b, x, w.r64.p, w);
  }
  st_.iters = it;
  st_.outer = 1;
  st_.rel_residual = rn / bref;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
}

void solve(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o, ocmp_solve_stats &st) {
  switch (o.solver) {
    case OCMP_SOLVER_CG: solve_cg(A, P, b, x, o, st); break;
    case OCMP_SOLVER_GMRES: solve_gmres(A, P, b, x, o, st); break;
    case OCMP_SOLVER_RICHARDSON: solve_richardson(A, P, b, x, o, st); break;
    default: throw Error(OCMP_ERR_UNSUPPORTED, "unknown solver");
  }
}

}  // namespace ocmp
