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
// Streaming kernels move 16 bytes per lane per load (float4 / double2).
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>

#include "internal.h"

#include <type_traits>

namespace ocmp {

constexpr int NTHR = 256;
constexpr int NXCD = 8;
constexpr int MAX_RED_BLOCKS = 2048;
constexpr int STREAM_BLOCKS = 1024;  // grid of the streaming reduction kernels (4 per CU)
constexpr int SCHUR_SCAL = 8;        // scalars at the head of the Schur inner solve's partial-sum buffer
// rows of the coarsest AMG level inverted densely at most (the inversion keeps one factor per row in
// LDS; the host stops coarsening at 48 rows and accepts up to 1200 when coarsening stalls)
constexpr int AMG_DENSE_MAX = 1200;
// a last level of at most this many rows is inverted in LDS (n x 2n doubles: 36 KB)
constexpr int AMG_LDS_MAX = 48;
constexpr int64_t AMG_TAIL_ROWS = 1200;  // levels (below the top) at or below this size end the V-cycle as one dense matrix
constexpr int64_t AMG_TAIL_LD = 1216;    // the largest leading dimension of that matrix (rows padded to 64)
constexpr double AMG_SMOOTH = 2.0 / 3.0;  // damping of the AMG levels' Jacobi smoothing

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

// 16-byte vector access
template <typename T>
struct Vec;
template <>
struct Vec<float> {
  using t = float4;
  static constexpr int w = 4;
};
template <>
struct Vec<double> {
  using t = double2;
  static constexpr int w = 2;
};

template <typename T>
__device__ __forceinline__ void vload(const T *__restrict__ p, T (&o)[Vec<T>::w]) {
  const typename Vec<T>::t v = *reinterpret_cast<const typename Vec<T>::t *>(p);
  if constexpr (Vec<T>::w == 4) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  } else {
    o[0] = v.x;
    o[1] = v.y;
  }
}

template <typename T>
__device__ __forceinline__ void vstore(T *__restrict__ p, const T (&o)[Vec<T>::w]) {
  typename Vec<T>::t v;
  if constexpr (Vec<T>::w == 4) {
    v.x = o[0];
    v.y = o[1];
    v.z = o[2];
    v.w = o[3];
  } else {
    v.x = o[0];
    v.y = o[1];
  }
  *reinterpret_cast<typename Vec<T>::t *>(p) = v;
}

static inline int red_grid(int64_t work_items) {
  int64_t g = (work_items + NTHR - 1) / NTHR;
  g = std::min<int64_t>(g, MAX_RED_BLOCKS);
  g = std::max<int64_t>(g, NXCD);
  g = (g + NXCD - 1) / NXCD * NXCD;
  return (int)g;
}

static inline int stream_grid(int64_t nvec) {
  int64_t g = (nvec + NTHR - 1) / NTHR;
  g = std::min<int64_t>(g, STREAM_BLOCKS);
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

// LDS-staged CSR SpMV. Block b owns rows [rblk[b], rblk[b+1]) whose nonzeros form one
// contiguous range of at most NNZB entries. The block first gathers x at its distinct
// columns (ucol, sorted: neighbouring lanes read neighbouring entries) into LDS; every thread
// then loads NNZB/256 (16-bit local column, value) pairs with fully coalesced loads and stages the
// products in LDS; rows are summed from LDS by 4 lanes each. The rows of a finite-element block
// share most of their columns, so the x gathers (the texture-unit bound part of a plain CSR
// SpMV) drop several-fold, and the 16-bit local columns save 2 bytes per nonzero.
//
// mode 0: y = A x
// mode 1: y = mask*(b - A x), partial sum y^2 (y == nullptr: the partial sum only)
// mode 2: y = mask*(A x), partial sum x.y                (CG: q = A p, p.q)
// mode 3: y = mask*(A (cs*sx*x))   cs: column scaling (right Jacobi), sx: device scalar  (GMRES on H1)
// mode 4: y = mask*(A (cs*x)) - dsh*x + csh*xp                  (a Newton-basis step of s-step GMRES)
// mode 5: y = xp - A x                                           (multigrid residual, b in xp)
// mode 6: y = xp + A x                                           (multigrid prolongation-add)

template <typename VT, typename XT>
__device__ __forceinline__ void spmv_epilogue(int mode, int64_t r, double s, XT *__restrict__ y,
                                              const XT *__restrict__ x, const double *__restrict__ b,
                                              const uint8_t *__restrict__ mask, XT scale, double &acc,
                                              XT dsh, const XT *__restrict__ xp, XT csh) {
  if (mode == 0) {
    y[r] = (XT)s;
  } else if (mode == 1) {
    const double v = (mask && !mask[r]) ? 0.0 : b[r] - s;
    if (y) y[r] = (XT)v;
    acc += v * v;
  } else if (mode == 2) {
    const XT v = (mask && !mask[r]) ? (XT)0 : (XT)s;
    y[r] = v;
    acc += (double)x[r] * (double)v;
  } else if (mode == 3) {
    y[r] = (mask && !mask[r]) ? (XT)0 : (XT)s * scale;
  } else if (mode == 5) {
    y[r] = xp[r] - (XT)s;
  } else if (mode == 6) {
    y[r] = xp[r] + (XT)s;
  } else {
    XT v = (XT)s - dsh * x[r];
    if (xp) v += csh * xp[r];
    y[r] = (mask && !mask[r]) ? (XT)0 : v;
  }
}

// One row block of a CSR product: the x values at the block's distinct columns (xl(c) loads the
// value at column c, xs(c, v) then scales it) are staged in LDS, each thread forms its share of the products, and 4 lanes
// per row sum them in a fixed order; ef(r, s) receives each row's sum. The LDS buffer holds first
// the staged x values, then (after every thread has read its operands) the products, so the
// kernel's LDS is exactly NNZB values (a power of two: whole blocks per CU).
template <typename VT, typename XT, int NNZB, typename XL, typename XS, typename EF>
__device__ __forceinline__ void csr_row_block(int lb, const int32_t *__restrict__ rblk, const int32_t *__restrict__ rp,
                                              const int32_t *__restrict__ col, const int32_t *__restrict__ ubase,
                                              const int32_t *__restrict__ ucol, const uint16_t *__restrict__ lcol,
                                              const VT *__restrict__ val, double *lds, XL xl, XS xsc, EF ef) {
  constexpr int SPMV_PER_THREAD = NNZB / NTHR;
  XT *xs = reinterpret_cast<XT *>(lds);
  XT *prod = xs;
  const int32_t r0 = rblk[lb], r1 = rblk[lb + 1];
  const int32_t j0 = rp[r0], j1 = rp[r1];
  const int nz = j1 - j0;
  const int nrows = r1 - r0;
  if (nz <= NNZB) {
    // rows of the first pass: fetch their bounds early so the loads overlap the staging
    const int ri0 = threadIdx.x >> 2, q = threadIdx.x & 3;
    int a0 = 0, e0 = 0;
    if (ri0 < nrows) {
      a0 = rp[r0 + ri0] - j0;
      e0 = rp[r0 + ri0 + 1] - j0;
    }
    int li[SPMV_PER_THREAD];
    VT vv[SPMV_PER_THREAD];
#pragma unroll
    for (int k = 0; k < SPMV_PER_THREAD; ++k) {
      const int t = threadIdx.x + k * NTHR;
      li[k] = -1;
      vv[k] = (VT)0;
      if (t < nz) {
        li[k] = lcol[j0 + t];
        vv[k] = val[j0 + t];
      }
    }
    // x at the block's distinct columns (nu >= 1 since nz >= 1; indices past nu are clamped
    // so both gathers of a round are unconditional)
    const int u0 = ubase[lb], nu = ubase[lb + 1] - u0;
    for (int base = 0; base < nu; base += 2 * NTHR) {
      const int t0 = base + threadIdx.x, t1 = t0 + NTHR;
      const int32_t c0 = ucol[u0 + min(t0, nu - 1)], c1 = ucol[u0 + min(t1, nu - 1)];
      XT x0 = xl(c0), x1 = xl(c1);
      x0 = xsc(c0, x0);
      x1 = xsc(c1, x1);
      if (t0 < nu) xs[t0] = x0;
      if (t1 < nu) xs[t1] = x1;
    }
    __syncthreads();
    XT pv[SPMV_PER_THREAD];
#pragma unroll
    for (int k = 0; k < SPMV_PER_THREAD; ++k) pv[k] = li[k] >= 0 ? (XT)vv[k] * xs[li[k]] : (XT)0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SPMV_PER_THREAD; ++k)
      if (li[k] >= 0) prod[threadIdx.x + k * NTHR] = pv[k];
    __syncthreads();
    for (int base = 0; base < nrows; base += NTHR / 4) {
      const int ri = base + ri0;
      int a = a0, e = e0;
      if (base > 0 && ri < nrows) {
        a = rp[r0 + ri] - j0;
        e = rp[r0 + ri + 1] - j0;
      }
      // partial row sums in the vector type (FP32 products summed in FP32, FP64 in FP64)
      XT s = (XT)0;
      if (ri < nrows)
        for (int t = a + q; t < e; t += 4) s += prod[t];
      s += __shfl_xor(s, 1, 4);
      s += __shfl_xor(s, 2, 4);
      if (ri < nrows && q == 0) ef((int64_t)(r0 + ri), (double)s);
    }
  } else {
    // a single row longer than the staging buffer
    double s = 0.0;
    for (int t = threadIdx.x; t < nz; t += NTHR) {
      const int32_t c = col[j0 + t];
      s += (double)((XT)val[j0 + t] * xsc(c, xl(c)));
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) ef((int64_t)r0, s);
  }
}

template <typename VT, typename XT, int NNZB>
__global__ void __launch_bounds__(NTHR) k_spmv_csr(int64_t nrblk, const int32_t *__restrict__ rblk,
                                                   const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                   const int32_t *__restrict__ ubase,
                                                   const int32_t *__restrict__ ucol,
                                                   const uint16_t *__restrict__ lcol,
                                                   const VT *__restrict__ val, const XT *__restrict__ x,
                                                   XT *__restrict__ y, const double *__restrict__ b,
                                                   const uint8_t *__restrict__ mask, const XT *__restrict__ cs,
                                                   const double *__restrict__ sx, int mode,
                                                   double *__restrict__ part, XT dsh, const XT *__restrict__ xp,
                                                   XT csh, const int *__restrict__ skip,
                                                   const int32_t *__restrict__ rowmap) {
  if (skip && *skip) return;  // a converged inner solve: the rest of its fixed launch sequence is empty
  __shared__ double lds[NNZB * sizeof(XT) / sizeof(double)];
  const int lb = xcd_block();
  double acc = 0.0;
  const XT scale = (mode == 3 && sx) ? (XT)sx[0] : (XT)1;
  if (lb < nrblk)
    csr_row_block<VT, XT, NNZB>(
        lb, rblk, rp, col, ubase, ucol, lcol, val, lds,
        [&](int32_t c) { return x[c]; },
        [&](int32_t c, XT v) { return (mode >= 3 && cs) ? v * cs[c] : v; },
        [&](int64_t r, double s) {
          // a reordered operator's row r is row rowmap[r] of the vectors
          spmv_epilogue<VT, XT>(mode, rowmap ? (int64_t)rowmap[r] : r, s, y, x, b, mask, scale, acc, dsh, xp, csh);
        });
  if (part) {
    __syncthreads();  // the block-sum scratch aliases the products
    acc = block_sum(acc, lds);
    if (threadIdx.x == 0) part[lb] = acc;
  }
}

// FP32 values and vectors use the larger row blocks (same 16 KB of LDS)
template <typename VT, typename XT>
static const RowBlocks &row_blocks(const Matrix *A) {
  return (sizeof(VT) == 4 && sizeof(XT) == 4 && A->rb32.nnzb) ? A->rb32 : A->rb64;
}

template <typename VT, typename XT>
static int spmv_grid(Matrix *A) {
  int64_t g = std::max<int64_t>(row_blocks<VT, XT>(A).nrblk, NXCD);
  g = (g + NXCD - 1) / NXCD * NXCD;
  return (int)g;
}
static int spmv_grid(Matrix *A) { return spmv_grid<double, double>(A); }

template <typename VT, typename XT>
static void spmv_dispatch(Matrix *A, const VT *val, const XT *x, XT *y, const double *b, const uint8_t *mask,
                          const XT *cs, const double *sx, int mode, double *part, int grid, hipStream_t st,
                          XT dsh = (XT)0, const XT *xp = nullptr, XT csh = (XT)0, const int *skip = nullptr) {
  const RowBlocks &rb = row_blocks<VT, XT>(A);
  OCMP_REQUIRE((grid == spmv_grid<VT, XT>(A)), "SpMV grid does not match the row blocks");
  if (rb.nnzb == SPMV_NNZB32)
    k_spmv_csr<VT, XT, SPMV_NNZB32><<<grid, NTHR, 0, st>>>(rb.nrblk, rb.d_rblk.p, A->d_row_ptr.p, A->d_col.p,
                                                          rb.d_ubase.p, rb.d_ucol.p, rb.d_lcol.p, val, x, y, b,
                                                          mask, cs, sx, mode, part, dsh, xp, csh, skip,
                                                          A->d_rowmap.p);
  else
    k_spmv_csr<VT, XT, SPMV_NNZB64><<<grid, NTHR, 0, st>>>(rb.nrblk, rb.d_rblk.p, A->d_row_ptr.p, A->d_col.p,
                                                          rb.d_ubase.p, rb.d_ucol.p, rb.d_lcol.p, val, x, y, b,
                                                          mask, cs, sx, mode, part, dsh, xp, csh, skip,
                                                          A->d_rowmap.p);
}

// algorithmic bytes of one CSR SpMV: values + 16-bit local columns + the blocks' distinct
// columns + row pointers + x once + y once
static double csr_bytes(const Matrix *A, int wv, int wx) {
  const RowBlocks &rb = (wv == 4 && wx == 4 && A->rb32.nnzb) ? A->rb32 : A->rb64;
  return (double)A->nnz * (wv + 2) + 4.0 * rb.nucol + 4.0 * (A->n + 1) + 2.0 * A->n * wx +
         (A->d_rowmap.p ? 4.0 * A->n : 0.0);
}

void spmv_f64(Matrix *A, const double *x, double *y) {
  Ctx *ctx = A->ctx;
  KTimer tm(ctx, ctx->cls("spmv_csr_f64"), csr_bytes(A, 8, 8));
  spmv_dispatch<double, double>(A, A->d_val.p, x, y, nullptr, nullptr, nullptr, nullptr, 0, nullptr, spmv_grid(A),
                                ctx->stream);
  OCMP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ DG operators (ELL layout)
//
// Ahat = A * blockdiag(A)^{-1}. Its diagonal blocks are the identity, so only
// the three facet-neighbour blocks are stored, column-major "ELL":
// ahat[(k*nb + j) * n + r] = Ahat(r, nbr_k*nb + j), r = e*nb + i, k = local facet.
// Boundary facets store zero blocks and point at the element itself.
// nbr[k*nel + e] is the neighbour element across facet k. A itself is kept in
// the same layout with four slots (slot 0 = diagonal block) for the FP64
// outer residual. Each lane handles VW consecutive rows so every matrix load
// moves 16 bytes per lane.

template <typename T, int NB>
__global__ void __launch_bounds__(NTHR) k_dg_ahat_spmv(int64_t nel, const int32_t *__restrict__ nbr,
                                                       const T *__restrict__ ahat, const T *__restrict__ x,
                                                       const double *__restrict__ sx, T *__restrict__ y,
                                                       T dcoef = (T)1, const T *__restrict__ xp = nullptr,
                                                       T pc = (T)0) {
  // y = sx * ((Ahat - (1 - dcoef) I) x + pc xp): dcoef = 1 is Ahat itself, dcoef = 0 its
  // off-diagonal part; xp (may be null) is the previous basis vector of a Newton-basis step
  constexpr int VW = Vec<T>::w;
  const int64_t n = nel * NB;
  const int64_t r0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * VW;
  if (r0 >= n) return;  // n % VW == 0 is checked by the host
  const T scale = sx ? (T)sx[0] : (T)1;
  T s[VW];
  vload(x + r0, s);
#pragma unroll
  for (int l = 0; l < VW; ++l) s[l] *= dcoef;
  if (xp) {
    T pv[VW];
    vload(xp + r0, pv);
#pragma unroll
    for (int l = 0; l < VW; ++l) s[l] += pc * pv[l];
  }
  int64_t cb[VW][3];
#pragma unroll
  for (int l = 0; l < VW; ++l) {
    const int64_t e = (r0 + l) / NB;
#pragma unroll
    for (int k = 0; k < 3; ++k) cb[l][k] = (int64_t)nbr[k * nel + e] * NB;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      T a[VW];
      vload(ahat + (int64_t)(k * NB + j) * n + r0, a);
#pragma unroll
      for (int l = 0; l < VW; ++l) s[l] += a[l] * x[cb[l][k] + j];
    }
#pragma unroll
  for (int l = 0; l < VW; ++l) s[l] *= scale;
  vstore(y + r0, s);
}

// r = b - A x with A in the 4-slot ELL layout; partial sums of r^2 (two rows per lane)
template <int NB>
__global__ void __launch_bounds__(NTHR) k_dg_ell_residual(int64_t nel, const int32_t *__restrict__ nbr,
                                                          const double *__restrict__ aell,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ b, double *__restrict__ r,
                                                          double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int64_t n = nel * NB;
  const int64_t nvec = n / 2;
  const int lb = xcd_block();
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  double acc = 0.0;
  for (int64_t v = i0 + threadIdx.x; v < i1; v += blockDim.x) {
    const int64_t r0 = 2 * v;
    double s[2];
    vload(b + r0, s);
    int64_t cb[2][4];
#pragma unroll
    for (int l = 0; l < 2; ++l) {
      const int64_t e = (r0 + l) / NB;
      cb[l][0] = e * NB;
#pragma unroll
      for (int k = 0; k < 3; ++k) cb[l][k + 1] = (int64_t)nbr[k * nel + e] * NB;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double a[2];
        vload(aell + (int64_t)(k * NB + j) * n + r0, a);
        s[0] -= a[0] * x[cb[0][k] + j];
        s[1] -= a[1] * x[cb[1][k] + j];
      }
    vstore(r + r0, s);
    acc += s[0] * s[0] + s[1] * s[1];
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[lb] = acc;
}

// invert the diagonal blocks (Gauss-Jordan, partial pivoting), one thread per element
template <int NB>
__global__ void k_dg_block_inverse(int64_t nel, const int32_t *__restrict__ rp, const int8_t *__restrict__ dg_pos,
                                   const double *__restrict__ val, double *__restrict__ binv) {
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
    for (int j = 0; j < NB; ++j) o[i * NB + j] = v[i][j];
}

// ahat(r, block k) = A(r, block nbr_k) * inv(D_nbr_k); aell = A in 4-slot ELL; one thread per row
template <int NB>
__global__ void k_dg_build_ell(int64_t nel, const int32_t *__restrict__ rp, const int8_t *__restrict__ dg_pos,
                               const int32_t *__restrict__ elem_nb, const double *__restrict__ val,
                               const double *__restrict__ binv, double *__restrict__ ahat,
                               float *__restrict__ ahat32, double *__restrict__ aell, int32_t *__restrict__ nbr) {
  const int64_t n = nel * NB;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t e = r / NB;
  const int i = (int)(r % NB);
  {
    const double *drow = val + rp[r] + dg_pos[e * 4] * NB;
#pragma unroll
    for (int j = 0; j < NB; ++j) aell[(int64_t)j * n + r] = drow[j];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int32_t nb = elem_nb[e * 3 + k];
    const int pos = dg_pos[e * 4 + 1 + k];
    double out[NB], raw[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) out[j] = raw[j] = 0.0;
    if (nb >= 0 && pos >= 0) {
      const double *arow = val + rp[r] + pos * NB;
      const double *bi = binv + (int64_t)nb * NB * NB;
#pragma unroll
      for (int l = 0; l < NB; ++l) {
        const double al = arow[l];
        raw[l] = al;
#pragma unroll
        for (int j = 0; j < NB; ++j) out[j] += al * bi[l * NB + j];
      }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (ahat) ahat[(int64_t)(k * NB + j) * n + r] = out[j];
      if (ahat32) ahat32[(int64_t)(k * NB + j) * n + r] = (float)out[j];
      aell[(int64_t)((k + 1) * NB + j) * n + r] = raw[j];
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

__global__ void k_jacobi_diag(int64_t n, const int32_t *__restrict__ dslot, const double *__restrict__ val,
                              const uint8_t *__restrict__ mask, double *__restrict__ dinv,
                              float *__restrict__ dinv32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int32_t j = dslot[r];
  const double d = j >= 0 ? val[j] : 0.0;
  const double v = (mask && !mask[r]) || d == 0.0 ? 0.0 : 1.0 / d;
  dinv[r] = v;
  if (dinv32) dinv32[r] = (float)v;
}

// Schur rows r (flag[r] = 1): dinv[r] = -1 / (-a_rr + sum_{j not Schur, free} a_rj^2 / a_jj),
// the diagonal of -(C + B diag(A)^{-1} B^T) for the system [[A, B^T], [B, -C]]. Reads
// dinv only on non-Schur columns, which k_jacobi_diag has filled.
// Schur-diagonal rows (flag set): s = -(C_ii) + sum_j B_ij^2 / A_jj over the velocity columns j.
// Pass 1 stores s in dinv and per-block maxima of s in part; k_schur_finish turns s into -1/s, or 0
// where s is not above SCHUR_REL times the largest s: pressure dofs that the phi weighting has all
// but removed (inside solid obstacles) stay out of the Krylov update instead of getting a huge scale.
constexpr double SCHUR_REL = 1e-10;
__global__ void __launch_bounds__(NTHR) k_schur_diag(int64_t n, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ col, const double *__restrict__ val,
                                                     const uint8_t *__restrict__ mask,
                                                     const uint8_t *__restrict__ flag, double *__restrict__ dinv,
                                                     double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double s = 0.0;
  if (r < n && flag[r]) {
    for (int32_t j = rp[r]; j < rp[r + 1]; ++j) {
      const int32_t c = col[j];
      const double a = val[j];
      if (c == r)
        s -= a;
      else if (!flag[c] && mask[c])
        s += a * a * dinv[c];
    }
    if (!mask[r] || !(s > 0.0)) s = 0.0;
    dinv[r] = s;
  }
  double m = s;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = sh[0];
    for (int w = 1; w < NTHR / 64; ++w) bm = fmax(bm, sh[w]);
    part[blockIdx.x] = bm;
  }
}

__global__ void __launch_bounds__(NTHR) k_max_reduce(const double *__restrict__ part, int nparts,
                                                     double *__restrict__ out) {
  __shared__ double sh[NTHR / 64];
  double m = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = fmax(m, part[b]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = sh[0];
    for (int w = 1; w < NTHR / 64; ++w) bm = fmax(bm, sh[w]);
    out[0] = bm;
  }
}

__global__ void k_schur_finish(int64_t n, const uint8_t *__restrict__ flag, const double *__restrict__ smax,
                               double *__restrict__ dinv, float *__restrict__ dinv32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !flag[r]) return;
  const double s = dinv[r];
  const double v = s > SCHUR_REL * smax[0] ? -1.0 / s : 0.0;
  dinv[r] = v;
  if (dinv32) dinv32[r] = (float)v;
}

// Velocity scale of S: dpos[u] = 1 / max(|a_uu|, 1e-2 max_j |a_uj|) on the free non-Schur rows, 0
// elsewhere. Positive by construction, so S stays symmetric positive semidefinite where the
// convection term has made a velocity diagonal small or negative (the P2 vertex functions' mass
// is small, and the discrete wind is only weakly divergence-free).
// DPOS_LANES lanes per row, each over every DPOS_LANES-th slot, then a max across the group
constexpr int DPOS_LANES = 16;
__global__ void __launch_bounds__(NTHR) k_schur_dpos(int64_t n, int64_t s0, int64_t s1,
                                                     const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ col,
                                                     const int32_t *__restrict__ dslot,
                                                     const double *__restrict__ val,
                                                     const uint8_t *__restrict__ mask, double *__restrict__ dpos) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / DPOS_LANES;
  const int q = threadIdx.x % DPOS_LANES;
  const bool on = r < n && mask[r] && (r < s0 || r >= s1);
  double mx = 0.0;
  if (on)
    for (int32_t j = rp[r] + q; j < rp[r + 1]; j += DPOS_LANES) {
      const int32_t c = col[j];
      if (c < s0 || c >= s1) mx = fmax(mx, fabs(val[j]));
    }
#pragma unroll
  for (int o = DPOS_LANES / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, DPOS_LANES));
  if (r < n && q == 0) {
    double v = 0.0;
    if (on) {
      const int32_t jd = dslot[r];
      const double d = jd >= 0 ? fabs(val[jd]) : 0.0;
      const double dd = fmax(d, 1e-2 * mx);
      v = dd > 0.0 ? 1.0 / dd : 0.0;
    }
    dpos[r] = v;
  }
}

// Values of S = C + B D^{-1} B^T on the Schur rows [s0, s1) of [[A, B^T], [B, -C]], D the
// positive velocity scale above: S_pq = -a_pq + sum_u a_pu a_uq dpos[u] over the non-Schur
// columns u of row p (dpos is 0 on Dirichlet dofs). One wavefront per row of S and one lane per entry: the
// lanes walk row p's columns u together and each looks its own column q up in the Schur part of
// row u (a few slots, read by every lane at once), so each entry is summed in one fixed order.
__global__ void __launch_bounds__(NTHR) k_schur_values(int64_t ns, int64_t s0, const int32_t *__restrict__ srp,
                                                       const int32_t *__restrict__ scol,
                                                       const int32_t *__restrict__ rp,
                                                       const int32_t *__restrict__ col,
                                                       const double *__restrict__ val,
                                                       const int32_t *__restrict__ sseg,
                                                       const double *__restrict__ dpos, double *__restrict__ sval) {
  const int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= ns) return;
  const int64_t p = s0 + row;
  const int32_t pb = sseg[2 * p], pe = sseg[2 * p + 1];
  const int32_t k1 = srp[row + 1];
  for (int32_t k0 = srp[row]; k0 < k1; k0 += 64) {
    const int32_t k = k0 + lane;
    const int32_t q = k < k1 ? (int32_t)(s0 + scol[k]) : -1;
    double acc = 0.0;
    for (int32_t j = pb; j < pe; ++j)
      if (col[j] == q) acc -= val[j];
    for (int pass = 0; pass < 2; ++pass) {
      const int32_t a = pass ? pe : rp[p], e = pass ? rp[p + 1] : pb;
      for (int32_t j = a; j < e; ++j) {
        const int32_t u = col[j];
        const double f = val[j] * dpos[u];
        if (f == 0.0) continue;
        const int32_t t1 = sseg[2 * u + 1];
        for (int32_t t = sseg[2 * u]; t < t1; ++t)
          if (col[t] == q) acc += f * val[t];
      }
    }
    if (k < k1) sval[k] = acc;
  }
}

// the block-triangular velocity rows: z_u[r] -= dinv[r] sum_k A_up[r, k] z_p[k], one thread per row
// (a velocity row couples to a handful of Schur columns, contiguous in A_up's row)
template <typename T>
__global__ void __launch_bounds__(NTHR) k_tri_rows(int64_t nu, const int32_t *__restrict__ rp,
                                                   const int32_t *__restrict__ col, const T *__restrict__ val,
                                                   const T *__restrict__ dinv, const T *__restrict__ zp,
                                                   T *__restrict__ zu, const int *__restrict__ skip) {
  if (skip && *skip) return;  // a converged FGMRES cycle: its remaining launches are empty
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nu) return;
  const T d = dinv[r];
  if (d == (T)0) return;  // Dirichlet rows (and rows left out) keep z_u = 0
  double s = 0.0;
  for (int32_t k = rp[r]; k < rp[r + 1]; ++k) s += (double)val[k] * (double)zp[col[k]];
  zu[r] = (T)((double)zu[r] - (double)d * s);
}

// dinv (and its FP32 copy) = 0 on the flagged rows
__global__ void k_zero_flagged(int64_t n, const uint8_t *__restrict__ flag, double *__restrict__ dinv,
                               float *__restrict__ dinv32) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || !flag[r]) return;
  dinv[r] = 0.0;
  if (dinv32) dinv32[r] = 0.0f;
}

// f[k] = a_pu dpos[u] for the slots base + k of the Schur rows (each product's first factor once)
__global__ void k_schur_bfac(int64_t nk, int64_t base, const int32_t *__restrict__ col,
                             const double *__restrict__ val, const double *__restrict__ dpos, double *__restrict__ f) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nk) f[k] = val[base + k] * dpos[col[base + k]];
}

// lanes per output of the pair-list kernels: each lane sums every PAIR_LANES-th pair of its
// output, then a fixed shuffle tree adds the lanes (a fixed order: bitwise repeatable)
constexpr int PAIR_LANES = 4;

// The same values from the product lists built with S's pattern: PAIR_LANES lanes per entry of S,
// -a_pq plus the products of its list (bf from k_schur_bfac: a_pu dpos[u] per slot)
__global__ void __launch_bounds__(NTHR) k_schur_values_pairs(int64_t snnz, const int32_t *__restrict__ cslot,
                                                             const int32_t *__restrict__ pptr,
                                                             const int32_t *__restrict__ pj,
                                                             const int32_t *__restrict__ pt, int64_t base,
                                                             const double *__restrict__ bf,
                                                             const double *__restrict__ val,
                                                             double *__restrict__ sval) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t k = t / PAIR_LANES;
  const int q = (int)(t % PAIR_LANES);
  double acc = 0.0;
  if (k < snnz) {
    const int32_t m1 = pptr[k + 1];
    for (int32_t m = pptr[k] + q; m < m1; m += PAIR_LANES) acc += bf[pj[m] - base] * val[pt[m]];
  }
#pragma unroll
  for (int o = 1; o < PAIR_LANES; o <<= 1) acc += __shfl_xor(acc, o, PAIR_LANES);
  if (k < snnz && q == 0) {
    const int32_t c = cslot[k];
    sval[k] = (c >= 0 ? -val[c] : 0.0) + acc;
  }
}

// Jacobi of S, in two passes like the Schur diagonal: S_ii of the free rows and per-block maxima,
// then 1/S_ii, or 0 where S_ii is not above SCHUR_REL times the largest (rows the phi weighting
// has all but removed stay out of the inner solve)
__global__ void __launch_bounds__(NTHR) k_schur_sdiag(int64_t ns, const int32_t *__restrict__ sdslot,
                                                      const double *__restrict__ sval,
                                                      const uint8_t *__restrict__ mask_s, double *__restrict__ sd,
                                                      double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double d = 0.0;
  if (i < ns) {
    const int32_t k = sdslot[i];
    d = k >= 0 ? sval[k] : 0.0;
    if (!mask_s[i] || !(d > 0.0)) d = 0.0;
    sd[i] = d;
  }
  double m = d;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = sh[0];
    for (int w = 1; w < NTHR / 64; ++w) bm = fmax(bm, sh[w]);
    part[blockIdx.x] = bm;
  }
}

__global__ void k_schur_sdinv(int64_t ns, const double *__restrict__ smax, double *__restrict__ sdinv,
                              float *__restrict__ sdinv32) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const double d = sdinv[i];
  const double v = d > SCHUR_REL * smax[0] ? 1.0 / d : 0.0;
  sdinv[i] = v;
  if (sdinv32) sdinv32[i] = (float)v;
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

// round to nearest even from the FP32 value (finite inputs)
__global__ void k_to_bf16(int64_t n, const double *__restrict__ a, bf16s *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t u = __float_as_uint((float)a[i]);
  u += 0x7FFFu + ((u >> 16) & 1u);
  bf16s o;
  o.b = (uint16_t)(u >> 16);
  b[i] = o;
}

// ------------------------------------------------------------------ vector kernels

template <typename T>
__global__ void k_fill_t(T *__restrict__ p, int64_t n, T v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// partial sums of a.b (masked)
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

constexpr int KC = 16;  // dot products per pass

// Partial sums of V[k].w for k in [k0, k0+KE), k <= nk (index nk: w.w), one row of `part` per k.
// KE is the number of accumulators, a compile-time count so that every row's load is issued
// without a branch (all KE loads of a thread in flight at once); rows past nk read w again (L1
// hits) and are not stored. (A single-kernel "last block reduces" variant needs device-scope
// fences in every block; on gfx950 each of those writes back the XCD's whole L2 and made this
// kernel 3x slower, so the partials are reduced by a separate small launch.)
template <typename T, int KE>
__global__ void __launch_bounds__(NTHR) k_mdot(int64_t n, const T *__restrict__ V, int64_t ldv,
                                               const T *__restrict__ w, int nk, double *__restrict__ part,
                                               const int *__restrict__ skip) {
  constexpr int VW = Vec<T>::w;
  __shared__ double sh[NTHR / 64][KE];
  if (skip && *skip) return;
  const int lb = xcd_block();
  const int64_t nvec = n / VW;
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  const int k0 = blockIdx.y * KE;
  const int kend = min(KE, nk + 1 - k0);  // live accumulators of this block row
  const T *rows[KE];
  double acc[KE];
#pragma unroll
  for (int c = 0; c < KE; ++c) {
    rows[c] = k0 + c < nk ? V + (int64_t)(k0 + c) * ldv : w;
    acc[c] = 0.0;
  }
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    T wv[VW];
    vload(w + i * VW, wv);
    T vv[KE][VW];
#pragma unroll
    for (int c = 0; c < KE; ++c) vload(rows[c] + i * VW, vv[c]);
#pragma unroll
    for (int c = 0; c < KE; ++c) {
      double s = 0.0;
#pragma unroll
      for (int l = 0; l < VW; ++l) s += (double)vv[c][l] * (double)wv[l];
      acc[c] += s;
    }
  }
  if (lb == gridDim.x - 1) {  // tail rows (n % VW)
    for (int64_t i = nvec * VW + threadIdx.x; i < n; i += blockDim.x) {
      const double wi = (double)w[i];
#pragma unroll
      for (int c = 0; c < KE; ++c) acc[c] += (double)rows[c][i] * wi;
    }
  }
#pragma unroll
  for (int c = 0; c < KE; ++c) {
    const double v = wave_sum(acc[c]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][c] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < kend) {
    double s = 0.0;
#pragma unroll
    for (int wv = 0; wv < NTHR / 64; ++wv) s += sh[wv][threadIdx.x];
    part[(int64_t)(k0 + threadIdx.x) * gridDim.x + lb] = s;
  }
}

// the grid of k_mdot<T, KE>: at most the blocks that are resident together (the kernel's occupancy,
// times the CUs the streaming grid assumes), so that no second, partial wave of blocks runs
template <typename T, int KE>
static int mdot_grid(int grid) {
  static const int per_cu = [] {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_mdot<T, KE>, NTHR, 0) != hipSuccess || b < 1) b = 1;
    return b;
  }();
  const int cap = per_cu * (STREAM_BLOCKS / 4) / NXCD * NXCD;
  return std::max(NXCD, std::min(grid, cap));
}

// the partial sums of V[k].w for k < nk and of w.w (k = nk), one row of `part` per k; returns the
// number of partials per row (the blocks of the launch, at most `grid`). One block row when the
// nk+1 dots fit KC accumulators, else block rows of KC.
template <typename T>
static int mdot_launch(int grid, int64_t n, const T *V, int64_t ldv, const T *w, int nk, double *part,
                       hipStream_t st, const int *skip = nullptr) {
  const int nd = nk + 1;
  if (nd > KC) {
    const int g = mdot_grid<T, KC>(grid);
    k_mdot<T, KC><<<dim3(g, (nd + KC - 1) / KC), NTHR, 0, st>>>(n, V, ldv, w, nk, part, skip);
    return g;
  }
  switch (nd) {
#define OCMP_MDOT(K)                                                   \
  case K: {                                                            \
    const int g = mdot_grid<T, K>(grid);                               \
    k_mdot<T, K><<<g, NTHR, 0, st>>>(n, V, ldv, w, nk, part, skip);          \
    return g;                                                          \
  }
    OCMP_MDOT(1) OCMP_MDOT(2) OCMP_MDOT(3) OCMP_MDOT(4) OCMP_MDOT(5) OCMP_MDOT(6) OCMP_MDOT(7) OCMP_MDOT(8)
    OCMP_MDOT(9) OCMP_MDOT(10) OCMP_MDOT(11) OCMP_MDOT(12) OCMP_MDOT(13) OCMP_MDOT(14) OCMP_MDOT(15)
    OCMP_MDOT(16)
#undef OCMP_MDOT
    default:
      return grid;
  }
}

// State of one GMRES cycle on the device: R column-major (ld = m+1), rotations, g, residual history.
struct GState {
  double *dot, *vs, *R, *csn, *g, *res;
  int m;
  double gscale;
  double *Hf = nullptr;  // if set: the unrotated Hessenberg columns (ld = m+1), for Ritz values
  // if set: raised to 1 by the first column whose residual estimate is not above `stop` (the host's
  // own test); the cycle's later launches then return at entry instead of computing iterations
  // that the host will discard
  int *done = nullptr;
  double stop = 0.0;
};

// Givens step for column j, run by one block: R(:,j) = [vs[k]*dot[k]]_k, h_{j+1,j} = hn;
// apply the previous rotations, form the new one, update g; vs[j+1] = 1/hn; res[j] = |g_{j+1}|*gscale.
__device__ void givens_step(const GState &G, int j, double hn) {
  __shared__ double col[256];
  __shared__ double rot[512];
  const int ld = G.m + 1;
  for (int k = threadIdx.x; k <= j; k += blockDim.x) col[k] = G.vs[k] * G.dot[k];
  for (int k = threadIdx.x; k < 2 * j; k += blockDim.x) rot[k] = G.csn[k];
  if (threadIdx.x == 0) col[j + 1] = hn;
  __syncthreads();
  if (G.Hf) {
    for (int k = threadIdx.x; k <= j + 1; k += blockDim.x) G.Hf[(int64_t)j * ld + k] = col[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < j; ++k) {
      const double c = rot[2 * k], s = rot[2 * k + 1];
      const double a = col[k], b = col[k + 1];
      col[k] = c * a + s * b;
      col[k + 1] = -s * a + c * b;
    }
    const double a = col[j], b = col[j + 1];
    const double den = hypot(a, b);
    const double c = den > 0.0 ? a / den : 1.0, s = den > 0.0 ? b / den : 0.0;
    G.csn[2 * j] = c;
    G.csn[2 * j + 1] = s;
    col[j] = den;
    col[j + 1] = 0.0;
    const double gj = G.g[j];
    G.g[j] = c * gj;
    G.g[j + 1] = -s * gj;
    G.res[j] = fabs(G.g[j + 1]) * G.gscale;
    if (G.done && !(G.res[j] > G.stop)) *G.done = 1;
    G.vs[j + 1] = hn > 0.0 ? 1.0 / hn : 0.0;
  }
  __syncthreads();
  double *R = G.R + (int64_t)j * ld;
  for (int k = threadIdx.x; k <= j + 1; k += blockDim.x) R[k] = col[k];
}

// w -= sum_k c_k V[k] with c_k = vs[k]^2 * dot[k] (classical Gram-Schmidt update); partials of |w|^2
template <typename T>
__global__ void __launch_bounds__(NTHR) k_gs_update(int64_t n, const T *__restrict__ V, int64_t ldv,
                                                    T *__restrict__ w, int nk, const double *__restrict__ dot,
                                                    const double *__restrict__ vs, double *__restrict__ part,
                                                    const int *__restrict__ skip) {
  if (skip && *skip) return;
  constexpr int VW = Vec<T>::w;
  __shared__ double sh[NTHR / 64];
  __shared__ double cks[256];
  for (int k = threadIdx.x; k < nk; k += blockDim.x) cks[k] = vs[k] * vs[k] * dot[k];
  __syncthreads();
  const int lb = xcd_block();
  const int64_t nvec = n / VW;
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    T wv[VW];
    vload(w + i * VW, wv);
    double s[VW];
#pragma unroll
    for (int l = 0; l < VW; ++l) s[l] = (double)wv[l];
#pragma unroll 4
    for (int k = 0; k < nk; ++k) {
      T vv[VW];
      vload(V + (int64_t)k * ldv + i * VW, vv);
      const double c = cks[k];
#pragma unroll
      for (int l = 0; l < VW; ++l) s[l] -= c * (double)vv[l];
    }
#pragma unroll
    for (int l = 0; l < VW; ++l) {
      wv[l] = (T)s[l];
      acc += (double)wv[l] * (double)wv[l];
    }
    vstore(w + i * VW, wv);
  }
  if (lb == gridDim.x - 1) {
    for (int64_t i = nvec * VW + threadIdx.x; i < n; i += blockDim.x) {
      double s = (double)w[i];
      for (int k = 0; k < nk; ++k) s -= cks[k] * (double)V[(int64_t)k * ldv + i];
      const T o = (T)s;
      w[i] = o;
      acc += (double)o * (double)o;
    }
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[lb] = acc;
}

// h_{j+1,j} = sqrt(sum of the update partials), then the Givens step of column j (one block)
__global__ void __launch_bounds__(256) k_norm_givens(const double *__restrict__ part, int nparts, int j, GState G) {
  __shared__ double sh[4];
  __shared__ double hn;
  if (G.done && *G.done) return;  // read by every thread before the first barrier; set only after it
  double s = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) s += part[b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) hn = sqrt(s);
  __syncthreads();
  givens_step(G, j, hn);
}

// back substitution R y = g (k columns) as a column sweep: once y_i is known every
// thread removes R(r,i) y_i from its g_r (column i of R is contiguous); then coef[l] = y_l * vs[l]
__global__ void __launch_bounds__(256) k_gmres_solve(int k, int m, const double *__restrict__ R,
                                                     const double *__restrict__ g, const double *__restrict__ vs,
                                                     double *__restrict__ coef) {
  __shared__ double gs[256];
  __shared__ double yi;
  const int ld = m + 1;
  for (int i = threadIdx.x; i < k; i += blockDim.x) gs[i] = g[i];
  __syncthreads();
  for (int i = k - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      yi = gs[i] / R[(int64_t)i * ld + i];
      coef[i] = yi;
    }
    __syncthreads();
    for (int r = threadIdx.x; r < i; r += blockDim.x) gs[r] -= R[(int64_t)i * ld + r] * yi;
    __syncthreads();
  }
  if (vs)
    for (int i = threadIdx.x; i < k; i += blockDim.x) coef[i] *= vs[i];
}

// z = dinv .* (s * v): the point-Jacobi half of the flexible preconditioner (dinv is 0 off the free dofs)
template <typename T>
__global__ void k_prec_jacobi(int64_t n, const T *__restrict__ dinv, const T *__restrict__ v,
                              const double *__restrict__ sv, T *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) z[i] = (T)((double)dinv[i] * (double)v[i] * sv[0]);
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
                             double inv_beta, int *__restrict__ done) {
  const int i = threadIdx.x;
  for (int k = i; k <= m; k += blockDim.x) g[k] = 0.0;
  if (i == 0) {
    g[0] = beta_scaled;
    vs[0] = inv_beta;
    if (done) *done = 0;
  }
}

// x += dinv .* u  (H1 right Jacobi)  or x += u (no preconditioner)
__global__ void k_diag_apply_add(int64_t n, const double *__restrict__ dinv, const double *__restrict__ u,
                                 double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += (dinv ? dinv[i] : 1.0) * u[i];
}

// y = a x + b y
__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

// ================================================================== host drivers

void axpby(Ctx *ctx, int64_t n, double a, const double *x, double b, double *y) {
  KTimer tm(ctx, ctx->cls("axpby"), 24.0 * n);
  k_axpby<<<nblk(n), NTHR, 0, ctx->stream>>>(n, a, x, b, y);
  OCMP_HIP(hipGetLastError());
}

struct Work {
  // scratch owned by the matrix' solver state; grown on demand
  DBuf<double> r64, u64, t64;
  DBuf<double> xbak;  // x before an s-step cycle (restored if the cycle does not lower the residual)
  DBuf<double> hf;    // unrotated Hessenberg matrix of a standard cycle (Ritz values for s-step shifts)
  DBuf<double> vd;   // FP64 work vectors (n * k)
  DBuf<float> vf;    // FP32 work vectors
  DBuf<double> part; // partial sums
  DBuf<double> small;// device scalars / Hessenberg etc.
  DBuf<int> done;     // the running FGMRES cycle's convergence flag (GState::done)
  double *hbuf = nullptr;  // pinned host
  size_t hcap = 0;
  DBuf<int32_t> nbr;       // DG neighbour table for the ELL operators
  std::vector<hipEvent_t> events;
  Work() = default;
  Work(const Work &) = delete;
  Work &operator=(const Work &) = delete;
  ~Work() {
    if (hbuf) (void)hipHostFree(hbuf);
    for (hipEvent_t e : events) (void)hipEventDestroy(e);
  }
};

void work_release(Work *w) { delete w; }

// the matrix owns its scratch: created on first use, freed with the matrix
static Work &work(Matrix *A) {
  if (!A->work) A->work = new Work();
  return *A->work;
}

static double *pinned(Work &w, size_t n) {
  if (w.hcap < n) {
    if (w.hbuf) (void)hipHostFree(w.hbuf);
    OCMP_HIP(hipHostMalloc((void **)&w.hbuf, n * sizeof(double), hipHostMallocDefault));
    w.hcap = n;
  }
  return w.hbuf;
}

static hipEvent_t work_event(Work &w, size_t i) {
  while (w.events.size() <= i) {
    hipEvent_t e;
    OCMP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    w.events.push_back(e);
  }
  return w.events[i];
}

static void ensure_part(Work &w, size_t n) {
  w.part.alloc(std::max<size_t>(w.part.n, std::max<size_t>(n, (size_t)68 * MAX_RED_BLOCKS)));
}

// first stage of a long reduction: block s sums part[s*seg, (s+1)*seg) into out[s]
__global__ void k_reduce_seg(const double *__restrict__ part, int64_t nparts, int64_t seg, double *__restrict__ out) {
  __shared__ double sh[NTHR / 64];
  const int64_t b0 = (int64_t)blockIdx.x * seg, b1 = min(nparts, b0 + seg);
  double s = 0.0;
  for (int64_t b = b0 + threadIdx.x; b < b1; b += blockDim.x) s += part[b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

constexpr int SEG_BLOCKS = 1024;

// reduce nvals partial rows of `grid` entries into dst (device), fixed order. A single long row
// (one partial per SpMV block) is reduced in two stages through part[grid, grid + SEG_BLOCKS).
static void reduce_to(double *part, int64_t grid, int nvals, double *dst, hipStream_t st, bool do_sqrt = false) {
  if (grid <= 8192 || nvals != 1) {
    k_reduce<<<nvals, NTHR, 0, st>>>(part, (int)grid, dst, do_sqrt ? 1 : 0);
    return;
  }
  const int64_t seg = (grid + SEG_BLOCKS - 1) / SEG_BLOCKS;
  double *tmp = part + grid;
  k_reduce_seg<<<SEG_BLOCKS, NTHR, 0, st>>>(part, grid, seg, tmp);
  k_reduce<<<1, NTHR, 0, st>>>(tmp, SEG_BLOCKS, dst, do_sqrt ? 1 : 0);
}

static bool dg_ell_ready(Matrix *A, Precond *P) {
  return A->dg && P && P->kind == OCMP_PRE_BLOCK_JACOBI && P->ready && P->aell.p && (A->n % 2 == 0);
}

// FP64 residual r = mask*(b - A x); returns ||r|| (syncs)
static double residual64(Matrix *A, Precond *P, const double *b, const double *x, double *r, Work &w) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  ensure_part(w, 0);
  int grid;
  // the ELL residual loads b and stores r two rows at a time (16-byte aligned)
  if (dg_ell_ready(A, P) && ((uintptr_t)b % 16 == 0) && ((uintptr_t)r % 16 == 0)) {
    const int64_t nel = A->space->mesh->nel;
    const int nb = A->dg_nb;
    grid = stream_grid(A->n / 2);
    KTimer tm(ctx, ctx->cls("dg_ell_residual_f64"), (double)A->n * 4 * nb * 8 + 12.0 * nel + 3.0 * A->n * 8);
    switch (nb) {
      case 1: k_dg_ell_residual<1><<<grid, NTHR, 0, st>>>(nel, w.nbr.p, P->aell.p, x, b, r, w.part.p); break;
      case 3: k_dg_ell_residual<3><<<grid, NTHR, 0, st>>>(nel, w.nbr.p, P->aell.p, x, b, r, w.part.p); break;
      case 6: k_dg_ell_residual<6><<<grid, NTHR, 0, st>>>(nel, w.nbr.p, P->aell.p, x, b, r, w.part.p); break;
      case 10: k_dg_ell_residual<10><<<grid, NTHR, 0, st>>>(nel, w.nbr.p, P->aell.p, x, b, r, w.part.p); break;
      default: throw Error(OCMP_ERR_UNSUPPORTED, "DG block size");
    }
  } else {
    grid = spmv_grid(A);
    ensure_part(w, (size_t)grid + SEG_BLOCKS + 64);
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
  ensure_part(w, 0);
  const int grid = red_grid(A->n);
  k_dot<double><<<grid, NTHR, 0, st>>>(A->n, v, v, A->space->d_free.p, w.part.p);
  reduce_to(w.part.p, grid, 1, ctx->scal.p + 101, st);
  double *h = pinned(w, 8);
  OCMP_HIP(hipMemcpyAsync(h, ctx->scal.p + 101, sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  return std::sqrt(h[0]);
}

// out[k] = src[slot[k]] (and its FP32 copy)
__global__ void k_gather_vals(int64_t nnz, const int32_t *__restrict__ slot, const double *__restrict__ src,
                              double *__restrict__ out, float *__restrict__ out32) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const double v = src[slot[k]];
  out[k] = v;
  if (out32) out32[k] = (float)v;
}

// d on the free Schur rows (0 elsewhere, and where d is not positive) and per-block maxima
__global__ void __launch_bounds__(NTHR) k_pdiag_mask(int64_t ns, const double *__restrict__ d,
                                                     const uint8_t *__restrict__ mask, double *__restrict__ out,
                                                     double *__restrict__ part) {
  __shared__ double sh[NTHR / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < ns) {
    v = mask[i] && d[i] > 0.0 ? d[i] : 0.0;
    out[i] = v;
  }
  double m = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double bm = sh[0];
    for (int w = 1; w < NTHR / 64; ++w) bm = fmax(bm, sh[w]);
    part[blockIdx.x] = bm;
  }
}

static size_t pcg_part_size(Matrix *M, const RowBlocks &rb);

// ------------------------------------------------------------------ AMG on S (device)

// P[e] = ident[e] - omega dinv[prow[e]] sum_{s in plist(e)} A[s]
__global__ void k_amg_pvals(int64_t np, const int32_t *__restrict__ prow, const int32_t *__restrict__ plptr,
                            const int32_t *__restrict__ plist, const uint8_t *__restrict__ ident,
                            const double *__restrict__ aval, const double *__restrict__ dinv, double omega,
                            double *__restrict__ pval, float *__restrict__ pval32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np) return;
  double s = 0.0;
  for (int32_t t = plptr[e]; t < plptr[e + 1]; ++t) s += aval[plist[t]];
  const double v = (double)ident[e] - omega * dinv[prow[e]] * s;
  pval[e] = v;
  if (pval32) pval32[e] = (float)v;
}

// out[o] = sum over the pairs [ptr[o], ptr[o+1]) of a[ia[t]] b[ib[t]] (PAIR_LANES lanes per output,
// fixed order)
__global__ void k_amg_pairs(int64_t no, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ia,
                            const int32_t *__restrict__ ib, const double *__restrict__ a, const double *__restrict__ b,
                            double *__restrict__ out, float *__restrict__ out32) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t o = g / PAIR_LANES;
  const int q = (int)(g % PAIR_LANES);
  double s = 0.0;
  if (o < no)
    for (int32_t t = ptr[o] + q; t < ptr[o + 1]; t += PAIR_LANES) s += a[ia[t]] * b[ib[t]];
#pragma unroll
  for (int w = 1; w < PAIR_LANES; w <<= 1) s += __shfl_xor(s, w, PAIR_LANES);
  if (o < no && q == 0) {
    out[o] = s;
    if (out32) out32[o] = (float)s;
  }
}

// the augmented dense matrix [A | I] (row-major, n x 2n) of the coarsest level. Rows left out of the
// solve (dinv 0: not free, or a negligible diagonal) become unit rows with a zero right-hand side and
// leave their columns out of the others, so their inverse rows and columns are zero.
__global__ void k_amg_dense(int64_t n, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                            const double *__restrict__ val, const double *__restrict__ dinv, double *__restrict__ m) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double *row = m + i * 2 * n;
  const bool use = dinv[i] != 0.0;
  for (int64_t j = 0; j < 2 * n; ++j) row[j] = (use && j == n + i) ? 1.0 : 0.0;
  if (!use) {
    row[i] = 1.0;
    return;
  }
  for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
    if (dinv[col[k]] != 0.0) row[col[k]] = val[k];
}

// Gauss-Jordan on [A | I] in one block (A symmetric positive definite: no pivoting), then the
// inverse (the right half) into inv and its FP32 copy
__global__ void __launch_bounds__(1024) k_amg_invert(int64_t n, double *__restrict__ m, double *__restrict__ inv,
                                                     float *__restrict__ inv32) {
  __shared__ double fac[AMG_DENSE_MAX];
  const int64_t w = 2 * n;
  for (int64_t k = 0; k < n; ++k) {
    const double piv = m[k * w + k];
    const double ip = piv != 0.0 ? 1.0 / piv : 0.0;
    __syncthreads();
    for (int64_t j = threadIdx.x; j < w; j += blockDim.x) m[k * w + j] *= ip;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) fac[i] = i == k ? 0.0 : m[i * w + k];
    __syncthreads();
    // 32 lanes along a row, 32 rows at a time (columns left of k are already zero in the rows)
    for (int64_t i = threadIdx.x >> 5; i < n; i += blockDim.x >> 5) {
      const double f = fac[i];
      if (f != 0.0)
        for (int64_t j = k + (threadIdx.x & 31); j < w; j += 32) m[i * w + j] -= f * m[k * w + j];
    }
    __syncthreads();
  }
  for (int64_t i = threadIdx.x >> 5; i < n; i += blockDim.x >> 5)
    for (int64_t j = threadIdx.x & 31; j < n; j += 32) {
      const double v = m[i * w + n + j];
      inv[i * n + j] = v;
      if (inv32) inv32[i * n + j] = (float)v;
    }
}

// the same inversion for a last level of at most AMG_LDS_MAX rows, entirely in LDS
__global__ void __launch_bounds__(1024) k_amg_invert_lds(int64_t n, const double *__restrict__ m0,
                                                         double *__restrict__ inv, float *__restrict__ inv32) {
  __shared__ double m[AMG_LDS_MAX * 2 * AMG_LDS_MAX];
  __shared__ double fac[AMG_LDS_MAX];
  const int w = 2 * (int)n;
  for (int t = threadIdx.x; t < (int)n * w; t += blockDim.x) m[t] = m0[t];
  __syncthreads();
  for (int k = 0; k < (int)n; ++k) {
    const double piv = m[k * w + k];
    const double ip = piv != 0.0 ? 1.0 / piv : 0.0;
    __syncthreads();
    for (int j = threadIdx.x; j < w; j += blockDim.x) m[k * w + j] *= ip;
    for (int i = threadIdx.x; i < (int)n; i += blockDim.x) fac[i] = i == k ? 0.0 : m[i * w + k];
    __syncthreads();
    for (int t = threadIdx.x; t < (int)n * w; t += blockDim.x) {
      const int i = t / w, j = t - i * w;
      if (i != k && j >= k) m[t] -= fac[i] * m[k * w + j];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < (int)(n * n); t += blockDim.x) {
    const int i = t / (int)n, j = t - i * (int)n;
    const double v = m[i * w + (int)n + j];
    inv[t] = v;
    if (inv32) inv32[t] = (float)v;
  }
}

// x = inv b (one wavefront per row)
template <typename T>
__global__ void k_amg_dense_mv(int64_t n, const T *__restrict__ inv, const T *__restrict__ b, T *__restrict__ x,
                               const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = blockIdx.x;
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 64) s += (double)inv[i * n + j] * (double)b[j];
  s = wave_sum(s);
  if (threadIdx.x == 0) x[i] = (T)s;
}

// W = P - w D^-1 Q entrywise on W's pattern (the smoothing folded into the prolongation)
__global__ void k_amg_wvals(int64_t nw, const int32_t *__restrict__ wrow, const int32_t *__restrict__ wps,
                            const int32_t *__restrict__ wqs, const double *__restrict__ pval,
                            const double *__restrict__ qval, const double *__restrict__ dinv, double w,
                            double *__restrict__ wval, float *__restrict__ wval32) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nw) return;
  const int32_t ps = wps[e], qs = wqs[e];
  const double v = (ps >= 0 ? pval[ps] : 0.0) - w * dinv[wrow[e]] * (qs >= 0 ? qval[qs] : 0.0);
  wval[e] = v;
  if (wval32) wval32[e] = (float)v;
}

// a CSR matrix with its row blocks, as one kernel argument
template <typename T>
struct RbCsr {
  int64_t nrblk;
  const int32_t *rblk, *rp, *col, *ubase, *ucol;
  const uint16_t *lcol;
  const T *val;
};
template <typename T>
static RbCsr<T> rb_csr(const Matrix *A, const T *val) {
  const RowBlocks &rb = A->rb64;
  return {rb.nrblk, rb.d_rblk.p, A->d_row_ptr.p, A->d_col.p, rb.d_ubase.p, rb.d_ucol.p, rb.d_lcol.p, val};
}

// The descending half of a V(1,1) level with the input b = g u (g = -sig[0], or 1 without sig):
// the first ga logical blocks form z = w D^-1 (2 b - w A D^-1 b) (the iterate after damped-Jacobi
// pre-smoothing from zero and post-smoothing without the coarse correction), the rest the
// restricted residual b_c = W^T b; bout (optional) receives b.
// the flexible step's Jacobi rows, carried by the top level's descending launch: z = dinv v sv[0]
// over all n rows (the Schur rows, where dinv is 0, are overwritten by the cycle's last launch)
template <typename T>
struct JacRows {
  int64_t n;
  int nblk;  // blocks of the launch that do it (0: none)
  const T *dinv, *v;
  const double *sv;
  T *z;
};

template <typename T, int NNZB>
__global__ void __launch_bounds__(NTHR) k_amg_down(int ga, RbCsr<T> A, RbCsr<T> Wt, const T *__restrict__ u,
                                                   const T *__restrict__ dinv, const double *__restrict__ sig,
                                                   T w, T *__restrict__ z, T *__restrict__ bc, T *__restrict__ bout,
                                                   JacRows<T> J, const int *__restrict__ skip) {
  __shared__ double lds[NNZB * sizeof(T) / sizeof(double)];
  if (skip && *skip) return;  // a converged FGMRES cycle: its remaining launches are empty
  const int lb = xcd_block();
  const double g = sig ? -sig[0] : 1.0;
  const int gw = (int)gridDim.x - ga - J.nblk;  // blocks of the W^T task
  if (lb >= ga + gw) {
    const double sv = J.sv[0];
    for (int64_t i = (int64_t)(lb - ga - gw) * blockDim.x + threadIdx.x; i < J.n; i += (int64_t)J.nblk * blockDim.x)
      J.z[i] = (T)((double)J.dinv[i] * (double)J.v[i] * sv);
    return;
  }
  if (lb < ga) {
    if (lb >= A.nrblk) return;
    csr_row_block<T, T, NNZB>(
        lb, A.rblk, A.rp, A.col, A.ubase, A.ucol, A.lcol, A.val, lds, [&](int32_t c) { return u[c]; },
        [&](int32_t c, T v) { return v * dinv[c]; },
        [&](int64_t r, double s) {
          const double b = g * (double)u[r];
          z[r] = (T)((double)w * (double)dinv[r] * (2.0 * b - (double)w * g * s));
          if (bout) bout[r] = (T)b;
        });
  } else {
    const int wb = lb - ga;
    if (wb >= Wt.nrblk) return;
    csr_row_block<T, T, NNZB>(
        wb, Wt.rblk, Wt.rp, Wt.col, Wt.ubase, Wt.ucol, Wt.lcol, Wt.val, lds, [&](int32_t c) { return u[c]; },
        [&](int32_t, T v) { return v; }, [&](int64_t r, double s) { bc[r] = (T)(g * s); });
  }
}

// b = -s v
template <typename T>
__global__ void k_neg_scale(int64_t n, const T *__restrict__ v, const double *__restrict__ sv, T *__restrict__ b,
                            const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (T)(-(double)v[i] * sv[0]);
}

// one level's V-cycle as a dense matrix, from the dense matrix Mc (nc x nc, row-major) of the levels
// below it: M = w D^-1 (2 I - w A D^-1) + W Mc W^T (the descending launch's z, plus the prolongated
// coarse solve of the restricted right-hand side W^T b); rows with dinv 0 have z = 0, as in that
// launch. M is written with leading dimension ldm >= n, zeros in the padding. One block per row i
// (n, nc <= AMG_TAIL_ROWS): A's row i scattered into LDS, r = W(i,:) Mc in LDS, then
// M(i,j) = z(i,j) + W(j,:) . r for every j.
__global__ void __launch_bounds__(256) k_amg_compose(int64_t n, const int32_t *__restrict__ arp,
                                                     const int32_t *__restrict__ acol,
                                                     const double *__restrict__ aval, const double *__restrict__ dinv,
                                                     double w, const int32_t *__restrict__ wrp,
                                                     const int32_t *__restrict__ wcol, const double *__restrict__ wval,
                                                     int64_t nc, const double *__restrict__ Mc, int64_t ldm,
                                                     double *__restrict__ M, float *__restrict__ M32) {
  __shared__ double arow[AMG_TAIL_ROWS];
  __shared__ double r[AMG_TAIL_ROWS];
  const int64_t i = blockIdx.x;
  for (int64_t j = threadIdx.x; j < n; j += blockDim.x) arow[j] = 0.0;
  __syncthreads();
  for (int32_t k = arp[i] + threadIdx.x; k < arp[i + 1]; k += blockDim.x) atomicAdd(&arow[acol[k]], aval[k]);
  for (int64_t c = threadIdx.x; c < nc; c += blockDim.x) {
    double t = 0.0;
    for (int32_t p = wrp[i]; p < wrp[i + 1]; ++p) t += wval[p] * Mc[(int64_t)wcol[p] * nc + c];
    r[c] = t;
  }
  __syncthreads();
  const double di = dinv[i];
  for (int64_t j = threadIdx.x; j < ldm; j += blockDim.x) {
    double v = 0.0;
    if (j < n) {
      const double z = di != 0.0 ? w * di * ((i == j ? 2.0 : 0.0) - w * arow[j] * dinv[j]) : 0.0;
      double s = 0.0;
      for (int32_t q = wrp[j]; q < wrp[j + 1]; ++q) s += r[wcol[q]] * wval[q];
      v = z + s;
    }
    M[i * ldm + j] = v;
    if (M32) M32[i * ldm + j] = (float)v;
  }
}

// x = M b for the composed tail (n rows, leading dimension ld: a multiple of 64, at most
// AMG_TAIL_LD, zeros past column n): b staged in LDS (zeros past n), one wavefront per row with
// 16-byte loads of the row
template <typename T>
__global__ void __launch_bounds__(256) k_amg_tail_mv(int64_t n, int64_t ld, const T *__restrict__ M,
                                                     const T *__restrict__ b, T *__restrict__ x,
                                                     const int *__restrict__ skip) {
  constexpr int VW = Vec<T>::w;
  __shared__ double bs[AMG_TAIL_LD];
  if (skip && *skip) return;
  for (int64_t j = threadIdx.x; j < ld; j += blockDim.x) bs[j] = j < n ? (double)b[j] : 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  for (int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); i < n; i += (int64_t)gridDim.x * 4) {
    const T *row = M + i * ld;
    double s = 0.0;
#pragma unroll 4
    for (int64_t j = (int64_t)lane * VW; j < ld; j += 64 * VW) {
      T a[VW];
      vload(row + j, a);
#pragma unroll
      for (int l = 0; l < VW; ++l) s += (double)a[l] * bs[j + l];
    }
    s = wave_sum(s);
    if (lane == 0) x[i] = (T)s;
  }
}

// the FP32 copy of an assembled matrix's values, made from its current values
static void refresh_val32(Matrix *A, hipStream_t st) {
  A->d_val32.alloc(A->nnz);
  k_to_f32<<<nblk(A->nnz), NTHR, 0, st>>>(A->nnz, A->d_val.p, A->d_val32.p);
  A->val32_version = A->version;
}

// true when FGMRES multiplies by A's Krylov operator (a separate matrix with its own FP32 values,
// current for A's block pairs) rather than by A: then A's FP32 copy is not read
static bool krylov_operator_separate(const Matrix *A) {
  return A->op && A->op_pairs == A->active_pairs && A->op->nnz > 0;
}

// the hierarchy's values from S's (and S's Jacobi) at every update
static void amg_update(Precond *Pc, bool mixed) {
  Amg &M = *Pc->amg;
  hipStream_t st = Pc->A->ctx->stream;
  for (size_t l = 0; l < M.lv.size(); ++l) {
    AmgLevel &L = *M.lv[l];
    Matrix *A = L.A;
    const int64_t n = A->n;
    if (l == 0) {
      L.dinv.alloc(n);
      OCMP_HIP(hipMemcpyAsync(L.dinv.p, Pc->sdinv.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
      ensure_diag_slots(A);
      L.dinv.alloc(n);
      if (mixed) L.dinv32.alloc(n);
      k_jacobi_diag<<<nblk(n), NTHR, 0, st>>>(n, A->d_dslot.p, A->d_val.p, nullptr, L.dinv.p,
                                              mixed ? L.dinv32.p : nullptr);
      if (mixed) {
        A->d_val32.alloc(A->nnz);
        k_to_f32<<<nblk(A->nnz), NTHR, 0, st>>>(A->nnz, A->d_val.p, A->d_val32.p);
      }
    }
    const size_t nv = (size_t)(l == 0 ? 5 : 3) * n;
    if (mixed) {
      L.dinv32.alloc(n);
      if (l == 0) OCMP_HIP(hipMemcpyAsync(L.dinv32.p, Pc->sdinv32.p, n * sizeof(float), hipMemcpyDeviceToDevice, st));
      L.vf.alloc(nv);
    } else {
      L.vd.alloc(nv);
    }
    if (l + 1 == M.lv.size()) {
      const int64_t nc = n;
      M.ncoarse = nc;
      M.cden.alloc((size_t)nc * 2 * nc);
      M.cinv.alloc((size_t)nc * nc);
      if (mixed) M.cinv32.alloc((size_t)nc * nc);
      k_amg_dense<<<nblk(nc), NTHR, 0, st>>>(nc, A->d_row_ptr.p, A->d_col.p, A->d_val.p, L.dinv.p, M.cden.p);
      if (nc <= AMG_LDS_MAX)
        k_amg_invert_lds<<<1, 1024, 0, st>>>(nc, M.cden.p, M.cinv.p, mixed ? M.cinv32.p : nullptr);
      else
        k_amg_invert<<<1, 1024, 0, st>>>(nc, M.cden.p, M.cinv.p, mixed ? M.cinv32.p : nullptr);
      // compose the levels from the tail's up to the last one (the inverse) into one matrix
      M.tail = l;
      if (nc <= AMG_TAIL_ROWS)
        while (M.tail > 1 && M.lv[M.tail - 1]->A->n <= AMG_TAIL_ROWS) --M.tail;
      if (M.tail < l) {
        const int64_t nt = M.lv[M.tail]->A->n, ni = M.lv[M.tail + 1]->A->n;  // tail, the largest level below it
        M.tld = (nt + 63) / 64 * 64;
        M.tinv.alloc((size_t)nt * M.tld);
        if (mixed) M.tinv32.alloc((size_t)nt * M.tld);
        M.tscr.alloc((size_t)2 * ni * ni);
        const double *mc = M.cinv.p;
        for (size_t k = l; k-- > M.tail;) {
          AmgLevel &K = *M.lv[k];
          const int64_t nk = K.A->n, nkc = M.lv[k + 1]->A->n;
          const bool top = k == M.tail;
          double *out = top ? M.tinv.p : M.tscr.p + (size_t)((l - k) & 1) * ni * ni;
          k_amg_compose<<<(unsigned)nk, 256, 0, st>>>(nk, K.A->d_row_ptr.p, K.A->d_col.p, K.A->d_val.p, K.dinv.p,
                                                      AMG_SMOOTH, K.W->d_row_ptr.p, K.W->d_col.p, K.W->d_val.p, nkc,
                                                      mc, top ? M.tld : nk, out,
                                                      top && mixed ? M.tinv32.p : nullptr);
          mc = out;
        }
      }
      break;
    }
    Matrix *P = L.P, *W = L.W, *Wt = L.Wt, *An = M.lv[l + 1]->A;
    if (mixed) {
      W->d_val32.alloc(W->nnz);
      Wt->d_val32.alloc(Wt->nnz);
    }
    k_amg_pvals<<<nblk(P->nnz), NTHR, 0, st>>>(P->nnz, L.d_prow.p, L.d_plptr.p, L.d_plist.p, L.d_pident.p, A->d_val.p,
                                               L.dinv.p, L.omega, P->d_val.p, nullptr);
    k_amg_pairs<<<nblk(L.nq * PAIR_LANES), NTHR, 0, st>>>(L.nq, L.d_qptr.p, L.d_qa.p, L.d_qp.p, A->d_val.p,
                                                          P->d_val.p, L.q.p, nullptr);
    k_amg_wvals<<<nblk(W->nnz), NTHR, 0, st>>>(W->nnz, L.d_wrow.p, L.d_wps.p, L.d_wqs.p, P->d_val.p, L.q.p, L.dinv.p,
                                               AMG_SMOOTH, W->d_val.p, mixed ? W->d_val32.p : nullptr);
    k_gather_vals<<<nblk(Wt->nnz), NTHR, 0, st>>>(Wt->nnz, L.d_wtslot.p, W->d_val.p, Wt->d_val.p,
                                                  mixed ? Wt->d_val32.p : nullptr);
    k_amg_pairs<<<nblk(An->nnz * PAIR_LANES), NTHR, 0, st>>>(An->nnz, L.d_cptr.p, L.d_cp.p, L.d_cq.p, P->d_val.p,
                                                             L.q.p, An->d_val.p, nullptr);
  }
}

// one V(1,1) cycle with damped Jacobi (weight AMG_SMOOTH) on level l: x ~ A_l^{-1} b with
// b = -sig[0] u (b = u without sig); bout (optional) receives b. Two launches per level (the
// smoothing steps ride on the grid transfers through W) and the dense solve on the last.
template <typename T>
static void amg_vcycle(Amg &M, size_t l, const T *u, const double *sig, T *x, T *bout, hipStream_t st,
                       JacRows<T> J = JacRows<T>{0, 0, nullptr, nullptr, nullptr, nullptr},
                       const int *skip = nullptr) {
  AmgLevel &L = *M.lv[l];
  const int64_t n = L.A->n;
  auto vecs = [](AmgLevel &V) -> T * {
    if constexpr (sizeof(T) == 8)
      return V.vd.p;
    else
      return V.vf.p;
  };
  auto vals = [](Matrix *A) -> const T * {
    if constexpr (sizeof(T) == 8)
      return A->d_val.p;
    else
      return A->d_val32.p;
  };
  if (l == M.tail) {
    const bool last = l + 1 == M.lv.size();
    const T *inv;
    if constexpr (sizeof(T) == 8)
      inv = last ? M.cinv.p : M.tinv.p;
    else
      inv = last ? M.cinv32.p : M.tinv32.p;
    if (!last) {  // the composed tail (no sig or bout below the top level)
      k_amg_tail_mv<T><<<(unsigned)std::min<int64_t>((n + 7) / 8, 512), 256, 0, st>>>(n, M.tld, inv, u, x, skip);
      return;
    }
    const T *b = u;
    if (sig) {
      T *bb = bout ? bout : vecs(L);
      k_neg_scale<T><<<nblk(n), NTHR, 0, st>>>(n, u, sig, bb, skip);
      b = bb;
    } else if (bout) {
      OCMP_HIP(hipMemcpyAsync(bout, u, n * sizeof(T), hipMemcpyDeviceToDevice, st));
    }
    k_amg_dense_mv<T><<<(unsigned)n, 64, 0, st>>>(n, inv, b, x, skip);
    return;
  }
  AmgLevel &C = *M.lv[l + 1];
  const T *dinv;
  if constexpr (sizeof(T) == 8)
    dinv = L.dinv.p;
  else
    dinv = L.dinv32.p;
  T *z = vecs(L) + 2 * n;
  T *bc = vecs(C), *xc = vecs(C) + C.A->n;
  const int ga = spmv_grid<double, double>(L.A), gw = spmv_grid<double, double>(L.Wt);
  const RbCsr<T> a = rb_csr<T>(L.A, vals(L.A)), wt = rb_csr<T>(L.Wt, vals(L.Wt));
  k_amg_down<T, SPMV_NNZB64><<<ga + gw + J.nblk, NTHR, 0, st>>>(ga, a, wt, u, dinv, sig, (T)AMG_SMOOTH, z, bc, bout,
                                                                 J, skip);
  amg_vcycle<T>(M, l + 1, bc, nullptr, xc, nullptr, st, JacRows<T>{0, 0, nullptr, nullptr, nullptr, nullptr}, skip);
  spmv_dispatch<T, T>(L.W, vals(L.W), xc, x, nullptr, nullptr, nullptr, nullptr, 6, nullptr, spmv_grid<T, T>(L.W), st,
                      (T)0, z, (T)0, skip);
}

// z[s0:s1] = -(cycles V-cycles on S) (s v)[s0:s1]: the first cycle from zero, each further one on the
// residual of the current iterate. Level 0's vectors: b, (unused), z (the cycle's own), then the
// residual and the correction of the further cycles.
template <typename T>
__global__ void k_add(int64_t n, const T *__restrict__ e, T *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += e[i];
}

// with jac, the top level's first launch also writes the Jacobi rows z = dinv v sv[0] (all rows,
// the Schur ones overwritten after); returns false (nothing done) when the hierarchy is a single
// dense level, which has no such launch
template <typename T>
static bool schur_amg_apply(Precond *Pc, const T *v, const double *sv, T *z, const T *jac_dinv = nullptr,
                            const int *skip = nullptr) {
  Amg &M = *Pc->amg;
  if (jac_dinv && M.lv.size() < 2) return false;
  Ctx *ctx = Pc->A->ctx;
  hipStream_t st = ctx->stream;
  AmgLevel &L0 = *M.lv[0];
  const int64_t n = L0.A->n;
  T *vb;
  const T *val;
  if constexpr (sizeof(T) == 8) {
    vb = L0.vd.p;
    val = L0.A->d_val.p;
  } else {
    vb = L0.vf.p;
    val = L0.A->d_val32.p;
  }
  T *b = vb, *rr = vb + 3 * n, *ee = vb + 4 * n;
  T *x = z + Pc->s0;
  KTimer tm(ctx, ctx->cls(sizeof(T) == 8 ? "schur_amg_vcycle_f64" : "schur_amg_vcycle_f32"));
  const bool more = Pc->schur_amg > 1;
  JacRows<T> J{0, 0, nullptr, nullptr, nullptr, nullptr};
  if (jac_dinv) J = JacRows<T>{Pc->A->n, STREAM_BLOCKS, jac_dinv, v, sv, z};
  amg_vcycle<T>(M, 0, v + Pc->s0, sv, x, more ? b : nullptr, st, J, skip);
  for (int c = 1; c < Pc->schur_amg; ++c) {
    spmv_dispatch<T, T>(L0.A, val, x, rr, nullptr, nullptr, nullptr, nullptr, 5, nullptr, spmv_grid<T, T>(L0.A), st,
                        (T)0, b, (T)0);
    amg_vcycle<T>(M, 0, rr, nullptr, ee, nullptr, st);
    k_add<T><<<nblk(n), NTHR, 0, st>>>(n, ee, x);
  }
  return true;
}

// (re)build A_uu and A_up from A's pattern and its current active block pairs
static void rebuild_velocity_blocks(Precond *P) {
  delete P->Auu;
  delete P->Aup;
  P->Auu = P->Aup = nullptr;
  build_velocity_blocks(P);
}

// the block-triangular step's blocks of A, their Jacobi, and 1/d on the Schur rows
static void btri_update(Precond *P, bool mixed) {
  Matrix *A = P->A;
  hipStream_t st = A->ctx->stream;
  if (!P->Auu || P->vel_pairs != A->active_pairs) rebuild_velocity_blocks(P);
  Matrix *Auu = P->Auu, *Aup = P->Aup;
  const int64_t nu = P->s0, ns = A->n - P->s0;
  if (mixed) {
    Auu->d_val32.alloc(Auu->nnz);
    Aup->d_val32.alloc(Aup->nnz);
  }
  k_gather_vals<<<nblk(Auu->nnz), NTHR, 0, st>>>(Auu->nnz, P->d_uuslot.p, A->d_val.p, Auu->d_val.p,
                                                  mixed ? Auu->d_val32.p : nullptr);
  k_gather_vals<<<nblk(Aup->nnz), NTHR, 0, st>>>(Aup->nnz, P->d_upslot.p, A->d_val.p, Aup->d_val.p,
                                                  mixed ? Aup->d_val32.p : nullptr);
  if (mixed && P->bf16) {
    Auu->d_valbf.alloc(Auu->nnz);
    k_to_bf16<<<nblk(Auu->nnz), NTHR, 0, st>>>(Auu->nnz, Auu->d_val.p, Auu->d_valbf.p);
  }
  ensure_diag_slots(Auu);
  P->vdinv.alloc(nu);
  if (mixed) P->vdinv32.alloc(nu);
  k_jacobi_diag<<<nblk(nu), NTHR, 0, st>>>(nu, Auu->d_dslot.p, Auu->d_val.p, A->space->d_free.p, P->vdinv.p,
                                           mixed ? P->vdinv32.p : nullptr);
  P->pdinv.alloc(ns);
  if (mixed) P->pdinv32.alloc(ns);
  {
    Work &w = work(A);
    const unsigned nbs = nblk(ns);
    ensure_part(w, (size_t)nbs + 8);
    double *smax = w.part.p + nbs;
    k_pdiag_mask<<<nbs, NTHR, 0, st>>>(ns, P->pdiag.p, A->space->d_free.p + nu, P->pdinv.p, w.part.p);
    k_max_reduce<<<1, NTHR, 0, st>>>(w.part.p, (int)nbs, smax);
    k_schur_sdinv<<<nbs, NTHR, 0, st>>>(ns, smax, P->pdinv.p, mixed ? P->pdinv32.p : nullptr);
  }
  if (mixed)
    P->vvf.alloc((size_t)9 * nu);
  else
    P->vvd.alloc((size_t)9 * nu);
  P->vpart.alloc(std::max(P->vpart.n, pcg_part_size(Auu, P->vrb)));
  P->vflag.alloc(1);
}

// the block-triangular velocity rows of a flexible Schur step: A_up's values (FP64, and FP32 for
// the mixed path) from A's
static void tri_update(Precond *P, bool mixed) {
  Matrix *A = P->A;
  hipStream_t st = A->ctx->stream;
  if (!P->Aup || P->vel_pairs != A->active_pairs) rebuild_velocity_blocks(P);
  Matrix *Aup = P->Aup;
  if (mixed) Aup->d_val32.alloc(Aup->nnz);
  k_gather_vals<<<nblk(Aup->nnz), NTHR, 0, st>>>(Aup->nnz, P->d_upslot.p, A->d_val.p, Aup->d_val.p,
                                                  mixed ? Aup->d_val32.p : nullptr);
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
    if (P->kind == OCMP_PRE_JACOBI) {
      ensure_diag_slots(A);
      k_jacobi_diag<<<nblk(A->n), NTHR, 0, st>>>(A->n, A->d_dslot.p, A->d_val.p, A->space->d_free.p, P->dinv.p,
                                                 mixed ? P->dinv32.p : nullptr);
    }
    if (P->kind == OCMP_PRE_JACOBI && P->schur_blocks) {
      if (!P->schur_flag.p) {
        std::vector<uint8_t> f(A->n, 0);
        const auto &bl = A->space->blocks;
        for (size_t k = 0; k < bl.size(); ++k)
          if (P->schur_blocks >> k & 1)
            for (int64_t i = 0; i < bl[k].n1; ++i) f[bl[k].dof_off + i] = 1;
        P->schur_flag.upload(f, st);
        OCMP_HIP(hipStreamSynchronize(st));
      }
      const bool schur_step = P->schur_iters > 0 || P->schur_amg > 0;
      if (schur_step) {
        // the flexible step solves the Schur rows itself: its Jacobi entries there are zero
        k_zero_flagged<<<nblk(A->n), NTHR, 0, st>>>(A->n, P->schur_flag.p, P->dinv.p, mixed ? P->dinv32.p : nullptr);
      } else {
        Work &w = work(A);
        const unsigned nb = nblk(A->n);
        ensure_part(w, (size_t)nb + 8);
        double *smax = w.part.p + nb;
        k_schur_diag<<<nb, NTHR, 0, st>>>(A->n, A->d_row_ptr.p, A->d_col.p, A->d_val.p, A->space->d_free.p,
                                          P->schur_flag.p, P->dinv.p, w.part.p);
        k_max_reduce<<<1, NTHR, 0, st>>>(w.part.p, (int)nb, smax);
        k_schur_finish<<<nb, NTHR, 0, st>>>(A->n, P->schur_flag.p, smax, P->dinv.p, mixed ? P->dinv32.p : nullptr);
      }
      if (schur_step) {
        if (!P->S) build_schur_matrix(P);
        Matrix *S = P->S;
        const int64_t ns = S->n;
        P->sdpos.alloc(A->n);
        k_schur_dpos<<<nblk(A->n * DPOS_LANES), NTHR, 0, st>>>(A->n, P->s0, P->s1, A->d_row_ptr.p, A->d_col.p,
                                                               A->d_dslot.p, A->d_val.p, A->space->d_free.p,
                                                               P->sdpos.p);
        if (P->d_spptr.p) {
          const int64_t base = A->row_ptr[P->s0], nk = A->row_ptr[P->s1] - base;
          P->sbf.alloc(std::max<int64_t>(nk, 1));
          k_schur_bfac<<<nblk(nk), NTHR, 0, st>>>(nk, base, A->d_col.p, A->d_val.p, P->sdpos.p, P->sbf.p);
          k_schur_values_pairs<<<nblk(S->nnz * PAIR_LANES), NTHR, 0, st>>>(
              S->nnz, P->d_scslot.p, P->d_spptr.p, P->d_spj.p, P->d_spt.p, base, P->sbf.p, A->d_val.p, S->d_val.p);
        } else
          k_schur_values<<<nblk(ns * 64), NTHR, 0, st>>>(ns, P->s0, S->d_row_ptr.p, S->d_col.p, A->d_row_ptr.p,
                                                         A->d_col.p, A->d_val.p, P->d_sseg.p, P->sdpos.p, S->d_val.p);
        ensure_diag_slots(S);
        P->sdinv.alloc(ns);
        if (mixed) P->sdinv32.alloc(ns);
        {
          Work &w = work(A);
          const unsigned nbs = nblk(ns);
          ensure_part(w, (size_t)nbs + 8);
          double *smax = w.part.p + nbs;
          k_schur_sdiag<<<nbs, NTHR, 0, st>>>(ns, S->d_dslot.p, S->d_val.p, A->space->d_free.p + P->s0, P->sdinv.p,
                                              w.part.p);
          k_max_reduce<<<1, NTHR, 0, st>>>(w.part.p, (int)nbs, smax);
          k_schur_sdinv<<<nbs, NTHR, 0, st>>>(ns, smax, P->sdinv.p, mixed ? P->sdinv32.p : nullptr);
        }
        // S is small (a pressure space): its FP32 SpMV keeps the 2048-nonzero row blocks for more
        // blocks per launch
        if (mixed) {
          S->d_val32.alloc(S->nnz);
          k_to_f32<<<nblk(S->nnz), NTHR, 0, st>>>(S->nnz, S->d_val.p, S->d_val32.p);
          if (P->bf16) {
            S->d_valbf.alloc(S->nnz);
            k_to_bf16<<<nblk(S->nnz), NTHR, 0, st>>>(S->nnz, S->d_val.p, S->d_valbf.p);
          }
          P->svf.alloc((size_t)8 * ns);
        } else {
          P->svd.alloc((size_t)8 * ns);
        }
        P->spart.alloc(std::max(P->spart.n, pcg_part_size(S, P->srb)));
        P->sflag.alloc(1);
        if (P->schur_amg > 0) {
          if (!P->amg) {
            // the hierarchy is built once, on the host, from this first set of values
            std::vector<double> hv(S->nnz), hd(ns);
            OCMP_HIP(hipMemcpyAsync(hv.data(), S->d_val.p, S->nnz * sizeof(double), hipMemcpyDeviceToHost, st));
            OCMP_HIP(hipMemcpyAsync(hd.data(), P->sdinv.p, ns * sizeof(double), hipMemcpyDeviceToHost, st));
            OCMP_HIP(hipStreamSynchronize(st));
            build_schur_amg(P, hv, hd);
          }
          amg_update(P, mixed);
        }
      }
      if (P->btri) btri_update(P, mixed);
      if (P->schur_tri && schur_step) tri_update(P, mixed);
    }
    if (mixed) {
      // a Schur step runs only inside FGMRES, which reads A's FP32 copy only when it multiplies by A
      // itself (and then refreshes a stale copy first): skip the conversion while a separate Krylov
      // operator is current
      const bool fgmres_only = P->schur_blocks && (P->schur_iters > 0 || P->schur_amg > 0);
      if (!(fgmres_only && krylov_operator_separate(A)))
        refresh_val32(A, st);
      else
        A->d_val32.alloc(A->nnz);
      ensure_row_blocks32(A);
    }
  } else {
    const Mesh &m = *A->space->mesh;
    const int nb = A->dg_nb;
    OCMP_REQUIRE(A->n % 4 == 0, "block Jacobi ELL operator needs a dof count divisible by 4");
    P->binv.alloc((size_t)m.nel * nb * nb);
    P->ahat.alloc(mixed ? 0 : (size_t)A->n * 3 * nb);
    if (mixed) P->ahat32.alloc((size_t)A->n * 3 * nb);
    P->aell.alloc((size_t)A->n * 4 * nb);
    Work &w = work(A);
    w.nbr.alloc((size_t)3 * m.nel);
#define OCMP_DG_CASE(NBV)                                                                                           \
  case NBV:                                                                                                         \
    k_dg_block_inverse<NBV><<<nblk(m.nel, 128), 128, 0, st>>>(m.nel, A->d_row_ptr.p, A->d_dg_pos.p, A->d_val.p,    \
                                                              P->binv.p);                                           \
    k_dg_build_ell<NBV><<<nblk(A->n), NTHR, 0, st>>>(m.nel, A->d_row_ptr.p, A->d_dg_pos.p, m.d_elem_nb.p,          \
                                                     A->d_val.p, P->binv.p, mixed ? nullptr : P->ahat.p,            \
                                                     mixed ? P->ahat32.p : nullptr, P->aell.p, w.nbr.p);            \
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

// The point-Jacobi rows (and the FP32 copy of A) from A's current values; the flexible Schur step
// (S, its AMG hierarchy or inner solve) stays as the last full update left it: a lagged step,
// still a fixed preconditioner for each FGMRES solve.
void precond_update_jacobi(Precond *P) {
  Matrix *A = P->A;
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const bool mixed = P->precision == OCMP_MIXED;
  OCMP_REQUIRE(P->ready, "a full preconditioner update must come first");
  OCMP_REQUIRE(P->kind == OCMP_PRE_JACOBI && P->schur_blocks && (P->schur_iters > 0 || P->schur_amg > 0) && !P->btri,
               "the Jacobi-only update is for the Schur inner solve and the Schur AMG step");
  KTimer tm(ctx, ctx->cls("precond_setup_jacobi"));
  k_jacobi_diag<<<nblk(A->n), NTHR, 0, st>>>(A->n, A->d_dslot.p, A->d_val.p, A->space->d_free.p, P->dinv.p,
                                             mixed ? P->dinv32.p : nullptr);
  k_zero_flagged<<<nblk(A->n), NTHR, 0, st>>>(A->n, P->schur_flag.p, P->dinv.p, mixed ? P->dinv32.p : nullptr);
  if (mixed && !krylov_operator_separate(A)) refresh_val32(A, st);  // FGMRES only: see precond_update
  if (P->schur_tri) tri_update(P, mixed);
  OCMP_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ CG

// Batches of `check` iterations are enqueued ahead; the host checks the
// residual history of batch b while batch b+1 runs, so the GPU never idles.
template <typename VT, typename T>
static int cg_inner(Matrix *A, const VT *val, const T *dinv, T *x, T *r, T *z, T *p, T *q, double tol_rr,
                    int maxit, int check_every, Work &w, int &iters_done, double *last_rr) {
  // x, r, z, p set by the caller: r = residual, z = dinv r, p = z, sc[0] = r.z
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  const int sgrid = spmv_grid<VT, T>(A);
  ensure_part(w, (size_t)sgrid + SEG_BLOCKS + 64);
  const int vgrid = red_grid(n);
  double *sc = ctx->scal.p;  // [0],[1]: rz ring, [2]: pq
  const int H = std::max(check_every, 1);
  w.small.alloc(std::max<size_t>(w.small.n, (size_t)2 * H + 64));
  double *hist = w.small.p;  // two halves of H entries
  double *hh = pinned(w, 2 * H + 8);
  const int c_spmv = ctx->cls(sizeof(VT) == 8 ? "spmv_csr_f64_cg" : "spmv_csr_f32_cg");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "cg_update_f64" : "cg_update_f32");
  struct Batch {
    int it0, nb, half;
  };
  std::deque<Batch> inflight;
  int it = 0, bidx = 0;
  auto enqueue = [&]() {
    const int nb = std::min(H, maxit - it);
    const int half = bidx & 1;
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
      reduce_to(w.part.p, vgrid, 1, sc + ((it + 1) & 1), st);
      reduce_to(w.part.p + vgrid, vgrid, 1, hist + half * H + s, st);
      k_cg_p<T><<<nblk(n), NTHR, 0, st>>>(n, p, z, sc, it);
    }
    OCMP_HIP(hipMemcpyAsync(hh + half * H, hist + half * H, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    OCMP_HIP(hipEventRecord(work_event(w, half), st));
    inflight.push_back({it - nb, nb, half});
    ++bidx;
  };
  while (true) {
    while (inflight.size() < 2 && it < maxit) enqueue();
    if (inflight.empty()) break;
    Batch b = inflight.front();
    inflight.pop_front();
    OCMP_HIP(hipEventSynchronize(work_event(w, b.half)));
    for (int s = 0; s < b.nb; ++s) {
      *last_rr = hh[b.half * H + s];
      if (!(*last_rr > tol_rr)) {  // also stops on NaN
        // a batch still in flight only refines x further; let it finish so r and x stay consistent
        OCMP_HIP(hipStreamSynchronize(st));
        iters_done = it;
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
  ensure_part(w, 0);
  w.r64.alloc(n);
  const double bn = norm_masked(A, b, w);
  double rn = residual64(A, P, b, x, w.r64.p, w);
  // ||b|| on free dofs is the reference; a problem driven only by Dirichlet data (b_F = 0)
  // measures against the initial residual instead
  const double bref = bn > 0.0 ? bn : (rn > 0.0 ? rn : 1.0);
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
  const int vgrid = red_grid(n);
  if (!mixed) {
    w.vd.alloc((size_t)4 * n);
    double *z = w.vd.p, *p = z + n, *q = p + n;
    double *r = w.r64.p;
    int outer = 0;
    while (true) {
      ++outer;
      // z = dinv r, p = z, rz = r.z   (r is the true residual)
      k_jacobi_apply<double, double><<<nblk(n), NTHR, 0, st>>>(n, P->dinv.p, r, 1.0, z, p);
      k_dot<double><<<vgrid, NTHR, 0, st>>>(n, r, z, nullptr, w.part.p);
      reduce_to(w.part.p, vgrid, 1, ctx->scal.p + 0, st);
      int done = 0;
      double last = 0.0;
      cg_inner<double, double>(A, A->d_val.p, P->dinv.p, x, r, z, p, q, o.tol * o.tol * bref * bref,
                               o.maxit - total, check, w, done, &last);
      total += done;
      rn = residual64(A, P, b, x, w.r64.p, w);  // true residual (also restarts r)
      st_.rel_residual = rn / bref;
      if (rn <= o.tol * bref || total >= o.maxit || !std::isfinite(rn) || done == 0) break;
    }
    st_.outer = outer;
  } else {
    OCMP_REQUIRE(P->dinv32.p && A->d_val32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
    if (A->val32_version != A->version) refresh_val32(A, st);  // an update that left the copy to FGMRES
    w.vf.alloc((size_t)5 * n);
    float *d = w.vf.p, *r = d + n, *z = r + n, *p = z + n, *q = p + n;
    // each inner FP32 CG reduces its residual by this factor (H1 P3, 7.2M dofs: 1e-5 took 560
    // iterations in 2 refinement steps, 1e-4 570 in 3, 1e-6 610, 1e-7 670)
    const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-5;
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
      rn = residual64(A, P, b, x, w.r64.p, w);
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
                       int maxit, int check, Work &w, bool capture_h = false) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  const bool dg = A->dg && P && P->kind == OCMP_PRE_BLOCK_JACOBI;
  const uint8_t *mask = A->dg ? nullptr : A->space->d_free.p;
  constexpr int VW = Vec<T>::w;
  // basis rows padded to a multiple of VW elements so every row is 16-byte aligned
  const int64_t ldv = (n + VW - 1) / VW * VW;
  DBuf<T> *Vbuf;
  if constexpr (sizeof(T) == 8) {
    w.vd.alloc(std::max<size_t>(w.vd.n, (size_t)(m + 1) * ldv));
    Vbuf = &w.vd;
  } else {
    w.vf.alloc(std::max<size_t>(w.vf.n, (size_t)(m + 1) * ldv));
    Vbuf = &w.vf;
  }
  T *V = Vbuf->p;
  // small device state: dot[m+2], vs[m+2], R[(m+1)*m], csn[2m], g[m+1], res[m+8], coef[m]
  const size_t ns = (size_t)(m + 2) * 2 + (size_t)(m + 1) * m + 2 * m + (m + 1) + (m + 8) + m;
  w.small.alloc(std::max(w.small.n, ns));
  double *dot = w.small.p, *vs = dot + (m + 2), *R = vs + (m + 2), *csn = R + (size_t)(m + 1) * m,
         *g = csn + 2 * m, *res = g + (m + 1), *coef = res + (m + 8);
  const int grid = stream_grid((n + VW - 1) / VW);
  ensure_part(w, (size_t)(m + 2 + KC) * grid);
  // v0 = r0 (scale 1/beta, stored unnormalised)
  k_scale_copy<double, T><<<nblk(n), NTHR, 0, st>>>(n, r0, 1.0, V);
  k_gmres_init<<<1, 64, 0, st>>>(m, 1.0, g, vs, 1.0 / beta, nullptr);
  double *hres = pinned(w, m + 8);
  const int c_op = ctx->cls(dg ? (sizeof(T) == 8 ? "dg_ahat_spmv_f64" : "dg_ahat_spmv_f32")
                               : (sizeof(T) == 8 ? "spmv_csr_f64_gmres" : "spmv_csr_f32_gmres"));
  const int c_dot = ctx->cls(sizeof(T) == 8 ? "gmres_mdot_f64" : "gmres_mdot_f32");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "gmres_update_f64" : "gmres_update_f32");
  const double gscale = beta;  // g is relative to beta
  const int limit = std::min(m, maxit);
  GState G{dot, vs, R, csn, g, res, m, gscale};
  if (capture_h) {  // keep the unrotated Hessenberg columns in w.hf (ld = m+1)
    w.hf.alloc(std::max<size_t>(w.hf.n, (size_t)(m + 1) * m));
    G.Hf = w.hf.p;
  }
  struct Batch {
    int j0, nb, ev;
  };
  std::deque<Batch> inflight;
  int j = 0, bidx = 0, k_used = -1;
  auto enqueue = [&]() {
    const int j_end = std::min(limit, j + check);
    for (int jj = j; jj < j_end; ++jj) {
      T *vj = V + (int64_t)jj * ldv;
      T *wv = V + (int64_t)(jj + 1) * ldv;
      {
        const double opb = dg ? (double)n * 3 * A->dg_nb * sizeof(T) + 12.0 * A->space->mesh->nel +
                                    2.0 * n * sizeof(T)
                              : csr_bytes(A, sizeof(T), sizeof(T));
        KTimer tm(ctx, c_op, opb);
        if (dg) {
          const T *ah;
          if constexpr (sizeof(T) == 8)
            ah = P->ahat.p;
          else
            ah = P->ahat32.p;
          const int64_t nel = A->space->mesh->nel;
          const unsigned gb = nblk((n + VW - 1) / VW);
          switch (A->dg_nb) {
            case 1: k_dg_ahat_spmv<T, 1><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 3: k_dg_ahat_spmv<T, 3><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 6: k_dg_ahat_spmv<T, 6><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
            case 10: k_dg_ahat_spmv<T, 10><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, vj, vs + jj, wv); break;
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
          spmv_dispatch<T, T>(A, vals, vj, wv, nullptr, mask, cs, vs + jj, 3, nullptr, spmv_grid<T, T>(A), st);
        }
      }
      int dgrid;
      {
        KTimer tm(ctx, c_dot, (double)(jj + 2) * n * sizeof(T));
        dgrid = mdot_launch<T>(grid, n, V, ldv, wv, jj + 1, w.part.p, st);
      }
      reduce_to(w.part.p, dgrid, jj + 1, dot, st);
      {
        KTimer tm(ctx, c_upd, (double)(jj + 3) * n * sizeof(T));
        k_gs_update<T><<<grid, NTHR, 0, st>>>(n, V, ldv, wv, jj + 1, dot, vs, w.part.p, nullptr);
      }
      k_norm_givens<<<1, 256, 0, st>>>(w.part.p, grid, jj, G);
    }
    const int nb = j_end - j;
    OCMP_HIP(hipMemcpyAsync(hres + j, res + j, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    const int ev = bidx & 1;
    OCMP_HIP(hipEventRecord(work_event(w, ev), st));
    inflight.push_back({j, nb, ev});
    ++bidx;
    j = j_end;
  };
  while (k_used < 0) {
    while (inflight.size() < 2 && j < limit) enqueue();
    if (inflight.empty()) {
      k_used = j;
      break;
    }
    Batch b = inflight.front();
    inflight.pop_front();
    OCMP_HIP(hipEventSynchronize(work_event(w, b.ev)));
    for (int s = 0; s < b.nb; ++s)
      if (!(hres[b.j0 + s] > stop_abs)) {
        k_used = b.j0 + s + 1;
        break;
      }
  }
  // x += M^{-1} sum_l y_l v_l   (columns beyond k_used computed by a batch still in flight are ignored)
  k_gmres_solve<<<1, 256, 0, st>>>(k_used, m, R, g, vs, coef);
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

// ------------------------------------------------------------------ s-step (communication-avoiding) GMRES
//
// One block of CA_S Arnoldi steps reads the basis once instead of CA_S times: the CA_S
// monomial Krylov vectors w_i = Ahat^i q_k are generated by CA_S operator applications, then
// orthogonalised against Q[0..k] with one block classical Gram-Schmidt pass (dot products in
// one kernel, the update fused with the Gram matrix of the result), and against each other by
// CholQR2 (Cholesky of the FP64 Gram matrix, twice). The Hessenberg columns follow from the
// change of basis: Ahat [q_k..q_{k+s-1}] = V (Rr[:,1:s+1] - [H_old Rr[0:k,0:s]; 0]) T^{-1}.
constexpr int CA_S = 5;
// basis vectors per y-tile of the block dot kernel: with FP32 data each thread's partial sums
// (a dozen products) accumulate in FP32, which leaves the registers for 16 rows; wave and block
// sums are FP64 in both cases
template <typename T>
struct CaTile {
  static constexpr int KQ = sizeof(T) == 4 ? 16 : 8;
  using Acc = typename std::conditional<sizeof(T) == 4, float, double>::type;
};
constexpr int CA_KQ_MAX = 16;
// The block-Jacobi-folded DG operator has identity diagonal blocks, so its spectrum sits around
// 1: the basis is built with (Ahat - CA_SIGMA I), which keeps the Gram matrix of a block of 5
// at condition ~1e4 instead of ~1e7 for plain powers (small enough for FP32 storage).
constexpr double CA_SIGMA = 1.0;
constexpr int CA_NG = CA_S * (CA_S + 1) / 2;
// largest restart of an s-step cycle: the block kernels size their LDS for k + 1 + CA_S <= CA_MMAX + 1
// basis rows (the host only takes the s-step path for m <= CA_MMAX)
constexpr int CA_MMAX = 50;

// Basis of one block: v_{t+1} = (Ahat - d[t] I) v_t + c[t] v_{t-1}, so that
// Ahat v_t = v_{t+1} + d[t] v_t - c[t] v_{t-1}. The first cycle of a solve uses the shifted
// monomial basis (d = CA_SIGMA, c = 0); later cycles use Leja-ordered Ritz values of the first
// cycle (a Newton basis), which keeps the block well conditioned for FP32 storage. d and c are
// held as the basis precision applies them.
struct CaShift {
  double d[CA_S];
  double c[CA_S];
};

// Newton shifts from the leading k x k block of an unrotated Hessenberg matrix (device, ld = m+1):
// Leja-ordered Ritz values, rounded to what the basis precision applies (the DG kernel applies
// 1 - d, the CSR kernel d). Returns false (sh unchanged) if the eigenvalues are not available.
static bool newton_from_hessenberg(const double *Hf, int m, int k, bool dg, bool fp32, CaShift &sh) {
  std::vector<double> Hh((size_t)(m + 1) * k), wr(k), wi(k);
  OCMP_HIP(hipMemcpy(Hh.data(), Hf, Hh.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (!hessenberg_eigvals(k, Hh.data(), m + 1, wr.data(), wi.data())) return false;
  CaShift out;
  newton_shifts(k, wr.data(), wi.data(), CA_S, out.d, out.c);
  if (fp32)
    for (int i = 0; i < CA_S; ++i) {
      out.d[i] = dg ? 1.0 - (double)(float)(1.0 - out.d[i]) : (double)(float)out.d[i];
      out.c[i] = (double)(float)out.c[i];
    }
  sh = out;
  return true;
}

// part[(q*S + i) * grid + lb] = partial of Q[q] . W[i],  q < nq
// Rows past nq in the last y-tile read row nq-1 again (no branch around the loads, so all of a
// thread's row loads are in flight together) and their sums are dropped.
template <typename T, int CA_KQ>
__global__ void __launch_bounds__(NTHR) k_bdot(int64_t n, const T *__restrict__ Q, int64_t ldv, int nq,
                                               const T *__restrict__ W, double *__restrict__ part) {
  constexpr int VW = Vec<T>::w;
  using Acc = typename CaTile<T>::Acc;
  __shared__ double sh[NTHR / 64][CA_KQ * CA_S];
  const int lb = xcd_block();
  const int64_t nvec = n / VW;
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  const int q0 = blockIdx.y * CA_KQ;
  const int nqt = min(CA_KQ, nq - q0);
  Acc acc[CA_KQ][CA_S];
#pragma unroll
  for (int c = 0; c < CA_KQ; ++c)
#pragma unroll
    for (int t = 0; t < CA_S; ++t) acc[c][t] = (Acc)0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    T wv[CA_S][VW];
#pragma unroll
    for (int t = 0; t < CA_S; ++t) vload(W + (int64_t)t * ldv + i * VW, wv[t]);
    T qv[CA_KQ][VW];
#pragma unroll
    for (int c = 0; c < CA_KQ; ++c) vload(Q + (int64_t)min(q0 + c, nq - 1) * ldv + i * VW, qv[c]);
#pragma unroll
    for (int c = 0; c < CA_KQ; ++c)
#pragma unroll
      for (int t = 0; t < CA_S; ++t) {
        Acc sum = (Acc)0;
#pragma unroll
        for (int l = 0; l < VW; ++l) sum += (Acc)qv[c][l] * (Acc)wv[t][l];
        acc[c][t] += sum;
      }
  }
  if (lb == gridDim.x - 1)
    for (int64_t i = nvec * VW + threadIdx.x; i < n; i += blockDim.x)
#pragma unroll
      for (int c = 0; c < CA_KQ; ++c)
#pragma unroll
        for (int t = 0; t < CA_S; ++t)
          acc[c][t] += (Acc)Q[(int64_t)min(q0 + c, nq - 1) * ldv + i] * (Acc)W[(int64_t)t * ldv + i];
#pragma unroll
  for (int c = 0; c < CA_KQ; ++c)
#pragma unroll
    for (int t = 0; t < CA_S; ++t) {
      const double v = wave_sum((double)acc[c][t]);
      if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][c * CA_S + t] = v;
    }
  __syncthreads();
  if ((int)threadIdx.x < nqt * CA_S) {
    double sum = 0.0;
#pragma unroll
    for (int wv = 0; wv < NTHR / 64; ++wv) sum += sh[wv][threadIdx.x];
    part[(int64_t)(q0 * CA_S + threadIdx.x) * gridDim.x + lb] = sum;
  }
}

// W[t] -= sum_q C[q][t] Q[q]  (C: nq x S, q-major), then partials of the Gram matrix of the
// rounded result, packed upper triangle: part[g * grid + lb]
template <typename T>
__global__ void __launch_bounds__(NTHR) k_bupdate_gram(int64_t n, const T *__restrict__ Q, int64_t ldv, int nq,
                                                       T *__restrict__ W, const double *__restrict__ C,
                                                       double *__restrict__ part) {
  constexpr int VW = Vec<T>::w;
  __shared__ double cq[(CA_MMAX + 1) * CA_S];
  __shared__ double sh[NTHR / 64][CA_NG];
  for (int e = threadIdx.x; e < nq * CA_S; e += blockDim.x) cq[e] = C[e];
  __syncthreads();
  const int lb = xcd_block();
  const int64_t nvec = n / VW;
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  double gacc[CA_NG];
#pragma unroll
  for (int g = 0; g < CA_NG; ++g) gacc[g] = 0.0;
  auto row = [&](double (&w)[CA_S], auto qload) {
    for (int q = 0; q < nq; ++q) {
      const double qv = qload(q);
#pragma unroll
      for (int t = 0; t < CA_S; ++t) w[t] -= cq[q * CA_S + t] * qv;
    }
  };
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    T wv[CA_S][VW];
    double w[VW][CA_S];
#pragma unroll
    for (int t = 0; t < CA_S; ++t) {
      vload(W + (int64_t)t * ldv + i * VW, wv[t]);
#pragma unroll
      for (int l = 0; l < VW; ++l) w[l][t] = (double)wv[t][l];
    }
    for (int q = 0; q < nq; ++q) {
      T qv[VW];
      vload(Q + (int64_t)q * ldv + i * VW, qv);
#pragma unroll
      for (int t = 0; t < CA_S; ++t) {
        const double c = cq[q * CA_S + t];
#pragma unroll
        for (int l = 0; l < VW; ++l) w[l][t] -= c * (double)qv[l];
      }
    }
#pragma unroll
    for (int t = 0; t < CA_S; ++t) {
#pragma unroll
      for (int l = 0; l < VW; ++l) wv[t][l] = (T)w[l][t];
      vstore(W + (int64_t)t * ldv + i * VW, wv[t]);
    }
#pragma unroll
    for (int l = 0; l < VW; ++l) {
      int g = 0;
#pragma unroll
      for (int a = 0; a < CA_S; ++a)
#pragma unroll
        for (int b = a; b < CA_S; ++b, ++g) gacc[g] += (double)wv[a][l] * (double)wv[b][l];
    }
  }
  if (lb == gridDim.x - 1)
    for (int64_t i = nvec * VW + threadIdx.x; i < n; i += blockDim.x) {
      double w[CA_S];
      for (int t = 0; t < CA_S; ++t) w[t] = (double)W[(int64_t)t * ldv + i];
      row(w, [&](int q) { return (double)Q[(int64_t)q * ldv + i]; });
      T o[CA_S];
      for (int t = 0; t < CA_S; ++t) {
        o[t] = (T)w[t];
        W[(int64_t)t * ldv + i] = o[t];
      }
      int g = 0;
      for (int a = 0; a < CA_S; ++a)
        for (int b = a; b < CA_S; ++b, ++g) gacc[g] += (double)o[a] * (double)o[b];
    }
#pragma unroll
  for (int g = 0; g < CA_NG; ++g) {
    const double v = wave_sum(gacc[g]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][g] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < CA_NG) {
    double sum = 0.0;
#pragma unroll
    for (int wv = 0; wv < NTHR / 64; ++wv) sum += sh[wv][threadIdx.x];
    part[(int64_t)threadIdx.x * gridDim.x + lb] = sum;
  }
}

// W <- W Rinv (Rinv upper triangular S x S, row-major), Gram partials of the result if part
template <typename T>
__global__ void __launch_bounds__(NTHR) k_btrsm_gram(int64_t n, T *__restrict__ W, int64_t ldv,
                                                     const double *__restrict__ Rinv, double *__restrict__ part) {
  constexpr int VW = Vec<T>::w;
  __shared__ double ri[CA_S * CA_S];
  __shared__ double sh[NTHR / 64][CA_NG];
  if ((int)threadIdx.x < CA_S * CA_S) ri[threadIdx.x] = Rinv[threadIdx.x];
  __syncthreads();
  const int lb = xcd_block();
  const int64_t nvec = n / VW;
  const int64_t chunk = (nvec + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(nvec, i0 + chunk);
  double gacc[CA_NG];
#pragma unroll
  for (int g = 0; g < CA_NG; ++g) gacc[g] = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    T wv[CA_S][VW];
#pragma unroll
    for (int t = 0; t < CA_S; ++t) vload(W + (int64_t)t * ldv + i * VW, wv[t]);
    T ov[CA_S][VW];
#pragma unroll
    for (int l = 0; l < VW; ++l)
#pragma unroll
      for (int t = 0; t < CA_S; ++t) {
        double sum = 0.0;
#pragma unroll
        for (int a = 0; a <= t; ++a) sum += (double)wv[a][l] * ri[a * CA_S + t];
        ov[t][l] = (T)sum;
      }
#pragma unroll
    for (int t = 0; t < CA_S; ++t) vstore(W + (int64_t)t * ldv + i * VW, ov[t]);
    if (part) {
#pragma unroll
      for (int l = 0; l < VW; ++l) {
        int g = 0;
#pragma unroll
        for (int a = 0; a < CA_S; ++a)
#pragma unroll
          for (int b = a; b < CA_S; ++b, ++g) gacc[g] += (double)ov[a][l] * (double)ov[b][l];
      }
    }
  }
  if (lb == gridDim.x - 1)
    for (int64_t i = nvec * VW + threadIdx.x; i < n; i += blockDim.x) {
      double w[CA_S];
      T o[CA_S];
      for (int t = 0; t < CA_S; ++t) w[t] = (double)W[(int64_t)t * ldv + i];
      for (int t = 0; t < CA_S; ++t) {
        double sum = 0.0;
        for (int a = 0; a <= t; ++a) sum += w[a] * ri[a * CA_S + t];
        o[t] = (T)sum;
        W[(int64_t)t * ldv + i] = o[t];
      }
      if (part) {
        int g = 0;
        for (int a = 0; a < CA_S; ++a)
          for (int b = a; b < CA_S; ++b, ++g) gacc[g] += (double)o[a] * (double)o[b];
      }
    }
  if (!part) return;
#pragma unroll
  for (int g = 0; g < CA_NG; ++g) {
    const double v = wave_sum(gacc[g]);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][g] = v;
  }
  __syncthreads();
  if ((int)threadIdx.x < CA_NG) {
    double sum = 0.0;
#pragma unroll
    for (int wv = 0; wv < NTHR / 64; ++wv) sum += sh[wv][threadIdx.x];
    part[(int64_t)threadIdx.x * gridDim.x + lb] = sum;
  }
}

// Cholesky G = R^T R of the packed Gram matrix; R and R^{-1} (row-major, upper). A non-positive
// pivot sets *flag (the block is then discarded by the host) and returns the identity. Every loop
// has a constant trip count so the 5x5 arrays stay in registers. ident, if given, is set to the
// identity (the second factor of a single-pass CholQR).
__global__ void k_ca_chol(const double *__restrict__ gp, double *__restrict__ R, double *__restrict__ Rinv,
                          double *__restrict__ flag, double *__restrict__ ident) {
  if (ident && threadIdx.x < CA_S * CA_S) ident[threadIdx.x] = (threadIdx.x / CA_S == threadIdx.x % CA_S) ? 1.0 : 0.0;
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double G[CA_S][CA_S], U[CA_S][CA_S], Ui[CA_S][CA_S];
#pragma unroll
  for (int a = 0; a < CA_S; ++a)
#pragma unroll
    for (int b = a; b < CA_S; ++b) G[a][b] = G[b][a] = gp[a * CA_S - a * (a - 1) / 2 + (b - a)];
  bool ok = true;
#pragma unroll
  for (int j = 0; j < CA_S; ++j) {
    double d = G[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) d -= U[k][j] * U[k][j];
    ok = ok && d > 0.0;
    U[j][j] = sqrt(d > 0.0 ? d : 1.0);
#pragma unroll
    for (int c = 0; c < j; ++c) U[j][c] = 0.0;
#pragma unroll
    for (int c = j + 1; c < CA_S; ++c) {
      double v = G[j][c];
#pragma unroll
      for (int k = 0; k < j; ++k) v -= U[k][j] * U[k][c];
      U[j][c] = v / U[j][j];
    }
  }
  if (!ok) {
    flag[0] = 1.0;
#pragma unroll
    for (int i = 0; i < CA_S * CA_S; ++i) {
      R[i] = (i / CA_S == i % CA_S) ? 1.0 : 0.0;
      Rinv[i] = (i / CA_S == i % CA_S) ? 1.0 : 0.0;
    }
    return;
  }
  // inverse of the upper triangular U, column by column (back substitution)
#pragma unroll
  for (int c = 0; c < CA_S; ++c)
#pragma unroll
    for (int r = CA_S - 1; r >= 0; --r) {
      double v = (r == c) ? 1.0 : 0.0;
#pragma unroll
      for (int k = r + 1; k < CA_S; ++k) v -= U[r][k] * Ui[k][c];
      Ui[r][c] = v / U[r][r];
    }
#pragma unroll
  for (int a = 0; a < CA_S; ++a)
#pragma unroll
    for (int b = 0; b < CA_S; ++b) {
      R[a * CA_S + b] = U[a][b];
      Rinv[a * CA_S + b] = (b < a) ? 0.0 : Ui[a][b];
    }
}

// New Hessenberg columns k..k+S-1 from the change of basis (Ahat v_t = v_{t+1} + d_t v_t -
// c_t v_{t-1}), then their Givens rotations.
// Hf: unrotated Hessenberg, Hr: rotated (both column-major, ld = m+1). C: (k+1) x S q-major.
// The residual estimates of the new columns and the CholQR breakdown flag also go straight to the
// host's pinned status buffer hst (res at [j], the flag at [m + 8 + slot]); the host reads them
// after the event recorded behind this kernel.
__global__ void __launch_bounds__(256) k_ca_hess(int k, int m, const double *__restrict__ C,
                                                 const double *__restrict__ R1, const double *__restrict__ R2,
                                                 double *__restrict__ Hf, GState G, CaShift B,
                                                 const double *__restrict__ flag, double *__restrict__ hst,
                                                 int slot) {
  constexpr int MAXR = CA_MMAX + 1;  // rows of the basis: k + 1 + CA_S <= m + 1
  __shared__ double Rr[MAXR * (CA_S + 1)];          // (k+1+S) x (S+1), column-major
  __shared__ double HC[(CA_MMAX + 1) * CA_S];       // H_old (k+1 x k) @ Rr[0:k, 0:S]
  __shared__ double Rt[CA_S * CA_S];
  __shared__ double Hn[MAXR * CA_S];
  __shared__ double rot[2 * MAXR];                  // the k rotations of the earlier columns
  __shared__ double tl[CA_S][CA_S + 2];             // rows k..k+c+1 of column k+c after those rotations
  const int ld = m + 1, nr = k + 1 + CA_S;
  const int tid = threadIdx.x;
  if (tid < CA_S * CA_S) {
    const int a = tid / CA_S, b = tid % CA_S;
    double v = 0.0;
#pragma unroll
    for (int t = 0; t < CA_S; ++t) v += R2[a * CA_S + t] * R1[t * CA_S + b];
    Rt[a * CA_S + b] = v;
  }
  for (int e = tid; e < 2 * k; e += blockDim.x) rot[e] = G.csn[e];
  __syncthreads();
  for (int e = tid; e < nr * (CA_S + 1); e += blockDim.x) {
    const int r = e % nr, c = e / nr;
    double v;
    if (c == 0)
      v = (r == k) ? 1.0 : 0.0;
    else if (r <= k)
      v = C[r * CA_S + (c - 1)];
    else
      v = Rt[(r - k - 1) * CA_S + (c - 1)];
    Rr[c * nr + r] = v;
  }
  __syncthreads();
  for (int e = tid; e < (k + 1) * CA_S; e += blockDim.x) {
    const int r = e % (k + 1), c = e / (k + 1);
    double v = 0.0;
    for (int t = (r > 0 ? r - 1 : 0); t < k; ++t) v += Hf[(int64_t)t * ld + r] * Rr[c * nr + t];  // H_old is upper Hessenberg
    HC[c * (k + 1) + r] = v;
  }
  __syncthreads();
  // Hnew = (Rr Bc - [HC; 0]) T^{-1},  T = Rr[k:k+S, 0:S] upper triangular, Bc the change of basis
  // (column t: 1 at row t+1, d[t] at row t, -c[t] at row t-1); one thread per row r solves x T = rhs
  for (int r = tid; r < nr; r += blockDim.x) {
    double x[CA_S];
#pragma unroll
    for (int c = 0; c < CA_S; ++c) {
      double v = Rr[(c + 1) * nr + r] + B.d[c] * Rr[c * nr + r] - (r <= k ? HC[c * (k + 1) + r] : 0.0);
      if (c > 0) v -= B.c[c] * Rr[(c - 1) * nr + r];
#pragma unroll
      for (int t = 0; t < c; ++t) v -= x[t] * Rr[c * nr + (k + t)];
      x[c] = v / Rr[c * nr + (k + c)];
    }
#pragma unroll
    for (int c = 0; c < CA_S; ++c) Hn[r * CA_S + c] = x[c];
  }
  __syncthreads();
  for (int e = tid; e < CA_S * ld; e += blockDim.x) {
    const int c = e / ld, r = e % ld;
    Hf[(int64_t)(k + c) * ld + r] = (r < nr) ? Hn[r * CA_S + c] : 0.0;
  }
  // rotated copy: thread c sweeps the k earlier rotations down column k+c (rows 0..k-1 are
  // final after them), carrying the running entry in a register
  if (tid < CA_S) {
    const int c = tid;
    double *Rc = G.R + (int64_t)(k + c) * ld;
    double a = Hn[c];
    for (int t = 0; t < k; ++t) {
      const double b = Hn[(t + 1) * CA_S + c], cs = rot[2 * t], sn = rot[2 * t + 1];
      Rc[t] = cs * a + sn * b;
      a = -sn * a + cs * b;
    }
    tl[c][0] = a;
#pragma unroll
    for (int i = 1; i < CA_S + 2; ++i) tl[c][i] = (i <= c + 1) ? Hn[(k + i) * CA_S + c] : 0.0;
  }
  __syncthreads();
  if (tid != 0) return;
  // the S new rotations, each applied to the later columns of the block
  double cb[CA_S], sb[CA_S];
  double gcur = G.g[k];
#pragma unroll
  for (int c = 0; c < CA_S; ++c) {
    const int j = k + c;
    double v[CA_S + 2];
#pragma unroll
    for (int i = 0; i < CA_S + 2; ++i) v[i] = tl[c][i];
#pragma unroll
    for (int i = 0; i < c; ++i) {
      const double a = v[i], b = v[i + 1];
      v[i] = cb[i] * a + sb[i] * b;
      v[i + 1] = -sb[i] * a + cb[i] * b;
    }
    const double a = v[c], b = v[c + 1];
    const double den = hypot(a, b);
    cb[c] = den > 0.0 ? a / den : 1.0;
    sb[c] = den > 0.0 ? b / den : 0.0;
    G.csn[2 * j] = cb[c];
    G.csn[2 * j + 1] = sb[c];
    v[c] = den;
    v[c + 1] = 0.0;
    G.g[j] = cb[c] * gcur;
    gcur = -sb[c] * gcur;
    G.g[j + 1] = gcur;
    G.res[j] = fabs(gcur) * G.gscale;
    hst[j] = G.res[j];
    double *Rc = G.R + (int64_t)j * ld + k;
#pragma unroll
    for (int i = 0; i <= c + 1; ++i) Rc[i] = v[i];
  }
  hst[m + 8 + slot] = flag[0];
}

// One CA-GMRES(m) cycle on the DG operator Ahat (block Jacobi folded in); same contract as
// gmres_cycle. Returns -1 (and changes nothing in x) if the first block breaks down.
template <typename T>
static int ca_gmres_cycle(Matrix *A, Precond *P, const double *r0, double beta, double *x, int m, double stop_abs,
                          int maxit, Work &w, bool &breakdown, CaShift &shift, bool &newton) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  constexpr int VW = Vec<T>::w;
  const int64_t ldv = (n + VW - 1) / VW * VW;
  DBuf<T> *Vbuf;
  if constexpr (sizeof(T) == 8) {
    w.vd.alloc(std::max<size_t>(w.vd.n, (size_t)(m + 1) * ldv));
    Vbuf = &w.vd;
  } else {
    w.vf.alloc(std::max<size_t>(w.vf.n, (size_t)(m + 1) * ldv));
    Vbuf = &w.vf;
  }
  T *Q = Vbuf->p;
  const size_t ns = (size_t)(m + 1) * CA_S + 4 * CA_S * CA_S + 2 * (size_t)(m + 1) * m + 2 * m + (m + 1) + (m + 8) +
                    m + 8;
  w.small.alloc(std::max(w.small.n, ns));
  double *Cm = w.small.p, *R1 = Cm + (size_t)(m + 1) * CA_S, *R1i = R1 + CA_S * CA_S, *R2 = R1i + CA_S * CA_S,
         *R2i = R2 + CA_S * CA_S, *Hf = R2i + CA_S * CA_S, *Hr = Hf + (size_t)(m + 1) * m, *csn = Hr + (size_t)(m + 1) * m,
         *g = csn + 2 * m, *res = g + (m + 1), *coef = res + (m + 8), *flag = coef + m, *one = flag + 1;
  const int grid = stream_grid((n + VW - 1) / VW);
  // the block dot reduces CA_KQ x CA_S sums per block: fewer, longer blocks amortise that
  // the block dot runs on a quarter of the streaming grid per y-tile (measured 88 us against
  // 91 us at half and 117 us at the full grid, DG-P2 512^2 FP32): fewer, longer row streams
  const int dgrid = std::max(NXCD, grid / 4 / NXCD * NXCD);
  ensure_part(w, (size_t)(m + 1) * CA_S * grid + 64);
  k_scale_copy<double, T><<<nblk(n), NTHR, 0, st>>>(n, r0, 1.0 / beta, Q);
  // g = e_0 (relative to beta), flag = 0, one = 1
  k_gmres_init<<<1, 64, 0, st>>>(m, 1.0, g, flag, 0.0, nullptr);
  k_fill_t<double><<<1, 64, 0, st>>>(one, 1, 1.0);
  double *hres = pinned(w, m + 16);
  double *hres_d = nullptr;  // the same pinned buffer as the device addresses it
  OCMP_HIP(hipHostGetDevicePointer((void **)&hres_d, hres, 0));
  GState G{nullptr, nullptr, Hr, csn, g, res, m, beta};
  const bool dg = A->dg;
  const int c_op = ctx->cls(dg ? (sizeof(T) == 8 ? "dg_ahat_spmv_f64" : "dg_ahat_spmv_f32")
                               : (sizeof(T) == 8 ? "spmv_csr_f64_gmres" : "spmv_csr_f32_gmres"));
  const int c_dot = ctx->cls(sizeof(T) == 8 ? "ca_bdot_f64" : "ca_bdot_f32");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "ca_bupdate_f64" : "ca_bupdate_f32");
  const int c_qr = ctx->cls(sizeof(T) == 8 ? "ca_cholqr_f64" : "ca_cholqr_f32");
  const int64_t nel = A->space->mesh->nel;
  const unsigned gb = nblk((n + VW - 1) / VW);
  // DG: the block-Jacobi-folded ELL operator; H1 / saddle point: the CSR matrix with the point
  // Jacobi (or Schur-diagonal) scaling applied to its columns (right preconditioning)
  const T *ah = nullptr, *vals = nullptr, *cs = nullptr;
  if constexpr (sizeof(T) == 8) {
    if (dg) ah = P->ahat.p;
    vals = A->d_val.p;
    if (P && P->kind == OCMP_PRE_JACOBI) cs = P->dinv.p;
  } else {
    if (dg) ah = P->ahat32.p;
    vals = A->d_val32.p;
    if (P && P->kind == OCMP_PRE_JACOBI) cs = P->dinv32.p;
  }
  const uint8_t *mask = dg ? nullptr : A->space->d_free.p;
  const int sgrid = dg ? 0 : spmv_grid<T, T>(A);
  struct Batch {
    int j0, ev;
  };
  std::deque<Batch> inflight;
  const int limit = std::min(m, std::max(CA_S, (maxit + CA_S - 1) / CA_S * CA_S)) / CA_S * CA_S;
  int k = 0, bidx = 0, k_used = -1;
  breakdown = false;
  auto enqueue = [&]() {
    for (int i = 0; i < CA_S; ++i) {
      // v_{i+1} = (Ahat - d_i I) v_i + c_i v_{i-1} (c_0 = 0 always)
      const T pc = (T)shift.c[i];
      const T *xp = (i > 0 && shift.c[i] != 0.0) ? Q + (int64_t)(k + i - 1) * ldv : nullptr;
      const T *xin = Q + (int64_t)(k + i) * ldv;
      T *yout = Q + (int64_t)(k + i + 1) * ldv;
      if (!dg) {
        KTimer tm(ctx, c_op, csr_bytes(A, sizeof(T), sizeof(T)) + (xp ? 2.0 : 1.0) * n * sizeof(T));
        spmv_dispatch<T, T>(A, vals, xin, yout, nullptr, mask, cs, nullptr, 4, nullptr, sgrid, st, (T)shift.d[i], xp,
                            pc);
        continue;
      }
      // the diagonal blocks of Ahat are the identity, so the kernel's diagonal coefficient is 1 - d_i
      const T dc = (T)(1.0 - shift.d[i]);
      KTimer tm(ctx, c_op, (double)n * 3 * A->dg_nb * sizeof(T) + 12.0 * nel + (xp ? 3.0 : 2.0) * n * sizeof(T));
      switch (A->dg_nb) {
        case 1: k_dg_ahat_spmv<T, 1><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, xin, one, yout, dc, xp, pc); break;
        case 3: k_dg_ahat_spmv<T, 3><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, xin, one, yout, dc, xp, pc); break;
        case 6: k_dg_ahat_spmv<T, 6><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, xin, one, yout, dc, xp, pc); break;
        case 10: k_dg_ahat_spmv<T, 10><<<gb, NTHR, 0, st>>>(nel, w.nbr.p, ah, xin, one, yout, dc, xp, pc); break;
        default: throw Error(OCMP_ERR_UNSUPPORTED, "DG block size");
      }
    }
    T *W = Q + (int64_t)(k + 1) * ldv;
    const int nq = k + 1;
    {
      // (an MFMA version of this kernel, v_mfma_f32_16x16x4_f32 tiles reading Q and W once, measured
      // 156 us against 121 us here: its 16-row operand gathers coalesce worse than these row streams)
      // tile height: the fewest y-tiles of at most KQ rows, each as short as the rows allow
      constexpr int KQ = CaTile<T>::KQ;
      const int ny = (nq + KQ - 1) / KQ, per = (nq + ny - 1) / ny;
      const int kq = per <= 4 ? 4 : per <= 8 ? 8 : per <= 12 ? 12 : 16;
      KTimer tm(ctx, c_dot, (double)(nq + CA_S * ny) * n * sizeof(T));
      dim3 gd(dgrid, (nq + kq - 1) / kq);
      switch (kq) {
        case 4: k_bdot<T, 4><<<gd, NTHR, 0, st>>>(n, Q, ldv, nq, W, w.part.p); break;
        case 8: k_bdot<T, 8><<<gd, NTHR, 0, st>>>(n, Q, ldv, nq, W, w.part.p); break;
        case 12: k_bdot<T, (KQ >= 12 ? 12 : KQ)><<<gd, NTHR, 0, st>>>(n, Q, ldv, nq, W, w.part.p); break;
        default: k_bdot<T, KQ><<<gd, NTHR, 0, st>>>(n, Q, ldv, nq, W, w.part.p); break;
      }
    }
    reduce_to(w.part.p, dgrid, nq * CA_S, Cm, st);
    {
      KTimer tm(ctx, c_upd, (double)(nq + 2 * CA_S) * n * sizeof(T));
      k_bupdate_gram<T><<<grid, NTHR, 0, st>>>(n, Q, ldv, nq, W, Cm, w.part.p);
    }
    double *gp = Cm + (size_t)(m + 1) * CA_S - CA_NG;  // packed Gram (the last rows of Cm are unused: nq <= m+1-S)
    reduce_to(w.part.p, grid, CA_NG, gp, st);
    {
      // CholQR, and a second pass (CholQR2) for FP64 bases only: the shifted blocks have Gram
      // condition ~1e4, so one pass leaves FP32 vectors orthogonal to ~1e-6, below their precision
      KTimer tm(ctx, c_qr, (sizeof(T) == 8 ? 4.0 : 2.0) * CA_S * n * sizeof(T));
      k_ca_chol<<<1, 64, 0, st>>>(gp, R1, R1i, flag, sizeof(T) == 8 ? nullptr : R2);
      if (sizeof(T) == 8) {
        k_btrsm_gram<T><<<grid, NTHR, 0, st>>>(n, W, ldv, R1i, w.part.p);
        reduce_to(w.part.p, grid, CA_NG, gp, st);
        k_ca_chol<<<1, 64, 0, st>>>(gp, R2, R2i, flag, nullptr);
        k_btrsm_gram<T><<<grid, NTHR, 0, st>>>(n, W, ldv, R2i, nullptr);
      } else {
        k_btrsm_gram<T><<<grid, NTHR, 0, st>>>(n, W, ldv, R1i, nullptr);
      }
    }
    k_ca_hess<<<1, 256, 0, st>>>(k, m, Cm, R1, R2, Hf, G, shift, flag, hres_d, bidx & 1);
    const int ev = bidx & 1;
    OCMP_HIP(hipEventRecord(work_event(w, ev), st));
    inflight.push_back({k, ev});
    ++bidx;
    k += CA_S;
  };
  while (k_used < 0) {
    while (inflight.size() < 2 && k < limit) enqueue();
    if (inflight.empty()) {
      k_used = k;
      break;
    }
    Batch b = inflight.front();
    inflight.pop_front();
    OCMP_HIP(hipEventSynchronize(work_event(w, b.ev)));
    if (hres[m + 8 + (b.ev)] != 0.0) {  // CholQR broke down in this block: keep the columns before it
      breakdown = true;
      k_used = b.j0;
      break;
    }
    for (int s = 0; s < CA_S; ++s)
      if (!(hres[b.j0 + s] > stop_abs)) {
        k_used = b.j0 + s + 1;
        break;
      }
  }
  // drain the look-ahead before the buffers are reused
  OCMP_HIP(hipStreamSynchronize(st));
  if (k_used <= 0) return breakdown ? -1 : 0;
  // Newton-basis shifts for the later cycles of this solve from this cycle's Ritz values
  if (!newton && !breakdown && k_used >= 2 * CA_S)
    newton = newton_from_hessenberg(Hf, m, k_used, A->dg, sizeof(T) == 4, shift);
  k_gmres_solve<<<1, 256, 0, st>>>(k_used, m, Hr, g, nullptr, coef);
  w.u64.alloc(n);
  k_combine<T><<<nblk(n), NTHR, 0, st>>>(n, Q, ldv, k_used, coef, beta, w.u64.p);
  if (!dg) {
    const double *dinv = (P && P->kind == OCMP_PRE_JACOBI) ? P->dinv.p : nullptr;
    k_diag_apply_add<<<nblk(n), NTHR, 0, st>>>(n, dinv, w.u64.p, x);
  } else {
    switch (A->dg_nb) {
      case 1: k_dg_block_apply_add<1><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 3: k_dg_block_apply_add<3><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 6: k_dg_block_apply_add<6><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
      case 10: k_dg_block_apply_add<10><<<nblk(n), NTHR, 0, st>>>(nel, P->binv.p, w.u64.p, x); break;
    }
  }
  OCMP_HIP(hipGetLastError());
  return k_used;
}

// Flexible GMRES cycle (right preconditioning that may vary between iterations):
// z_j = M_j v_j with M = D^{-1} (+ the learned correction when `learned`), w = A z_j,
// x += Z y. Same device-side Gram-Schmidt / Givens machinery as gmres_cycle.
// ------------------------------------------------------------------ Schur inner PCG
//
// Jacobi PCG on S for the Schur rows of an FGMRES preconditioner step, in the single-reduction
// form of Chronopoulos and Gear (J. Comput. Appl. Math. 25 (1989) 153-168): with
// gamma = (r, u), delta = (S u, u), beta = gamma / gamma_old and
// alpha = gamma / (delta - beta gamma / alpha_old), one launch per iteration updates
// p = u + beta p, s = w + beta s (= S p), x += alpha p, r -= alpha s, u = dinv r and computes the
// next w = S u with both partial dot products (k_scg_fused). Each block sums the partial dot
// products it needs itself (the same order in every block, so all blocks use the same alpha and
// beta); the (gamma, alpha) pair of the previous iteration alternates between two slots by
// iteration. A zero denominator gives a zero step.
// The host enqueues max_iters iterations without waiting. With a tolerance, the launch that finds
// gamma <= tol^2 gamma_0 sets a device flag instead of updating, and every later launch of the
// sequence returns on reading it.
//
// Scalars (sc): [0..3] (gamma, alpha) slots, [4] gamma_0, [5] iterations done (a running count
// over the solve).

// x = 0, r = sign s v - sub (sub may be NULL), u = dinv r; partials of r.u; clears the converged flag
template <typename T>
__global__ void __launch_bounds__(NTHR) k_scg_init(int64_t n, const T *__restrict__ v, const double *__restrict__ sv,
                                                   double sign, const T *__restrict__ sub,
                                                   const T *__restrict__ dinv, T *__restrict__ x, T *__restrict__ r,
                                                   T *__restrict__ u, double *__restrict__ gpart,
                                                   int *__restrict__ flag, const int *__restrict__ outer) {
  __shared__ double sh[NTHR / 64];
  if (outer && *outer) {  // the outer cycle converged: this solve starts converged, so its launches are empty
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 1;
    return;
  }
  const int lb = xcd_block();
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t i0 = lb * chunk, i1 = min(n, i0 + chunk);
  const double sc = sign * sv[0];
  double acc = 0.0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    x[i] = (T)0;
    T ri = (T)((double)v[i] * sc);
    if (sub) ri -= sub[i];
    r[i] = ri;
    const T ui = dinv[i] * ri;
    u[i] = ui;
    acc += (double)ri * (double)ui;
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    gpart[lb] = acc;
    if (blockIdx.x == 0) flag[0] = 0;
  }
}

// the two sums of sum_parts in one pass: both sets of loads in flight together, one shared round
__device__ __forceinline__ void sum_parts2(const double *__restrict__ pa, int na, const double *__restrict__ pb,
                                           int nb, double *sh, double &sa, double &sb) {
  double a[8], b[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int i = threadIdx.x + k * blockDim.x;
    const double va = pa[min(i, na - 1)], vb = pb[min(i, nb - 1)];
    a[k] = i < na ? va : 0.0;
    b[k] = i < nb ? vb : 0.0;
  }
  for (int i = threadIdx.x + 8 * blockDim.x; i < na; i += blockDim.x) a[0] += pa[i];
  for (int i = threadIdx.x + 8 * blockDim.x; i < nb; i += blockDim.x) b[0] += pb[i];
  double x = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  double y = ((b[0] + b[1]) + (b[2] + b[3])) + ((b[4] + b[5]) + (b[6] + b[7]));
  x = wave_sum(x);
  y = wave_sum(y);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  if (lane == 0) {
    sh[wid] = x;
    sh[nw + wid] = y;
  }
  __syncthreads();
  x = 0.0;
  y = 0.0;
  for (int w = 0; w < nw; ++w) {
    x += sh[w];
    y += sh[nw + w];
  }
  __syncthreads();
  sa = x;
  sb = y;
}

// One Chronopoulos-Gear iteration in one launch (row blocks of SCG_NNZB = 4096 nonzeros, 512 threads:
// measured 7 % faster per launch than 8192/1024, 4096/256 and 2048/256 on the 200,198 pressure rows
// of the INS channel). The previous launch left r_i, w_i = S u_i,
// s_{i-1} and the partials of gamma_i = (r_i, u_i), delta_i = (w_i, u_i); this one forms alpha_i
// and beta_i from them and, for the rows of its CSR row block, updates
//   s_i = w_i + beta s_{i-1},  p_i = u_i + beta p_{i-1},  x += alpha p_i,
//   r_{i+1} = r_i - alpha s_i,  u_{i+1} = dinv r_{i+1}
// and computes w_{i+1} = S u_{i+1} with the partials of gamma_{i+1} and delta_{i+1}. The SpMV needs
// u_{i+1} at columns owned by other blocks, so each block recomputes it there from r_i, w_i and
// s_{i-1} (pointwise, by the same inline code as the owner); r, s and w are double-buffered by
// iteration parity, so every value a block reads was written by an earlier launch. u is never
// stored and p, x are only touched by their owner. u_0 = dinv r_0 (it = 0: beta = 0, no s_{-1}).
template <typename T>
struct ScgNext {
  T s, r, u;
};
template <typename T>
__device__ __forceinline__ ScgNext<T> scg_next(const T *__restrict__ rin, const T *__restrict__ win,
                                               const T *__restrict__ sin, const T *__restrict__ dinv, int64_t c,
                                               T a, T b, bool first) {
  ScgNext<T> o;
  o.s = first ? win[c] : win[c] + b * sin[c];
  o.r = rin[c] - a * o.s;
  o.u = dinv[c] * o.r;
  return o;
}

template <typename T, typename VT, int NNZB, int NT>
__global__ void __launch_bounds__(NT) k_scg_fused(
    int64_t nrblk, const int32_t *__restrict__ rblk, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
    const int32_t *__restrict__ ubase, const int32_t *__restrict__ ucol, const uint16_t *__restrict__ lcol,
    const VT *__restrict__ val, int it, double tol2, T *__restrict__ x, T *__restrict__ p, const T *__restrict__ rin,
    T *__restrict__ rout, const T *__restrict__ sin, T *__restrict__ sout, const T *__restrict__ win,
    T *__restrict__ wout, const T *__restrict__ dinv, const double *__restrict__ gin, int ng,
    const double *__restrict__ din, int nd, double *__restrict__ gout, double *__restrict__ dout,
    double *__restrict__ sc, int *__restrict__ flag) {
  constexpr int PER = NNZB / NT;
  __shared__ double lds[NNZB * sizeof(T) / sizeof(double)];
  T *xs = reinterpret_cast<T *>(lds);
  T *prod = xs;
  double *sh = lds;
  if (*flag) return;
  double g, d;
  sum_parts2(gin, ng, din, nd, sh, g, d);
  const double g0 = it == 0 ? g : sc[4];
  if (!(g > tol2 * g0)) {  // converged (or a zero right-hand side): the same decision in every block
    if (blockIdx.x == 0 && threadIdx.x == 0) flag[0] = 1;
    return;
  }
  double beta = 0.0, den = d;
  if (it > 0) {
    const double gp = sc[2 * ((it + 1) & 1)], ap = sc[2 * ((it + 1) & 1) + 1];
    beta = gp != 0.0 ? g / gp : 0.0;
    if (ap != 0.0) den = d - beta * g / ap;
  }
  const double alpha = den != 0.0 ? g / den : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[2 * (it & 1)] = g;
    sc[2 * (it & 1) + 1] = alpha;
    if (it == 0) sc[4] = g;
    sc[5] += 1.0;
  }
  const T a = (T)alpha, b = (T)beta;
  const bool first = it == 0;
  const int lb = xcd_block();
  double ag = 0.0, ad = 0.0;
  // the owner's update of row k, given its new S u value
  auto finish = [&](int64_t k, double sv) {
    const ScgNext<T> o = scg_next<T>(rin, win, sin, dinv, k, a, b, first);
    const T uo = dinv[k] * rin[k];
    const T pk = first ? uo : uo + b * p[k];
    p[k] = pk;
    sout[k] = o.s;
    x[k] += a * pk;
    rout[k] = o.r;
    const T wk = (T)sv;
    wout[k] = wk;
    ag += (double)o.r * (double)o.u;
    ad += (double)wk * (double)o.u;
  };
  // a grid of at most SCG_MAX_GRID blocks walks all row blocks, so the partials every block sums
  // stay few however large the matrix is
  for (int64_t rbk = lb; rbk < nrblk; rbk += gridDim.x) {
    const int32_t r0 = rblk[rbk], r1 = rblk[rbk + 1];
    const int32_t j0 = rp[r0], j1 = rp[r1];
    const int nz = j1 - j0;
    const int nrows = r1 - r0;
    __syncthreads();  // sh (the partial sums above) and the last row block's products alias the staging buffer
    if (nz <= NNZB) {
      const int ri0 = threadIdx.x >> 2, q = threadIdx.x & 3;
      int a0 = 0, e0 = 0;
      if (ri0 < nrows) {
        a0 = rp[r0 + ri0] - j0;
        e0 = rp[r0 + ri0 + 1] - j0;
      }
      int li[PER];
      T vv[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int t = threadIdx.x + k * NT;
        li[k] = -1;
        vv[k] = (T)0;
        if (t < nz) {
          li[k] = lcol[j0 + t];
          vv[k] = (T)val[j0 + t];
        }
      }
      const int u0 = ubase[rbk], nu = ubase[rbk + 1] - u0;
      for (int base = 0; base < nu; base += 2 * NT) {
        const int t0 = base + threadIdx.x, t1 = t0 + NT;
        const int32_t c0 = ucol[u0 + min(t0, nu - 1)], c1 = ucol[u0 + min(t1, nu - 1)];
        const T x0 = scg_next<T>(rin, win, sin, dinv, c0, a, b, first).u;
        const T x1 = scg_next<T>(rin, win, sin, dinv, c1, a, b, first).u;
        if (t0 < nu) xs[t0] = x0;
        if (t1 < nu) xs[t1] = x1;
      }
      __syncthreads();
      T pv[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) pv[k] = li[k] >= 0 ? vv[k] * xs[li[k]] : (T)0;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PER; ++k)
        if (li[k] >= 0) prod[threadIdx.x + k * NT] = pv[k];
      __syncthreads();
      for (int base = 0; base < nrows; base += NT / 4) {
        const int ri = base + ri0;
        int ja = a0, je = e0;
        if (base > 0 && ri < nrows) {
          ja = rp[r0 + ri] - j0;
          je = rp[r0 + ri + 1] - j0;
        }
        T sv = (T)0;  // partial row sums in the vector type, as in csr_row_block
        if (ri < nrows)
          for (int t = ja + q; t < je; t += 4) sv += prod[t];
        sv += __shfl_xor(sv, 1, 4);
        sv += __shfl_xor(sv, 2, 4);
        if (ri < nrows && q == 0) finish(r0 + ri, (double)sv);
      }
    } else {
      // a single row longer than the staging buffer
      double sv = 0.0;
      for (int t = threadIdx.x; t < nz; t += NT) {
        const int32_t c = col[j0 + t];
        sv += (double)((T)val[j0 + t] * scg_next<T>(rin, win, sin, dinv, c, a, b, first).u);
      }
      sv = block_sum(sv, sh);
      if (threadIdx.x == 0) finish(r0, sv);
    }
  }
  __syncthreads();  // sh aliases the products
  ag = block_sum(ag, sh);
  ad = block_sum(ad, sh);
  if (threadIdx.x == 0) {
    gout[lb] = ag;
    dout[lb] = ad;
  }
}

// blocks of the fused iteration: one per row block, at most SCG_MAX_GRID (a multiple of the XCD count)
constexpr int SCG_MAX_GRID = 1024;
static int scg_grid(const RowBlocks &rb) {
  const int64_t g = (std::max<int64_t>(rb.nrblk, NXCD) + NXCD - 1) / NXCD * NXCD;
  return (int)std::min<int64_t>(g, SCG_MAX_GRID);
}

// One inner Jacobi-PCG solve M x = sign s v - sub from x = 0: its matrix (values in T), row blocks
// for the fused iteration, Jacobi, vectors (8 n), scalars + partial sums, converged flag, and the
// timing classes of its two kernels.
template <typename T, typename VT>
struct PcgArgs {
  Matrix *M;
  const RowBlocks *rb;
  const VT *val;
  const T *dinv;
  T *buf;
  double *sc;
  int *flag;
  double tol;
  int c_op, c_it;
  const int *outer = nullptr;  // a converged outer FGMRES cycle (every launch of this solve returns at entry)
};

// at most K iterations: the start (x = 0, r_0, u_0, gamma_0 partials; then w_0 = M u_0 with the
// delta_0 partials) and K fused iterations
template <typename T, typename VT>
static void inner_pcg(const PcgArgs<T, VT> &q, const T *v, const double *sv, double sign, const T *sub, T *x,
                      int K) {
  Matrix *S = q.M;
  Ctx *ctx = S->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = S->n;
  const VT *val = q.val;
  const T *dinv = q.dinv;
  T *buf = q.buf;
  T *r[2] = {buf, buf + n}, *s[2] = {buf + 2 * n, buf + 3 * n}, *w[2] = {buf + 4 * n, buf + 5 * n};
  T *p = buf + 6 * n, *u = buf + 7 * n;
  const int vg = std::min(red_grid(n), 256), sg = spmv_grid<VT, T>(S);
  const RowBlocks &rb = *q.rb;
  const int fg = scg_grid(rb);
  const int pm = std::max(sg, fg);
  double *sc = q.sc;
  double *g0 = sc + SCHUR_SCAL;
  double *gq[2] = {g0 + vg, g0 + vg + pm}, *dq[2] = {g0 + vg + 2 * pm, g0 + vg + 3 * pm};
  int *flag = q.flag;
  const double tol2 = q.tol * q.tol;
  const double it_bytes = csr_bytes(S, sizeof(VT), sizeof(T)) + 10.0 * n * sizeof(T);
  k_scg_init<T><<<vg, NTHR, 0, st>>>(n, v, sv, sign, sub, dinv, x, r[0], u, g0, flag, q.outer);
  {
    KTimer tm(ctx, q.c_op, csr_bytes(S, sizeof(VT), sizeof(T)));
    spmv_dispatch<VT, T>(S, val, u, w[0], nullptr, nullptr, nullptr, nullptr, 2, dq[0], sg, st, (T)0, nullptr, (T)0,
                         flag);
  }
  // delta_0: one value (sc[6]) from the SpMV's partials, one per row block of the plain SpMV
  reduce_to(dq[0], sg, 1, sc + 6, st);
  for (int it = 0; it < K; ++it) {
    KTimer tm(ctx, q.c_it, it_bytes);
    const int a = it & 1, b = a ^ 1;
    const double *gin = it == 0 ? g0 : gq[a];
    const double *din = it == 0 ? sc + 6 : dq[a];
    const int ng = it == 0 ? vg : fg, nd = it == 0 ? 1 : fg;
    k_scg_fused<T, VT, SCG_NNZB, SCG_NT><<<fg, SCG_NT, 0, st>>>(
        rb.nrblk, rb.d_rblk.p, S->d_row_ptr.p, S->d_col.p, rb.d_ubase.p, rb.d_ucol.p, rb.d_lcol.p, val, it, tol2, x,
        p, r[a], r[b], s[a], s[b], w[a], w[b], dinv, gin, ng, din, nd, gq[b], dq[b], sc, flag);
  }
}

// scalars + partials an inner PCG on M needs (the two-stage reduction of the first SpMV's partials
// uses SEG_BLOCKS more behind them)
static size_t pcg_part_size(Matrix *M, const RowBlocks &rb) {
  const int sg = std::max(spmv_grid<double, double>(M), spmv_grid<float, float>(M));
  return (size_t)SCHUR_SCAL + red_grid(M->n) + 4 * std::max(sg, scg_grid(rb)) + SEG_BLOCKS;
}

// z[s0:s1] = -S^{-1} (s v)[s0:s1] by at most K PCG iterations from zero
template <typename T, typename VT>
static void schur_inner_t(Precond *P, const VT *val, const T *v, const double *sv, T *z, int K,
                          const int *outer) {
  Ctx *ctx = P->S->ctx;
  PcgArgs<T, VT> q;
  q.outer = outer;
  q.M = P->S;
  q.rb = &P->srb;
  q.val = val;
  if constexpr (sizeof(T) == 8) {
    q.dinv = P->sdinv.p;
    q.buf = P->svd.p;
  } else {
    q.dinv = P->sdinv32.p;
    q.buf = P->svf.p;
  }
  q.sc = P->spart.p;
  q.flag = P->sflag.p;
  q.tol = P->schur_tol;
  q.c_op = ctx->cls(sizeof(T) == 8 ? "spmv_schur_f64" : (sizeof(VT) == 2 ? "spmv_schur_bf16" : "spmv_schur_f32"));
  q.c_it = ctx->cls(sizeof(T) == 8 ? "schur_pcg_fused_f64"
                                   : (sizeof(VT) == 2 ? "schur_pcg_fused_bf16" : "schur_pcg_fused_f32"));
  inner_pcg<T, VT>(q, v + P->s0, sv, -1.0, nullptr, z + P->s0, K);
}

template <typename T>
static void schur_inner(Precond *P, const T *v, const double *sv, T *z, int K, const int *outer = nullptr) {
  if constexpr (sizeof(T) == 8) {
    schur_inner_t<double, double>(P, P->S->d_val.p, v, sv, z, K, outer);
  } else {
    if (P->bf16)
      schur_inner_t<float, bf16s>(P, P->S->d_valbf.p, v, sv, z, K, outer);
    else
      schur_inner_t<float, float>(P, P->S->d_val32.p, v, sv, z, K, outer);
  }
}

// z_p = -(s v_p) pdinv on the Schur rows
template <typename T>
__global__ void k_btri_p(int64_t ns, const T *__restrict__ pdinv, const T *__restrict__ v,
                         const double *__restrict__ sv, T *__restrict__ z, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) z[i] = (T)(-(double)pdinv[i] * (double)v[i] * sv[0]);
}

// the block-triangular step: z_p = -(s v_p) / d_p, then z_u = A_uu^{-1} (s v_u - A_up z_p)
template <typename T, typename VT>
static void btri_apply_t(Precond *P, const VT *uuval, const T *v, const double *sv, T *z, int K,
                         const int *outer) {
  Ctx *ctx = P->A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t nu = P->s0, ns = P->A->n - P->s0;
  const T *pdinv, *upval;
  PcgArgs<T, VT> q;
  q.M = P->Auu;
  q.rb = &P->vrb;
  q.val = uuval;
  if constexpr (sizeof(T) == 8) {
    pdinv = P->pdinv.p;
    upval = P->Aup->d_val.p;
    q.dinv = P->vdinv.p;
    q.buf = P->vvd.p;
  } else {
    pdinv = P->pdinv32.p;
    upval = P->Aup->d_val32.p;
    q.dinv = P->vdinv32.p;
    q.buf = P->vvf.p;
  }
  q.sc = P->vpart.p;
  q.flag = P->vflag.p;
  q.tol = P->vel_tol;
  q.c_op = ctx->cls(sizeof(T) == 8 ? "spmv_vel_f64" : (sizeof(VT) == 2 ? "spmv_vel_bf16" : "spmv_vel_f32"));
  q.c_it = ctx->cls(sizeof(T) == 8 ? "vel_pcg_fused_f64" : (sizeof(VT) == 2 ? "vel_pcg_fused_bf16" : "vel_pcg_fused_f32"));
  q.outer = outer;
  T *y = q.buf + 8 * nu;
  k_btri_p<T><<<nblk(ns), NTHR, 0, st>>>(ns, pdinv, v + nu, sv, z + nu, outer);
  {
    KTimer tm(ctx, ctx->cls(sizeof(T) == 8 ? "spmv_coupling_f64" : "spmv_coupling_f32"),
              csr_bytes(P->Aup, sizeof(T), sizeof(T)));
    spmv_dispatch<T, T>(P->Aup, upval, z + nu, y, nullptr, nullptr, nullptr, nullptr, 0, nullptr,
                        spmv_grid<T, T>(P->Aup), st, (T)0, nullptr, (T)0, outer);
  }
  inner_pcg<T, VT>(q, v, sv, 1.0, y, z, K);
}

template <typename T>
static void btri_apply(Precond *P, const T *v, const double *sv, T *z, int K, const int *outer = nullptr) {
  if constexpr (sizeof(T) == 8) {
    btri_apply_t<double, double>(P, P->Auu->d_val.p, v, sv, z, K, outer);
  } else {
    if (P->bf16)
      btri_apply_t<float, bf16s>(P, P->Auu->d_valbf.p, v, sv, z, K, outer);
    else
      btri_apply_t<float, float>(P, P->Auu->d_val32.p, v, sv, z, K, outer);
  }
}

template <typename T>
static int fgmres_cycle(Matrix *A, Matrix *K, Precond *P, bool learned, const double *r0, double beta, double *x,
                        int m, double stop_abs, int maxit, int check, Work &w) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  const int64_t n = A->n;
  const uint8_t *mask = A->space->d_free.p;
  constexpr int VW = Vec<T>::w;
  const int64_t ldv = (n + VW - 1) / VW * VW;
  DBuf<T> *Vbuf;
  if constexpr (sizeof(T) == 8) {
    w.vd.alloc(std::max<size_t>(w.vd.n, (size_t)(2 * m + 1) * ldv));
    Vbuf = &w.vd;
  } else {
    w.vf.alloc(std::max<size_t>(w.vf.n, (size_t)(2 * m + 1) * ldv));
    Vbuf = &w.vf;
  }
  T *V = Vbuf->p;
  T *Z = V + (size_t)(m + 1) * ldv;
  const size_t ns = (size_t)(m + 2) * 2 + (size_t)(m + 1) * m + 2 * m + (m + 1) + (m + 8) + m;
  w.small.alloc(std::max(w.small.n, ns));
  double *dot = w.small.p, *vs = dot + (m + 2), *R = vs + (m + 2), *csn = R + (size_t)(m + 1) * m,
         *g = csn + 2 * m, *res = g + (m + 1), *coef = res + (m + 8);
  const int grid = stream_grid((n + VW - 1) / VW);
  ensure_part(w, (size_t)(m + 2 + KC) * grid);
  k_scale_copy<double, T><<<nblk(n), NTHR, 0, st>>>(n, r0, 1.0, V);
  // the convergence flag lets the look-ahead batch's launches past the converged column return at
  // entry; it stays off while kernels are timed, so that every timed launch does its full work
  w.done.alloc(1);
  const int *skip = ctx->timing ? nullptr : w.done.p;
  k_gmres_init<<<1, 64, 0, st>>>(m, 1.0, g, vs, 1.0 / beta, w.done.p);
  double *hres = pinned(w, m + 8);
  const T *dinv;
  const T *vals;
  if constexpr (sizeof(T) == 8) {
    dinv = P->dinv.p;
    vals = K->d_val.p;
  } else {
    dinv = P->dinv32.p;
    vals = K->d_val32.p;
  }
  const int c_op = ctx->cls(sizeof(T) == 8 ? "spmv_csr_f64_fgmres" : "spmv_csr_f32_fgmres");
  const int c_pre = ctx->cls(sizeof(T) == 8 ? "fgmres_jacobi_f64" : "fgmres_jacobi_f32");
  const int c_dot = ctx->cls(sizeof(T) == 8 ? "gmres_mdot_f64" : "gmres_mdot_f32");
  const int c_upd = ctx->cls(sizeof(T) == 8 ? "gmres_update_f64" : "gmres_update_f32");
  const int limit = std::min(m, maxit);
  GState G{dot, vs, R, csn, g, res, m, beta};
  if (skip) {
    G.done = w.done.p;
    G.stop = stop_abs;
  }
  struct Batch {
    int j0, nb, ev, cap;
  };
  std::deque<Batch> inflight;
  int j = 0, bidx = 0, k_used = -1;
  // inner solve (Schur PCG, or the velocity PCG of the block-triangular step) with a tolerance:
  // each step enqueues `cap` iterations (the rest of a converged inner solve are empty launches);
  // after each batch the host reads how many ran and sets the cap to 1.3x that plus 10, doubling
  // it when a step used them all
  const bool btri = P->btri;
  const int in_iters = btri ? P->vel_iters : P->schur_iters;
  const bool adapt = in_iters > 0 && (btri ? P->vel_tol : P->schur_tol) > 0.0;
  int &in_cap = btri ? P->vel_cap : P->schur_cap;
  double &in_seen = btri ? P->vel_seen : P->schur_seen;
  const double *in_count = btri ? P->vpart.p + 5 : P->spart.p + 5;
  auto enqueue = [&]() {
    const int j_end = std::min(limit, j + check);
    const int cap = in_iters > 0 ? (adapt ? in_cap : in_iters) : 0;
    for (int jj = j; jj < j_end; ++jj) {
      T *vj = V + (int64_t)jj * ldv;
      T *zj = Z + (int64_t)jj * ldv;
      T *wv = V + (int64_t)(jj + 1) * ldv;
      if (btri) {
        btri_apply<T>(P, vj, vs + jj, zj, cap, skip);
      } else if (P->schur_amg > 0 && !learned && schur_amg_apply<T>(P, vj, vs + jj, zj, dinv, skip)) {
        P->amg_cycles += P->schur_amg;  // the Jacobi rows rode on the cycle's first launch
      } else {
        {
          KTimer tm(ctx, c_pre, 3.0 * n * sizeof(T));
          k_prec_jacobi<T><<<nblk(n), NTHR, 0, st>>>(n, dinv, vj, vs + jj, zj);
        }
        if (P->schur_amg > 0) {
          schur_amg_apply<T>(P, vj, vs + jj, zj, nullptr, skip);
          P->amg_cycles += P->schur_amg;
        } else if (P->schur_iters > 0) {
          schur_inner<T>(P, vj, vs + jj, zj, cap, skip);
        }
      }
      if (P->schur_tri && !btri) {
        // upper block-triangular velocity rows: z_u -= D^-1 A_up z_p (z_u holds D^-1 s v_u)
        const T *upval;
        if constexpr (sizeof(T) == 8)
          upval = P->Aup->d_val.p;
        else
          upval = P->Aup->d_val32.p;
        const int64_t nu = P->s0;
        KTimer tm(ctx, ctx->cls(sizeof(T) == 8 ? "schur_tri_f64" : "schur_tri_f32"),
                  (double)P->Aup->nnz * (4 + sizeof(T)) + 4.0 * (nu + 1) + 3.0 * nu * sizeof(T) +
                      (double)(P->A->n - nu) * sizeof(T));
        k_tri_rows<T><<<nblk(nu), NTHR, 0, st>>>(nu, P->Aup->d_row_ptr.p, P->Aup->d_col.p, upval, dinv, zj + nu, zj, skip);
      }
      if (learned) unet_correct<T>(P->unet, mask, vj, vs + jj, zj, 1);
      {
        KTimer tm(ctx, c_op, csr_bytes(K, sizeof(T), sizeof(T)));
        spmv_dispatch<T, T>(K, vals, zj, wv, nullptr, mask, nullptr, nullptr, 3, nullptr, spmv_grid<T, T>(K), st,
                            (T)0, nullptr, (T)0, skip);
      }
      int dgrid;
      {
        KTimer tm(ctx, c_dot, (double)(jj + 2) * n * sizeof(T));
        dgrid = mdot_launch<T>(grid, n, V, ldv, wv, jj + 1, w.part.p, st, skip);
      }
      reduce_to(w.part.p, dgrid, jj + 1, dot, st);
      {
        KTimer tm(ctx, c_upd, (double)(jj + 3) * n * sizeof(T));
        k_gs_update<T><<<grid, NTHR, 0, st>>>(n, V, ldv, wv, jj + 1, dot, vs, w.part.p, skip);
      }
      k_norm_givens<<<1, 256, 0, st>>>(w.part.p, grid, jj, G);
    }
    const int nb = j_end - j;
    OCMP_HIP(hipMemcpyAsync(hres + j, res + j, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    const int ev = bidx & 1;
    if (adapt)
      OCMP_HIP(hipMemcpyAsync(hres + m + 2 + ev, in_count, sizeof(double), hipMemcpyDeviceToHost, st));
    OCMP_HIP(hipEventRecord(work_event(w, ev), st));
    inflight.push_back({j, nb, ev, cap});
    ++bidx;
    j = j_end;
  };
  while (k_used < 0) {
    while (inflight.size() < 2 && j < limit) enqueue();
    if (inflight.empty()) {
      k_used = j;
      break;
    }
    Batch b = inflight.front();
    inflight.pop_front();
    OCMP_HIP(hipEventSynchronize(work_event(w, b.ev)));
    int conv = -1;
    for (int s = 0; s < b.nb; ++s)
      if (!(hres[b.j0 + s] > stop_abs)) {
        conv = s;
        break;
      }
    if (adapt) {
      // with the flag on, the batch's columns past the converged one ran no inner iterations
      const int nb_run = (skip && conv >= 0) ? conv + 1 : b.nb;
      const double cnt = hres[m + 2 + b.ev];
      const double per = (cnt - in_seen) / nb_run;
      in_seen = cnt;
      in_cap = per >= b.cap - 0.5 ? std::min(in_iters, 2 * b.cap) : std::min(in_iters, (int)(1.3 * per) + 10);
    }
    if (conv >= 0) k_used = b.j0 + conv + 1;
  }
  // x += Z y  (the z_j already carry the scale of v_j)
  k_gmres_solve<<<1, 256, 0, st>>>(k_used, m, R, g, nullptr, coef);
  w.u64.alloc(n);
  k_combine<T><<<nblk(n), NTHR, 0, st>>>(n, Z, ldv, k_used, coef, beta, w.u64.p);
  k_diag_apply_add<<<nblk(n), NTHR, 0, st>>>(n, nullptr, w.u64.p, x);
  OCMP_HIP(hipGetLastError());
  return k_used;
}

// FGMRES with the learned correction and its fallback: the first cycle runs with the
// learned preconditioner, the second with the plain one; the rest use whichever reduced
// the FP64 residual faster per iteration. A learned cycle that does not reduce the
// residual switches to the plain preconditioner for good.
// The operator FGMRES multiplies by: A without the entries of block pairs no term of the last
// assembly touched (structural zeros: the SpMV streams fewer bytes, same products otherwise), its
// FP64 and FP32 values gathered from A's after every assembly; A itself when every pair is live.
// The FP64 refinement residual uses it too: the dropped entries are exact zeros.
static Matrix *krylov_operator(Matrix *A, bool mixed) {
  if (A->dg) return A;
  const int nb = (int)A->space->blocks.size();
  if (nb < 2 || nb > 8) return A;
  const uint64_t all = nb * nb == 64 ? ~0ull : (1ull << (nb * nb)) - 1;
  if (A->active_pairs == 0 || (A->active_pairs & all) == all) return A;
  if (!A->op || A->op_pairs != A->active_pairs) build_solver_operator(A);
  Matrix *K = A->op;
  if (K->nnz == 0) return A;
  hipStream_t st = A->ctx->stream;
  if (mixed) {
    K->d_val32.alloc(std::max<int64_t>(K->nnz, 1));
    ensure_row_blocks32(K);
  }
  if (A->op_version != A->version) {
    k_gather_vals<<<nblk(K->nnz), NTHR, 0, st>>>(K->nnz, K->d_src.p, A->d_val.p, K->d_val.p,
                                                  mixed ? K->d_val32.p : nullptr);
    A->op_version = A->version;
    if (mixed) A->op_version32 = A->version;
  } else if (mixed && A->op_version32 != A->version) {
    k_to_f32<<<nblk(K->nnz), NTHR, 0, st>>>(K->nnz, K->d_val.p, K->d_val32.p);
    A->op_version32 = A->version;
  }
  return K;
}

static void solve_fgmres(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                         ocmp_solve_stats &st_) {
  OCMP_REQUIRE(P && P->kind == OCMP_PRE_JACOBI, "FGMRES needs a point-Jacobi (or Schur-diagonal) preconditioner");
  OCMP_REQUIRE(!A->dg, "FGMRES is for H1 systems");
  OCMP_REQUIRE(P->schur_iters == 0 || (P->S && P->sdinv.p && (o.precision != OCMP_MIXED || P->sdinv32.p)),
               "update the preconditioner after turning the Schur inner solve on");
  OCMP_REQUIRE(!P->btri || (P->Auu && P->vdinv.p && P->pdinv.p && (o.precision != OCMP_MIXED || P->vdinv32.p)),
               "update the preconditioner after turning the block-triangular step on");
  OCMP_REQUIRE(P->schur_amg == 0 || (P->amg && (o.precision != OCMP_MIXED || P->amg->lv[0]->dinv32.p)),
               "update the preconditioner after turning the Schur AMG step on");
  OCMP_REQUIRE(!P->schur_tri || (P->Aup && (o.precision != OCMP_MIXED || P->Aup->d_val32.p)),
               "update the preconditioner after turning the triangular velocity rows on");
  P->amg_cycles = 0;
  OCMP_REQUIRE(!P->bf16 || o.precision != OCMP_MIXED ||
                   ((P->schur_iters == 0 || P->btri || P->S->d_valbf.p) && (!P->btri || P->Auu->d_valbf.p)),
               "update the preconditioner after turning BF16 storage on");
  if (P->schur_iters > 0 && !P->btri) {
    k_fill_t<double><<<1, 64, 0, A->ctx->stream>>>(P->spart.p + 5, 1, 0.0);
    P->schur_seen = 0.0;
    if (P->schur_cap <= 0 || P->schur_cap > P->schur_iters) P->schur_cap = P->schur_iters;
  }
  if (P->btri) {
    k_fill_t<double><<<1, 64, 0, A->ctx->stream>>>(P->vpart.p + 5, 1, 0.0);
    P->vel_seen = 0.0;
    if (P->vel_cap <= 0 || P->vel_cap > P->vel_iters) P->vel_cap = P->vel_iters;
  }
  Work &w = work(A);
  const int64_t n = A->n;
  ensure_part(w, 0);
  w.r64.alloc(n);
  const bool mixed = o.precision == OCMP_MIXED;
  if (mixed) OCMP_REQUIRE(P->dinv32.p && A->d_val32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
  Matrix *K = krylov_operator(A, mixed);
  if (mixed && K == A && A->val32_version != A->version) refresh_val32(A, A->ctx->stream);
  const double bn = norm_masked(A, b, w);
  double rn = residual64(K, P, b, x, w.r64.p, w);
  const double bref = bn > 0.0 ? bn : (rn > 0.0 ? rn : 1.0);
  st_.initial_rel_residual = rn / bref;
  st_.rel_residual = rn / bref;
  // with a Schur inner solve a step costs milliseconds: poll every iteration so that at most a
  // step or two run past convergence
  const int check = o.check_every > 0 ? o.check_every : (P->schur_iters > 0 || P->btri ? 1 : 5);
  const int m = o.restart;
  int total = 0, outer = 0;
  const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-5;
  const bool have_learned = P->unet != nullptr;
  double rate_learned = 0.0, rate_plain = 0.0;  // log residual reduction per iteration (< 0 is progress)
  bool known_l = false, known_p = false, fell_back = false;
  // the two probe cycles (learned, then plain) are at most FGMRES_PROBE iterations long, so that a
  // correction that does not help costs little; full cycles follow with the faster one
  constexpr int FGMRES_PROBE = 40;
  while (rn > o.tol * bref && total < o.maxit && std::isfinite(rn)) {
    ++outer;
    bool use = have_learned && !fell_back;
    if (use && known_l && !known_p) use = false;          // the probe cycle with the plain preconditioner
    if (use && known_l && known_p && rate_plain < rate_learned) use = false, fell_back = true;
    const bool probe = have_learned && (!known_l || !known_p) && !fell_back;
    const int lim = probe ? std::min(o.maxit - total, std::min(m, FGMRES_PROBE)) : o.maxit - total;
    const double stop_abs = mixed ? std::max(o.tol * bref, inner * rn) : o.tol * bref;
    int used;
    if (mixed)
      used = fgmres_cycle<float>(A, K, P, use, w.r64.p, rn, x, m, stop_abs, lim, check, w);
    else
      used = fgmres_cycle<double>(A, K, P, use, w.r64.p, rn, x, m, stop_abs, lim, check, w);
    total += used;
    const double rn_old = rn;
    rn = residual64(K, P, b, x, w.r64.p, w);
    st_.rel_residual = rn / bref;
    if (used > 0 && rn > 0 && rn_old > 0) {
      const double rate = std::log(rn / rn_old) / used;
      if (use) {
        rate_learned = rate, known_l = true;
        if (!(rn < rn_old)) fell_back = true;
      } else {
        rate_plain = rate, known_p = true;
      }
    }
    if (used == 0 || (mixed && o.max_outer > 0 && outer >= o.max_outer)) break;
    if (!(rn < rn_old) && used < std::min(m, lim) && !use) break;  // no progress, the cycle thought it had converged
  }
  st_.iters = total;
  st_.outer = outer;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
  st_.fallback = have_learned && fell_back ? 1 : 0;
  if (P->schur_amg > 0) {
    st_.precond_iters = P->amg_cycles;
  } else if (P->schur_iters > 0 || P->btri) {
    double cnt = 0.0;
    OCMP_HIP(hipMemcpyAsync(&cnt, (P->btri ? P->vpart.p : P->spart.p) + 5, sizeof(double), hipMemcpyDeviceToHost,
                            A->ctx->stream));
    OCMP_HIP(hipStreamSynchronize(A->ctx->stream));
    st_.precond_iters = (int64_t)cnt;
  }
}

static void solve_gmres(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                        ocmp_solve_stats &st_) {
  Work &w = work(A);
  const int64_t n = A->n;
  ensure_part(w, 0);
  w.r64.alloc(n);
  const bool mixed = o.precision == OCMP_MIXED;
  if (mixed && P && P->kind == OCMP_PRE_JACOBI) {
    OCMP_REQUIRE(P->dinv32.p && A->d_val32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
    if (A->val32_version != A->version) refresh_val32(A, A->ctx->stream);  // an update that left it to FGMRES
  }
  if (mixed && P && P->kind == OCMP_PRE_BLOCK_JACOBI)
    OCMP_REQUIRE(P->ahat32.p, "mixed precision needs a preconditioner built with OCMP_MIXED");
  if (!mixed && P && P->kind == OCMP_PRE_BLOCK_JACOBI)
    OCMP_REQUIRE(P->ahat.p, "FP64 GMRES needs a preconditioner built with OCMP_FP64");
  if (mixed && !P) {
    // plain GMRES in FP32 needs an FP32 copy of A
    refresh_val32(A, A->ctx->stream);
    ensure_row_blocks32(A);
  }
  const double bn = norm_masked(A, b, w);
  double rn = residual64(A, P, b, x, w.r64.p, w);
  // ||b|| on free dofs is the reference; a problem driven only by Dirichlet data (b_F = 0)
  // measures against the initial residual instead
  const double bref = bn > 0.0 ? bn : (rn > 0.0 ? rn : 1.0);
  st_.initial_rel_residual = rn / bref;
  st_.rel_residual = rn / bref;
  const int check = o.check_every > 0 ? o.check_every : 5;
  const int m = o.restart;
  int total = 0, outer = 0;
  const double inner = o.inner_tol > 0 ? o.inner_tol : 1e-5;
  // s-step cycles when asked for and the restart fits the block size: on the DG block-Jacobi
  // operator, and on a CSR matrix with point-Jacobi / Schur-diagonal (or no) preconditioning
  const bool ca_dg = A->dg && P && P->kind == OCMP_PRE_BLOCK_JACOBI;
  const bool ca_csr = !A->dg && (!P || P->kind == OCMP_PRE_JACOBI);
  bool ca = o.s_step == CA_S && (ca_dg || ca_csr) && m % CA_S == 0 && m <= CA_MMAX;
  // DG: the first s-step cycle uses the shifted monomial basis (the spectrum of the folded operator
  // clusters around 1) and leaves Newton shifts for the rest. CSR: a first standard cycle records
  // its Hessenberg matrix for the shifts.
  CaShift ca_shift;
  for (int i = 0; i < CA_S; ++i) {
    ca_shift.d[i] = CA_SIGMA;
    ca_shift.c[i] = 0.0;
  }
  bool newton = false;
  while (rn > o.tol * bref && total < o.maxit && std::isfinite(rn)) {
    ++outer;
    const double stop_abs = mixed ? std::max(o.tol * bref, inner * rn) : o.tol * bref;
    int used = -1;
    if (ca && ca_csr && !newton) {
      if (mixed)
        used = gmres_cycle<float>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w, true);
      else
        used = gmres_cycle<double>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w, true);
      if (used >= 2 * CA_S) newton = newton_from_hessenberg(w.hf.p, m, used, false, mixed, ca_shift);
    }
    const bool ran_ca = ca && used < 0;
    if (ran_ca) {
      w.xbak.alloc(A->n);
      OCMP_HIP(hipMemcpyAsync(w.xbak.p, x, A->n * sizeof(double), hipMemcpyDeviceToDevice, A->ctx->stream));
      bool broke = false;
      if (mixed)
        used = ca_gmres_cycle<float>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, w, broke, ca_shift, newton);
      else
        used = ca_gmres_cycle<double>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, w, broke, ca_shift,
                                      newton);
      if (broke) ca = false, st_.fallback = 1;  // CholQR lost definiteness: standard cycles from here on
    }
    if (used < 0) {
      if (mixed)
        used = gmres_cycle<float>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w);
      else
        used = gmres_cycle<double>(A, P, w.r64.p, rn, x, m, stop_abs, o.maxit - total, check, w);
    }
    total += used;
    const double rn_old = rn;
    rn = residual64(A, P, b, x, w.r64.p, w);
    if (ran_ca && ca && used > 0 && !(rn < rn_old)) {
      // the s-step cycle did not lower the FP64 residual (its basis lost too much to rounding):
      // undo it and continue with standard cycles
      OCMP_HIP(hipMemcpyAsync(x, w.xbak.p, A->n * sizeof(double), hipMemcpyDeviceToDevice, A->ctx->stream));
      rn = residual64(A, P, b, x, w.r64.p, w);
      ca = false;
      st_.fallback = 1;
      st_.rel_residual = rn / bref;
      continue;
    }
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
  ensure_part(w, 0);
  w.r64.alloc(n);
  const double bn = norm_masked(A, b, w);
  double rn = residual64(A, P, b, x, w.r64.p, w);
  // ||b|| on free dofs is the reference; a problem driven only by Dirichlet data (b_F = 0)
  // measures against the initial residual instead
  const double bref = bn > 0.0 ? bn : (rn > 0.0 ? rn : 1.0);
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
    rn = residual64(A, P, b, x, w.r64.p, w);
  }
  st_.iters = it;
  st_.outer = 1;
  st_.rel_residual = rn / bref;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
}

// r = mask (b - A x); returns ||r||_2 (FP64)
double residual_vector(Matrix *A, const double *b, const double *x, double *r) {
  Work &w = work(A);
  return residual64(A, nullptr, b, x, r, w);
}

// ------------------------------------------------------------------ MINRES
//
// Preconditioned MINRES (Paige and Saunders, SIAM J. Numer. Anal. 12 (1975) 617-629; the recurrence
// of Elman, Silvester and Wathen, Finite Elements and Fast Iterative Solvers, Alg. 2.4) for a
// symmetric A on the free dofs, FP64. The preconditioner must be symmetric positive definite: the
// Jacobi preconditioner's |D|^{-1} (on a saddle point with the Schur-diagonal preconditioner that is
// the positive block-diagonal preconditioner diag(|A_uu|)^{-1}, |S_diag|^{-1}), the identity without
// one. The scalar recurrence runs on the host: each iteration reads two dot products. |eta| (the
// preconditioned residual norm) decides when to check the FP64 true residual; a cycle whose true
// residual is still above the tolerance restarts from the current x.

// v2 = q - a v1 - b v0; z2 = |dinv| v2 (free rows only: dinv is 0 elsewhere; no preconditioner: mask)
__global__ void k_minres_v(int64_t n, const double *__restrict__ q, const double *__restrict__ v1,
                           const double *__restrict__ v0, double a, double b, const double *__restrict__ dinv,
                           const uint8_t *__restrict__ mask, double *__restrict__ v2, double *__restrict__ z2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double v = q[i] - a * v1[i] - b * v0[i];
  v2[i] = v;
  z2[i] = dinv ? fabs(dinv[i]) * v : (mask[i] ? v : 0.0);
}

// w2 = (z / gamma - a3 w0 - a2 w1) / a1;  x += t w2
__global__ void k_minres_w(int64_t n, const double *__restrict__ z, double ig, const double *__restrict__ w0,
                           const double *__restrict__ w1, double a1, double a2, double a3, double *__restrict__ w2,
                           double t, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double w = (z[i] * ig - a3 * w0[i] - a2 * w1[i]) / a1;
  w2[i] = w;
  x[i] += t * w;
}

static void solve_minres(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o,
                         ocmp_solve_stats &st_) {
  Ctx *ctx = A->ctx;
  hipStream_t st = ctx->stream;
  Work &w = work(A);
  const int64_t n = A->n;
  OCMP_REQUIRE(o.precision == OCMP_FP64, "MinRes runs in FP64");
  OCMP_REQUIRE(!P || P->kind == OCMP_PRE_JACOBI || P->kind == OCMP_PRE_NONE,
               "MinRes needs a point-Jacobi (or Schur-diagonal) preconditioner, or none");
  OCMP_REQUIRE(!P || (P->schur_iters == 0 && !P->btri), "MinRes needs a fixed preconditioner (no inner solve)");
  const double *dinv = P && P->kind == OCMP_PRE_JACOBI ? P->dinv.p : nullptr;
  const uint8_t *mask = A->space->d_free.p;
  ensure_part(w, 0);
  w.r64.alloc(n);
  w.vd.alloc(std::max<size_t>(w.vd.n, (size_t)10 * n));
  double *base = w.vd.p;
  double *V[3] = {base, base + n, base + 2 * n};
  double *Wv[3] = {base + 3 * n, base + 4 * n, base + 5 * n};
  double *Z[2] = {base + 6 * n, base + 7 * n};
  double *xs = base + 8 * n, *aq = base + 9 * n;
  const double bn = norm_masked(A, b, w);
  double rn = residual64(A, P, b, x, w.r64.p, w);
  const double bref = bn > 0.0 ? bn : (rn > 0.0 ? rn : 1.0);
  st_.initial_rel_residual = rn / bref;
  st_.rel_residual = rn / bref;
  const int grid = red_grid(n);
  double *sc = ctx->scal.p;
  double *h = pinned(w, 8);
  auto dot = [&](const double *a, const double *c) {
    k_dot<double><<<grid, NTHR, 0, st>>>(n, a, c, mask, w.part.p);
    reduce_to(w.part.p, grid, 1, sc + 102, st);
    OCMP_HIP(hipMemcpyAsync(h, sc + 102, sizeof(double), hipMemcpyDeviceToHost, st));
    OCMP_HIP(hipStreamSynchronize(st));
    return h[0];
  };
  const int sgrid = spmv_grid(A);
  int total = 0, outer = 0;
  while (rn > o.tol * bref && total < o.maxit && std::isfinite(rn)) {
    ++outer;
    // v1 = r (the true residual, masked), z = M v1, gamma1 = sqrt(v1.z); v0 = w0 = w1 = 0
    k_fill_t<double><<<nblk(n), NTHR, 0, st>>>(V[0], n, 0.0);
    k_fill_t<double><<<nblk(2 * n), NTHR, 0, st>>>(Wv[0], 2 * n, 0.0);
    k_minres_v<<<nblk(n), NTHR, 0, st>>>(n, w.r64.p, V[0], V[0], 0.0, 0.0, dinv, mask, V[1], Z[0]);
    double g1 = std::sqrt(std::max(dot(V[1], Z[0]), 0.0)), g0 = 1.0;
    if (!(g1 > 0.0)) break;
    double eta = g1, c0 = 1.0, c1 = 1.0, s0 = 0.0, s1 = 0.0;
    const double eta0 = g1;
    // the true residual is about rn |eta| / eta0: stop the cycle at half the tolerance
    const double eta_stop = 0.5 * o.tol * bref / rn * eta0;
    int it = 0, zc = 0;
    double *v0 = V[0], *v1 = V[1], *v2 = V[2], *w0 = Wv[0], *w1 = Wv[1], *w2 = Wv[2];
    while (total + it < o.maxit) {
      double *z = Z[zc], *zn = Z[zc ^ 1];
      k_scale_copy<double, double><<<nblk(n), NTHR, 0, st>>>(n, z, 1.0 / g1, xs);
      spmv_dispatch<double, double>(A, A->d_val.p, xs, aq, nullptr, mask, nullptr, nullptr, 3, nullptr, sgrid, st);
      const double delta = dot(aq, xs);
      k_minres_v<<<nblk(n), NTHR, 0, st>>>(n, aq, v1, v0, delta / g1, it == 0 ? 0.0 : g1 / g0, dinv, mask, v2, zn);
      const double g2 = std::sqrt(std::max(dot(v2, zn), 0.0));
      const double a0 = c1 * delta - c0 * s1 * g1;
      const double a1 = std::sqrt(a0 * a0 + g2 * g2);
      const double a2 = s1 * delta + c0 * c1 * g1;
      const double a3 = s0 * g1;
      if (!(a1 > 0.0)) break;
      const double c2 = a0 / a1, s2 = g2 / a1;
      k_minres_w<<<nblk(n), NTHR, 0, st>>>(n, z, 1.0 / g1, w0, w1, a1, a2, a3, w2, c2 * eta, x);
      eta = -s2 * eta;
      ++it;
      double *tv = v0;
      v0 = v1, v1 = v2, v2 = tv;
      double *tw = w0;
      w0 = w1, w1 = w2, w2 = tw;
      zc ^= 1;
      g0 = g1, g1 = g2, c0 = c1, c1 = c2, s0 = s1, s1 = s2;
      if (!(g1 > 0.0) || std::fabs(eta) <= eta_stop) break;
    }
    total += it;
    const double rn_old = rn;
    rn = residual64(A, P, b, x, w.r64.p, w);
    st_.rel_residual = rn / bref;
    if (it == 0 || !(rn < rn_old)) break;
  }
  st_.iters = total;
  st_.outer = outer;
  st_.converged = st_.rel_residual <= o.tol ? 1 : 0;
}

// out[0] = ||b - A x||_2, out[1] = ||b||_2, both over the free dofs in FP64
// the matrix a residual b - A x is formed with: the Krylov operator when it exists for A's current
// block pairs (its FP64 values, and the FP32 copy if it has one, refreshed in place: no allocation),
// else A. The operator leaves out only exact zeros, so the residual is A's up to summation order.
static Matrix *residual_operator(Matrix *A) {
  Matrix *K = A->op;
  if (A->dg || !K || A->op_pairs != A->active_pairs || K->nnz == 0 || K->rb64.nnzb == 0) return A;
  const int nb = (int)A->space->blocks.size();
  const uint64_t all = nb * nb == 64 ? ~0ull : (1ull << (nb * nb)) - 1;
  if (A->active_pairs == 0 || (A->active_pairs & all) == all) return A;
  if (A->op_version != A->version) {
    const bool f32 = K->d_val32.p && K->d_val32.n == (size_t)std::max<int64_t>(K->nnz, 1);
    k_gather_vals<<<nblk(K->nnz), NTHR, 0, A->ctx->stream>>>(K->nnz, K->d_src.p, A->d_val.p, K->d_val.p,
                                                              f32 ? K->d_val32.p : nullptr);
    A->op_version = A->version;
    if (f32) A->op_version32 = A->version;
  }
  return K;
}

void residual_norms(Matrix *A, const double *b, const double *x, double *out) {
  Work &w = work(A);
  w.r64.alloc(A->n);
  out[0] = residual64(residual_operator(A), nullptr, b, x, w.r64.p, w);
  out[1] = norm_masked(A, b, w);
}

void solve(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o, ocmp_solve_stats &st) {
  OCMP_REQUIRE(!P || (P->schur_iters == 0 && P->schur_amg == 0) || o.solver == OCMP_SOLVER_FGMRES,
               "the Schur inner solve and the Schur AMG step run inside FGMRES's preconditioner step");
  switch (o.solver) {
    case OCMP_SOLVER_CG: solve_cg(A, P, b, x, o, st); break;
    case OCMP_SOLVER_GMRES: solve_gmres(A, P, b, x, o, st); break;
    case OCMP_SOLVER_RICHARDSON: solve_richardson(A, P, b, x, o, st); break;
    case OCMP_SOLVER_FGMRES: solve_fgmres(A, P, b, x, o, st); break;
    case OCMP_SOLVER_MINRES: solve_minres(A, P, b, x, o, st); break;
    default: throw Error(OCMP_ERR_UNSUPPORTED, "unknown solver");
  }
}

// ------------------------------------------------------------------ allocation-free entry points
//
// Callers that may not allocate device memory (code under stream capture, serving loops,
// harnesses that forbid hipMalloc) get variants whose partial sums live in one static device
// array. Each GPU the library runs on holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for that device (the array's
// name is a host-side stand-in, not a device address). Calls on one GPU share the array, so
// calls that overlap on one GPU race; every call writes the partial sums it reads.

constexpr int64_t NOALLOC_SCRATCH_DOUBLES = (int64_t)1 << 19;  // 4 MiB per device
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_addr;  // device -> address of that device's copy

double *noalloc_scratch_init(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  auto it = g_scratch_addr.find(device);
  if (it != g_scratch_addr.end()) return it->second;
  int cur = -1;
  OCMP_HIP(hipGetDevice(&cur));
  OCMP_REQUIRE(cur == device, "noalloc_scratch_init: the device must be current");
  void *ptr = nullptr;
  OCMP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  OCMP_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address for the scratch array");
  g_scratch_addr[device] = static_cast<double *>(ptr);
  return static_cast<double *>(ptr);
}

// out_dev[0] = ||b - A x||_2, out_dev[1] = ||b||_2 over the free dofs (FP64), written on the
// context stream without allocating or synchronising. The same kernels as residual_norms,
// except that the residual vector is not stored (mode 1 of the SpMV with y = nullptr).
void residual_norms_noalloc(Matrix *A0, const double *b, const double *x, double *out_dev) {
  Matrix *A = residual_operator(A0);
  Ctx *ctx = A->ctx;
  double *sp = ctx->noalloc_scratch;
  OCMP_REQUIRE(sp != nullptr, "the context has no allocation-free scratch");
  OCMP_REQUIRE(A->rb64.nnzb != 0, "the matrix has no SpMV row blocks");
  hipStream_t st = ctx->stream;
  const int grid = spmv_grid(A);
  const int dgrid = red_grid(A->n);
  // [0, grid): SpMV partials; [grid, grid + SEG_BLOCKS): second-stage sums of a long reduction;
  // then dgrid partials of ||b||^2
  const int64_t need = (int64_t)grid + SEG_BLOCKS + dgrid;
  if (need > NOALLOC_SCRATCH_DOUBLES)
    throw Error(OCMP_ERR_ARG, "residual_norms_noalloc: the matrix needs " + std::to_string(need) +
                                  " doubles of scratch, more than the " + std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                                  " of the static array (at most about " +
                                  std::to_string((NOALLOC_SCRATCH_DOUBLES - SEG_BLOCKS - MAX_RED_BLOCKS) * SPMV_NNZB64) +
                                  " nonzeros)");
  spmv_dispatch<double, double>(A, A->d_val.p, x, (double *)nullptr, b, A->space->d_free.p, (const double *)nullptr,
                                nullptr, 1, sp, grid, st);
  reduce_to(sp, grid, 1, out_dev, st, true);
  double *bp = sp + grid + SEG_BLOCKS;
  k_dot<double><<<dgrid, NTHR, 0, st>>>(A->n, b, b, A->space->d_free.p, bp);
  k_reduce<<<1, NTHR, 0, st>>>(bp, dgrid, out_dev + 1, 1);
  OCMP_HIP(hipGetLastError());
}

}  // namespace ocmp
