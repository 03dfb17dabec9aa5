// Per-DOF MLP initial guess for transient solves: inference and training kernels.
//
// Network: x in R^4 -> 128 -> 128 -> 128 -> 1 with ReLU, weights stored in bf16,
// every sum accumulated in fp32.  The two 128x128 layers run on MFMA
// (v_mfma_f32_16x16x16_bf16) in a transposed formulation: a wave holds the
// activations of 32 rows (DOF samples) as the B operand (K = input feature,
// N = row) and the weight tile as the A operand (M = output feature), so the
// 16x16 fp32 result tile of output block m is, lane for lane, the B operand of
// K-block m of the next layer: activations never leave registers between layers.
//
// Training (one Adam step on a batch of B samples, loss 0.5/B sum (y-t)^2):
//   k_mlp_fwd<1>   forward, stores a1..a3 (bf16)
//   k_mlp_bwd      dz3 = (d4*W4) o [a3>0], dz2 = (W3^T dz3) o [a2>0], dz1 = (W2^T dz2) o [a1>0]
//                  (dz rounded to bf16: they are MFMA operands and are stored)
//   k_mlp_wgrad    gW2 = dz2^T a1, gW3 = dz3^T a2 on MFMA, per-block partials (K = rows)
//   k_mlp_small    gW1, b1..b3, gW4, b4 and the loss, per-block partials (VALU)
//   k_mlp_adam     fixed-order sum of partials, Adam update, bf16 mirror of the weights
// Every reduction has a fixed order, so results are bitwise repeatable.
// This component has no counterpart in the reference (it feeds the initial
// guess of the Oseen/implicit-Euler step, opencmp/models/ins.py:323-359).
#include "internal.h"

#include <cmath>

namespace ocmp {
namespace {

constexpr int H = OCMP_MLP_HIDDEN, NIN = 4;
constexpr int OW1 = 0, OB1 = OW1 + H * NIN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + H, NP = OB4 + 1;
static_assert(NP == OCMP_MLP_NPARAM, "parameter layout");
constexpr int LDW = H + 4;       // padded LDS row of a 128x128 bf16 weight tile (264 B: 8-byte aligned)
constexpr int ROWS_BLK = 128;    // rows per block iteration: 4 waves x 32 rows
constexpr int CH = 64;           // rows per staged chunk of the weight-gradient kernel
constexpr int LDC = CH + 4;      // padded LDS row (136 B)
constexpr int NSMALL = H * NIN + 4 * H + 2;  // gW1, gb1, gb2, gb3, gW4, gb4, loss
constexpr int MAXSEG = OCMP_MLP_MAXSEG;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

__device__ __forceinline__ u16 f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ float bfs(short h) { return bf2f((u16)h); }

struct SegArgs {
  int nseg;
  int64_t off[MAXSEG + 1];
  float s[MAXSEG], U[MAXSEG];
};

struct GuessArgs {
  const double *u0, *u1, *u2, *u3;
  double *out;
  int base;  // 1: the network predicts the correction to the linear extrapolation 2 u0 - u1
};

__device__ __forceinline__ int seg_of(const SegArgs &sg, int64_t i) {
  int k = 0;
  while (k + 1 < sg.nseg && i >= sg.off[k + 1]) ++k;
  return k;
}

// features of one DOF from its last four states: scaled increments and the scaled state
__device__ __forceinline__ float4 features(double u0, double u1, double u2, double u3, float s, float U) {
  const double is = 1.0 / (double)s;
  return make_float4((float)((u0 - u1) * is), (float)((u1 - u2) * is), (float)((u2 - u3) * is),
                     (float)(u0 / (double)U));
}

// copy a 128x128 bf16 weight (row-major [out][in]) into padded LDS, transposed on request
__device__ void stage_w(u16 *lds, const u16 *__restrict__ w, bool transpose) {
  for (int e = threadIdx.x * 4; e < H * H; e += blockDim.x * 4) {
    const int o = e / H, i = e % H;
    const ushort4 v = *(const ushort4 *)(w + e);
    if (!transpose) {
      *(ushort4 *)(lds + o * LDW + i) = v;
    } else {
      lds[(i + 0) * LDW + o] = v.x;
      lds[(i + 1) * LDW + o] = v.y;
      lds[(i + 2) * LDW + o] = v.z;
      lds[(i + 3) * LDW + o] = v.w;
    }
  }
}

// acc[m][nt] = W[16m.., :] * act[:][nt] (+ bias); act holds the B operand of every K block
__device__ __forceinline__ void dense_mfma(const u16 *sW, const s4 (&act)[8][2], f4 (&acc)[8][2], int r, int g) {
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const s4 a = *(const s4 *)(sW + (16 * m + r) * LDW + 16 * kk + 4 * g);
      acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[kk][0], acc[m][0], 0, 0, 0);
      acc[m][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, act[kk][1], acc[m][1], 0, 0, 0);
    }
  }
}

__device__ __forceinline__ s4 relu_bf(const f4 v) {
  s4 o;
  o.x = (short)f2bf(fmaxf(v.x, 0.f));
  o.y = (short)f2bf(fmaxf(v.y, 0.f));
  o.z = (short)f2bf(fmaxf(v.z, 0.f));
  o.w = (short)f2bf(fmaxf(v.w, 0.f));
  return o;
}

// MODE 0: Y = f(X); MODE 1: same and store a1..a3;
// MODE 2: out = u0 + base * (u0 - u1) + s * f(features(u0..u3))
template <int MODE>
__global__ __launch_bounds__(256) void k_mlp_fwd(const u16 *__restrict__ pb, int64_t n, int64_t npad,
                                                 const float4 *__restrict__ X, float *__restrict__ Y,
                                                 u16 *__restrict__ act, SegArgs sg, GuessArgs ga) {
  __shared__ u16 sW2[H * LDW], sW3[H * LDW];
  __shared__ float sW1[H * NIN], sB1[H], sB2[H], sB3[H], sW4[H], sB4;
  stage_w(sW2, pb + OW2, false);
  stage_w(sW3, pb + OW3, false);
  for (int e = threadIdx.x; e < H * NIN; e += blockDim.x) sW1[e] = bf2f(pb[OW1 + e]);
  for (int e = threadIdx.x; e < H; e += blockDim.x) {
    sB1[e] = bf2f(pb[OB1 + e]);
    sB2[e] = bf2f(pb[OB2 + e]);
    sB3[e] = bf2f(pb[OB3 + e]);
    sW4[e] = bf2f(pb[OW4 + e]);
  }
  if (threadIdx.x == 0) sB4 = bf2f(pb[OB4]);
  __syncthreads();

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
  for (int64_t base = (int64_t)blockIdx.x * ROWS_BLK; base < npad; base += (int64_t)gridDim.x * ROWS_BLK) {
    const int64_t rb = base + wv * 32;
    s4 act_[8][2];
    float su[2] = {1.f, 1.f};
    double u0v[2] = {0.0, 0.0}, u1v[2] = {0.0, 0.0};
    // layer 1 (4 -> 128) on VALU, written straight into the B-operand layout
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const int64_t row = rb + 16 * nt + r;
      float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
      if (row < n) {
        if (MODE == 2) {
          const int k = seg_of(sg, row);
          u0v[nt] = ga.u0[row];
          u1v[nt] = ga.u1[row];
          su[nt] = sg.s[k];
          x = features(u0v[nt], u1v[nt], ga.u2[row], ga.u3[row], sg.s[k], sg.U[k]);
        } else {
          x = X[row];
        }
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        float z[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int o = 16 * m + 4 * g + i;
          z[i] = sB1[o] + sW1[o * 4 + 0] * x.x + sW1[o * 4 + 1] * x.y + sW1[o * 4 + 2] * x.z + sW1[o * 4 + 3] * x.w;
        }
        act_[m][nt] = relu_bf(f4{z[0], z[1], z[2], z[3]});
        if (MODE == 1) *(s4 *)(act + (size_t)row * H + 16 * m + 4 * g) = act_[m][nt];
      }
    }
    // layers 2 and 3 on MFMA
#pragma unroll
    for (int L = 0; L < 2; ++L) {
      const u16 *sW = L == 0 ? sW2 : sW3;
      const float *sB = L == 0 ? sB2 : sB3;
      f4 acc[8][2];
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        const f4 b = {sB[16 * m + 4 * g], sB[16 * m + 4 * g + 1], sB[16 * m + 4 * g + 2], sB[16 * m + 4 * g + 3]};
        acc[m][0] = b;
        acc[m][1] = b;
      }
      dense_mfma(sW, act_, acc, r, g);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          act_[m][nt] = relu_bf(acc[m][nt]);
          if (MODE == 1)
            *(s4 *)(act + (size_t)(L + 1) * npad * H + (size_t)(rb + 16 * nt + r) * H + 16 * m + 4 * g) =
                act_[m][nt];
        }
      }
    }
    // layer 4 (128 -> 1): lane partial over its 32 features, then across the 4 lane groups
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      float p = 0.f;
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        p += sW4[16 * m + 4 * g + 0] * bfs(act_[m][nt].x);
        p += sW4[16 * m + 4 * g + 1] * bfs(act_[m][nt].y);
        p += sW4[16 * m + 4 * g + 2] * bfs(act_[m][nt].z);
        p += sW4[16 * m + 4 * g + 3] * bfs(act_[m][nt].w);
      }
      p += __shfl_xor(p, 16);
      p += __shfl_xor(p, 32);
      const int64_t row = rb + 16 * nt + r;
      if (g == 0 && row < n) {
        const float y = p + sB4;
        if (MODE == 2)
          ga.out[row] = u0v[nt] + (double)ga.base * (u0v[nt] - u1v[nt]) + (double)su[nt] * (double)y;
        else
          Y[row] = y;
      }
    }
  }
}

// backward through layers 4..1 for 32 rows per wave; writes dz1..dz3 (bf16) and d4 = dL/dy
__global__ __launch_bounds__(256) void k_mlp_bwd(const u16 *__restrict__ pb, int64_t n, int64_t npad,
                                                 const float *__restrict__ Y, const float *__restrict__ T, float invB,
                                                 const u16 *__restrict__ act, u16 *__restrict__ dz,
                                                 float *__restrict__ d4) {
  __shared__ u16 sW2T[H * LDW], sW3T[H * LDW];
  __shared__ float sW4[H];
  stage_w(sW2T, pb + OW2, true);
  stage_w(sW3T, pb + OW3, true);
  for (int e = threadIdx.x; e < H; e += blockDim.x) sW4[e] = bf2f(pb[OW4 + e]);
  __syncthreads();

  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
  const size_t LS = (size_t)npad * H;
  for (int64_t base = (int64_t)blockIdx.x * ROWS_BLK; base < npad; base += (int64_t)gridDim.x * ROWS_BLK) {
    const int64_t rb = base + wv * 32;
    s4 d_[8][2];
    // dz3 = (d4 * W4) o [a3 > 0]
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
      const int64_t row = rb + 16 * nt + r;
      const float d = row < n ? (Y[row] - T[row]) * invB : 0.f;
      if (g == 0) d4[row] = d;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) {
        const s4 a = *(const s4 *)(act + 2 * LS + (size_t)row * H + 16 * kk + 4 * g);
        const int o = 16 * kk + 4 * g;
        s4 v;
        v.x = bfs(a.x) > 0.f ? (short)f2bf(d * sW4[o + 0]) : (short)0;
        v.y = bfs(a.y) > 0.f ? (short)f2bf(d * sW4[o + 1]) : (short)0;
        v.z = bfs(a.z) > 0.f ? (short)f2bf(d * sW4[o + 2]) : (short)0;
        v.w = bfs(a.w) > 0.f ? (short)f2bf(d * sW4[o + 3]) : (short)0;
        d_[kk][nt] = v;
        *(s4 *)(dz + 2 * LS + (size_t)row * H + o) = v;
      }
    }
    // dz2 = (W3^T dz3) o [a2 > 0], then dz1 = (W2^T dz2) o [a1 > 0]
#pragma unroll
    for (int L = 1; L >= 0; --L) {
      const u16 *sWT = L == 1 ? sW3T : sW2T;
      f4 acc[8][2];
#pragma unroll
      for (int m = 0; m < 8; ++m) acc[m][0] = acc[m][1] = f4{0.f, 0.f, 0.f, 0.f};
      dense_mfma(sWT, d_, acc, r, g);
#pragma unroll
      for (int m = 0; m < 8; ++m) {
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
          const int64_t row = rb + 16 * nt + r;
          const size_t at = (size_t)L * LS + (size_t)row * H + 16 * m + 4 * g;
          const s4 a = *(const s4 *)(act + at);
          s4 v;
          v.x = bfs(a.x) > 0.f ? (short)f2bf(acc[m][nt].x) : (short)0;
          v.y = bfs(a.y) > 0.f ? (short)f2bf(acc[m][nt].y) : (short)0;
          v.z = bfs(a.z) > 0.f ? (short)f2bf(acc[m][nt].z) : (short)0;
          v.w = bfs(a.w) > 0.f ? (short)f2bf(acc[m][nt].w) : (short)0;
          d_[m][nt] = v;
          *(s4 *)(dz + at) = v;
        }
      }
    }
  }
}

// gW_{L} = dz_L^T a_{L-1} for the two 128x128 layers (blockIdx.y = 0: W2, 1: W3); K = rows.
// Block b sums rows [b*rpb, (b+1)*rpb); wave w owns output rows 32w..32w+31.
__global__ __launch_bounds__(256) void k_mlp_wgrad(int64_t npad, int64_t rpb, const u16 *__restrict__ act,
                                                   const u16 *__restrict__ dz, float *__restrict__ part) {
  __shared__ u16 sD[H * LDC], sA[H * LDC];
  const int L = blockIdx.y;
  const size_t LS = (size_t)npad * H;
  const u16 *D = dz + (size_t)(L + 1) * LS;
  const u16 *A = act + (size_t)L * LS;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
  f4 acc[2][8];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nn = 0; nn < 8; ++nn) acc[mi][nn] = f4{0.f, 0.f, 0.f, 0.f};
  const int64_t r0 = (int64_t)blockIdx.x * rpb, r1 = r0 + rpb < npad ? r0 + rpb : npad;
  for (int64_t c0 = r0; c0 < r1; c0 += CH) {
    __syncthreads();
    // stage 64 rows of dz_L and a_{L-1}, transposed to [feature][row]
    for (int e = threadIdx.x * 4; e < CH * H; e += blockDim.x * 4) {
      const int rr = e / H, f = e % H;
      const ushort4 dv = *(const ushort4 *)(D + (size_t)(c0 + rr) * H + f);
      const ushort4 av = *(const ushort4 *)(A + (size_t)(c0 + rr) * H + f);
      sD[(f + 0) * LDC + rr] = dv.x;
      sD[(f + 1) * LDC + rr] = dv.y;
      sD[(f + 2) * LDC + rr] = dv.z;
      sD[(f + 3) * LDC + rr] = dv.w;
      sA[(f + 0) * LDC + rr] = av.x;
      sA[(f + 1) * LDC + rr] = av.y;
      sA[(f + 2) * LDC + rr] = av.z;
      sA[(f + 3) * LDC + rr] = av.w;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < CH / 16; ++ks) {
      s4 a[2];
#pragma unroll
      for (int mi = 0; mi < 2; ++mi) a[mi] = *(const s4 *)(sD + (32 * wv + 16 * mi + r) * LDC + 16 * ks + 4 * g);
#pragma unroll
      for (int nn = 0; nn < 8; ++nn) {
        const s4 b = *(const s4 *)(sA + (16 * nn + r) * LDC + 16 * ks + 4 * g);
        acc[0][nn] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[0], b, acc[0][nn], 0, 0, 0);
        acc[1][nn] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[1], b, acc[1][nn], 0, 0, 0);
      }
    }
  }
  float *P = part + ((size_t)blockIdx.x * 2 + L) * H * H;
#pragma unroll
  for (int mi = 0; mi < 2; ++mi)
#pragma unroll
    for (int nn = 0; nn < 8; ++nn) {
      const int o = 32 * wv + 16 * mi + 4 * g, in = 16 * nn + r;
      P[(o + 0) * H + in] = acc[mi][nn].x;
      P[(o + 1) * H + in] = acc[mi][nn].y;
      P[(o + 2) * H + in] = acc[mi][nn].z;
      P[(o + 3) * H + in] = acc[mi][nn].w;
    }
}

// per-block partials of gW1 (dz1 x), gb1..gb3, gW4 (d4 a3), gb4 and sum (y-t)^2
__global__ __launch_bounds__(256) void k_mlp_small(int64_t n, int64_t npad, int64_t rpb,
                                                   const float4 *__restrict__ X, const float *__restrict__ Y,
                                                   const float *__restrict__ T, const float *__restrict__ d4,
                                                   const u16 *__restrict__ act, const u16 *__restrict__ dz,
                                                   float *__restrict__ part) {
  __shared__ float sh[128 * 8 + 2];
  const size_t LS = (size_t)npad * H;
  const int o = threadIdx.x & 127, half = threadIdx.x >> 7;
  float w1[4] = {0.f, 0.f, 0.f, 0.f}, b1 = 0.f, b2 = 0.f, b3 = 0.f, w4 = 0.f, b4 = 0.f, ls = 0.f;
  const int64_t r0 = (int64_t)blockIdx.x * rpb, r1 = r0 + rpb < n ? r0 + rpb : n;
  for (int64_t row = r0 + half; row < r1; row += 2) {
    const float4 x = X[row];
    const float z1 = bf2f(dz[(size_t)row * H + o]);
    w1[0] += z1 * x.x;
    w1[1] += z1 * x.y;
    w1[2] += z1 * x.z;
    w1[3] += z1 * x.w;
    b1 += z1;
    b2 += bf2f(dz[LS + (size_t)row * H + o]);
    b3 += bf2f(dz[2 * LS + (size_t)row * H + o]);
    const float d = d4[row];
    w4 += d * bf2f(act[2 * LS + (size_t)row * H + o]);
    if (o == 0) {
      b4 += d;
      const float e = Y[row] - T[row];
      ls += e * e;
    }
  }
  if (half == 1) {
    float *q = sh + o * 8;
    q[0] = w1[0]; q[1] = w1[1]; q[2] = w1[2]; q[3] = w1[3];
    q[4] = b1; q[5] = b2; q[6] = b3; q[7] = w4;
    if (o == 0) { sh[1024] = b4; sh[1025] = ls; }
  }
  __syncthreads();
  if (half == 0) {
    const float *q = sh + o * 8;
    float *P = part + (size_t)blockIdx.x * NSMALL;
    for (int c = 0; c < 4; ++c) P[o * 4 + c] = w1[c] + q[c];
    P[H * NIN + o] = b1 + q[4];
    P[H * NIN + H + o] = b2 + q[5];
    P[H * NIN + 2 * H + o] = b3 + q[6];
    P[H * NIN + 3 * H + o] = w4 + q[7];
    if (o == 0) {
      P[H * NIN + 4 * H] = b4 + sh[1024];
      P[H * NIN + 4 * H + 1] = ls + sh[1025];
    }
  }
}

// fixed-order sum of the partials into the gradient, optional Adam step and bf16 mirror
__global__ __launch_bounds__(256) void k_mlp_adam(int nbw, int nbs, const float *__restrict__ pw,
                                                  const float *__restrict__ ps, float *__restrict__ grad, float *p,
                                                  float *m, float *v, u16 *pb, float lr, float c1, float c2,
                                                  int update, double *loss, double half_invB) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > NP) return;
  if (j == NP) {  // loss
    double s = 0.0;
    for (int b = 0; b < nbs; ++b) s += (double)ps[(size_t)b * NSMALL + NSMALL - 1];
    *loss = s * half_invB;
    return;
  }
  float gsum = 0.f;
  if (j >= OW2 && j < OB2) {
    for (int b = 0; b < nbw; ++b) gsum += pw[((size_t)b * 2 + 0) * H * H + (j - OW2)];
  } else if (j >= OW3 && j < OB3) {
    for (int b = 0; b < nbw; ++b) gsum += pw[((size_t)b * 2 + 1) * H * H + (j - OW3)];
  } else {
    int k;
    if (j < OB1) k = j - OW1;
    else if (j < OW2) k = H * NIN + (j - OB1);
    else if (j < OW3) k = H * NIN + H + (j - OB2);
    else if (j < OW4) k = H * NIN + 2 * H + (j - OB3);
    else if (j < OB4) k = H * NIN + 3 * H + (j - OW4);
    else k = H * NIN + 4 * H;
    for (int b = 0; b < nbs; ++b) gsum += ps[(size_t)b * NSMALL + k];
  }
  if (grad) grad[j] = gsum;
  if (update) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mj = b1 * m[j] + (1.f - b1) * gsum;
    const float vj = b2 * v[j] + (1.f - b2) * gsum * gsum;
    m[j] = mj;
    v[j] = vj;
    const float pj = p[j] - lr * (mj * c1) / (sqrtf(vj * c2) + eps);
    p[j] = pj;
    pb[j] = f2bf(pj);
  }
}

__global__ void k_mlp_mirror(const float *__restrict__ p, u16 *__restrict__ pb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < NP) pb[j] = f2bf(p[j]);
}

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// random (snapshot j, dof i) samples: X = features of u^j..u^{j-3},
// T = (u^{j+1} - u^j - base (u^j - u^{j-1})) / s
__global__ void k_mlp_sample(const double *__restrict__ snap, int nsnap, int64_t ld, int64_t n, SegArgs sg,
                             uint64_t seed, int batch, int base, float4 *__restrict__ X,
                             float *__restrict__ T) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const uint64_t h1 = splitmix(seed * 0x2545F4914F6CDD1Dull + (uint64_t)b);
  const uint64_t h2 = splitmix(h1);
  const int j = 3 + (int)(h1 % (uint64_t)(nsnap - 4));
  const int64_t i = (int64_t)(h2 % (uint64_t)n);
  const int k = seg_of(sg, i);
  const double u0 = snap[(size_t)j * ld + i], u1 = snap[(size_t)(j - 1) * ld + i];
  X[b] = features(u0, u1, snap[(size_t)(j - 2) * ld + i], snap[(size_t)(j - 3) * ld + i], sg.s[k], sg.U[k]);
  T[b] = (float)((snap[(size_t)(j + 1) * ld + i] - u0 - (double)base * (u0 - u1)) / (double)sg.s[k]);
}

// per (segment, block) sums of squared increments and squared values over the snapshots
__global__ __launch_bounds__(256) void k_mlp_scales(const double *__restrict__ snap, int nsnap, int64_t ld,
                                                    SegArgs sg, double *__restrict__ part) {
  __shared__ double sh[2][256];
  const int k = blockIdx.y;
  const int64_t o = sg.off[k], len = sg.off[k + 1] - sg.off[k];
  const int64_t tot = (int64_t)(nsnap - 1) * len;
  double si = 0.0, sv = 0.0;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int j = 1 + (int)(q / len);
    const int64_t i = o + q % len;
    const double u = snap[(size_t)j * ld + i], d = u - snap[(size_t)(j - 1) * ld + i];
    si += d * d;
    sv += u * u;
  }
  sh[0][threadIdx.x] = si;
  sh[1][threadIdx.x] = sv;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[((size_t)k * gridDim.x + blockIdx.x) * 2 + 0] = sh[0][0];
    part[((size_t)k * gridDim.x + blockIdx.x) * 2 + 1] = sh[1][0];
  }
}

constexpr int FWD_GRID_MAX = 1024;
constexpr int WG_BLOCKS_MAX = 128;

int fwd_grid(int64_t npad) {
  const int64_t g = npad / ROWS_BLK;
  return (int)(g < FWD_GRID_MAX ? g : FWD_GRID_MAX);
}

}  // namespace

struct Mlp {
  Ctx *ctx;
  int64_t max_batch = 0, bpad = 0;
  DBuf<float> p, m, v, grad, partw, parts, y, d4;
  DBuf<u16> pb, act, dz;
  DBuf<double> loss, spart;
  int64_t step = 0;
  SegArgs sg{};
  int base = 0;  // guess relative to u0 (0) or to the linear extrapolation 2 u0 - u1 (1)
};

static void mlp_alloc(Mlp *M, int64_t max_batch) {
  M->max_batch = max_batch;
  M->bpad = (max_batch + ROWS_BLK - 1) / ROWS_BLK * ROWS_BLK;
  M->p.alloc(NP);
  M->m.alloc(NP);
  M->v.alloc(NP);
  M->grad.alloc(NP);
  M->pb.alloc(NP + 8);
  M->loss.alloc(1);
  M->act.alloc((size_t)3 * M->bpad * H);
  M->dz.alloc((size_t)3 * M->bpad * H);
  M->y.alloc(M->bpad);
  M->d4.alloc(M->bpad);
  M->partw.alloc((size_t)WG_BLOCKS_MAX * 2 * H * H);
  M->parts.alloc((size_t)WG_BLOCKS_MAX * NSMALL);
  M->sg.nseg = 1;
  M->sg.off[0] = 0;
  M->sg.off[1] = INT64_MAX;
  M->sg.s[0] = M->sg.U[0] = 1.f;
}

Mlp *mlp_create(Ctx *ctx, const float *params, int64_t max_batch) {
  OCMP_REQUIRE(max_batch >= 1 && max_batch <= ((int64_t)1 << 26), "mlp: max_batch out of range [1, 2^26]");
  Mlp *M = new Mlp();
  M->ctx = ctx;
  try {
    mlp_alloc(M, max_batch);
    mlp_set_params(M, params);
  } catch (...) {
    delete M;
    throw;
  }
  return M;
}

void mlp_destroy(Mlp *M) { delete M; }

void mlp_set_params(Mlp *M, const float *params) {
  hipStream_t s = M->ctx->stream;
  OCMP_REQUIRE(params, "mlp: null parameter array");
  OCMP_HIP(hipMemcpyAsync(M->p.p, params, NP * sizeof(float), hipMemcpyHostToDevice, s));
  OCMP_HIP(hipMemsetAsync(M->m.p, 0, NP * sizeof(float), s));
  OCMP_HIP(hipMemsetAsync(M->v.p, 0, NP * sizeof(float), s));
  k_mlp_mirror<<<(NP + 255) / 256, 256, 0, s>>>(M->p.p, M->pb.p);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipStreamSynchronize(s));
  M->step = 0;
}

void mlp_get_params(Mlp *M, float *params) {
  OCMP_REQUIRE(params, "mlp: null output array");
  OCMP_HIP(hipMemcpyAsync(params, M->p.p, NP * sizeof(float), hipMemcpyDeviceToHost, M->ctx->stream));
  OCMP_HIP(hipStreamSynchronize(M->ctx->stream));
}

static void set_segments(SegArgs &sg, int nseg, const int64_t *off) {
  OCMP_REQUIRE(nseg >= 1 && nseg <= MAXSEG, "mlp: nseg out of range [1, OCMP_MLP_MAXSEG]");
  OCMP_REQUIRE(off, "mlp: null segment offsets");
  OCMP_REQUIRE(off[0] == 0, "mlp: segment offsets must start at 0");
  for (int k = 0; k < nseg; ++k) OCMP_REQUIRE(off[k + 1] > off[k], "mlp: segment offsets must increase");
  sg.nseg = nseg;
  for (int k = 0; k <= nseg; ++k) sg.off[k] = off[k];
}

void mlp_set_scales(Mlp *M, int nseg, const int64_t *off, const double *s, const double *U) {
  set_segments(M->sg, nseg, off);
  for (int k = 0; k < nseg; ++k) {
    OCMP_REQUIRE(s[k] > 0 && U[k] > 0 && std::isfinite(s[k]) && std::isfinite(U[k]), "mlp: scales must be > 0");
    M->sg.s[k] = (float)s[k];
    M->sg.U[k] = (float)U[k];
  }
}

void mlp_get_scales(Mlp *M, int *nseg, int64_t *off, double *s, double *U) {
  *nseg = M->sg.nseg;
  for (int k = 0; k <= M->sg.nseg; ++k) off[k] = M->sg.off[k];
  for (int k = 0; k < M->sg.nseg; ++k) {
    s[k] = M->sg.s[k];
    U[k] = M->sg.U[k];
  }
}

void mlp_fit_scales(Mlp *M, const double *snap, int nsnap, int64_t ld, int nseg, const int64_t *off) {
  OCMP_REQUIRE(snap && nsnap >= 2, "mlp: need at least two snapshots");
  SegArgs sg{};
  set_segments(sg, nseg, off);
  OCMP_REQUIRE(off[nseg] <= ld, "mlp: segments exceed the snapshot stride");
  const int nb = 256;
  M->spart.alloc((size_t)nseg * nb * 2);
  hipStream_t st = M->ctx->stream;
  k_mlp_scales<<<dim3(nb, nseg), 256, 0, st>>>(snap, nsnap, ld, sg, M->spart.p);
  OCMP_HIP(hipGetLastError());
  std::vector<double> h((size_t)nseg * nb * 2);
  OCMP_HIP(hipMemcpyAsync(h.data(), M->spart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  std::vector<double> s(nseg), U(nseg);
  for (int k = 0; k < nseg; ++k) {
    double si = 0.0, sv = 0.0;
    for (int b = 0; b < nb; ++b) {
      si += h[((size_t)k * nb + b) * 2];
      sv += h[((size_t)k * nb + b) * 2 + 1];
    }
    const double cnt = (double)(nsnap - 1) * (double)(off[k + 1] - off[k]);
    s[k] = std::sqrt(si / cnt);
    U[k] = std::sqrt(sv / cnt);
    if (!(s[k] > 1e-300)) s[k] = 1.0;  // a component that does not move: features stay 0
    if (!(U[k] > 1e-300)) U[k] = 1.0;
  }
  mlp_set_scales(M, nseg, off, s.data(), U.data());
}

void mlp_sample(Mlp *M, const double *snap, int nsnap, int64_t ld, int64_t n, uint64_t seed, int64_t batch,
                float *X, float *T) {
  OCMP_REQUIRE(snap && X && T, "mlp: null buffer");
  OCMP_REQUIRE(nsnap >= 5, "mlp: sampling needs at least five snapshots");
  OCMP_REQUIRE(n >= 1 && n <= ld, "mlp: n out of range");
  OCMP_REQUIRE(M->sg.off[M->sg.nseg] >= n || M->sg.off[M->sg.nseg] == INT64_MAX,
               "mlp: scale segments do not cover the dofs");
  OCMP_REQUIRE(batch >= 1 && batch <= ((int64_t)1 << 30), "mlp: batch out of range");
  hipStream_t st = M->ctx->stream;
  k_mlp_sample<<<(unsigned)((batch + 255) / 256), 256, 0, st>>>(snap, nsnap, ld, n, M->sg, seed, (int)batch,
                                                               M->base, (float4 *)X, T);
  OCMP_HIP(hipGetLastError());
}

void mlp_forward(Mlp *M, int64_t n, const float *X, float *Y) {
  OCMP_REQUIRE(X && Y, "mlp: null buffer");
  OCMP_REQUIRE(n >= 1, "mlp: empty input");
  OCMP_REQUIRE(((uintptr_t)X & 15) == 0, "mlp: X must be 16-byte aligned (n x 4 float32)");
  const int64_t npad = (n + ROWS_BLK - 1) / ROWS_BLK * ROWS_BLK;
  Ctx *c = M->ctx;
  KTimer t(c, c->cls("mlp_forward"), (double)n * 20.0 + NP * 2.0);
  k_mlp_fwd<0><<<fwd_grid(npad), 256, 0, c->stream>>>(M->pb.p, n, npad, (const float4 *)X, Y, nullptr, M->sg,
                                                       GuessArgs{});
  OCMP_HIP(hipGetLastError());
}

void mlp_guess(Mlp *M, int64_t n, const double *u0, const double *u1, const double *u2, const double *u3,
               double *out) {
  OCMP_REQUIRE(u0 && u1 && u2 && u3 && out, "mlp: null history vector");
  OCMP_REQUIRE(n >= 1, "mlp: empty input");
  OCMP_REQUIRE(M->sg.off[M->sg.nseg] >= n || M->sg.off[M->sg.nseg] == INT64_MAX,
               "mlp: scale segments do not cover the dofs");
  const int64_t npad = (n + ROWS_BLK - 1) / ROWS_BLK * ROWS_BLK;
  Ctx *c = M->ctx;
  KTimer t(c, c->cls("mlp_guess"), (double)n * 40.0 + NP * 2.0);
  k_mlp_fwd<2><<<fwd_grid(npad), 256, 0, c->stream>>>(M->pb.p, n, npad, nullptr, nullptr, nullptr, M->sg,
                                                       GuessArgs{u0, u1, u2, u3, out, M->base});
  OCMP_HIP(hipGetLastError());
}

void mlp_set_base(Mlp *M, int base) {
  OCMP_REQUIRE(base == 0 || base == 1, "mlp: base must be 0 (previous state) or 1 (linear extrapolation)");
  M->base = base;
}

int mlp_get_base(const Mlp *M) { return M->base; }

// gradient of the batch loss (and optionally one Adam step); loss (device) written to M->loss
static void mlp_backprop(Mlp *M, int64_t n, const float *X, const float *T, float *grad, bool update, float lr) {
  OCMP_REQUIRE(X && T, "mlp: null buffer");
  OCMP_REQUIRE(n >= 1 && n <= M->max_batch, "mlp: batch larger than the max_batch given at creation");
  OCMP_REQUIRE(((uintptr_t)X & 15) == 0, "mlp: X must be 16-byte aligned (n x 4 float32)");
  Ctx *c = M->ctx;
  hipStream_t st = c->stream;
  const int64_t npad = (n + ROWS_BLK - 1) / ROWS_BLK * ROWS_BLK;
  const int grid = fwd_grid(npad);
  const float invB = 1.0f / (float)n;
  {
    KTimer t(c, c->cls("mlp_train_fwd"), (double)n * 20.0 + (double)npad * 3 * H * 2);
    k_mlp_fwd<1><<<grid, 256, 0, st>>>(M->pb.p, n, npad, (const float4 *)X, M->y.p, M->act.p, M->sg, GuessArgs{});
    OCMP_HIP(hipGetLastError());
  }
  {
    KTimer t(c, c->cls("mlp_train_bwd"), (double)npad * (8.0 + 6 * H * 2));
    k_mlp_bwd<<<grid, 256, 0, st>>>(M->pb.p, n, npad, M->y.p, T, invB, M->act.p, M->dz.p, M->d4.p);
    OCMP_HIP(hipGetLastError());
  }
  // rows per weight-gradient block: a multiple of the staged chunk, at most WG_BLOCKS_MAX blocks
  int64_t rpb = (npad + WG_BLOCKS_MAX - 1) / WG_BLOCKS_MAX;
  rpb = (rpb + CH - 1) / CH * CH;
  const int nbw = (int)((npad + rpb - 1) / rpb);
  int64_t rps = (n + WG_BLOCKS_MAX - 1) / WG_BLOCKS_MAX;
  const int nbs = (int)((n + rps - 1) / rps);
  {
    KTimer t(c, c->cls("mlp_train_wgrad"), (double)npad * 4 * H * 2 + (double)nbw * 2 * H * H * 4);
    k_mlp_wgrad<<<dim3(nbw, 2), 256, 0, st>>>(npad, rpb, M->act.p, M->dz.p, M->partw.p);
    OCMP_HIP(hipGetLastError());
    k_mlp_small<<<nbs, 256, 0, st>>>(n, npad, rps, (const float4 *)X, M->y.p, T, M->d4.p, M->act.p, M->dz.p,
                                     M->parts.p);
    OCMP_HIP(hipGetLastError());
  }
  float c1 = 1.f, c2 = 1.f;
  if (update) {
    M->step += 1;
    c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)M->step)));
    c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)M->step)));
  }
  {
    KTimer t(c, c->cls("mlp_adam"), (double)nbw * 2 * H * H * 4 + (double)nbs * NSMALL * 4 + NP * 26.0);
    k_mlp_adam<<<(NP + 1 + 255) / 256, 256, 0, st>>>(nbw, nbs, M->partw.p, M->parts.p, grad ? grad : M->grad.p,
                                                     M->p.p, M->m.p, M->v.p, M->pb.p, lr, c1, c2, update ? 1 : 0,
                                                     M->loss.p, 0.5 / (double)n);
    OCMP_HIP(hipGetLastError());
  }
}

void mlp_grad(Mlp *M, int64_t n, const float *X, const float *T, float *grad, double *loss) {
  mlp_backprop(M, n, X, T, grad, false, 0.f);
  if (loss) {
    OCMP_HIP(hipMemcpyAsync(loss, M->loss.p, sizeof(double), hipMemcpyDeviceToHost, M->ctx->stream));
    OCMP_HIP(hipStreamSynchronize(M->ctx->stream));
  }
}

void mlp_train_step(Mlp *M, int64_t n, const float *X, const float *T, float lr, double *loss) {
  OCMP_REQUIRE(lr > 0.f && std::isfinite(lr), "mlp: learning rate must be > 0");
  mlp_backprop(M, n, X, T, nullptr, true, lr);
  if (loss) {
    OCMP_HIP(hipMemcpyAsync(loss, M->loss.p, sizeof(double), hipMemcpyDeviceToHost, M->ctx->stream));
    OCMP_HIP(hipStreamSynchronize(M->ctx->stream));
  }
}

}  // namespace ocmp
