// Learned preconditioner correction: a 3-level U-net on the vertex grid of a
// structured (N^3 cubes x 6 tets) mesh, used inside flexible GMRES.
//
// The correction acts on the vertex dofs of four scalar blocks (the velocity
// components and the pressure of the diffuse-interface Stokes system): a vector
// v is restricted to a 4-channel grid (vertex id = grid index, x fastest), the
// network maps it to a 4-channel grid y, and y is added back onto the same
// vertex dofs. As a preconditioner, M v = D^{-1} v + E U(R v) (D the point-Jacobi
// diagonal). The network is trained on the GPU to minimise
// 0.5 ||v - A M v||^2 over vectors v from plain solves (self-supervised: only
// products with A are needed), by the kernels below and Adam.
//
// Layers (3x3x3 kernels, zero padding), bf16 weights and activations, fp32 sums:
//   a0 = relu(conv (x;  4 -> 8))        a1 = relu(conv2(a0; 8 -> 16))
//   a2 = relu(conv (a1; 16 -> 16))      a3 = relu(conv2(a2; 16 -> 32))
//   a4 = relu(conv (a3; 32 -> 32))      a5 = relu(tconv2(a4; 32 -> 16) + a2)
//   a6 = relu(tconv2(a5; 16 -> 8) + a0) y  = conv (a6; 8 -> 4)
// conv stride s: out[o,z] = b[o] + sum_{i,d} W[o,i,d] in[i, s z + d - 1]
// tconv stride 2: out[o,z] = b[o] + sum_{i,d} W[i,o,d] in[i,y] over 2y + d - 1 = z
// Weight-gradient reductions use fixed partitions and fixed-order sums: bitwise
// repeatable. No counterpart in the reference (its preconditioners are
// opencmp/models/base_model.py:365-383).
#include "internal.h"

#include <cmath>

namespace ocmp {
namespace {

typedef unsigned short u16;

__device__ __forceinline__ u16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }

constexpr int NL = 8;          // layers
constexpr int MAXC = 32;       // widest layer
constexpr int COB = 8;         // output channels per thread in the conv kernel
constexpr int NT = 256;

struct LayerDef {
  int cin, cout, stride, transposed, level_in, level_out;
};
// level 0: n0 = N+1, level 1: n1 = N/2+1, level 2: n2 = N/4+1
constexpr LayerDef LD[NL] = {{4, 8, 1, 0, 0, 0},   {8, 16, 2, 0, 0, 1},  {16, 16, 1, 0, 1, 1}, {16, 32, 2, 0, 1, 2},
                             {32, 32, 1, 0, 2, 2}, {32, 16, 2, 1, 2, 1}, {16, 8, 2, 1, 1, 0},  {8, 4, 1, 0, 0, 0}};

constexpr int layer_w(int l) { return LD[l].cin * LD[l].cout * 27; }
constexpr int layer_p(int l) { return layer_w(l) + LD[l].cout; }
constexpr int nparam() {
  int s = 0;
  for (int l = 0; l < NL; ++l) s += layer_p(l);
  return s;
}
constexpr int NP = nparam();
static_assert(NP == OCMP_UNET_NPARAM, "parameter layout");

int param_off(int l) {
  int s = 0;
  for (int k = 0; k < l; ++k) s += layer_p(k);
  return s;
}

struct Conv {
  const void *in;
  int in_bf16, cin, nin;
  void *out;
  int out_bf16, cout, nout;
  const u16 *W;
  int so, si;               // w(o, i, k) = W[(o*so + i*si)*27 + k]
  const u16 *bias;          // may be null
  const void *skip;         // added before the activation, may be null
  int skip_bf16;
  const u16 *mask;          // multiply by [mask > 0] (backward), may be null
  int relu, stride, transposed;
};

__device__ __forceinline__ float ld_any(const void *p, int bf, size_t i) {
  return bf ? bf2f(((const u16 *)p)[i]) : ((const float *)p)[i];
}

// one thread per output voxel, COB output channels (blockIdx.y picks the chunk)
__global__ __launch_bounds__(NT) void k_conv3(Conv c) {
  __shared__ float wl[27 * MAXC * COB];
  const int o0 = blockIdx.y * COB;
  const int nco = min(COB, c.cout - o0);
  for (int e = threadIdx.x; e < 27 * c.cin * COB; e += NT) {
    const int o = e % COB, i = (e / COB) % c.cin, k = e / (COB * c.cin);
    wl[e] = o < nco ? bf2f(c.W[((size_t)(o0 + o) * c.so + (size_t)i * c.si) * 27 + k]) : 0.f;
  }
  __syncthreads();
  const int64_t zout = (int64_t)c.nout * c.nout * c.nout;
  const int64_t zin = (int64_t)c.nin * c.nin * c.nin;
  const int64_t z = (int64_t)blockIdx.x * NT + threadIdx.x;
  if (z >= zout) return;
  const int zx = (int)(z % c.nout), zy = (int)((z / c.nout) % c.nout), zz = (int)(z / ((int64_t)c.nout * c.nout));
  float acc[COB];
#pragma unroll
  for (int o = 0; o < COB; ++o) acc[o] = 0.f;
  for (int k = 0; k < 27; ++k) {
    const int dz = k / 9, dy = (k / 3) % 3, dx = k % 3;
    int sx, sy, sz;
    if (!c.transposed) {
      sx = c.stride * zx + dx - 1;
      sy = c.stride * zy + dy - 1;
      sz = c.stride * zz + dz - 1;
    } else {
      const int tx = zx - dx + 1, ty = zy - dy + 1, tz = zz - dz + 1;
      if (c.stride == 2 && ((tx | ty | tz) & 1)) continue;
      sx = c.stride == 2 ? tx >> 1 : tx;
      sy = c.stride == 2 ? ty >> 1 : ty;
      sz = c.stride == 2 ? tz >> 1 : tz;
    }
    if (sx < 0 || sy < 0 || sz < 0 || sx >= c.nin || sy >= c.nin || sz >= c.nin) continue;
    const int64_t src = ((int64_t)sz * c.nin + sy) * c.nin + sx;
    for (int i = 0; i < c.cin; ++i) {
      const float v = ld_any(c.in, c.in_bf16, (size_t)i * zin + src);
      const float *w = wl + (k * c.cin + i) * COB;
#pragma unroll
      for (int o = 0; o < COB; ++o) acc[o] += w[o] * v;
    }
  }
  for (int o = 0; o < nco; ++o) {
    const size_t at = (size_t)(o0 + o) * zout + z;
    float v = acc[o];
    if (c.bias) v += bf2f(c.bias[o0 + o]);
    if (c.skip) v += ld_any(c.skip, c.skip_bf16, at);
    if (c.relu) v = fmaxf(v, 0.f);
    if (c.mask && !(bf2f(c.mask[at]) > 0.f)) v = 0.f;
    if (c.out_bf16)
      ((u16 *)c.out)[at] = f2bf(v);
    else
      ((float *)c.out)[at] = v;
  }
}

// gW[a][b][k] partial over the voxels of chunk blockIdx.y:  sum_z P[a][z] Q[b][s z + d_k - 1]
struct WGrad {
  const void *P;
  int p_bf16, na, np;
  const void *Q;
  int q_bf16, nb, nq;
  int stride;
  int64_t chunk;
  float *part;  // [nchunks][na*nb][27]
};

__global__ __launch_bounds__(NT) void k_conv_wgrad(WGrad g) {
  __shared__ float sh[NT / 64][27];
  const int pair = blockIdx.x;
  const int a = pair / g.nb, b = pair % g.nb;
  const int64_t zp = (int64_t)g.np * g.np * g.np, zq = (int64_t)g.nq * g.nq * g.nq;
  const int64_t z0 = (int64_t)blockIdx.y * g.chunk, z1 = min(zp, z0 + g.chunk);
  float acc[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) acc[k] = 0.f;
  for (int64_t z = z0 + threadIdx.x; z < z1; z += NT) {
    const float pv = ld_any(g.P, g.p_bf16, (size_t)a * zp + z);
    if (pv == 0.f) continue;
    const int zx = (int)(z % g.np), zy = (int)((z / g.np) % g.np), zz = (int)(z / ((int64_t)g.np * g.np));
#pragma unroll
    for (int k = 0; k < 27; ++k) {
      const int dz = k / 9, dy = (k / 3) % 3, dx = k % 3;
      const int sx = g.stride * zx + dx - 1, sy = g.stride * zy + dy - 1, sz = g.stride * zz + dz - 1;
      if (sx >= 0 && sy >= 0 && sz >= 0 && sx < g.nq && sy < g.nq && sz < g.nq)
        acc[k] += pv * ld_any(g.Q, g.q_bf16, (size_t)b * zq + ((int64_t)sz * g.nq + sy) * g.nq + sx);
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    float v = acc[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) sh[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 27) {
    float s = 0.f;
    for (int w = 0; w < NT / 64; ++w) s += sh[w][threadIdx.x];
    g.part[((size_t)blockIdx.y * gridDim.x + pair) * 27 + threadIdx.x] = s;
  }
}

// per-channel partial sums of a fp32 gradient over voxel chunks: part[chunk][c]
__global__ __launch_bounds__(NT) void k_chan_sum(const float *__restrict__ G, int nc, int64_t zn, int64_t chunk,
                                                 float *__restrict__ part) {
  __shared__ float sh[NT / 64];
  const int c = blockIdx.x;
  const int64_t z0 = (int64_t)blockIdx.y * chunk, z1 = min(zn, z0 + chunk);
  float s = 0.f;
  for (int64_t z = z0 + threadIdx.x; z < z1; z += NT) s += G[(size_t)c * zn + z];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.f;
    for (int w = 0; w < NT / 64; ++w) t += sh[w];
    part[(size_t)blockIdx.y * nc + c] = t;
  }
}

// grad[off + j] (+)= sum over chunks of part[chunk][j], j < len
__global__ void k_chunk_reduce(const float *__restrict__ part, int nchunks, int len, float *__restrict__ grad,
                               int accumulate) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= len) return;
  float s = 0.f;
  for (int c = 0; c < nchunks; ++c) s += part[(size_t)c * len + j];
  grad[j] = accumulate ? grad[j] + s : s;
}

__global__ void k_adam_step(int n, const float *__restrict__ grad, float gscale, float *p, float *m, float *v,
                            u16 *pb, float lr, float c1, float c2) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const float g = grad[j] * gscale;
  const float mj = 0.9f * m[j] + 0.1f * g;
  const float vj = 0.999f * v[j] + 0.001f * g * g;
  m[j] = mj;
  v[j] = vj;
  const float pj = p[j] - lr * (mj * c1) / (sqrtf(vj * c2) + 1e-8f);
  p[j] = pj;
  pb[j] = f2bf(pj);
}

__global__ void k_mirror_bf16(int n, const float *__restrict__ p, u16 *__restrict__ pb) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) pb[j] = f2bf(p[j]);
}

struct Chan4 {
  int64_t off[4];   // dof offset of each channel's scalar block (its vertex dofs come first)
  float scale[4];
};

// x[c][p] = bf16(sv * v[off_c + p] * scale_c)
template <typename T>
__global__ void k_restrict(const T *__restrict__ v, const double *__restrict__ sv, int64_t nv, Chan4 ch,
                           u16 *__restrict__ x) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nv) return;
  const float s = sv ? (float)sv[0] : 1.f;
#pragma unroll
  for (int c = 0; c < 4; ++c) x[(size_t)c * nv + p] = f2bf((float)v[ch.off[c] + p] * s * ch.scale[c]);
}

// g[c][p] = scale_c * v[off_c + p]   (fp32, gradient path)
__global__ void k_restrict_f32(const double *__restrict__ v, int64_t nv, Chan4 ch, float *__restrict__ g) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nv) return;
#pragma unroll
  for (int c = 0; c < 4; ++c) g[(size_t)c * nv + p] = (float)v[ch.off[c] + p] * ch.scale[c];
}

// z[off_c + p] (+)= scale_c * y[c][p] on free dofs
template <typename T>
__global__ void k_embed(const float *__restrict__ y, int64_t nv, Chan4 ch, const uint8_t *__restrict__ freem,
                        T *__restrict__ z, int accumulate) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nv) return;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int64_t i = ch.off[c] + p;
    const T v = freem[i] ? (T)(ch.scale[c] * y[(size_t)c * nv + p]) : (T)0;
    z[i] = accumulate ? (T)(z[i] + v) : v;
  }
}

// r = v - A D^{-1} v helpers on FP64 dof vectors
__global__ void k_mul_diag(int64_t n, const double *__restrict__ d, const double *__restrict__ v,
                           double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = d[i] * v[i];
}

__global__ void k_sub_masked(int64_t n, const double *__restrict__ a, const double *__restrict__ b,
                             const uint8_t *__restrict__ freem, double *__restrict__ out, double *__restrict__ part) {
  __shared__ double sh[NT / 64];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = freem[i] ? a[i] - b[i] : 0.0;
    out[i] = v;
    s += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += sh[w];
    part[blockIdx.x] = t;
  }
}

__global__ void k_neg_masked(int64_t n, const uint8_t *__restrict__ freem, double *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = freem[i] ? -v[i] : 0.0;
}

// per-block partial sums over samples and vertex points of (v_c)^2 and (dinv v)_c^2, c < 4
__global__ __launch_bounds__(NT) void k_chan_stats(const double *__restrict__ S, int ns, int64_t ld, int64_t nv,
                                                   Chan4 ch, const double *__restrict__ dinv,
                                                   double *__restrict__ part) {
  __shared__ double sh[8][NT / 64];
  double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int64_t tot = (int64_t)ns * nv;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < tot; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = q / nv, p = q % nv;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t i = ch.off[c] + p;
      const double v = S[(size_t)s * ld + i];
      const double z = dinv[i] * v;
      a[c] += v * v;
      a[4 + c] += z * z;
    }
  }
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    double v = a[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[c][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double t = 0.0;
    for (int w = 0; w < NT / 64; ++w) t += sh[threadIdx.x][w];
    part[(size_t)blockIdx.x * 8 + threadIdx.x] = t;
  }
}

constexpr int SUB_GRID = 1024;

}  // namespace

struct Unet {
  Ctx *ctx;
  int N;
  int64_t n[3], zn[3];
  DBuf<float> p, m, v, grad, part;
  DBuf<u16> pb;
  DBuf<u16> x, act[7];          // x, a0..a6 (bf16)
  DBuf<float> y, gy, g[7], gtmp; // y, dL/dy, gradients of a0..a6 (pre-activation)
  Chan4 in_ch{}, out_ch{};
  int64_t step = 0;
  // training state
  DBuf<double> rj, tmp, res, zf, spart;
};

static void conv_launch(Unet *U, const Conv &c, hipStream_t st) {
  const int64_t zout = (int64_t)c.nout * c.nout * c.nout;
  dim3 grid((unsigned)((zout + NT - 1) / NT), (unsigned)((c.cout + COB - 1) / COB));
  k_conv3<<<grid, NT, 0, st>>>(c);
  OCMP_HIP(hipGetLastError());
}

// forward over the bf16 input U->x; keeps every activation
static void unet_forward(Unet *U, hipStream_t st) {
  const u16 *W[NL], *B[NL];
  for (int l = 0; l < NL; ++l) {
    W[l] = U->pb.p + param_off(l);
    B[l] = W[l] + layer_w(l);
  }
  const void *in = U->x.p;
  for (int l = 0; l < NL; ++l) {
    const LayerDef &L = LD[l];
    Conv c{};
    c.in = (l == 0) ? (const void *)U->x.p : (const void *)U->act[l - 1].p;
    c.in_bf16 = 1;
    c.cin = L.cin;
    c.nin = (int)U->n[L.level_in];
    c.cout = L.cout;
    c.nout = (int)U->n[L.level_out];
    c.W = W[l];
    if (!L.transposed) {
      c.so = L.cin;
      c.si = 1;
    } else {
      c.so = 1;
      c.si = L.cout;
    }
    c.bias = B[l];
    c.stride = L.stride;
    c.transposed = L.transposed;
    if (l == 5) c.skip = U->act[2].p, c.skip_bf16 = 1;
    if (l == 6) c.skip = U->act[0].p, c.skip_bf16 = 1;
    if (l < NL - 1) {
      c.out = U->act[l].p;
      c.out_bf16 = 1;
      c.relu = 1;
    } else {
      c.out = U->y.p;
      c.out_bf16 = 0;
      c.relu = 0;
    }
    (void)in;
    conv_launch(U, c, st);
  }
}

static void wgrad_launch(Unet *U, int l, const void *P, int p_bf16, int na, int np, const void *Q, int q_bf16,
                         int nb, int nq, int stride, float *grad_w, int accumulate, hipStream_t st) {
  const int64_t zp = (int64_t)np * np * np;
  const int pairs = na * nb;
  int nch = (int)std::max<int64_t>(1, std::min<int64_t>((zp + 2047) / 2048, std::max(1, 2048 / pairs)));
  const int64_t chunk = (zp + nch - 1) / nch;
  nch = (int)((zp + chunk - 1) / chunk);
  OCMP_REQUIRE((size_t)nch * pairs * 27 <= U->part.n, "unet: weight-gradient scratch too small");
  WGrad g{P, p_bf16, na, np, Q, q_bf16, nb, nq, stride, chunk, U->part.p};
  k_conv_wgrad<<<dim3(pairs, nch), NT, 0, st>>>(g);
  OCMP_HIP(hipGetLastError());
  const int len = pairs * 27;
  k_chunk_reduce<<<(len + 255) / 256, 256, 0, st>>>(U->part.p, nch, len, grad_w, accumulate);
  OCMP_HIP(hipGetLastError());
  (void)l;
}

static void bias_grad(Unet *U, const float *G, int nc, int n, float *grad_b, int accumulate, hipStream_t st) {
  const int64_t zn = (int64_t)n * n * n;
  const int nch = (int)std::max<int64_t>(1, std::min<int64_t>(64, (zn + 4095) / 4096));
  const int64_t chunk = (zn + nch - 1) / nch;
  OCMP_REQUIRE((size_t)nch * nc <= U->gtmp.n, "unet: bias-gradient scratch too small");
  k_chan_sum<<<dim3(nc, nch), NT, 0, st>>>(G, nc, zn, chunk, U->gtmp.p);
  OCMP_HIP(hipGetLastError());
  k_chunk_reduce<<<(nc + 255) / 256, 256, 0, st>>>(U->gtmp.p, nch, nc, grad_b, accumulate);
  OCMP_HIP(hipGetLastError());
}

// backward from U->gy (dL/dy, fp32): parameter gradients into U->grad (accumulated if asked)
static void unet_backward(Unet *U, int accumulate, hipStream_t st) {
  const u16 *W[NL];
  float *GW[NL], *GB[NL];
  for (int l = 0; l < NL; ++l) {
    W[l] = U->pb.p + param_off(l);
    GW[l] = U->grad.p + param_off(l);
    GB[l] = GW[l] + layer_w(l);
  }
  const int n0 = (int)U->n[0], n1 = (int)U->n[1], n2 = (int)U->n[2];
  // adjoint (grad-input) of layer l applied to gout, masked by the layer input's activation
  auto adj = [&](int l, const float *gout, float *gin, const float *skipg, const u16 *mask) {
    const LayerDef &L = LD[l];
    Conv c{};
    c.in = gout;
    c.in_bf16 = 0;
    c.cin = L.cout;
    c.nin = (int)U->n[L.level_out];
    c.out = gin;
    c.out_bf16 = 0;
    c.cout = L.cin;
    c.nout = (int)U->n[L.level_in];
    c.W = W[l];
    if (!L.transposed) {  // w(o = ci, i = co) = W[co][ci]
      c.so = 1;
      c.si = L.cin;
      c.transposed = 1;   // adjoint of a gather conv is a scatter (transposed) conv
    } else {              // layer weight W[ci][co]: w(o = ci, i = co) = W[ci][co]
      c.so = L.cout;
      c.si = 1;
      c.transposed = 0;
    }
    c.stride = L.stride;
    c.skip = skipg;
    c.skip_bf16 = 0;
    c.mask = mask;
    conv_launch(U, c, st);
  };
  // OUT (conv a6 -> y)
  wgrad_launch(U, 7, U->gy.p, 0, 4, n0, U->act[6].p, 1, 8, n0, 1, GW[7], accumulate, st);
  bias_grad(U, U->gy.p, 4, n0, GB[7], accumulate, st);
  adj(7, U->gy.p, U->g[6].p, nullptr, U->act[6].p);                     // g6
  // U0 (tconv a5 -> a6): gW[ci][co] = sum_y a5[ci][y] g6[co][2y+d-1]
  wgrad_launch(U, 6, U->act[5].p, 1, 16, n1, U->g[6].p, 0, 8, n0, 2, GW[6], accumulate, st);
  bias_grad(U, U->g[6].p, 8, n0, GB[6], accumulate, st);
  adj(6, U->g[6].p, U->g[5].p, nullptr, U->act[5].p);                   // g5
  // U1 (tconv a4 -> a5)
  wgrad_launch(U, 5, U->act[4].p, 1, 32, n2, U->g[5].p, 0, 16, n1, 2, GW[5], accumulate, st);
  bias_grad(U, U->g[5].p, 16, n1, GB[5], accumulate, st);
  adj(5, U->g[5].p, U->g[4].p, nullptr, U->act[4].p);                   // g4
  // L2 (conv a3 -> a4)
  wgrad_launch(U, 4, U->g[4].p, 0, 32, n2, U->act[3].p, 1, 32, n2, 1, GW[4], accumulate, st);
  bias_grad(U, U->g[4].p, 32, n2, GB[4], accumulate, st);
  adj(4, U->g[4].p, U->g[3].p, nullptr, U->act[3].p);                   // g3
  // D2 (conv stride 2 a2 -> a3)
  wgrad_launch(U, 3, U->g[3].p, 0, 32, n2, U->act[2].p, 1, 16, n1, 2, GW[3], accumulate, st);
  bias_grad(U, U->g[3].p, 32, n2, GB[3], accumulate, st);
  adj(3, U->g[3].p, U->g[2].p, U->g[5].p, U->act[2].p);                 // g2 (+ skip from a5)
  // L1 (conv a1 -> a2)
  wgrad_launch(U, 2, U->g[2].p, 0, 16, n1, U->act[1].p, 1, 16, n1, 1, GW[2], accumulate, st);
  bias_grad(U, U->g[2].p, 16, n1, GB[2], accumulate, st);
  adj(2, U->g[2].p, U->g[1].p, nullptr, U->act[1].p);                   // g1
  // D1 (conv stride 2 a0 -> a1)
  wgrad_launch(U, 1, U->g[1].p, 0, 16, n1, U->act[0].p, 1, 8, n0, 2, GW[1], accumulate, st);
  bias_grad(U, U->g[1].p, 16, n1, GB[1], accumulate, st);
  adj(1, U->g[1].p, U->g[0].p, U->g[6].p, U->act[0].p);                 // g0 (+ skip from a6)
  // L0 (conv x -> a0)
  wgrad_launch(U, 0, U->g[0].p, 0, 8, n0, U->x.p, 1, 4, n0, 1, GW[0], accumulate, st);
  bias_grad(U, U->g[0].p, 8, n0, GB[0], accumulate, st);
}

Unet *unet_create(Ctx *ctx, int N, const float *params) {
  OCMP_REQUIRE(N >= 4 && N % 4 == 0 && N <= 512, "unet: N must be a multiple of 4 in [4, 512]");
  OCMP_REQUIRE(params, "unet: null parameters");
  auto *U = new Unet();
  U->ctx = ctx;
  U->N = N;
  U->n[0] = N + 1;
  U->n[1] = N / 2 + 1;
  U->n[2] = N / 4 + 1;
  for (int l = 0; l < 3; ++l) U->zn[l] = U->n[l] * U->n[l] * U->n[l];
  try {
    U->p.alloc(NP);
    U->m.alloc(NP);
    U->v.alloc(NP);
    U->grad.alloc(NP);
    U->pb.alloc(NP);
    U->x.alloc((size_t)4 * U->zn[0]);
    for (int l = 0; l < 7; ++l) {
      const LayerDef &L = LD[l];
      U->act[l].alloc((size_t)L.cout * U->zn[L.level_out]);
      U->g[l].alloc((size_t)L.cout * U->zn[L.level_out]);
    }
    U->y.alloc((size_t)4 * U->zn[0]);
    U->gy.alloc((size_t)4 * U->zn[0]);
    // reduction scratch at its largest size (pairs x chunks <= 2048, channels x chunks <= 32 x 64)
    U->part.alloc((size_t)2048 * 27);
    U->gtmp.alloc((size_t)32 * 64);
    unet_set_params(U, params);
  } catch (...) {
    delete U;
    throw;
  }
  for (int c = 0; c < 4; ++c) U->in_ch.scale[c] = U->out_ch.scale[c] = 1.f;
  return U;
}

void unet_destroy(Unet *U) { delete U; }

int64_t unet_grid_points(const Unet *U) { return U->zn[0]; }

void unet_set_params(Unet *U, const float *params) {
  hipStream_t st = U->ctx->stream;
  OCMP_HIP(hipMemcpyAsync(U->p.p, params, NP * sizeof(float), hipMemcpyHostToDevice, st));
  OCMP_HIP(hipMemsetAsync(U->m.p, 0, NP * sizeof(float), st));
  OCMP_HIP(hipMemsetAsync(U->v.p, 0, NP * sizeof(float), st));
  k_mirror_bf16<<<(NP + 255) / 256, 256, 0, st>>>(NP, U->p.p, U->pb.p);
  OCMP_HIP(hipGetLastError());
  OCMP_HIP(hipStreamSynchronize(st));
  U->step = 0;
}

void unet_get_params(Unet *U, float *params) {
  OCMP_HIP(hipMemcpyAsync(params, U->p.p, NP * sizeof(float), hipMemcpyDeviceToHost, U->ctx->stream));
  OCMP_HIP(hipStreamSynchronize(U->ctx->stream));
}

void unet_set_channels(Unet *U, const int64_t *off, const double *in_scale, const double *out_scale) {
  for (int c = 0; c < 4; ++c) {
    OCMP_REQUIRE(off[c] >= 0, "unet: negative channel offset");
    OCMP_REQUIRE(in_scale[c] > 0 && out_scale[c] > 0 && std::isfinite(in_scale[c]) && std::isfinite(out_scale[c]),
                 "unet: channel scales must be positive");
    U->in_ch.off[c] = U->out_ch.off[c] = off[c];
    U->in_ch.scale[c] = (float)in_scale[c];
    U->out_ch.scale[c] = (float)out_scale[c];
  }
}

int64_t unet_max_index(const Unet *U) {
  int64_t m = 0;
  for (int c = 0; c < 4; ++c) m = std::max(m, U->in_ch.off[c] + U->zn[0]);
  return m;
}

// channel scales from samples (device f64 rows, stride ld): in_c = 1/rms(v_c), out_c = rms((D^{-1} v)_c)
void unet_fit_channels(Unet *U, Precond *P, const double *samples, int ns, int64_t ld, const int64_t *off) {
  OCMP_REQUIRE(P && P->kind == OCMP_PRE_JACOBI && P->ready, "unet: needs an updated Jacobi preconditioner");
  for (int c = 0; c < 4; ++c) {
    OCMP_REQUIRE(off[c] >= 0 && off[c] + U->zn[0] <= P->A->n, "unet: channel dofs out of range");
    U->in_ch.off[c] = U->out_ch.off[c] = off[c];
  }
  hipStream_t st = U->ctx->stream;
  const int nb = 512;
  DBuf<double> part;
  part.alloc((size_t)nb * 8);
  k_chan_stats<<<nb, NT, 0, st>>>(samples, ns, ld, U->zn[0], U->in_ch, P->dinv.p, part.p);
  OCMP_HIP(hipGetLastError());
  std::vector<double> h((size_t)nb * 8);
  OCMP_HIP(hipMemcpyAsync(h.data(), part.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  const double cnt = (double)ns * (double)U->zn[0];
  for (int c = 0; c < 4; ++c) {
    double sv = 0.0, sz = 0.0;
    for (int b = 0; b < nb; ++b) {
      sv += h[(size_t)b * 8 + c];
      sz += h[(size_t)b * 8 + 4 + c];
    }
    const double rv = std::sqrt(sv / cnt), rz = std::sqrt(sz / cnt);
    U->in_ch.scale[c] = rv > 0 ? (float)(1.0 / rv) : 1.f;
    U->out_ch.scale[c] = rz > 0 ? (float)rz : 1.f;
  }
}

void unet_get_channels(const Unet *U, int64_t *off, double *in_scale, double *out_scale) {
  for (int c = 0; c < 4; ++c) {
    off[c] = U->in_ch.off[c];
    in_scale[c] = U->in_ch.scale[c];
    out_scale[c] = U->out_ch.scale[c];
  }
}

// copy one stored activation (layer -1: the bf16 input, 0..6: a0..a6, 7: y) as float32
__global__ void k_bf16_to_f32(int64_t n, const u16 *__restrict__ a, float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bf2f(a[i]);
}

int64_t unet_activation(Unet *U, int layer, float *out) {
  hipStream_t st = U->ctx->stream;
  if (layer == 7) {
    const int64_t n = 4 * U->zn[0];
    if (out) OCMP_HIP(hipMemcpyAsync(out, U->y.p, n * sizeof(float), hipMemcpyDeviceToDevice, st));
    return n;
  }
  const u16 *src = layer < 0 ? U->x.p : U->act[layer].p;
  const int64_t n = layer < 0 ? 4 * U->zn[0] : (int64_t)LD[layer].cout * U->zn[LD[layer].level_out];
  if (out) {
    k_bf16_to_f32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, src, out);
    OCMP_HIP(hipGetLastError());
  }
  return n;
}

// y = U(x): grid in/out (device, x: float32 [4][n0^3] rounded to bf16 on entry, y: float32)
void unet_forward_grid(Unet *U, const float *x, float *y) {
  hipStream_t st = U->ctx->stream;
  Ctx *c = U->ctx;
  KTimer t(c, c->cls("unet_forward"), 0.0);
  const int64_t n4 = 4 * U->zn[0];
  k_mirror_bf16<<<(unsigned)((n4 + 255) / 256), 256, 0, st>>>((int)n4, x, U->x.p);
  OCMP_HIP(hipGetLastError());
  unet_forward(U, st);
  OCMP_HIP(hipMemcpyAsync(y, U->y.p, n4 * sizeof(float), hipMemcpyDeviceToDevice, st));
}

// parameter gradient of sum(gy * U(x)) (grid, device float32) into grad (device float32[NP])
void unet_grad_grid(Unet *U, const float *x, const float *gy, float *grad) {
  hipStream_t st = U->ctx->stream;
  const int64_t n4 = 4 * U->zn[0];
  k_mirror_bf16<<<(unsigned)((n4 + 255) / 256), 256, 0, st>>>((int)n4, x, U->x.p);
  OCMP_HIP(hipGetLastError());
  unet_forward(U, st);
  OCMP_HIP(hipMemcpyAsync(U->gy.p, gy, n4 * sizeof(float), hipMemcpyDeviceToDevice, st));
  unet_backward(U, 0, st);
  OCMP_HIP(hipMemcpyAsync(grad, U->grad.p, NP * sizeof(float), hipMemcpyDeviceToDevice, st));
}

// z (+)= E U(R (s v)) on the free dofs of A's space (the learned half of the preconditioner)
template <typename T>
void unet_correct(Unet *U, const uint8_t *freem, const T *v, const double *sv, T *z, int accumulate) {
  hipStream_t st = U->ctx->stream;
  Ctx *c = U->ctx;
  KTimer t(c, c->cls(sizeof(T) == 8 ? "unet_correction_f64" : "unet_correction_f32"), 0.0);
  const int64_t nv = U->zn[0];
  k_restrict<T><<<(unsigned)((nv + 255) / 256), 256, 0, st>>>(v, sv, nv, U->in_ch, U->x.p);
  OCMP_HIP(hipGetLastError());
  unet_forward(U, st);
  k_embed<T><<<(unsigned)((nv + 255) / 256), 256, 0, st>>>(U->y.p, nv, U->out_ch, freem, z, accumulate);
  OCMP_HIP(hipGetLastError());
}
template void unet_correct<float>(Unet *, const uint8_t *, const float *, const double *, float *, int);
template void unet_correct<double>(Unet *, const uint8_t *, const double *, const double *, double *, int);

// One Adam step of the self-supervised loss  L = 1/(2B) sum_b ||r_b - A E U(R v_b)||^2,
// r_b = v_b - A D^{-1} v_b, over the batch of samples v_b = samples[idx[b]] (device f64 rows
// of length A->n). Returns the batch loss (host, syncs).
double unet_train_step(Unet *U, Matrix *A, Precond *P, const double *samples, int64_t ld, const int *idx, int nb,
                       float lr) {
  OCMP_REQUIRE(P && P->kind == OCMP_PRE_JACOBI && P->ready, "unet training needs an updated Jacobi preconditioner");
  OCMP_REQUIRE(nb >= 1, "unet: empty batch");
  hipStream_t st = U->ctx->stream;
  const int64_t n = A->n;
  const uint8_t *fm = A->space->d_free.p;
  U->rj.alloc(n);
  U->tmp.alloc(n);
  U->res.alloc(n);
  U->zf.alloc(n);
  U->spart.alloc((size_t)SUB_GRID * (nb + 1));  // row nb: scratch partials of r_J
  const int64_t nv = U->zn[0];
  const unsigned gb = (unsigned)((n + 255) / 256);
  for (int b = 0; b < nb; ++b) {
    const double *vb = samples + (size_t)idx[b] * ld;
    // r_J = v - A D^{-1} v
    k_mul_diag<<<gb, 256, 0, st>>>(n, P->dinv.p, vb, U->tmp.p);
    spmv_f64(A, U->tmp.p, U->res.p);
    k_sub_masked<<<SUB_GRID, 256, 0, st>>>(n, vb, U->res.p, fm, U->rj.p, U->spart.p + (size_t)nb * SUB_GRID);
    // z = E U(R v) (zero elsewhere), res = r_J - A z, loss partials
    OCMP_HIP(hipMemsetAsync(U->zf.p, 0, n * sizeof(double), st));
    unet_correct<double>(U, fm, vb, nullptr, U->zf.p, 1);
    spmv_f64(A, U->zf.p, U->tmp.p);
    k_sub_masked<<<SUB_GRID, 256, 0, st>>>(n, U->rj.p, U->tmp.p, fm, U->res.p, U->spart.p + (size_t)b * SUB_GRID);
    // dL/dz = -A^T res = -A res (the Stokes systems are symmetric); restricted to the grid
    spmv_f64(A, U->res.p, U->tmp.p);
    k_neg_masked<<<gb, 256, 0, st>>>(n, fm, U->tmp.p);
    k_restrict_f32<<<(unsigned)((nv + 255) / 256), 256, 0, st>>>(U->tmp.p, nv, U->out_ch, U->gy.p);
    OCMP_HIP(hipGetLastError());
    // U->x still holds R v of this sample, the activations its forward pass
    unet_backward(U, b > 0 ? 1 : 0, st);
  }
  U->step += 1;
  const float c1 = (float)(1.0 / (1.0 - std::pow(0.9, (double)U->step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow(0.999, (double)U->step)));
  k_adam_step<<<(NP + 255) / 256, 256, 0, st>>>(NP, U->grad.p, 1.f / (float)nb, U->p.p, U->m.p, U->v.p, U->pb.p, lr,
                                                c1, c2);
  OCMP_HIP(hipGetLastError());
  std::vector<double> h((size_t)SUB_GRID * nb);  // rows 0..nb-1: the loss partials
  OCMP_HIP(hipMemcpyAsync(h.data(), U->spart.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  OCMP_HIP(hipStreamSynchronize(st));
  double loss = 0.0;
  for (double v : h) loss += v;
  return 0.5 * loss / nb;
}

}  // namespace ocmp
