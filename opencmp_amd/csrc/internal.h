// Internal declarations of the MI355X assemble+solve library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <memory>
#include <vector>

#include "../../include/ocmp.h"

namespace ocmp {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string &msg);

#define OCMP_HIP(call)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ::ocmp::Error(OCMP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define OCMP_REQUIRE(cond, msg)                                         \
  do {                                                                  \
    if (!(cond)) throw ::ocmp::Error(OCMP_ERR_ARG, std::string(msg));   \
  } while (0)

// Device buffer owned by the library.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count) OCMP_HIP(hipMalloc((void **)&p, count * sizeof(T)));
    n = count;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) OCMP_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// ------------------------------------------------------------------ context

struct TimingRec {
  int cls;
  hipEvent_t a, b;
  double bytes;  // algorithmic bytes of the timed launch(es)
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  bool timing = false;
  std::vector<std::string> class_names;
  std::map<std::string, int> class_ids;
  std::vector<TimingRec> pending;
  std::vector<hipEvent_t> event_pool;
  std::map<int, std::pair<long, double>> acc;  // class -> (launches, ms)
  std::map<int, double> acc_bytes;               // class -> algorithmic bytes
  // reduction scratch (device) and pinned host mirror
  DBuf<double> red;        // partial sums
  DBuf<double> scal;       // device scalars
  double *host_scal = nullptr;  // pinned
  // this device's copy of the allocation-free entry points' static scratch array (device address)
  double *noalloc_scratch = nullptr;
  int cls(const std::string &name);
  void tic(int c, hipEvent_t *a);
  void toc(int c, hipEvent_t a, double bytes);
  hipEvent_t get_event();
  void flush_timing();
};

// RAII kernel-class timer: records events around launches when timing is on.
// `bytes` is the algorithmic HBM traffic of the timed launch (what the
// kernel must move at minimum), accumulated per class for roofline reports.
struct KTimer {
  Ctx *ctx;
  int c;
  double bytes;
  hipEvent_t a = nullptr;
  KTimer(Ctx *x, int cls, double alg_bytes = 0.0) : ctx(x), c(cls), bytes(alg_bytes) {
    if (ctx->timing) ctx->tic(c, &a);
  }
  ~KTimer() {
    if (ctx->timing && a) ctx->toc(c, a, bytes);
  }
};

// ------------------------------------------------------------------ mesh

struct Mesh {
  Ctx *ctx;
  int dim;
  int64_t nv, nel, nedge = 0, nface = 0, nfacet = 0, ninterior = 0, nbnd = 0;
  std::vector<double> coords;          // nv*dim
  std::vector<int32_t> elems;          // nel*(dim+1), sorted per element
  std::vector<int32_t> edges;          // nedge*2
  std::vector<int32_t> elem_edges;     // nel*nle
  std::vector<int32_t> faces;          // nface*3 (3D)
  std::vector<int32_t> elem_faces;     // nel*4 (3D)
  std::vector<int32_t> elem_facets;    // nel*(dim+1): facet id of local facet k (opposite vertex k)
  std::vector<int32_t> facet_verts;    // nfacet*dim
  std::vector<int32_t> f_e0, f_k0, f_e1, f_k1, f_marker;  // per facet
  // per element-facet neighbour: nb[e*(dim+1)+k] = neighbour element or -(marker+1) on boundary
  std::vector<int32_t> elem_nb;
  std::vector<int8_t> elem_nbk;       // neighbour's local facet index
  DBuf<double> d_coords;
  DBuf<int32_t> d_elems;
  DBuf<int32_t> d_elem_nb;
  DBuf<int8_t> d_elem_nbk;
};

// ------------------------------------------------------------------ space

struct Block {
  int kind, order, dim;
  int64_t dof_off;  // offset in compound numbering
  int64_t n1;       // dofs in this block
  int loc_off;      // offset in compound element dof list
  int nloc;         // local dofs of this block
  uint64_t dirichlet_mask;
  int comp;         // index of the ocmp_fe_desc it came from
};

struct Space {
  Mesh *mesh;
  bool dgjumps = false;
  std::vector<Block> blocks;
  int64_t ndof = 0;
  int nloc = 0;
  std::vector<int32_t> edofs;   // nel*nloc (compound numbering)
  std::vector<uint8_t> free;    // ndof
  DBuf<int32_t> d_edofs;
  DBuf<uint8_t> d_free;
  // Dirichlet lift support per block: dirichlet vertex/edge/face lists with marker masks
  // (computed lazily per (block, marker mask))
  struct Lift {
    DBuf<int32_t> verts;   // vertex ids
    DBuf<int32_t> edges;   // edge ids
    DBuf<int32_t> faces;   // face ids (3D)
    DBuf<int32_t> fedges;  // per listed face: edge ids of (a,b), (a,c), (b,c), a < b < c
    int64_t nverts = 0, nedges = 0, nfaces = 0;
  };
  std::map<std::pair<int, uint64_t>, Lift *> lifts;
  DBuf<int32_t> d_edges;   // mesh edges (for lifts)
  // whole-domain set of order-3 H1 blocks: the triangles carrying a face-bubble dof (3D faces,
  // 2D cells), their vertex triples and edge ids (built on first use)
  DBuf<int32_t> d_tri_verts, d_tri_edges;
  int64_t ntri = 0;
  // deterministic element schedule for H1 scatter: elements sharing a vertex get different colours
  int ncolours = 0;
  std::vector<int64_t> colour_off;   // ncolours+1
  DBuf<int32_t> d_colour_elems;      // elements sorted by colour
  std::vector<int64_t> bnd_colour_off;  // ncolours+1: the elements with a boundary facet, per colour
  DBuf<int32_t> d_bnd_colour_elems;
  // per boundary-marker set: the elements with a facet on those markers, coloured among themselves
  // (elements sharing a vertex differ), for boundary-only assembly calls
  struct BndSet {
    std::vector<int64_t> off;
    DBuf<int32_t> elems;
  };
  std::map<uint64_t, std::unique_ptr<BndSet>> bnd_sets;
  // cached kernel tables keyed by a signature
  std::map<std::string, DBuf<double> *> tables;
  ~Space() {
    for (auto &kv : tables) delete kv.second;
    for (auto &kv : lifts) delete kv.second;
  }
};

// ------------------------------------------------------------------ reference tables

struct Rule {
  int dim = 0, nq = 0;
  std::vector<double> lam;  // nq*(dim+1)
  std::vector<double> w;    // nq
};
Rule simplex_rule(int dim, int degree);
void line_rule(int degree, std::vector<double> &t, std::vector<double> &w);
int nloc1(int kind, int dim, int order);
// values phi[nl] and reference gradients g[nl*dim] at one barycentric point
void eval_basis(int kind, int dim, int order, const double *lam, double *phi, double *g);

// ------------------------------------------------------------------ matrix

// A bfloat16 storage value: the upper 16 bits of an FP32 number (conversion to float is exact).
struct bf16s {
  uint16_t b;
  bf16s() = default;
  __host__ __device__ explicit bf16s(int) : b(0) {}
  __host__ __device__ operator float() const {
    union {
      uint32_t u;
      float f;
    } c;
    c.u = (uint32_t)b << 16;
    return c.f;
  }
};

// Rows grouped into blocks of at most nnzb nonzeros (block b = rows [rblk[b], rblk[b+1])), the
// distinct columns of each block (ucol[ubase[b] .. ubase[b+1]), sorted) and every nonzero's
// 16-bit position in its block's list.
struct RowBlocks {
  int nnzb = 0;  // 0: not built
  std::vector<int32_t> rblk;
  DBuf<int32_t> d_rblk, d_ubase, d_ucol;
  DBuf<uint16_t> d_lcol;
  int64_t nrblk = 0, nucol = 0;
};

struct Work;                  // a matrix' Krylov scratch (linalg.hip), created on first solve
void work_release(Work *w);   // frees it (device buffers, pinned host buffer, events)

struct Matrix {
  Space *space;
  Ctx *ctx;
  int64_t n = 0, nnz = 0;
  bool dg = false;
  std::vector<int32_t> row_ptr, col;
  DBuf<int32_t> d_row_ptr, d_col;
  DBuf<double> d_val;
  // H1 boundary-facet scatter: element -> CSR slot map (colours live in the space), built on the
  // first assembly with boundary terms
  DBuf<int32_t> d_slot;              // nel*nloc*nloc
  // H1 row-owner gather (volume terms): per CSR row its element entries in element order
  // (patch), every entry's compound local row index and the position in the row of each of the
  // element's nloc columns (1 byte when every row has at most 256 nonzeros, else 2)
  DBuf<int32_t> d_pptr;              // n+1
  DBuf<int32_t> d_pe;                // entries: element
  DBuf<uint8_t> d_pslot8;            // entries x nloc (maxrow <= 256)
  DBuf<uint16_t> d_pslot16;          // entries x nloc (maxrow > 256)
  int64_t npatch = 0;
  int maxrow = 0;
  // row blocks of the gather: rows [gblk[b], gblk[b+1]) with at most gemax element entries and
  // gaccmax nonzeros (staged in LDS); empty when some row alone would not fit. Within a block the
  // entries are stored sorted by local row index (d_pe, d_pslot in that order): d_gloff[b*(nl+1)+L]
  // is where local row L starts (relative to the block's first entry), d_gpos[p] the sorted position
  // of the p-th entry in row order (pptr indexes row order)
  std::vector<int32_t> gblk;
  DBuf<int32_t> d_gblk;
  DBuf<uint16_t> d_gpos, d_gloff;
  DBuf<uint8_t> d_pli;               // sorted entries: local row index
  int gemax = 0, gaccmax = 0;
  DBuf<double> d_factors;            // nel x (factors per element) of the last assembly
  DBuf<int32_t> d_dslot;             // n: CSR slot of the diagonal entry (-1: none), built on first use
  DBuf<char> d_gdesc;                // the gather's kernel descriptors
  // DG: per element-facet block position in the block row (0..3), diag block position
  DBuf<int8_t> d_dg_pos;             // nel*(dim+2): pos of diag block, then of each facet neighbour (-1 none)
  DBuf<int32_t> d_dg_bcol;           // nel*4 block columns (-1 padded), sorted
  int dg_nb = 0;                     // local dofs per element (block size)
  // LDS-staged SpMV row blocks: FP64 values use blocks of SPMV_NNZB64 nonzeros, FP32 values
  // (built with the FP32 copy) blocks of SPMV_NNZB32
  RowBlocks rb64, rb32;
  bool assembled = false;
  // mixed precision copy
  DBuf<float> d_val32;
  DBuf<bf16s> d_valbf;   // BF16 copy (the inner solves of a preconditioner stored in BF16)
  // solver scratch owned by this matrix (one per matrix, so distinct matrices never share it)
  Work *work = nullptr;
  // block pairs (test block * nblocks + trial block; nblocks <= 8) that the terms since the last
  // non-accumulating assembly touched, and a counter bumped by every assembly
  uint64_t active_pairs = 0, version = 0;
  uint64_t val32_version = ~0ull;  // the `version` whose values d_val32 holds (the FP32 copy of d_val)
  // the Krylov operator: this matrix without the entries of untouched block pairs (structural
  // zeros, e.g. the two velocity components' cross blocks), values gathered from d_val when stale
  Matrix *op = nullptr;
  uint64_t op_pairs = 0, op_version = ~0ull, op_version32 = ~0ull;
  DBuf<int32_t> d_src;               // in an operator: each entry's slot in the parent's values
  DBuf<int32_t> d_rowmap;            // in an operator: the parent's row of each (reordered) row
  Matrix() = default;
  Matrix(const Matrix &) = delete;
  Matrix &operator=(const Matrix &) = delete;
  ~Matrix() {
    work_release(work);
    delete op;
  }
};

struct Precond {
  Matrix *A;
  int kind, precision;
  // saddle-point systems: rows of these scalar blocks get the diagonal Schur complement
  // approximation -1/(C_ii + sum_j B_ij^2 / A_jj) instead of 1/a_ii (kind stays Jacobi)
  uint64_t schur_blocks = 0;
  struct Unet *unet = nullptr;  // learned correction added to the Jacobi part by FGMRES
  DBuf<uint8_t> schur_flag;  // n: 1 on rows of the Schur blocks
  DBuf<double> dinv;     // Jacobi: n (0 on Dirichlet dofs)
  DBuf<float> dinv32;
  DBuf<double> binv;     // block Jacobi: nel*nb*nb
  DBuf<float> binv32;
  // right-preconditioned DG operator: Ahat = A * blockdiag^{-1}, off-diagonal blocks only
  DBuf<double> ahat;     // n*3*nb, column-major ELL (facet-neighbour blocks)
  DBuf<float> ahat32;
  // A itself in the same ELL layout with the diagonal block as slot 0 (FP64 outer residual)
  DBuf<double> aell;     // n*4*nb
  bool ready = false;
  // Schur inner solve (FGMRES only; 0: off): on the Schur rows [s0, s1) the flexible
  // preconditioner takes schur_iters PCG iterations on S = C + B D^{-1} B^T in place of the
  // Schur-diagonal step, D_uu = max(|a_uu|, 1e-2 max_j |a_uj|). S's pattern is built once from
  // A's, its values at every update.
  int schur_iters = 0;     // at most this many PCG iterations per step
  double schur_tol = 0.0;  // stop when (r, dinv r) <= tol^2 (r0, dinv r0) (0: always schur_iters)
  int schur_cap = 0;       // with a tolerance: iterations enqueued per step, adapted to those needed
  double schur_seen = 0.0; // inner iterations of the running solve already accounted for
  int64_t s0 = 0, s1 = 0;
  Matrix *S = nullptr;
  DBuf<int32_t> d_sseg;  // per row of A: first and one-past-last CSR slot of its columns in [s0, s1)
  DBuf<double> sdpos;    // n: the positive velocity scale S is built with (0 off the free velocity rows)
  DBuf<double> sbf;      // per slot of the Schur rows: a_pu sdpos[u], the first factor of S's products
  // S's values as lists of products: for entry k of S, the A slot of its C entry (d_scslot, -1:
  // none) and the (a_pu slot, a_uq slot) pairs [d_spptr[k], d_spptr[k+1]) in row order of A;
  // empty (values by search instead) when the lists would exceed SCHUR_MAX_PAIRS
  DBuf<int32_t> d_scslot, d_spptr, d_spj, d_spt;
  RowBlocks srb;         // S's row blocks for the fused inner iteration (SCG_NNZB nonzeros each)
  DBuf<double> sdinv;    // Jacobi of S: 1/S_ii, 0 on rows left out (non-free, or S_ii negligible)
  DBuf<float> sdinv32;
  DBuf<double> svd;      // PCG vectors: r, s, w (two each, by iteration parity), p, u (8 ns)
  DBuf<float> svf;
  DBuf<double> spart;    // PCG scalars and partial sums
  DBuf<int> sflag;       // converged flag of the running inner solve
  // Block-triangular step (FGMRES; the Schur rows are the last ones, [s0, n)): on the Schur rows
  // z_p = -s v_p / d_p with d the caller's diagonal Schur approximation (rows where d is not above
  // 1e-10 max d, or not free, get 0), then z_u = A_uu^{-1} (s v_u - A_up z_p) by at most vel_iters
  // Jacobi-PCG iterations on the velocity block, stopped at (r, D^-1 r) <= vel_tol^2 (r0, D^-1 r0).
  bool btri = false;
  int vel_iters = 0, vel_cap = 0;
  double vel_tol = 0.0, vel_seen = 0.0;
  DBuf<double> pdiag;    // the caller's d on the Schur rows
  DBuf<double> pdinv;    // 1/d (0: left out)
  DBuf<float> pdinv32;
  Matrix *Auu = nullptr, *Aup = nullptr;  // A's velocity block and velocity-pressure coupling
  DBuf<int32_t> d_uuslot, d_upslot;        // the A slot of each of their values
  RowBlocks vrb;                           // Auu's row blocks for the fused iteration
  DBuf<double> vdinv;    // Jacobi of Auu (0 on Dirichlet rows)
  DBuf<float> vdinv32;
  DBuf<double> vvd;      // PCG vectors (8 n_u) and A_up z_p (n_u)
  DBuf<float> vvf;
  DBuf<double> vpart;    // PCG scalars and partial sums
  DBuf<int> vflag;
  // mixed path: keep the inner solves' matrices (S, the velocity block) in BF16 (FP32 vectors, FP64
  // sums); they are part of the preconditioner, which FGMRES lets vary
  bool bf16 = false;
  // Schur rows by smoothed-aggregation AMG V-cycles on S (FGMRES; replaces the inner PCG): the
  // hierarchy (aggregates, patterns, product lists) is built on the host at the first update with S's
  // values, its values on the device at every update
  int schur_amg = 0;       // V-cycles per preconditioner step (0: off)
  // with a flexible Schur step: the velocity rows as the upper block-triangular step,
  // z_u = D^-1 (s v_u - A_up z_p) after z_p, instead of D^-1 s v_u
  bool schur_tri = false;
  uint64_t vel_pairs = 0;  // A's active block pairs when Auu / Aup were built (rebuilt when they change)
  struct Amg *amg = nullptr;
  int64_t amg_cycles = 0;  // V-cycles of the running solve
  Precond() = default;
  Precond(const Precond &) = delete;
  Precond &operator=(const Precond &) = delete;
  ~Precond();
};

// ------------------------------------------------------------------ kernels (launchers)

void launch_assemble_h1(Matrix *A, const std::vector<ocmp_term> &terms, const std::vector<ocmp_field> &fields,
                        bool accumulate);
void launch_assemble_dg(Matrix *A, const std::vector<ocmp_term> &terms, bool accumulate);
void launch_assemble_linear(Space *s, double *b, const std::vector<ocmp_lterm> &terms,
                            const std::vector<ocmp_field> &fields, bool accumulate);
void launch_dirichlet_set(Space *s, int block, double *x, int fn, const double *c, uint64_t mask);
// GridFunction.Set(g) on one scalar block over the whole domain: projection-based interpolation
// (vertex values, edge and face projections) on H1, element-wise L2 projection on L2
void launch_set(Space *s, int block, double *x, int fn, const double *c);
// config expressions (OCMP_FN_EXPR): register a validated postfix program on ctx's device (returns
// its id), and check a closed-form function id and its parameters before a launch
int expr_register(Ctx *ctx, int n, const double *code);
void check_fn(const Ctx *ctx, int fn, const double *c);
void spmv_f64(Matrix *A, const double *x, double *y);
void solve(Matrix *A, Precond *P, const double *b, double *x, const ocmp_solve_opts &o, ocmp_solve_stats &st);
void precond_update(Precond *P);
void precond_update_jacobi(Precond *P);
void residual_norms(Matrix *A, const double *b, const double *x, double *out);
bool hessenberg_eigvals(int n, const double *H, int ldh, double *wr, double *wi);
void newton_shifts(int n, const double *wr, const double *wi, int s, double *d, double *c);
double residual_vector(Matrix *A, const double *b, const double *x, double *r);
void error_norms(Space *s, int block, const double *x, const double *ref, int fn, const double *c, double *out);
void divergence_norm(Space *s, int block0, const double *x, double *out);
void error_norms_vec(Space *s, int nb, const int *blocks, const double *x, const double *ref, const int *fns,
                     const double *cs, double *out);
void l2_pair_norms(Space *s, int nb, const int *blocks, const double *a, const double *b, double *out);
void linf_norm(Space *s, int block, const double *x, const double *ref, double *out);
void axpby(Ctx *ctx, int64_t n, double a, const double *x, double b, double *y);
// allocation-free entry points: the device address of the static scratch array on `device`
// (looked up once per device, with that device current), and the residual norms computed in it
double *noalloc_scratch_init(int device);
void residual_norms_noalloc(Matrix *A, const double *b, const double *x, double *out_dev);

// per-DOF MLP initial guess (mlp.hip)
struct Mlp;
Mlp *mlp_create(Ctx *ctx, const float *params, int64_t max_batch);
void mlp_destroy(Mlp *M);
void mlp_set_params(Mlp *M, const float *params);
void mlp_get_params(Mlp *M, float *params);
void mlp_set_scales(Mlp *M, int nseg, const int64_t *off, const double *s, const double *U);
void mlp_get_scales(Mlp *M, int *nseg, int64_t *off, double *s, double *U);
void mlp_fit_scales(Mlp *M, const double *snap, int nsnap, int64_t ld, int nseg, const int64_t *off);
void mlp_sample(Mlp *M, const double *snap, int nsnap, int64_t ld, int64_t n, uint64_t seed, int64_t batch,
                float *X, float *T);
void mlp_forward(Mlp *M, int64_t n, const float *X, float *Y);
void mlp_grad(Mlp *M, int64_t n, const float *X, const float *T, float *grad, double *loss);
void mlp_train_step(Mlp *M, int64_t n, const float *X, const float *T, float lr, double *loss);
void mlp_set_base(Mlp *M, int base);
int mlp_get_base(const Mlp *M);
void mlp_guess(Mlp *M, int64_t n, const double *u0, const double *u1, const double *u2, const double *u3,
               double *out);

// learned U-net preconditioner correction (unet.hip)
struct Unet;
Unet *unet_create(Ctx *ctx, int N, const float *params);
void unet_destroy(Unet *U);
int64_t unet_grid_points(const Unet *U);  // (N+1)^3
int64_t unet_max_index(const Unet *U);    // 1 + the largest dof index a channel touches
void unet_fit_channels(Unet *U, Precond *P, const double *samples, int ns, int64_t ld, const int64_t *off);
void unet_get_channels(const Unet *U, int64_t *off, double *in_scale, double *out_scale);
int64_t unet_activation(Unet *U, int layer, float *out);  // element count; copies if out
void unet_set_params(Unet *U, const float *params);
void unet_get_params(Unet *U, float *params);
void unet_set_channels(Unet *U, const int64_t *off, const double *in_scale, const double *out_scale);
void unet_forward_grid(Unet *U, const float *x, float *y);
void unet_grad_grid(Unet *U, const float *x, const float *gy, float *grad);
template <typename T>
void unet_correct(Unet *U, const uint8_t *freem, const T *v, const double *sv, T *z, int accumulate);
double unet_train_step(Unet *U, Matrix *A, Precond *P, const double *samples, int64_t ld, const int *idx, int nb,
                       float lr);

// greedy colouring + slot map (host)
void build_h1_schedule(Matrix *A);
void build_h1_gather(Matrix *A);
void build_colouring(Space &s);
void build_pattern(Matrix *A);
// nonzeros per row block of the LDS-staged SpMV (LDS: 16 KB per block in both precisions;
// measured on H1 P3, 7.2M dofs: FP64 1.01 ms at 2048 against 1.04 at 3072 and 1.20 at 4096,
// FP32 0.69 ms at 4096 against 0.83 at 2048 and 0.85 at 8192)
constexpr int SPMV_NNZB64 = 2048;
constexpr int SPMV_NNZB32 = 4096;
// the fused Schur PCG iteration: row blocks of S of this many nonzeros, SCG_NT threads a block
constexpr int SCG_NNZB = 4096;
constexpr int SCG_NT = 512;
void build_row_blocks(Matrix *A, RowBlocks &rb, int nnzb);
void build_solver_operator(Matrix *A);
void ensure_row_blocks32(Matrix *A);
// the pattern of a Schur preconditioner's S = C + B diag(A)^{-1} B^T and the per-row Schur
// column ranges of A (host, once per preconditioner)
void build_schur_matrix(Precond *P);

// ------------------------------------------------------------------ AMG on the Schur matrix
//
// Level l: matrix A (level 0 is P->S, not owned), prolongation P (n x nc) and restriction R = P^T
// (nc x n) as CSR matrices, the Jacobi of A, and the device lists that give P, R and the next
// level's A from A's values: P[e] = ident[e] - omega dinv[prow[e]] sum_{s in plist(e)} A[s];
// R[k] = P[rslot[k]]; Q = A P with Q[q] = sum over (A slot, P slot) pairs; A_next = P^T Q with
// A_next[c] = sum over (P slot, Q slot) pairs. The last level is solved by a dense inverse.
struct AmgLevel {
  // W = (I - w D^-1 A) P (w the smoother's damping) folds the damped-Jacobi smoothing steps of a
  // V(1,1) cycle into the grid transfers: the restricted residual after pre-smoothing from zero is
  // W^T b, and the post-smoothed iterate is z + W x_c with z = w D^-1 (2 b - w A D^-1 b)
  Matrix *A = nullptr, *P = nullptr, *W = nullptr, *Wt = nullptr;
  bool own_A = false;
  double omega = 0.0;                       // prolongation smoothing weight 4 / (3 lambda_max(D^-1 A))
  int64_t nq = 0;                           // nonzeros of Q = A P
  DBuf<int32_t> d_prow, d_plptr, d_plist;   // per P entry: row, list range; A slots
  DBuf<uint8_t> d_pident;                   // per P entry: 1 where the column is the row's own aggregate
  DBuf<int32_t> d_wrow, d_wps, d_wqs;       // per W entry: row, P slot and Q slot (-1: none)
  DBuf<int32_t> d_wtslot;                   // per W^T entry: its W slot
  DBuf<int32_t> d_qptr, d_qa, d_qp;         // Q entries' (A slot, P slot) pair lists
  DBuf<int32_t> d_cptr, d_cp, d_cq;         // next A's (P slot, Q slot) pair lists
  DBuf<double> q;                           // Q values
  DBuf<double> dinv;                        // Jacobi of A (0 on rows left out)
  DBuf<float> dinv32;
  DBuf<double> vd;                          // V-cycle vectors: b, x, z (3 n; level 0 also the
                                            // further cycles' residual and correction)
  DBuf<float> vf;
};
struct Amg {
  std::vector<std::unique_ptr<AmgLevel>> lv;
  int64_t ncoarse = 0;                      // rows of the last level
  DBuf<double> cinv;                        // its dense inverse (row-major)
  DBuf<float> cinv32;
  DBuf<double> cden;                        // scratch for the inversion (ncoarse x 2 ncoarse)
  // the levels from `tail` down, as one dense matrix of level tail's rows: the V-cycle from there
  // on is linear in its right-hand side (fixed Jacobi weights, the dense inverse at the bottom), so
  // it is composed once per update and applied by one dense product; tail = the last level when no
  // level above it is small enough
  size_t tail = 0;
  int64_t tld = 0;                          // leading dimension of tail's matrix (rows padded with zeros)
  DBuf<double> tinv, tscr;                  // tail's matrix (row-major); composition scratch
  DBuf<float> tinv32;
  ~Amg();
};
void build_schur_amg(Precond *P, const std::vector<double> &sval, const std::vector<double> &sdinv);
void ensure_diag_slots(Matrix *A);
void build_velocity_blocks(Precond *P);
constexpr int64_t SCHUR_MAX_PAIRS = (int64_t)1 << 28;

}  // namespace ocmp

struct ocmp_ctx { ocmp::Ctx c; };
struct ocmp_mesh { ocmp::Mesh m; };
struct ocmp_space { ocmp::Space s; };
struct ocmp_matrix { ocmp::Matrix A; };
struct ocmp_precond { ocmp::Precond P; };
struct ocmp_mlp { ocmp::Mlp *M; ocmp::Ctx *ctx; };
struct ocmp_unet { ocmp::Unet *U; ocmp::Ctx *ctx; };
