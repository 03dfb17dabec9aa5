/*
 * ocmp.h -- C ABI of the MI355X (gfx950) assemble+solve backend for OpenCMP models.
 *
 * The reference has no FFI of its own: OpenCMP calls NGSolve through pybind11
 * (BilinearForm.Assemble, LinearForm.Assemble, Preconditioner, solvers.CG/GMRes,
 * GridFunction.Set). Each entry point below names the reference call site it
 * replaces. Paths are relative to the OpenCMP source tree.
 *
 * Conventions
 *  - Every function returns int: OCMP_OK (0) or a negative error code;
 *    ocmp_last_error() returns a thread-local message for the last failure.
 *  - Host arrays passed in are borrowed for the duration of the call.
 *  - Pointers documented as "device" are HIP device pointers owned by the caller
 *    (e.g. the storage of a torch tensor on cuda:N); the library never frees them.
 *  - Each context owns one HIP stream; all work is enqueued on it. Calls are
 *    synchronous with respect to the host only where documented (solve, copies
 *    to host). A handle is not thread-safe; distinct contexts are independent.
 */
#ifndef OCMP_H
#define OCMP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OCMP_OK = 0,
  OCMP_ERR_ARG = -1,         /* invalid argument (shape, index, NULL)      */
  OCMP_ERR_HIP = -2,         /* HIP runtime failure                         */
  OCMP_ERR_STATE = -3,       /* call sequence error (e.g. not assembled)    */
  OCMP_ERR_UNSUPPORTED = -4, /* valid request outside the implemented set   */
  OCMP_ERR_INTERNAL = -5,
  OCMP_ERR_NUMERIC = -6      /* an iteration did not converge               */
};

typedef struct ocmp_ctx ocmp_ctx;
typedef struct ocmp_mesh ocmp_mesh;
typedef struct ocmp_space ocmp_space;
typedef struct ocmp_matrix ocmp_matrix;
typedef struct ocmp_precond ocmp_precond;

/* ------------------------------------------------------------------ context */

/* Replaces the ngcore.TaskManager() scope of opencmp/run.py:57-61.
 * hip_stream: the hipStream_t to run on; NULL selects the device's null
 * stream (PyTorch's default stream). The caller keeps ownership. */
int ocmp_ctx_create(int device, void *hip_stream, ocmp_ctx **out);
int ocmp_ctx_destroy(ocmp_ctx *ctx);
int ocmp_ctx_sync(ocmp_ctx *ctx);
void *ocmp_ctx_stream(ocmp_ctx *ctx);
/* Per-kernel-class HIP event timing (no reference counterpart; the reference
 * has no timers, opencmp/run.py:58). */
int ocmp_ctx_timing(ocmp_ctx *ctx, int enable);
/* Writes a JSON object {"class": [launches, total_ms, algorithmic_bytes], ...} and resets. */
int ocmp_ctx_timing_report(ocmp_ctx *ctx, char *buf, int64_t cap);
const char *ocmp_last_error(void);
int ocmp_version(void);

/* --------------------------------------------------------------------- mesh */

/* Replaces ngs.Mesh(...) from load_mesh (opencmp/helpers/io.py:69-110).
 * coords: host (nv, dim) f64; elems: host (nel, dim+1) i32 vertex ids;
 * bnd_facets: host (nbnd, dim) i32; bnd_markers: host (nbnd) i32 in [0, 64).
 * dim is 2 (triangles) or 3 (tetrahedra). */
int ocmp_mesh_create(ocmp_ctx *ctx, int dim, int64_t nv, const double *coords, int64_t nel,
                     const int32_t *elems, int64_t nbnd, const int32_t *bnd_facets,
                     const int32_t *bnd_markers, ocmp_mesh **out);
/* out[8] = {nv, nedge, nface, nel, nfacet, n_interior_facets, nbnd, dim} */
int ocmp_mesh_info(const ocmp_mesh *mesh, int64_t *out);
int ocmp_mesh_destroy(ocmp_mesh *mesh);

/* -------------------------------------------------------------------- space */

enum { OCMP_H1 = 0, OCMP_L2 = 1 };

typedef struct {
  int32_t kind;            /* OCMP_H1 or OCMP_L2                                  */
  int32_t order;           /* H1: 1..3; L2: 0..3 (triangles)                       */
  int32_t ncomp;           /* 1, or dim for VectorH1                               */
  int32_t reserved;
  uint64_t dirichlet_mask; /* bit m set: boundary marker m is Dirichlet (H1 only)  */
} ocmp_fe_desc;

/* Replaces Model._construct_fes: ngs.FESpace([H1|VectorH1|L2 ...], dgjumps=...)
 * (opencmp/models/poisson.py:60-64, opencmp/models/ins.py:95-128,
 * opencmp/models/multi_component_ins.py:79-102). Components are numbered in
 * "blocks": a VectorH1 component expands into ncomp consecutive scalar blocks. */
int ocmp_space_create(ocmp_mesh *mesh, int ncomp, const ocmp_fe_desc *comps, int dgjumps,
                      ocmp_space **out);
/* out[4] = {ndof, nloc (compound element dofs), nblocks, nfree} */
int ocmp_space_info(const ocmp_space *s, int64_t *out);
/* out[nblocks*4] = {dof offset, ndof, element-local offset, nloc} per block */
int ocmp_space_blocks(const ocmp_space *s, int64_t *out);
/* Replaces fes.FreeDofs() (opencmp/models/base_model.py:906): host u8[ndof]. */
int ocmp_space_freedofs(const ocmp_space *s, uint8_t *mask);
/* host i32[nel*nloc]: the compound element dof table (for parity checks). */
int ocmp_space_element_dofs(const ocmp_space *s, int32_t *out);
int ocmp_space_destroy(ocmp_space *s);

/* ------------------------------------------------------------------- matrix */

/* Builds the NGSolve-style CSR pattern (full element cliques, plus facet
 * neighbours when dgjumps) and the deterministic assembly schedule: element
 * colouring plus element->CSR slot map (H1) or owner-computes block rows (L2).
 * Replaces the pattern half of BilinearForm(fes) + Assemble. */
int ocmp_matrix_create(ocmp_space *s, ocmp_matrix **out);
/* out[4] = {n, nnz, ncolours, is_dg} */
int ocmp_matrix_info(const ocmp_matrix *A, int64_t *out);
int ocmp_matrix_pattern(const ocmp_matrix *A, int32_t *row_ptr, int32_t *col);  /* host */
int ocmp_matrix_values(const ocmp_matrix *A, double *vals);                    /* host, syncs */
int ocmp_matrix_destroy(ocmp_matrix *A);

/* --------------------------------------------------------------- integrators */

enum {
  OCMP_BFI_GRADGRAD = 0, /* coef * w * grad u . grad v                         */
  OCMP_BFI_MASS = 1,     /* coef * w * u v                                     */
  OCMP_BFI_CONV = 2,     /* coef * w * u (wind . grad v)                       */
  OCMP_BFI_DTRIAL = 3,   /* coef * w * d_dir u * v                             */
  OCMP_BFI_DTEST = 4,    /* coef * w * u * d_dir v                             */
  OCMP_BFI_BND_UPWIND = 5, /* coef * max(wind.n, 0) * u v on boundary facets whose marker is in
                              dirichlet_mask (H1, trial == test): the no-backflow part of the stress
                              boundary condition, opencmp/models/ins.py:210-216 */
  OCMP_BFI_BND_MASS = 6,   /* coef * u v on boundary facets whose marker is in dirichlet_mask
                              (H1, trial == test): the Robin term, opencmp/models/poisson.py:94-96 */
  OCMP_BFI_BND_FLUX = 7,   /* coef * (wind.n) * u v on boundary facets whose marker is in
                              dirichlet_mask (H1, trial == test): the convective part of a Neumann
                              boundary of a transported species, opencmp/models/multi_component_ins.py:277-285 */
  OCMP_BFI_DG_FACET = 16 /* interior + boundary facet terms of an L2 block, see below */
};

enum { OCMP_WEIGHT_NONE = 0, OCMP_WEIGHT_FIELD = 1, OCMP_WEIGHT_ONE_MINUS_FIELD = 2 };

/* A grid function on one scalar H1 block of a space (device dof vector of that space). */
typedef struct {
  const ocmp_space *space;
  int32_t block;
  int32_t reserved;
  const double *vec; /* device, length = ndof of `space` */
} ocmp_field;

/* One bilinear term. Volume terms: trial/test are scalar block indices. The
 * weight w is 1, a field phi, or 1-phi (phi-weighting of the diffuse-interface
 * models, opencmp/models/stokes_dim.py:42-87). elem_coef (device, per element)
 * multiplies coef when non-NULL (per-element diffusivity). The wind is the
 * constant `wind` unless wind_field >= 0, in which case fields
 * wind_field .. wind_field+dim-1 are its components (Oseen wind).
 *
 * OCMP_BFI_DG_FACET (L2 block, trial == test), the skeleton terms of
 * opencmp/models/multi_component_ins.py:185-197 and :287-301:
 *  interior: eps*(alpha/h [c][r] - {grad c}.n [r] - {grad r}.n [c])      if sipg
 *            + [r] (c max(w.n,0) + c.Other() min(w.n,0))                  if upwind
 *  boundary, marker in dirichlet_mask: eps*(alpha/h c r - r grad c.n - c grad r.n)  if sipg
 *  every boundary facet: r c max(w.n,0)                                 if upwind
 * with alpha = ipc*p^2 (opencmp/models/base_model.py:150) and h the element size
 * |det J|^(1/dim) (facet: mean of both sides). */
typedef struct {
  int32_t kind;
  int32_t trial, test;
  int32_t dir;
  double coef;
  const double *elem_coef;
  double wind[3];
  int32_t wind_field;
  int32_t weight_mode;
  int32_t weight_field;
  int32_t sipg;
  int32_t upwind;
  int32_t reserved;
  double eps;
  double alpha;
  uint64_t dirichlet_mask;
} ocmp_term;

/* Replaces BilinearForm.Assemble() (opencmp/solvers/base_solver.py:368-377,
 * opencmp/models/ins.py:339-342). Writes (accumulate=0) or adds (1) the terms
 * into A's device CSR values. Deterministic: no floating-point atomics. */
int ocmp_assemble_bilinear(ocmp_matrix *A, int nterms, const ocmp_term *terms, int nfields,
                           const ocmp_field *fields, int accumulate);

enum { OCMP_LFI_SOURCE = 0, OCMP_LFI_DG_DIRICHLET = 1,
       OCMP_LFI_BND_SOURCE = 2 /* coef * g * v on boundary facets whose marker is in dirichlet_mask
                                  (H1, closed-form g): Neumann and Robin data,
                                  opencmp/models/poisson.py:151-158 */ };
enum {
  OCMP_FN_CONST = 0,     /* c[0]                                               */
  OCMP_FN_SINPROD = 1,   /* c[0] * prod_d sin(pi x_d)                           */
  OCMP_FN_POLY2 = 2,     /* 2D: c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2;
                            3D: c0 + c1 x + c2 y + c3 z + c4 x^2 + c5 y^2 + c6 z^2 + c7 xy + c8 xz + c9 yz */
  OCMP_FN_FIELD = 3,     /* the field `field` times c[0]                        */
  OCMP_FN_EXPR = 4       /* c[2] * P(x, y, z, t = c[1]) + c[3] for the program P registered as
                            c[0] by ocmp_expr_create (the config expressions of
                            opencmp/config_functions/parse_arithmetic.py:37-268)  */
};

/* Postfix programs of OCMP_FN_EXPR: one double per opcode, OCMP_OP_CONST followed by its value.
 * Unary ops replace the top of the stack, binary ops pop b then a and push a op b; HS2 pops
 * (t, w), RAMP4 pops (t, p1, p2, tr). sig, tanh, H and ramp are opencmp/helpers/math.py:23-131
 * (tanh clamps its argument at 350, H_s shifts by 9.21025, ramp is ramp_cos); sgn uses the
 * 1e-12 dead band of parse_arithmetic.py; round is round-half-even, as Python's. */
enum {
  OCMP_OP_END = 0, OCMP_OP_CONST = 1, OCMP_OP_X = 2, OCMP_OP_Y = 3, OCMP_OP_Z = 4, OCMP_OP_T = 5,
  OCMP_OP_NEG = 10, OCMP_OP_ADD = 11, OCMP_OP_SUB = 12, OCMP_OP_MUL = 13, OCMP_OP_DIV = 14, OCMP_OP_POW = 15,
  OCMP_OP_SIN = 20, OCMP_OP_COS = 21, OCMP_OP_TAN = 22, OCMP_OP_EXP = 23, OCMP_OP_SQRT = 24,
  OCMP_OP_TANH = 25, OCMP_OP_SIG = 26, OCMP_OP_ABS = 27, OCMP_OP_SGN = 28, OCMP_OP_TRUNC = 29,
  OCMP_OP_ROUND = 30,
  OCMP_OP_HS = 40, OCMP_OP_HS2 = 41,    /* H(t), H(t, w)                         */
  OCMP_OP_RAMP = 50, OCMP_OP_RAMP4 = 53 /* ramp(t), ramp(t, p1, p2, tr)          */
};

/* Registers a postfix program (n doubles ending in OCMP_OP_END, at most 8 stack slots deep) on the
 * context's device and returns its id for c[0] of OCMP_FN_EXPR. Replaces the evaluation of
 * parsed config expressions into NGSolve CoefficientFunctions
 * (opencmp/config_functions/parse_arithmetic.py:122-268) for the GPU path: the kernels evaluate the
 * program at their own quadrature and interpolation points. Programs stay registered for the
 * life of the process (64K doubles per device in all). */
int ocmp_expr_create(ocmp_ctx *ctx, int32_t n, const double *code, int32_t *id);

/* One linear-form term: coef * w * f * v (SOURCE), or the DG Dirichlet data
 * terms eps*g*(alpha/h v - grad v.n) - v g min(w.n,0) on boundary facets in
 * dirichlet_mask (opencmp/models/multi_component_ins.py:371-386). */
typedef struct {
  int32_t kind;
  int32_t test;
  int32_t fn;
  int32_t field;
  double c[10];
  double coef;
  int32_t weight_mode;
  int32_t weight_field;
  double eps;
  double alpha;
  double wind[3];
  uint64_t dirichlet_mask;
} ocmp_lterm;

/* Replaces LinearForm.Assemble(). b: device f64[ndof]. */
int ocmp_assemble_linear(ocmp_space *s, double *b, int nterms, const ocmp_lterm *terms, int nfields,
                         const ocmp_field *fields, int accumulate);

/* Replaces gfu.Set(g, definedon=mesh.Boundaries(...)) of
 * Model.apply_dirichlet_bcs_to (opencmp/models/base_model.py:321-341) for one
 * scalar H1 block: vertex dofs on the given markers take g's vertex values,
 * edge dofs the 1D L2 projection of g minus its vertex interpolant, 3D face
 * dofs (order 3) the L2 projection over the face of g minus its vertex and edge
 * parts onto the face bubble. g: closed form or expression (OCMP_FN_EXPR).
 * x: device f64[ndof]. */
int ocmp_dirichlet_set(ocmp_space *s, int block, double *x, int fn, const double *c,
                       uint64_t marker_mask);

/* gfu.Set(g) on one scalar block over the whole domain (the initial conditions of
 * opencmp/models/base_model.py:216-221 and the interpolation of _l_inf, opencmp/helpers/error.py:
 * 80-88): on H1 blocks (order 1-3) the vertex values, then per edge and per triangle with a
 * face-bubble dof (3D faces, 2D cells) the L2 projection of what the lower entities leave; on
 * L2 blocks (triangles) the element-wise L2 projection. g: closed form or expression. Syncs. */
int ocmp_set(ocmp_space *s, int block, double *x, int fn, const double *c);

/* y = A x; x, y device f64[n]. Replaces a.mat * vec (opencmp/models/base_model.py:920). */
int ocmp_matvec(ocmp_matrix *A, const double *x, double *y);

/* y = a x + b y on the context stream; x, y device f64[n]. Composes time-scheme
 * right-hand sides such as dt*f + M u_n (opencmp/solvers/time_integration_schemes.py:663-685). */
int ocmp_axpby(ocmp_ctx *ctx, int64_t n, double a, const double *x, double b, double *y);

/* out[0] = sqrt(int (u_h - g)^2), out[1] = sqrt(int |grad(u_h - g)|^2) for scalar
 * block `block` of the device vector x. g is the grid function `ref` (device,
 * same space) when ref != NULL, else the closed form (fn, c[10]). Replaces the
 * L2 norms of opencmp/helpers/error.py:54-128 (norm, mean, _l_2) used by
 * calc_error (:238-301) and by the Picard test of opencmp/models/ins.py:346-355.
 * The H1-seminorm has no counterpart in the reference. Synchronises. */
int ocmp_error_norms(ocmp_space *s, int block, const double *x, const double *ref, int fn, const double *c,
                     double *out);

/* Norms of one vector component (nblocks = 1..3 scalar blocks of one kind and order) against
 * closed forms or expressions (fns[k], c[10 k .. 10 k + 9] per block) or, with ref != NULL,
 * against the vector ref on the same space: out[0] = sqrt(int |u_h - g|^2) (_l_2), out[1] =
 * int |u_h - g| with the pointwise Euclidean norm (_l_1 of opencmp/helpers/error.py:70-72,
 * listed by calc_error :252), out[2] = sqrt(int |grad(u_h - g)|^2) (no reference counterpart;
 * NaN against an expression, whose derivative is not available). Synchronises. */
int ocmp_error_norms_vec(ocmp_space *s, int nblocks, const int32_t *blocks, const double *x, const double *ref,
                         const int32_t *fns, const double *c, double *out);

/* The Picard stop test of opencmp/models/ins.py:346-355 in one pass: out[0] =
 * sqrt(int |a - b|^2) (norm("l2_norm", W, u), error.py:75-77) and out[1] = sqrt(int |b|^2)
 * (mean(u), error.py:54-67) over one vector component (nblocks = 1..3 scalar blocks of one
 * kind and order) of two vectors on the space, integrated exactly (rule of degree 2p).
 * Synchronises. */
int ocmp_l2_pair_norms(ocmp_space *s, int nblocks, const int32_t *blocks, const double *a, const double *b,
                       double *out);

/* out[0] = sqrt(int (div u_h)^2) for the vector field in the dim consecutive scalar
 * blocks block0 .. block0+dim-1 of x (one H1 vector component): _divergence of
 * opencmp/helpers/error.py:151-167 (mass conservation check). Synchronises. */
int ocmp_divergence_norm(ocmp_space *s, int block0, const double *x, double *out);

/* out[0] = max_i |x_i - ref_i| over the dofs of scalar block `block` (ref NULL: max |x_i|):
 * _l_inf of opencmp/helpers/error.py:80-88, which interpolates the difference into the
 * space and takes the largest coefficient; for a reference grid function on the same
 * space that is exactly this. Synchronises. */
int ocmp_linf_norm(ocmp_space *s, int block, const double *x, const double *ref, double *out);

/* out[0] = ||b - A x||_2 and out[1] = ||b||_2 over the free dofs, FP64 (host
 * doubles, syncs). The residual a solve starts from: NGSolve's solvers compute
 * it inside (opencmp/models/base_model.py:886-947) without exposing it; the
 * learned-guess fallback compares it between candidate initial guesses. */
int ocmp_residual_norms(ocmp_matrix *A, const double *b, const double *x, double *out);

/* Allocation-free variant of ocmp_residual_norms. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc). out: DEVICE f64[2],
 * out[0] = ||b - A x||_2, out[1] = ||b||_2 over the free dofs, written on the
 * context stream; the call neither allocates nor synchronises (the caller
 * synchronises before reading out). Partial sums live in a static 4 MiB device
 * array per GPU (one copy per device, held until the process ends), so matrices
 * with more than about 1.07e9 nonzeros are rejected with OCMP_ERR_ARG. Calls that
 * overlap on one GPU share that array and race. */
int ocmp_residual_norms_noalloc(ocmp_matrix *A, const double *b, const double *x, double *out);

/* r = b - A x on the free dofs, 0 on the Dirichlet dofs (device f64), *norm = ||r||_2 (host, syncs). */
int ocmp_residual(ocmp_matrix *A, const double *b, const double *x, double *r, double *norm);

/* ----------------------------------------------------- preconditioner, solve */

enum { OCMP_PRE_NONE = 0, OCMP_PRE_JACOBI = 1, OCMP_PRE_BLOCK_JACOBI = 2 };
enum { OCMP_FP64 = 0, OCMP_MIXED = 1 };

/* Replaces ngs.Preconditioner(a, 'local') (opencmp/models/base_model.py:365-383):
 * point Jacobi on free dofs (H1), or element-block Jacobi (L2). With
 * precision OCMP_MIXED the library also keeps the FP32 copies used by the
 * inner solve of iterative refinement. */
int ocmp_precond_create(ocmp_matrix *A, int kind, int precision, ocmp_precond **out);
/* Diagonal preconditioner for the saddle-point systems of the Stokes/INS models
 * [[A, B^T], [B, -C]] (velocity blocks A, pressure blocks C = 1e-10 p q,
 * opencmp/models/ins.py:178-270): 1/a_ii on the velocity rows and
 * -1/(C_ii + sum_j B_ij^2/A_jj) on the rows of the scalar blocks set in
 * schur_blocks (bit k = block k), the diagonal of the Schur complement
 * -(C + B diag(A)^{-1} B^T). The reference solves these systems with a direct
 * factorisation (its 'direct' default); this replaces it on the Krylov path. */
int ocmp_precond_create_schur(ocmp_matrix *A, int precision, uint64_t schur_blocks, ocmp_precond **out);
/* Schur inner solve for a preconditioner from ocmp_precond_create_schur (same
 * replacement of the reference's direct factorisation of the saddle-point system,
 * opencmp/models/ins.py:178-270 with opencmp/solvers/base_solver.py:711-719):
 * max_iters > 0 makes the flexible step of OCMP_SOLVER_FGMRES solve the Schur
 * rows with Jacobi-PCG on S = C + B diag(A)^{-1} B^T (the sign of the Schur
 * diagonal, i.e. -S^{-1} on those rows) instead of scaling them by the Schur
 * diagonal: at most max_iters iterations, fewer once the preconditioned residual
 * (r, diag(S)^{-1} r) has fallen by tol^2 (tol 0: always max_iters); 0 turns it
 * off. The Schur blocks must be adjacent in the dof numbering. S's pattern is
 * built at the next ocmp_precond_update (ocmp_solve runs one first if none has
 * run since S was turned on) and its values at every update. The step is
 * nonlinear in its input, so the other solvers refuse such a preconditioner
 * (OCMP_ERR_ARG). */
int ocmp_precond_set_schur_solve(ocmp_precond *P, int max_iters, double tol);
/* Block-triangular step for a preconditioner from ocmp_precond_create_schur
 * (another replacement of the reference's direct factorisation of the saddle-point
 * system, opencmp/models/stokes_dim.py:28-153 with
 * opencmp/models/base_model.py:365-383 and :886-947), for systems whose velocity
 * block is not diagonally dominated (steady Stokes, diffuse-interface penalties).
 * The Schur blocks must be the last dofs [s0, n). The flexible step of
 * OCMP_SOLVER_FGMRES becomes z_p = -v_p / d_p on the Schur rows (d: schur_diag,
 * a device array of n - s0 positive entries approximating the Schur complement,
 * e.g. the lumped pressure mass over the viscosity, copied at this call; rows
 * where d is not above 1e-10 max d, or not free, get 0), then
 * z_u = A_uu^{-1} (v_u - A_up z_p) by at most vel_iters Jacobi-PCG iterations on
 * the velocity block A_uu, fewer once (r, diag(A_uu)^{-1} r) has fallen by
 * vel_tol^2. The velocity blocks' pattern is built at the next
 * ocmp_precond_update, their values at every update. Replaces the Schur inner
 * solve (turn that off first); vel_iters = 0 turns it off. */
int ocmp_precond_set_block_triangular(ocmp_precond *P, const double *schur_diag, int vel_iters, double vel_tol);
/* BF16 storage of the inner solves' matrices (the Schur matrix S of the Schur
 * inner solve, the velocity block of the block-triangular step) for a
 * preconditioner built with OCMP_MIXED: the inner Jacobi-PCG of the FP32 FGMRES
 * cycles reads BF16 values (rounded to nearest even from FP64, widened exactly
 * to FP32; FP32 vectors, FP64 sums). This realises the BF16-stored
 * preconditioner copy inside iterative refinement: the outer FP64 residual and
 * the FGMRES operator keep FP32/FP64 values, and FGMRES admits a preconditioner
 * that differs from the exact inner solve. No reference counterpart (NGSolve
 * runs in FP64). on = 0 turns it off; the copies are made at the next update. */
int ocmp_precond_set_bf16(ocmp_precond *P, int on);
/* Schur AMG step for a preconditioner from ocmp_precond_create_schur (the same
 * replacement of the reference's direct factorisation as ocmp_precond_set_schur_solve,
 * opencmp/models/ins.py:178-270 with opencmp/solvers/base_solver.py:711-719): the
 * flexible step of OCMP_SOLVER_FGMRES applies vcycles V(1,1) cycles of smoothed-
 * aggregation algebraic multigrid on S = C + B D^{-1} B^T to the Schur rows (the sign
 * of the Schur diagonal, i.e. -S^{-1}) instead of the inner PCG; the velocity rows
 * keep 1/a_ii. The hierarchy (aggregates, prolongations, Galerkin patterns) is built
 * on the host at the next ocmp_precond_update from S's values then, and its values
 * are recomputed on the device at every update. The step is a fixed linear operator.
 * Replaces the Schur inner solve (turn that off first); vcycles = 0 turns it off. */
int ocmp_precond_set_schur_amg(ocmp_precond *P, int vcycles);
/* Upper block-triangular velocity rows for a flexible Schur step (ocmp_precond_set_schur_solve
 * or _set_schur_amg; the same replacement of the reference's factorisation,
 * opencmp/models/ins.py:178-270 with base_solver.py:711-719): after the Schur rows
 * z_p, the velocity rows become z_u = diag(A)^{-1} (v_u - B^T z_p) instead of
 * diag(A)^{-1} v_u, with B^T the velocity rows' Schur columns of A (the Schur blocks
 * last). Costs one SpMV with B^T per step. on = 0 turns it off; the next update gathers
 * B^T's values. */
int ocmp_precond_set_schur_triangular(ocmp_precond *P, int on);
/* The Schur AMG hierarchy built by the last update (no counterpart in the reference,
 * whose preconditioners are NGSolve objects: this reports what ocmp_precond_set_schur_amg
 * built). *levels gets the number of levels (0: none built); rows[l] and nnz[l] the size
 * of level l's operator for l < max_levels (0 past the last level). */
int ocmp_precond_schur_amg_levels(const ocmp_precond *P, int64_t *rows, int64_t *nnz, int max_levels,
                                  int *levels);
/* Replaces Preconditioner.Update() (opencmp/solvers/base_solver.py:711-719):
 * recompute from A's current values. */
int ocmp_precond_update(ocmp_precond *P);
/* A partial Update() for the Picard iterations after the first (opencmp/models/ins.py:
 * 338-344 updates the preconditioner at each one): the point-Jacobi rows and the FP32
 * copy of A from A's current values, while the flexible Schur step (ocmp_precond_set_
 * schur_solve or _set_schur_amg) keeps the Schur matrix and hierarchy of the last full
 * ocmp_precond_update. The solution still meets each solve's tolerance; only the
 * preconditioner lags. Needs a full update first. */
int ocmp_precond_update_jacobi(ocmp_precond *P);
int ocmp_precond_destroy(ocmp_precond *P);

/* OCMP_SOLVER_MINRES replaces ngs.solvers.MinRes (opencmp/models/base_model.py:929-932):
 * preconditioned MINRES for symmetric A, FP64 only, with |D|^{-1} of a point-Jacobi or
 * Schur-diagonal preconditioner (symmetric positive definite, as MINRES needs) or none. */
enum { OCMP_SOLVER_CG = 0, OCMP_SOLVER_GMRES = 1, OCMP_SOLVER_RICHARDSON = 2, OCMP_SOLVER_FGMRES = 3,
       OCMP_SOLVER_MINRES = 4 };

typedef struct {
  int32_t solver;      /* OCMP_SOLVER_*                                          */
  int32_t precision;   /* OCMP_FP64 or OCMP_MIXED (FP32 inner, FP64 refinement)  */
  int32_t restart;     /* GMRES restart length m                                 */
  int32_t maxit;       /* total inner iterations                                 */
  double tol;          /* ||b - A x||_2 / ||b||_2 on free dofs, computed in FP64
                          (relative to the initial residual when b is 0 on the free dofs) */
  double inner_tol;    /* mixed: relative reduction asked of each inner solve
                          (0: 1e-5)                                             */
  int32_t check_every; /* CG: residual polled by the host every k iterations     */
  int32_t max_outer;   /* mixed: refinement steps                                */
  int32_t s_step;      /* GMRES on the DG block-Jacobi operator: 5 = s-step cycles
                          (blocks of 5 Arnoldi steps, one basis pass per block;
                          restart a multiple of 5, at most 50), else standard   */
  int32_t reserved;
} ocmp_solve_opts;

typedef struct {
  int32_t iters;        /* Krylov iterations (all inner solves)           */
  int32_t outer;        /* refinement steps (1 for FP64)                  */
  int32_t converged;
  int32_t fallback;     /* FGMRES with a learned correction: 1 if it fell back to the plain one;
                           s-step GMRES: 1 if a block lost definiteness and standard cycles took over */
  double rel_residual;  /* final ||b-Ax||/||b||, FP64, free dofs          */
  double initial_rel_residual;
  int64_t precond_iters; /* FGMRES with a Schur inner solve: PCG iterations done in all its steps */
} ocmp_solve_stats;

/* Replaces Model.linear_solve (opencmp/models/base_model.py:886-947): solves
 * A x = b on the free dofs starting from x (which holds the Dirichlet lift),
 * leaving Dirichlet entries of x unchanged. b, x: device f64[n]. Returns OCMP_OK
 * also when not converged; check stats->converged. */
int ocmp_solve(ocmp_matrix *A, ocmp_precond *P, const double *b, double *x,
               const ocmp_solve_opts *opts, ocmp_solve_stats *stats);

/* s-step GMRES basis shifts. No counterpart in the reference (NGSolve's GMRes
 * has no s-step variant); ocmp_solve uses them internally, they are exported
 * so that callers and tests can inspect them. Host-only, no device calls.
 *
 * Eigenvalues (wr + i wi) of the n x n upper Hessenberg matrix H, column-major
 * with leading dimension ldh (the Ritz values of a GMRES cycle): Francis
 * double-shift QR. Returns OCMP_ERR_NUMERIC if it does not converge. */
int ocmp_hessenberg_eigvals(int32_t n, const double *H, int32_t ldh, double *wr, double *wi);
/* s Newton-basis shifts from n Ritz values in modified Leja order, complex
 * conjugates adjacent: basis step t is v_{t+1} = (A - d[t] I) v_t + c[t] v_{t-1},
 * with c[t] = (Im z)^2 on the second member of a pair and 0 otherwise. */
int ocmp_newton_shifts(int32_t n, const double *wr, const double *wi, int32_t s, double *d, double *c);

/* ----------------------------------------------------- learned initial guess */

/* Per-DOF MLP predicting the next time step of a transient solve from the last
 * four states of each DOF. The reference has no counterpart: it starts every
 * solve of a time step from the previous solution (the GridFunction carries it,
 * opencmp/models/ins.py:323-359, opencmp/solvers/time_integration_schemes.py:26-80).
 *
 * Network: 4 -> 128 -> 128 -> 128 -> 1, ReLU, bf16 weights, fp32 accumulation.
 * Parameters (float32, this order): W1[128][4], b1[128], W2[128][128], b2[128],
 * W3[128][128], b3[128], W4[128], b4[1].
 * Features of DOF i in segment k (a segment = the DOFs of one scalar block):
 *   x = ((u0-u1)/s_k, (u1-u2)/s_k, (u2-u3)/s_k, u0/U_k),  guess = u0 + base (u0-u1) + s_k * f(x)
 * with u0 the newest state. Training minimises 0.5/B sum (f(x) - t)^2 with
 * t = (u_next - u0 - base (u0 - u1))/s_k by Adam (beta 0.9/0.999, eps 1e-8); base 0 unless set.
 * X: device float32[n][4] (16-byte aligned), T/Y: device float32[n]. */
#define OCMP_MLP_HIDDEN 128
#define OCMP_MLP_NPARAM 33793
#define OCMP_MLP_MAXSEG 8
typedef struct ocmp_mlp ocmp_mlp;

/* params: host float32[OCMP_MLP_NPARAM]; max_batch bounds the training batch. */
int ocmp_mlp_create(ocmp_ctx *ctx, const float *params, int64_t max_batch, ocmp_mlp **out);
int ocmp_mlp_destroy(ocmp_mlp *mlp);
/* host copies; set also resets the Adam moments and step count */
int ocmp_mlp_get_params(ocmp_mlp *mlp, float *params);
int ocmp_mlp_set_params(ocmp_mlp *mlp, const float *params);
/* feature scales per segment: seg_off[nseg+1] (host, seg_off[0] = 0), s/U[nseg] > 0 */
int ocmp_mlp_set_scales(ocmp_mlp *mlp, int nseg, const int64_t *seg_off, const double *s, const double *U);
int ocmp_mlp_get_scales(ocmp_mlp *mlp, int *nseg, int64_t *seg_off, double *s, double *U);
/* s_k = rms of the increments, U_k = rms of the states over snapshots
 * snap[j*ld + i] (device f64, j < nsnap); a segment that never moves gets 1 */
int ocmp_mlp_fit_scales(ocmp_mlp *mlp, const double *snap, int nsnap, int64_t ld, int nseg,
                        const int64_t *seg_off);
/* batch random samples (snapshot j in [3, nsnap-2], dof i < n) of (X, T), seeded; T follows the base */
int ocmp_mlp_sample(ocmp_mlp *mlp, const double *snap, int nsnap, int64_t ld, int64_t n, uint64_t seed,
                    int64_t batch, float *X, float *T);
/* Y = f(X) */
int ocmp_mlp_forward(ocmp_mlp *mlp, int64_t n, const float *X, float *Y);
/* gradient of the loss on (X, T) into grad (device float32[OCMP_MLP_NPARAM]); loss: host, may be NULL */
int ocmp_mlp_grad(ocmp_mlp *mlp, int64_t n, const float *X, const float *T, float *grad, double *loss);
/* one Adam step on (X, T); loss (host, may be NULL: then nothing synchronises) is the pre-step loss */
int ocmp_mlp_train_step(ocmp_mlp *mlp, int64_t n, const float *X, const float *T, float lr, double *loss);
/* What the network predicts the increment from: base 0 (default) the newest state u0, base 1 the
 * linear extrapolation 2 u0 - u1. With base 1 training targets are (u_next - 2 u0 + u1)/s_k and the
 * guess is 2 u0 - u1 + s_k f(x): an untrained (near-zero) network already extrapolates. */
int ocmp_mlp_set_base(ocmp_mlp *mlp, int base);
int ocmp_mlp_get_base(ocmp_mlp *mlp, int *base);
/* out = u0 + base (u0 - u1) + s * f(features(u0, u1, u2, u3)) over n dofs (device f64); out may
 * alias none of u0..u3 */
int ocmp_mlp_guess(ocmp_mlp *mlp, int64_t n, const double *u0, const double *u1, const double *u2,
                   const double *u3, double *out);

/* ----------------------------------------------- learned preconditioner correction */

/* 3-level U-net on the vertex grid of an N^3 x 6-tet structured mesh (vertex id =
 * grid index, x fastest), 4 channels: the vertex dofs of four scalar blocks (the
 * velocity components and the pressure of the diffuse-interface Stokes model,
 * opencmp/models/stokes_dim.py:32-153). Used by OCMP_SOLVER_FGMRES as the
 * preconditioner M v = D^{-1} v + E U(R v) when attached to a point-Jacobi /
 * Schur-diagonal preconditioner; FGMRES falls back to D^{-1} alone when the
 * learned cycles reduce the residual more slowly. The reference has no
 * counterpart (its preconditioners: opencmp/models/base_model.py:365-383).
 * Layers (3x3x3, padding 1): conv 4->8, conv/2 8->16, conv 16->16, conv/2 16->32,
 * conv 32->32, tconv*2 32->16 (+skip), tconv*2 16->8 (+skip), conv 8->4; ReLU
 * except the last; bf16 weights and activations, fp32 sums. Parameters float32,
 * per layer W then b, W [out][in][27] (conv) or [in][out][27] (tconv). */
#define OCMP_UNET_NPARAM 70980
typedef struct ocmp_unet ocmp_unet;

int ocmp_unet_create(ocmp_ctx *ctx, int N, const float *params, ocmp_unet **out);
int ocmp_unet_destroy(ocmp_unet *u);
int ocmp_unet_get_params(ocmp_unet *u, float *params);
int ocmp_unet_set_params(ocmp_unet *u, const float *params);
/* channel c reads/writes dofs dof_off[c] + vertex; input x_c = in_scale[c] * v, output z += out_scale[c] * y_c */
int ocmp_unet_set_channels(ocmp_unet *u, const int64_t *dof_off, const double *in_scale, const double *out_scale);
/* channel offsets and scales fitted to samples (device f64 rows, stride ld, e.g. residuals of
 * plain solves): in_scale = 1/rms(v_c), out_scale = rms((D^{-1} v)_c); P: A's Jacobi preconditioner */
int ocmp_unet_fit_channels(ocmp_unet *u, ocmp_precond *P, const double *samples, int nsamples, int64_t ld,
                           const int64_t *dof_off);
int ocmp_unet_get_channels(ocmp_unet *u, int64_t *dof_off, double *in_scale, double *out_scale);
/* grid-level forward / parameter gradient of sum(gy * y) (device float32 [4][(N+1)^3]; grad [NPARAM]) */
int ocmp_unet_forward(ocmp_unet *u, const float *x, float *y);
int ocmp_unet_grad(ocmp_unet *u, const float *x, const float *gy, float *grad);
/* the activations of the last forward pass as float32 (device out, may be NULL to query *count):
 * layer -1 the bf16-rounded input, 0..6 a0..a6 ([C][n_level^3]), 7 the output y */
int ocmp_unet_activation(ocmp_unet *u, int layer, float *out, int64_t *count);
/* one Adam step on 0.5/B sum_b ||r_b - A E U(R v_b)||^2, r_b = v_b - A D^{-1} v_b, v_b = samples row idx[b]
 * (device f64, row stride ld); P the (updated) Jacobi preconditioner of A; loss: host */
int ocmp_unet_train_step(ocmp_unet *u, ocmp_matrix *A, ocmp_precond *P, const double *samples, int nsamples,
                         int64_t ld, const int32_t *idx, int nbatch, float lr, double *loss);
/* attach (or, with u = NULL, detach) the learned correction used by OCMP_SOLVER_FGMRES */
int ocmp_precond_attach_unet(ocmp_precond *P, ocmp_unet *u);

#ifdef __cplusplus
}
#endif
#endif /* OCMP_H */
