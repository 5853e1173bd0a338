/*
 * msfem.h -- C ABI of the MI355X MsFEM basis engine (libmsfem_hip.so).
 *
 * The engine replaces the per-coarse-cell local loop of
 * konsim83/MPI-Parallel-Multiscale-Diffusion-FEM:
 *
 *   DiffusionProblemBasis<dim>::run()              include/base/diffusion_problem_basis.tpp:440-474
 *     make_grid / setup_system / assemble_system   diffusion_problem_basis.tpp:92-242
 *     2^dim constrained basis solves (CG+SSOR)     diffusion_problem_basis.tpp:295-317, :450-465
 *     assemble_global_element_matrix               diffusion_problem_basis.tpp:247-285
 *   get_global_element_matrix / _rhs               diffusion_problem_basis.tpp:320-334
 *   called per locally owned cell from
 *   DiffusionProblemMultiscale::initialize_and_compute_basis  include/base/diffusion_problem_ms.tpp:42-88
 *
 * and the coarse system around it:
 *   DiffusionProblemMultiscale::assemble_system    diffusion_problem_ms.tpp:161-255
 *   DiffusionProblemMultiscale::solve_iterative    diffusion_problem_ms.tpp:260-296
 *   send_global_weights_to_cell / set_global_weights  diffusion_problem_ms.tpp:301-325,
 *                                                  diffusion_problem_basis.tpp:354-377
 *
 * One call builds a whole batch of coarse cells on one GPU.  All functions
 * return an int status (MSFEM_OK on success), never throw, and keep the
 * message of the last failure in msfem_last_error(ctx).  Caller buffers are
 * borrowed for the duration of the call.  One host thread per context.
 *
 * Data conventions (shared by every entry point):
 *   - dim is 2 or 3; nb = 2^dim bases per coarse cell; basis i belongs to
 *     coarse vertex i, lexicographic with x fastest (deal.II vertex order).
 *   - a coarse cell is an axis-aligned box: box[c] = {lo_x, hi_x, lo_y, hi_y
 *     [, lo_z, hi_z]}, or (msfem_batch.vertices) a general cell given by its
 *     2^dim vertices in deal.II order: a convex quadrilateral (2D) or a
 *     hexahedron (3D); its fine grid is the bi/trilinear image of the uniform
 *     grid (what GridGenerator::general_cell + refine_global produce,
 *     include/base/diffusion_problem_basis.tpp:92-99), every fine cell is
 *     integrated with the bi/trilinear mapping, and the Dirichlet data is
 *     BasisQ1's N_i(x) = sum_m c_mi mono_m(x), mono = (1, x, y, xy) in 2D and
 *     (1, x, y, z, xy, yz, xz, xyz) in 3D, with N_i(v_j) = delta_ij
 *     (include/coefficients/basis_q1.tpp:26-46, :86-96) at the mapped nodes.
 *   - a cell with nf fine cells per side has (nf+1)^dim fine nodes,
 *     lexicographic, x fastest.  Its bases are stored as [nb][(nf+1)^dim],
 *     cells concatenated in batch order.
 *   - coefficient tiles are concatenated in batch order, one per cell, AoS per
 *     fine cell (fine cells lexicographic):
 *       MSFEM_COEF_SCALAR_CELL  [nf^d]            one value per fine cell
 *       MSFEM_COEF_TENSOR_CELL  [nf^d][ns]        symmetric tensor per fine cell
 *       MSFEM_COEF_SCALAR_QP    [nf^d][2^d]       scalar at the QGauss(2) points
 *       MSFEM_COEF_TENSOR_QP    [nf^d][2^d][ns]   tensor at the QGauss(2) points
 *     with ns = 3 (a11,a12,a22) in 2D and 6 (a11,a12,a13,a22,a23,a33) in 3D,
 *     quadrature points tensor-ordered, x fastest.  The QP kinds reproduce the
 *     reference's MatrixCoeff::value_list at the quadrature points
 *     (include/coefficients/matrix_coeff.tpp:66-91).
 */
#ifndef MSFEM_H_
#define MSFEM_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSFEM_API_VERSION 3

/* status codes */
#define MSFEM_OK 0
#define MSFEM_ERR_ARG 1         /* invalid argument or unsupported shape */
#define MSFEM_ERR_HIP 2         /* a HIP runtime call failed */
#define MSFEM_ERR_NOCONV 3      /* some local or coarse solve did not converge */
#define MSFEM_ERR_NODEVICE 4    /* no usable GPU */

/* coefficient kinds */
#define MSFEM_COEF_SCALAR_CELL 0
#define MSFEM_COEF_TENSOR_CELL 1
#define MSFEM_COEF_SCALAR_QP 2
#define MSFEM_COEF_TENSOR_QP 3

/* right-hand side kinds (load vector b, hence F^K) */
#define MSFEM_RHS_CONST 0       /* f = rhs_value (reference default: 2.0,
                                   include/coefficients/right_hand_side.tpp:17-40) */
#define MSFEM_RHS_SINSIN 1      /* f = prod_k sin(pi x_k), evaluated at QGauss(2) points */

/* arithmetic of the local solves */
#define MSFEM_PREC_FP64 0         /* plain FP64 CG */
#define MSFEM_PREC_FP32_REFINE 1  /* FP32 CG inside FP64 residual refinement */

/* coarse boundary conditions */
#define MSFEM_BC_ZERO_DIRICHLET 0 /* u = 0 on every face */
#define MSFEM_BC_REFERENCE 1      /* reference default: Dirichlet g=(x-1/2)^2+(y-1/2)^2 on
                                     x=0,y=0(,z=0), Neumann h=cos(2 PI_D x)cos(2 PI_D y) on
                                     x=1,y=1(,z=1); diffusion_problem_ms.tpp:128-137,:206-235 */

typedef struct msfem_ctx msfem_ctx;

/* Local-solve options.  Stopping rule of every local solve: the true FP64
 * residual of the interior equations satisfies
 *     ||r||_2 <= max(abs_tol, rel_tol * ||b||_2),   b = -K_IB g_i.
 * The reference stops at ||r||_2 <= 1e-12 (absolute), 1000 iterations
 * (include/base/diffusion_problem_basis.tpp:297). */
typedef struct msfem_solver_opts {
  int32_t precision;          /* MSFEM_PREC_* */
  int32_t max_iter;           /* cap on inner CG iterations per basis, summed over cycles */
  int32_t max_refine;         /* cap on FP64 residual-refinement cycles */
  int32_t partition_of_unity; /* 1: solve nb-1 bases, phi_{nb-1} = 1 - sum of the others */
  double rel_tol;
  double abs_tol;
  double inner_rel_tol;       /* FP32 inner solve: residual reduction per cycle */
  int32_t variant;            /* launch-shape table entry: 0 = built-in default,
                                 >0 = explicit variant id (see msfem_variants) */
  int32_t graph;              /* 1: run the build as a HIP graph, one concurrent branch
                                 per size class, instantiated on first use and replayed
                                 by later calls with the same batch layout, pointers and
                                 options (msfem_ctx_profile is not updated; time it with
                                 msfem_ctx_mark) */
  int32_t coarse_space;       /* 1 (default): two-level preconditioner (Jacobi plus the
                                 bilinear coarse space of an 8 x 8 coarse grid per local
                                 problem, solved exactly) for the 64^2 and 128^2 classes at
                                 the default launch shapes; on 64^2 classes also the
                                 intermediate hat levels of spacing 2 and 4, on FP32 128^2
                                 classes those of spacing 4 and 8, scaled by their diagonals
                                 (additive);
                                 2: the two-level form without the
                                 intermediate levels; 0: Jacobi only.  Same converged
                                 values, fewer iterations. */
  int32_t q1_start;           /* without x0: 1 (default) = the interior starts at the Q1
                                 interpolant of the Dirichlet data (box cells: the nodal
                                 Q1 function itself; general cells: the transfinite
                                 interpolant of the boundary values over the parametric
                                 node grid), 0 = zero interior.  The stopping test stays
                                 relative to ||b|| = ||K_IB g_i|| either way; the start
                                 only shortens the solve. */
  double energy_tol;          /* > 0 (box cells): replaces the residual test above by one
                                 that guarantees ||phi - phi*||_K <= energy_tol ||phi*||_K
                                 for every basis (phi* the exact discrete solution, K the
                                 fine stiffness): stop when ||r||_2 <= energy_tol
                                 sqrt(lam_lb e_lb), with lam_lb = (min over the tile of
                                 lambda_min(A)) x (smallest eigenvalue of the
                                 unit-coefficient stiffness) <= lambda_min(K_II) and
                                 e_lb = (int |grad N_i|^2)^2 / int grad N_i^T A^-1 grad N_i
                                 <= ||phi*||_K^2 (complementary energy with the
                                 divergence-free flux grad N_i).  0 (default): off. */
  int32_t reserved[2];
} msfem_solver_opts;

typedef struct msfem_batch {
  int32_t dim;
  int32_t n_cells;
  const int32_t *n_fine;      /* HOST array [n_cells]: fine cells per side (2D: 1..1024,
                                 3D: 1..64; the reference's 2^n_refine_local) */
  const double *box;          /* [n_cells][dim][2] */
  int32_t coef_kind;          /* MSFEM_COEF_* */
  const double *coef;         /* concatenated coefficient tiles */
  int32_t rhs_kind;           /* MSFEM_RHS_* */
  double rhs_value;
  const double *x0;           /* optional initial bases (bases layout) or NULL: interior
                                 values only are read; a non-finite value is taken as 0.
                                 A problem whose initial FP64 residual already meets the
                                 tolerance takes 0 iterations (accepted as given). */
  const double *vertices;     /* optional general cells (box is then ignored and may be
                                 NULL), deal.II vertex order (vertex v at the parametric
                                 corner (v&1, (v>>1)&1, (v>>2)&1)):
                                 2D [n_cells][4][2] quadrilaterals, each convex with
                                 positive orientation;
                                 3D [n_cells][8][3] hexahedra whose trilinear map's
                                 Jacobian determinant has positive Bernstein
                                 coefficients (a sufficient test for a positive
                                 determinant everywhere; msfem_batch_build checks both,
                                 the device entry points trust the caller).
                                 NULL: the cells are the boxes. */
} msfem_batch;

typedef struct msfem_batch_out {
  double *Ae;       /* [n_cells][nb][nb]  A^K_ij = phi_i^T K phi_j */
  double *Fe;       /* [n_cells][nb]      F^K_i  = phi_i^T b       */
  double *bases;    /* optional (NULL): [sum_c nb (nf_c+1)^dim] basis nodal values */
  int32_t *iters;   /* optional: [n_cells][nb] inner CG iterations */
  int32_t *status;  /* optional: [n_cells][nb] 0 converged, 1 not converged */
  double *resid;    /* optional: [n_cells][nb] final true residual norm */
  double *energy_bound; /* optional: [n_cells][nb] a posteriori bound of the relative
                           energy-norm error ||phi - phi*||_K / ||phi*||_K of each
                           basis: sqrt(q / (A^K_ii - q)), q = ||r||_2^2 / lam_lb
                           (||phi||_K^2 = ||phi*||_K^2 + ||phi - phi*||_K^2 since
                           K phi* = 0 on the interior); NaN for general
                           quadrilaterals.  No counterpart in the reference. */
} msfem_batch_out;

/* ---- context --------------------------------------------------------- */
/* Replaces: the per-rank state of DiffusionProblemMultiscale (diffusion_problem_ms.tpp:18-37). */
int msfem_ctx_create(int device, msfem_ctx **out);
int msfem_ctx_destroy(msfem_ctx *ctx);
const char *msfem_last_error(const msfem_ctx *ctx); /* ctx may be NULL (creation errors) */
void msfem_default_opts(msfem_solver_opts *opts);
/* 1 if a class of nf fine cells per side with these options, coefficient kind
 * and geometry (general: vertex cells) runs the two-level local solve on this
 * context (its launch-shape table applied), 2 if its preconditioner also takes
 * the intermediate hat levels (spacing 2 and 4 at 64^2, 4 and 8 at 128^2), else
 * 0.  No counterpart in the
 * reference; for callers that account for the solver's work. */
int msfem_uses_coarse_space(msfem_ctx *ctx, int dim, int nf, const msfem_solver_opts *opts, int coef_kind,
                            int general);
int msfem_version(void);

/* Kernel timing of the batch builds (HIP events on the build's stream around
 * the local-solve launches -- after each class's preparation: conductances,
 * general cells' boundary data, the two-level coarse setup -- and around the
 * element launches).  After msfem_ctx_set_profiling(ctx, 1),
 * msfem_ctx_profile returns the device milliseconds of the last completed
 * build (it waits for that build's events).  No counterpart in the reference,
 * whose only timer is the TimerOutput scope "basis initialization and
 * computation" (include/base/diffusion_problem_ms.tpp:44-45). */
int msfem_ctx_set_profiling(msfem_ctx *ctx, int on);
int msfem_ctx_profile(msfem_ctx *ctx, double *ms_local_solve, double *ms_element);

/* Timing marks (no counterpart in the reference): msfem_ctx_mark records HIP
 * event `slot` (0 <= slot < MSFEM_N_MARKS) on the context stream;
 * msfem_ctx_elapsed waits for mark `to` and returns the device milliseconds
 * between two recorded marks (e.g. around a surrogate inference). */
#define MSFEM_N_MARKS 8
int msfem_ctx_mark(msfem_ctx *ctx, int slot);
int msfem_ctx_elapsed(msfem_ctx *ctx, int from, int to, double *ms);
/* Two-stream ordering (no counterpart in the reference): msfem_ctx_aux_stream
 * returns the context's second stream (created at default priority on first
 * use; msfem_ctx_synchronize also waits for it); msfem_ctx_mark_on records
 * mark `slot` on `stream` (NULL = the context stream); msfem_ctx_wait_mark
 * makes `stream` (NULL = the context stream) wait for the last recording of
 * mark `slot`.  Lets a caller overlap work on the two streams, e.g. the
 * surrogate's inference of one chunk of cells with the builds of the previous
 * chunk. */
int msfem_ctx_aux_stream(msfem_ctx *ctx, void **stream);
int msfem_ctx_mark_on(msfem_ctx *ctx, int slot, void *stream);
int msfem_ctx_wait_mark(msfem_ctx *ctx, int slot, void *stream);

/* Device-memory plumbing for callers without an allocator of their own
 * (no counterpart in the reference, whose data never leaves the host).
 * Memory comes from hipMalloc on the context's device; copies are synchronous;
 * msfem_ctx_stream returns the context's own stream for the *_device entry points. */
int msfem_dev_alloc(msfem_ctx *ctx, uint64_t bytes, void **ptr);
int msfem_dev_free(msfem_ctx *ctx, void *ptr);
int msfem_memcpy_h2d(msfem_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int msfem_memcpy_d2h(msfem_ctx *ctx, void *dst, const void *src, uint64_t bytes);
void *msfem_ctx_stream(msfem_ctx *ctx);
int msfem_ctx_synchronize(msfem_ctx *ctx);

/* Launch-shape table of the local solves (filled by the autotuner,
 * mpi-parallel-multiscale-diffusion-fem_amd/autotune.py).  Every resolution
 * class has msfem_variant_count(dim, nf) compiled launch shapes with ids
 * 1..count; msfem_set_variant records the id used for (dim, nf, precision,
 * coefficient kind) in this context when opts->variant is 0 (id 0 = the
 * built-in default).  No counterpart in the reference. */
int msfem_variant_count(int dim, int nf);
int msfem_set_variant(msfem_ctx *ctx, int dim, int nf, int precision, int coef_kind, int variant);

/* ---- the hot path ------------------------------------------------------ */
/* Builds every cell of the batch: fine assembly, nb basis solves, A^K, F^K.
 * Replaces DiffusionProblemBasis<dim>::run() for all cells
 * (include/base/diffusion_problem_basis.tpp:440-474, driven from
 * diffusion_problem_ms.tpp:84-87) and the getters
 * get_global_element_matrix / get_global_element_rhs (:320-334).
 * Host pointers; the call is synchronous.  Returns MSFEM_ERR_NOCONV (after
 * writing every output) when a basis solve hit its iteration cap, where the
 * reference throws SolverControl::NoConvergence.  An empty batch (n_cells =
 * 0) is a no-op returning MSFEM_OK; its arrays and outputs may be NULL. */
int msfem_batch_build(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *opts,
                      msfem_batch_out *out);

/* Same, with every array except in->n_fine in device memory and the work
 * enqueued on `stream` (a hipStream_t, NULL = default stream) without a host
 * synchronisation.  out->bases may be NULL (the context keeps scratch).
 * Convergence is reported only through out->status. */
int msfem_batch_build_device(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *opts,
                             msfem_batch_out *out, void *stream);

/* Relative energy-norm error of every basis of a batch against reference bases
 * (e.g. a mixed-precision or surrogate-started build against the plain FP64
 * build): err[c][i] = ||phi_i - phi_ref,i||_K / ||phi_ref,i||_K with the cell's
 * fine stiffness K (the element kernel's phi^T K phi on the difference and on
 * the reference).  in gives the layout and the coefficients (its x0, rhs and
 * outputs are unused); every array except in->n_fine in device memory, work
 * enqueued on `stream`.  No counterpart in the reference: the measure of the
 * error the mixed-precision and surrogate forms state against the FP64 path. */
int msfem_energy_error_device(msfem_ctx *ctx, const msfem_batch *in, const double *bases, const double *bases_ref,
                              double *err, void *stream);

/* ---- coarse system ------------------------------------------------------ */
/* Unit square/cube split into nc^dim coarse cells, lexicographic, x fastest
 * (GridGenerator::hyper_cube + refine_global, diffusion_problem_ms.tpp:93-103).
 * Assembles the coarse matrix in structured stencil form Kc[(nc+1)^dim][3^dim]
 * (entry t couples node n with n + (t%3-1, t/3%3-1[, t/9-1])), the load Fc
 * (Neumann terms included) and the Dirichlet data (cmask, cval).
 * Replaces DiffusionProblemMultiscale::assemble_system (diffusion_problem_ms.tpp:161-255)
 * with setup_system's constraints (:108-156).  Host pointers. */
int msfem_coarse_assemble(msfem_ctx *ctx, int dim, int nc, const double *Ae, const double *Fe, int bc_kind,
                          double *Kc, double *Fc, uint8_t *cmask, double *cval);
int msfem_coarse_assemble_device(msfem_ctx *ctx, int dim, int nc, const double *Ae, const double *Fe,
                                 int bc_kind, double *Kc, double *Fc, uint8_t *cmask, double *cval,
                                 void *stream);

/* FP64 CG on the constrained coarse system (u = cval where cmask), stops at
 * ||r||_2 <= rel_tol*||b||_2 or max_iter.  Replaces
 * DiffusionProblemMultiscale::solve_iterative (diffusion_problem_ms.tpp:260-296),
 * which uses AMG-CG to an absolute 1e-12.  *iters receives the iteration count.
 * Host pointers; returns MSFEM_ERR_NOCONV if max_iter was reached. */
int msfem_coarse_solve(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc,
                       const uint8_t *cmask, const double *cval, double rel_tol, int max_iter, double *u,
                       int32_t *iters);
int msfem_coarse_solve_device(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc,
                              const uint8_t *cmask, const double *cval, double rel_tol, int max_iter,
                              double *u, int32_t *iters, void *stream);

/* The same solve with no device allocation: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The kernels of
 * msfem_coarse_solve_device run with their scratch in a fixed 16 MiB device
 * array of the library (one per GPU, held until the process ends, its address
 * looked up when a context is created on that GPU); the call enqueues one
 * kernel on `stream` and calls no hipMalloc, hipFree or other allocating
 * routine.  Systems whose scratch exceeds the array (for power-of-two sizes:
 * 2D beyond 256 x 256 cells, 3D beyond 32^3) are rejected with MSFEM_ERR_ARG.  Every call on one
 * GPU shares that GPU's array, so calls that overlap on one GPU (different
 * streams, no ordering between them) race; the kernels write every scratch
 * value before reading it, so nothing needs clearing between calls. */
int msfem_coarse_solve_device_noalloc(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc,
                                      const uint8_t *cmask, const double *cval, double rel_tol, int max_iter,
                                      double *u, int32_t *iters, void *stream);

/* Per-cell weights w[c][i] = u[global node of vertex i of cell c]
 * (send_global_weights_to_cell, diffusion_problem_ms.tpp:301-325).  Host pointers. */
int msfem_gather_weights(int dim, int nc, const double *u, double *w);

/* Fine reconstruction u|_K = sum_i w_i phi_i for every cell
 * (DiffusionProblemBasis::set_global_weights, diffusion_problem_basis.tpp:354-377).
 * bases in the bases layout, w [n_cells][nb], u in the layout of one basis
 * per cell ([sum_c (nf_c+1)^dim]).  Device pointers except n_fine. */
int msfem_reconstruct_device(msfem_ctx *ctx, int dim, int n_cells, const int32_t *n_fine,
                             const double *bases, const double *w, double *u, void *stream);

/* ---- fine-scale standard FEM ---------------------------------------------
 * Replaces: DiffusionProblem<2> (include/base/diffusion_problem.tpp:40-279),
 * the single-scale Q1 solve on the fine mesh that the reference's main runs
 * beside the MsFEM (source/main.cxx:33-35).  Q1 on the n x n fine mesh of the
 * unit square with the coefficient tile `coef` (host; one cell of nf = n in
 * the batch layout of coef_kind), load f (rhs_kind, rhs_value) and boundary
 * data bc_kind (MSFEM_BC_*); FP64 Jacobi-PCG over many workgroups, stopped at
 * ||r||_2 <= rel_tol ||r_0||_2 (r_0: residual of the Dirichlet lift) or
 * max_iter iterations (then MSFEM_ERR_NOCONV, u still written).  u: host
 * [(n+1)^2] nodal values, lexicographic; 2D only (dim = 2). */
int msfem_fine_solve(msfem_ctx *ctx, int dim, int n, int coef_kind, const double *coef, int rhs_kind,
                     double rhs_value, int bc_kind, double rel_tol, int max_iter, double *u, int32_t *iters);

/* ---- synthetic inputs (host, FP64, bit-identical for every consumer) ----- */
/* Counter-based Philox-4x32-10 (Salmon et al., SC'11): out = Philox(ctr, key). */
void msfem_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Boxes of the nc^dim coarse cells of the unit square/cube -> box[nc^dim][dim][2]. */
int msfem_gen_unit_boxes(int dim, int nc, double *box);
/* Reference default coefficient MatrixCoeff<dim> (PI_D = 3.14592653509793218403,
 * k = 57, scale 0.9999, rotation by PI_D/3; include/coefficients/matrix_coeff.hpp:45-49,
 * matrix_coeff.tpp:17-91) at the QGauss(2) points -> TENSOR_QP tiles. */
int msfem_gen_coef_reference(int dim, int n_cells, const int32_t *n_fine, const double *box, double *coef);
/* Unit square split into nc x nc general quadrilaterals: the interior vertices
 * of the uniform mesh are moved by amp/nc * (2u - 1) per axis, u uniform from
 * Philox(ctr={vertex, 0, 2, 0}, key=seed) (0 <= amp < 0.25 keeps every cell
 * convex) -> vertices[nc*nc][4][2], deal.II vertex order, cells x fastest.
 * Test and benchmark meshes for the general-cell path. */
int msfem_gen_perturbed_quads(int nc, double amp, uint64_t seed, double *vertices);
/* MatrixCoeff<2> at the mapped QGauss(2) points of the fine cells of general
 * quadrilaterals (vertices[n_cells][4][2]) -> TENSOR_QP tiles. */
int msfem_gen_coef_reference_quad(int n_cells, const int32_t *n_fine, const double *vertices, double *coef);
/* Unit cube split into nc^3 general hexahedra: the interior vertices of the
 * uniform mesh are moved by amp/nc * (2u_k - 1) along axis k, u_0..u_2 uniform
 * from Philox(ctr={vertex, 0, 3, 0}, key=seed) (0 <= amp < 1/6 keeps every
 * Jacobian matrix strictly diagonally dominant, so every cell is positively
 * oriented and passes the build's orientation test) -> vertices[nc^3][8][3],
 * deal.II vertex order, cells x fastest.  Test meshes for general hexahedra. */
int msfem_gen_perturbed_hexes(int nc, double amp, uint64_t seed, double *vertices);
/* MatrixCoeff<3> at the mapped QGauss(2) points of the fine cells of general
 * hexahedra (vertices[n_cells][8][3]) -> TENSOR_QP tiles. */
int msfem_gen_coef_reference_hex(int n_cells, const int32_t *n_fine, const double *vertices, double *coef);
/* alpha = 2 + sin(2 pi x/eps) cos(2 pi y/eps) at the QGauss(2) points -> SCALAR_QP tiles (2D). */
int msfem_gen_coef_periodic(int n_cells, const int32_t *n_fine, const double *box, double eps, double *coef);
/* alpha = exp(clamp(g, -clip, clip)), g ~ N(0,1) i.i.d. per fine cell -> SCALAR_CELL tiles.
 * The normal of fine cell j of cell c comes from Philox(ctr={j, c, 0, 0}, key={seed_lo, seed_hi}). */
int msfem_gen_coef_lognormal(int dim, int n_cells, const int32_t *n_fine, uint64_t seed, double clip,
                             double *coef);
/* Uneven batch: per cell a 16^2 probe field of SPD tensors; its contrast
 * log(max lambda / min lambda) binned at the quartiles of all cells picks the
 * fine resolution classes[0..3] (low to high contrast) -> n_fine[nc^2],
 * contrast[nc^2] (may be NULL).  2D. */
int msfem_gen_resolution_classes(int nc, uint64_t seed, double sigma, const int32_t classes[4], int32_t *n_fine,
                                 double *contrast);
/* SPD tensor per fine cell R(theta) diag(l1,l2) R(theta)^T, theta ~ U[0,pi),
 * log l ~ N(0, sigma) -> TENSOR_CELL tiles.  2D. */
int msfem_gen_coef_tensor(int n_cells, const int32_t *n_fine, uint64_t seed, double sigma, double *coef);
/* Gaussian random field log-coefficient tiles (nf x nf per cell): white
 * noise convolved with a Gaussian kernel so that the covariance is
 * exp(-r^2 / (2 corr_len^2)), unit variance, clamped to [-clip, clip].
 * Writes float32 g [n_cells][nf*nf]; alpha = exp(g). */
int msfem_gen_grf_log(int n_cells, int nf, uint64_t seed, double corr_len, double clip, float *g);

/* ---- learned basis surrogate: network primitives (device pointers) -------
 * No counterpart in the reference.  Building blocks of the convolutional
 * surrogate that predicts a cell's bases from its log-coefficient tile and
 * serves them as the initial guess of msfem_batch_build_device (x0), whose
 * FP64 residual test accepts the prediction or continues the solve.
 * Activations FP32 in [channel][sample][y][x]; BF16 values are raw uint16. */
/* C[M][N] (+)= op(A)[M][K] op(B)[K][N]; A [M][K] (transA: [K][M]), B [K][N]
 * (transB: [N][K]) BF16 row-major; C FP32 row-major; MFMA with FP32 accumulate. */
int msfem_nn_gemm_bf16(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                       float *C, int ldc, int accumulate, void *stream);
/* The same with the K range split over `splits` workgroup layers (for small
 * M x N with long K, e.g. weight gradients): work holds splits*M*N floats,
 * the partial products are summed in a fixed order (deterministic). */
int msfem_nn_gemm_bf16_splitk(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B,
                              int ldb, float *C, int ldc, int accumulate, int splits, float *work, void *stream);
/* col[(c*kh+ky)*kw+kx][(b*Ho+y)*Wo+x] = bf16(in[c][b][y+ky-pad][x+kx-pad]) (0 outside) */
int msfem_nn_im2col(int C, int B, int H, int W, int kh, int kw, int pad, const float *in, void *col, void *stream);
/* Implicit-GEMM convolution: out[Cout][B Ho Wo] = w[Cout][Cin kh kw] x
 * im2col(in) with the column matrix gathered inside the GEMM's tile loads
 * (never stored); the same values as msfem_nn_im2col + msfem_nn_gemm_bf16.
 * Requires Wo % 8 == 0 (MSFEM_ERR_ARG otherwise).  splits > 1: split K with
 * a fixed-order reduction through work[splits][Cout][B Ho Wo] (FP32).
 * bias (or NULL): out = act(out + bias[row]) as msfem_nn_bias_act, fused
 * into the GEMM's epilogue (relu: 1 = ReLU, 0 = identity). */
int msfem_nn_conv_bf16(int Cin, int B, int H, int W, int kh, int kw, int pad, const void *w, int Cout,
                       const float *in, float *out, int splits, float *work, const float *bias, int relu,
                       void *stream);
/* Direct 3x3 convolution (padding 1): the input rows of a 128-pixel output
 * tile are staged once per 32-channel block in LDS and shared by the 9 taps.
 * Requires Cin % 32 == 0, Cout % 32 == 0, W in {16, 32, 64} with
 * H % (128 / W) == 0, or W = H in {4, 8} with B H W % 128 == 0 (whole images
 * per tile), and 16-byte aligned wp and
 * in (MSFEM_ERR_ARG otherwise).
 * wp = the BF16 weights [Cout][Cin 3 3] regrouped by msfem_nn_conv3x3_pack
 * (Cout * Cin * 9 values).  Same result as msfem_nn_conv_bf16 with kh = kw = 3,
 * pad = 1 up to FP32 rounding of the sums (the k order differs).
 * upsample = 1: in is [Cin][B][H/2][W/2] (8-byte aligned), read as its
 * nearest upsampling to H x W (the values of msfem_nn_up2, never stored).
 * add (or NULL): out = act(out + bias[row]) + add, one FP32 addition per
 * value (a skip connection; bitwise what msfem_nn_axpby(1, y, 1, add) gives). */
int msfem_nn_conv3x3_pack(int Cin, int Cout, const void *w, void *wp, void *stream);
int msfem_nn_conv3x3_bf16(int Cin, int B, int H, int W, const void *wp, int Cout, const float *in, int upsample,
                          float *out, const float *bias, int relu, const float *add, void *stream);
/* The same, also writing pool (or nothing for NULL) = the 2x2 average pooling
 * of out, [Cout][B][H/2][W/2] FP32, bitwise what msfem_nn_pool2 of out gives
 * (an encoder layer and the pooling after it in one pass; H even), with the
 * activations optionally stored as BF16 (io bit 0: in holds BF16 values, 4-
 * or 8-byte aligned; bit 1: out receives each result rounded to BF16, to
 * nearest even).  A consumer that rounds its input to BF16 anyway (the 3x3
 * convolutions here, msfem_nn_head_bases) computes bitwise the same values
 * from the BF16 store as from the FP32 one. */
int msfem_nn_conv3x3_bf16_ex(int Cin, int B, int H, int W, const void *wp, int Cout, const void *in, int upsample,
                             void *out, const float *bias, int relu, const float *add, float *pool, int io,
                             void *stream);
/* A network's first two layers in one pass: h = act0(conv3x3(in, w0) + b0)
 * (one input channel -> 32, padding 1; w0 BF16 [32][9] as msfem_nn_conv_bf16
 * takes it) is formed in the staging of out = act(conv3x3(h) + bias)
 * (32 -> Cout = 32, wp packed by msfem_nn_conv3x3_pack), with the optional
 * 2x2 average pooling of out as msfem_nn_conv3x3_bf16_pool.  in: [B][H][64]
 * FP32, 16-byte aligned, H even.  Bitwise equal to msfem_nn_conv_bf16 (no
 * split) followed by msfem_nn_conv3x3_bf16_ex, without h in memory. */
int msfem_nn_conv3x3_first_bf16(int B, int H, int W, const void *w0, const float *b0, int relu0, const float *in,
                                const void *wp, int Cout, float *out, const float *bias, int relu, float *pool,
                                void *stream);
/* Direct convolution for few output channels (Cout <= 8: a network's head)
 * on the VALU: exact BF16 x BF16 products of the rounded input and weights w
 * [Cout][Cin kh kw] summed in FP32; bias/relu as msfem_nn_conv_bf16.  The same
 * values as the im2col + GEMM form up to FP32 rounding of the sums.  Requires
 * kh = kw in {1, 2, 3} and Cin kh kw C <= 16384 with C = 4 (Cout <= 4) or 8. */
int msfem_nn_conv_small(int Cin, int B, int H, int W, int kh, int kw, int pad, const void *w, int Cout,
                        const float *in, float *out, const float *bias, int relu, void *stream);
/* The basis surrogate's head written as bases: bases[b][c][Y (nf+1) + X] =
 * N_c(X, Y) + (double)(scale * h_c) at the interior nodes of cell b, h = the
 * 2x2 valid convolution (Cin -> 4, BF16 weights w [4][Cin 2 2]) + bias (+
 * ReLU) of the activations in [Cin][B][nf][nf] at pixel (X-1, Y-1), and N_c
 * on the boundary: bitwise msfem_nn_conv_small, msfem_nn_axpby(scale, h, 0)
 * and msfem_nn_delta_to_bases in one pass.  in_bf16: in holds BF16 values. */
int msfem_nn_head_bases(int Cin, int B, int nf, const void *w, const void *in, int in_bf16, const float *bias,
                        int relu, float scale, double *bases, void *stream);
/* adjoint of im2col (FP32, gather form) */
int msfem_nn_col2im(int C, int B, int H, int W, int kh, int kw, int pad, const float *dcol, float *dx, void *stream);
int msfem_nn_bias_act(int C, uint64_t N, const float *bias, float *y, int relu, void *stream);
int msfem_nn_act_bwd(uint64_t n, const float *y, float *dy, void *dy_bf16, int relu, void *stream);
int msfem_nn_bias_grad(int C, uint64_t N, const float *d, float *db, void *stream);
int msfem_nn_pool2(int planes, int H, int W, const float *in, float *out, int backward, void *stream);
int msfem_nn_up2(int planes, int H, int W, const float *in, float *out, int backward, void *stream);
int msfem_nn_axpy(uint64_t n, float a, const float *x, float *y, void *stream);
/* out = a x + b y (y unread when b == 0) */
int msfem_nn_axpby(uint64_t n, float a, const float *x, float b, const float *y, float *out, void *stream);
int msfem_nn_to_bf16(uint64_t n, const float *x, void *y, void *stream);
/* grad = 2 (pred - target) / n; partial[b] = per-block sums of squares (n_partial blocks) */
int msfem_nn_mse(uint64_t n, const float *pred, const float *target, float *grad, float *partial, int n_partial,
                 void *stream);
int msfem_nn_adam(uint64_t n, float *w, const float *g, float *m, float *v, void *w_bf16, float lr, float beta1,
                  float beta2, float eps, int step, float weight_decay, void *stream);
/* bases [cell][4][(nf+1)^2] FP64 = N_i + delta_i (interior), delta [4][cells][(nf-1)^2] FP32; and the inverse */
int msfem_nn_delta_to_bases(int n_cells, int nf, const float *delta, double *bases, void *stream);
int msfem_nn_bases_to_delta(int n_cells, int nf, const double *bases, float *delta, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* MSFEM_H_ */
