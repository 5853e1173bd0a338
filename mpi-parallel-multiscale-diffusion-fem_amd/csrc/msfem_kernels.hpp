// Host-side launchers of the MsFEM kernels (implemented in the .hip files).
#pragma once

#include <hip/hip_runtime.h>

#include "msfem_device.hpp"

namespace msfem {

// local2d.hip
int local2d_variants(int nf);
// whether a class (resolution, storage mode, launch-shape id) runs the two-level form
bool local2d_coarse_space(int nf, int mode, int variant);
// inverse coarse matrices of the two-level form (nf = 64, 128; boxes or general
// quadrilaterals): cinv [cells in class][49 * 49]
hipError_t launch_coarse_setup_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                                  const double *verts, const double *coef, const long long *coef_off, double *cinv,
                                  hipStream_t st);
// the intermediate levels' inverse diagonals of a 64^2 two-level class
// (mlinv [cells in class][kMlFloats])
hipError_t launch_ml_diag_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                             const double *verts, const double *coef, const long long *coef_off, float *mlinv,
                             hipStream_t st);
hipError_t launch_local_solve_2d(bool fp32, int mode, int nf, int variant, const LocalArgs &a, hipStream_t st);
// verts: general quadrilaterals [n_cells][4][2] or nullptr (boxes)
hipError_t launch_prep_couplings_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                                   const double *verts, const double *coef, const long long *coef_off,
                                   double *cond, const long long *cond_off, bool want32, hipStream_t st);
hipError_t launch_prep_cond_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *verts, const double *coef, const long long *coef_off, double *cond,
                               const long long *cond_off, hipStream_t st);
// Dirichlet data N_i (BasisQ1 by monomial inversion) of general quadrilaterals
// at their boundary fine nodes, written into every basis block.
hipError_t launch_quad_dirichlet_2d(int n_cells_class, const int *cells, int nf, const double *verts, double *bases,
                                    const long long *bas_off, hipStream_t st);

// local3d.hip
int local3d_variants(int nf);
hipError_t launch_local_solve_3d(bool fp32, int nf, int variant, const LocalArgs &a, hipStream_t st);

// generic.hip: the local solve for resolutions and coefficient kinds outside the
// tuned kernels.  Element matrices [nf^d][2^d][2^d] per cell (offsets elem_off,
// generic_elem_doubles each); work: generic_work_doubles per problem.
size_t generic_elem_doubles(int dim, int nf);
size_t generic_work_doubles(int dim, int nf);
hipError_t launch_generic_prep(int dim, int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *verts, const double *coef, const long long *coef_off, double *elem,
                               const long long *elem_off, hipStream_t st);
hipError_t launch_generic_solve(int dim, int nf, const LocalArgs &a, double *work, hipStream_t st);
// Dirichlet data N_i (BasisQ1 by monomial inversion) of general hexahedra at
// their boundary fine nodes, written into every basis block.
hipError_t launch_hex_dirichlet_3d(int n_cells_class, const int *cells, int nf, const double *verts, double *bases,
                                   const long long *bas_off, hipStream_t st);

// element.hip: A^K, F^K of every cell of one size class; partition-of-unity fix-up.
struct ElementArgs {
  int n_cells_class;
  const int *cells;
  int dim, nf, nb;
  int coef_kind;              // coefficient kind of the batch (MSFEM_COEF_*)
  const double *box;
  const double *verts;        // general cells [n_cells][2^d][d] or nullptr
  const double *coef;
  const long long *coef_off;
  const double *cond;         // 2D conductance tiles when the class used kModeGlobal (else nullptr)
  const long long *cond_off;
  const double *bases;
  const long long *bas_off;
  int rhs_kind;
  double rhs_value;
  double *Ae, *Fe;
};
hipError_t launch_element(const ElementArgs &a, hipStream_t st);
hipError_t launch_pou_fixup(int n_cells_class, const int *cells, int dim, int nf, double *bases,
                            const long long *bas_off, hipStream_t st);

// energy.hip: energy-norm control of the local solves (box cells).  lam_lb
// [n_cells]: lower bound of lambda_min(K_II); ptol [n_cells][nb] (or nullptr):
// tau sqrt(lam_lb e_lb), the residual norm that guarantees a relative
// energy-norm error <= tau; bound [n_cells][nb]: the a posteriori bound from
// the final residual norms and A^K_ii (lam_lb nullptr: NaN).
hipError_t launch_energy_terms(int dim, int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *coef, const long long *coef_off, double tau, double *lam_lb,
                               double *ptol, hipStream_t st);
hipError_t launch_energy_bound(int n_cells_class, const int *cells, int nb, int pou, const double *lam_lb,
                               const double *resid, const double *Ae, double *bound, hipStream_t st);
// d = x - y (n doubles); err[c][i] = sqrt(max(Ad_ii, 0) / Ar_ii)
hipError_t launch_sub(size_t n, const double *x, const double *y, double *d, hipStream_t st);
hipError_t launch_energy_ratio(int n_cells, int nb, const double *Ad, const double *Ar, double *err, hipStream_t st);

// coarse.hip
// Scratch layout of the multigrid-preconditioned coarse CG (offsets in doubles).
constexpr int kMgMaxLevels = 12;
constexpr int kMgMaxDense = 343; // largest coarsest level solved densely (7^3 or 18^2 nodes)
struct MgLayout {
  int levels;                        // 0: no multigrid for this size (Jacobi-PCG instead)
  int n[kMgMaxLevels];               // cells per side of level l (n[0] = nc)
  long long nn[kMgMaxLevels];        // nodes of level l
  long long off_x[kMgMaxLevels], off_t[kMgMaxLevels], off_d[kMgMaxLevels], off_b[kMgMaxLevels],
    off_A[kMgMaxLevels], off_u[kMgMaxLevels];
  long long off_mask[kMgMaxLevels];  // level l >= 1: constrained-node flags, bytes at (char *)(w + off_mask[l])
  long long off_cx, off_cr, off_cp, off_cq; // CG vectors on level 0
  long long off_dense, off_dmap;     // coarsest Cholesky factor [F][F], free-node numbers
  long long total;
};
MgLayout mg_layout(int dim, int nc);
hipError_t launch_coarse_assemble(int dim, int nc, const double *Ae, const double *Fe, int bc_kind, double *Kc,
                                  double *Fc, unsigned char *cmask, double *cval, hipStream_t st);
hipError_t launch_coarse_solve(int dim, int nc, const double *Kc, const double *Fc, const unsigned char *cmask,
                               const double *cval, double rel_tol, int max_iter, double *u, int *iters,
                               double *work, hipStream_t st);
size_t coarse_solve_work_doubles(int dim, int nc);
hipError_t launch_reconstruct(int n_cells, int nb, const long long *bas_off, const long long *out_off,
                              const int *n_nodes, const double *bases, const double *w, double *u, hipStream_t st);

// noalloc.hip: the fixed device scratch of the allocation-free entry point
// (one copy per GPU).  noalloc_register_device looks up this device's address
// (the device must be current); noalloc_scratch returns it (nullptr before).
hipError_t noalloc_register_device(int device);
double *noalloc_scratch(int device);
size_t noalloc_scratch_bytes();

// fine.hip: fine-scale standard FEM (element matrices + multi-workgroup FP64 CG
// on a structured stencil matrix, scalars in device memory).
struct FineCgState {
  double rz, alpha, beta, rnorm, tol, rel_tol;
  int iters, max_iter, done, converged;
};
struct FineCgArgs {
  int dim, np, nn;
  const double *Kc, *Fc;
  const unsigned char *cmask;
  const double *cval;
  double *x, *r, *p, *q, *dinv, *part; // part: 2 * fine_cg_blocks(nn) doubles
  FineCgState *state;
};
int fine_cg_blocks(int nn);
hipError_t launch_fine_elements_2d(int n, int coef_kind, const double *coef, int rhs_kind, double rhs_value,
                                   double *Ae, double *Fe, hipStream_t st);
hipError_t launch_fine_cg_start(const FineCgArgs &a, hipStream_t st);
hipError_t launch_fine_cg_steps(const FineCgArgs &a, int steps, hipStream_t st);

} // namespace msfem
