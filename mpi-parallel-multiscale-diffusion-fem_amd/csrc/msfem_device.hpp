// Device-side building blocks shared by the MsFEM kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msfem {

constexpr int kWave = 64;

// Coefficient storage modes of the local solve kernels.
enum CondMode : int {
  kModeScalar = 0, // one scalar per fine cell, cached in registers
  kModeTensor = 1, // 2D symmetric tensor per fine cell, 4 conductances cached in registers
  kModeGlobal = 2  // per-fine-cell element conductances read from global memory (prep buffer)
};

// Arguments of the local basis solves of one size class (one fine resolution).
struct LocalArgs {
  int n_problems;             // cells_in_class * nb_solve
  int nb_solve;               // bases solved per cell
  int nb;                     // 2^dim
  const int *cells;           // [cells_in_class] batch cell index
  const double *box;          // [n_cells][dim][2]
  const double *coef;         // concatenated coefficient tiles
  const long long *coef_off;  // [n_cells] tile offsets (elements)
  const double *cond;         // kModeGlobal: concatenated conductance tiles
  const long long *cond_off;  // [n_cells]
  double *bases;              // concatenated basis blocks (working FP64 solution + output)
  const long long *bas_off;   // [n_cells]
  const double *x0;           // optional initial guesses (bases layout) or nullptr
  int *iters;                 // [n_cells][nb] or nullptr
  int *status;                // [n_cells][nb] or nullptr
  double *resid;              // [n_cells][nb] or nullptr
  double rel_tol, abs_tol, inner_rel_tol;
  int max_iter, max_refine;
};

// Sum over the LW lanes of a lane group (xor butterfly: every lane gets the
// bitwise identical result, so loop exits stay uniform).
template <int LW>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
  for (int m = 1; m < LW; m <<= 1)
    v += __shfl_xor(v, m, kWave);
  return v;
}

// Q1 nodal function N_i of the coarse cell at fine node (X, Y[, Z]) of an
// nf-grid: tensor product of 1D hats, exact for axis-aligned boxes (the
// closed form of include/coefficients/basis_q1.tpp:26-158 on boxes).
__device__ __forceinline__ double nodal_q1(int i, int dim, int X, int Y, int Z, int nf)
{
  const double inv = 1.0 / nf;
  const double tx = X * inv, ty = Y * inv, tz = Z * inv;
  double v = ((i & 1) ? tx : 1.0 - tx) * ((i & 2) ? ty : 1.0 - ty);
  if (dim == 3)
    v *= (i & 4) ? tz : 1.0 - tz;
  return v;
}

// 2D conductances of a fine box cell (element matrix K_c as a graph Laplacian
// on the 4 vertices: y_a = sum_b c_ab (p_a - p_b)).  rho = hy / hx.
//   horizontal edges (0,1),(2,3): ch ; vertical edges (0,2),(1,3): cv ;
//   diagonal (0,3): cd ; anti-diagonal (1,2): ca.
struct Cond4 {
  double ch, cv, cd, ca;
};

__device__ __forceinline__ Cond4 cond_scalar(double a, double rho)
{
  const double ir = 1.0 / rho;
  return {a * (rho * (1.0 / 3.0) - ir * (1.0 / 6.0)), a * (ir * (1.0 / 3.0) - rho * (1.0 / 6.0)),
          a * ((rho + ir) * (1.0 / 6.0)), a * ((rho + ir) * (1.0 / 6.0))};
}

// Tensor [[a, b], [b, c]]: K_c = a rho Kxx + c/rho Kyy + b Kxy with the unit
// square Q1 matrices Kxx, Kyy and Kxy = (s t^T + t s^T)/4.
__device__ __forceinline__ Cond4 cond_tensor(double a, double b, double c, double rho)
{
  const double ir = 1.0 / rho;
  const double d = a * rho * (1.0 / 6.0) + c * ir * (1.0 / 6.0);
  return {a * rho * (1.0 / 3.0) - c * ir * (1.0 / 6.0), c * ir * (1.0 / 3.0) - a * rho * (1.0 / 6.0), d + 0.5 * b,
          d - 0.5 * b};
}

} // namespace msfem
