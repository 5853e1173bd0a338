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
  kModeGlobal = 2, // per-fine-cell element conductances read from global memory (prep buffer)
  kModeStream = 3, // FP32 inner CG: 4 FP32 conductances per fine cell streamed from global memory
                   // (planar prep buffer), the CG product vector w kept in LDS
  kModeGeneric = 4 // any resolution / coefficient kind: QGauss(2) element matrices in global memory,
                   // one workgroup per problem, FP64 (generic.hip)
};

// kModeStream node couplings: four planes (aE, aN, aNE, aNW: the couplings of
// node (X, Y) with (X+1, Y), (X, Y+1), (X+1, Y+1), (X-1, Y+1); K = sum over
// couplings a (e_C - e_k)(e_C - e_k)^T), entry Y (nf+1) + X, plane stride
// below (padding keeps the kernels' unguarded reads in bounds); FP64 planes
// first, then FP32 copies: 6 planes' worth of doubles per cell.
__host__ __device__ constexpr int coupling_plane_stride(int nf) { return (nf + 1) * (nf + 1) + 2 * (nf + 1); }

// FP64 values of q per problem in the FP64 streamed kernel's global work
// array: one per owned node slot (R rows x CPL columns x threads), which is
// nf^2 for every launch shape with SW * R = nf and nf = 64 CPL lanes.
__host__ __device__ constexpr int stream_q_slots(int nf) { return nf * nf; }

// Two-level preconditioner of the 64^2 local problems: bilinear coarse space
// of the 8 x 8 coarse grid (spacing kCoarseH fine cells) of the tile, its
// kCoarseSide^2 interior nodes (zero Dirichlet on the tile boundary).
constexpr int kCoarseH = 8, kCoarseSide = 7, kCoarseDofs = kCoarseSide * kCoarseSide;
// Intermediate levels of the 64^2 two-level preconditioner: inverse diagonals
// of the spacing-2 (31 x 31 interior nodes) and spacing-4 (15 x 15) hat levels
// per cell, FP32, padded to a multiple of 4 floats.
constexpr int kMl2 = 31 * 31, kMl4 = 15 * 15, kMlFloats = (kMl2 + kMl4 + 3) / 4 * 4;

// Workgroup -> block index such that the blocks one XCD receives (hardware
// dispatch is round-robin over the 8 XCDs) hold consecutive problems: the
// bases of one coarse cell then run on one XCD at the same time and share its
// L2.  grid must be a multiple of 8; a bijection on [0, grid).
__device__ __forceinline__ int xcd_block(int b, int grid)
{
  return (b & 7) * (grid >> 3) + (b >> 3);
}

// Arguments of the local basis solves of one size class (one fine resolution).
struct LocalArgs {
  int n_problems;             // cells_in_class * nb_solve
  int nb_solve;               // bases solved per cell
  int nb;                     // 2^dim
  const int *cells;           // [cells_in_class] batch cell index
  const double *box;          // [n_cells][dim][2] (unused for general cells)
  const double *verts;        // [n_cells][2^d][d] general quadrilaterals / hexahedra or nullptr: boxes
  const double *coef;         // concatenated coefficient tiles
  const long long *coef_off;  // [n_cells] tile offsets (elements)
  const double *cond;         // kModeGlobal: concatenated conductance tiles
  const float *cond32;        // kModeStream: planar FP32 conductance tiles (same offsets, in floats)
  double *work;               // kModeStream FP64: per-problem q slots (R CPL threads doubles each)
  const long long *cond_off;  // [n_cells]
  double *bases;              // concatenated basis blocks (working FP64 solution + output)
  const long long *bas_off;   // [n_cells]
  const double *x0;           // optional initial guesses (bases layout) or nullptr
  int q1_start;               // without x0: 1 = interior starts at the Q1 interpolant of g
  const double *cinv;         // two-level: [cells in class][kCoarseDofs^2] inverse coarse matrices, or nullptr
  const float *mlinv;         // 64^2 two-level: [cells in class][kMlFloats] inverse diagonals of the
                              // intermediate hat levels (spacing 2 and 4), or nullptr: two levels only
  int *iters;                 // [n_cells][nb] or nullptr
  int *status;                // [n_cells][nb] or nullptr
  double *resid;              // [n_cells][nb] or nullptr
  const double *ptol;         // [n_cells][nb] absolute residual tolerance per problem (energy
                              // stopping rule) or nullptr: max(abs_tol, rel_tol ||b||)
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

// DPP move of a double (two 32-bit v_mov_dpp); CTRL is a GFX9 dpp_ctrl code.
// Every control used here (quad_perm, row_ror) reads a valid lane of the same
// row and all rows and banks are enabled, so no lane keeps an old value: the
// move needs no initialised destination (bound_ctrl set, old undefined), which
// saves a v_mov per DPP step.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v)
{
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float readlane_f32(float v, int lane)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <typename S, int CTRL>
__device__ __forceinline__ S dpp_any(S v)
{
  if constexpr (sizeof(S) == 8)
    return dpp_f64<CTRL>(v);
  else
    return dpp_f32<CTRL>(v);
}

template <typename S>
__device__ __forceinline__ S readlane_any(S v, int lane)
{
  if constexpr (sizeof(S) == 8)
    return readlane_f64(v, lane);
  else
    return readlane_f32(v, lane);
}

// Sum of the values of lanes 0..N-1 of the wave (N <= 16, a power of two),
// returned identically in every lane: DPP steps inside row 0, then readlane.
template <typename S, int N>
__device__ __forceinline__ S first_lanes_sum(S v)
{
  static_assert(N >= 1 && N <= 16 && (N & (N - 1)) == 0, "N must be a power of two <= 16");
  if constexpr (N >= 2)
    v += dpp_any<S, 0xB1>(v); // pairs
  if constexpr (N >= 4)
    v += dpp_any<S, 0x4E>(v); // quads
  if constexpr (N >= 8)
    {
      // two row rotations: every lane of row 0 then holds the whole row
      // (lanes N..15 contribute zeros), whichever way the rotation turns
      v += dpp_any<S, 0x124>(v);
      v += dpp_any<S, 0x128>(v);
    }
  return readlane_any<S>(v, 0);
}

// Generic-type version of group_sum_dpp (FP32 partial sums move with one DPP
// instruction per step instead of two).
template <typename S, int LW>
__device__ __forceinline__ S group_sum_any(S v)
{
  v += dpp_any<S, 0xB1>(v);
  v += dpp_any<S, 0x4E>(v);
  v += dpp_any<S, 0x124>(v);
  v += dpp_any<S, 0x128>(v);
  if constexpr (LW == 64)
    return (readlane_any<S>(v, 0) + readlane_any<S>(v, 16)) + (readlane_any<S>(v, 32) + readlane_any<S>(v, 48));
  else if constexpr (LW == 32)
    {
      const int base = (int)(threadIdx.x & (kWave - 1)) & ~31;
      return __shfl(v, base, kWave) + __shfl(v, base + 16, kWave);
    }
  else
    return __shfl(v, (int)(threadIdx.x & (kWave - 1)) & ~15, kWave);
}

// Sum over a lane group of LW = 16, 32 or 64 lanes, identical in every lane of
// the group: DPP within each 16-lane row (quad_perm [1,0,3,2], [2,3,0,1],
// row_ror 4, row_ror 8), then the row totals of one lane per row combined in a
// fixed order (readlane for a whole wave, ds_bpermute otherwise).
template <int LW>
__device__ __forceinline__ double group_sum_dpp(double v)
{
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  if constexpr (LW == 64)
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
  else if constexpr (LW == 32)
    {
      const int base = (int)(threadIdx.x & (kWave - 1)) & ~31;
      return __shfl(v, base, kWave) + __shfl(v, base + 16, kWave);
    }
  else
    return __shfl(v, (int)(threadIdx.x & (kWave - 1)) & ~15, kWave);
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

// Sum over each group of 8 consecutive lanes, in every lane of the group
// (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror); bitwise identical in the 8
// lanes since every step adds the same two values in either order.
template <typename S>
__device__ __forceinline__ S oct_sum_any(S v)
{
  v += dpp_any<S, 0xB1>(v);
  v += dpp_any<S, 0x4E>(v);
  v += dpp_any<S, 0x141>(v);
  return v;
}

// Sum over the 16 lanes of each DPP row, in every lane of the row.
template <typename S>
__device__ __forceinline__ S row_sum_any(S v)
{
  v += dpp_any<S, 0xB1>(v);
  v += dpp_any<S, 0x4E>(v);
  v += dpp_any<S, 0x124>(v);
  v += dpp_any<S, 0x128>(v);
  return v;
}

// Shared-memory slots (type S) that problem_sums<S, LW, SW, PPW, NV> uses.
template <int SW, int PPW, int NV>
__host__ __device__ constexpr int problem_sums_slots()
{
  return (PPW == 1 && SW > 1) ? 4 * SW * NV : SW * PPW * (NV + 1);
}

// Per-problem sums of NV values (type S) over LW lanes x SW waves, plus the OR
// of a per-problem flag over the whole workgroup (the loop-continuation test).
// red: problem_sums_slots<SW, PPW, NV>() values of type S.  Identical results
// in every lane.
template <typename S, int LW, int SW, int PPW, int NV>
__device__ __forceinline__ bool problem_sums(S (&v)[NV], bool flag, S *red, int wave, int slot, int pl)
{
  if constexpr (PPW == 1 && SW > 1)
    {
      // One problem per workgroup: the flag is the same in every lane and is
      // returned as it is.  The 4 x SW DPP-row totals go to LDS straight from
      // the rows (no cross-row step inside the wave); lanes 0..15 then add the
      // totals of rows l, l + 16, ... and a DPP sum over row 0 plus a
      // readlane gives every lane the same value, in a fixed order.
      static_assert(LW == kWave && SW <= 16, "whole waves, at most 16");
#pragma unroll
      for (int k = 0; k < NV; ++k)
        v[k] = row_sum_any<S>(v[k]);
      if ((pl & 15) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          red[(wave * 4 + (pl >> 4)) * NV + k] = v[k];
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NV; ++k)
        {
          S part = 0;
#pragma unroll
          for (int j = 0; j < 4 * SW; j += 16)
            if (pl < 16 && j + pl < 4 * SW)
              part += red[(j + pl) * NV + k];
          v[k] = first_lanes_sum<S, 16>(part);
        }
      return flag;
    }
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = group_sum_any<S, LW>(v[k]);
  if constexpr (SW == 1)
    return __syncthreads_or(flag);
  else
    {
      if (pl == 0)
        {
#pragma unroll
          for (int k = 0; k < NV; ++k)
            red[(wave * PPW + slot) * (NV + 1) + k] = v[k];
          red[(wave * PPW + slot) * (NV + 1) + NV] = flag ? (S)1 : (S)0;
        }
      __syncthreads();
      // lane j of the group holds wave j's partials (j < SW <= 16 <= LW), then a
      // DPP sum over those lanes: few live registers, identical totals in every lane
      static_assert(SW <= 16, "at most 16 stacked waves");
      const bool has = pl < SW;
      S f = 0;
#pragma unroll
      for (int k = 0; k <= NV; ++k)
        {
          const S part = has ? red[(pl * PPW + slot) * (NV + 1) + k] : (S)0;
          S tot;
          if constexpr (PPW == 1)
            tot = first_lanes_sum<S, SW>(part);
          else
            tot = group_sum_any<S, LW>(part);
          if (k < NV)
            v[k] = tot;
          else
            f = tot;
        }
      return __ballot(f != (S)0) != 0;
    }
}

// a / b for the CG scalars.  FP32: a times the hardware reciprocal (v_rcp_f32,
// within 1 ulp) -- two instructions instead of the scaled IEEE division
// sequence; the CG coefficients tolerate the last-bit difference.
template <typename S>
__device__ __forceinline__ S quotient(S a, S b)
{
  if constexpr (sizeof(S) == 4)
    return a * __builtin_amdgcn_rcpf(b);
  else
    return a / b;
}

// 1 / b for the CG scalars (b > 0 and normal there).  FP32: v_rcp_f32.
// FP64: the hardware estimate (v_rcp_f64) refined by two Newton steps
// r <- r + r (1 - b r), each of which squares the relative error, so the
// result is correctly rounded or one ulp off -- five instructions instead of
// the scaled IEEE division sequence.
template <typename S>
__device__ __forceinline__ S reciprocal(S b)
{
  if constexpr (sizeof(S) == 4)
    return __builtin_amdgcn_rcpf(b);
  else
    {
      double r = __builtin_amdgcn_rcp(b);
      double e = __builtin_fma(-b, r, 1.0);
      r = __builtin_fma(r, e, r);
      e = __builtin_fma(-b, r, 1.0);
      return __builtin_fma(r, e, r);
    }
}


// QGauss(2) element matrix of fine cell f of a 2D coefficient tile (any kind) on
// an hx x hy cell (diffusion_problem_basis.tpp:207-222); vertex order (0,0),
// (1,0), (0,1), (1,1).
__device__ inline void qg_element_2d(int coef_kind, const double *cf, int f, double hx, double hy, double (&K)[4][4])
{
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v)
      K[u][v] = 0;
  const double s = 0.5 / sqrt(3.0);
  for (int q = 0; q < 4; ++q)
    {
      const double xi = (q & 1) ? 0.5 + s : 0.5 - s, eta = (q & 2) ? 0.5 + s : 0.5 - s;
      double A11, A12, A22;
      switch (coef_kind)
        {
          case 0: A11 = A22 = cf[f]; A12 = 0; break;
          case 1: A11 = cf[3 * f]; A12 = cf[3 * f + 1]; A22 = cf[3 * f + 2]; break;
          case 2: A11 = A22 = cf[4 * f + q]; A12 = 0; break;
          default: A11 = cf[12 * f + 3 * q]; A12 = cf[12 * f + 3 * q + 1]; A22 = cf[12 * f + 3 * q + 2]; break;
        }
      double gx[4], gy[4];
      for (int v = 0; v < 4; ++v)
        {
          const double fx = (v & 1) ? xi : 1.0 - xi, fy = (v & 2) ? eta : 1.0 - eta;
          gx[v] = ((v & 1) ? 1.0 : -1.0) * fy / hx;
          gy[v] = ((v & 2) ? 1.0 : -1.0) * fx / hy;
        }
      const double jxw = 0.25 * hx * hy;
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v)
          K[u][v] += (gx[u] * (A11 * gx[v] + A12 * gy[v]) + gy[u] * (A12 * gx[v] + A22 * gy[v])) * jxw;
    }
}

// Bilinear map F(xi, eta) = v0 + xi a + eta b + xi eta c of a quadrilateral
// with vertices v[4][2] in deal.II order (a = v1 - v0, b = v2 - v0,
// c = v3 - v2 - v1 + v0).
struct QuadMap {
  double x0, y0, ax, ay, bx, by, cx, cy;
  __device__ explicit QuadMap(const double *v)
    : x0(v[0]), y0(v[1]), ax(v[2] - v[0]), ay(v[3] - v[1]), bx(v[4] - v[0]), by(v[5] - v[1]),
      cx(v[6] - v[4] - v[2] + v[0]), cy(v[7] - v[5] - v[3] + v[1])
  {
  }
  __device__ void point(double xi, double eta, double &x, double &y) const
  {
    x = x0 + xi * ax + eta * bx + xi * eta * cx;
    y = y0 + xi * ay + eta * by + xi * eta * cy;
  }
};

// QGauss(2) element matrix and load of fine cell f of a general quadrilateral
// refined nf times per side (diffusion_problem_basis.tpp:161-242 on the mesh of
// GridGenerator::general_cell + refine_global, :92-99).  The fine cell is the
// image of the parametric square [fx, fx+1] x [fy, fy+1] / nf under the coarse
// bilinear map F, which is also its own bilinear mapping, so its Jacobian at
// reference point (s, t) is dF/d(xi, eta) / nf at xi = (fx + s)/nf,
// eta = (fy + t)/nf; grad psi = J^{-T} grad_(s,t) psi, JxW = det J / 4;
// coefficient and f at the mapped point.  be may be nullptr.
__device__ inline void qg_element_quad(int coef_kind, const double *cf, int f, int nf, const QuadMap &m,
                                       double (&K)[4][4], int rhs_kind, double rhs_value, double *be)
{
  for (int u = 0; u < 4; ++u)
    for (int v = 0; v < 4; ++v)
      K[u][v] = 0;
  if (be)
    for (int u = 0; u < 4; ++u)
      be[u] = 0;
  const double gs = 0.5 / sqrt(3.0), inv = 1.0 / nf;
  const int fx = f % nf, fy = f / nf;
  for (int q = 0; q < 4; ++q)
    {
      const double s = (q & 1) ? 0.5 + gs : 0.5 - gs, t = (q & 2) ? 0.5 + gs : 0.5 - gs;
      const double xi = (fx + s) * inv, eta = (fy + t) * inv;
      const double J00 = (m.ax + eta * m.cx) * inv, J01 = (m.bx + xi * m.cx) * inv;
      const double J10 = (m.ay + eta * m.cy) * inv, J11 = (m.by + xi * m.cy) * inv;
      const double det = J00 * J11 - J01 * J10, rdet = 1.0 / det;
      double A11, A12, A22;
      switch (coef_kind)
        {
          case 0: A11 = A22 = cf[f]; A12 = 0; break;
          case 1: A11 = cf[3 * f]; A12 = cf[3 * f + 1]; A22 = cf[3 * f + 2]; break;
          case 2: A11 = A22 = cf[4 * f + q]; A12 = 0; break;
          default: A11 = cf[12 * f + 3 * q]; A12 = cf[12 * f + 3 * q + 1]; A22 = cf[12 * f + 3 * q + 2]; break;
        }
      double gx[4], gy[4], val[4];
      for (int v = 0; v < 4; ++v)
        {
          const double fs = (v & 1) ? s : 1.0 - s, ft = (v & 2) ? t : 1.0 - t;
          const double ds = ((v & 1) ? 1.0 : -1.0) * ft, dt = ((v & 2) ? 1.0 : -1.0) * fs;
          gx[v] = (J11 * ds - J10 * dt) * rdet;
          gy[v] = (J00 * dt - J01 * ds) * rdet;
          val[v] = fs * ft;
        }
      const double jxw = 0.25 * det;
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v)
          K[u][v] += (gx[u] * (A11 * gx[v] + A12 * gy[v]) + gy[u] * (A12 * gx[v] + A22 * gy[v])) * jxw;
      if (be)
        {
          double fv = rhs_value;
          if (rhs_kind == 1)
            {
              double x, y;
              m.point(xi, eta, x, y);
              fv = sin(M_PI * x) * sin(M_PI * y);
            }
          for (int u = 0; u < 4; ++u)
            be[u] += val[u] * fv * jxw;
        }
    }
}

// Trilinear map of a hexahedron with vertices v[8][3] in deal.II order (vertex
// w at the parametric corner (w&1, (w>>1)&1, (w>>2)&1)), in monomial form
// F = c0 + xi c1 + eta c2 + zeta c3 + xi eta c4 + eta zeta c5 + xi zeta c6
//     + xi eta zeta c7.
struct HexMap {
  double c[8][3];
  __device__ explicit HexMap(const double *v)
  {
    for (int k = 0; k < 3; ++k)
      {
        const double v0 = v[k], v1 = v[3 + k], v2 = v[6 + k], v3 = v[9 + k];
        const double v4 = v[12 + k], v5 = v[15 + k], v6 = v[18 + k], v7 = v[21 + k];
        c[0][k] = v0;
        c[1][k] = v1 - v0;
        c[2][k] = v2 - v0;
        c[3][k] = v4 - v0;
        c[4][k] = v3 - v2 - v1 + v0;
        c[5][k] = v6 - v4 - v2 + v0;
        c[6][k] = v5 - v4 - v1 + v0;
        c[7][k] = v7 - v6 - v5 - v3 + v4 + v2 + v1 - v0;
      }
  }
  __device__ void point(double xi, double eta, double zeta, double (&x)[3]) const
  {
    for (int k = 0; k < 3; ++k)
      x[k] = c[0][k] + xi * c[1][k] + eta * c[2][k] + zeta * c[3][k] + xi * eta * c[4][k] + eta * zeta * c[5][k] +
             xi * zeta * c[6][k] + xi * eta * zeta * c[7][k];
  }
  // J[k][l] = d x_k / d xi_l
  __device__ void jacobian(double xi, double eta, double zeta, double (&J)[3][3]) const
  {
    for (int k = 0; k < 3; ++k)
      {
        J[k][0] = c[1][k] + eta * c[4][k] + zeta * c[6][k] + eta * zeta * c[7][k];
        J[k][1] = c[2][k] + xi * c[4][k] + zeta * c[5][k] + xi * zeta * c[7][k];
        J[k][2] = c[3][k] + eta * c[5][k] + xi * c[6][k] + xi * eta * c[7][k];
      }
  }
};

// QGauss(2) element matrix (and optionally load) of fine cell f of a general
// hexahedron refined nf times per side (diffusion_problem_basis.tpp:161-242 on
// the mesh of GridGenerator::general_cell + refine_global, :92-99): the fine
// cell is the image of the parametric box [fx, fx+1] x [fy, fy+1] x [fz, fz+1]
// / nf under the coarse trilinear map F, also its own trilinear mapping, so its
// Jacobian at reference point s is dF/dxi / nf at xi = (f + s) / nf;
// grad psi = J^{-T} grad_s psi (J^{-1} by cofactors), JxW = det J / 8;
// coefficient and f at the mapped point.  K: [8 x 8] row-major.  be may be nullptr.
__device__ inline void qg_element_hex(int coef_kind, const double *cf, int f, int nf, const HexMap &m,
                                      double (&K)[64], int rhs_kind, double rhs_value, double *be)
{
#pragma unroll
  for (int uv = 0; uv < 64; ++uv)
    K[uv] = 0;
  if (be)
    for (int u = 0; u < 8; ++u)
      be[u] = 0;
  const double gs = 0.5 / sqrt(3.0), inv = 1.0 / nf;
  const int fi[3] = {f % nf, (f / nf) % nf, f / (nf * nf)};
  for (int q = 0; q < 8; ++q)
    {
      double s[3], xi[3];
#pragma unroll
      for (int k = 0; k < 3; ++k)
        {
          s[k] = ((q >> k) & 1) ? 0.5 + gs : 0.5 - gs;
          xi[k] = (fi[k] + s[k]) * inv;
        }
      double J[3][3];
      m.jacobian(xi[0], xi[1], xi[2], J);
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          J[k][l] *= inv;
      double C[3][3]; // cofactors: J^{-1} = C^T / det
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          {
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3, l1 = (l + 1) % 3, l2 = (l + 2) % 3;
            C[k][l] = J[k1][l1] * J[k2][l2] - J[k1][l2] * J[k2][l1];
          }
      const double det = J[0][0] * C[0][0] + J[0][1] * C[0][1] + J[0][2] * C[0][2], rdet = 1.0 / det;
      double A[3][3];
      const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          switch (coef_kind)
            {
              case 0: A[k][l] = k == l ? cf[f] : 0.0; break;
              case 1: A[k][l] = cf[6 * (size_t)f + sidx[k][l]]; break;
              case 2: A[k][l] = k == l ? cf[8 * (size_t)f + q] : 0.0; break;
              default: A[k][l] = cf[48 * (size_t)f + 6 * q + sidx[k][l]]; break;
            }
      double g[8][3], val[8];
#pragma unroll
      for (int v = 0; v < 8; ++v)
        {
          double f1[3], gr[3];
#pragma unroll
          for (int k = 0; k < 3; ++k)
            f1[k] = ((v >> k) & 1) ? s[k] : 1.0 - s[k];
          gr[0] = ((v & 1) ? 1.0 : -1.0) * f1[1] * f1[2];
          gr[1] = ((v & 2) ? 1.0 : -1.0) * f1[0] * f1[2];
          gr[2] = ((v & 4) ? 1.0 : -1.0) * f1[0] * f1[1];
          val[v] = f1[0] * f1[1] * f1[2];
#pragma unroll
          for (int k = 0; k < 3; ++k)
            g[v][k] = (C[k][0] * gr[0] + C[k][1] * gr[1] + C[k][2] * gr[2]) * rdet;
        }
      const double jxw = 0.125 * det;
#pragma unroll
      for (int u = 0; u < 8; ++u)
        {
          double Ag[3];
#pragma unroll
          for (int k = 0; k < 3; ++k)
            Ag[k] = (A[k][0] * g[u][0] + A[k][1] * g[u][1] + A[k][2] * g[u][2]) * jxw;
#pragma unroll
          for (int v = 0; v < 8; ++v)
            K[u * 8 + v] += Ag[0] * g[v][0] + Ag[1] * g[v][1] + Ag[2] * g[v][2];
        }
      if (be)
        {
          double fv = rhs_value;
          if (rhs_kind == 1)
            {
              double x[3];
              m.point(xi[0], xi[1], xi[2], x);
              fv = sin(M_PI * x[0]) * sin(M_PI * x[1]) * sin(M_PI * x[2]);
            }
          for (int u = 0; u < 8; ++u)
            be[u] += val[u] * fv * jxw;
        }
    }
}

} // namespace msfem
