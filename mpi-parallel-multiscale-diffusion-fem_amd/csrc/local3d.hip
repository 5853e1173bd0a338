// 3D local basis solves (DiffusionProblemBasis<3>, include/base/diffusion_problem_basis.tpp:440-474):
// K_II x = -K_IB g_i for the 8 trilinear nodal functions g_i, scalar
// coefficient per fine cell, matrix-free 27-point Q1 operator.
//
// One workgroup owns one problem.  Thread t owns the node column
// (X, Y) = (1 + t % NF, 1 + (t / NF) % NF) and R consecutive z-levels of
// stack t / NF^2; the preconditioned residual z of the cell lives in a
// zero-padded LDS grid, the coefficients of the 4 x (R+1) fine cells around
// the column and the CG vectors of the owned nodes in registers.  Same FP64
// refinement structure and the same single-reduction Jacobi PCG (Chronopoulos
// and Gear) as the 2D kernel: one stencil apply and one fused reduction of
// (r.z, w.z, r.r) per iteration, FP32 sums and scalars in the FP32 kernel.
//
// Stencil: (K z)_C = sum over the 8 cells c around C of alpha_c sum over the
// other vertices b of c of u(v_C xor v_b) (z_C - z_b), with the box's unit
// conductances u(d).  On cubic fine cells the face conductances u(1), u(2),
// u(4) vanish, and the kernel then skips those three of the seven terms per
// cell (a uniform branch per problem).
#include <type_traits>

#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

// Unit conductances c(delta) = -K_unit(a, a xor delta) of a box fine cell with
// sides h: K_unit = sum_k (V/h_k^2) D(delta_k) prod_{m!=k} M(delta_m),
// D(0)=1, D(1)=-1, M(0)=1/3, M(1)=1/6 (trilinear Q1 on a box).
__device__ __forceinline__ void unit_cond3(const double h[3], double c[8])
{
  const double V = h[0] * h[1] * h[2];
  c[0] = 0;
  for (int d = 1; d < 8; ++d)
    {
      double k = 0;
      for (int ax = 0; ax < 3; ++ax)
        {
          double t = V / (h[ax] * h[ax]) * (((d >> ax) & 1) ? -1.0 : 1.0);
          for (int m = 0; m < 3; ++m)
            if (m != ax)
              t *= ((d >> m) & 1) ? (1.0 / 6.0) : (1.0 / 3.0);
          k += t;
        }
      c[d] = -k;
    }
}

} // namespace

template <typename T, int NF, int SZ, int R>
__global__ void __launch_bounds__(NF *NF *SZ) local_solve_3d(LocalArgs a)
{
  constexpr int NP = NF + 1;
  constexpr int NT = NF * NF * SZ;
  constexpr int NW = NT / kWave;
  static_assert(NT % kWave == 0, "workgroup must be whole waves");
  static_assert(SZ * R >= NF - 1, "z-levels not covered");
  constexpr int GX = NF + 2, GXY = GX * GX;
  constexpr int GZ = (SZ * R + 2 > NP) ? SZ * R + 2 : NP;

  __shared__ T sgrid[GXY * GZ];
  __shared__ T sred[problem_sums_slots<NW, 1, 3>()];       // inner-CG partial sums
  __shared__ double sred64[problem_sums_slots<NW, 1, 1>()]; // FP64 residual norms

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int prob = xcd_block(blockIdx.x, gridDim.x);
  const bool pvalid = prob < a.n_problems;
  int cell = 0, basis = 0;
  if (pvalid)
    {
      cell = a.cells[prob / a.nb_solve];
      basis = prob % a.nb_solve;
    }
  const int X = 1 + tid % NF, Y = 1 + (tid / NF) % NF, Z0 = 1 + (tid / (NF * NF)) * R;
  const bool colact = pvalid && X <= NF - 1 && Y <= NF - 1;
  T *gwin = sgrid + (Z0 - 1) * GXY + (Y - 1) * GX + (X - 1); // window origin
  T *gown = sgrid + Z0 * GXY + Y * GX + X;

  const double *bx = a.box + (size_t)cell * 6;
  double h[3];
  for (int k = 0; k < 3; ++k)
    h[k] = (bx[2 * k + 1] - bx[2 * k]) / NF;
  double uc64[8];
  unit_cond3(h, uc64);
  T uc[8];
  double ctot64 = 0;
#pragma unroll
  for (int d = 0; d < 8; ++d)
    {
      uc[d] = (T)uc64[d];
      ctot64 += uc64[d];
    }
  const T ctot = (T)ctot64;
  // cubic fine cells: the face conductances vanish (exactly for an exact cube)
  const double umax = fmax(fabs(uc64[3]), fmax(fabs(uc64[5]), fmax(fabs(uc64[6]), fabs(uc64[7]))));
  const bool cube = fmax(fabs(uc64[1]), fmax(fabs(uc64[2]), fabs(uc64[4]))) <= 1e-14 * umax;
  const double *cf = a.coef + a.coef_off[cell];
  double *phi = a.bases + a.bas_off[cell] + (size_t)basis * NP * NP * NP;

  // cells (X-1+j, Y-1+k, Z0-1+i), j,k in {0,1}
  T cc[R + 1][2][2];
#pragma unroll
  for (int i = 0; i <= R; ++i)
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        {
          const int cx = X - 1 + j, cy = Y - 1 + k, cz = Z0 - 1 + i;
          cc[i][k][j] = (pvalid && cx < NF && cy < NF && cz < NF) ? (T)cf[(cz * NF + cy) * NF + cx] : (T)0;
        }
  // Jacobi diagonal; 0 on slots that are not interior nodes of a valid problem
  T dinv[R];
#pragma unroll
  for (int i = 0; i < R; ++i)
    {
      const bool act = colact && Z0 + i <= NF - 1;
      T s = 0;
#pragma unroll
      for (int dz = 0; dz < 2; ++dz)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
          for (int dx = 0; dx < 2; ++dx)
            s += cc[i + dz][dy][dx];
      dinv[i] = act ? (T)1 / (ctot * s) : (T)0;
    }

  for (int k = tid; k < GXY * GZ; k += NT)
    sgrid[k] = 0;
  if (pvalid)
    for (int k = tid; k < NP * NP * NP; k += NT)
      {
        const int XX = k % NP, YY = (k / NP) % NP, ZZ = k / (NP * NP);
        const bool bnd = XX == 0 || XX == NF || YY == 0 || YY == NF || ZZ == 0 || ZZ == NF;
        // interior: the given start, else the Q1 interpolant of the boundary data, else 0
        const double v = bnd || (!a.x0 && a.q1_start)
                           ? nodal_q1(basis, 3, XX, YY, ZZ, NF)
                           : (a.x0 ? a.x0[a.bas_off[cell] + (size_t)basis * NP * NP * NP + k] : 0.0);
        phi[k] = isfinite(v) ? v : 0.0; // a non-finite initial value starts from 0
      }
  __syncthreads();

  // (K phi) at node (X, Y, Z) in FP64 from global memory; bonly: interior values taken as 0
  auto apply64 = [&](int Z, bool bonly) -> double {
    auto v = [&](int XX, int YY, int ZZ) -> double {
      const bool bnd = XX == 0 || XX == NF || YY == 0 || YY == NF || ZZ == 0 || ZZ == NF;
      return (bonly && !bnd) ? 0.0 : phi[(ZZ * NP + YY) * NP + XX];
    };
    const double pc = v(X, Y, Z);
    double y = 0;
    for (int c = 0; c < 8; ++c) // cell = node + (dx,dy,dz), d in {-1,0}
      {
        const int dx = (c & 1) - 1, dy = ((c >> 1) & 1) - 1, dz = ((c >> 2) & 1) - 1;
        const double al = cf[((Z + dz) * NF + (Y + dy)) * NF + (X + dx)];
        const int av = (-dx) | ((-dy) << 1) | ((-dz) << 2); // vertex of the node in the cell
        double s = 0;
        for (int b = 0; b < 8; ++b)
          {
            if (b == av)
              continue;
            const int ox = dx + (b & 1), oy = dy + ((b >> 1) & 1), oz = dz + ((b >> 2) & 1);
            s += uc64[av ^ b] * (pc - v(X + ox, Y + oy, Z + oz));
          }
        y += al * s;
      }
    return y;
  };

  // with a given start the tolerance is still relative to ||b|| = ||K_IB g||
  const bool given_start = a.x0 || a.q1_start;
  double tol_eff = 0.0, bnorm = 0.0;
  if (given_start)
    {
      double s[1] = {0};
#pragma unroll 1
      for (int i = 0; i < R; ++i)
        if (colact && Z0 + i <= NF - 1)
          {
            const double b = apply64(Z0 + i, true);
            s[0] += b * b;
          }
      problem_sums<double, kWave, NW, 1, 1>(s, false, sred64, wave, 0, lane);
      bnorm = sqrt(s[0]);
      tol_eff = a.ptol ? a.ptol[(size_t)cell * a.nb + basis] : fmax(a.abs_tol, a.rel_tol * bnorm);
      __syncthreads();
    }

  T xr[R], rr[R], pr[R], sr[R];
  int it_total = 0, state = pvalid ? 0 : 1, conv = 0;
  double rnorm = 0;
  for (int cyc = 0;; ++cyc)
    {
      // ---- true FP64 residual of the current phi (own slots of the grid carry it to rr)
      double s[1] = {0};
#pragma unroll 1
      for (int i = 0; i < R; ++i)
        if (colact && Z0 + i <= NF - 1)
          {
            const double r64 = -apply64(Z0 + i, false);
            gown[i * GXY] = (T)r64;
            s[0] += r64 * r64;
          }
#pragma unroll
      for (int i = 0; i < R; ++i)
        rr[i] = dinv[i] != (T)0 ? gown[i * GXY] : (T)0;
      problem_sums<double, kWave, NW, 1, 1>(s, false, sred64, wave, 0, lane);
      rnorm = sqrt(s[0]);
      if (cyc == 0 && !given_start)
        {
          bnorm = rnorm;
          tol_eff = a.ptol ? a.ptol[(size_t)cell * a.nb + basis] : fmax(a.abs_tol, a.rel_tol * bnorm);
        }
      if (state == 0)
        {
          if (rnorm <= tol_eff)
            {
              state = 1;
              conv = 1;
            }
          else if (it_total >= a.max_iter || cyc >= a.max_refine)
            state = 1;
        }
      if (state != 0) // uniform over the workgroup (one problem)
        break;
      double inner_tol = 0.5 * tol_eff;
      if (sizeof(T) < sizeof(double))
        inner_tol = fmax(a.inner_rel_tol * rnorm, inner_tol);
      const T inner_tol2 = (T)(inner_tol * inner_tol);
#pragma unroll
      for (int i = 0; i < R; ++i)
        {
          xr[i] = 0;
          pr[i] = 0;
          sr[i] = 0;
        }
      bool irun = true, first = true;
      T gamma_prev = 1, alpha_prev = 1;
      while (true)
        {
          // z = M r into the LDS grid (zero on padding/boundary slots); irun is
          // uniform (one problem per workgroup) and true here
#pragma unroll
          for (int i = 0; i < R; ++i)
            gown[i * GXY] = rr[i] * dinv[i];
          __syncthreads();
          // w = K z with the z-levels slid through a 3x3x3 window; fused partial sums
          T wv[R];
          T sums[3] = {0, 0, 0}; // (r, z), (w, z), (r, r)
          auto stencil = [&](auto cube_tag) {
            constexpr bool CUBE = decltype(cube_tag)::value;
            T w[3][3][3]; // [z][y][x]
#pragma unroll
            for (int zz = 0; zz < 2; ++zz)
#pragma unroll
              for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                for (int xx = 0; xx < 3; ++xx)
                  w[zz][yy][xx] = gwin[zz * GXY + yy * GX + xx];
#pragma unroll
            for (int i = 0; i < R; ++i)
              {
#pragma unroll
                for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                  for (int xx = 0; xx < 3; ++xx)
                    w[2][yy][xx] = gwin[(i + 2) * GXY + yy * GX + xx];
                const T pc = w[1][1][1];
                T y = 0, sa = 0;
#pragma unroll
                for (int c = 0; c < 8; ++c)
                  {
                    const int dx = (c & 1) - 1, dy = ((c >> 1) & 1) - 1, dz = ((c >> 2) & 1) - 1;
                    const T al = cc[i + 1 + dz][1 + dy][1 + dx];
                    const int av = (-dx) | ((-dy) << 1) | ((-dz) << 2);
                    T s2 = 0;
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                      {
                        const int d = av ^ b;
                        if (b == av || (CUBE && (d == 1 || d == 2 || d == 4)))
                          continue;
                        const int ox = dx + (b & 1), oy = dy + ((b >> 1) & 1), oz = dz + ((b >> 2) & 1);
                        s2 += uc[d] * w[1 + oz][1 + oy][1 + ox];
                      }
                    y -= al * s2;
                    sa += al;
                  }
                y += ctot * sa * pc;
                wv[i] = dinv[i] != (T)0 ? y : (T)0;
                sums[0] += rr[i] * pc;
                sums[1] += wv[i] * pc;
                sums[2] += rr[i] * rr[i];
#pragma unroll
                for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                  for (int xx = 0; xx < 3; ++xx)
                    {
                      w[0][yy][xx] = w[1][yy][xx];
                      w[1][yy][xx] = w[2][yy][xx];
                    }
                // keeps the next levels' LDS reads from being hoisted (registers)
                __builtin_amdgcn_sched_barrier(0);
              }
          };
          if (cube) // uniform per problem
            stencil(std::true_type());
          else
            stencil(std::false_type());
          // one problem per workgroup: irun is uniform, the loop ends with it
          problem_sums<T, kWave, NW, 1, 3>(sums, irun, sred, wave, 0, lane);
          if (irun && (sums[2] <= inner_tol2 || it_total >= a.max_iter))
            irun = false;
          if (irun)
            {
              const T gamma = sums[0], delta = sums[1];
              T beta = 0, alpha;
              if (first)
                alpha = quotient<T>(gamma, delta);
              else
                {
                  beta = quotient<T>(gamma, gamma_prev);
                  alpha = quotient<T>(gamma, delta - beta * quotient<T>(gamma, alpha_prev));
                }
#pragma unroll
              for (int i = 0; i < R; ++i)
                {
                  const T z = rr[i] * dinv[i];
                  pr[i] = z + beta * pr[i];
                  sr[i] = wv[i] + beta * sr[i];
                  xr[i] += alpha * pr[i];
                  rr[i] -= alpha * sr[i];
                }
              gamma_prev = gamma;
              alpha_prev = alpha;
              first = false;
              ++it_total;
            }
          if (!irun)
            break;
        }
#pragma unroll
      for (int i = 0; i < R; ++i)
        if (dinv[i] != (T)0)
          phi[((Z0 + i) * NP + Y) * NP + X] += (double)xr[i];
      __syncthreads();
    }
  if (pvalid && tid == 0)
    {
      const size_t o = (size_t)cell * a.nb + basis;
      if (a.iters)
        a.iters[o] = it_total;
      if (a.status)
        a.status[o] = conv ? 0 : 1;
      if (a.resid)
        a.resid[o] = rnorm;
    }
}

namespace {
template <typename T, int NF, int SZ, int R>
hipError_t launch3(const LocalArgs &a, hipStream_t st)
{
  if (a.n_problems == 0)
    return hipSuccess;
  // a multiple of 8 blocks (xcd_block is a bijection on it); blocks past the
  // last problem find no valid problem and leave at the first residual test
  const int grid = (a.n_problems + 7) / 8 * 8;
  hipLaunchKernelGGL((local_solve_3d<T, NF, SZ, R>), dim3(grid), dim3(NF * NF * SZ), 0, st, a);
  return hipGetLastError();
}
template <typename T>
hipError_t launch3t(int nf, int variant, const LocalArgs &a, hipStream_t st)
{
  switch (nf)
    {
      case 8: return launch3<T, 8, 1, 7>(a, st);
      case 16: return variant == 1 ? launch3<T, 16, 1, 15>(a, st) : launch3<T, 16, 2, 8>(a, st);
      default: return hipErrorInvalidValue;
    }
}
} // namespace

int local3d_variants(int nf) { return nf == 16 ? 2 : 1; }

hipError_t launch_local_solve_3d(bool fp32, int nf, int variant, const LocalArgs &a, hipStream_t st)
{
  // ids: 16 -> 1 = (2 stacks of 8 z-levels), 2 = (1 stack of 15); 8 -> 1
  variant = variant == 2 ? 1 : 0;
  return fp32 ? launch3t<float>(nf, variant, a, st) : launch3t<double>(nf, variant, a, st);
}

} // namespace msfem
