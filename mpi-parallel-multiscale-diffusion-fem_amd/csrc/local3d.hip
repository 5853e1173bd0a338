// 3D local basis solves (DiffusionProblemBasis<3>, include/base/diffusion_problem_basis.tpp:440-474):
// K_II x = -K_IB g_i for the 8 trilinear nodal functions g_i, scalar
// coefficient per fine cell, matrix-free 27-point Q1 operator.
//
// One workgroup owns one problem.  Thread t owns the node column
// (X, Y) = (1 + t % NF, 1 + (t / NF) % NF) and R consecutive z-levels of
// stack t / NF^2; the search direction of the cell lives in a zero-padded LDS
// grid, the coefficients of the 4 x (R+1) fine cells around the column in
// registers.  Same FP64 refinement structure as the 2D kernel.
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

template <int NW>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red)
{
  a = group_sum<kWave>(a);
  b = group_sum<kWave>(b);
  if constexpr (NW > 1)
    {
      const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
      if (lane == 0)
        {
          red[2 * wave] = a;
          red[2 * wave + 1] = b;
        }
      __syncthreads();
      a = 0;
      b = 0;
#pragma unroll
      for (int w = 0; w < NW; ++w)
        {
          a += red[2 * w];
          b += red[2 * w + 1];
        }
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
  __shared__ double sred[3][2 * NW];

  const int tid = threadIdx.x;
  const int prob = blockIdx.x;
  const int cell = a.cells[prob / a.nb_solve];
  const int basis = prob % a.nb_solve;
  const int X = 1 + tid % NF, Y = 1 + (tid / NF) % NF, Z0 = 1 + (tid / (NF * NF)) * R;
  const bool colact = X <= NF - 1 && Y <= NF - 1;
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
          cc[i][k][j] = (cx < NF && cy < NF && cz < NF) ? (T)cf[(cz * NF + cy) * NF + cx] : (T)0;
        }
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
  for (int k = tid; k < NP * NP * NP; k += NT)
    {
      const int XX = k % NP, YY = (k / NP) % NP, ZZ = k / (NP * NP);
      const bool bnd = XX == 0 || XX == NF || YY == 0 || YY == NF || ZZ == 0 || ZZ == NF;
      double v = bnd ? nodal_q1(basis, 3, XX, YY, ZZ, NF)
                     : (a.x0 ? a.x0[a.bas_off[cell] + (size_t)basis * NP * NP * NP + k] : 0.0);
      phi[k] = isfinite(v) ? v : 0.0; // a non-finite initial value starts from 0
    }
  __syncthreads();

  // (K phi) at node (X, Y, Z) in FP64 from global memory
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

  double tol_eff = 0.0, bnorm = 0.0;
  if (a.x0)
    {
      double s = 0, dummy = 0;
#pragma unroll 1
      for (int i = 0; i < R; ++i)
        if (colact && Z0 + i <= NF - 1)
          {
            const double b = apply64(Z0 + i, true);
            s += b * b;
          }
      block_sum2<NW>(s, dummy, sred[2]);
      bnorm = sqrt(s);
      tol_eff = fmax(a.abs_tol, a.rel_tol * bnorm);
      __syncthreads();
    }

  T xr[R], rr[R], pr[R], q[R];
  int it_total = 0, state = 0, conv = 0;
  double rnorm = 0;
  for (int cyc = 0;; ++cyc)
    {
      double s = 0, dummy = 0;
#pragma unroll 1
      for (int i = 0; i < R; ++i)
        if (colact && Z0 + i <= NF - 1)
          {
            const double r64 = -apply64(Z0 + i, false);
            gown[i * GXY] = (T)r64;
            s += r64 * r64;
          }
#pragma unroll
      for (int i = 0; i < R; ++i)
        rr[i] = dinv[i] != (T)0 ? gown[i * GXY] : (T)0;
      block_sum2<NW>(s, dummy, sred[2]);
      rnorm = sqrt(s);
      if (cyc == 0 && !a.x0)
        {
          bnorm = rnorm;
          tol_eff = fmax(a.abs_tol, a.rel_tol * bnorm);
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
      double rz = 0, t0 = 0;
#pragma unroll
      for (int i = 0; i < R; ++i)
        {
          xr[i] = 0;
          const T z = rr[i] * dinv[i];
          pr[i] = z;
          rz += (double)(rr[i] * z);
          gown[i * GXY] = z;
        }
      block_sum2<NW>(rz, t0, sred[1]);
      __syncthreads();
      while (true)
        {
          T w[3][3][3]; // [z][y][x] window around the current node
#pragma unroll
          for (int zz = 0; zz < 2; ++zz)
#pragma unroll
            for (int yy = 0; yy < 3; ++yy)
#pragma unroll
              for (int xx = 0; xx < 3; ++xx)
                w[zz][yy][xx] = gwin[zz * GXY + yy * GX + xx];
          double pq = 0;
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
                      if (b == av)
                        continue;
                      const int ox = dx + (b & 1), oy = dy + ((b >> 1) & 1), oz = dz + ((b >> 2) & 1);
                      s2 += uc[av ^ b] * w[1 + oz][1 + oy][1 + ox];
                    }
                  y -= al * s2;
                  sa += al;
                }
              y += ctot * sa * pc;
              q[i] = dinv[i] != (T)0 ? y : (T)0;
              pq += (double)(pr[i] * q[i]);
#pragma unroll
              for (int yy = 0; yy < 3; ++yy)
#pragma unroll
                for (int xx = 0; xx < 3; ++xx)
                  {
                    w[0][yy][xx] = w[1][yy][xx];
                    w[1][yy][xx] = w[2][yy][xx];
                  }
            }
          block_sum2<NW>(pq, t0, sred[0]);
          const T alpha = (T)(rz / pq);
          double rz_new = 0, rn2 = 0;
#pragma unroll
          for (int i = 0; i < R; ++i)
            {
              xr[i] += alpha * pr[i];
              rr[i] -= alpha * q[i];
              const T z = rr[i] * dinv[i];
              rz_new += (double)(rr[i] * z);
              rn2 += (double)(rr[i] * rr[i]);
              q[i] = z;
            }
          block_sum2<NW>(rz_new, rn2, sred[1]);
          ++it_total;
          if (sqrt(rn2) <= inner_tol || it_total >= a.max_iter)
            break; // uniform: one problem per workgroup
          const T beta = (T)(rz_new / rz);
          rz = rz_new;
#pragma unroll
          for (int i = 0; i < R; ++i)
            {
              pr[i] = q[i] + beta * pr[i];
              gown[i * GXY] = pr[i];
            }
          __syncthreads();
        }
#pragma unroll
      for (int i = 0; i < R; ++i)
        if (dinv[i] != (T)0)
          phi[((Z0 + i) * NP + Y) * NP + X] += (double)xr[i];
      __syncthreads();
    }
  if (tid == 0)
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
  hipLaunchKernelGGL((local_solve_3d<T, NF, SZ, R>), dim3(a.n_problems), dim3(NF * NF * SZ), 0, st, a);
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
