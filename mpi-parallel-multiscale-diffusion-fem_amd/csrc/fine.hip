// Fine-scale standard FEM on the unit square: the reference's
// DiffusionProblem<dim> (include/base/diffusion_problem.tpp:40-279), the
// single-scale Q1 solve on the fine mesh that the MsFEM approximates.
//   * element matrices and loads of every fine cell (QGauss(2), :150-183);
//   * assembly with the reference's boundary data: the coarse assembly kernel
//     applied to the fine cells (Dirichlet g on x_k = 0, Neumann h on x_k = 1);
//   * solve: FP64 Jacobi-PCG over many workgroups on the structured stencil
//     matrix (the reference: CG + AMG, :243-279), one launch per CG step with
//     the scalars in device memory, so that a block of iterations can be
//     captured once as a HIP graph and replayed until convergence.
// The structured CG also serves large coarse systems.
#include "msfem.h"
#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

constexpr int kFineThreads = 256;

// QGauss(2) element matrix and load vector of fine cell f of an n x n mesh of
// the unit square (coefficient tile in the batch layout of one cell).
__global__ void fine_elements_2d_kernel(int n, int coef_kind, const double *coef, int rhs_kind, double rhs_value,
                                        double *Ae, double *Fe)
{
  const double h = 1.0 / n, gs = 0.5 / sqrt(3.0), jxw = 0.25 * h * h;
  for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < n * n; f += gridDim.x * blockDim.x)
    {
      double K[4][4];
      qg_element_2d(coef_kind, coef, f, h, h, K);
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v)
          Ae[(size_t)f * 16 + u * 4 + v] = K[u][v];
      const int fx = f % n, fy = f / n;
      double b[4] = {0, 0, 0, 0};
      for (int q = 0; q < 4; ++q)
        {
          const double xi = (q & 1) ? 0.5 + gs : 0.5 - gs, eta = (q & 2) ? 0.5 + gs : 0.5 - gs;
          double fv = rhs_value;
          if (rhs_kind == MSFEM_RHS_SINSIN)
            fv = sin(M_PI * (fx + xi) * h) * sin(M_PI * (fy + eta) * h);
          for (int v = 0; v < 4; ++v)
            b[v] += ((v & 1) ? xi : 1.0 - xi) * ((v & 2) ? eta : 1.0 - eta) * fv * jxw;
        }
      for (int v = 0; v < 4; ++v)
        Fe[(size_t)f * 4 + v] = b[v];
    }
}

// Block sum of one value in a fixed order (deterministic); thread 0 gets it.
__device__ __forceinline__ double block_sum_fixed(double v, double *red)
{
  v = group_sum<kWave>(v);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  if (lane == 0)
    red[wave] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / kWave); ++w)
      s += red[w];
  __syncthreads();
  return s;
}

template <int DIM>
__device__ __forceinline__ double stencil_row(int np, const double *Kc, const double *v, int n)
{
  constexpr int NS3 = DIM == 2 ? 9 : 27;
  const int X = n % np, Y = (n / np) % np, Z = DIM == 3 ? n / (np * np) : 0;
  double s = 0;
#pragma unroll
  for (int t = 0; t < NS3; ++t)
    {
      const int x = X + t % 3 - 1, y = Y + (t / 3) % 3 - 1, z = Z + (DIM == 3 ? t / 9 - 1 : 0);
      if (x < 0 || x >= np || y < 0 || y >= np || z < 0 || z >= (DIM == 3 ? np : 1))
        continue;
      s += Kc[(size_t)n * NS3 + t] * v[x + np * (y + np * z)];
    }
  return s;
}

// x = Dirichlet values / 0, diagonal inverse, p = 0.
__global__ void scg_init_x(int nn, int ns3, const double *Kc, const unsigned char *cmask, const double *cval,
                           double *x, double *dinv, double *p)
{
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      x[n] = cmask[n] ? cval[n] : 0.0;
      dinv[n] = cmask[n] ? 0.0 : 1.0 / Kc[(size_t)n * ns3 + ns3 / 2];
      p[n] = 0.0;
    }
}

// r = F - K x on free nodes, z = D^-1 r, p = z; partials (r.z, r.r).
template <int DIM>
__global__ void __launch_bounds__(kFineThreads) scg_init_r(int np, int nn, const double *Kc, const double *Fc,
                                                            const unsigned char *cmask, const double *x,
                                                            const double *dinv, double *r, double *p, double *part)
{
  __shared__ double red[kFineThreads / kWave];
  double rz = 0, rr = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      const double rv = cmask[n] ? 0.0 : Fc[n] - stencil_row<DIM>(np, Kc, x, n);
      r[n] = rv;
      const double z = rv * dinv[n];
      p[n] = z;
      rz += rv * z;
      rr += rv * rv;
    }
  rz = block_sum_fixed(rz, red);
  rr = block_sum_fixed(rr, red);
  if (threadIdx.x == 0)
    {
      part[2 * blockIdx.x] = rz;
      part[2 * blockIdx.x + 1] = rr;
    }
}

// q = K p on free nodes; partial p.q.
template <int DIM>
__global__ void __launch_bounds__(kFineThreads) scg_spmv(int np, int nn, const double *Kc, const unsigned char *cmask,
                                                          const double *p, double *q, double *part,
                                                          const FineCgState *st)
{
  __shared__ double red[kFineThreads / kWave];
  if (st->done)
    return;
  double pq = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      const double qv = cmask[n] ? 0.0 : stencil_row<DIM>(np, Kc, p, n);
      q[n] = qv;
      pq += p[n] * qv;
    }
  pq = block_sum_fixed(pq, red);
  if (threadIdx.x == 0)
    part[2 * blockIdx.x] = pq;
}

// x += alpha p, r -= alpha q, partials (r.z, r.r) of the new residual.
__global__ void __launch_bounds__(kFineThreads) scg_update(int nn, const double *dinv, const double *p, const double *q,
                                                            double *x, double *r, double *part, const FineCgState *st)
{
  __shared__ double red[kFineThreads / kWave];
  if (st->done)
    return;
  const double alpha = st->alpha;
  double rz = 0, rr = 0;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      x[n] += alpha * p[n];
      const double rv = r[n] - alpha * q[n];
      r[n] = rv;
      rz += rv * rv * dinv[n];
      rr += rv * rv;
    }
  rz = block_sum_fixed(rz, red);
  rr = block_sum_fixed(rr, red);
  if (threadIdx.x == 0)
    {
      part[2 * blockIdx.x] = rz;
      part[2 * blockIdx.x + 1] = rr;
    }
}

// p = D^-1 r + beta p
__global__ void scg_direction(int nn, const double *dinv, const double *r, double *p, const FineCgState *st)
{
  if (st->done)
    return;
  const double beta = st->beta;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    p[n] = r[n] * dinv[n] + beta * p[n];
}

// One workgroup: sums the per-block partials in a fixed order and advances the
// CG scalars.  mode 0: start (rz, ||r0|| -> tolerance), 1: alpha, 2: beta and
// the convergence test.
__global__ void __launch_bounds__(kFineThreads) scg_scalars(int nblk, const double *part, int mode, FineCgState *st)
{
  __shared__ double red[kFineThreads / kWave];
  if (mode != 0 && st->done)
    return;
  double a = 0, b = 0;
  for (int k = threadIdx.x; k < nblk; k += blockDim.x)
    {
      a += part[2 * k];
      b += part[2 * k + 1];
    }
  a = block_sum_fixed(a, red);
  b = block_sum_fixed(b, red);
  if (threadIdx.x != 0)
    return;
  if (mode == 0)
    {
      st->rz = a;
      st->rnorm = sqrt(b);
      st->tol = st->rel_tol * st->rnorm;
      st->iters = 0;
      st->converged = st->rnorm <= st->tol ? 1 : 0;
      st->done = (st->converged || st->max_iter <= 0) ? 1 : 0;
    }
  else if (mode == 1)
    st->alpha = st->rz / a;
  else
    {
      st->iters += 1;
      st->rnorm = sqrt(b);
      st->beta = a / st->rz;
      st->rz = a;
      if (st->rnorm <= st->tol)
        {
          st->converged = 1;
          st->done = 1;
        }
      else if (st->iters >= st->max_iter)
        st->done = 1;
    }
}

} // namespace

int fine_cg_blocks(int nn)
{
  int b = (nn + kFineThreads - 1) / kFineThreads;
  return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}

hipError_t launch_fine_elements_2d(int n, int coef_kind, const double *coef, int rhs_kind, double rhs_value,
                                   double *Ae, double *Fe, hipStream_t st)
{
  int grid = (n * n + kFineThreads - 1) / kFineThreads;
  if (grid > 65535)
    grid = 65535;
  hipLaunchKernelGGL(fine_elements_2d_kernel, dim3(grid), dim3(kFineThreads), 0, st, n, coef_kind, coef, rhs_kind,
                     rhs_value, Ae, Fe);
  return hipGetLastError();
}

hipError_t launch_fine_cg_start(const FineCgArgs &a, hipStream_t st)
{
  const int ns3 = a.dim == 2 ? 9 : 27, nblk = fine_cg_blocks(a.nn);
  hipLaunchKernelGGL(scg_init_x, dim3(nblk), dim3(kFineThreads), 0, st, a.nn, ns3, a.Kc, a.cmask, a.cval, a.x,
                     a.dinv, a.p);
  if (a.dim == 2)
    hipLaunchKernelGGL(scg_init_r<2>, dim3(nblk), dim3(kFineThreads), 0, st, a.np, a.nn, a.Kc, a.Fc, a.cmask, a.x,
                       a.dinv, a.r, a.p, a.part);
  else
    hipLaunchKernelGGL(scg_init_r<3>, dim3(nblk), dim3(kFineThreads), 0, st, a.np, a.nn, a.Kc, a.Fc, a.cmask, a.x,
                       a.dinv, a.r, a.p, a.part);
  hipLaunchKernelGGL(scg_scalars, dim3(1), dim3(kFineThreads), 0, st, nblk, a.part, 0, a.state);
  return hipGetLastError();
}

hipError_t launch_fine_cg_steps(const FineCgArgs &a, int steps, hipStream_t st)
{
  const int nblk = fine_cg_blocks(a.nn);
  for (int k = 0; k < steps; ++k)
    {
      if (a.dim == 2)
        hipLaunchKernelGGL(scg_spmv<2>, dim3(nblk), dim3(kFineThreads), 0, st, a.np, a.nn, a.Kc, a.cmask, a.p, a.q,
                           a.part, a.state);
      else
        hipLaunchKernelGGL(scg_spmv<3>, dim3(nblk), dim3(kFineThreads), 0, st, a.np, a.nn, a.Kc, a.cmask, a.p, a.q,
                           a.part, a.state);
      hipLaunchKernelGGL(scg_scalars, dim3(1), dim3(kFineThreads), 0, st, nblk, a.part, 1, a.state);
      hipLaunchKernelGGL(scg_update, dim3(nblk), dim3(kFineThreads), 0, st, a.nn, a.dinv, a.p, a.q, a.x, a.r, a.part,
                         a.state);
      hipLaunchKernelGGL(scg_scalars, dim3(1), dim3(kFineThreads), 0, st, nblk, a.part, 2, a.state);
      hipLaunchKernelGGL(scg_direction, dim3(nblk), dim3(kFineThreads), 0, st, a.nn, a.dinv, a.r, a.p, a.state);
    }
  return hipGetLastError();
}

} // namespace msfem
