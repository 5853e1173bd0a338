// Coarse system of the MsFEM on the unit square/cube with nc cells per side:
//   assembly     DiffusionProblemMultiscale::assemble_system
//                (include/base/diffusion_problem_ms.tpp:161-255; Dirichlet
//                constraints :128-137, Neumann face terms :206-235)
//   solve        DiffusionProblemMultiscale::solve_iterative (:260-296): FP64 CG
//   weights      send_global_weights_to_cell (:301-325) and the fine
//                reconstruction of DiffusionProblemBasis::set_global_weights
//                (include/base/diffusion_problem_basis.tpp:354-377)
// The coarse matrix is kept in structured stencil form Kc[node][3^dim].
#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

constexpr double kPiD = 3.14592653509793218403; // PI_D of include/coefficients/coefficients.h:21

__device__ __forceinline__ double dirichlet_g(double x, double y) { return (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5); }
__device__ __forceinline__ double neumann_h(double x, double y) { return cos(2 * kPiD * x) * cos(2 * kPiD * y); }

// One thread per coarse node: gathers A^K, F^K of the <= 2^dim adjacent cells
// in a fixed order (deterministic), adds Neumann face integrals (QGauss(2) on
// faces) and marks the Dirichlet nodes.
template <int DIM>
__global__ void coarse_assemble_kernel(int nc, const double *Ae, const double *Fe, int bc_kind, double *Kc,
                                       double *Fc, unsigned char *cmask, double *cval)
{
  constexpr int NV = 1 << DIM, NS3 = DIM == 2 ? 9 : 27;
  const int np = nc + 1;
  const int nn = DIM == 2 ? np * np : np * np * np;
  const double H = 1.0 / nc, gs = 0.5 / sqrt(3.0);
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      const int I[3] = {n % np, (n / np) % np, DIM == 3 ? n / (np * np) : 0};
      double k[NS3];
#pragma unroll
      for (int t = 0; t < NS3; ++t)
        k[t] = 0;
      double f = 0;
      for (int a = 0; a < NV; ++a) // the node is vertex a of cell I - offset(a)
        {
          int ci[3];
          bool ok = true;
          for (int m = 0; m < DIM; ++m)
            {
              ci[m] = I[m] - ((a >> m) & 1);
              ok = ok && ci[m] >= 0 && ci[m] < nc;
            }
          if (!ok)
            continue;
          const size_t c = DIM == 2 ? (size_t)ci[0] + (size_t)nc * ci[1]
                                    : (size_t)ci[0] + (size_t)nc * (ci[1] + (size_t)nc * ci[2]);
          for (int b = 0; b < NV; ++b)
            {
              int t = 0, mul = 1;
              for (int m = 0; m < DIM; ++m)
                {
                  t += (((b >> m) & 1) - ((a >> m) & 1) + 1) * mul;
                  mul *= 3;
                }
              k[t] += Ae[c * NV * NV + a * NV + b];
            }
          f += Fe[c * NV + a];
          if (bc_kind == 1)
            for (int m = 0; m < DIM; ++m) // face x_m = 1 (boundary ids 1, 3, 5)
              {
                if (ci[m] != nc - 1 || !((a >> m) & 1))
                  continue;
                for (int q = 0; q < (1 << (DIM - 1)); ++q)
                  {
                    double xq[3], w = 1.0, psi = 1.0;
                    int tq = 0;
                    for (int mm = 0; mm < DIM; ++mm)
                      {
                        if (mm == m)
                          {
                            xq[mm] = 1.0;
                            continue;
                          }
                        const double xi = ((q >> tq) & 1) ? 0.5 + gs : 0.5 - gs;
                        ++tq;
                        xq[mm] = (ci[mm] + xi) * H;
                        w *= 0.5 * H;
                        psi *= ((a >> mm) & 1) ? xi : 1.0 - xi;
                      }
                    f += neumann_h(xq[0], xq[1]) * psi * w;
                  }
              }
        }
#pragma unroll
      for (int t = 0; t < NS3; ++t)
        Kc[(size_t)n * NS3 + t] = k[t];
      Fc[n] = f;
      bool con = false;
      for (int m = 0; m < DIM; ++m)
        con = con || I[m] == 0 || (bc_kind == 0 && I[m] == nc);
      cmask[n] = con ? 1 : 0;
      cval[n] = (con && bc_kind == 1) ? dirichlet_g(I[0] * H, I[1] * H) : 0.0;
    }
}

constexpr int kCoarseThreads = 1024;

__device__ __forceinline__ double block_sum(double v, double *red)
{
  v = group_sum<kWave>(v);
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0)
    red[wave] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < kCoarseThreads / kWave; ++w)
    s += red[w];
  return s;
}

// Jacobi-preconditioned FP64 CG on the free coarse nodes, one workgroup;
// vectors in global memory (work: x, r, p, q, dinv).
template <int DIM>
__global__ void __launch_bounds__(kCoarseThreads) coarse_cg_kernel(int nc, const double *Kc, const double *Fc,
                                                                    const unsigned char *cmask, const double *cval,
                                                                    double rel_tol, int max_iter, double *u,
                                                                    int *iters, double *work)
{
  constexpr int NS3 = DIM == 2 ? 9 : 27;
  __shared__ double red[kCoarseThreads / kWave];
  const int np = nc + 1;
  const int nn = DIM == 2 ? np * np : np * np * np;
  double *x = work, *r = work + nn, *p = work + 2 * (size_t)nn, *q = work + 3 * (size_t)nn,
         *dinv = work + 4 * (size_t)nn;

  auto nbr = [&](int n, int t) -> int {
    const int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = DIM == 3 ? t / 9 - 1 : 0;
    const int X = n % np + ox, Y = (n / np) % np + oy, Z = (DIM == 3 ? n / (np * np) : 0) + oz;
    if (X < 0 || X >= np || Y < 0 || Y >= np || Z < 0 || Z >= (DIM == 3 ? np : 1))
      return -1;
    return X + np * (Y + np * Z);
  };
  auto matvec = [&](const double *v, int n) -> double { // (K v)_n over all neighbours
    double s = 0;
    for (int t = 0; t < NS3; ++t)
      {
        const int m = nbr(n, t);
        if (m >= 0)
          s += Kc[(size_t)n * NS3 + t] * v[m];
      }
    return s;
  };

  for (int n = threadIdx.x; n < nn; n += blockDim.x)
    {
      x[n] = cmask[n] ? cval[n] : 0.0;
      p[n] = 0.0;
      dinv[n] = cmask[n] ? 0.0 : 1.0 / Kc[(size_t)n * NS3 + NS3 / 2];
    }
  __syncthreads();
  double bb = 0, rz = 0;
  for (int n = threadIdx.x; n < nn; n += blockDim.x)
    {
      const double rv = cmask[n] ? 0.0 : Fc[n] - matvec(x, n);
      r[n] = rv;
      bb += rv * rv;
      const double z = rv * dinv[n];
      p[n] = z;
      rz += rv * z;
    }
  bb = block_sum(bb, red);
  rz = block_sum(rz, red);
  const double tol = rel_tol * sqrt(bb);
  int it = 0;
  double rn = sqrt(bb);
  while (rn > tol && it < max_iter)
    {
      double pq = 0;
      for (int n = threadIdx.x; n < nn; n += blockDim.x)
        {
          const double qv = cmask[n] ? 0.0 : matvec(p, n);
          q[n] = qv;
          pq += p[n] * qv;
        }
      pq = block_sum(pq, red);
      const double alpha = rz / pq;
      double rz_new = 0, rr = 0;
      for (int n = threadIdx.x; n < nn; n += blockDim.x)
        {
          x[n] += alpha * p[n];
          const double rv = r[n] - alpha * q[n];
          r[n] = rv;
          rz_new += rv * rv * dinv[n];
          rr += rv * rv;
        }
      rz_new = block_sum(rz_new, red);
      rr = block_sum(rr, red);
      const double beta = rz_new / rz;
      rz = rz_new;
      for (int n = threadIdx.x; n < nn; n += blockDim.x)
        p[n] = r[n] * dinv[n] + beta * p[n];
      rn = sqrt(rr);
      ++it;
      __syncthreads();
    }
  for (int n = threadIdx.x; n < nn; n += blockDim.x)
    u[n] = x[n];
  if (threadIdx.x == 0)
    *iters = it;
}

__global__ void reconstruct_kernel(int n_cells, int nb, const long long *bas_off, const long long *out_off,
                                   const int *n_nodes, const double *bases, const double *w, double *u)
{
  const int cell = blockIdx.y;
  if (cell >= n_cells)
    return;
  const int nn = n_nodes[cell];
  const double *phi = bases + bas_off[cell];
  double wi[8];
  for (int i = 0; i < nb; ++i)
    wi[i] = w[(size_t)cell * nb + i];
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < nn; n += gridDim.x * blockDim.x)
    {
      double s = 0;
      for (int i = 0; i < nb; ++i)
        s += wi[i] * phi[(size_t)i * nn + n];
      u[out_off[cell] + n] = s;
    }
}

} // namespace

hipError_t launch_coarse_assemble(int dim, int nc, const double *Ae, const double *Fe, int bc_kind, double *Kc,
                                  double *Fc, unsigned char *cmask, double *cval, hipStream_t st)
{
  const int nn = dim == 2 ? (nc + 1) * (nc + 1) : (nc + 1) * (nc + 1) * (nc + 1);
  const int grid = (nn + 255) / 256;
  if (dim == 2)
    hipLaunchKernelGGL(coarse_assemble_kernel<2>, dim3(grid), dim3(256), 0, st, nc, Ae, Fe, bc_kind, Kc, Fc, cmask,
                       cval);
  else
    hipLaunchKernelGGL(coarse_assemble_kernel<3>, dim3(grid), dim3(256), 0, st, nc, Ae, Fe, bc_kind, Kc, Fc, cmask,
                       cval);
  return hipGetLastError();
}

size_t coarse_solve_work_doubles(int dim, int nc)
{
  const size_t np = nc + 1;
  return 5 * (dim == 2 ? np * np : np * np * np);
}

hipError_t launch_coarse_solve(int dim, int nc, const double *Kc, const double *Fc, const unsigned char *cmask,
                               const double *cval, double rel_tol, int max_iter, double *u, int *iters,
                               double *work, hipStream_t st)
{
  if (dim == 2)
    hipLaunchKernelGGL(coarse_cg_kernel<2>, dim3(1), dim3(kCoarseThreads), 0, st, nc, Kc, Fc, cmask, cval, rel_tol,
                       max_iter, u, iters, work);
  else
    hipLaunchKernelGGL(coarse_cg_kernel<3>, dim3(1), dim3(kCoarseThreads), 0, st, nc, Kc, Fc, cmask, cval, rel_tol,
                       max_iter, u, iters, work);
  return hipGetLastError();
}

hipError_t launch_reconstruct(int n_cells, int nb, const long long *bas_off, const long long *out_off,
                              const int *n_nodes, const double *bases, const double *w, double *u, hipStream_t st)
{
  if (n_cells == 0)
    return hipSuccess;
  hipLaunchKernelGGL(reconstruct_kernel, dim3(16, n_cells), dim3(256), 0, st, n_cells, nb, bas_off, out_off, n_nodes,
                     bases, w, u);
  return hipGetLastError();
}

} // namespace msfem
