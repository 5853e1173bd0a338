// Generic local basis solve: the constrained problems of
// DiffusionProblemBasis<dim>::run() (include/base/diffusion_problem_basis.tpp:440-474)
// for every case the tuned kernels do not take -- any number of local
// refinements (the reference refines n_refine_local times,
// diffusion_problem_basis.tpp:92-99: nf = 2^n_refine_local fine cells per
// side; here 2D nf <= 1024, 3D nf <= 64) and, in 3D, every coefficient kind
// (the reference's MatrixCoeff<3> at the QGauss(2) points,
// include/coefficients/matrix_coeff.tpp:28-41,66-91).
//
// prep: every fine cell's QGauss(2) element matrix (diffusion_problem_basis.tpp:
// 207-222; general quadrilaterals and hexahedra with the bi/trilinear mapping)
// into the class's element buffer, shared by the cell's 2^d bases.
// solve: one 256-thread workgroup per (cell, basis) problem; FP64
// Jacobi-preconditioned CG on the interior equations with the iterate in the
// bases block and r, p, q, D^-1 in per-problem scratch; the operator gathers,
// per node, the rows of the 2^d adjacent element matrices.  Same stopping rule
// and refinement structure as the tuned kernels: the true FP64 residual
// ||r|| <= max(abs_tol, rel_tol ||K_IB g||) (or the energy rule's per-problem
// tolerance), the CG of each cycle run to half of it.  Both precisions compute
// in FP64 here.
#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

constexpr int kGenThreads = 256;

// QGauss(2) element matrix of fine cell f of a 3D box tile (side lengths h),
// any coefficient kind; vertex v = x + 2 y + 4 z bits, quadrature points
// tensor-ordered with x fastest.
__device__ inline void qg_element_3d(int kind, const double *cf, int f, const double (&h)[3], double (&K)[8][8])
{
  for (int u = 0; u < 8; ++u)
    for (int v = 0; v < 8; ++v)
      K[u][v] = 0;
  const double gs = 0.5 / sqrt(3.0), jxw = 0.125 * h[0] * h[1] * h[2];
  for (int q = 0; q < 8; ++q)
    {
      double xi[3];
      for (int k = 0; k < 3; ++k)
        xi[k] = ((q >> k) & 1) ? 0.5 + gs : 0.5 - gs;
      double A[3][3];
      const int sidx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          switch (kind)
            {
              case 0: A[k][l] = k == l ? cf[f] : 0.0; break;
              case 1: A[k][l] = cf[6 * (size_t)f + sidx[k][l]]; break;
              case 2: A[k][l] = k == l ? cf[8 * (size_t)f + q] : 0.0; break;
              default: A[k][l] = cf[48 * (size_t)f + 6 * q + sidx[k][l]]; break;
            }
      double g[8][3];
      for (int v = 0; v < 8; ++v)
        for (int k = 0; k < 3; ++k)
          {
            double d = ((v >> k) & 1) ? 1.0 : -1.0;
            for (int m = 0; m < 3; ++m)
              if (m != k)
                d *= ((v >> m) & 1) ? xi[m] : 1.0 - xi[m];
            g[v][k] = d / h[k];
          }
      for (int u = 0; u < 8; ++u)
        for (int v = 0; v < 8; ++v)
          {
            double e = 0;
            for (int k = 0; k < 3; ++k)
              for (int l = 0; l < 3; ++l)
                e += g[u][k] * A[k][l] * g[v][l];
            K[u][v] += e * jxw;
          }
    }
}

constexpr int kGenPrepThreads = 256;

template <int DIM>
__global__ void __launch_bounds__(kGenPrepThreads) generic_prep_kernel(int n_cells_class, const int *cells, int nf, int kind, const double *box,
                                    const double *verts, const double *coef, const long long *coef_off, double *elem,
                                    const long long *elem_off)
{
  constexpr int NV = 1 << DIM;
  const long long per = DIM == 2 ? (long long)nf * nf : (long long)nf * nf * nf;
  const long long total = (long long)n_cells_class * per;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x)
    {
      const int cell = cells[t / per];
      const int f = (int)(t % per);
      const double *cf = coef + coef_off[cell];
      double *E = elem + elem_off[cell] + (size_t)f * NV * NV;
      if constexpr (DIM == 2)
        {
          double K[4][4];
          if (verts)
            qg_element_quad(kind, cf, f, nf, QuadMap(verts + (size_t)cell * 8), K, 0, 0.0, nullptr);
          else
            {
              const double *bx = box + (size_t)cell * 4;
              qg_element_2d(kind, cf, f, (bx[1] - bx[0]) / nf, (bx[3] - bx[2]) / nf, K);
            }
          for (int u = 0; u < 4; ++u)
            for (int v = 0; v < 4; ++v)
              E[u * 4 + v] = K[u][v];
        }
      else if (verts)
        {
          double K[64];
          qg_element_hex(kind, cf, f, nf, HexMap(verts + (size_t)cell * 24), K, 0, 0.0, nullptr);
          for (int uv = 0; uv < 64; ++uv)
            E[uv] = K[uv];
        }
      else
        {
          const double *bx = box + (size_t)cell * 6;
          const double h[3] = {(bx[1] - bx[0]) / nf, (bx[3] - bx[2]) / nf, (bx[5] - bx[4]) / nf};
          double K[8][8];
          qg_element_3d(kind, cf, f, h, K);
          for (int u = 0; u < 8; ++u)
            for (int v = 0; v < 8; ++v)
              E[u * 8 + v] = K[u][v];
        }
    }
}

// Fixed-order workgroup sum of NVAL <= 2 values (every thread gets the same totals).
template <int NVAL>
__device__ __forceinline__ void gen_block_sum(double (&v)[NVAL], double (*red)[2])
{
  static_assert(NVAL <= 2, "two reduction slots per wave");
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NVAL; ++k)
    v[k] = group_sum<kWave>(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NVAL; ++k)
      red[wave][k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NVAL; ++k)
    {
      double s = 0;
      for (int w = 0; w < kGenThreads / kWave; ++w)
        s += red[w][k];
      v[k] = s;
    }
  __syncthreads();
}

template <int DIM>
__global__ void __launch_bounds__(kGenThreads) generic_solve_kernel(LocalArgs a, int nf, double *work)
{
  constexpr int NV = 1 << DIM;
  __shared__ double red[kGenThreads / kWave][2];
  const int prob = blockIdx.x;
  if (prob >= a.n_problems)
    return; // uniform over the workgroup
  const int cell = a.cells[prob / a.nb_solve], basis = prob % a.nb_solve;
  const int np = nf + 1;
  const int nn = DIM == 2 ? np * np : np * np * np;
  const bool general = a.verts != nullptr;
  double *phi = a.bases + a.bas_off[cell] + (size_t)basis * nn;
  const double *E = a.cond + a.cond_off[cell];
  double *r = work + (size_t)prob * 4 * nn, *p = r + nn, *q = p + nn, *dinv = q + nn;

  auto coords = [&](int n, int (&I)[3]) {
    I[0] = n % np;
    I[1] = (n / np) % np;
    I[2] = DIM == 3 ? n / (np * np) : 0;
  };
  auto interior = [&](const int (&I)[3]) {
    for (int k = 0; k < DIM; ++k)
      if (I[k] == 0 || I[k] == nf)
        return false;
    return true;
  };
  // (K x)_n from the rows of the adjacent element matrices; bonly: interior values taken as 0
  auto apply = [&](const double *x, int n, bool bonly) -> double {
    int I[3];
    coords(n, I);
    double y = 0;
    for (int o = 0; o < NV; ++o) // the node is vertex o of fine cell I - o
      {
        int c[3] = {0, 0, 0};
        bool ok = true;
        for (int k = 0; k < DIM; ++k)
          {
            c[k] = I[k] - ((o >> k) & 1);
            ok = ok && c[k] >= 0 && c[k] < nf;
          }
        if (!ok)
          continue;
        const size_t f = DIM == 2 ? (size_t)c[1] * nf + c[0] : ((size_t)c[2] * nf + c[1]) * nf + c[0];
        const double *row = E + f * NV * NV + (size_t)o * NV;
        for (int b = 0; b < NV; ++b)
          {
            int J[3] = {c[0] + (b & 1), c[1] + ((b >> 1) & 1), DIM == 3 ? c[2] + ((b >> 2) & 1) : 0};
            const int m = J[0] + np * (J[1] + np * J[2]);
            if (bonly && interior(J))
              continue;
            y += row[b] * x[m];
          }
      }
    return y;
  };

  // boundary values (box cells: the nodal Q1 function; general cells: written by
  // the BasisQ1 kernel before), then the start of the interior
  if (!general)
    for (int n = threadIdx.x; n < nn; n += kGenThreads)
      {
        int I[3];
        coords(n, I);
        if (!interior(I))
          phi[n] = nodal_q1(basis, DIM, I[0], I[1], I[2], nf);
      }
  __syncthreads();
  for (int n = threadIdx.x; n < nn; n += kGenThreads)
    {
      int I[3];
      coords(n, I);
      if (!interior(I))
        {
          dinv[n] = 0.0;
          continue;
        }
      double v = 0.0;
      if (a.x0)
        v = a.x0[a.bas_off[cell] + (size_t)basis * nn + n];
      else if (a.q1_start && (!general || DIM == 3)) // (hexahedra: the parametric nodal Q1 function)
        v = nodal_q1(basis, DIM, I[0], I[1], I[2], nf);
      else if (a.q1_start) // transfinite interpolant of the boundary values over the parametric grid
        {
          const double t = (double)I[0] / nf, s = (double)I[1] / nf;
          const double bs = phi[I[0]], bn = phi[nf * np + I[0]], bw = phi[I[1] * np], be = phi[I[1] * np + nf];
          const double c00 = phi[0], c10 = phi[nf], c01 = phi[nf * np], c11 = phi[nf * np + nf];
          v = (1 - s) * bs + s * bn + (1 - t) * bw + t * be -
              ((1 - t) * (1 - s) * c00 + t * (1 - s) * c10 + (1 - t) * s * c01 + t * s * c11);
        }
      phi[n] = isfinite(v) ? v : 0.0;
      double d = 0;
      for (int o = 0; o < NV; ++o)
        {
          int c[3] = {0, 0, 0};
          for (int k = 0; k < DIM; ++k)
            c[k] = I[k] - ((o >> k) & 1);
          const size_t f = DIM == 2 ? (size_t)c[1] * nf + c[0] : ((size_t)c[2] * nf + c[1]) * nf + c[0];
          d += E[f * NV * NV + (size_t)o * NV + o];
        }
      dinv[n] = 1.0 / d;
    }
  __syncthreads();

  // ||b|| = ||K_IB g|| for the relative test
  double bb[1] = {0};
  for (int n = threadIdx.x; n < nn; n += kGenThreads)
    if (dinv[n] != 0.0)
      {
        const double b = apply(phi, n, true);
        bb[0] += b * b;
      }
  gen_block_sum<1>(bb, red);
  const double bnorm = sqrt(bb[0]);
  const double tol = a.ptol ? a.ptol[(size_t)cell * a.nb + basis] : fmax(a.abs_tol, a.rel_tol * bnorm);

  int it_total = 0, conv = 0;
  double rnorm = 0;
  for (int cyc = 0;; ++cyc)
    {
      // true FP64 residual r = -(K phi)_I
      double rr[1] = {0};
      for (int n = threadIdx.x; n < nn; n += kGenThreads)
        {
          const double v = dinv[n] != 0.0 ? -apply(phi, n, false) : 0.0;
          r[n] = v;
          rr[0] += v * v;
        }
      gen_block_sum<1>(rr, red);
      rnorm = sqrt(rr[0]);
      if (rnorm <= tol)
        {
          conv = 1;
          break;
        }
      if (it_total >= a.max_iter || cyc >= a.max_refine)
        break;
      // PCG on K d = r from d = 0, added into phi
      const double itol2 = 0.25 * tol * tol;
      double rz[1] = {0};
      for (int n = threadIdx.x; n < nn; n += kGenThreads)
        {
          const double z = dinv[n] * r[n];
          p[n] = z;
          rz[0] += r[n] * z;
        }
      gen_block_sum<1>(rz, red);
      double rzv = rz[0];
      while (it_total < a.max_iter)
        {
          double pq[1] = {0};
          for (int n = threadIdx.x; n < nn; n += kGenThreads)
            {
              const double v = dinv[n] != 0.0 ? apply(p, n, false) : 0.0;
              q[n] = v;
              pq[0] += p[n] * v;
            }
          gen_block_sum<1>(pq, red);
          const double alpha = rzv / pq[0];
          double s2[2] = {0, 0};
          for (int n = threadIdx.x; n < nn; n += kGenThreads)
            {
              if (dinv[n] == 0.0)
                continue;
              phi[n] += alpha * p[n];
              const double v = r[n] - alpha * q[n];
              r[n] = v;
              s2[0] += v * v;
              s2[1] += v * v * dinv[n];
            }
          gen_block_sum<2>(s2, red);
          ++it_total;
          if (s2[0] <= itol2 || !(s2[1] > 0))
            break;
          const double beta = s2[1] / rzv;
          rzv = s2[1];
          for (int n = threadIdx.x; n < nn; n += kGenThreads)
            p[n] = dinv[n] * r[n] + beta * p[n];
          __syncthreads();
        }
      __syncthreads();
    }
  if (threadIdx.x == 0)
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

// Dirichlet data of general hexahedra: BasisQ1 (include/coefficients/basis_q1.tpp:
// 26-46, :86-96) is N_i(x) = sum_m C[m][i] mono_m(x), mono = (1, x, y, z, xy,
// yz, xz, xyz), C = M^{-1}, M[j][m] = mono_m(v_j).  One workgroup per cell:
// one thread inverts M (Gauss-Jordan with partial pivoting, as deal.II's
// FullMatrix::invert) in LDS, then the workgroup evaluates N_i at the mapped
// boundary nodes F(I / nf) and writes them into every basis block.  Boundary
// node k: the planes z = 0 and z = nf whole (2 (nf+1)^2 nodes), then the ring
// of each inner plane z = 1 .. nf-1 (4 nf nodes each).
constexpr int kHexBcThreads = 256;

__global__ void __launch_bounds__(kHexBcThreads) hex_dirichlet_3d(const int *cells, int nf, const double *verts,
                                                                   double *bases, const long long *bas_off)
{
  __shared__ double a[8][16];
  const int cell = cells[blockIdx.x];
  const double *v = verts + (size_t)cell * 24;
  if (threadIdx.x == 0)
    {
      for (int r = 0; r < 8; ++r)
        {
          const double x = v[3 * r], y = v[3 * r + 1], z = v[3 * r + 2];
          const double mono[8] = {1.0, x, y, z, x * y, y * z, x * z, x * y * z};
          for (int c = 0; c < 8; ++c)
            {
              a[r][c] = mono[c];
              a[r][8 + c] = r == c ? 1.0 : 0.0;
            }
        }
      for (int c = 0; c < 8; ++c)
        {
          int piv = c;
          for (int r = c + 1; r < 8; ++r)
            if (fabs(a[r][c]) > fabs(a[piv][c]))
              piv = r;
          if (piv != c)
            for (int m = 0; m < 16; ++m)
              {
                const double t = a[c][m];
                a[c][m] = a[piv][m];
                a[piv][m] = t;
              }
          const double d = 1.0 / a[c][c];
          for (int m = 0; m < 16; ++m)
            a[c][m] *= d;
          for (int r = 0; r < 8; ++r)
            if (r != c)
              {
                const double fct = a[r][c];
                for (int m = 0; m < 16; ++m)
                  a[r][m] -= fct * a[c][m];
              }
        }
    }
  __syncthreads();
  const HexMap map(v);
  const int np = nf + 1;
  const size_t nn = (size_t)np * np * np;
  const int nplane = np * np, nring = 4 * nf;
  const int nbnd = 2 * nplane + (nf - 1) * nring;
  double *phi = bases + bas_off[cell];
  for (int k = threadIdx.x; k < nbnd; k += kHexBcThreads)
    {
      int X, Y, Z;
      if (k < 2 * nplane)
        {
          Z = k < nplane ? 0 : nf;
          X = (k % nplane) % np;
          Y = (k % nplane) / np;
        }
      else
        {
          const int kk = k - 2 * nplane;
          Z = 1 + kk / nring;
          const int side = kk % nring / nf, j = kk % nring % nf;
          X = side == 0 ? j : side == 1 ? nf : side == 2 ? nf - j : 0;
          Y = side == 0 ? 0 : side == 1 ? j : side == 2 ? nf : nf - j;
        }
      double x[3];
      map.point((double)X / nf, (double)Y / nf, (double)Z / nf, x);
      const double mono[8] = {1.0, x[0], x[1], x[2], x[0] * x[1], x[1] * x[2], x[0] * x[2], x[0] * x[1] * x[2]};
      const size_t node = (size_t)X + (size_t)np * (Y + (size_t)np * Z);
      for (int i = 0; i < 8; ++i)
        {
          double val = 0;
          for (int m = 0; m < 8; ++m)
            val += a[m][8 + i] * mono[m];
          phi[(size_t)i * nn + node] = val;
        }
    }
}

} // namespace

size_t generic_elem_doubles(int dim, int nf)
{
  const size_t nv = (size_t)1 << dim, nc = dim == 2 ? (size_t)nf * nf : (size_t)nf * nf * nf;
  return nc * nv * nv;
}

size_t generic_work_doubles(int dim, int nf)
{
  const size_t np = nf + 1;
  return 4 * (dim == 2 ? np * np : np * np * np);
}

hipError_t launch_generic_prep(int dim, int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *verts, const double *coef, const long long *coef_off, double *elem,
                               const long long *elem_off, hipStream_t st)
{
  const long long total = (long long)n_cells_class * (dim == 2 ? (long long)nf * nf : (long long)nf * nf * nf);
  if (total == 0)
    return hipSuccess;
  long long grid = (total + kGenPrepThreads - 1) / kGenPrepThreads;
  if (grid > 65536)
    grid = 65536;
  if (dim == 2)
    hipLaunchKernelGGL(generic_prep_kernel<2>, dim3((unsigned)grid), dim3(kGenPrepThreads), 0, st, n_cells_class,
                       cells, nf, coef_kind, box, verts, coef, coef_off, elem, elem_off);
  else
    hipLaunchKernelGGL(generic_prep_kernel<3>, dim3((unsigned)grid), dim3(kGenPrepThreads), 0, st, n_cells_class,
                       cells, nf, coef_kind, box, verts, coef, coef_off, elem, elem_off);
  return hipGetLastError();
}

hipError_t launch_hex_dirichlet_3d(int n_cells_class, const int *cells, int nf, const double *verts, double *bases,
                                  const long long *bas_off, hipStream_t st)
{
  if (n_cells_class == 0)
    return hipSuccess;
  hipLaunchKernelGGL(hex_dirichlet_3d, dim3(n_cells_class), dim3(kHexBcThreads), 0, st, cells, nf, verts, bases,
                     bas_off);
  return hipGetLastError();
}

hipError_t launch_generic_solve(int dim, int nf, const LocalArgs &a, double *work, hipStream_t st)
{
  if (a.n_problems == 0)
    return hipSuccess;
  if (dim == 2)
    hipLaunchKernelGGL(generic_solve_kernel<2>, dim3(a.n_problems), dim3(kGenThreads), 0, st, a, nf, work);
  else
    hipLaunchKernelGGL(generic_solve_kernel<3>, dim3(a.n_problems), dim3(kGenThreads), 0, st, a, nf, work);
  return hipGetLastError();
}

} // namespace msfem
