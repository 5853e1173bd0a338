// Energy-norm control of the local solves (no counterpart in the reference,
// whose local solves stop at an absolute residual of 1e-12,
// include/base/diffusion_problem_basis.tpp:297).
//
// For basis i of a box cell with interior residual r of the computed phi:
//   ||phi - phi*||_K^2 = r^T K_II^-1 r <= ||r||^2 / lam_lb,
//   lam_lb = (min over the tile of lambda_min(A)) * mu_min   (K(A) >= a K(I)),
//   mu_min = smallest eigenvalue of the unit-coefficient Q1 stiffness on the
//            interior nodes: sum_a k_a prod_{b != a} m_b over the separable
//            modes sin(j pi x), k = (2 - 2 cos t)/h, m = h (4 + 2 cos t)/6,
//   ||phi*||_K^2 >= e_lb = e_unit^2 / D_i   (complementary energy with the
//            divergence-free flux t grad N_i: N_i is multilinear, so harmonic),
//   e_unit = int |grad N_i|^2,  D_i = int grad N_i^T A^-1 grad N_i,
// where the discrete energy of a Q1 function equals the exact integral for
// coefficients constant per fine cell (QGauss(2) is exact there) and bounds the
// integral with A replaced by its smallest eigenvalue per cell for the
// quadrature-point kinds.  energy_terms_kernel computes lam_lb per cell and,
// for the stopping rule, ptol = tau sqrt(lam_lb e_lb) per problem;
// energy_bound_kernel turns the final residual norms and A^K_ii = ||phi||_K^2
// into the a posteriori bound sqrt(q / (A^K_ii - q)), q = ||r||^2 / lam_lb.
#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

constexpr int kEnergyThreads = 256;

// Smallest eigenvalue of a symmetric 3x3 matrix (trigonometric closed form of
// the characteristic cubic), shrunk by 1e-10 relative so that its rounding
// cannot lift it above the true value.
__device__ inline double sym3_lambda_min(double a11, double a12, double a13, double a22, double a23, double a33)
{
  const double p1 = a12 * a12 + a13 * a13 + a23 * a23;
  const double q = (a11 + a22 + a33) / 3.0;
  const double b11 = a11 - q, b22 = a22 - q, b33 = a33 - q;
  const double p2 = b11 * b11 + b22 * b22 + b33 * b33 + 2.0 * p1;
  if (!(p2 > 0))
    return q;
  const double p = sqrt(p2 / 6.0), ip = 1.0 / p;
  const double c11 = b11 * ip, c22 = b22 * ip, c33 = b33 * ip, c12 = a12 * ip, c13 = a13 * ip, c23 = a23 * ip;
  const double det = c11 * (c22 * c33 - c23 * c23) - c12 * (c12 * c33 - c23 * c13) + c13 * (c12 * c23 - c22 * c13);
  const double r = fmin(1.0, fmax(-1.0, 0.5 * det));
  const double phi = acos(r) / 3.0;
  const double lmin = q + 2.0 * p * cos(phi + 2.0 * M_PI / 3.0);
  return lmin - 1e-10 * fabs(q) - 1e-10 * p;
}

// lambda_min of fine cell f and the entries of A^-1 (2D: b11, b12, b22; 3D:
// the scalar 1/lambda on the diagonal, an upper bound of A^-1) used in D_i.
template <int DIM>
__device__ __forceinline__ double cell_lambda(int kind, const double *cf, int f, double (&B)[3])
{
  constexpr int NQ = 1 << DIM, NS = DIM == 2 ? 3 : 6;
  double lam = 0;
  if (kind == 0) // scalar per fine cell
    lam = cf[f];
  else if (kind == 2) // scalar at the quadrature points: their minimum
    {
      lam = cf[(size_t)f * NQ];
      for (int q = 1; q < NQ; ++q)
        lam = fmin(lam, cf[(size_t)f * NQ + q]);
    }
  else if (DIM == 2 && kind == 1) // SPD tensor per fine cell: exact inverse
    {
      const double a = cf[3 * f], b = cf[3 * f + 1], c = cf[3 * f + 2];
      const double det = a * c - b * b, m = 0.5 * (a + c), d = sqrt(0.25 * (a - c) * (a - c) + b * b);
      lam = det / (m + d); // the smaller root, without cancellation
      B[0] = c / det;
      B[1] = -b / det;
      B[2] = a / det;
      return lam;
    }
  else if (DIM == 2 && kind == 3) // tensor at the quadrature points: smallest eigenvalue
    {
      lam = 1e300;
      for (int q = 0; q < NQ; ++q)
        {
          const double *t = cf + ((size_t)f * NQ + q) * NS;
          const double a = t[0], b = t[1], c = t[2];
          const double det = a * c - b * b, m = 0.5 * (a + c), d = sqrt(0.25 * (a - c) * (a - c) + b * b);
          lam = fmin(lam, det / (m + d));
        }
    }
  else if (DIM == 3 && kind == 1) // tensor per fine cell (a11, a12, a13, a22, a23, a33)
    {
      const double *t = cf + (size_t)f * NS;
      lam = sym3_lambda_min(t[0], t[1], t[2], t[3], t[4], t[5]);
    }
  else if (DIM == 3 && kind == 3) // tensor at the quadrature points
    {
      lam = 1e300;
      for (int q = 0; q < NQ; ++q)
        {
          const double *t = cf + ((size_t)f * NQ + q) * NS;
          lam = fmin(lam, sym3_lambda_min(t[0], t[1], t[2], t[3], t[4], t[5]));
        }
    }
  B[0] = 1.0 / lam;
  B[1] = 0.0;
  B[2] = 1.0 / lam;
  return lam;
}

// Normalised integrals over fine interval j (of nf) of l^2 and l, l = t (bit 1)
// or 1 - t (bit 0).
__device__ __forceinline__ double jint2(int bit, int j, int nf)
{
  const int k = bit ? j : nf - 1 - j;
  const double a = (double)k / nf, b = (double)(k + 1) / nf;
  return (b * b * b - a * a * a) / 3.0;
}
__device__ __forceinline__ double jint1(int bit, int j, int nf)
{
  const int k = bit ? j : nf - 1 - j;
  const double a = (double)k / nf, b = (double)(k + 1) / nf;
  return 0.5 * (b * b - a * a);
}

// Fixed-order block reduction of NV values at once over kEnergyThreads
// threads (values 0 and 1 by minimum, the others by sum), one barrier per
// tree level.
template <int NV>
__device__ __forceinline__ void block_reduce_n(double (&v)[NV], double (*red)[kEnergyThreads])
{
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    red[k][tid] = v[k];
  __syncthreads();
  for (int s = kEnergyThreads / 2; s > 0; s >>= 1)
    {
      if (tid < s)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          red[k][tid] = k < 2 ? fmin(red[k][tid], red[k][tid + s]) : red[k][tid] + red[k][tid + s];
      __syncthreads();
    }
#pragma unroll
  for (int k = 0; k < NV; ++k)
    v[k] = red[k][0];
}

// TPC threads per cell of a size class (kEnergyThreads: one workgroup; kWave:
// one wave, 4 cells per workgroup, the reductions inside the wave and no
// workgroup barrier): lam_lb[cell] and, when tau > 0, ptol[cell][i] = tau
// sqrt(lam_lb e_lb_i).
// (resolutions up to kEnergyMaxNF per side tabulate the interval integrals
// and the mode cosines once per cell instead of per fine cell and mode: the
// divisions and cospi calls were most of the kernel's time)
constexpr int kEnergyMaxNF = 1024;

template <int DIM, int TPC>
__global__ void __launch_bounds__(kEnergyThreads)
  energy_terms_kernel(int n_cells_class, const int *cells, int nf, int kind, const double *box, const double *coef,
                      const long long *coef_off, double tau, double *lam_lb, double *ptol)
{
  constexpr int NB = 1 << DIM, CPB = kEnergyThreads / TPC;
  static_assert(TPC == kEnergyThreads || TPC == kWave, "a workgroup or a wave per cell");
  __shared__ double redn[TPC == kEnergyThreads ? 2 + NB : 1][kEnergyThreads];
  // dynamic, per cell of the workgroup: [jint2 bit 0, jint2 bit 1, jint1 bit 0,
  // jint1 bit 1][nf], then the mode factors k_a(j), m_a(j) per axis [DIM][2][nf]
  // (j = 1 .. nf - 1); none beyond kEnergyMaxNF
  extern __shared__ double stab_all[];
  const int tid = threadIdx.x % TPC, slot = threadIdx.x / TPC;
  const int k = blockIdx.x * CPB + slot;
  if (k >= n_cells_class)
    return;
  double *stab = stab_all + (size_t)slot * (4 + 2 * DIM) * nf;
  auto sync = [&]() {
    if constexpr (TPC == kEnergyThreads)
      __syncthreads();
    else
      {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
  };
  const int cell = cells[k];
  const double *bx = box + (size_t)cell * 2 * DIM;
  const double *cf = coef + coef_off[cell];
  double H[3] = {1, 1, 1}, h[3] = {1, 1, 1};
  for (int a = 0; a < DIM; ++a)
    {
      H[a] = bx[2 * a + 1] - bx[2 * a];
      h[a] = H[a] / nf;
    }
  int nfd = nf * nf;
  if (DIM == 3)
    nfd *= nf;
  const bool tab = nf <= kEnergyMaxNF; // (uniform)
  double *skm = stab + 4 * nf;
  // mode factors along axis a: k = (2 - 2 cos t) / h_a, m = h_a (4 + 2 cos t) / 6, t = j pi / nf
  auto ha = [&](int a) -> double { return a == 0 ? h[0] : a == 1 ? h[1] : h[2]; }; // (no indexed array)
  auto kfac = [&](int a, int j) -> double { return (2.0 - 2.0 * cospi((double)j / nf)) / ha(a); };
  auto mfac = [&](int a, int j) -> double { return ha(a) * (4.0 + 2.0 * cospi((double)j / nf)) / 6.0; };
  if (tab)
    {
      for (int t = tid; t < 4 * nf; t += TPC)
        {
          const int w = t / nf, j = t % nf;
          stab[t] = w < 2 ? jint2(w, j, nf) : jint1(w - 2, j, nf);
        }
      for (int t = tid; t < DIM * 2 * nf; t += TPC)
        {
          const int a = t / (2 * nf), km = t / nf % 2, j = t % nf;
          skm[t] = j == 0 ? 0.0 : km == 0 ? kfac(a, j) : mfac(a, j);
        }
      sync();
    }
  auto J2 = [&](int bit, int j) -> double { return tab ? stab[bit * nf + j] : jint2(bit, j, nf); };
  auto J1 = [&](int bit, int j) -> double { return tab ? stab[(2 + bit) * nf + j] : jint1(bit, j, nf); };
  // the constant factors of G (2D: gxx = H_y / (H_x nf) J2_y, gyy = H_x / (H_y nf) J2_x;
  // 3D: gxx = H_y H_z / (H_x nf) J2_y J2_z, ...)
  double cg[3];
  if (DIM == 2)
    {
      cg[0] = H[1] / (H[0] * nf);
      cg[1] = H[0] / (H[1] * nf);
      cg[2] = 0;
    }
  else
    {
      cg[0] = H[1] * H[2] / (H[0] * nf);
      cg[1] = H[0] * H[2] / (H[1] * nf);
      cg[2] = H[0] * H[1] / (H[2] * nf);
    }
  double lam = 1e300, D[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
    D[i] = 0;
  if constexpr (DIM == 2)
    {
      // D_i = sum_f B0 cg0 J2(by, jy) + B2 cg1 J2(bx, jx) + 2 B1 sx sy J1(bx, jx) J1(by, jy)
      // (bit bx, by of basis i; sx, sy = +-1 by bit): separable, so the fine
      // cells only accumulate P[b] = sum B0 J2(b, jy), Q[b] = sum B2 J2(b, jx)
      // and R[bx][by] = sum B1 J1(bx, jx) J1(by, jy)
      double P[2] = {0, 0}, Q[2] = {0, 0}, Rm[2][2] = {{0, 0}, {0, 0}};
      for (int f = tid; f < nfd; f += TPC)
        {
          double B[3];
          lam = fmin(lam, cell_lambda<DIM>(kind, cf, f, B));
          const int jx = f % nf, jy = f / nf;
#pragma unroll
          for (int b = 0; b < 2; ++b)
            {
              P[b] += B[0] * J2(b, jy);
              Q[b] += B[2] * J2(b, jx);
            }
          if (B[1] != 0.0)
            {
              const double x0 = J1(0, jx), x1 = J1(1, jx), y0 = J1(0, jy), y1 = J1(1, jy);
              Rm[0][0] += B[1] * (x0 * y0);
              Rm[0][1] += B[1] * (x0 * y1);
              Rm[1][0] += B[1] * (x1 * y0);
              Rm[1][1] += B[1] * (x1 * y1);
            }
        }
#pragma unroll
      for (int i = 0; i < NB; ++i)
        {
          const int bxi = i & 1, byi = (i >> 1) & 1;
          const double sxy = (bxi == byi) ? 1.0 : -1.0;
          D[i] = cg[0] * P[byi] + cg[1] * Q[bxi] + 2.0 * sxy * Rm[bxi][byi];
        }
    }
  else
    for (int f = tid; f < nfd; f += TPC)
      {
        double B[3];
        lam = fmin(lam, cell_lambda<DIM>(kind, cf, f, B));
        const int j[3] = {f % nf, (f / nf) % nf, DIM == 3 ? f / (nf * nf) : 0};
  #pragma unroll
        for (int i = 0; i < NB; ++i)
          {
            int bit[3];
            double s[3], j2[3], j1[3];
  #pragma unroll
            for (int a = 0; a < DIM; ++a)
              {
                bit[a] = (i >> a) & 1;
                s[a] = bit[a] ? 1.0 : -1.0;
                j2[a] = J2(bit[a], j[a]);
                j1[a] = J1(bit[a], j[a]);
              }
            // G[k][m] = int over the fine cell of d_k N_i d_m N_i
            double g = 0;
            if (DIM == 2)
              {
                const double gxx = cg[0] * j2[1], gyy = cg[1] * j2[0];
                const double gxy = s[0] * s[1] * j1[0] * j1[1];
                g = B[0] * gxx + 2.0 * B[1] * gxy + B[2] * gyy;
              }
            else
              {
                // scalar coefficients only in 3D: A^-1 = I / lam
                const double gxx = cg[0] * j2[1] * j2[2];
                const double gyy = cg[1] * j2[0] * j2[2];
                const double gzz = cg[2] * j2[0] * j2[1];
                g = B[0] * (gxx + gyy + gzz);
              }
            D[i] += g;
          }
      }
  // mu_min over the separable modes j = 1 .. nf - 1 per axis
  double mu = 1e300;
  const int nm = nf - 1;
  const int nmodes = DIM == 2 ? nm * nm : nm * nm * nm;
  for (int q = tid; q < nmodes; q += TPC)
    {
      const int jm[3] = {q % nm + 1, (q / nm) % nm + 1, DIM == 3 ? q / (nm * nm) + 1 : 1};
      double kk[3], mm[3];
#pragma unroll
      for (int a = 0; a < DIM; ++a)
        {
          kk[a] = tab ? skm[(2 * a) * nf + jm[a]] : kfac(a, jm[a]);
          mm[a] = tab ? skm[(2 * a + 1) * nf + jm[a]] : mfac(a, jm[a]);
        }
      double v;
      if (DIM == 2)
        v = kk[0] * mm[1] + mm[0] * kk[1];
      else
        v = kk[0] * mm[1] * mm[2] + mm[0] * kk[1] * mm[2] + mm[0] * mm[1] * kk[2];
      mu = fmin(mu, v);
    }
  double Dsum[NB];
  if constexpr (TPC == kEnergyThreads)
    {
      // one fixed-order tree over the workgroup for all 2 + NB values (min, min, sums)
      double vals[2 + NB] = {lam, mu};
#pragma unroll
      for (int i = 0; i < NB; ++i)
        vals[2 + i] = D[i];
      block_reduce_n<2 + NB>(vals, redn);
      lam = vals[0];
      mu = vals[1];
#pragma unroll
      for (int i = 0; i < NB; ++i)
        Dsum[i] = vals[2 + i];
    }
  else
    {
      // over the wave: butterfly minima (exact), fixed-order DPP sums
      for (int off = kWave / 2; off > 0; off >>= 1)
        {
          lam = fmin(lam, __shfl_xor(lam, off, kWave));
          mu = fmin(mu, __shfl_xor(mu, off, kWave));
        }
#pragma unroll
      for (int i = 0; i < NB; ++i)
        Dsum[i] = group_sum<kWave>(D[i]);
    }
  const double llb = lam * mu;
  if (tid == 0)
    {
      lam_lb[cell] = llb;
      if (ptol)
        {
          double vol = 1, inv = 0;
          for (int a = 0; a < DIM; ++a)
            {
              vol *= H[a];
              inv += 1.0 / (H[a] * H[a]);
            }
          const double e_unit = vol * inv / (DIM == 2 ? 3.0 : 9.0);
          for (int i = 0; i < NB; ++i)
            {
              const double e_lb = e_unit * e_unit / Dsum[i];
              ptol[(size_t)cell * NB + i] = tau * sqrt(llb * e_lb);
            }
        }
    }
}

// One thread per (cell of the class, basis): the a posteriori bound.  With the
// partition-of-unity option the last basis is 1 - sum of the others, whose
// interior residual is minus the sum of theirs: its norm is bounded by the sum.
__global__ void energy_bound_kernel(int n_cells_class, const int *cells, int nb, int pou, const double *lam_lb,
                                    const double *resid, const double *Ae, double *bound)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_cells_class * nb)
    return;
  const int cell = cells[t / nb], i = t % nb;
  const size_t o = (size_t)cell * nb;
  double eta = resid[o + i];
  if (pou && i == nb - 1)
    {
      eta = 0;
      for (int j = 0; j < nb - 1; ++j)
        eta += resid[o + j];
    }
  const double lam = lam_lb ? lam_lb[cell] : __builtin_nan("");
  const double q = eta * eta / lam, e = Ae[o * nb + (size_t)i * nb + i];
  bound[o + i] = e > q ? sqrt(q / (e - q)) : (lam > 0 ? __builtin_inf() : __builtin_nan(""));
}

__global__ void sub_kernel(size_t n, const double *x, const double *y, double *d)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = x[i] - y[i];
}

// err[c][i] = sqrt(max(Ad_ii, 0) / Ar_ii) for every cell and basis.
__global__ void energy_ratio_kernel(int n_cells, int nb, const double *Ad, const double *Ar, double *err)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_cells * nb)
    return;
  const size_t ii = (size_t)(t / nb) * nb * nb + (size_t)(t % nb) * (nb + 1);
  err[t] = sqrt(fmax(Ad[ii], 0.0) / Ar[ii]);
}

} // namespace

hipError_t launch_sub(size_t n, const double *x, const double *y, double *d, hipStream_t st)
{
  if (n == 0)
    return hipSuccess;
  size_t grid = (n + 255) / 256;
  if (grid > 8192)
    grid = 8192;
  hipLaunchKernelGGL(sub_kernel, dim3((unsigned)grid), dim3(256), 0, st, n, x, y, d);
  return hipGetLastError();
}

hipError_t launch_energy_ratio(int n_cells, int nb, const double *Ad, const double *Ar, double *err, hipStream_t st)
{
  const int n = n_cells * nb;
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(energy_ratio_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n_cells, nb, Ad, Ar, err);
  return hipGetLastError();
}

hipError_t launch_energy_terms(int dim, int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *coef, const long long *coef_off, double tau, double *lam_lb,
                               double *ptol, hipStream_t st)
{
  if (n_cells_class == 0)
    return hipSuccess;
  const size_t lds = nf <= kEnergyMaxNF ? (size_t)(4 + 2 * dim) * nf * sizeof(double) : 0;
  // up to 128 fine cells per side one wave per cell (4 cells per workgroup,
  // their tables side by side), beyond that a workgroup per cell
  const bool wave = nf <= 128;
  const unsigned grid = wave ? (unsigned)((n_cells_class + 3) / 4) : (unsigned)n_cells_class;
  const size_t wl = wave ? 4 * lds : lds;
  if (dim == 2 && wave)
    hipLaunchKernelGGL((energy_terms_kernel<2, kWave>), dim3(grid), dim3(kEnergyThreads), wl, st, n_cells_class,
                       cells, nf, coef_kind, box, coef, coef_off, tau, lam_lb, ptol);
  else if (dim == 2)
    hipLaunchKernelGGL((energy_terms_kernel<2, kEnergyThreads>), dim3(grid), dim3(kEnergyThreads), wl, st,
                       n_cells_class, cells, nf, coef_kind, box, coef, coef_off, tau, lam_lb, ptol);
  else if (wave)
    hipLaunchKernelGGL((energy_terms_kernel<3, kWave>), dim3(grid), dim3(kEnergyThreads), wl, st, n_cells_class,
                       cells, nf, coef_kind, box, coef, coef_off, tau, lam_lb, ptol);
  else
    hipLaunchKernelGGL((energy_terms_kernel<3, kEnergyThreads>), dim3(grid), dim3(kEnergyThreads), wl, st,
                       n_cells_class, cells, nf, coef_kind, box, coef, coef_off, tau, lam_lb, ptol);
  return hipGetLastError();
}

hipError_t launch_energy_bound(int n_cells_class, const int *cells, int nb, int pou, const double *lam_lb,
                               const double *resid, const double *Ae, double *bound, hipStream_t st)
{
  const int n = n_cells_class * nb;
  if (n == 0)
    return hipSuccess;
  hipLaunchKernelGGL(energy_bound_kernel, dim3((n + 255) / 256), dim3(256), 0, st, n_cells_class, cells, nb, pou,
                     lam_lb, resid, Ae, bound);
  return hipGetLastError();
}

} // namespace msfem
