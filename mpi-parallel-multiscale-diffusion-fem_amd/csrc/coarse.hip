// Coarse system of the MsFEM on the unit square/cube with nc cells per side:
//   assembly     DiffusionProblemMultiscale::assemble_system
//                (include/base/diffusion_problem_ms.tpp:161-255; Dirichlet
//                constraints :128-137, Neumann face terms :206-235)
//   solve        DiffusionProblemMultiscale::solve_iterative (:260-296): FP64 CG
//   weights      send_global_weights_to_cell (:301-325) and the fine
//                reconstruction of DiffusionProblemBasis::set_global_weights
//                (include/base/diffusion_problem_basis.tpp:354-377)
// The coarse matrix is kept in structured stencil form Kc[node][3^dim].
#include <algorithm>

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

// ---- multigrid-preconditioned CG ---------------------------------------------------
// The coarse system plays the role of the reference's AMG-preconditioned CG
// (diffusion_problem_ms.tpp:267-289).  Here the grid is structured, so the
// preconditioner is one geometric V(2,2) cycle: levels n, n/2, ... (while the
// cells per side are even and >= 4), bilinear/trilinear prolongation P, full
// weighting P^T, Galerkin operators A_{l+1} = P^T A_l P on the free nodes (a
// coarse node is constrained when the fine node at its position is), damped
// Jacobi smoothing (omega = 0.7), and a dense Cholesky solve on the coarsest
// level.  The cycle is symmetric, so CG applies.  One workgroup runs the whole
// solve; every vector lives in the scratch `w` (L2-resident at these sizes) and
// every pass ends in a workgroup barrier.  Layout: MgLayout (host-computed).
constexpr double kMgOmega = 0.7;

template <int DIM>
struct MgGrid {
  int n; // cells per side; nodes (n+1)^DIM, x fastest
  __device__ __forceinline__ int np() const { return n + 1; }
  // node / p by a float reciprocal and an exact one-step fix-up (node < 2^24)
  __device__ __forceinline__ static int divp(int node, int p)
  {
    int q = (int)((float)node * (1.0f / (float)p));
    if (q * p > node)
      --q;
    else if ((q + 1) * p <= node)
      ++q;
    return q;
  }
  __device__ __forceinline__ void coords(int node, int (&I)[3]) const
  {
    const int p = n + 1;
    const int q = divp(node, p);
    I[0] = node - q * p;
    if (DIM == 2)
      {
        I[1] = q;
        I[2] = 0;
      }
    else
      {
        const int q2 = divp(q, p);
        I[1] = q - q2 * p;
        I[2] = q2;
      }
  }
  __device__ __forceinline__ int index(const int (&I)[3]) const
  {
    const int p = n + 1;
    return I[0] + p * (I[1] + (DIM == 3 ? p * I[2] : 0));
  }
  __device__ __forceinline__ bool inside(const int (&I)[3]) const
  {
    for (int k = 0; k < DIM; ++k)
      if (I[k] < 0 || I[k] > n)
        return false;
    return true;
  }
};

// Constrained node of multigrid level l: the level-0 node at its position is
// constrained (level l node I sits at level-0 node I * 2^l).
template <int DIM>
__device__ __forceinline__ bool mg_masked_from0(const MgLayout &g, int nc, const unsigned char *cmask, int l,
                                                int node)
{
  if (l == 0)
    return cmask[node] != 0;
  MgGrid<DIM> Gl{g.n[l]}, G0{nc};
  int I[3];
  Gl.coords(node, I);
  for (int k = 0; k < DIM; ++k)
    I[k] <<= l;
  return cmask[G0.index(I)] != 0;
}

// The flags of levels 1.. as byte arrays in the work buffer (one pass before
// the Galerkin products, which then read one byte per test).
template <int DIM>
__global__ void __launch_bounds__(256) mg_mask_kernel(MgLayout g, int nc, const unsigned char *cmask, double *w)
{
  for (int l = 1; l < g.levels; ++l)
    {
      unsigned char *mk = (unsigned char *)(w + g.off_mask[l]);
      for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < (int)g.nn[l]; n += gridDim.x * blockDim.x)
        mk[n] = mg_masked_from0<DIM>(g, nc, cmask, l, n) ? 1 : 0;
    }
}

__device__ __forceinline__ bool mg_masked(const MgLayout &g, const unsigned char *cmask, const double *w, int l,
                                          int node)
{
  return (l == 0 ? cmask[node] : ((const unsigned char *)(w + g.off_mask[l]))[node]) != 0;
}

// Galerkin operator of level l >= 1, A_l = P^T A_{l-1} P with P the bilinear /
// trilinear interpolation between consecutive levels, on the free nodes (rows
// and columns of constrained nodes dropped), stored as a 3^DIM stencil per
// node.  One thread per (coarse node, fine node of its support): each sums the
// contributions of its fine node's row into a private stencil, the workgroup's
// 3^DIM partial stencils of a coarse node are then added in a fixed order.  The
// levels run one launch after another, each reading the one before.
template <int DIM>
constexpr int mg_galerkin_nodes() { return DIM == 2 ? 28 : 9; } // coarse nodes per workgroup

template <int DIM>
__global__ void __launch_bounds__(mg_galerkin_nodes<DIM>() * (DIM == 2 ? 9 : 27))
  mg_galerkin_kernel(MgLayout g, int l, const double *Kc, const unsigned char *cmask, double *w)
{
  constexpr int S = DIM == 2 ? 9 : 27, NPB = mg_galerkin_nodes<DIM>();
  __shared__ double part[NPB][S][S];
  MgGrid<DIM> Gf{g.n[l - 1]}, Gc{g.n[l]};
  const double *Af = l == 1 ? Kc : w + g.off_A[l - 1];
  double *Ac = w + g.off_A[l];
  const int cn = threadIdx.x / S, o = threadIdx.x % S;
  const int c = blockIdx.x * NPB + cn;
  double acc[S];
#pragma unroll
  for (int t = 0; t < S; ++t)
    acc[t] = 0;
  if (c < (int)g.nn[l] && !mg_masked(g, cmask, w, l, c))
    {
      int I[3];
      Gc.coords(c, I);
      // fine node i = 2I + o in the support of coarse node I
      const int od[3] = {o % 3 - 1, (o / 3) % 3 - 1, DIM == 3 ? o / 9 - 1 : 0};
      int i[3] = {2 * I[0] + od[0], 2 * I[1] + od[1], 2 * I[2] + od[2]};
      if (Gf.inside(i) && !mg_masked(g, cmask, w, l - 1, Gf.index(i)))
        {
          const int fi = Gf.index(i);
          double wi = 1;
          for (int k = 0; k < DIM; ++k)
            wi *= od[k] ? 0.5 : 1.0;
          for (int t = 0; t < S; ++t) // fine neighbour j = i + s
            {
              int j[3] = {i[0] + t % 3 - 1, i[1] + (t / 3) % 3 - 1, DIM == 3 ? i[2] + t / 9 - 1 : 0};
              if (!Gf.inside(j) || mg_masked(g, cmask, w, l - 1, Gf.index(j)))
                continue;
              const double a = wi * Af[(size_t)fi * S + t];
              // coarse parents J of j: per axis one (even) or two (odd) with weight 1/2
              int nodd = 0;
              for (int k = 0; k < DIM; ++k)
                nodd += j[k] & 1;
              for (int comb = 0; comb < (1 << nodd); ++comb)
                {
                  int J[3] = {0, 0, 0}, bitpos = 0;
                  double wj = 1;
                  for (int k = 0; k < DIM; ++k)
                    if (j[k] & 1)
                      {
                        J[k] = (j[k] >> 1) + ((comb >> bitpos) & 1);
                        ++bitpos;
                        wj *= 0.5;
                      }
                    else
                      J[k] = j[k] >> 1;
                  if (mg_masked(g, cmask, w, l, Gc.index(J)))
                    continue;
                  int tt = 0, mul = 1;
                  for (int k = 0; k < DIM; ++k)
                    {
                      tt += (J[k] - I[k] + 1) * mul;
                      mul *= 3;
                    }
                  acc[tt] += a * wj;
                }
            }
        }
    }
#pragma unroll
  for (int t = 0; t < S; ++t)
    part[cn][o][t] = acc[t];
  __syncthreads();
  if (c < (int)g.nn[l])
    {
      const int t = o; // this thread adds stencil entry t of coarse node c over the support
      double v = 0;
      for (int k = 0; k < S; ++k)
        v += part[cn][k][t];
      Ac[(size_t)c * S + t] = v;
    }
}

template <int DIM>
__global__ void __launch_bounds__(kCoarseThreads)
  coarse_mgcg_kernel(MgLayout g, int nc, const double *Kc, const double *Fc, const unsigned char *cmask,
                     const double *cval, double rel_tol, int max_iter, double *u, int *iters, double *w)
{
  constexpr int S = DIM == 2 ? 9 : 27;
  __shared__ double red[kCoarseThreads / kWave];
  const int tid = threadIdx.x;
  const int L = g.levels;
  MgGrid<DIM> G0{nc};
  const int nn0 = (int)g.nn[0];
  double *cx = w + g.off_cx, *cr = w + g.off_cr, *cp = w + g.off_cp, *cq = w + g.off_cq;

  // constrained node of level l: the level-0 node at its position is constrained
  auto masked = [&](int l, int node) -> bool { return mg_masked(g, cmask, w, l, node); };
  auto Aof = [&](int l) -> const double * { return l == 0 ? Kc : w + g.off_A[l]; };
  // constrained node of level l after the set-up: its Jacobi entry is exactly 0
  auto fixed = [&](int l, int node) -> bool { return w[g.off_d[l] + node] == 0.0; };
  // (A_l v)_node over the neighbours inside the grid (rows of constrained nodes: 0);
  // after the set-up
  auto apply = [&](int l, const double *v, int node) -> double {
    if (fixed(l, node))
      return 0.0;
    MgGrid<DIM> Gl{g.n[l]};
    int I[3];
    Gl.coords(node, I);
    const double *A = Aof(l) + (size_t)node * S;
    double s = 0;
    for (int t = 0; t < S; ++t)
      {
        int J[3] = {I[0] + t % 3 - 1, I[1] + (t / 3) % 3 - 1, DIM == 3 ? I[2] + t / 9 - 1 : 0};
        if (Gl.inside(J))
          s += A[t] * v[Gl.index(J)];
      }
    return s;
  };

  // (A_l D_l v)_node: the operator applied to the Jacobi-scaled vector (d = 0 on
  // constrained nodes; rows of constrained nodes: 0); after the set-up
  auto apply_dv = [&](int l, const double *v, int node) -> double {
    if (fixed(l, node))
      return 0.0;
    MgGrid<DIM> Gl{g.n[l]};
    int I[3];
    Gl.coords(node, I);
    const double *A = Aof(l) + (size_t)node * S;
    const double *d = w + g.off_d[l];
    double s = 0;
    for (int t = 0; t < S; ++t)
      {
        int J[3] = {I[0] + t % 3 - 1, I[1] + (t / 3) % 3 - 1, DIM == 3 ? I[2] + t / 9 - 1 : 0};
        if (Gl.inside(J))
          {
            const int m = Gl.index(J);
            s += A[t] * (d[m] * v[m]);
          }
      }
    return s;
  };

  // ---- set-up: Jacobi diagonals, coarsest Cholesky (the Galerkin operators of
  // levels 1.. come from mg_galerkin_kernel launches before this kernel) ------
  for (int l = 0; l < L; ++l)
    {
      double *d = w + g.off_d[l];
      const double *A = Aof(l);
      for (int n = tid; n < (int)g.nn[l]; n += kCoarseThreads)
        d[n] = masked(l, n) ? 0.0 : 1.0 / A[(size_t)n * S + S / 2];
    }
  // coarsest level: free-node numbering (thread 0), dense matrix, Cholesky in place
  const int lc = L - 1, Nc = (int)g.nn[lc];
  double *M = w + g.off_dense, *dmap = w + g.off_dmap;
  __shared__ int s_nfree;
  if (tid == 0)
    {
      int f = 0;
      for (int n = 0; n < Nc; ++n)
        dmap[n] = masked(lc, n) ? -1.0 : (double)f++;
      s_nfree = f;
    }
  __syncthreads();
  const int F = s_nfree;
  for (int e = tid; e < F * F; e += kCoarseThreads)
    M[e] = 0.0;
  __syncthreads();
  {
    MgGrid<DIM> Gl{g.n[lc]};
    const double *A = Aof(lc);
    for (int n = tid; n < Nc; n += kCoarseThreads)
      {
        const int fa = (int)dmap[n];
        if (fa < 0)
          continue;
        int I[3];
        Gl.coords(n, I);
        for (int t = 0; t < S; ++t)
          {
            int J[3] = {I[0] + t % 3 - 1, I[1] + (t / 3) % 3 - 1, DIM == 3 ? I[2] + t / 9 - 1 : 0};
            if (!Gl.inside(J))
              continue;
            const int fb = (int)dmap[Gl.index(J)];
            if (fb >= 0)
              M[(size_t)fa * F + fb] = A[(size_t)n * S + t];
          }
      }
  }
  __syncthreads();
  for (int k = 0; k < F; ++k)
    {
      if (tid == 0)
        M[(size_t)k * F + k] = sqrt(M[(size_t)k * F + k]);
      __syncthreads();
      for (int i = k + 1 + tid; i < F; i += kCoarseThreads)
        M[(size_t)i * F + k] /= M[(size_t)k * F + k];
      __syncthreads();
      const int m = F - k - 1; // update the trailing lower triangle, rows/cols k+1..F-1
      for (int e = tid; e < m * m; e += kCoarseThreads)
        {
          const int i = k + 1 + e / m, j = k + 1 + e % m;
          if (j <= i)
            M[(size_t)i * F + j] -= M[(size_t)i * F + k] * M[(size_t)j * F + k];
        }
      __syncthreads();
    }

  // ---- one V(2,2) cycle: z = V(b0) into level 0's x ------------------------------
  auto vcycle = [&](const double *b0) {
    for (int l = 0; l + 1 < L; ++l)
      {
        const double *b = l == 0 ? b0 : w + g.off_b[l];
        double *x = w + g.off_x[l], *t = w + g.off_t[l];
        const double *d = w + g.off_d[l];
        double *u = w + g.off_u[l];
        const int nn = (int)g.nn[l];
        // two damped Jacobi sweeps from x = 0 in two passes, each residual from
        // the previous one's neighbours (x1 = w D b, t1 = b - A x1 = b - w A D b;
        // x2 = x1 + w D t1, t2 = b - A x2 = t1 - w A D t1): the values of the
        // four-pass form up to rounding, two barriers fewer
        for (int n = tid; n < nn; n += kCoarseThreads)
          {
            x[n] = kMgOmega * d[n] * b[n];
            t[n] = fixed(l, n) ? 0.0 : b[n] - kMgOmega * apply_dv(l, b, n);
          }
        __syncthreads();
        for (int n = tid; n < nn; n += kCoarseThreads)
          {
            x[n] += kMgOmega * d[n] * t[n];
            u[n] = fixed(l, n) ? 0.0 : t[n] - kMgOmega * apply_dv(l, t, n);
          }
        __syncthreads();
        t = u; // the residual restricted below
        MgGrid<DIM> Gf{g.n[l]}, Gc{g.n[l + 1]};
        double *bc = w + g.off_b[l + 1];
        for (int c = tid; c < (int)g.nn[l + 1]; c += kCoarseThreads)
          {
            double sres = 0;
            if (!fixed(l + 1, c))
              {
                int I[3];
                Gc.coords(c, I);
                for (int o = 0; o < S; ++o)
                  {
                    const int od[3] = {o % 3 - 1, (o / 3) % 3 - 1, DIM == 3 ? o / 9 - 1 : 0};
                    int i[3] = {2 * I[0] + od[0], 2 * I[1] + od[1], 2 * I[2] + od[2]};
                    if (!Gf.inside(i))
                      continue;
                    double wi = 1;
                    for (int k = 0; k < DIM; ++k)
                      wi *= od[k] ? 0.5 : 1.0;
                    sres += wi * t[Gf.index(i)];
                  }
              }
            bc[c] = sres;
          }
        __syncthreads();
      }
    // coarsest: dense triangular solves by one thread
    {
      const double *b = lc == 0 ? b0 : w + g.off_b[lc];
      double *x = w + g.off_x[lc], *y = w + g.off_t[lc];
      if (tid == 0)
        {
          for (int n = 0; n < Nc; ++n)
            if (dmap[n] >= 0)
              y[(int)dmap[n]] = b[n];
          for (int i = 0; i < F; ++i)
            {
              double v = y[i];
              for (int j = 0; j < i; ++j)
                v -= M[(size_t)i * F + j] * y[j];
              y[i] = v / M[(size_t)i * F + i];
            }
          for (int i = F - 1; i >= 0; --i)
            {
              double v = y[i];
              for (int j = i + 1; j < F; ++j)
                v -= M[(size_t)j * F + i] * y[j];
              y[i] = v / M[(size_t)i * F + i];
            }
        }
      __syncthreads();
      for (int n = tid; n < Nc; n += kCoarseThreads)
        x[n] = dmap[n] >= 0 ? y[(int)dmap[n]] : 0.0;
      __syncthreads();
    }
    for (int l = L - 2; l >= 0; --l)
      {
        const double *b = l == 0 ? b0 : w + g.off_b[l];
        double *x = w + g.off_x[l], *t = w + g.off_t[l];
        const double *d = w + g.off_d[l];
        const double *xc = w + g.off_x[l + 1];
        const int nn = (int)g.nn[l];
        MgGrid<DIM> Gf{g.n[l]}, Gc{g.n[l + 1]};
        for (int n = tid; n < nn; n += kCoarseThreads)
          {
            if (fixed(l, n))
              continue;
            int j[3];
            Gf.coords(n, j);
            int nodd = 0;
            for (int k = 0; k < DIM; ++k)
              nodd += j[k] & 1;
            double v = 0;
            for (int comb = 0; comb < (1 << nodd); ++comb)
              {
                int J[3] = {0, 0, 0}, bitpos = 0;
                double wj = 1;
                for (int k = 0; k < DIM; ++k)
                  if (j[k] & 1)
                    {
                      J[k] = (j[k] >> 1) + ((comb >> bitpos) & 1);
                      ++bitpos;
                      wj *= 0.5;
                    }
                  else
                    J[k] = j[k] >> 1;
                v += wj * xc[Gc.index(J)];
              }
            x[n] += v;
          }
        __syncthreads();
        // two damped Jacobi sweeps in two passes (t = b - A x, x1 = x + w D t;
        // x2 = x1 + w D (b - A x1) = x1 + w D (t - w A D t)), x1 in u
        double *u = w + g.off_u[l];
        for (int n = tid; n < nn; n += kCoarseThreads)
          {
            const double tv = fixed(l, n) ? 0.0 : b[n] - apply(l, x, n);
            t[n] = tv;
            u[n] = x[n] + kMgOmega * d[n] * tv;
          }
        __syncthreads();
        for (int n = tid; n < nn; n += kCoarseThreads)
          {
            const double sv = fixed(l, n) ? 0.0 : t[n] - kMgOmega * apply_dv(l, t, n);
            x[n] = u[n] + kMgOmega * d[n] * sv;
          }
        __syncthreads();
      }
  };

  // ---- CG on the constrained system (x = cval on the constrained nodes) -----------
  for (int n = tid; n < nn0; n += kCoarseThreads)
    {
      cx[n] = cmask[n] ? cval[n] : 0.0;
      cp[n] = 0.0;
    }
  __syncthreads();
  double bb = 0;
  for (int n = tid; n < nn0; n += kCoarseThreads)
    {
      double rv = 0.0;
      if (!cmask[n]) // F - K x with the Dirichlet values in x (every neighbour counted)
        {
          int I[3];
          G0.coords(n, I);
          double s = 0;
          for (int t = 0; t < S; ++t)
            {
              int J[3] = {I[0] + t % 3 - 1, I[1] + (t / 3) % 3 - 1, DIM == 3 ? I[2] + t / 9 - 1 : 0};
              if (G0.inside(J))
                s += Kc[(size_t)n * S + t] * cx[G0.index(J)];
            }
          rv = Fc[n] - s;
        }
      cr[n] = rv;
      bb += rv * rv;
    }
  bb = block_sum(bb, red);
  const double tol = rel_tol * sqrt(bb);
  double rn = sqrt(bb), rz = 0;
  int it = 0;
  double *z = w + g.off_x[0];
  if (rn > tol)
    {
      vcycle(cr);
      for (int n = tid; n < nn0; n += kCoarseThreads)
        {
          cp[n] = z[n];
          rz += cr[n] * z[n];
        }
      rz = block_sum(rz, red);
    }
  while (rn > tol && it < max_iter)
    {
      double pq = 0;
      for (int n = tid; n < nn0; n += kCoarseThreads)
        {
          const double qv = apply(0, cp, n);
          cq[n] = qv;
          pq += cp[n] * qv;
        }
      pq = block_sum(pq, red);
      const double alpha = rz / pq;
      double rr = 0;
      for (int n = tid; n < nn0; n += kCoarseThreads)
        {
          cx[n] += alpha * cp[n];
          const double rv = cr[n] - alpha * cq[n];
          cr[n] = rv;
          rr += rv * rv;
        }
      rr = block_sum(rr, red);
      rn = sqrt(rr);
      ++it;
      if (rn <= tol || it >= max_iter)
        break;
      vcycle(cr);
      double rz_new = 0;
      for (int n = tid; n < nn0; n += kCoarseThreads)
        rz_new += cr[n] * z[n];
      rz_new = block_sum(rz_new, red);
      const double beta = rz_new / rz;
      rz = rz_new;
      for (int n = tid; n < nn0; n += kCoarseThreads)
        cp[n] = z[n] + beta * cp[n];
      __syncthreads();
    }
  __syncthreads();
  for (int n = tid; n < nn0; n += kCoarseThreads)
    u[n] = cx[n];
  if (tid == 0)
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

MgLayout mg_layout(int dim, int nc)
{
  MgLayout g = {};
  auto nodes = [dim](long long n) { return dim == 2 ? (n + 1) * (n + 1) : (n + 1) * (n + 1) * (n + 1); };
  int n = nc, L = 0;
  g.n[L++] = n;
  while (n % 2 == 0 && n >= 4 && L < kMgMaxLevels)
    {
      n /= 2;
      g.n[L++] = n;
    }
  if (nodes(g.n[L - 1]) > kMgMaxDense)
    return MgLayout{}; // the grid does not coarsen to a small dense problem: no multigrid
  g.levels = L;
  const long long S = dim == 2 ? 9 : 27;
  long long off = 0;
  for (int l = 0; l < L; ++l)
    g.nn[l] = nodes(g.n[l]);
  g.off_cx = off, off += g.nn[0];
  g.off_cr = off, off += g.nn[0];
  g.off_cp = off, off += g.nn[0];
  g.off_cq = off, off += g.nn[0];
  for (int l = 0; l < L; ++l)
    {
      g.off_x[l] = off, off += g.nn[l];
      g.off_t[l] = off, off += g.nn[l];
      g.off_u[l] = off, off += g.nn[l];
      g.off_d[l] = off, off += g.nn[l];
      g.off_b[l] = off, off += l > 0 ? g.nn[l] : 0; // level 0's right-hand side is the CG residual
      g.off_A[l] = off, off += l > 0 ? g.nn[l] * S : 0; // level 0's operator is the coarse matrix itself
      g.off_mask[l] = off, off += l > 0 ? (g.nn[l] + 7) / 8 : 0; // level 0's flags are the coarse mask
    }
  const long long nd = g.nn[L - 1];
  g.off_dense = off, off += nd * nd;
  g.off_dmap = off, off += nd;
  g.total = off;
  return g;
}

size_t coarse_solve_work_doubles(int dim, int nc)
{
  const MgLayout g = mg_layout(dim, nc);
  if (g.levels > 0)
    return (size_t)g.total;
  const size_t np = nc + 1;
  return 5 * (dim == 2 ? np * np : np * np * np);
}

hipError_t launch_coarse_solve(int dim, int nc, const double *Kc, const double *Fc, const unsigned char *cmask,
                               const double *cval, double rel_tol, int max_iter, double *u, int *iters,
                               double *work, hipStream_t st)
{
  const MgLayout g = mg_layout(dim, nc);
  if (g.levels > 0)
    {
      if (dim == 2)
        hipLaunchKernelGGL(mg_mask_kernel<2>, dim3(64), dim3(256), 0, st, g, nc, cmask, work);
      else
        hipLaunchKernelGGL(mg_mask_kernel<3>, dim3(64), dim3(256), 0, st, g, nc, cmask, work);
      for (int l = 1; l < g.levels; ++l) // Galerkin operators, level by level, over the whole GPU
        {
          if (dim == 2)
            hipLaunchKernelGGL(mg_galerkin_kernel<2>,
                               dim3((unsigned)((g.nn[l] + mg_galerkin_nodes<2>() - 1) / mg_galerkin_nodes<2>())),
                               dim3(mg_galerkin_nodes<2>() * 9), 0, st, g, l, Kc, cmask, work);
          else
            hipLaunchKernelGGL(mg_galerkin_kernel<3>,
                               dim3((unsigned)((g.nn[l] + mg_galerkin_nodes<3>() - 1) / mg_galerkin_nodes<3>())),
                               dim3(mg_galerkin_nodes<3>() * 27), 0, st, g, l, Kc, cmask, work);
        }
      if (dim == 2)
        hipLaunchKernelGGL(coarse_mgcg_kernel<2>, dim3(1), dim3(kCoarseThreads), 0, st, g, nc, Kc, Fc, cmask, cval,
                           rel_tol, max_iter, u, iters, work);
      else
        hipLaunchKernelGGL(coarse_mgcg_kernel<3>, dim3(1), dim3(kCoarseThreads), 0, st, g, nc, Kc, Fc, cmask, cval,
                           rel_tol, max_iter, u, iters, work);
      return hipGetLastError();
    }
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
