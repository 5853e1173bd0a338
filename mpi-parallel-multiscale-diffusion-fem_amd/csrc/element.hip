// Coarse element matrix and load of every coarse cell from its converged bases:
//     A^K_ij = phi_i^T K phi_j,    F^K_i = phi_i^T b,
// DiffusionProblemBasis<dim>::assemble_global_element_matrix
// (include/base/diffusion_problem_basis.tpp:247-285), evaluated fine cell by
// fine cell as sum_c phi_i|_c^T K_c phi_j|_c with the QGauss(2) element
// matrices of assemble_system (:161-242).  One workgroup per coarse cell; the
// per-thread partial sums are combined in a fixed order (deterministic).
#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

constexpr int kElemThreads = 256;
constexpr int kElemMaxNF = 128; // largest fine resolution of a cell (2D)
constexpr int kElem3MaxNF = 16; // and in 3D

__device__ __forceinline__ int sym_index(int dim, int i, int j)
{
  if (i > j)
    {
      const int t = i;
      i = j;
      j = t;
    }
  if (dim == 2)
    return i == 0 ? j : 2;
  if (i == 0)
    return j;
  if (i == 1)
    return 2 + j;
  return 5;
}

// Upper-triangle index of (i, j), i <= j, of an nv x nv symmetric matrix.
template <int NVV>
__device__ __forceinline__ constexpr int symi_of(int i, int j)
{
  return i * NVV - i * (i - 1) / 2 + (j - i);
}

// QUAD: general cells (a.verts: quadrilaterals, hexahedra), geometry per fine cell
// QPK: the coefficient kind may be one given at the quadrature points; false
// compiles only the per-fine-cell kinds (their Gauss sums are folded into unit
// matrices), which keeps the common instance's registers (and so its occupancy)
// free of the quadrature-point path.
template <int DIM, bool QUAD, bool QPK = true>
__global__ void __launch_bounds__(kElemThreads, DIM == 2 ? 2 : 1) element_kernel(ElementArgs a)
{
  constexpr int NV = 1 << DIM, NQ = 1 << DIM, NS = DIM == 2 ? 3 : 6;
  // A^K is symmetric: the running sums hold its upper triangle (NSYM entries,
  // index symi(i, j), i <= j), then F^K
  constexpr int NSYM = NV * (NV + 1) / 2, NVAL = NSYM + NV;
  __shared__ double sM[NS][NV * NV]; // unit element matrices of each tensor component
  __shared__ double sred[kElemThreads / kWave][NVAL];
  auto symi = [](int i, int j) { return symi_of<NV>(i, j); };
  // f = prod_k sin(pi x_k) is separable: sin(pi x_k) at every fine cell's two
  // Gauss coordinates along each axis, tabulated once per cell (box cells)
  __shared__ double ssin[QUAD ? 1 : DIM][QUAD ? 1 : 2 * kElemMaxNF];

  const int cell = a.cells[blockIdx.x];
  const int nf = a.nf, np = nf + 1;
  // general cell: geometry per fine cell from the bi/trilinear map; the box
  // data below is then unused (unit placeholders)
  const double *vq = QUAD ? a.verts + (size_t)cell * NV * DIM : nullptr;
  const double unit_box[6] = {0, 1, 0, 1, 0, 1};
  const double *bx = vq ? unit_box : a.box + (size_t)cell * 2 * DIM;
  double h[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < DIM; ++k)
    {
      lo[k] = bx[2 * k];
      h[k] = (bx[2 * k + 1] - bx[2 * k]) / nf;
    }
  const double jxw = (DIM == 2 ? h[0] * h[1] : h[0] * h[1] * h[2]) / NQ;
  const double gs = 0.5 / sqrt(3.0);

  // d psi_a / d x_k at quadrature point q
  auto grad = [&](int aa, int k, int q) -> double {
    double g = ((aa >> k) & 1) ? 1.0 : -1.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m)
      if (m != k)
        {
          const double xi = ((q >> m) & 1) ? 0.5 + gs : 0.5 - gs;
          g *= ((aa >> m) & 1) ? xi : 1.0 - xi;
        }
    return g / h[k];
  };
  auto shape = [&](int aa, int q) -> double {
    double v = 1.0;
#pragma unroll
    for (int m = 0; m < DIM; ++m)
      {
        const double xi = ((q >> m) & 1) ? 0.5 + gs : 0.5 - gs;
        v *= ((aa >> m) & 1) ? xi : 1.0 - xi;
      }
    return v;
  };

  // Unit matrices: M_s[a][b] = sum_q jxw (d_k psi_a d_l psi_b + [k != l] d_l psi_a d_k psi_b)
  // for the tensor component s = (k, l), k <= l.
  for (int t = threadIdx.x; t < NS * NV * NV; t += blockDim.x)
    {
      const int s = t / (NV * NV), ab = t % (NV * NV), aa = ab / NV, bb = ab % NV;
      int k = 0, l = 0;
      for (int kk = 0; kk < DIM; ++kk)
        for (int ll = kk; ll < DIM; ++ll)
          if (sym_index(DIM, kk, ll) == s)
            {
              k = kk;
              l = ll;
            }
      double v = 0;
      for (int q = 0; q < NQ; ++q)
        {
          double e = grad(aa, k, q) * grad(bb, l, q);
          if (k != l)
            e += grad(aa, l, q) * grad(bb, k, q);
          v += e * jxw;
        }
      sM[s][ab] = v;
    }
  // (beyond kElemMaxNF fine cells per side the factors are evaluated in place)
  const bool sin_table = nf <= kElemMaxNF;
  if constexpr (!QUAD)
    if (a.rhs_kind == 1 && sin_table)
      for (int t = threadIdx.x; t < DIM * 2 * nf; t += blockDim.x)
        {
          const int k = t / (2 * nf), fi = t % (2 * nf) / 2;
          const double xi = (t & 1) ? 0.5 + gs : 0.5 - gs;
          ssin[k][t % (2 * nf)] = sin(M_PI * (lo[k] + (bx[2 * k + 1] - bx[2 * k]) * ((double)fi / nf) + h[k] * xi));
        }
  __syncthreads();

  const double *cf = a.coef + a.coef_off[cell];
  const double *phi = a.bases + a.bas_off[cell];
  const size_t nn = DIM == 2 ? (size_t)np * np : (size_t)np * np * np;
  const int ncell = DIM == 2 ? nf * nf : nf * nf * nf;

  double acc[NVAL];
#pragma unroll
  for (int v = 0; v < NVAL; ++v)
    acc[v] = 0;

  for (int f = threadIdx.x; f < ncell; f += blockDim.x)
    {
      const int fx = f % nf, fy = (f / nf) % nf, fz = DIM == 3 ? f / (nf * nf) : 0;
      double Ke[NV * NV];
      double be[NV];
      if constexpr (QUAD && DIM == 2)
        {
          double K4[4][4];
          qg_element_quad(a.coef_kind, cf, f, nf, QuadMap(vq), K4, a.rhs_kind, a.rhs_value, be);
#pragma unroll
          for (int ab = 0; ab < 16; ++ab)
            Ke[ab] = K4[ab / 4][ab % 4];
        }
      else if constexpr (QUAD)
        qg_element_hex(a.coef_kind, cf, f, nf, HexMap(vq), Ke, a.rhs_kind, a.rhs_value, be);
      else if (!QPK || a.coef_kind == 0 || a.coef_kind == 1)
        {
          double cs[NS];
          if (a.coef_kind == 0)
            {
#pragma unroll
              for (int s = 0; s < NS; ++s)
                cs[s] = 0;
              cs[sym_index(DIM, 0, 0)] = cs[sym_index(DIM, 1, 1)] = cf[f];
              if (DIM == 3)
                cs[sym_index(DIM, 2, 2)] = cf[f];
            }
          else
            {
#pragma unroll
              for (int s = 0; s < NS; ++s)
                cs[s] = cf[(size_t)NS * f + s];
            }
#pragma unroll
          for (int ab = 0; ab < NV * NV; ++ab)
            {
              double v = 0;
#pragma unroll
              for (int s = 0; s < NS; ++s)
                v += cs[s] * sM[s][ab];
              Ke[ab] = v;
            }
        }
      else if constexpr (QPK)
        {
          // coefficient at the quadrature points: Gauss sum per fine cell
#pragma unroll
          for (int ab = 0; ab < NV * NV; ++ab)
            Ke[ab] = 0;
          for (int q = 0; q < NQ; ++q)
            {
              double A[3][3];
              for (int k = 0; k < DIM; ++k)
                for (int l = 0; l < DIM; ++l)
                  {
                    if (a.coef_kind == 2)
                      A[k][l] = k == l ? cf[(size_t)NQ * f + q] : 0.0;
                    else
                      A[k][l] = cf[((size_t)NQ * f + q) * NS + sym_index(DIM, k, l)];
                  }
              double g[NV][3];
              for (int aa = 0; aa < NV; ++aa)
                for (int k = 0; k < DIM; ++k)
                  g[aa][k] = grad(aa, k, q);
              for (int aa = 0; aa < NV; ++aa)
                for (int bb = 0; bb < NV; ++bb)
                  {
                    double e = 0;
                    for (int k = 0; k < DIM; ++k)
                      for (int l = 0; l < DIM; ++l)
                        e += g[aa][k] * A[k][l] * g[bb][l];
                    Ke[aa * NV + bb] += e * jxw;
                  }
            }
        }
      // load vector of the fine cell (box cells)
      if constexpr (!QUAD)
        {
#pragma unroll
          for (int aa = 0; aa < NV; ++aa)
            be[aa] = 0;
          for (int q = 0; q < NQ; ++q)
            {
              double fv = a.rhs_value;
              if (a.rhs_kind == 1)
                {
                  fv = 1.0;
                  const int fi[3] = {fx, fy, fz};
                  for (int k = 0; k < DIM; ++k)
                    fv *= sin_table ? ssin[k][2 * fi[k] + ((q >> k) & 1)]
                                    : sin(M_PI * (lo[k] + (bx[2 * k + 1] - bx[2 * k]) * ((double)fi[k] / nf) +
                                                  h[k] * (((q >> k) & 1) ? 0.5 + gs : 0.5 - gs)));
                }
#pragma unroll
              for (int aa = 0; aa < NV; ++aa)
                be[aa] += shape(aa, q) * fv * jxw;
            }
        }
      // bases at the fine cell's vertices
      double pv[NV][NV]; // [basis][vertex]
#pragma unroll
      for (int aa = 0; aa < NV; ++aa)
        {
          const size_t node = (size_t)(fx + (aa & 1)) +
                              (size_t)np * ((fy + ((aa >> 1) & 1)) + (size_t)np * (fz + ((aa >> 2) & 1)));
#pragma unroll
          for (int i = 0; i < NV; ++i)
            pv[i][aa] = phi[(size_t)i * nn + node];
        }
#pragma unroll
      for (int j = 0; j < NV; ++j)
        {
          double y[NV];
#pragma unroll
          for (int aa = 0; aa < NV; ++aa)
            {
              double s = 0;
#pragma unroll
              for (int bb = 0; bb < NV; ++bb)
                s += Ke[aa * NV + bb] * pv[j][bb];
              y[aa] = s;
            }
#pragma unroll
          for (int i = 0; i <= j; ++i)
            {
              double s = 0;
#pragma unroll
              for (int aa = 0; aa < NV; ++aa)
                s += pv[i][aa] * y[aa];
              acc[symi(i, j)] += s;
            }
        }
#pragma unroll
      for (int i = 0; i < NV; ++i)
        {
          double s = 0;
#pragma unroll
          for (int aa = 0; aa < NV; ++aa)
            s += pv[i][aa] * be[aa];
          acc[NSYM + i] += s;
        }
    }

  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int v = 0; v < NVAL; ++v)
    {
      const double s = group_sum<kWave>(acc[v]);
      if (lane == 0)
        sred[wave][v] = s;
    }
  __syncthreads();
  for (int v = threadIdx.x; v < NV * NV + NV; v += blockDim.x)
    {
      // entry (i, j) of A^K from the upper triangle, then F^K
      int k = NSYM + v - NV * NV;
      if (v < NV * NV)
        {
          const int i = v / NV, j = v % NV;
          k = i <= j ? symi(i, j) : symi(j, i);
        }
      double s = 0;
      for (int w = 0; w < kElemThreads / kWave; ++w)
        s += sred[w][k];
      if (v < NV * NV)
        a.Ae[(size_t)cell * NV * NV + v] = s;
      else
        a.Fe[(size_t)cell * NV + v - NV * NV] = s;
    }
}

// 3D (scalar coefficient per fine cell, the only 3D kind):
//     A^K = sum_c alpha_c P_c^T L P_c,   F^K = sum_c P_c^T b_c,
// P_c[a][i] = phi_i at vertex a of fine cell c, L the QGauss(2) Q1 element
// matrix of the unit coefficient on the fine box (in LDS).  Row by row of
// Z = alpha L P, only the upper triangle of A^K is accumulated: 44 running
// sums per thread instead of a full 8x8 element matrix, 8x8 vertex values
// and 72 sums (which spilled).
__global__ void __launch_bounds__(kElemThreads) element3d_kernel(ElementArgs a)
{
  constexpr int NV = 8, NQ = 8, NU = NV * (NV + 1) / 2, NACC = NU + NV;
  __shared__ double sL[NV * NV];
  __shared__ double sred[kElemThreads / kWave][NACC];
  // the load's factors: sin(pi x_k) at the two Gauss coordinates of every fine
  // cell along each axis (sin sin sin load), or 1 (constant load: f = rhs_value)
  __shared__ double sfac[3][2 * kElem3MaxNF];

  const int cell = a.cells[blockIdx.x];
  const int nf = a.nf, np = nf + 1;
  const double *bx = a.box + (size_t)cell * 6;
  double h[3], lo[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    {
      lo[k] = bx[2 * k];
      h[k] = (bx[2 * k + 1] - bx[2 * k]) / nf;
    }
  const double jxw = h[0] * h[1] * h[2] / NQ;
  const double gs = 0.5 / sqrt(3.0);
  auto gauss = [&](int q, int m) -> double { return ((q >> m) & 1) ? 0.5 + gs : 0.5 - gs; };
  auto grad = [&](int aa, int k, int q) -> double {
    double g = ((aa >> k) & 1) ? 1.0 : -1.0;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      if (m != k)
        g *= ((aa >> m) & 1) ? gauss(q, m) : 1.0 - gauss(q, m);
    return g / h[k];
  };
  for (int t = threadIdx.x; t < NV * NV; t += blockDim.x)
    {
      const int aa = t / NV, bb = t % NV;
      double v = 0;
      for (int q = 0; q < NQ; ++q)
        for (int k = 0; k < 3; ++k)
          v += grad(aa, k, q) * grad(bb, k, q) * jxw;
      sL[t] = v;
    }
  for (int t = threadIdx.x; t < 3 * 2 * nf; t += blockDim.x)
    {
      const int k = t / (2 * nf), fi = t % (2 * nf) / 2;
      sfac[k][t % (2 * nf)] =
          a.rhs_kind == 1 ? sin(M_PI * (lo[k] + (bx[2 * k + 1] - bx[2 * k]) * ((double)fi / nf) + h[k] * gauss(t & 1, 0)))
                          : 1.0;
    }
  __syncthreads();
  const double fconst = a.rhs_kind == 1 ? 1.0 : a.rhs_value;

  const double *cf = a.coef + a.coef_off[cell];
  const double *phi = a.bases + a.bas_off[cell];
  const size_t nn = (size_t)np * np * np;
  const int ncell = nf * nf * nf;
  double acc[NACC];
#pragma unroll
  for (int v = 0; v < NACC; ++v)
    acc[v] = 0;
  for (int f = threadIdx.x; f < ncell; f += blockDim.x)
    {
      const int fx = f % nf, fy = (f / nf) % nf, fz = f / (nf * nf);
      const double alpha = cf[f];
      double P[NV][NV]; // [vertex][basis]
#pragma unroll
      for (int aa = 0; aa < NV; ++aa)
        {
          const size_t node = (size_t)(fx + (aa & 1)) + (size_t)np * ((fy + ((aa >> 1) & 1)) +
                                                                       (size_t)np * (fz + ((aa >> 2) & 1)));
#pragma unroll
          for (int i = 0; i < NV; ++i)
            P[aa][i] = phi[(size_t)i * nn + node];
        }
      // load vector of the fine cell
      double be[NV];
#pragma unroll
      for (int aa = 0; aa < NV; ++aa)
        be[aa] = 0;
      for (int q = 0; q < NQ; ++q)
        {
          const double fv =
              fconst * sfac[0][2 * fx + (q & 1)] * sfac[1][2 * fy + ((q >> 1) & 1)] * sfac[2][2 * fz + ((q >> 2) & 1)];
#pragma unroll
          for (int aa = 0; aa < NV; ++aa)
            {
              double sh = 1.0;
#pragma unroll
              for (int m = 0; m < 3; ++m)
                sh *= ((aa >> m) & 1) ? gauss(q, m) : 1.0 - gauss(q, m);
              be[aa] += sh * fv * jxw;
            }
        }
#pragma unroll
      for (int aa = 0; aa < NV; ++aa)
        {
          double z[NV];
#pragma unroll
          for (int j = 0; j < NV; ++j)
            z[j] = 0;
#pragma unroll
          for (int bb = 0; bb < NV; ++bb)
            {
              const double l = alpha * sL[aa * NV + bb];
#pragma unroll
              for (int j = 0; j < NV; ++j)
                z[j] += l * P[bb][j];
            }
          int u = 0;
#pragma unroll
          for (int i = 0; i < NV; ++i)
            {
#pragma unroll
              for (int j = i; j < NV; ++j)
                acc[u++] += P[aa][i] * z[j];
              acc[NU + i] += P[aa][i] * be[aa];
            }
        }
    }
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int v = 0; v < NACC; ++v)
    {
      const double s = group_sum<kWave>(acc[v]);
      if (lane == 0)
        sred[wave][v] = s;
    }
  __syncthreads();
  for (int v = threadIdx.x; v < NACC; v += blockDim.x)
    {
      double s = 0;
      for (int w = 0; w < kElemThreads / kWave; ++w)
        s += sred[w][v];
      if (v < NU)
        {
          int i = 0, r = v;
          while (r >= NV - i)
            {
              r -= NV - i;
              ++i;
            }
          const int j = i + r;
          a.Ae[(size_t)cell * NV * NV + i * NV + j] = s;
          a.Ae[(size_t)cell * NV * NV + j * NV + i] = s;
        }
      else
        a.Fe[(size_t)cell * NV + v - NU] = s;
    }
}

// 2D box cells of 64^2 fine cells with per-fine-cell scalar (KIND 0) or
// tensor (KIND 1) coefficients, the same A^K and F^K as element_kernel: one
// wave per cell, lane = fine column fx sweeping the 64 rows, so the bases at
// a row of nodes are loaded once (two nodes per lane) and kept for the next
// row, instead of four times per node.  K_e is the exact QGauss(2) box
// element matrix ka Kxx + kc Kyy + kb Kxy (ka = a rho / 6, kc = c / (6 rho),
// kb = b / 4, rho = h_y / h_x; setup_element_2d's form), the load vector
// sum_q psi_a(q) f(q) jxw.  The upper triangle of A^K and F^K are summed per
// lane over its column, then over the wave in a fixed order.
template <int KIND>
__global__ void __launch_bounds__(256) element_rows_kernel(ElementArgs a)
{
  constexpr int NF = 64, NP = NF + 1, NV = 4, NSYM = 10, NVAL = NSYM + NV;
  __shared__ double ssy[4][2 * NF]; // per wave: sin(pi y) at the two Gauss rows of every fine row
  const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const int cc = blockIdx.x * 4 + wave;
  if (cc >= a.n_cells_class)
    return;
  const int cell = a.cells[cc];
  const double *bx = a.box + (size_t)cell * 4;
  const double hx = (bx[1] - bx[0]) / NF, hy = (bx[3] - bx[2]) / NF, rho = hy / hx;
  const double jxw = hx * hy / 4.0, gs = 0.5 / sqrt(3.0);
  const double *cf = a.coef + a.coef_off[cell];
  const double *phi = a.bases + a.bas_off[cell];
  constexpr size_t nn = (size_t)NP * NP;
  const int fx = lane;
  // sin-sin load: the column's two Gauss x-points, the rows' from the table
  double sxq[2] = {1.0, 1.0};
  if (a.rhs_kind == 1)
    {
      for (int k = 0; k < 2; ++k)
        {
          const double xi = k ? 0.5 + gs : 0.5 - gs;
          sxq[k] = sin(M_PI * (bx[0] + (bx[1] - bx[0]) * ((double)fx / NF) + hx * xi));
          ssy[wave][2 * lane + k] = sin(M_PI * (bx[2] + (bx[3] - bx[2]) * ((double)lane / NF) + hy * xi));
        }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  auto shape = [&](int aa, int q) -> double { // psi_a at Gauss point q (bit 0: x, bit 1: y)
    const double xi = (q & 1) ? 0.5 + gs : 0.5 - gs, eta = (q & 2) ? 0.5 + gs : 0.5 - gs;
    return ((aa & 1) ? xi : 1.0 - xi) * ((aa & 2) ? eta : 1.0 - eta);
  };
  double acc[NVAL];
#pragma unroll
  for (int v = 0; v < NVAL; ++v)
    acc[v] = 0;
  // bases at the nodes (fx, fy) and (fx + 1, fy) of the current bottom row
  double bl[NV], br[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i)
    {
      bl[i] = phi[(size_t)i * nn + fx];
      br[i] = phi[(size_t)i * nn + fx + 1];
    }
#pragma unroll 1
  for (int fy = 0; fy < NF; ++fy)
    {
      double tl[NV], tr[NV];
#pragma unroll
      for (int i = 0; i < NV; ++i)
        {
          tl[i] = phi[(size_t)i * nn + (size_t)(fy + 1) * NP + fx];
          tr[i] = phi[(size_t)i * nn + (size_t)(fy + 1) * NP + fx + 1];
        }
      const int f = fy * NF + fx;
      double ka, kc, kb = 0.0;
      if constexpr (KIND == 0)
        {
          const double al = cf[f];
          ka = al * rho * (1.0 / 6.0);
          kc = al / rho * (1.0 / 6.0);
        }
      else
        {
          ka = cf[3 * f] * rho * (1.0 / 6.0);
          kc = cf[3 * f + 2] / rho * (1.0 / 6.0);
          kb = 0.25 * cf[3 * f + 1];
        }
      // K_e[u][v] (vertex u = (u & 1, u >> 1))
      auto Ke = [&](int u, int v) -> double {
        const double sx = ((u ^ v) & 1) ? -1.0 : 1.0, sy = ((u ^ v) & 2) ? -1.0 : 1.0;
        const double ny = ((u ^ v) & 2) ? 1.0 : 2.0, nx = ((u ^ v) & 1) ? 1.0 : 2.0;
        const double su = (u & 1) ? 1.0 : -1.0, tu = (u & 2) ? 1.0 : -1.0;
        const double sv = (v & 1) ? 1.0 : -1.0, tv = (v & 2) ? 1.0 : -1.0;
        return ka * (sx * ny) + kc * (sy * nx) + (KIND == 0 ? 0.0 : kb * (su * tv + tu * sv));
      };
      double pv[NV][NV]; // [basis][vertex]
#pragma unroll
      for (int i = 0; i < NV; ++i)
        {
          pv[i][0] = bl[i];
          pv[i][1] = br[i];
          pv[i][2] = tl[i];
          pv[i][3] = tr[i];
        }
#pragma unroll
      for (int j = 0; j < NV; ++j)
        {
          double y[NV];
#pragma unroll
          for (int u = 0; u < NV; ++u)
            {
              double t = 0;
#pragma unroll
              for (int v = 0; v < NV; ++v)
                t += Ke(u, v) * pv[j][v];
              y[u] = t;
            }
#pragma unroll
          for (int i = 0; i <= j; ++i)
            {
              double t = 0;
#pragma unroll
              for (int u = 0; u < NV; ++u)
                t += pv[i][u] * y[u];
              acc[symi_of<NV>(i, j)] += t;
            }
        }
      // load vector of the fine cell
      double be[NV] = {0, 0, 0, 0};
#pragma unroll
      for (int q = 0; q < 4; ++q)
        {
          const double fv = a.rhs_kind == 1 ? sxq[q & 1] * ssy[wave][2 * fy + (q >> 1)] : a.rhs_value;
#pragma unroll
          for (int u = 0; u < NV; ++u)
            be[u] += shape(u, q) * fv * jxw;
        }
#pragma unroll
      for (int i = 0; i < NV; ++i)
        {
          double t = 0;
#pragma unroll
          for (int u = 0; u < NV; ++u)
            t += pv[i][u] * be[u];
          acc[NSYM + i] += t;
        }
#pragma unroll
      for (int i = 0; i < NV; ++i)
        {
          bl[i] = tl[i];
          br[i] = tr[i];
        }
    }
#pragma unroll
  for (int v = 0; v < NVAL; ++v)
    acc[v] = group_sum<kWave>(acc[v]);
  if (lane < NV * NV + NV)
    {
      // lane v writes A^K entry v (from the upper triangle) or F^K entry v - 16
      double out = 0;
#pragma unroll
      for (int v = 0; v < NV * NV + NV; ++v)
        if (lane == v)
          {
            int k = NSYM + v - NV * NV;
            if (v < NV * NV)
              {
                const int i = v / NV, j = v % NV;
                k = i <= j ? symi_of<NV>(i, j) : symi_of<NV>(j, i);
              }
            out = acc[k];
          }
      if (lane < NV * NV)
        a.Ae[(size_t)cell * NV * NV + lane] = out;
      else
        a.Fe[(size_t)cell * NV + lane - NV * NV] = out;
    }
}

__global__ void pou_fixup_kernel(int n_cells_class, const int *cells, int nb, size_t nn, double *bases,
                                 const long long *bas_off)
{
  const size_t total = (size_t)n_cells_class * nn;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int cell = cells[t / nn];
      const size_t n = t % nn;
      double *phi = bases + bas_off[cell];
      double s = 0;
      for (int i = 0; i < nb - 1; ++i)
        s += phi[(size_t)i * nn + n];
      phi[(size_t)(nb - 1) * nn + n] = 1.0 - s;
    }
}

} // namespace

hipError_t launch_element(const ElementArgs &a, hipStream_t st)
{
  if (a.n_cells_class == 0)
    return hipSuccess;
  if (a.dim == 2 && a.verts)
    hipLaunchKernelGGL((element_kernel<2, true>), dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  else if (a.verts) // general hexahedra
    hipLaunchKernelGGL((element_kernel<3, true>), dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  else if (a.dim == 2 && a.coef_kind <= 1 && a.nf == 64 && a.nb == 4) // per-fine-cell coefficient, 64^2
    {
      if (a.coef_kind == 0)
        hipLaunchKernelGGL((element_rows_kernel<0>), dim3((a.n_cells_class + 3) / 4), dim3(256), 0, st, a);
      else
        hipLaunchKernelGGL((element_rows_kernel<1>), dim3((a.n_cells_class + 3) / 4), dim3(256), 0, st, a);
    }
  else if (a.dim == 2 && a.coef_kind <= 1) // per-fine-cell coefficient
    hipLaunchKernelGGL((element_kernel<2, false, false>), dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  else if (a.dim == 2)
    hipLaunchKernelGGL((element_kernel<2, false>), dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  else if (a.coef_kind == 0 && a.nf <= kElem3MaxNF) // scalar per fine cell
    hipLaunchKernelGGL(element3d_kernel, dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  else
    hipLaunchKernelGGL((element_kernel<3, false>), dim3(a.n_cells_class), dim3(kElemThreads), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_pou_fixup(int n_cells_class, const int *cells, int dim, int nf, double *bases,
                            const long long *bas_off, hipStream_t st)
{
  if (n_cells_class == 0)
    return hipSuccess;
  const size_t nn = dim == 2 ? (size_t)(nf + 1) * (nf + 1) : (size_t)(nf + 1) * (nf + 1) * (nf + 1);
  const size_t total = (size_t)n_cells_class * nn;
  size_t grid = (total + 255) / 256;
  if (grid > 65536)
    grid = 65536;
  hipLaunchKernelGGL(pou_fixup_kernel, dim3((unsigned)grid), dim3(256), 0, st, n_cells_class, cells, 1 << dim, nn,
                     bases, bas_off);
  return hipGetLastError();
}

} // namespace msfem
