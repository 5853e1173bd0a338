/*
 * msfem_oracle.c -- CPU FP64 restatement of the MsFEM basis pipeline of
 * konsim83/MPI-Parallel-Multiscale-Diffusion-FEM.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: only tests/, the smoke
 * check and the CPU-baseline leg of bench.py load it.  The product library
 * never links, loads or calls it.
 *
 * What it restates (reference files, relative to the reference root):
 *   - Q1 basis Dirichlet data by monomial inversion
 *       include/coefficients/basis_q1.tpp:26-75 (ctor), :86-113 (value)
 *   - local fine mesh on one coarse cell, Q1 stiffness with QGauss(2), load
 *       include/base/diffusion_problem_basis.tpp:92-99 (make_grid),
 *       :161-242 (assemble_system)
 *   - 2^dim basis problems with zero rhs, Dirichlet data N_i on the cell
 *       boundary, solved (a) exactly by banded Cholesky on K_II (the values any
 *       converged solve must reach) and (b) by the reference's own recipe:
 *       condensed full system, CG + SSOR(omega), absolute tolerance, iteration
 *       cap  -- diffusion_problem_basis.tpp:295-317, :440-474
 *   - coarse element matrix A_ij = phi_i^T K phi_j and load F_i = phi_i^T b
 *       diffusion_problem_basis.tpp:247-285
 *   - coarse assembly with Dirichlet (ids 0,2,4) and Neumann (ids 1,3,5) data
 *       include/base/diffusion_problem_ms.tpp:108-156, :161-255,
 *       include/coefficients/dirichlet_bc.tpp:17-43, neumann_bc.tpp:17-41
 *   - coarse solve (direct; the reference uses AMG-CG to 1e-12)
 *       diffusion_problem_ms.tpp:260-296
 *   - fine reconstruction u|_K = sum_i w_i phi_i
 *       diffusion_problem_basis.tpp:354-377
 *   - general (non-box) coarse cells, quadrilaterals (2D) and hexahedra (3D):
 *       the fine mesh of GridGenerator::general_cell + refine_global
 *       (diffusion_problem_basis.tpp:92-99), FE_Q(1) with the bi/trilinear
 *       mapping and QGauss(2) on every fine cell (:161-242), Dirichlet data N_i
 *       of BasisQ1 at the mapped boundary nodes (basis_q1.tpp:26-46, :86-96)
 *   - the reference default coefficient MatrixCoeff (with PI_D, k=57,
 *       scale 0.9999) in long double: include/coefficients/matrix_coeff.hpp:45-49,
 *       matrix_coeff.tpp:17-91, coefficients.h:21
 *
 * Conventions shared with the product library:
 *   nodes of a cell's fine grid are lexicographic, x fastest:
 *     n = X + (nf+1)*(Y + (nf+1)*Z);   fine cells likewise with nf.
 *   vertex v of a (coarse or fine) cell has offsets (v&1, (v>>1)&1, (v>>2)&1)
 *   (deal.II lexicographic vertex order, x fastest).
 *   quadrature point q of a fine cell: QGauss(2) tensor points, x fastest.
 *   symmetric tensors are stored as upper triangles:
 *     2D (a11,a12,a22); 3D (a11,a12,a13,a22,a23,a33).
 *   coefficient tiles (one per coarse cell, contiguous, AoS per fine cell):
 *     SCALAR_CELL [nf^d]; TENSOR_CELL [nf^d][ns]; SCALAR_QP [nf^d][2^d];
 *     TENSOR_QP [nf^d][2^d][ns].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_COEF_SCALAR_CELL 0
#define ORC_COEF_TENSOR_CELL 1
#define ORC_COEF_SCALAR_QP 2
#define ORC_COEF_TENSOR_QP 3

#define ORC_RHS_CONST 0
#define ORC_RHS_SINSIN 1

#define ORC_SOLVER_DIRECT 0
#define ORC_SOLVER_SSOR_PCG 1

/* PI_D of the reference (include/coefficients/coefficients.h:21): not pi. */
static const long double ORC_PI_D_L = 3.14592653509793218403L;
static const double ORC_PI_D = 3.14592653509793218403;

static double gauss_pt(int k) /* QGauss<1>(2) on [0,1] */
{
  const double s = 0.5 / sqrt(3.0);
  return k == 0 ? 0.5 - s : 0.5 + s;
}

static int ipow(int b, int e)
{
  int r = 1;
  while (e-- > 0)
    r *= b;
  return r;
}

int orc_n_sym(int dim) { return dim == 2 ? 3 : 6; }

int orc_coef_stride(int dim, int kind)
{
  const int nq = 1 << dim, ns = orc_n_sym(dim);
  switch (kind)
    {
      case ORC_COEF_SCALAR_CELL: return 1;
      case ORC_COEF_TENSOR_CELL: return ns;
      case ORC_COEF_SCALAR_QP: return nq;
      case ORC_COEF_TENSOR_QP: return nq * ns;
      default: return -1;
    }
}

/* symmetric index in the upper-triangle storage */
static int sym_idx(int dim, int i, int j)
{
  if (i > j)
    {
      int t = i;
      i = j;
      j = t;
    }
  if (dim == 2)
    return i == 0 ? j : 2;
  /* 3D: (0,0)0 (0,1)1 (0,2)2 (1,1)3 (1,2)4 (2,2)5 */
  if (i == 0)
    return j;
  if (i == 1)
    return 2 + j;
  return 5;
}

/* ------------------------------------------------------------------ */
/* BasisQ1 by monomial inversion  (basis_q1.tpp:26-75, :86-113)        */
/* ------------------------------------------------------------------ */
static void monomials(int dim, const double *p, double *m)
{
  if (dim == 2)
    {
      m[0] = 1;
      m[1] = p[0];
      m[2] = p[1];
      m[3] = p[0] * p[1];
    }
  else
    {
      m[0] = 1;
      m[1] = p[0];
      m[2] = p[1];
      m[3] = p[2];
      m[4] = p[0] * p[1];
      m[5] = p[1] * p[2];
      m[6] = p[0] * p[2];
      m[7] = p[0] * p[1] * p[2];
    }
}

/* verts: [2^d][d]; coeff: [2^d][2^d] row-major, coeff[m][i] = coefficient of
 * monomial m in N_i (columns of the inverse, basis_q1.tpp:45-46). Returns 0 on
 * success, 1 if the monomial matrix is singular. */
int orc_basis_q1_coeffs(int dim, const double *verts, double *coeff)
{
  const int n = 1 << dim;
  double a[8][16];
  for (int i = 0; i < n; ++i)
    {
      double m[8];
      monomials(dim, verts + i * dim, m);
      for (int j = 0; j < n; ++j)
        {
          a[i][j]     = m[j];
          a[i][n + j] = (i == j) ? 1.0 : 0.0;
        }
    }
  for (int c = 0; c < n; ++c)
    {
      int piv = c;
      for (int r = c + 1; r < n; ++r)
        if (fabs(a[r][c]) > fabs(a[piv][c]))
          piv = r;
      if (a[piv][c] == 0.0)
        return 1;
      if (piv != c)
        for (int j = 0; j < 2 * n; ++j)
          {
            double t    = a[c][j];
            a[c][j]     = a[piv][j];
            a[piv][j]   = t;
          }
      const double d = a[c][c];
      for (int j = 0; j < 2 * n; ++j)
        a[c][j] /= d;
      for (int r = 0; r < n; ++r)
        if (r != c && a[r][c] != 0.0)
          {
            const double f = a[r][c];
            for (int j = 0; j < 2 * n; ++j)
              a[r][j] -= f * a[c][j];
          }
    }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      coeff[i * n + j] = a[i][n + j];
  return 0;
}

double orc_basis_q1_value(int dim, const double *coeff, int index, const double *p)
{
  const int n = 1 << dim;
  double m[8];
  monomials(dim, p, m);
  double v = 0;
  for (int k = 0; k < n; ++k)
    v += coeff[k * n + index] * m[k];
  return v;
}

/* box: [d][2] (lo, hi per axis) -> deal.II-ordered vertices [2^d][d] */
void orc_box_vertices(int dim, const double *box, double *verts)
{
  for (int v = 0; v < (1 << dim); ++v)
    for (int k = 0; k < dim; ++k)
      verts[v * dim + k] = ((v >> k) & 1) ? box[2 * k + 1] : box[2 * k];
}

/* Multilinear map of a general cell with vertices verts[2^d][d] in deal.II
 * order (vertex v sits at the parametric corner (v&1, (v>>1)&1, (v>>2)&1)) at
 * the parametric point xi[d] in [0,1]^d: bilinear for a quadrilateral,
 * trilinear for a hexahedron. */
static void cell_map(int dim, const double *verts, const double *xi, double *x)
{
  for (int k = 0; k < dim; ++k)
    x[k] = 0;
  for (int v = 0; v < (1 << dim); ++v)
    {
      double w = 1;
      for (int m = 0; m < dim; ++m)
        w *= ((v >> m) & 1) ? xi[m] : 1 - xi[m];
      for (int k = 0; k < dim; ++k)
        x[k] += w * verts[v * dim + k];
    }
}

/* Fine node I[d] of a general cell refined nf times per side.
 * GridGenerator::general_cell + refine_global (diffusion_problem_basis.tpp:
 * 92-99) put each new vertex at an edge midpoint, at the mean of a face's four
 * vertices or (3D) at the mean of a cell's eight vertices; on a multilinear
 * cell these are the map's values at the parametric midpoints, and a child of
 * a multilinear cell is the image of its parametric sub-box under the parent's
 * map (the restriction is again multilinear), so the refined vertices are the
 * map's values on the uniform parametric grid. */
void orc_cell_node(int dim, const double *verts, int nf, const int *I, double *x)
{
  double xi[3];
  for (int k = 0; k < dim; ++k)
    xi[k] = (double)I[k] / nf;
  cell_map(dim, verts, xi, x);
}

/* The same for a quadrilateral (2D): fine node (X, Y). */
void orc_quad_node(const double *quad, int nf, int X, int Y, double *x)
{
  const int I[2] = {X, Y};
  orc_cell_node(2, quad, nf, I, x);
}

/* ------------------------------------------------------------------ */
/* Reference default coefficient in long double                        */
/* (matrix_coeff.hpp:45-49, matrix_coeff.tpp:17-63)                     */
/* ------------------------------------------------------------------ */
void orc_matrix_coeff_ref(int dim, const double *p, double *A_sym)
{
  const long double alpha = ORC_PI_D_L / 3, beta = ORC_PI_D_L / 6, gamma = ORC_PI_D_L / 4;
  long double R[3][3] = {{0}};
  if (dim == 2)
    {
      R[0][0] = cosl(alpha);
      R[0][1] = sinl(alpha);
      R[1][0] = -sinl(alpha);
      R[1][1] = cosl(alpha);
    }
  else
    {
      R[0][0] = cosl(alpha) * cosl(gamma) - sinl(alpha) * cosl(beta) * sinl(gamma);
      R[0][1] = -cosl(alpha) * sinl(gamma) - sinl(alpha) * cosl(beta) * cosl(gamma);
      R[0][2] = sinl(alpha) * sinl(beta);
      R[1][0] = sinl(alpha) * cosl(gamma) + cosl(alpha) * cosl(beta) * sinl(gamma);
      R[1][1] = -sinl(alpha) * sinl(gamma) + cosl(alpha) * cosl(beta) * cosl(gamma);
      R[1][2] = -cosl(alpha) * sinl(beta);
      R[2][0] = sinl(beta) * sinl(gamma);
      R[2][1] = sinl(beta) * cosl(gamma);
      R[2][2] = cosl(beta);
    }
  const long double k = 57, s = 0.9999L;
  const long double dval =
    1.0L - s * (0.5L * sinl(2 * ORC_PI_D_L * k * (long double)p[0]) +
                0.5L * sinl(2 * ORC_PI_D_L * k * (long double)p[1]));
  for (int i = 0; i < dim; ++i)
    for (int j = i; j < dim; ++j)
      {
        long double a = 0;
        for (int m = 0; m < dim; ++m)
          a += R[i][m] * dval * R[j][m];
        A_sym[sym_idx(dim, i, j)] = (double)a;
      }
}

/* ------------------------------------------------------------------ */
/* Local assembly (diffusion_problem_basis.tpp:161-242)                */
/* K in structured stencil form: Kst[node][3^d], offset index          */
/* s = (ox+1) + 3(oy+1) + 9(oz+1).                                      */
/* ------------------------------------------------------------------ */
static void qp_tensor(int dim, int kind, const double *cdata, int q, double A[3][3])
{
  const int ns = orc_n_sym(dim);
  double s[6];
  switch (kind)
    {
      case ORC_COEF_SCALAR_CELL:
        for (int i = 0; i < ns; ++i)
          s[i] = 0;
        s[sym_idx(dim, 0, 0)] = s[sym_idx(dim, 1, 1)] = cdata[0];
        if (dim == 3)
          s[sym_idx(dim, 2, 2)] = cdata[0];
        break;
      case ORC_COEF_TENSOR_CELL:
        for (int i = 0; i < ns; ++i)
          s[i] = cdata[i];
        break;
      case ORC_COEF_SCALAR_QP:
        for (int i = 0; i < ns; ++i)
          s[i] = 0;
        s[sym_idx(dim, 0, 0)] = s[sym_idx(dim, 1, 1)] = cdata[q];
        if (dim == 3)
          s[sym_idx(dim, 2, 2)] = cdata[q];
        break;
      default: /* TENSOR_QP */
        for (int i = 0; i < ns; ++i)
          s[i] = cdata[q * ns + i];
        break;
    }
  for (int i = 0; i < dim; ++i)
    for (int j = 0; j < dim; ++j)
      A[i][j] = s[sym_idx(dim, i, j)];
}

static double rhs_eval(int dim, int kind, double value, const double *x)
{
  if (kind == ORC_RHS_SINSIN)
    {
      double f = 1.0;
      for (int k = 0; k < dim; ++k)
        f *= sin(M_PI * x[k]);
      return f;
    }
  return value;
}

/* Element matrix and load of one fine box cell of sides h[] with lower corner
 * x0[], QGauss(2), FE_Q(1)  (diffusion_problem_basis.tpp:207-222). */
void orc_fine_element(int dim, const double *x0, const double *h, int coef_kind,
                      const double *cdata, int rhs_kind, double rhs_value,
                      double *Ke /* [nv][nv] */, double *be /* [nv] */)
{
  const int nv = 1 << dim, nq = 1 << dim;
  double jxw = 1.0;
  for (int k = 0; k < dim; ++k)
    jxw *= h[k];
  jxw /= nq; /* QGauss(2) weights are 1/2 per axis */
  memset(Ke, 0, sizeof(double) * nv * nv);
  memset(be, 0, sizeof(double) * nv);
  for (int q = 0; q < nq; ++q)
    {
      double xi[3], xq[3];
      for (int k = 0; k < dim; ++k)
        {
          xi[k] = gauss_pt((q >> k) & 1);
          xq[k] = x0[k] + h[k] * xi[k];
        }
      double A[3][3];
      qp_tensor(dim, coef_kind, cdata, q, A);
      const double f = rhs_eval(dim, rhs_kind, rhs_value, xq);
      double val[8], grad[8][3];
      for (int a = 0; a < nv; ++a)
        {
          double v = 1.0;
          for (int k = 0; k < dim; ++k)
            v *= ((a >> k) & 1) ? xi[k] : 1.0 - xi[k];
          val[a] = v;
          for (int k = 0; k < dim; ++k)
            {
              double g = ((a >> k) & 1) ? 1.0 : -1.0;
              for (int m = 0; m < dim; ++m)
                if (m != k)
                  g *= ((a >> m) & 1) ? xi[m] : 1.0 - xi[m];
              grad[a][k] = g / h[k];
            }
        }
      for (int a = 0; a < nv; ++a)
        {
          for (int b = 0; b < nv; ++b)
            {
              double s = 0;
              for (int i = 0; i < dim; ++i)
                for (int j = 0; j < dim; ++j)
                  s += grad[a][i] * A[i][j] * grad[b][j];
              Ke[a * nv + b] += s * jxw;
            }
          be[a] += val[a] * f * jxw;
        }
    }
}

int orc_n_nodes(int dim, int nf) { return ipow(nf + 1, dim); }

/* Kst: [(nf+1)^d][3^d], b: [(nf+1)^d]; box [d][2]; tile: the cell's coefficient tile */
void orc_local_assemble(int dim, int nf, const double *box, int coef_kind, const double *tile,
                        int rhs_kind, double rhs_value, double *Kst, double *b)
{
  const int np = nf + 1, ns3 = ipow(3, dim), nv = 1 << dim;
  const int nn = ipow(np, dim), ncell = ipow(nf, dim);
  const int stride = orc_coef_stride(dim, coef_kind);
  double h[3];
  for (int k = 0; k < dim; ++k)
    h[k] = (box[2 * k + 1] - box[2 * k]) / nf;
  memset(Kst, 0, sizeof(double) * (size_t)nn * ns3);
  memset(b, 0, sizeof(double) * (size_t)nn);
  for (int c = 0; c < ncell; ++c)
    {
      int ci[3] = {c % nf, (c / nf) % nf, dim == 3 ? c / (nf * nf) : 0};
      double x0[3];
      for (int k = 0; k < dim; ++k)
        x0[k] = box[2 * k] + (box[2 * k + 1] - box[2 * k]) * ((double)ci[k] / nf);
      double Ke[64], be[8];
      orc_fine_element(dim, x0, h, coef_kind, tile + (size_t)c * stride, rhs_kind, rhs_value, Ke, be);
      int node[8];
      for (int a = 0; a < nv; ++a)
        {
          int X = ci[0] + (a & 1), Y = ci[1] + ((a >> 1) & 1), Z = ci[2] + ((a >> 2) & 1);
          node[a] = X + np * (Y + np * Z);
        }
      for (int a = 0; a < nv; ++a)
        {
          for (int bb = 0; bb < nv; ++bb)
            {
              int ox = (bb & 1) - (a & 1), oy = ((bb >> 1) & 1) - ((a >> 1) & 1);
              int oz = ((bb >> 2) & 1) - ((a >> 2) & 1);
              int s = (ox + 1) + 3 * (oy + 1) + (dim == 3 ? 9 * (oz + 1) : 0);
              Kst[(size_t)node[a] * ns3 + s] += Ke[a * nv + bb];
            }
          b[node[a]] += be[a];
        }
    }
}

/* Element matrix and load of one fine quadrilateral with vertices fv[4][2]
 * (deal.II order): FE_Q(1) with the bilinear mapping (deal.II's default
 * MappingQ1), QGauss(2) (diffusion_problem_basis.tpp:163-222).  At each
 * quadrature point J_kl = sum_a fv[a][k] d psi_a / d xi_l, grad psi_a =
 * J^{-T} grad_xi psi_a, JxW = det J / 4; coefficient and f at the mapped point. */
void orc_fine_element_quad(const double *fv, int coef_kind, const double *cdata, int rhs_kind, double rhs_value,
                           double *Ke /* [4][4] */, double *be /* [4] */)
{
  memset(Ke, 0, sizeof(double) * 16);
  memset(be, 0, sizeof(double) * 4);
  for (int q = 0; q < 4; ++q)
    {
      const double xi = gauss_pt(q & 1), eta = gauss_pt((q >> 1) & 1);
      double val[4], gr[4][2];
      for (int a = 0; a < 4; ++a)
        {
          const double fx = (a & 1) ? xi : 1 - xi, fy = (a & 2) ? eta : 1 - eta;
          val[a] = fx * fy;
          gr[a][0] = ((a & 1) ? 1.0 : -1.0) * fy;
          gr[a][1] = ((a & 2) ? 1.0 : -1.0) * fx;
        }
      double J[2][2] = {{0, 0}, {0, 0}}, xq[2] = {0, 0};
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 2; ++k)
          {
            xq[k] += val[a] * fv[2 * a + k];
            for (int l = 0; l < 2; ++l)
              J[k][l] += fv[2 * a + k] * gr[a][l];
          }
      const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      const double Ji[2][2] = {{J[1][1] / det, -J[0][1] / det}, {-J[1][0] / det, J[0][0] / det}};
      double g[4][2];
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 2; ++k)
          g[a][k] = Ji[0][k] * gr[a][0] + Ji[1][k] * gr[a][1];
      double A[3][3];
      qp_tensor(2, coef_kind, cdata, q, A);
      const double f = rhs_eval(2, rhs_kind, rhs_value, xq), jxw = det * 0.25;
      for (int a = 0; a < 4; ++a)
        {
          for (int b = 0; b < 4; ++b)
            {
              double t = 0;
              for (int i = 0; i < 2; ++i)
                for (int j = 0; j < 2; ++j)
                  t += g[a][i] * A[i][j] * g[b][j];
              Ke[a * 4 + b] += t * jxw;
            }
          be[a] += val[a] * f * jxw;
        }
    }
}

/* Local system of a general quadrilateral (2D), as orc_local_assemble. */
void orc_local_assemble_quad(int nf, const double *quad, int coef_kind, const double *tile, int rhs_kind,
                             double rhs_value, double *Kst, double *b)
{
  const int np = nf + 1, nn = np * np;
  const int stride = orc_coef_stride(2, coef_kind);
  memset(Kst, 0, sizeof(double) * (size_t)nn * 9);
  memset(b, 0, sizeof(double) * (size_t)nn);
  for (int c = 0; c < nf * nf; ++c)
    {
      const int cx = c % nf, cy = c / nf;
      double fv[8], Ke[16], be[4];
      for (int a = 0; a < 4; ++a)
        orc_quad_node(quad, nf, cx + (a & 1), cy + ((a >> 1) & 1), fv + 2 * a);
      orc_fine_element_quad(fv, coef_kind, tile + (size_t)c * stride, rhs_kind, rhs_value, Ke, be);
      for (int a = 0; a < 4; ++a)
        {
          const int na = (cx + (a & 1)) + np * (cy + ((a >> 1) & 1));
          for (int bb = 0; bb < 4; ++bb)
            {
              const int ox = (bb & 1) - (a & 1), oy = ((bb >> 1) & 1) - ((a >> 1) & 1);
              Kst[(size_t)na * 9 + (ox + 1) + 3 * (oy + 1)] += Ke[a * 4 + bb];
            }
          b[na] += be[a];
        }
    }
}

/* Element matrix and load of one fine hexahedron with vertices fv[8][3]
 * (deal.II order): FE_Q(1) with the trilinear mapping (MappingQ1), QGauss(2)
 * (diffusion_problem_basis.tpp:163-222).  At each quadrature point
 * J_kl = sum_a fv[a][k] d psi_a / d xi_l, grad psi_a = J^{-T} grad_xi psi_a
 * (J^{-1} by cofactors), JxW = det J / 8; coefficient and f at the mapped point. */
void orc_fine_element_hex(const double *fv, int coef_kind, const double *cdata, int rhs_kind, double rhs_value,
                          double *Ke /* [8][8] */, double *be /* [8] */)
{
  memset(Ke, 0, sizeof(double) * 64);
  memset(be, 0, sizeof(double) * 8);
  for (int q = 0; q < 8; ++q)
    {
      double xi[3];
      for (int k = 0; k < 3; ++k)
        xi[k] = gauss_pt((q >> k) & 1);
      double val[8], gr[8][3];
      for (int a = 0; a < 8; ++a)
        {
          double f1[3], d1[3];
          for (int k = 0; k < 3; ++k)
            {
              f1[k] = ((a >> k) & 1) ? xi[k] : 1 - xi[k];
              d1[k] = ((a >> k) & 1) ? 1.0 : -1.0;
            }
          val[a]   = f1[0] * f1[1] * f1[2];
          gr[a][0] = d1[0] * f1[1] * f1[2];
          gr[a][1] = f1[0] * d1[1] * f1[2];
          gr[a][2] = f1[0] * f1[1] * d1[2];
        }
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, xq[3] = {0, 0, 0};
      for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k)
          {
            xq[k] += val[a] * fv[3 * a + k];
            for (int l = 0; l < 3; ++l)
              J[k][l] += fv[3 * a + k] * gr[a][l];
          }
      /* cofactor matrix C_kl = (-1)^(k+l) minor_kl; J^{-1} = C^T / det */
      double Cf[3][3];
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          {
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3, l1 = (l + 1) % 3, l2 = (l + 2) % 3;
            Cf[k][l] = J[k1][l1] * J[k2][l2] - J[k1][l2] * J[k2][l1];
          }
      const double det = J[0][0] * Cf[0][0] + J[0][1] * Cf[0][1] + J[0][2] * Cf[0][2];
      /* (J^{-T} g)_k = sum_l (J^{-1})_lk g_l = sum_l C_kl g_l / det */
      double g[8][3];
      for (int a = 0; a < 8; ++a)
        for (int k = 0; k < 3; ++k)
          g[a][k] = (Cf[k][0] * gr[a][0] + Cf[k][1] * gr[a][1] + Cf[k][2] * gr[a][2]) / det;
      double A[3][3];
      qp_tensor(3, coef_kind, cdata, q, A);
      const double f = rhs_eval(3, rhs_kind, rhs_value, xq), jxw = det * 0.125;
      for (int a = 0; a < 8; ++a)
        {
          for (int b = 0; b < 8; ++b)
            {
              double t = 0;
              for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j)
                  t += g[a][i] * A[i][j] * g[b][j];
              Ke[a * 8 + b] += t * jxw;
            }
          be[a] += val[a] * f * jxw;
        }
    }
}

/* Local system of a general hexahedron (3D), as orc_local_assemble. */
void orc_local_assemble_hex(int nf, const double *hex, int coef_kind, const double *tile, int rhs_kind,
                            double rhs_value, double *Kst, double *b)
{
  const int np = nf + 1, nn = np * np * np;
  const int stride = orc_coef_stride(3, coef_kind);
  memset(Kst, 0, sizeof(double) * (size_t)nn * 27);
  memset(b, 0, sizeof(double) * (size_t)nn);
  for (int c = 0; c < nf * nf * nf; ++c)
    {
      const int ci[3] = {c % nf, (c / nf) % nf, c / (nf * nf)};
      double fv[24], Ke[64], be[8];
      for (int a = 0; a < 8; ++a)
        {
          const int I[3] = {ci[0] + (a & 1), ci[1] + ((a >> 1) & 1), ci[2] + ((a >> 2) & 1)};
          orc_cell_node(3, hex, nf, I, fv + 3 * a);
        }
      orc_fine_element_hex(fv, coef_kind, tile + (size_t)c * stride, rhs_kind, rhs_value, Ke, be);
      for (int a = 0; a < 8; ++a)
        {
          const int na = (ci[0] + (a & 1)) + np * ((ci[1] + ((a >> 1) & 1)) + np * (ci[2] + ((a >> 2) & 1)));
          for (int bb = 0; bb < 8; ++bb)
            {
              const int ox = (bb & 1) - (a & 1), oy = ((bb >> 1) & 1) - ((a >> 1) & 1);
              const int oz = ((bb >> 2) & 1) - ((a >> 2) & 1);
              Kst[(size_t)na * 27 + (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)] += Ke[a * 8 + bb];
            }
          b[na] += be[a];
        }
    }
}

/* y = K x on all nodes */
void orc_stencil_apply(int dim, int nf, const double *Kst, const double *x, double *y)
{
  const int np = nf + 1, ns3 = ipow(3, dim);
  const int nn = ipow(np, dim);
  for (int n = 0; n < nn; ++n)
    {
      int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
      double s = 0;
      for (int t = 0; t < ns3; ++t)
        {
          int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
          int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
          if (XX < 0 || XX > nf || YY < 0 || YY > nf || ZZ < 0 || ZZ > (dim == 3 ? nf : 0))
            continue;
          s += Kst[(size_t)n * ns3 + t] * x[XX + np * (YY + np * ZZ)];
        }
      y[n] = s;
    }
}

static int is_boundary_node(int dim, int nf, int n)
{
  const int np = nf + 1;
  int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 1;
  if (X == 0 || X == nf || Y == 0 || Y == nf)
    return 1;
  if (dim == 3 && (Z == 0 || Z == nf))
    return 1;
  return 0;
}

/* Dirichlet data g_i at all nodes (interior entries = 0): BasisQ1 by monomial
 * inversion evaluated at the fine node coordinates (diffusion_problem_basis.tpp:119-135).
 * verts: a general cell's vertices [2^d][d], or NULL for the box. */
static void dirichlet_data(int dim, int nf, const double *box, const double *gverts, int index, double *g)
{
  double verts[24], coeff[64];
  if (gverts)
    memcpy(verts, gverts, sizeof(double) * (size_t)(1 << dim) * dim);
  else
    orc_box_vertices(dim, box, verts);
  orc_basis_q1_coeffs(dim, verts, coeff);
  const int np = nf + 1, nn = ipow(np, dim);
  for (int n = 0; n < nn; ++n)
    {
      if (!is_boundary_node(dim, nf, n))
        {
          g[n] = 0;
          continue;
        }
      int I[3] = {n % np, (n / np) % np, dim == 3 ? n / (np * np) : 0};
      double p[3];
      if (gverts)
        orc_cell_node(dim, gverts, nf, I, p);
      else
        for (int k = 0; k < dim; ++k)
          p[k] = box[2 * k] + (box[2 * k + 1] - box[2 * k]) * ((double)I[k] / nf);
      g[n] = orc_basis_q1_value(dim, coeff, index, p);
    }
}

void orc_dirichlet_data(int dim, int nf, const double *box, int index, double *g)
{
  dirichlet_data(dim, nf, box, NULL, index, g);
}

/* The same at the nodes of a general quadrilateral's fine grid (2D). */
void orc_dirichlet_data_quad(int nf, const double *quad, int index, double *g)
{
  dirichlet_data(2, nf, NULL, quad, index, g);
}

/* ... and of a general hexahedron's (3D). */
void orc_dirichlet_data_hex(int nf, const double *hex, int index, double *g)
{
  dirichlet_data(3, nf, NULL, hex, index, g);
}

/* ------------------------------------------------------------------ */
/* Banded Cholesky (direct solves: the values a converged CG reaches)  */
/* ------------------------------------------------------------------ */
typedef struct
{
  int n, bw;
  double *L; /* row i, column j in [i-bw, i] at L[i*(bw+1) + j-i+bw] */
} band_t;

static int band_factor(band_t *B)
{
  const int n = B->n, bw = B->bw, w = bw + 1;
  double *L = B->L;
  for (int i = 0; i < n; ++i)
    {
      int j0 = i - bw < 0 ? 0 : i - bw;
      for (int j = j0; j <= i; ++j)
        {
          double s  = L[(size_t)i * w + j - i + bw];
          int    k0 = j0 > j - bw ? j0 : j - bw;
          for (int k = k0; k < j; ++k)
            s -= L[(size_t)i * w + k - i + bw] * L[(size_t)j * w + k - j + bw];
          if (j == i)
            {
              if (!(s > 0))
                return 1;
              L[(size_t)i * w + bw] = sqrt(s);
            }
          else
            L[(size_t)i * w + j - i + bw] = s / L[(size_t)j * w + bw];
        }
    }
  return 0;
}

static void band_solve(const band_t *B, double *x)
{
  const int n = B->n, bw = B->bw, w = bw + 1;
  const double *L = B->L;
  for (int i = 0; i < n; ++i)
    {
      double s  = x[i];
      int    k0 = i - bw < 0 ? 0 : i - bw;
      for (int k = k0; k < i; ++k)
        s -= L[(size_t)i * w + k - i + bw] * x[k];
      x[i] = s / L[(size_t)i * w + bw];
    }
  for (int i = n - 1; i >= 0; --i)
    {
      double s  = x[i];
      int    k1 = i + bw > n - 1 ? n - 1 : i + bw;
      for (int k = i + 1; k <= k1; ++k)
        s -= L[(size_t)k * w + i - k + bw] * x[k];
      x[i] = s / L[(size_t)i * w + bw];
    }
}

/* Basis problems by direct solve: phi[i][n], i < 2^d.  Returns 0 or 1 (not SPD). */
static int local_bases_direct(int dim, int nf, const double *box, const double *verts, const double *Kst,
                              double *phi)
{
  const int np = nf + 1, m = nf - 1, ns3 = ipow(3, dim), nb = 1 << dim;
  const int nn = ipow(np, dim), ni = ipow(m, dim);
  const int bw = dim == 2 ? m + 1 : m * m + m + 1;
  int *imap = (int *)malloc(sizeof(int) * nn);
  int *inode = (int *)malloc(sizeof(int) * (ni > 0 ? ni : 1));
  int cnt = 0;
  for (int n = 0; n < nn; ++n)
    {
      if (is_boundary_node(dim, nf, n))
        imap[n] = -1;
      else
        {
          imap[n]      = cnt;
          inode[cnt++] = n;
        }
    }
  band_t B;
  B.n  = ni;
  B.bw = bw;
  B.L  = (double *)calloc((size_t)ni * (bw + 1) + 1, sizeof(double));
  for (int i = 0; i < ni; ++i)
    {
      int n = inode[i];
      int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
      for (int t = 0; t < ns3; ++t)
        {
          int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
          int nn2 = (X + ox) + np * ((Y + oy) + np * (Z + oz));
          int j   = imap[nn2];
          if (j < 0 || j > i)
            continue;
          B.L[(size_t)i * (bw + 1) + j - i + bw] = Kst[(size_t)n * ns3 + t];
        }
    }
  int rc = band_factor(&B);
  double *g = (double *)malloc(sizeof(double) * nn);
  double *x = (double *)malloc(sizeof(double) * (ni > 0 ? ni : 1));
  for (int ib = 0; ib < nb && rc == 0; ++ib)
    {
      dirichlet_data(dim, nf, box, verts, ib, g);
      for (int i = 0; i < ni; ++i)
        {
          int n = inode[i];
          int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
          double s = 0;
          for (int t = 0; t < ns3; ++t)
            {
              int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
              int nn2 = (X + ox) + np * ((Y + oy) + np * (Z + oz));
              if (imap[nn2] < 0)
                s += Kst[(size_t)n * ns3 + t] * g[nn2];
            }
          x[i] = -s;
        }
      band_solve(&B, x);
      double *out = phi + (size_t)ib * nn;
      for (int n = 0; n < nn; ++n)
        out[n] = imap[n] < 0 ? g[n] : x[imap[n]];
    }
  free(B.L);
  free(g);
  free(x);
  free(imap);
  free(inode);
  return rc;
}

int orc_local_bases_direct(int dim, int nf, const double *box, const double *Kst, double *phi)
{
  return local_bases_direct(dim, nf, box, NULL, Kst, phi);
}

int orc_local_bases_direct_quad(int nf, const double *quad, const double *Kst, double *phi)
{
  return local_bases_direct(2, nf, NULL, quad, Kst, phi);
}

int orc_local_bases_direct_hex(int nf, const double *hex, const double *Kst, double *phi)
{
  return local_bases_direct(3, nf, NULL, hex, Kst, phi);
}

/* ------------------------------------------------------------------ */
/* The reference's own recipe: condense, CG + SSOR(omega), absolute    */
/* tolerance, iteration cap (diffusion_problem_basis.tpp:295-317,      */
/* :450-465).  Returns the number of bases that did not converge.      */
/* ------------------------------------------------------------------ */
static int local_bases_ssor_pcg(int dim, int nf, const double *box, const double *verts, const double *Kst,
                                double omega, double abs_tol, int max_iter, double *phi, int *iters)
{
  const int np = nf + 1, ns3 = ipow(3, dim), nb = 1 << dim, nn = ipow(np, dim);
  const int center = dim == 2 ? 4 : 13;
  double *A  = (double *)malloc(sizeof(double) * (size_t)nn * ns3); /* condensed */
  double *g  = (double *)malloc(sizeof(double) * nn);
  double *rh = (double *)malloc(sizeof(double) * nn);
  double *r  = (double *)malloc(sizeof(double) * nn);
  double *z  = (double *)malloc(sizeof(double) * nn);
  double *p  = (double *)malloc(sizeof(double) * nn);
  double *q  = (double *)malloc(sizeof(double) * nn);
  char   *bd = (char *)malloc(nn);
  for (int n = 0; n < nn; ++n)
    bd[n] = (char)is_boundary_node(dim, nf, n);
  /* condensed matrix: rows/columns of constrained nodes zeroed, diagonal kept */
  for (int n = 0; n < nn; ++n)
    {
      int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
      for (int t = 0; t < ns3; ++t)
        {
          int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
          int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
          double v = Kst[(size_t)n * ns3 + t];
          if (XX < 0 || XX > nf || YY < 0 || YY > nf || ZZ < 0 || ZZ > (dim == 3 ? nf : 0))
            v = 0;
          else if (t != center && (bd[n] || bd[XX + np * (YY + np * ZZ)]))
            v = 0;
          A[(size_t)n * ns3 + t] = v;
        }
    }
  int failures = 0;
  for (int ib = 0; ib < nb; ++ib)
    {
      double *x = phi + (size_t)ib * nn;
      dirichlet_data(dim, nf, box, verts, ib, g);
      /* condensed rhs (diffusion_problem_basis.tpp:455,461): -K_IB g on free
       * rows, diag*g on constrained rows */
      for (int n = 0; n < nn; ++n)
        {
          if (bd[n])
            {
              rh[n] = Kst[(size_t)n * ns3 + center] * g[n];
              continue;
            }
          int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
          double s = 0;
          for (int t = 0; t < ns3; ++t)
            {
              int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
              int nn2 = (X + ox) + np * ((Y + oy) + np * (Z + oz));
              if (bd[nn2])
                s += Kst[(size_t)n * ns3 + t] * g[nn2];
            }
          rh[n] = -s;
        }
      for (int n = 0; n < nn; ++n)
        {
          x[n] = 0;
          r[n] = rh[n];
        }
      double rn = 0;
      for (int n = 0; n < nn; ++n)
        rn += r[n] * r[n];
      rn      = sqrt(rn);
      int it  = 0;
      double rz = 0;
      while (rn > abs_tol && it < max_iter)
        {
          /* z = M^{-1} r, SSOR: (D + wL) y = w(2-w) r ; y <- D y ; (D + wU) z = y */
          for (int n = 0; n < nn; ++n)
            {
              int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
              double s = omega * (2 - omega) * r[n];
              for (int t = 0; t < center; ++t)
                {
                  int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
                  int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
                  if (XX < 0 || XX > nf || YY < 0 || YY > nf || ZZ < 0)
                    continue;
                  s -= omega * A[(size_t)n * ns3 + t] * z[XX + np * (YY + np * ZZ)];
                }
              z[n] = s / A[(size_t)n * ns3 + center];
            }
          for (int n = 0; n < nn; ++n)
            z[n] *= A[(size_t)n * ns3 + center];
          for (int n = nn - 1; n >= 0; --n)
            {
              int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
              double s = z[n];
              for (int t = center + 1; t < ns3; ++t)
                {
                  int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
                  int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
                  if (XX < 0 || XX > nf || YY < 0 || YY > nf || ZZ > (dim == 3 ? nf : 0))
                    continue;
                  s -= omega * A[(size_t)n * ns3 + t] * z[XX + np * (YY + np * ZZ)];
                }
              z[n] = s / A[(size_t)n * ns3 + center];
            }
          double rz_new = 0;
          for (int n = 0; n < nn; ++n)
            rz_new += r[n] * z[n];
          if (it == 0)
            for (int n = 0; n < nn; ++n)
              p[n] = z[n];
          else
            {
              const double beta = rz_new / rz;
              for (int n = 0; n < nn; ++n)
                p[n] = z[n] + beta * p[n];
            }
          rz = rz_new;
          /* q = A p */
          for (int n = 0; n < nn; ++n)
            {
              int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
              double s = 0;
              for (int t = 0; t < ns3; ++t)
                {
                  int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
                  int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
                  if (XX < 0 || XX > nf || YY < 0 || YY > nf || ZZ < 0 || ZZ > (dim == 3 ? nf : 0))
                    continue;
                  s += A[(size_t)n * ns3 + t] * p[XX + np * (YY + np * ZZ)];
                }
              q[n] = s;
            }
          double pq = 0;
          for (int n = 0; n < nn; ++n)
            pq += p[n] * q[n];
          const double alpha = rz / pq;
          rn = 0;
          for (int n = 0; n < nn; ++n)
            {
              x[n] += alpha * p[n];
              r[n] -= alpha * q[n];
              rn += r[n] * r[n];
            }
          rn = sqrt(rn);
          ++it;
        }
      /* distribute (diffusion_problem_basis.tpp:308) */
      for (int n = 0; n < nn; ++n)
        if (bd[n])
          x[n] = g[n];
      if (iters)
        iters[ib] = it;
      if (rn > abs_tol)
        ++failures;
    }
  free(A);
  free(g);
  free(rh);
  free(r);
  free(z);
  free(p);
  free(q);
  free(bd);
  return failures;
}

int orc_local_bases_ssor_pcg(int dim, int nf, const double *box, const double *Kst, double omega,
                             double abs_tol, int max_iter, double *phi, int *iters)
{
  return local_bases_ssor_pcg(dim, nf, box, NULL, Kst, omega, abs_tol, max_iter, phi, iters);
}

/* A_ij = phi_i^T K phi_j, F_i = phi_i^T b  (diffusion_problem_basis.tpp:247-285) */
void orc_element(int dim, int nf, const double *Kst, const double *b, const double *phi, double *Ae,
                 double *Fe)
{
  const int nb = 1 << dim, nn = orc_n_nodes(dim, nf);
  double *tmp = (double *)malloc(sizeof(double) * nn);
  for (int j = 0; j < nb; ++j)
    {
      orc_stencil_apply(dim, nf, Kst, phi + (size_t)j * nn, tmp);
      for (int i = 0; i < nb; ++i)
        {
          double s = 0;
          const double *pi = phi + (size_t)i * nn;
          for (int n = 0; n < nn; ++n)
            s += pi[n] * tmp[n];
          Ae[i * nb + j] = s;
        }
    }
  for (int i = 0; i < nb; ++i)
    {
      double s = 0;
      const double *pi = phi + (size_t)i * nn;
      for (int n = 0; n < nn; ++n)
        s += pi[n] * b[n];
      Fe[i] = s;
    }
  free(tmp);
}

/* One coarse cell: DiffusionProblemBasis<dim>::run() (diffusion_problem_basis.tpp:440-474).
 * phi (optional, [2^d][(nf+1)^d]) receives the bases; iters (optional, [2^d]).
 * Returns 0 on success. */
int orc_cell_build(int dim, int nf, const double *box, int coef_kind, const double *tile, int rhs_kind,
                   double rhs_value, int solver, double omega, double abs_tol, int max_iter, double *phi,
                   double *Ae, double *Fe, int *iters)
{
  const int nn = orc_n_nodes(dim, nf), ns3 = ipow(3, dim), nb = 1 << dim;
  double *Kst = (double *)malloc(sizeof(double) * (size_t)nn * ns3);
  double *b   = (double *)malloc(sizeof(double) * nn);
  double *ph  = phi ? phi : (double *)malloc(sizeof(double) * (size_t)nb * nn);
  orc_local_assemble(dim, nf, box, coef_kind, tile, rhs_kind, rhs_value, Kst, b);
  int rc;
  if (solver == ORC_SOLVER_DIRECT)
    {
      rc = orc_local_bases_direct(dim, nf, box, Kst, ph);
      if (iters)
        for (int i = 0; i < nb; ++i)
          iters[i] = 0;
    }
  else
    rc = orc_local_bases_ssor_pcg(dim, nf, box, Kst, omega, abs_tol, max_iter, ph, iters);
  orc_element(dim, nf, Kst, b, ph, Ae, Fe);
  free(Kst);
  free(b);
  if (!phi)
    free(ph);
  return rc;
}

/* One general coarse cell (quadrilateral in 2D, hexahedron in 3D), as
 * orc_cell_build; verts [2^d][d]. */
static int cell_build_general(int dim, int nf, const double *verts, int coef_kind, const double *tile,
                              int rhs_kind, double rhs_value, int solver, double omega, double abs_tol,
                              int max_iter, double *phi, double *Ae, double *Fe, int *iters)
{
  const int nn = orc_n_nodes(dim, nf), ns3 = ipow(3, dim), nb = 1 << dim;
  double *Kst = (double *)malloc(sizeof(double) * (size_t)nn * ns3);
  double *b   = (double *)malloc(sizeof(double) * nn);
  double *ph  = phi ? phi : (double *)malloc(sizeof(double) * (size_t)nb * nn);
  if (dim == 2)
    orc_local_assemble_quad(nf, verts, coef_kind, tile, rhs_kind, rhs_value, Kst, b);
  else
    orc_local_assemble_hex(nf, verts, coef_kind, tile, rhs_kind, rhs_value, Kst, b);
  int rc;
  if (solver == ORC_SOLVER_DIRECT)
    {
      rc = local_bases_direct(dim, nf, NULL, verts, Kst, ph);
      if (iters)
        for (int i = 0; i < nb; ++i)
          iters[i] = 0;
    }
  else
    rc = local_bases_ssor_pcg(dim, nf, NULL, verts, Kst, omega, abs_tol, max_iter, ph, iters);
  orc_element(dim, nf, Kst, b, ph, Ae, Fe);
  free(Kst);
  free(b);
  if (!phi)
    free(ph);
  return rc;
}

/* Exported form of cell_build_general (general quadrilateral or hexahedron). */
int orc_cell_build_general(int dim, int nf, const double *verts, int coef_kind, const double *tile, int rhs_kind,
                           double rhs_value, int solver, double omega, double abs_tol, int max_iter, double *phi,
                           double *Ae, double *Fe, int *iters)
{
  return cell_build_general(dim, nf, verts, coef_kind, tile, rhs_kind, rhs_value, solver, omega, abs_tol, max_iter,
                            phi, Ae, Fe, iters);
}

int orc_cell_build_quad(int nf, const double *quad, int coef_kind, const double *tile, int rhs_kind,
                        double rhs_value, int solver, double omega, double abs_tol, int max_iter, double *phi,
                        double *Ae, double *Fe, int *iters)
{
  return cell_build_general(2, nf, quad, coef_kind, tile, rhs_kind, rhs_value, solver, omega, abs_tol, max_iter,
                            phi, Ae, Fe, iters);
}

/* Many cells, OpenMP over cells (each cell is independent, as the reference's
 * per-rank cell loop diffusion_problem_ms.tpp:84-87).
 * box [n][d][2]; nf [n]; coef: concatenated tiles in cell order;
 * bases (optional): concatenated [2^d][(nf+1)^d] blocks in cell order.
 * Returns the number of cells whose solve failed. */
int orc_batch_build(int dim, int n_cells, const double *box, const int *nf, int coef_kind,
                    const double *coef, int rhs_kind, double rhs_value, int solver, double omega,
                    double abs_tol, int max_iter, int n_threads, double *bases, double *Ae, double *Fe,
                    int *iters)
{
  const int nb = 1 << dim, stride = orc_coef_stride(dim, coef_kind);
  size_t *coff = (size_t *)malloc(sizeof(size_t) * (n_cells + 1));
  size_t *boff = (size_t *)malloc(sizeof(size_t) * (n_cells + 1));
  coff[0] = boff[0] = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      coff[c + 1] = coff[c] + (size_t)ipow(nf[c], dim) * stride;
      boff[c + 1] = boff[c] + (size_t)nb * orc_n_nodes(dim, nf[c]);
    }
  int fails = 0;
#ifdef _OPENMP
  if (n_threads > 0)
    omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fails)
#endif
  for (int c = 0; c < n_cells; ++c)
    {
      int rc = orc_cell_build(dim, nf[c], box + (size_t)c * 2 * dim, coef_kind, coef + coff[c], rhs_kind,
                              rhs_value, solver, omega, abs_tol, max_iter, bases ? bases + boff[c] : NULL,
                              Ae + (size_t)c * nb * nb, Fe + (size_t)c * nb, iters ? iters + (size_t)c * nb : NULL);
      if (rc)
        fails += 1;
    }
  free(coff);
  free(boff);
  return fails;
}

/* The same for general cells: quadrilaterals (2D, verts [n][4][2]) or
 * hexahedra (3D, verts [n][8][3]). */
int orc_batch_build_general(int dim, int n_cells, const double *verts, const int *nf, int coef_kind,
                            const double *coef, int rhs_kind, double rhs_value, int solver, double omega,
                            double abs_tol, int max_iter, int n_threads, double *bases, double *Ae, double *Fe,
                            int *iters)
{
  const int nb = 1 << dim, stride = orc_coef_stride(dim, coef_kind);
  size_t *coff = (size_t *)malloc(sizeof(size_t) * (n_cells + 1));
  size_t *boff = (size_t *)malloc(sizeof(size_t) * (n_cells + 1));
  coff[0] = boff[0] = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      coff[c + 1] = coff[c] + (size_t)ipow(nf[c], dim) * stride;
      boff[c + 1] = boff[c] + (size_t)nb * orc_n_nodes(dim, nf[c]);
    }
  int fails = 0;
#ifdef _OPENMP
  if (n_threads > 0)
    omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : fails)
#endif
  for (int c = 0; c < n_cells; ++c)
    {
      int rc = cell_build_general(dim, nf[c], verts + (size_t)c * nb * dim, coef_kind, coef + coff[c], rhs_kind,
                                  rhs_value, solver, omega, abs_tol, max_iter, bases ? bases + boff[c] : NULL,
                                  Ae + (size_t)c * nb * nb, Fe + (size_t)c * nb,
                                  iters ? iters + (size_t)c * nb : NULL);
      if (rc)
        fails += 1;
    }
  free(coff);
  free(boff);
  return fails;
}

int orc_batch_build_quad(int n_cells, const double *quad, const int *nf, int coef_kind, const double *coef,
                         int rhs_kind, double rhs_value, int solver, double omega, double abs_tol, int max_iter,
                         int n_threads, double *bases, double *Ae, double *Fe, int *iters)
{
  return orc_batch_build_general(2, n_cells, quad, nf, coef_kind, coef, rhs_kind, rhs_value, solver, omega,
                                 abs_tol, max_iter, n_threads, bases, Ae, Fe, iters);
}

int orc_max_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Coarse system on the unit square/cube with nc cells per side        */
/* (diffusion_problem_ms.tpp:93-103 mesh, :108-156 constraints,        */
/*  :161-255 assembly, :260-296 solve)                                  */
/* bc_kind 0: homogeneous Dirichlet on every face.                      */
/* bc_kind 1: the reference default: Dirichlet g=(x-1/2)^2+(y-1/2)^2 on  */
/*   boundary ids 0,2,4 (x=0,y=0,z=0), Neumann h=cos(2 PI_D x)cos(2 PI_D y)*/
/*   on ids 1,3,5 (x=1,y=1,z=1), QGauss(2) on faces.                    */
/* Kc: [(nc+1)^d][3^d] unconstrained; Fc: [(nc+1)^d] incl. Neumann;      */
/* cmask/cval: constrained nodes and their values.                       */
/* ------------------------------------------------------------------ */
static double dirichlet_g(const double *x) { return (x[0] - 0.5) * (x[0] - 0.5) + (x[1] - 0.5) * (x[1] - 0.5); }
static double neumann_h(const double *x) { return cos(2 * ORC_PI_D * x[0]) * cos(2 * ORC_PI_D * x[1]); }

void orc_coarse_assemble(int dim, int nc, const double *Ae, const double *Fe, int bc_kind, double *Kc, double *Fc,
                         unsigned char *cmask, double *cval)
{
  const int np = nc + 1, ns3 = ipow(3, dim), nv = 1 << dim;
  const int nn = ipow(np, dim), ncell = ipow(nc, dim);
  const double H = 1.0 / nc;
  memset(Kc, 0, sizeof(double) * (size_t)nn * ns3);
  memset(Fc, 0, sizeof(double) * nn);
  for (int c = 0; c < ncell; ++c)
    {
      int ci[3] = {c % nc, (c / nc) % nc, dim == 3 ? c / (nc * nc) : 0};
      double fe[8];
      for (int a = 0; a < nv; ++a)
        fe[a] = Fe[(size_t)c * nv + a];
      if (bc_kind == 1)
        {
          /* faces on x=1 (id 1), y=1 (id 3), z=1 (id 5) */
          for (int k = 0; k < dim; ++k)
            {
              if (ci[k] != nc - 1)
                continue;
              const int nq = 1 << (dim - 1);
              for (int q = 0; q < nq; ++q)
                {
                  double xi[3], xq[3];
                  int t = 0;
                  double w = 1.0;
                  for (int m = 0; m < dim; ++m)
                    {
                      if (m == k)
                        {
                          xi[m] = 1.0;
                          continue;
                        }
                      xi[m] = gauss_pt((q >> t) & 1);
                      ++t;
                      w *= 0.5 * H;
                    }
                  for (int m = 0; m < dim; ++m)
                    xq[m] = (ci[m] + xi[m]) * H;
                  const double hv = neumann_h(xq);
                  for (int a = 0; a < nv; ++a)
                    {
                      double v = 1.0;
                      for (int m = 0; m < dim; ++m)
                        v *= ((a >> m) & 1) ? xi[m] : 1.0 - xi[m];
                      fe[a] += hv * v * w;
                    }
                }
            }
        }
      int node[8];
      for (int a = 0; a < nv; ++a)
        {
          int X = ci[0] + (a & 1), Y = ci[1] + ((a >> 1) & 1), Z = ci[2] + ((a >> 2) & 1);
          node[a] = X + np * (Y + np * Z);
        }
      for (int a = 0; a < nv; ++a)
        {
          for (int b = 0; b < nv; ++b)
            {
              int ox = (b & 1) - (a & 1), oy = ((b >> 1) & 1) - ((a >> 1) & 1), oz = ((b >> 2) & 1) - ((a >> 2) & 1);
              int s = (ox + 1) + 3 * (oy + 1) + (dim == 3 ? 9 * (oz + 1) : 0);
              Kc[(size_t)node[a] * ns3 + s] += Ae[(size_t)c * nv * nv + a * nv + b];
            }
          Fc[node[a]] += fe[a];
        }
    }
  for (int n = 0; n < nn; ++n)
    {
      int I[3] = {n % np, (n / np) % np, dim == 3 ? n / (np * np) : 0};
      double x[3];
      for (int m = 0; m < dim; ++m)
        x[m] = I[m] * H;
      int con = 0;
      for (int m = 0; m < dim; ++m)
        {
          if (I[m] == 0)
            con = 1;
          if (bc_kind == 0 && I[m] == nc)
            con = 1;
        }
      cmask[n] = (unsigned char)con;
      cval[n]  = con ? (bc_kind == 1 ? dirichlet_g(x) : 0.0) : 0.0;
    }
}

/* Solve the constrained coarse system exactly: u_D = g_D, A_FF u_F = F_F - A_FD g_D. */
int orc_coarse_solve(int dim, int nc, const double *Kc, const double *Fc, const unsigned char *cmask,
                     const double *cval, double *u)
{
  const int np = nc + 1, ns3 = ipow(3, dim), nn = ipow(np, dim);
  int *fmap = (int *)malloc(sizeof(int) * nn);
  int nfree = 0;
  for (int n = 0; n < nn; ++n)
    fmap[n] = cmask[n] ? -1 : nfree++;
  int bw = 0;
  for (int n = 0; n < nn; ++n)
    {
      if (fmap[n] < 0)
        continue;
      int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
      for (int t = 0; t < ns3; ++t)
        {
          int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
          int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
          if (XX < 0 || XX > nc || YY < 0 || YY > nc || ZZ < 0 || ZZ > (dim == 3 ? nc : 0))
            continue;
          int j = fmap[XX + np * (YY + np * ZZ)];
          if (j >= 0 && abs(j - fmap[n]) > bw)
            bw = abs(j - fmap[n]);
        }
    }
  band_t B;
  B.n  = nfree;
  B.bw = bw;
  B.L  = (double *)calloc((size_t)nfree * (bw + 1) + 1, sizeof(double));
  double *rhs = (double *)malloc(sizeof(double) * (nfree > 0 ? nfree : 1));
  for (int n = 0; n < nn; ++n)
    {
      int i = fmap[n];
      if (i < 0)
        continue;
      int X = n % np, Y = (n / np) % np, Z = dim == 3 ? n / (np * np) : 0;
      double s = Fc[n];
      for (int t = 0; t < ns3; ++t)
        {
          int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = dim == 3 ? t / 9 - 1 : 0;
          int XX = X + ox, YY = Y + oy, ZZ = Z + oz;
          if (XX < 0 || XX > nc || YY < 0 || YY > nc || ZZ < 0 || ZZ > (dim == 3 ? nc : 0))
            continue;
          int n2 = XX + np * (YY + np * ZZ);
          int j  = fmap[n2];
          if (j < 0)
            s -= Kc[(size_t)n * ns3 + t] * cval[n2];
          else if (j <= i)
            B.L[(size_t)i * (bw + 1) + j - i + bw] = Kc[(size_t)n * ns3 + t];
        }
      rhs[i] = s;
    }
  int rc = band_factor(&B);
  if (rc == 0)
    band_solve(&B, rhs);
  for (int n = 0; n < nn; ++n)
    u[n] = fmap[n] < 0 ? cval[n] : rhs[fmap[n]];
  free(B.L);
  free(rhs);
  free(fmap);
  return rc;
}

/* u|_K = sum_i w_i phi_i  (diffusion_problem_basis.tpp:354-377) */
void orc_reconstruct(int dim, int nf, const double *phi, const double *w, double *u)
{
  const int nb = 1 << dim, nn = orc_n_nodes(dim, nf);
  for (int n = 0; n < nn; ++n)
    {
      double s = 0;
      for (int i = 0; i < nb; ++i)
        s += w[i] * phi[(size_t)i * nn + n];
      u[n] = s;
    }
}
