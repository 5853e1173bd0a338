// Synthetic inputs for the MsFEM engine: counter-based random fields and the
// reference's analytic coefficient, generated on the host in FP64 so that every
// consumer (GPU engine, CPU oracle) sees bit-identical inputs.
//
// Reference coefficient: include/coefficients/matrix_coeff.hpp:45-49,
// include/coefficients/matrix_coeff.tpp:17-91, include/coefficients/coefficients.h:21.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "msfem.h"

namespace {

constexpr double kPiD = 3.14592653509793218403; // PI_D of the reference: not pi
constexpr double kTwoM32 = 1.0 / 4294967296.0;

inline void philox_round(uint32_t c[4], const uint32_t k[2])
{
  const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c[1] ^ k[0], n1 = lo1, n2 = hi0 ^ c[3] ^ k[1], n3 = lo0;
  c[0] = n0;
  c[1] = n1;
  c[2] = n2;
  c[3] = n3;
}

inline void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  uint32_t k[2] = {key[0], key[1]};
  for (int r = 0; r < 10; ++r)
    {
      if (r > 0)
        {
          k[0] += 0x9E3779B9u;
          k[1] += 0xBB67AE85u;
        }
      philox_round(c, k);
    }
  std::memcpy(out, c, sizeof(c));
}

inline double unit_open(uint32_t v) { return ((double)v + 0.5) * kTwoM32; } // in (0,1)

// Two independent N(0,1) draws (Box-Muller) and two spare uniforms from one Philox block.
struct Draw
{
  double z0, z1, u2, u3;
};

inline Draw draw(uint64_t seed, uint32_t a, uint32_t b, uint32_t stream)
{
  const uint32_t ctr[4] = {a, b, stream, 0u};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox(ctr, key, o);
  const double u0 = unit_open(o[0]), u1 = unit_open(o[1]);
  const double rad = std::sqrt(-2.0 * std::log(u0));
  const double ang = 2.0 * M_PI * u1;
  return {rad * std::cos(ang), rad * std::sin(ang), unit_open(o[2]), unit_open(o[3])};
}

inline int ipow(int b, int e)
{
  int r = 1;
  while (e-- > 0)
    r *= b;
  return r;
}

inline double gauss_pt(int k)
{
  const double s = 0.5 / std::sqrt(3.0);
  return k == 0 ? 0.5 - s : 0.5 + s;
}

// Symmetric tensor R diag(d..d) R^T of the reference default coefficient.
void matrix_coeff(int dim, const double *p, double *A_sym)
{
  const double alpha = kPiD / 3, beta = kPiD / 6, gamma = kPiD / 4;
  double R[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (dim == 2)
    {
      R[0][0] = std::cos(alpha);
      R[0][1] = std::sin(alpha);
      R[1][0] = -std::sin(alpha);
      R[1][1] = std::cos(alpha);
    }
  else
    {
      R[0][0] = std::cos(alpha) * std::cos(gamma) - std::sin(alpha) * std::cos(beta) * std::sin(gamma);
      R[0][1] = -std::cos(alpha) * std::sin(gamma) - std::sin(alpha) * std::cos(beta) * std::cos(gamma);
      R[0][2] = std::sin(alpha) * std::sin(beta);
      R[1][0] = std::sin(alpha) * std::cos(gamma) + std::cos(alpha) * std::cos(beta) * std::sin(gamma);
      R[1][1] = -std::sin(alpha) * std::sin(gamma) + std::cos(alpha) * std::cos(beta) * std::cos(gamma);
      R[1][2] = -std::cos(alpha) * std::sin(beta);
      R[2][0] = std::sin(beta) * std::sin(gamma);
      R[2][1] = std::sin(beta) * std::cos(gamma);
      R[2][2] = std::cos(beta);
    }
  const double k = 57, s = 0.9999;
  const double d = 1.0 - s * (0.5 * std::sin(2 * kPiD * k * p[0]) + 0.5 * std::sin(2 * kPiD * k * p[1]));
  int idx = 0;
  for (int i = 0; i < dim; ++i)
    for (int j = i; j < dim; ++j)
      {
        double a = 0;
        for (int m = 0; m < dim; ++m)
          a += R[i][m] * d * R[j][m];
        A_sym[idx++] = a;
      }
}

bool valid_dim(int dim) { return dim == 2 || dim == 3; }

} // namespace

extern "C" {

void msfem_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

int msfem_gen_unit_boxes(int dim, int nc, double *box)
{
  if (!valid_dim(dim) || nc <= 0 || !box)
    return MSFEM_ERR_ARG;
  const int n = ipow(nc, dim);
  for (int c = 0; c < n; ++c)
    {
      int ci[3] = {c % nc, (c / nc) % nc, c / (nc * nc)};
      for (int k = 0; k < dim; ++k)
        {
          box[(size_t)c * 2 * dim + 2 * k] = (double)ci[k] / nc;
          box[(size_t)c * 2 * dim + 2 * k + 1] = (double)(ci[k] + 1) / nc;
        }
    }
  return MSFEM_OK;
}

int msfem_gen_coef_reference(int dim, int n_cells, const int32_t *n_fine, const double *box, double *coef)
{
  if (!valid_dim(dim) || n_cells < 0 || !n_fine || !box || !coef)
    return MSFEM_ERR_ARG;
  const int ns = dim == 2 ? 3 : 6, nq = 1 << dim;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int nf = n_fine[c];
      const double *bx = box + (size_t)c * 2 * dim;
      double h[3];
      for (int k = 0; k < dim; ++k)
        h[k] = (bx[2 * k + 1] - bx[2 * k]) / nf;
      const int ncell = ipow(nf, dim);
      for (int f = 0; f < ncell; ++f)
        {
          const int fi[3] = {f % nf, (f / nf) % nf, f / (nf * nf)};
          for (int q = 0; q < nq; ++q)
            {
              double p[3] = {0, 0, 0};
              for (int k = 0; k < dim; ++k)
                p[k] = bx[2 * k] + (bx[2 * k + 1] - bx[2 * k]) * ((double)fi[k] / nf) + h[k] * gauss_pt((q >> k) & 1);
              matrix_coeff(dim, p, coef + off);
              off += ns;
            }
        }
    }
  return MSFEM_OK;
}

int msfem_gen_perturbed_quads(int nc, double amp, uint64_t seed, double *vertices)
{
  if (nc <= 0 || !(amp >= 0) || !(amp < 0.25) || !vertices)
    return MSFEM_ERR_ARG;
  const int np = nc + 1;
  std::vector<double> xy((size_t)np * np * 2);
  for (int j = 0; j < np; ++j)
    for (int i = 0; i < np; ++i)
      {
        double x = (double)i / nc, y = (double)j / nc;
        if (i > 0 && i < nc && j > 0 && j < nc)
          {
            const Draw d = draw(seed, (uint32_t)(j * np + i), 0u, 2u);
            x += amp * (2.0 * d.u2 - 1.0) / nc;
            y += amp * (2.0 * d.u3 - 1.0) / nc;
          }
        xy[2 * ((size_t)j * np + i)] = x;
        xy[2 * ((size_t)j * np + i) + 1] = y;
      }
  for (int c = 0; c < nc * nc; ++c)
    {
      const int cx = c % nc, cy = c / nc;
      for (int v = 0; v < 4; ++v)
        {
          const size_t n = (size_t)(cy + (v >> 1)) * np + cx + (v & 1);
          vertices[(size_t)c * 8 + 2 * v] = xy[2 * n];
          vertices[(size_t)c * 8 + 2 * v + 1] = xy[2 * n + 1];
        }
    }
  return MSFEM_OK;
}

int msfem_gen_coef_reference_quad(int n_cells, const int32_t *n_fine, const double *vertices, double *coef)
{
  if (n_cells < 0 || !n_fine || !vertices || !coef)
    return MSFEM_ERR_ARG;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int nf = n_fine[c];
      const double *v = vertices + (size_t)c * 8;
      for (int f = 0; f < nf * nf; ++f)
        for (int q = 0; q < 4; ++q)
          {
            const double xi = ((f % nf) + gauss_pt(q & 1)) / nf, eta = ((f / nf) + gauss_pt(q >> 1)) / nf;
            double p[2];
            for (int k = 0; k < 2; ++k)
              p[k] = (1 - xi) * (1 - eta) * v[k] + xi * (1 - eta) * v[2 + k] + (1 - xi) * eta * v[4 + k] +
                     xi * eta * v[6 + k];
            matrix_coeff(2, p, coef + off);
            off += 3;
          }
    }
  return MSFEM_OK;
}

int msfem_gen_perturbed_hexes(int nc, double amp, uint64_t seed, double *vertices)
{
  if (nc <= 0 || !(amp >= 0) || !(amp < 1.0 / 6.0) || !vertices)
    return MSFEM_ERR_ARG;
  const int np = nc + 1;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  std::vector<double> xyz((size_t)np * np * np * 3);
  for (int l = 0; l < np; ++l)
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i)
        {
          const size_t n = ((size_t)l * np + j) * np + i;
          double x[3] = {(double)i / nc, (double)j / nc, (double)l / nc};
          if (i > 0 && i < nc && j > 0 && j < nc && l > 0 && l < nc)
            {
              const uint32_t ctr[4] = {(uint32_t)n, 0u, 3u, 0u};
              uint32_t u[4];
              philox(ctr, key, u);
              for (int k = 0; k < 3; ++k)
                x[k] += amp * (2.0 * unit_open(u[k]) - 1.0) / nc;
            }
          for (int k = 0; k < 3; ++k)
            xyz[3 * n + k] = x[k];
        }
  for (int c = 0; c < nc * nc * nc; ++c)
    {
      const int cx = c % nc, cy = (c / nc) % nc, cz = c / (nc * nc);
      for (int v = 0; v < 8; ++v)
        {
          const size_t n = ((size_t)(cz + (v >> 2)) * np + cy + ((v >> 1) & 1)) * np + cx + (v & 1);
          for (int k = 0; k < 3; ++k)
            vertices[(size_t)c * 24 + 3 * v + k] = xyz[3 * n + k];
        }
    }
  return MSFEM_OK;
}

int msfem_gen_coef_reference_hex(int n_cells, const int32_t *n_fine, const double *vertices, double *coef)
{
  if (n_cells < 0 || !n_fine || !vertices || !coef)
    return MSFEM_ERR_ARG;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int nf = n_fine[c];
      const double *v = vertices + (size_t)c * 24;
      for (int f = 0; f < nf * nf * nf; ++f)
        {
          const int fi[3] = {f % nf, (f / nf) % nf, f / (nf * nf)};
          for (int q = 0; q < 8; ++q)
            {
              double xi[3], p[3] = {0, 0, 0};
              for (int k = 0; k < 3; ++k)
                xi[k] = (fi[k] + gauss_pt((q >> k) & 1)) / nf;
              for (int w = 0; w < 8; ++w)
                {
                  double wt = 1;
                  for (int m = 0; m < 3; ++m)
                    wt *= ((w >> m) & 1) ? xi[m] : 1 - xi[m];
                  for (int k = 0; k < 3; ++k)
                    p[k] += wt * v[3 * w + k];
                }
              matrix_coeff(3, p, coef + off);
              off += 6;
            }
        }
    }
  return MSFEM_OK;
}

int msfem_gen_coef_periodic(int n_cells, const int32_t *n_fine, const double *box, double eps, double *coef)
{
  if (n_cells < 0 || !n_fine || !box || !coef || !(eps > 0))
    return MSFEM_ERR_ARG;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int nf = n_fine[c];
      const double *bx = box + (size_t)c * 4;
      const double hx = (bx[1] - bx[0]) / nf, hy = (bx[3] - bx[2]) / nf;
      for (int f = 0; f < nf * nf; ++f)
        {
          const int fx = f % nf, fy = f / nf;
          for (int q = 0; q < 4; ++q)
            {
              const double x = bx[0] + (bx[1] - bx[0]) * ((double)fx / nf) + hx * gauss_pt(q & 1);
              const double y = bx[2] + (bx[3] - bx[2]) * ((double)fy / nf) + hy * gauss_pt(q >> 1);
              coef[off++] = 2.0 + std::sin(2.0 * M_PI * x / eps) * std::cos(2.0 * M_PI * y / eps);
            }
        }
    }
  return MSFEM_OK;
}

int msfem_gen_coef_lognormal(int dim, int n_cells, const int32_t *n_fine, uint64_t seed, double clip, double *coef)
{
  if (!valid_dim(dim) || n_cells < 0 || !n_fine || !coef || !(clip > 0))
    return MSFEM_ERR_ARG;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int ncell = ipow(n_fine[c], dim);
      for (int f = 0; f < ncell; ++f)
        {
          const double g = draw(seed, (uint32_t)f, (uint32_t)c, 0u).z0;
          coef[off++] = std::exp(std::min(clip, std::max(-clip, g)));
        }
    }
  return MSFEM_OK;
}

int msfem_gen_resolution_classes(int nc, uint64_t seed, double sigma, const int32_t classes[4], int32_t *n_fine,
                                 double *contrast)
{
  if (nc <= 0 || !classes || !n_fine || !(sigma > 0))
    return MSFEM_ERR_ARG;
  const int n = nc * nc;
  std::vector<double> con(n);
  for (int c = 0; c < n; ++c)
    {
      double zmin = 1e300, zmax = -1e300;
      for (int j = 0; j < 256; ++j)
        {
          const Draw d = draw(seed, (uint32_t)j, (uint32_t)c, 1u);
          zmin = std::min(zmin, std::min(d.z0, d.z1));
          zmax = std::max(zmax, std::max(d.z0, d.z1));
        }
      con[c] = sigma * (zmax - zmin); // log(max lambda / min lambda) over the probe
    }
  std::vector<double> s(con);
  std::sort(s.begin(), s.end());
  const double q1 = s[(size_t)n / 4], q2 = s[(size_t)n / 2], q3 = s[(size_t)3 * n / 4];
  for (int c = 0; c < n; ++c)
    {
      const int k = (con[c] >= q1) + (con[c] >= q2) + (con[c] >= q3);
      n_fine[c] = classes[k];
      if (contrast)
        contrast[c] = con[c];
    }
  return MSFEM_OK;
}

int msfem_gen_coef_tensor(int n_cells, const int32_t *n_fine, uint64_t seed, double sigma, double *coef)
{
  if (n_cells < 0 || !n_fine || !coef || !(sigma > 0))
    return MSFEM_ERR_ARG;
  size_t off = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      const int nf = n_fine[c];
      for (int f = 0; f < nf * nf; ++f)
        {
          const Draw d = draw(seed, (uint32_t)f, (uint32_t)c, 0u);
          const double l1 = std::exp(sigma * d.z0), l2 = std::exp(sigma * d.z1);
          const double th = M_PI * d.u2;
          const double cs = std::cos(th), sn = std::sin(th);
          coef[off++] = l1 * cs * cs + l2 * sn * sn;
          coef[off++] = (l1 - l2) * cs * sn;
          coef[off++] = l1 * sn * sn + l2 * cs * cs;
        }
    }
  return MSFEM_OK;
}

int msfem_gen_grf_log(int n_cells, int nf, uint64_t seed, double corr_len, double clip, float *g)
{
  if (n_cells < 0 || nf <= 0 || !g || !(corr_len > 0) || !(clip > 0))
    return MSFEM_ERR_ARG;
  // White noise convolved with exp(-r^2/(2 s^2)), s = corr_len/sqrt(2), has
  // covariance proportional to exp(-r^2/(2 corr_len^2)).
  const double s = corr_len / std::sqrt(2.0);
  const int rad = (int)std::ceil(3.0 * s);
  const int kw = 2 * rad + 1, pw = nf + 2 * rad;
  std::vector<double> w1(kw);
  double norm1 = 0;
  for (int i = 0; i < kw; ++i)
    {
      const double d = i - rad;
      w1[i] = std::exp(-d * d / (2 * s * s));
      norm1 += w1[i] * w1[i];
    }
  const double norm = 1.0 / norm1; // 2D separable kernel: sum w^2 = norm1^2
  std::vector<double> noise((size_t)pw * pw), tmp((size_t)pw * nf);
  for (int c = 0; c < n_cells; ++c)
    {
      for (int j = 0; j < pw * pw; ++j)
        noise[j] = draw(seed, (uint32_t)j, (uint32_t)c, 2u).z0;
      for (int y = 0; y < pw; ++y) // convolve along x
        for (int x = 0; x < nf; ++x)
          {
            double a = 0;
            for (int k = 0; k < kw; ++k)
              a += w1[k] * noise[(size_t)y * pw + x + k];
            tmp[(size_t)y * nf + x] = a;
          }
      for (int y = 0; y < nf; ++y) // along y
        for (int x = 0; x < nf; ++x)
          {
            double a = 0;
            for (int k = 0; k < kw; ++k)
              a += w1[k] * tmp[(size_t)(y + k) * nf + x];
            a *= norm;
            g[(size_t)c * nf * nf + (size_t)y * nf + x] = (float)std::min(clip, std::max(-clip, a));
          }
    }
  return MSFEM_OK;
}

} // extern "C"
