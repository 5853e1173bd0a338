// msfem_vtu.hpp -- VTK XML output of fine-scale fields on a coarse cell.
//
// Host code only (no GPU).  Replaces the reference's deal.II DataOut output of
// the per-cell fine solution and bases and of the parallel master record:
//   DiffusionProblemBasis::output_basis / output_global_solution_in_cell
//     (include/base/diffusion_problem_basis.tpp:393-435)
//   DiffusionProblemMultiscale::output_global_fine
//     (include/base/diffusion_problem_ms.tpp:384-420, write_pvtu_record)
//   DiffusionProblemMultiscale::output_global_coarse
//     (include/base/diffusion_problem_ms.tpp:330-380)
// A cell with nf fine cells per side is written as an UnstructuredGrid of
// nf^dim VTK_QUAD (2D) or VTK_HEXAHEDRON (3D) elements on its (nf+1)^dim
// lexicographic fine nodes (x fastest), with the given nodal fields as
// Float64 PointData, ASCII.  The cell is a box or (2D) a general quadrilateral
// whose fine nodes are the bilinear image of the uniform grid.
#ifndef MSFEM_VTU_HPP_
#define MSFEM_VTU_HPP_

#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace msfem {
namespace vtu {

struct Field {
  std::string name;
  const double *values; // (nf+1)^dim nodal values, lexicographic
};

// box: {lo_x, hi_x, lo_y, hi_y[, lo_z, hi_z]}; verts (or nullptr): a general
// cell's 2^dim vertices {x0, y0[, z0], ...} in deal.II order, used instead of
// the box (the fine nodes are the cell's bi/trilinear map on the parametric grid).
inline void write_cell(const std::string &path, int dim, const double *box, int nf, const std::vector<Field> &fields,
                       const double *verts = nullptr)
{
  if ((dim != 2 && dim != 3) || nf <= 0)
    throw std::invalid_argument("vtu::write_cell: dim must be 2 or 3 and nf > 0");
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f)
    throw std::runtime_error("vtu::write_cell: cannot open " + path);
  const int np = nf + 1;
  const long long npts = dim == 2 ? (long long)np * np : (long long)np * np * np;
  const long long ncells = dim == 2 ? (long long)nf * nf : (long long)nf * nf * nf;
  const int nv = dim == 2 ? 4 : 8;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n");
  std::fprintf(f, "<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", npts, ncells);
  std::fprintf(f, "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (long long k = 0; k < npts; ++k)
    {
      const int X = (int)(k % np), Y = (int)((k / np) % np), Z = dim == 3 ? (int)(k / ((long long)np * np)) : 0;
      double x, y, z = 0.0;
      if (verts) // the cell's multilinear map at the parametric node
        {
          const double xi[3] = {(double)X / nf, (double)Y / nf, (double)Z / nf};
          double p[3] = {0, 0, 0};
          for (int v = 0; v < nv; ++v)
            {
              double w = 1;
              for (int m = 0; m < dim; ++m)
                w *= ((v >> m) & 1) ? xi[m] : 1 - xi[m];
              for (int k = 0; k < dim; ++k)
                p[k] += w * verts[v * dim + k];
            }
          x = p[0];
          y = p[1];
          z = p[2];
        }
      else
        {
          x = box[0] + X * ((box[1] - box[0]) / nf);
          y = box[2] + Y * ((box[3] - box[2]) / nf);
          if (dim == 3)
            z = box[4] + Z * ((box[5] - box[4]) / nf);
        }
      std::fprintf(f, "%.17g %.17g %.17g\n", x, y, z);
    }
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" "
                  "format=\"ascii\">\n");
  for (long long c = 0; c < ncells; ++c)
    {
      const long long cx = c % nf, cy = (c / nf) % nf, cz = dim == 3 ? c / ((long long)nf * nf) : 0;
      auto node = [&](int dx, int dy, int dz) { return (cx + dx) + np * ((cy + dy) + (long long)np * (cz + dz)); };
      // VTK vertex order: counter-clockwise in each x-y face, bottom face first
      if (dim == 2)
        std::fprintf(f, "%lld %lld %lld %lld\n", node(0, 0, 0), node(1, 0, 0), node(1, 1, 0), node(0, 1, 0));
      else
        std::fprintf(f, "%lld %lld %lld %lld %lld %lld %lld %lld\n", node(0, 0, 0), node(1, 0, 0), node(1, 1, 0),
                     node(0, 1, 0), node(0, 0, 1), node(1, 0, 1), node(1, 1, 1), node(0, 1, 1));
    }
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (long long c = 1; c <= ncells; ++c)
    std::fprintf(f, "%lld\n", c * nv);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (long long c = 0; c < ncells; ++c)
    std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);
  std::fprintf(f, "</DataArray>\n</Cells>\n<PointData>\n");
  for (const Field &fd : fields)
    {
      std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", fd.name.c_str());
      for (long long k = 0; k < npts; ++k)
        std::fprintf(f, "%.17g\n", fd.values[k]);
      std::fprintf(f, "</DataArray>\n");
    }
  std::fprintf(f, "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  if (std::fclose(f) != 0)
    throw std::runtime_error("vtu::write_cell: write failed for " + path);
}

// A structured grid of n^dim cells (VTK_QUAD / VTK_HEXAHEDRON) on (n+1)^dim
// nodes given by their coordinates points[(n+1)^dim][3] (x fastest), with
// nodal fields: the coarse mesh of DiffusionProblemMultiscale::output_global_coarse.
inline void write_grid(const std::string &path, int dim, int n, const double *points, const std::vector<Field> &fields)
{
  if ((dim != 2 && dim != 3) || n <= 0)
    throw std::invalid_argument("vtu::write_grid: dim must be 2 or 3 and n > 0");
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f)
    throw std::runtime_error("vtu::write_grid: cannot open " + path);
  const int np = n + 1;
  const long long npts = dim == 2 ? (long long)np * np : (long long)np * np * np;
  const long long ncells = dim == 2 ? (long long)n * n : (long long)n * n * n;
  const int nv = dim == 2 ? 4 : 8;
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<UnstructuredGrid>\n");
  std::fprintf(f, "<Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n", npts, ncells);
  std::fprintf(f, "<Points>\n<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (long long k = 0; k < npts; ++k)
    std::fprintf(f, "%.17g %.17g %.17g\n", points[3 * k], points[3 * k + 1], points[3 * k + 2]);
  std::fprintf(f, "</DataArray>\n</Points>\n<Cells>\n<DataArray type=\"Int64\" Name=\"connectivity\" "
                  "format=\"ascii\">\n");
  for (long long c = 0; c < ncells; ++c)
    {
      const long long cx = c % n, cy = (c / n) % n, cz = dim == 3 ? c / ((long long)n * n) : 0;
      auto node = [&](int dx, int dy, int dz) { return (cx + dx) + np * ((cy + dy) + (long long)np * (cz + dz)); };
      if (dim == 2)
        std::fprintf(f, "%lld %lld %lld %lld\n", node(0, 0, 0), node(1, 0, 0), node(1, 1, 0), node(0, 1, 0));
      else
        std::fprintf(f, "%lld %lld %lld %lld %lld %lld %lld %lld\n", node(0, 0, 0), node(1, 0, 0), node(1, 1, 0),
                     node(0, 1, 0), node(0, 0, 1), node(1, 0, 1), node(1, 1, 1), node(0, 1, 1));
    }
  std::fprintf(f, "</DataArray>\n<DataArray type=\"Int64\" Name=\"offsets\" format=\"ascii\">\n");
  for (long long c = 1; c <= ncells; ++c)
    std::fprintf(f, "%lld\n", c * nv);
  std::fprintf(f, "</DataArray>\n<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n");
  for (long long c = 0; c < ncells; ++c)
    std::fprintf(f, "%d\n", dim == 2 ? 9 : 12);
  std::fprintf(f, "</DataArray>\n</Cells>\n<PointData>\n");
  for (const Field &fd : fields)
    {
      std::fprintf(f, "<DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", fd.name.c_str());
      for (long long k = 0; k < npts; ++k)
        std::fprintf(f, "%.17g\n", fd.values[k]);
      std::fprintf(f, "</DataArray>\n");
    }
  std::fprintf(f, "</PointData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  if (std::fclose(f) != 0)
    throw std::runtime_error("vtu::write_grid: write failed for " + path);
}

// Parallel master record naming the per-cell pieces (paths relative to it).
inline void write_pvtu(const std::string &path, const std::vector<std::string> &field_names,
                       const std::vector<std::string> &pieces)
{
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f)
    throw std::runtime_error("vtu::write_pvtu: cannot open " + path);
  std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<PUnstructuredGrid GhostLevel=\"0\">\n<PPointData>\n");
  for (const std::string &n : field_names)
    std::fprintf(f, "<PDataArray type=\"Float64\" Name=\"%s\"/>\n", n.c_str());
  std::fprintf(f, "</PPointData>\n<PPoints>\n<PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n</PPoints>\n");
  for (const std::string &p : pieces)
    std::fprintf(f, "<Piece Source=\"%s\"/>\n", p.c_str());
  std::fprintf(f, "</PUnstructuredGrid>\n</VTKFile>\n");
  if (std::fclose(f) != 0)
    throw std::runtime_error("vtu::write_pvtu: write failed for " + path);
}

} // namespace vtu
} // namespace msfem

#endif // MSFEM_VTU_HPP_
