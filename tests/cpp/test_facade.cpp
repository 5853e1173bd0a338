// Drives the C++ facade the way the reference's main drives its classes
// (source/main.cxx:37-39): DiffusionProblemMultiscale<dim>(n_refine, n_refine_local).run()
// on the reference default problem, then writes the coarse solution, the
// element matrices and CG iteration counts for the Python side to check.
// Usage: test_facade [--dim=3] <n_refine> <n_refine_local> <out.bin> [<vtu_dir> [<distortion>]]
// (with vtu_dir other than "-": also the fine and coarse output,
// DiffusionProblemMultiscale::output_global_fine / output_global_coarse; with
// distortion: the coarse mesh's interior vertices moved by up to
// distortion/nc, seed 1, so every cell is a general quadrilateral / hexahedron)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "msfem_basis.hpp"

template <int dim>
int drive(int argc, char **argv)
{
  const unsigned nr = std::atoi(argv[1]), nrl = std::atoi(argv[2]);
  msfem::Engine engine(0);
  msfem::DiffusionProblemMultiscale<dim> ms(engine, nr, nrl);
  if (argc == 6)
    ms.set_mesh_distortion(std::atof(argv[5]), 1);
  ms.run();
  const auto &u = ms.coarse_solution();
  std::FILE *f = std::fopen(argv[3], "wb");
  if (!f)
    return 3;
  const long long nu = (long long)u.size(), nc = (long long)ms.cells().size();
  std::fwrite(&nu, sizeof nu, 1, f);
  std::fwrite(u.data(), sizeof(double), u.size(), f);
  std::fwrite(&nc, sizeof nc, 1, f);
  for (const auto &kv : ms.cells())
    {
      const auto &A = kv.second.get_global_element_matrix();
      const auto &F = kv.second.get_global_element_rhs();
      std::fwrite(A.data(), sizeof(double), A.size(), f);
      std::fwrite(F.data(), sizeof(double), F.size(), f);
    }
  std::fclose(f);
  if (argc >= 5 && std::string(argv[4]) != "-")
    {
      ms.output_global_fine(argv[4]);
      ms.output_global_coarse(argv[4]);
    }
  std::printf("facade ok: dim %d, %lld cells, coarse CG %d iterations\n", dim, nc, ms.coarse_iterations());
  // a getter before run() must throw, as the reference asserts on misuse
  msfem::BasisBatch<dim> batch(engine);
  std::array<double, 2 * dim> box;
  for (int k = 0; k < dim; ++k)
    {
      box[2 * k] = 0.0;
      box[2 * k + 1] = 0.5;
    }
  msfem::DiffusionProblemBasis<dim> cell(3, box, 0, batch);
  try
    {
      cell.get_global_element_matrix();
      std::printf("missing run() not detected\n");
      return 4;
    }
  catch (const msfem::Error &)
    {
    }
  return 0;
}

int main(int argc, char **argv)
{
  int dim = 2;
  if (argc > 1 && std::string(argv[1]) == "--dim=3")
    {
      dim = 3;
      --argc;
      ++argv;
    }
  if (argc < 4 || argc > 6)
    {
      std::fprintf(stderr, "usage: %s [--dim=3] n_refine n_refine_local out.bin [vtu_dir [distortion]]\n", argv[0]);
      return 2;
    }
  try
    {
      return dim == 2 ? drive<2>(argc, argv) : drive<3>(argc, argv);
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "exception: %s\n", e.what());
      return 1;
    }
}
