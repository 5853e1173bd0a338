// Drives the C++ facade the way the reference's main drives its classes
// (source/main.cxx:37-39): DiffusionProblemMultiscale<2>(n_refine, n_refine_local).run()
// on the reference default problem, then writes the coarse solution, the
// element matrices and CG iteration counts for the Python side to check.
// Usage: test_facade <n_refine> <n_refine_local> <out.bin> [<vtu_dir> [<distortion>]]
// (with vtu_dir other than "-": also the fine output,
// DiffusionProblemMultiscale::output_global_fine; with distortion: the coarse
// mesh's interior vertices moved by up to distortion/nc, seed 1, so every cell
// is a general quadrilateral)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "msfem_basis.hpp"

int main(int argc, char **argv)
{
  if (argc < 4 || argc > 6)
    {
      std::fprintf(stderr, "usage: %s n_refine n_refine_local out.bin [vtu_dir [distortion]]\n", argv[0]);
      return 2;
    }
  const unsigned nr = std::atoi(argv[1]), nrl = std::atoi(argv[2]);
  try
    {
      msfem::Engine engine(0);
      msfem::DiffusionProblemMultiscale<2> ms(engine, nr, nrl);
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
        ms.output_global_fine(argv[4]);
      std::printf("facade ok: %lld cells, coarse CG %d iterations\n", nc, ms.coarse_iterations());
      // a getter before run() must throw, as the reference asserts on misuse
      msfem::BasisBatch<2> batch(engine);
      msfem::DiffusionProblemBasis<2> cell(3, {0.0, 0.5, 0.0, 0.5}, 0, batch);
      try
        {
          cell.get_global_element_matrix();
          std::printf("missing run() not detected\n");
          return 4;
        }
      catch (const msfem::Error &)
        {
        }
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "exception: %s\n", e.what());
      return 1;
    }
  return 0;
}
