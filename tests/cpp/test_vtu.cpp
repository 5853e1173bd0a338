// Writes small VTU/PVTU files with msfem_vtu.hpp for tests/test_vtu.py to
// check (no GPU).  Usage: test_vtu <out_dir>
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "msfem_vtu.hpp"

int main(int argc, char **argv)
{
  if (argc != 2)
    return 2;
  const std::string dir = argv[1];
  // 2D: box [0.5, 1] x [0, 0.25], 3 fine cells per side, field x + 10 y
  {
    const double box[4] = {0.5, 1.0, 0.0, 0.25};
    const int nf = 3, np = nf + 1;
    std::vector<double> v(np * np);
    for (int k = 0; k < np * np; ++k)
      v[k] = (box[0] + (k % np) * (box[1] - box[0]) / nf) + 10 * (box[2] + (k / np) * (box[3] - box[2]) / nf);
    msfem::vtu::write_cell(dir + "/cell2d.vtu", 2, box, nf, {{"solution", v.data()}, {"twice", v.data()}});
  }
  // 3D: unit cube, 2 fine cells per side, field z
  {
    const double box[6] = {0, 1, 0, 1, 0, 1};
    const int nf = 2, np = nf + 1;
    std::vector<double> v(np * np * np);
    for (int k = 0; k < np * np * np; ++k)
      v[k] = (k / (np * np)) * 0.5;
    msfem::vtu::write_cell(dir + "/cell3d.vtu", 3, box, nf, {{"solution", v.data()}});
  }
  // 2D general quadrilateral (deal.II vertex order), 4 fine cells per side, field x
  {
    const double quad[8] = {0.1, 0.2, 0.9, 0.1, 0.25, 0.8, 0.7, 1.05};
    const int nf = 4, np = nf + 1;
    std::vector<double> v(np * np);
    for (int k = 0; k < np * np; ++k)
      {
        const double xi = (double)(k % np) / nf, eta = (double)(k / np) / nf;
        v[k] = (1 - xi) * (1 - eta) * quad[0] + xi * (1 - eta) * quad[2] + (1 - xi) * eta * quad[4] + xi * eta * quad[6];
      }
    msfem::vtu::write_cell(dir + "/quad2d.vtu", 2, nullptr, nf, {{"solution", v.data()}}, quad);
  }
  msfem::vtu::write_pvtu(dir + "/master.pvtu", {"solution"}, {"cell2d.vtu", "cell3d.vtu"});
  std::printf("vtu ok\n");
  return 0;
}
