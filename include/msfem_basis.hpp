// msfem_basis.hpp -- C++ facade with the shape of the reference's per-cell
// interface, over the C ABI of msfem.h (header-only; link libmsfem_hip.so).
//
//   DiffusionProblemBasis<dim>       include/base/diffusion_problem_basis.hpp:68-285
//   DiffusionProblemMultiscale<dim>  include/base/diffusion_problem_ms.hpp:91-231
//
// The reference constructs one DiffusionProblemBasis per locally owned coarse
// cell, calls run() on each in a serial loop and then reads the element matrix
// and load through getters (include/base/diffusion_problem_ms.tpp:50-87,
// :200-201).  Here the constructor registers the cell with a BasisBatch, run()
// only marks it, and the first getter (or BasisBatch::flush()) builds every
// registered cell in one msfem_batch_build call on the GPU.  The call sequence
// of the reference driver is unchanged.
//
// Differences forced by the data path: a cell is described by its box, or by
// its 2^dim vertices for a general quadrilateral / hexahedron, instead of a deal.II cell
// iterator (the reference reads the vertices from the iterator,
// diffusion_problem_basis.tpp:43-49), the communicator argument is gone (the local
// problem never used it, diffusion_problem_basis.hpp:198), FullMatrix<double>
// and Vector<double> become row-major std::vector<double>, and a failed solve
// throws msfem::NoConvergence (the reference throws deal.II's
// SolverControl::NoConvergence out of run()).
#ifndef MSFEM_BASIS_HPP_
#define MSFEM_BASIS_HPP_

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "msfem.h"
#include "msfem_vtu.hpp"

namespace msfem {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
struct NoConvergence : Error {
  explicit NoConvergence(const std::string &m) : Error(MSFEM_ERR_NOCONV, m) {}
};

inline void check(int rc, const msfem_ctx *ctx)
{
  if (rc == MSFEM_OK)
    return;
  const std::string msg = msfem_last_error(ctx);
  if (rc == MSFEM_ERR_NOCONV)
    throw NoConvergence(msg);
  throw Error(rc, msg);
}

// Owns the GPU context (one per process and device, as one MPI rank owns its cells).
class Engine {
public:
  explicit Engine(int device = 0) { check(msfem_ctx_create(device, &ctx_), nullptr); }
  ~Engine() { msfem_ctx_destroy(ctx_); }
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;
  msfem_ctx *get() const { return ctx_; }

private:
  msfem_ctx *ctx_ = nullptr;
};

// Coefficient and load of the cells of a batch.  The default is the
// reference's own problem: MatrixCoeff with PI_D at the quadrature points
// (include/coefficients/matrix_coeff.tpp:66-91) and f = 2
// (include/coefficients/right_hand_side.tpp:27-40).
struct ProblemData {
  int coef_kind = MSFEM_COEF_TENSOR_QP;
  int rhs_kind = MSFEM_RHS_CONST;
  double rhs_value = 2.0;
  // Fills the concatenated coefficient tiles of the given cells; nullptr =
  // the reference coefficient.
  void (*fill)(int dim, int n_cells, const int32_t *n_fine, const double *box, double *coef) = nullptr;
  // The same for a batch of general cells (vertices [n][2^dim][dim]); a batch
  // with such cells and a custom coefficient needs it.
  void (*fill_general)(int dim, int n_cells, const int32_t *n_fine, const double *vertices, double *coef) = nullptr;
};

template <int dim>
class BasisBatch {
public:
  static constexpr int nb = 1 << dim;

  BasisBatch(Engine &engine, ProblemData data = ProblemData(), msfem_solver_opts opts = default_opts())
    : engine_(engine), data_(data), opts_(opts)
  {
  }

  static msfem_solver_opts default_opts()
  {
    msfem_solver_opts o;
    msfem_default_opts(&o);
    return o;
  }

  int add(unsigned n_refine_local, const std::array<double, 2 * dim> &box)
  {
    boxes_.push_back(box);
    std::array<double, nb * dim> v;
    for (int i = 0; i < nb; ++i)
      for (int k = 0; k < dim; ++k)
        v[i * dim + k] = box[2 * k + ((i >> k) & 1)];
    verts_.push_back(v);
    n_fine_.push_back(1 << n_refine_local);
    built_ = false;
    return (int)n_fine_.size() - 1;
  }
  // A general cell -- quadrilateral (2D) or hexahedron (3D) -- by its 2^dim
  // vertices in deal.II order.  The whole batch is then built as general cells
  // (boxes as their 2^dim vertices).
  int add_general(unsigned n_refine_local, const std::vector<std::array<double, dim>> &corner_points)
  {
    if ((int)corner_points.size() != nb)
      throw Error(MSFEM_ERR_ARG, "general cells: 2^dim corner points");
    std::array<double, 2 * dim> bbox;
    std::array<double, nb * dim> v;
    for (int k = 0; k < dim; ++k)
      {
        bbox[2 * k] = bbox[2 * k + 1] = corner_points[0][k];
        for (int i = 0; i < nb; ++i)
          {
            v[i * dim + k] = corner_points[i][k];
            bbox[2 * k] = std::min(bbox[2 * k], corner_points[i][k]);
            bbox[2 * k + 1] = std::max(bbox[2 * k + 1], corner_points[i][k]);
          }
      }
    boxes_.push_back(bbox); // bounding box: unused by a general batch
    verts_.push_back(v);
    n_fine_.push_back(1 << n_refine_local);
    general_ = true;
    built_ = false;
    return (int)n_fine_.size() - 1;
  }

  // One GPU build for every registered cell.
  void flush()
  {
    if (built_)
      return;
    const int n = (int)n_fine_.size();
    std::vector<double> box(boxes_.size() * 2 * dim);
    for (size_t c = 0; c < boxes_.size(); ++c)
      for (int k = 0; k < 2 * dim; ++k)
        box[c * 2 * dim + k] = boxes_[c][k];
    const int ns = dim == 2 ? 3 : 6, nq = 1 << dim;
    const int stride = data_.coef_kind == MSFEM_COEF_SCALAR_CELL   ? 1
                       : data_.coef_kind == MSFEM_COEF_TENSOR_CELL ? ns
                       : data_.coef_kind == MSFEM_COEF_SCALAR_QP   ? nq
                                                                   : nq * ns;
    size_t ncoef = 0, nbases = 0;
    offsets_.assign(n + 1, 0);
    for (int c = 0; c < n; ++c)
      {
        ncoef += (size_t)std::pow(n_fine_[c], dim) * stride;
        offsets_[c + 1] = offsets_[c] + (size_t)nb * (size_t)std::pow(n_fine_[c] + 1, dim);
      }
    nbases = offsets_[n];
    std::vector<double> coef(ncoef), vert;
    if (general_)
      {
        vert.resize((size_t)n * nb * dim);
        for (int c = 0; c < n; ++c)
          std::copy(verts_[c].begin(), verts_[c].end(), vert.begin() + (size_t)c * nb * dim);
        if (data_.fill_general)
          data_.fill_general(dim, n, n_fine_.data(), vert.data(), coef.data());
        else if (data_.fill)
          throw Error(MSFEM_ERR_ARG, "a batch of general cells needs ProblemData::fill_general");
        else if (dim == 2)
          check(msfem_gen_coef_reference_quad(n, n_fine_.data(), vert.data(), coef.data()), nullptr);
        else
          check(msfem_gen_coef_reference_hex(n, n_fine_.data(), vert.data(), coef.data()), nullptr);
      }
    else if (data_.fill)
      data_.fill(dim, n, n_fine_.data(), box.data(), coef.data());
    else
      check(msfem_gen_coef_reference(dim, n, n_fine_.data(), box.data(), coef.data()), nullptr);
    Ae_.assign((size_t)n * nb * nb, 0.0);
    Fe_.assign((size_t)n * nb, 0.0);
    bases_.assign(nbases, 0.0);
    iters_.assign((size_t)n * nb, 0);
    msfem_batch in{dim, n, n_fine_.data(), box.data(), data_.coef_kind, coef.data(), data_.rhs_kind, data_.rhs_value,
                   nullptr, general_ ? vert.data() : nullptr};
    msfem_batch_out out{Ae_.data(), Fe_.data(), bases_.data(), iters_.data(), nullptr, nullptr, nullptr};
    check(msfem_batch_build(engine_.get(), &in, &opts_, &out), engine_.get());
    built_ = true;
  }

  bool built() const { return built_; }
  bool general() const { return general_; }
  int n_fine(int id) const { return n_fine_[id]; }
  const double *vertices(int id) const { return verts_[id].data(); }
  const double *element_matrix(int id) { return Ae_.data() + (size_t)id * nb * nb; }
  const double *element_rhs(int id) { return Fe_.data() + (size_t)id * nb; }
  const double *bases(int id) { return bases_.data() + offsets_[id]; }
  const int32_t *iterations(int id) { return iters_.data() + (size_t)id * nb; }

private:
  Engine &engine_;
  ProblemData data_;
  msfem_solver_opts opts_;
  std::vector<std::array<double, 2 * dim>> boxes_;
  std::vector<std::array<double, nb * dim>> verts_;
  bool general_ = false;
  std::vector<int32_t> n_fine_;
  std::vector<size_t> offsets_;
  std::vector<double> Ae_, Fe_, bases_;
  std::vector<int32_t> iters_;
  bool built_ = false;
};

// Same public surface as the reference class (diffusion_problem_basis.hpp:77-138).
template <int dim>
class DiffusionProblemBasis {
public:
  static constexpr int nb = 1 << dim;

  DiffusionProblemBasis() = delete;
  DiffusionProblemBasis(unsigned n_refine_local, const std::array<double, 2 * dim> &box, unsigned local_subdomain,
                        BasisBatch<dim> &batch)
    : batch_(&batch), id_(batch.add(n_refine_local, box)), local_subdomain_(local_subdomain), box_(box)
  {
  }
  // A general cell (quadrilateral or hexahedron): its vertices in deal.II
  // order, as corner_points in the reference (diffusion_problem_basis.tpp:43-49).
  DiffusionProblemBasis(unsigned n_refine_local, const std::vector<std::array<double, dim>> &corner_points,
                        unsigned local_subdomain, BasisBatch<dim> &batch)
    : batch_(&batch), id_(batch.add_general(n_refine_local, corner_points)), local_subdomain_(local_subdomain),
      box_()
  {
  }
  DiffusionProblemBasis(const DiffusionProblemBasis &) = default;

  // diffusion_problem_basis.tpp:440-474.  The work runs batched on the first getter.
  void run() { ran_ = true; }

  // diffusion_problem_basis.tpp:320-334 (row-major nb x nb and nb)
  const std::vector<double> &get_global_element_matrix() const
  {
    ensure();
    if (Ae_.empty())
      Ae_.assign(batch_->element_matrix(id_), batch_->element_matrix(id_) + nb * nb);
    return Ae_;
  }
  const std::vector<double> &get_global_element_rhs() const
  {
    ensure();
    if (Fe_.empty())
      Fe_.assign(batch_->element_rhs(id_), batch_->element_rhs(id_) + nb);
    return Fe_;
  }

  // diffusion_problem_basis.tpp:354-377: u|_K = sum_i w_i phi_i on the fine nodes.
  void set_global_weights(const std::vector<double> &weights)
  {
    if ((int)weights.size() != nb)
      throw Error(MSFEM_ERR_ARG, "set_global_weights: expected one weight per basis");
    ensure();
    weights_ = weights;
    const int nn = (int)std::pow(batch_->n_fine(id_) + 1, dim);
    const double *phi = batch_->bases(id_);
    solution_.assign(nn, 0.0);
    for (int i = 0; i < nb; ++i)
      for (int k = 0; k < nn; ++k)
        solution_[k] += weights_[i] * phi[(size_t)i * nn + k];
    weights_set_ = true;
  }
  const std::vector<double> &get_global_solution() const
  {
    if (!weights_set_)
      throw Error(MSFEM_ERR_ARG, "global weights must be set first");
    return solution_;
  }
  // The basis nodal values [nb][(nf+1)^dim], lexicographic (x fastest).
  const double *get_basis_values() const
  {
    ensure();
    return batch_->bases(id_);
  }
  const int32_t *get_iterations() const
  {
    ensure();
    return batch_->iterations(id_);
  }

  // diffusion_problem_basis.tpp:380-389 (file name of the cell's fine output)
  const std::string &get_filename_global()
  {
    if (filename_.empty())
      {
        char buf[64];
        snprintf(buf, sizeof buf, "%05u", local_subdomain_);
        filename_ = std::string(dim == 2 ? "solution-ms_fine-2d" : "solution-ms_fine-3d") + "." + buf + ".cell-" +
                    std::to_string(id_) + ".vtu";
      }
    return filename_;
  }
  void set_output_flag(bool flag) { output_flag_ = flag; }
  bool output_flag() const { return output_flag_; }

  // diffusion_problem_basis.tpp:393-414: the cell's bases as one VTU file
  // "2d-basis.<subdomain>.cell-<id>.vtu" (3d-...) in directory `dir`.
  void output_basis(const std::string &dir = "") const
  {
    ensure();
    char buf[64];
    snprintf(buf, sizeof buf, "%05u", local_subdomain_);
    const std::string name = std::string(dim == 2 ? "2d-" : "3d-") + "basis." + buf + ".cell-" + std::to_string(id_) +
                             ".vtu";
    const int nf = batch_->n_fine(id_);
    const size_t nn = (size_t)std::pow(nf + 1, dim);
    std::vector<vtu::Field> fields;
    for (int i = 0; i < nb; ++i)
      fields.push_back({"basis_" + std::to_string(i), batch_->bases(id_) + (size_t)i * nn});
    vtu::write_cell(join(dir, name), dim, box_.data(), nf, fields, geometry());
  }

  // diffusion_problem_basis.tpp:418-431: the fine solution of the cell
  // (after set_global_weights) as get_filename_global() in directory `dir`.
  void output_global_solution_in_cell(const std::string &dir = "")
  {
    if (!weights_set_)
      throw Error(MSFEM_ERR_ARG, "global weights must be set first");
    vtu::write_cell(join(dir, get_filename_global()), dim, box_.data(), batch_->n_fine(id_),
                    {{"solution", solution_.data()}}, geometry());
  }

  static std::string join(const std::string &dir, const std::string &name)
  {
    return dir.empty() ? name : (dir.back() == '/' ? dir + name : dir + "/" + name);
  }

private:
  // vertices of a general cell for the VTU writer (nullptr: the box)
  const double *geometry() const { return batch_->general() ? batch_->vertices(id_) : nullptr; }

  void ensure() const
  {
    if (!ran_)
      throw Error(MSFEM_ERR_ARG, "run() was not called");
    batch_->flush();
  }

  BasisBatch<dim> *batch_;
  int id_;
  unsigned local_subdomain_;
  std::array<double, 2 * dim> box_;
  bool ran_ = false, output_flag_ = false, weights_set_ = false;
  mutable std::vector<double> Ae_, Fe_;
  std::vector<double> weights_, solution_;
  std::string filename_;
};

// The multiscale problem around the hot path (include/base/diffusion_problem_ms.tpp:447-491):
// coarse mesh of the unit square/cube, one basis object
// per cell, coarse assembly with the reference boundary data, FP64 coarse
// solve, weights back to the cells.
template <int dim>
class DiffusionProblemMultiscale {
public:
  DiffusionProblemMultiscale(Engine &engine, unsigned n_refine, unsigned n_refine_local,
                             ProblemData data = ProblemData(), int bc_kind = MSFEM_BC_REFERENCE)
    : engine_(engine), n_refine_(n_refine), n_refine_local_(n_refine_local), batch_(engine, data), bc_kind_(bc_kind)
  {
  }

  void run()
  {
    const int nc = 1 << n_refine_, n = (int)std::pow(nc, dim), nb = 1 << dim;
    std::vector<double> box((size_t)n * 2 * dim);
    std::vector<double> &verts = verts_;
    verts.clear();
    check(msfem_gen_unit_boxes(dim, nc, box.data()), nullptr);
    if (distortion_ > 0)
      {
        verts.resize((size_t)n * nb * dim);
        if (dim == 2)
          check(msfem_gen_perturbed_quads(nc, distortion_, distortion_seed_, verts.data()), nullptr);
        else
          check(msfem_gen_perturbed_hexes(nc, distortion_, distortion_seed_, verts.data()), nullptr);
      }
    cells_.clear();
    for (int c = 0; c < n; ++c) // initialize_and_compute_basis (diffusion_problem_ms.tpp:42-88)
      {
        if (distortion_ > 0)
          {
            std::vector<std::array<double, dim>> corners(nb);
            for (int i = 0; i < nb; ++i)
              for (int k = 0; k < dim; ++k)
                corners[i][k] = verts[((size_t)c * nb + i) * dim + k];
            cells_.emplace(c, DiffusionProblemBasis<dim>(n_refine_local_, corners, 0, batch_));
          }
        else
          {
            std::array<double, 2 * dim> b;
            for (int k = 0; k < 2 * dim; ++k)
              b[k] = box[(size_t)c * 2 * dim + k];
            cells_.emplace(c, DiffusionProblemBasis<dim>(n_refine_local_, b, 0, batch_));
          }
        cells_.at(c).set_output_flag(output_all_);
      }
    for (auto &kv : cells_)
      kv.second.run();
    std::vector<double> Ae((size_t)n * nb * nb), Fe((size_t)n * nb); // assemble_system (:161-255)
    for (auto &kv : cells_)
      {
        const auto &A = kv.second.get_global_element_matrix();
        const auto &F = kv.second.get_global_element_rhs();
        std::copy(A.begin(), A.end(), Ae.begin() + (size_t)kv.first * nb * nb);
        std::copy(F.begin(), F.end(), Fe.begin() + (size_t)kv.first * nb);
      }
    const size_t nn = (size_t)std::pow(nc + 1, dim), ns3 = (size_t)std::pow(3, dim);
    std::vector<double> Kc(nn * ns3), Fc(nn), cval(nn);
    std::vector<uint8_t> cmask(nn);
    check(msfem_coarse_assemble(engine_.get(), dim, nc, Ae.data(), Fe.data(), bc_kind_, Kc.data(), Fc.data(),
                                cmask.data(), cval.data()),
          engine_.get());
    solution_.assign(nn, 0.0); // solve_iterative (:260-296)
    check(msfem_coarse_solve(engine_.get(), dim, nc, Kc.data(), Fc.data(), cmask.data(), cval.data(), 1e-14, 100000,
                             solution_.data(), &coarse_iterations_),
          engine_.get());
    std::vector<double> w((size_t)n * nb); // send_global_weights_to_cell (:301-325)
    check(msfem_gather_weights(dim, nc, solution_.data(), w.data()), nullptr);
    for (auto &kv : cells_)
      kv.second.set_global_weights(std::vector<double>(w.begin() + (size_t)kv.first * nb,
                                                       w.begin() + (size_t)(kv.first + 1) * nb));
    for (auto &kv : cells_) // run() of a cell with the output flag writes its bases
      if (kv.second.output_flag())
        kv.second.output_basis(output_dir_);
  }

  // diffusion_problem_ms.tpp:384-420: every cell's fine solution plus the
  // master record "solution-ms_fine-2d.pvtu" (3d) naming them, in `dir`.
  void output_global_fine(const std::string &dir = "")
  {
    std::vector<std::string> pieces;
    for (auto &kv : cells_)
      {
        kv.second.output_global_solution_in_cell(dir);
        pieces.push_back(kv.second.get_filename_global());
      }
    vtu::write_pvtu(DiffusionProblemBasis<dim>::join(dir, dim == 2 ? "solution-ms_fine-2d.pvtu"
                                                                   : "solution-ms_fine-3d.pvtu"),
                    {"solution"}, pieces);
  }

  // diffusion_problem_ms.tpp:330-380: the coarse solution "u" on the coarse mesh,
  // "solution-ms_coarse-<d>d_refinements-<n_refine>.0000.vtu" (one process:
  // subdomain 0) and its master record "..._refinements-<n_refine>.pvtu", in
  // `dir`.  After run().
  void output_global_coarse(const std::string &dir = "") const
  {
    if (solution_.empty())
      throw Error(MSFEM_ERR_ARG, "output_global_coarse before run()");
    const int nc = 1 << n_refine_, np = nc + 1;
    const size_t nn = (size_t)std::pow(np, dim);
    std::vector<double> pts(3 * nn, 0.0);
    for (size_t k = 0; k < nn; ++k)
      {
        const int I[3] = {(int)(k % np), (int)((k / np) % np), dim == 3 ? (int)(k / ((size_t)np * np)) : 0};
        for (int m = 0; m < dim; ++m)
          pts[3 * k + m] = (double)I[m] / nc;
        if (!verts_.empty()) // moved vertices: vertex of a cell that has this node
          {
            int cc[3] = {0, 0, 0}, v = 0;
            for (int m = 0; m < dim; ++m)
              {
                cc[m] = I[m] < nc ? I[m] : nc - 1;
                v += (I[m] - cc[m]) << m;
              }
            const size_t cell = ((size_t)cc[2] * nc + cc[1]) * nc + cc[0];
            for (int m = 0; m < dim; ++m)
              pts[3 * k + m] = verts_[(cell * (1 << dim) + v) * dim + m];
          }
      }
    const std::string base = std::string(dim == 2 ? "solution-ms_coarse-2d" : "solution-ms_coarse-3d") +
                             "_refinements-" + std::to_string(n_refine_);
    vtu::write_grid(DiffusionProblemBasis<dim>::join(dir, base + ".0000.vtu"), dim, nc, pts.data(),
                    {{"u", solution_.data()}});
    vtu::write_pvtu(DiffusionProblemBasis<dim>::join(dir, base + ".pvtu"), {"u"}, {base + ".0000.vtu"});
  }

  // Coarse mesh with its interior vertices moved by up to amp/nc per axis
  // (msfem_gen_perturbed_quads, 0 < amp < 0.25, or msfem_gen_perturbed_hexes,
  // 0 < amp < 1/6): every cell is then a general quadrilateral / hexahedron.
  // The boundary stays the unit square's / cube's, so the coarse boundary data
  // is unchanged.  amp = 0: the uniform mesh.
  void set_mesh_distortion(double amp, uint64_t seed)
  {
    const double amax = dim == 2 ? 0.25 : 1.0 / 6.0;
    if (amp != 0 && !(amp > 0 && amp < amax))
      throw Error(MSFEM_ERR_ARG, "mesh distortion: 0 < amp < 0.25 (2D) or 1/6 (3D)");
    distortion_ = amp;
    distortion_seed_ = seed;
  }

  // Bases of every cell written by run() (the reference's per-cell output flag).
  void set_output_flag(bool flag, const std::string &dir = "")
  {
    output_all_ = flag;
    output_dir_ = dir;
  }

  const std::vector<double> &coarse_solution() const { return solution_; }
  int coarse_iterations() const { return coarse_iterations_; }
  const std::map<int, DiffusionProblemBasis<dim>> &cells() const { return cells_; }

private:
  Engine &engine_;
  unsigned n_refine_, n_refine_local_;
  BasisBatch<dim> batch_;
  int bc_kind_;
  std::map<int, DiffusionProblemBasis<dim>> cells_;
  std::vector<double> solution_;
  int32_t coarse_iterations_ = 0;
  bool output_all_ = false;
  std::string output_dir_;
  double distortion_ = 0;
  uint64_t distortion_seed_ = 0;
  std::vector<double> verts_; // moved coarse cells' vertices of the last run (general cells)
};

} // namespace msfem

#endif // MSFEM_BASIS_HPP_
