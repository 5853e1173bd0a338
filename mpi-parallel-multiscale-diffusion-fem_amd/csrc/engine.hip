// Host side of the MsFEM engine: context, batch planning by fine resolution,
// kernel sequencing and the C ABI of include/msfem.h.
//
// One msfem_batch_build call does, for all cells at once, what the reference
// does per cell in DiffusionProblemBasis<dim>::run()
// (include/base/diffusion_problem_basis.tpp:440-474), driven by the cell loop
// of DiffusionProblemMultiscale::initialize_and_compute_basis
// (include/base/diffusion_problem_ms.tpp:42-88).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "msfem.h"
#include "msfem_kernels.hpp"

using namespace msfem;

namespace {

thread_local std::string g_create_error;

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

// Batch layout derived from n_fine: classes by fine resolution and offsets.
struct Plan {
  int dim = 0, coef_kind = -1, nb = 0;
  bool general = false;                    // general quadrilaterals (msfem_batch.vertices)
  std::vector<int32_t> n_fine;
  std::vector<int> class_nf;               // distinct resolutions
  std::vector<std::vector<int>> class_cells;
  std::vector<long long> coef_off, bas_off, cond_off;
  long long coef_total = 0, bas_total = 0, cond_total = 0;
  size_t cells_dev_off[8] = {0};           // offsets (ints) of each class list in d_cells
  bool uploaded = false;
};

} // namespace

constexpr int kSideStreams = 4;

struct msfem_ctx {
  int device = 0;
  std::string err;
  hipStream_t stream = nullptr;
  std::map<std::string, DevBuf> bufs;
  std::map<long long, int> variants; // (dim, nf, precision, coef_kind) -> launch-shape id
  Plan plan;
  bool profiling = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr}; // solve begin/end, element begin/end
  bool ev_recorded = false;
  hipEvent_t marks[MSFEM_N_MARKS] = {}; // user marks (timing, cross-stream order)
  hipStream_t aux = nullptr;            // the caller-visible second stream
  // graph builds: side streams (one per size class), fork/join events, and the
  // instantiated graphs; `generation` changes whenever a buffer a graph may
  // hold is reallocated or the plan changes, which retires every cached graph
  hipStream_t side[kSideStreams] = {};
  hipEvent_t fork = nullptr, join[kSideStreams] = {};
  std::map<std::string, hipGraphExec_t> graphs;
  unsigned long long generation = 0, graph_generation = 0;
  // The plan metadata and the scratch buffers are shared by every build of the
  // context: `scratch_done` is recorded on a build's stream after its last
  // kernel, so a later build on another stream waits for it, and a rewrite of
  // the metadata or a reallocation waits for it on the host first.
  hipEvent_t scratch_done = nullptr;
  hipStream_t scratch_stream = nullptr;
  bool scratch_recorded = false;
  // the current build computes the energy terms (stopping rule or bound), and
  // whether its last basis comes from the partition of unity
  bool energy_terms = false;
  int energy_pou = 0;
};

namespace {

int fail(msfem_ctx *ctx, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_create_error = buf;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                                  \
  do                                                                                                        \
    {                                                                                                       \
      hipError_t e_ = (expr);                                                                               \
      if (e_ != hipSuccess)                                                                                 \
        return fail(ctx, MSFEM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                              \
    }                                                                                                       \
  while (0)

// Host wait until the last build that used the context's shared scratch and
// metadata has finished (before they are freed or rewritten).
int wait_scratch_idle(msfem_ctx *ctx)
{
  if (ctx->scratch_recorded)
    HIP_TRY(ctx, hipEventSynchronize(ctx->scratch_done));
  return MSFEM_OK;
}

// Orders work about to be enqueued on `st` after the last user of the shared
// scratch when that ran on another stream (not under the caller's graph capture,
// where an outside event cannot be waited on; a capturing caller keeps to one
// stream).
int order_after_scratch(msfem_ctx *ctx, hipStream_t st, bool *capturing)
{
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  HIP_TRY(ctx, hipStreamIsCapturing(st, &cap));
  *capturing = cap != hipStreamCaptureStatusNone;
  if (ctx->scratch_recorded && ctx->scratch_stream != st && !*capturing)
    HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->scratch_done, 0));
  return MSFEM_OK;
}

// Marks the end of the shared scratch's use by the work just enqueued on `st`.
int retire_scratch(msfem_ctx *ctx, hipStream_t st, bool capturing)
{
  if (capturing)
    return MSFEM_OK;
  if (!ctx->scratch_done)
    HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->scratch_done, hipEventDisableTiming));
  HIP_TRY(ctx, hipEventRecord(ctx->scratch_done, st));
  ctx->scratch_stream = st;
  ctx->scratch_recorded = true;
  return MSFEM_OK;
}

// Grow-only device scratch owned by the context.
int scratch(msfem_ctx *ctx, const char *name, size_t bytes, void **out)
{
  DevBuf &b = ctx->bufs[name];
  if (b.bytes < bytes)
    {
      ++ctx->generation;
      if (b.p)
        {
          int rc = wait_scratch_idle(ctx);
          if (rc)
            return rc;
          HIP_TRY(ctx, hipFree(b.p));
        }
      b.p = nullptr;
      b.bytes = 0;
      HIP_TRY(ctx, hipMalloc(&b.p, bytes > 0 ? bytes : 1));
      b.bytes = bytes;
    }
  *out = b.p;
  return MSFEM_OK;
}

int coef_stride(int dim, int kind)
{
  const int nq = 1 << dim, ns = dim == 2 ? 3 : 6;
  switch (kind)
    {
      case MSFEM_COEF_SCALAR_CELL: return 1;
      case MSFEM_COEF_TENSOR_CELL: return ns;
      case MSFEM_COEF_SCALAR_QP: return nq;
      case MSFEM_COEF_TENSOR_QP: return nq * ns;
      default: return -1;
    }
}

long long ipow(long long b, int e)
{
  long long r = 1;
  while (e-- > 0)
    r *= b;
  return r;
}

long long variant_key(int dim, int nf, int precision, int coef_kind)
{
  return ((long long)dim * 1024 + nf) * 64 + precision * 8 + coef_kind;
}

// Resolutions with tuned kernels (launch-shape variants) ...
bool nf_supported(int dim, int nf)
{
  if (dim == 2)
    return nf == 16 || nf == 32 || nf == 64 || nf == 128;
  return nf == 8 || nf == 16;
}

// ... and every resolution the build accepts (the others run the generic solve).
bool nf_valid(int dim, int nf) { return nf >= 1 && nf <= (dim == 2 ? 1024 : 64); }

// Local-solve storage mode of a class.  Scalar and tensor coefficients per
// fine cell of box cells are cached in registers (up to 64^2); everything
// else -- QP coefficients, general quadrilaterals (whose geometry lives in the
// precomputed couplings) and the 128^2 tensor tiles -- streams node couplings
// from planes (64^2 and 128^2) or reads per-fine-cell conductances (16^2,
// 32^2, whose lane groups hold several problems per workgroup).
int class_mode(int dim, int coef_kind, int nf, bool general)
{
  if (!nf_supported(dim, nf) || (dim == 3 && (coef_kind != MSFEM_COEF_SCALAR_CELL || general)))
    return kModeGeneric;
  if (dim == 3)
    return kModeScalar;
  if (!general && coef_kind == MSFEM_COEF_SCALAR_CELL)
    return kModeScalar;
  if (!general && coef_kind == MSFEM_COEF_TENSOR_CELL && nf <= 64)
    return kModeTensor;
  if (nf >= 64)
    return kModeStream;
  return kModeGlobal;
}

// A quadrilateral (deal.II vertex order) whose bilinear map has a positive
// Jacobian determinant everywhere: the determinant is affine in each parametric
// variable, so its four corner values decide (convex, positively oriented).
bool quad_ok(const double *v)
{
  const double ax = v[2] - v[0], ay = v[3] - v[1], bx = v[4] - v[0], by = v[5] - v[1];
  const double cx = v[6] - v[4] - v[2] + v[0], cy = v[7] - v[5] - v[3] + v[1];
  auto det = [&](double xi, double eta) { return (ax + eta * cx) * (by + xi * cy) - (ay + eta * cy) * (bx + xi * cx); };
  const double d[4] = {det(0, 0), det(1, 0), det(0, 1), det(1, 1)};
  for (double x : d)
    if (!(x > 0) || !std::isfinite(x))
      return false;
  return true;
}

// A hexahedron (deal.II vertex order) whose trilinear map has a positive
// Jacobian determinant everywhere, certified by a sufficient test: det J is of
// degree <= 2 in each parametric variable, and its 27 Bernstein coefficients on
// [0,1]^3 are all positive (then so is det J, a convex combination of them).
// Column l of J is multilinear in the other two variables with the cell's four
// edge vectors along l as Bernstein coefficients, so the coefficient at
// (a, b, c) in {0,1,2}^3 is the weighted mean of the triple products
// e_x(j,k) . (e_y(i,k') x e_z(i',j')) over i + i' = a, j + j' = b, k + k' = c
// (weight 1/2 for each mixed pair, from t(1-t) = B_1^2 / 2).  The eight corner
// coefficients are the corner determinants, so every negatively oriented or
// degenerate corner is rejected.
bool hex_ok(const double *v)
{
  auto P = [&](int x, int y, int z) { return v + 3 * (x + 2 * y + 4 * z); };
  auto edge = [&](const double *p, const double *q, double (&e)[3]) {
    for (int k = 0; k < 3; ++k)
      e[k] = q[k] - p[k];
  };
  double ex[2][2][3], ey[2][2][3], ez[2][2][3]; // ex[j][k]: edge along x at y = j, z = k; etc.
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      {
        edge(P(0, i, j), P(1, i, j), ex[i][j]);
        edge(P(i, 0, j), P(i, 1, j), ey[i][j]);
        edge(P(i, j, 0), P(i, j, 1), ez[i][j]);
      }
  for (int a = 0; a <= 2; ++a)
    for (int b = 0; b <= 2; ++b)
      for (int c = 0; c <= 2; ++c)
        {
          double sum = 0;
          for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
              for (int k = 0; k < 2; ++k)
                {
                  const int i2 = a - i, j2 = b - j, k2 = c - k;
                  if (i2 < 0 || i2 > 1 || j2 < 0 || j2 > 1 || k2 < 0 || k2 > 1)
                    continue;
                  const double *e1 = ex[j][k], *e2 = ey[i][k2], *e3 = ez[i2][j2];
                  const double t = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                                   e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
                  sum += t * (i != i2 ? 0.5 : 1.0) * (j != j2 ? 0.5 : 1.0) * (k != k2 ? 0.5 : 1.0);
                }
          if (!(sum > 0) || !std::isfinite(sum))
            return false;
        }
  return true;
}

int validate(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o)
{
  if (!in || !o)
    return fail(ctx, MSFEM_ERR_ARG, "null batch or options");
  if (in->dim != 2 && in->dim != 3)
    return fail(ctx, MSFEM_ERR_ARG, "dim must be 2 or 3 (got %d)", in->dim);
  if (in->n_cells < 0 || (in->n_cells > 0 && (!in->n_fine || (!in->box && !in->vertices) || !in->coef)))
    return fail(ctx, MSFEM_ERR_ARG, "n_cells < 0 or a required array is NULL");
  if (coef_stride(in->dim, in->coef_kind) < 0)
    return fail(ctx, MSFEM_ERR_ARG, "unknown coefficient kind %d", in->coef_kind);
  if (in->rhs_kind != MSFEM_RHS_CONST && in->rhs_kind != MSFEM_RHS_SINSIN)
    return fail(ctx, MSFEM_ERR_ARG, "unknown rhs kind %d", in->rhs_kind);
  if (o->precision != MSFEM_PREC_FP64 && o->precision != MSFEM_PREC_FP32_REFINE)
    return fail(ctx, MSFEM_ERR_ARG, "unknown precision %d", o->precision);
  if (o->variant < 0)
    return fail(ctx, MSFEM_ERR_ARG, "variant must be >= 0");
  for (int c = 0; c < in->n_cells && o->variant > 0; ++c)
    {
      if (class_mode(in->dim, in->coef_kind, in->n_fine[c], in->vertices != nullptr) == kModeGeneric)
        continue; // the generic solve has one launch shape
      const int n = in->dim == 2 ? local2d_variants(in->n_fine[c]) : local3d_variants(in->n_fine[c]);
      if (o->variant > n)
        return fail(ctx, MSFEM_ERR_ARG, "variant %d does not exist for %d fine cells per side", o->variant,
                    in->n_fine[c]);
    }
  if (o->max_iter < 1 || o->max_refine < 1 || !(o->rel_tol >= 0) || !(o->abs_tol >= 0) ||
      (o->rel_tol == 0 && o->abs_tol == 0 && !(o->energy_tol > 0)) || !(o->energy_tol >= 0) ||
      !(o->energy_tol < 1))
    return fail(ctx, MSFEM_ERR_ARG, "invalid tolerances or iteration caps");
  if (o->energy_tol > 0 && in->vertices)
    return fail(ctx, MSFEM_ERR_ARG, "the energy stopping rule needs box cells (its eigenvalue bound is for boxes)");
  for (int c = 0; c < in->n_cells; ++c)
    if (!nf_valid(in->dim, in->n_fine[c]))
      return fail(ctx, MSFEM_ERR_ARG, "cell %d: unsupported fine resolution %d for dim %d (2D: 1..1024, 3D: 1..64)",
                  c, in->n_fine[c], in->dim);
  return MSFEM_OK;
}

int make_plan(msfem_ctx *ctx, const msfem_batch *in)
{
  Plan &p = ctx->plan;
  const bool general = in->vertices != nullptr;
  const bool same = p.uploaded && p.dim == in->dim && p.coef_kind == in->coef_kind && p.general == general &&
                    p.n_fine.size() == (size_t)in->n_cells &&
                    std::equal(p.n_fine.begin(), p.n_fine.end(), in->n_fine);
  if (same)
    return MSFEM_OK;
  ++ctx->generation;
  p = Plan();
  p.dim = in->dim;
  p.coef_kind = in->coef_kind;
  p.general = general;
  p.nb = 1 << in->dim;
  p.n_fine.assign(in->n_fine, in->n_fine + in->n_cells);
  const int stride = coef_stride(in->dim, in->coef_kind);
  p.coef_off.resize(in->n_cells);
  p.bas_off.resize(in->n_cells);
  p.cond_off.resize(in->n_cells);
  std::map<int, std::vector<int>> cls;
  for (int c = 0; c < in->n_cells; ++c)
    {
      const int nf = in->n_fine[c];
      p.coef_off[c] = p.coef_total;
      p.coef_total += ipow(nf, in->dim) * stride;
      p.bas_off[c] = p.bas_total;
      p.bas_total += (long long)p.nb * ipow(nf + 1, in->dim);
      p.cond_off[c] = p.cond_total;
      const int mode = class_mode(in->dim, in->coef_kind, nf, general);
      if (mode == kModeGlobal) // 6 FP64 conductances per fine cell
        p.cond_total += 6 * ipow(nf, 2);
      if (mode == kModeGeneric) // QGauss(2) element matrix per fine cell
        p.cond_total += (long long)generic_elem_doubles(in->dim, nf);
      if (mode == kModeStream) // 4 FP64 + 4 FP32 coupling planes
        p.cond_total += 6 * (long long)coupling_plane_stride(nf);
      cls[nf].push_back(c);
    }
  std::vector<int> all_cells;
  int k = 0;
  for (auto &kv : cls)
    {
      p.class_nf.push_back(kv.first);
      p.cells_dev_off[k++] = all_cells.size();
      all_cells.insert(all_cells.end(), kv.second.begin(), kv.second.end());
      p.class_cells.push_back(kv.second);
    }
  // device metadata: cells | coef_off | bas_off | cond_off
  const size_t n = in->n_cells;
  const size_t bytes = n * sizeof(int) + 3 * n * sizeof(long long) + 64;
  void *d;
  int rc = scratch(ctx, "meta", bytes, &d);
  if (rc)
    return rc;
  std::vector<char> host(bytes);
  size_t off = 0;
  memcpy(host.data(), all_cells.data(), n * sizeof(int));
  off = ((n * sizeof(int) + 7) / 8) * 8;
  memcpy(host.data() + off, p.coef_off.data(), n * sizeof(long long));
  memcpy(host.data() + off + n * sizeof(long long), p.bas_off.data(), n * sizeof(long long));
  memcpy(host.data() + off + 2 * n * sizeof(long long), p.cond_off.data(), n * sizeof(long long));
  // an earlier build (possibly on another stream) may still read the old
  // metadata: wait for it, then copy synchronously
  rc = wait_scratch_idle(ctx);
  if (rc)
    return rc;
  HIP_TRY(ctx, hipMemcpy(d, host.data(), bytes, hipMemcpyHostToDevice));
  p.uploaded = true;
  return MSFEM_OK;
}

const int *d_cells(msfem_ctx *ctx) { return (const int *)ctx->bufs["meta"].p; }
const long long *d_off(msfem_ctx *ctx, int which) // 0 coef, 1 bas, 2 cond
{
  const size_t n = ctx->plan.n_fine.size();
  const size_t off = ((n * sizeof(int) + 7) / 8) * 8;
  return (const long long *)((const char *)ctx->bufs["meta"].p + off + which * n * sizeof(long long));
}

__global__ void fill_kernel(void *p, size_t n_words, uint32_t v)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_words; i += (size_t)gridDim.x * blockDim.x)
    ((uint32_t *)p)[i] = v;
}

hipError_t fill_zero(void *p, size_t bytes, hipStream_t st)
{
  if (!p || bytes == 0)
    return hipSuccess;
  const size_t words = bytes / 4;
  size_t grid = (words + 255) / 256;
  if (grid > 4096)
    grid = 4096;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)grid), dim3(256), 0, st, p, words, 0u);
  return hipGetLastError();
}

// Local solves (coefficient preparation, solve, partition-of-unity fix-up) of size class k.
// Storage mode of size class nf in the current batch build.
int build_mode(const msfem_batch *in, int nf)
{
  return class_mode(in->dim, in->coef_kind, nf, in->vertices != nullptr);
}

// Launch-shape id of size class k (the explicit option, else the table's entry).
int class_variant(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, int nf)
{
  int variant = o->variant;
  if (variant == 0)
    {
      auto it = ctx->variants.find(variant_key(in->dim, nf, o->precision, in->coef_kind));
      if (it != ctx->variants.end())
        variant = it->second;
    }
  return variant;
}

// Whether size class nf runs the two-level local solve.
bool class_coarse(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, int nf)
{
  return o->coarse_space && in->dim == 2 && build_mode(in, nf) != kModeGeneric &&
         local2d_coarse_space(nf, build_mode(in, nf), class_variant(ctx, in, o, nf));
}

// Whether size class nf's two-level solve also takes the intermediate hat
// levels (coarse_space 1): spacing 2 and 4 on the 64^2 classes, spacing 4 and
// 8 on the FP32 128^2 classes (streamed form).
bool class_ml(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, int nf)
{
  const int mode = build_mode(in, nf);
  const bool fits = (nf == 64 && (mode == kModeScalar || mode == kModeTensor || mode == kModeStream)) ||
                    (nf == 128 && mode == kModeStream && o->precision == MSFEM_PREC_FP32_REFINE);
  return o->coarse_space == 1 && fits && class_coarse(ctx, in, o, nf);
}

// Floats of the intermediate levels' inverse diagonals before size class k_end.
size_t ml_floats(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, size_t k_end)
{
  const Plan &p = ctx->plan;
  size_t n = 0;
  for (size_t k = 0; k < k_end && k < p.class_nf.size(); ++k)
    if (class_ml(ctx, in, o, p.class_nf[k]))
      n += p.class_cells[k].size() * (size_t)kMlFloats;
  return n;
}

// Doubles of the two-level classes' inverse coarse matrices before size class
// k_end (each such class owns a slice: concurrent graph branches share none).
size_t coarse_inv_doubles(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, size_t k_end)
{
  const Plan &p = ctx->plan;
  size_t n = 0;
  for (size_t k = 0; k < k_end && k < p.class_nf.size(); ++k)
    if (class_coarse(ctx, in, o, p.class_nf[k]))
      n += p.class_cells[k].size() * (size_t)kCoarseDofs * kCoarseDofs;
  return n;
}

// Doubles of the FP64 streamed classes' q slots before size class k_end (each
// such class owns a slice of one work array, so concurrent graph branches never
// share slots).
size_t stream_work_doubles(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, size_t k_end)
{
  const Plan &p = ctx->plan;
  size_t n = 0;
  if (o->precision != MSFEM_PREC_FP64)
    return 0;
  for (size_t k = 0; k < k_end && k < p.class_nf.size(); ++k)
    if (build_mode(in, p.class_nf[k]) == kModeStream)
      n += p.class_cells[k].size() * (size_t)(o->partition_of_unity ? p.nb - 1 : p.nb) *
           stream_q_slots(p.class_nf[k]);
  return n;
}

// Doubles of the generic classes' per-problem work vectors before size class k_end.
size_t generic_work_total(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, size_t k_end)
{
  const Plan &p = ctx->plan;
  size_t n = 0;
  for (size_t k = 0; k < k_end && k < p.class_nf.size(); ++k)
    if (build_mode(in, p.class_nf[k]) == kModeGeneric)
      n += p.class_cells[k].size() * (size_t)(o->partition_of_unity ? p.nb - 1 : p.nb) *
           generic_work_doubles(in->dim, p.class_nf[k]);
  return n;
}

// Per-class preparation before the solve: conductances / node couplings,
// general cells' boundary data, the two-level inverse coarse matrices.
int launch_class_prepare(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, size_t k, double *bases,
                         double *cond, hipStream_t st)
{
  Plan &p = ctx->plan;
  const int nf = p.class_nf[k];
  const int ncls = (int)p.class_cells[k].size();
  const int *cells = d_cells(ctx) + p.cells_dev_off[k];
  const bool fp32 = o->precision == MSFEM_PREC_FP32_REFINE;
  const int mode = build_mode(in, nf);
  if (mode == kModeGlobal)
    HIP_TRY(ctx, launch_prep_cond_2d(ncls, cells, nf, in->coef_kind, in->box, in->vertices, in->coef, d_off(ctx, 0),
                                     cond, d_off(ctx, 2), st));
  if (mode == kModeStream)
    HIP_TRY(ctx, launch_prep_couplings_2d(ncls, cells, nf, in->coef_kind, in->box, in->vertices, in->coef,
                                          d_off(ctx, 0), cond, d_off(ctx, 2), fp32, st));
  if (mode == kModeGeneric)
    HIP_TRY(ctx, launch_generic_prep(in->dim, ncls, cells, nf, in->coef_kind, in->box, in->vertices, in->coef,
                                     d_off(ctx, 0), cond, d_off(ctx, 2), st));
  if (in->vertices && in->dim == 2) // the solve kernels read the boundary data from the bases
    HIP_TRY(ctx, launch_quad_dirichlet_2d(ncls, cells, nf, in->vertices, bases, d_off(ctx, 1), st));
  else if (in->vertices)
    HIP_TRY(ctx, launch_hex_dirichlet_3d(ncls, cells, nf, in->vertices, bases, d_off(ctx, 1), st));
  if (class_coarse(ctx, in, o, nf))
    {
      double *base = (double *)ctx->bufs["coarse_inv"].p;
      if (!base)
        return fail(ctx, MSFEM_ERR_HIP, "two-level scratch missing");
      HIP_TRY(ctx, launch_coarse_setup_2d(ncls, cells, nf, in->coef_kind, in->box, in->vertices, in->coef,
                                          d_off(ctx, 0), base + coarse_inv_doubles(ctx, in, o, k), st));
    }
  if (class_ml(ctx, in, o, nf))
    {
      float *mb = (float *)ctx->bufs["ml_inv"].p;
      if (!mb)
        return fail(ctx, MSFEM_ERR_HIP, "intermediate-level scratch missing");
      HIP_TRY(ctx, launch_ml_diag_2d(ncls, cells, nf, in->coef_kind, in->box, in->vertices, in->coef, d_off(ctx, 0),
                                     mb + ml_floats(ctx, in, o, k), st));
    }
  if (ctx->energy_terms && !in->vertices)
    HIP_TRY(ctx, launch_energy_terms(in->dim, ncls, cells, nf, in->coef_kind, in->box, in->coef, d_off(ctx, 0),
                                     o->energy_tol, (double *)ctx->bufs["lam_lb"].p,
                                     o->energy_tol > 0 ? (double *)ctx->bufs["ptol"].p : nullptr, st));
  return MSFEM_OK;
}

// The local solves of class k (after launch_class_prepare), then the
// partition-of-unity fix-up.
int launch_class_solve(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, msfem_batch_out *out,
                       size_t k, double *bases, double *cond, double *work, hipStream_t st)
{
  Plan &p = ctx->plan;
  const int nb = p.nb;
  const bool pou = o->partition_of_unity != 0;
  const int nf = p.class_nf[k];
  const int ncls = (int)p.class_cells[k].size();
  const int *cells = d_cells(ctx) + p.cells_dev_off[k];
  const bool fp32 = o->precision == MSFEM_PREC_FP32_REFINE;
  const int mode = build_mode(in, nf);
  LocalArgs a;
  a.nb_solve = pou ? nb - 1 : nb;
  a.n_problems = ncls * a.nb_solve;
  a.nb = nb;
  a.cells = cells;
  a.box = in->box;
  a.verts = in->vertices;
  a.coef = in->coef;
  a.coef_off = d_off(ctx, 0);
  a.cond = cond;
  a.cond32 = (const float *)cond;
  a.work = work ? work + stream_work_doubles(ctx, in, o, k) : nullptr;
  a.cond_off = d_off(ctx, 2);
  a.bases = bases;
  a.bas_off = d_off(ctx, 1);
  a.x0 = in->x0;
  a.q1_start = o->q1_start;
  a.iters = out->iters;
  a.status = out->status;
  a.resid = out->resid;
  a.rel_tol = o->rel_tol;
  a.abs_tol = o->abs_tol;
  a.inner_rel_tol = o->inner_rel_tol;
  a.max_iter = o->max_iter;
  a.max_refine = o->max_refine;
  a.ptol = o->energy_tol > 0 ? (const double *)ctx->bufs["ptol"].p : nullptr;
  const int variant = class_variant(ctx, in, o, nf);
  a.cinv = nullptr;
  if (class_coarse(ctx, in, o, nf))
    a.cinv = (const double *)ctx->bufs["coarse_inv"].p + coarse_inv_doubles(ctx, in, o, k);
  a.mlinv = nullptr;
  if (class_ml(ctx, in, o, nf))
    a.mlinv = (const float *)ctx->bufs["ml_inv"].p + ml_floats(ctx, in, o, k);
  if (mode == kModeGeneric)
    {
      double *gw = (double *)ctx->bufs["gen_work"].p;
      if (!gw)
        return fail(ctx, MSFEM_ERR_HIP, "generic-solve scratch missing");
      HIP_TRY(ctx, launch_generic_solve(in->dim, nf, a, gw + generic_work_total(ctx, in, o, k), st));
    }
  else if (in->dim == 2)
    HIP_TRY(ctx, launch_local_solve_2d(fp32, mode, nf, variant, a, st));
  else
    HIP_TRY(ctx, launch_local_solve_3d(fp32, nf, variant, a, st));
  if (pou)
    HIP_TRY(ctx, launch_pou_fixup(ncls, cells, in->dim, nf, bases, d_off(ctx, 1), st));
  return MSFEM_OK;
}

// A^K, F^K of size class k from its bases.
int launch_class_element(msfem_ctx *ctx, const msfem_batch *in, msfem_batch_out *out, size_t k, double *bases,
                         double *cond, hipStream_t st)
{
  Plan &p = ctx->plan;
  ElementArgs e;
  e.n_cells_class = (int)p.class_cells[k].size();
  e.cells = d_cells(ctx) + p.cells_dev_off[k];
  e.dim = in->dim;
  e.nf = p.class_nf[k];
  e.nb = p.nb;
  e.coef_kind = in->coef_kind;
  e.box = in->box;
  e.verts = in->vertices;
  e.coef = in->coef;
  e.coef_off = d_off(ctx, 0);
  e.cond = cond;
  e.cond_off = d_off(ctx, 2);
  e.bases = bases;
  e.bas_off = d_off(ctx, 1);
  e.rhs_kind = in->rhs_kind;
  e.rhs_value = in->rhs_value;
  e.Ae = out->Ae;
  e.Fe = out->Fe;
  HIP_TRY(ctx, launch_element(e, st));
  if (out->energy_bound)
    HIP_TRY(ctx, launch_energy_bound(e.n_cells_class, e.cells, p.nb, ctx->energy_pou,
                                     in->vertices ? nullptr : (const double *)ctx->bufs["lam_lb"].p, out->resid,
                                     out->Ae, out->energy_bound, st));
  return MSFEM_OK;
}

void drop_graphs(msfem_ctx *ctx)
{
  for (auto &kv : ctx->graphs)
    hipGraphExecDestroy(kv.second);
  ctx->graphs.clear();
}

// The whole build as one HIP graph: a branch per size class (solve, then
// element matrices), all branches concurrent, joined on `st`.  Instantiated
// once per (plan, scratch buffers, arguments, options) and replayed after.
int run_graph(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, msfem_batch_out *out,
              double *bases, double *cond, double *work, hipStream_t st)
{
  Plan &p = ctx->plan;
  if (ctx->graph_generation != ctx->generation)
    {
      drop_graphs(ctx);
      ctx->graph_generation = ctx->generation;
    }
  // the key is built from the named fields (never from struct bytes, whose
  // padding a C caller may leave uninitialised); the n_fine contents enter
  // through the generation, which changes with the plan
  std::string key;
  auto add = [&key](const auto &v) { key.append((const char *)&v, sizeof v); };
  add(ctx->generation);
  add(in->dim), add(in->n_cells), add(in->box), add(in->coef_kind), add(in->coef), add(in->rhs_kind);
  add(in->rhs_value), add(in->x0), add(in->vertices);
  add(o->precision), add(o->max_iter), add(o->max_refine), add(o->partition_of_unity), add(o->rel_tol);
  add(o->abs_tol), add(o->inner_rel_tol), add(o->variant), add(o->coarse_space), add(o->q1_start);
  add(o->energy_tol);
  add(out->Ae), add(out->Fe), add(out->bases), add(out->iters), add(out->status), add(out->resid);
  add(out->energy_bound);
  add(bases), add(st);
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end())
    {
      const size_t nk = p.class_nf.size();
      for (size_t k = 0; k < nk && k < kSideStreams; ++k)
        if (!ctx->side[k])
          HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->side[k], hipStreamNonBlocking));
      if (!ctx->fork)
        {
          HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->fork, hipEventDisableTiming));
          for (auto &e : ctx->join)
            HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
      HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = MSFEM_OK;
      hipError_t e = hipEventRecord(ctx->fork, st);
      const bool pou = o->partition_of_unity != 0;
      const size_t nstat = (size_t)in->n_cells * p.nb;
      if (e == hipSuccess && pou)
        {
          e = fill_zero(out->iters, nstat * sizeof(int32_t), st);
          if (e == hipSuccess)
            e = fill_zero(out->status, nstat * sizeof(int32_t), st);
          if (e == hipSuccess)
            e = fill_zero(out->resid, nstat * sizeof(double), st);
          if (e == hipSuccess)
            e = hipEventRecord(ctx->fork, st);
        }
      for (size_t k = 0; k < nk && e == hipSuccess && rc == MSFEM_OK; ++k)
        {
          hipStream_t s = ctx->side[k % kSideStreams];
          e = hipStreamWaitEvent(s, ctx->fork, 0);
          if (e != hipSuccess)
            break;
          rc = launch_class_prepare(ctx, in, o, k, bases, cond, s);
          if (rc == MSFEM_OK)
            rc = launch_class_solve(ctx, in, o, out, k, bases, cond, work, s);
          if (rc == MSFEM_OK)
            rc = launch_class_element(ctx, in, out, k, bases, cond, s);
          if (rc == MSFEM_OK)
            e = hipEventRecord(ctx->join[k % kSideStreams], s);
          if (e == hipSuccess && rc == MSFEM_OK)
            e = hipStreamWaitEvent(st, ctx->join[k % kSideStreams], 0);
        }
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &g);
      if (rc != MSFEM_OK)
        {
          if (g)
            hipGraphDestroy(g);
          return rc;
        }
      if (e != hipSuccess || ec != hipSuccess)
        {
          if (g)
            hipGraphDestroy(g);
          return fail(ctx, MSFEM_ERR_HIP, "graph capture failed: %s",
                      hipGetErrorString(e != hipSuccess ? e : ec));
        }
      hipGraphExec_t x = nullptr;
      e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
      hipGraphDestroy(g);
      if (e != hipSuccess)
        return fail(ctx, MSFEM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      if (ctx->graphs.size() >= 16)
        drop_graphs(ctx);
      it = ctx->graphs.emplace(key, x).first;
    }
  ctx->ev_recorded = false; // graph builds are timed from outside (msfem_ctx_mark)
  HIP_TRY(ctx, hipGraphLaunch(it->second, st));
  return MSFEM_OK;
}

int enqueue_build(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, msfem_batch_out *out,
                  hipStream_t st);

int build_device(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, msfem_batch_out *out,
                 hipStream_t st)
{
  int rc = validate(ctx, in, o);
  if (rc)
    return rc;
  if (in->n_cells == 0) // nothing to build or write (the outputs may be NULL)
    return MSFEM_OK;
  if (!out || !out->Ae || !out->Fe)
    return fail(ctx, MSFEM_ERR_ARG, "Ae and Fe outputs are required");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  bool capturing = false;
  rc = order_after_scratch(ctx, st, &capturing);
  if (rc)
    return rc;
  rc = enqueue_build(ctx, in, o, out, st);
  if (rc)
    return rc;
  return retire_scratch(ctx, st, capturing);
}

// The build's kernels on `st` (arguments validated, device current).
int enqueue_build(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *o, msfem_batch_out *out,
                  hipStream_t st)
{
  int rc = make_plan(ctx, in);
  if (rc)
    return rc;
  Plan &p = ctx->plan;
  double *bases = out->bases;
  if (!bases)
    {
      rc = scratch(ctx, "bases", p.bas_total * sizeof(double), (void **)&bases);
      if (rc)
        return rc;
    }
  double *cond = nullptr;
  if (p.cond_total > 0)
    {
      rc = scratch(ctx, "cond", p.cond_total * sizeof(double), (void **)&cond);
      if (rc)
        return rc;
    }
  // inverse coarse matrices of the two-level classes (allocated before any capture)
  const size_t cinv_doubles = coarse_inv_doubles(ctx, in, o, p.class_nf.size());
  if (cinv_doubles > 0)
    {
      void *d;
      rc = scratch(ctx, "coarse_inv", cinv_doubles * sizeof(double), &d);
      if (rc)
        return rc;
    }
  // inverse diagonals of the intermediate levels (allocated before any capture)
  const size_t mlf = ml_floats(ctx, in, o, p.class_nf.size());
  if (mlf > 0)
    {
      void *d;
      rc = scratch(ctx, "ml_inv", mlf * sizeof(float), &d);
      if (rc)
        return rc;
    }
  // per-problem vectors of the generic classes (allocated before any capture)
  const size_t gen_doubles = generic_work_total(ctx, in, o, p.class_nf.size());
  if (gen_doubles > 0)
    {
      void *d;
      rc = scratch(ctx, "gen_work", gen_doubles * sizeof(double), &d);
      if (rc)
        return rc;
    }
  // energy terms (stopping rule and/or a posteriori bound; allocated before any capture)
  msfem_batch_out eout = *out;
  out = &eout;
  ctx->energy_terms = (o->energy_tol > 0 || out->energy_bound) && !in->vertices;
  ctx->energy_pou = o->partition_of_unity != 0;
  if (ctx->energy_terms)
    {
      void *d;
      rc = scratch(ctx, "lam_lb", (size_t)in->n_cells * sizeof(double), &d);
      if (rc == MSFEM_OK && o->energy_tol > 0)
        rc = scratch(ctx, "ptol", (size_t)in->n_cells * p.nb * sizeof(double), &d);
      if (rc)
        return rc;
    }
  if (out->energy_bound && !out->resid) // the bound needs the final residual norms
    {
      rc = scratch(ctx, "resid_e", (size_t)in->n_cells * p.nb * sizeof(double), (void **)&eout.resid);
      if (rc)
        return rc;
    }
  // global q slots of the FP64 streamed classes (allocated before any capture)
  double *work = nullptr;
  const size_t work_doubles = stream_work_doubles(ctx, in, o, p.class_nf.size());
  if (work_doubles > 0)
    {
      rc = scratch(ctx, "stream_q", work_doubles * sizeof(double), (void **)&work);
      if (rc)
        return rc;
    }
  if (o->graph)
    return run_graph(ctx, in, o, out, bases, cond, work, st);
  const bool pou = o->partition_of_unity != 0;
  const size_t nstat = (size_t)in->n_cells * p.nb;
  if (pou)
    {
      HIP_TRY(ctx, fill_zero(out->iters, nstat * sizeof(int32_t), st));
      HIP_TRY(ctx, fill_zero(out->status, nstat * sizeof(int32_t), st));
      HIP_TRY(ctx, fill_zero(out->resid, nstat * sizeof(double), st));
    }
  // all classes' preparation, then their local solves (the profiled span), then
  // all element matrices
  for (size_t k = 0; k < p.class_nf.size(); ++k)
    {
      rc = launch_class_prepare(ctx, in, o, k, bases, cond, st);
      if (rc)
        return rc;
    }
  if (ctx->profiling)
    HIP_TRY(ctx, hipEventRecord(ctx->ev[0], st));
  for (size_t k = 0; k < p.class_nf.size(); ++k)
    {
      rc = launch_class_solve(ctx, in, o, out, k, bases, cond, work, st);
      if (rc)
        return rc;
    }
  if (ctx->profiling)
    {
      HIP_TRY(ctx, hipEventRecord(ctx->ev[1], st));
      HIP_TRY(ctx, hipEventRecord(ctx->ev[2], st));
    }
  for (size_t k = 0; k < p.class_nf.size(); ++k)
    {
      rc = launch_class_element(ctx, in, out, k, bases, cond, st);
      if (rc)
        return rc;
    }
  if (ctx->profiling)
    {
      HIP_TRY(ctx, hipEventRecord(ctx->ev[3], st));
      ctx->ev_recorded = true;
    }
  return MSFEM_OK;
}

struct HostStage {
  msfem_ctx *ctx;
  int rc = MSFEM_OK;
  template <typename T>
  T *up(const char *name, const T *src, size_t n)
  {
    void *d = nullptr;
    if (rc || !src)
      return nullptr;
    rc = scratch(ctx, name, n * sizeof(T), &d);
    if (rc)
      return nullptr;
    hipError_t e = hipMemcpyAsync(d, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess)
      rc = fail(ctx, MSFEM_ERR_HIP, "hipMemcpyAsync H2D: %s", hipGetErrorString(e));
    return (T *)d;
  }
  template <typename T>
  T *alloc(const char *name, size_t n, bool want)
  {
    void *d = nullptr;
    if (rc || !want)
      return nullptr;
    rc = scratch(ctx, name, n * sizeof(T), &d);
    return rc ? nullptr : (T *)d;
  }
  template <typename T>
  void down(T *dst, const T *src, size_t n)
  {
    if (rc || !dst || !src)
      return;
    hipError_t e = hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream);
    if (e != hipSuccess)
      rc = fail(ctx, MSFEM_ERR_HIP, "hipMemcpyAsync D2H: %s", hipGetErrorString(e));
  }
};

} // namespace

extern "C" {

int msfem_version(void) { return MSFEM_API_VERSION; }

const char *msfem_last_error(const msfem_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int msfem_uses_coarse_space(msfem_ctx *ctx, int dim, int nf, const msfem_solver_opts *opts, int coef_kind,
                            int general)
{
  if (!ctx || !opts)
    return 0;
  msfem_batch b = {};
  b.dim = dim;
  b.coef_kind = coef_kind;
  static const double one_vertex_set[8] = {0, 0, 1, 0, 0, 1, 1, 1};
  b.vertices = general ? one_vertex_set : nullptr;
  return class_ml(ctx, &b, opts, nf) ? 2 : class_coarse(ctx, &b, opts, nf) ? 1 : 0;
}

void msfem_default_opts(msfem_solver_opts *o)
{
  if (!o)
    return;
  memset(o, 0, sizeof(*o));
  o->precision = MSFEM_PREC_FP64;
  o->max_iter = 20000;
  o->max_refine = 20;
  o->partition_of_unity = 0;
  o->rel_tol = 1e-13;
  o->abs_tol = 0.0;
  o->inner_rel_tol = 1e-4;
  o->variant = 0;
  o->coarse_space = 1;
  o->q1_start = 1;
}

int msfem_ctx_create(int device, msfem_ctx **out)
{
  if (!out)
    return fail(nullptr, MSFEM_ERR_ARG, "out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0)
    return fail(nullptr, MSFEM_ERR_NODEVICE, "no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= n)
    return fail(nullptr, MSFEM_ERR_ARG, "device %d out of range (%d devices)", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess)
    return fail(nullptr, MSFEM_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  // this device's copy of the allocation-free entry point's scratch (once per device)
  e = noalloc_register_device(device);
  if (e != hipSuccess)
    return fail(nullptr, MSFEM_ERR_HIP, "hipGetSymbolAddress (fixed scratch): %s", hipGetErrorString(e));
  msfem_ctx *ctx = new msfem_ctx;
  ctx->device = device;
  e = hipStreamCreate(&ctx->stream);
  if (e != hipSuccess)
    {
      delete ctx;
      return fail(nullptr, MSFEM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
  *out = ctx;
  return MSFEM_OK;
}

int msfem_dev_alloc(msfem_ctx *ctx, uint64_t bytes, void **ptr)
{
  if (!ctx || !ptr)
    return fail(ctx, MSFEM_ERR_ARG, "null argument");
  *ptr = nullptr;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMalloc(ptr, bytes > 0 ? bytes : 1));
  return MSFEM_OK;
}

int msfem_dev_free(msfem_ctx *ctx, void *ptr)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if (!ptr)
    return MSFEM_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  HIP_TRY(ctx, hipFree(ptr));
  return MSFEM_OK;
}

int msfem_memcpy_h2d(msfem_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
  if (!ctx || (bytes && (!dst || !src)))
    return fail(ctx, MSFEM_ERR_ARG, "null argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MSFEM_OK;
}

int msfem_memcpy_d2h(msfem_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
  if (!ctx || (bytes && (!dst || !src)))
    return fail(ctx, MSFEM_ERR_ARG, "null argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MSFEM_OK;
}

void *msfem_ctx_stream(msfem_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int msfem_ctx_synchronize(msfem_ctx *ctx)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->aux)
    HIP_TRY(ctx, hipStreamSynchronize(ctx->aux));
  return MSFEM_OK;
}

int msfem_ctx_aux_stream(msfem_ctx *ctx, void **stream)
{
  if (!ctx || !stream)
    return fail(ctx, MSFEM_ERR_ARG, "null argument");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!ctx->aux)
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking)); // default priority
  *stream = ctx->aux;
  return MSFEM_OK;
}

int msfem_ctx_mark_on(msfem_ctx *ctx, int slot, void *stream)
{
  if (!ctx || slot < 0 || slot >= MSFEM_N_MARKS)
    return fail(ctx, MSFEM_ERR_ARG, "mark slot %d outside [0, %d)", slot, MSFEM_N_MARKS);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!ctx->marks[slot])
    HIP_TRY(ctx, hipEventCreate(&ctx->marks[slot]));
  HIP_TRY(ctx, hipEventRecord(ctx->marks[slot], stream ? (hipStream_t)stream : ctx->stream));
  return MSFEM_OK;
}

int msfem_ctx_wait_mark(msfem_ctx *ctx, int slot, void *stream)
{
  if (!ctx || slot < 0 || slot >= MSFEM_N_MARKS)
    return fail(ctx, MSFEM_ERR_ARG, "mark slot %d outside [0, %d)", slot, MSFEM_N_MARKS);
  if (!ctx->marks[slot])
    return fail(ctx, MSFEM_ERR_ARG, "mark slot %d never recorded", slot);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, hipStreamWaitEvent(stream ? (hipStream_t)stream : ctx->stream, ctx->marks[slot], 0));
  return MSFEM_OK;
}

int msfem_variant_count(int dim, int nf)
{
  if (!nf_supported(dim, nf))
    return 0;
  return dim == 2 ? local2d_variants(nf) : local3d_variants(nf);
}

int msfem_set_variant(msfem_ctx *ctx, int dim, int nf, int precision, int coef_kind, int variant)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if (!nf_supported(dim, nf) || variant < 0 || variant > msfem_variant_count(dim, nf) ||
      (precision != MSFEM_PREC_FP64 && precision != MSFEM_PREC_FP32_REFINE) || coef_kind < 0 || coef_kind > 3)
    return fail(ctx, MSFEM_ERR_ARG, "invalid variant table entry");
  const long long key = variant_key(dim, nf, precision, coef_kind);
  ++ctx->generation;
  if (variant == 0)
    ctx->variants.erase(key);
  else
    ctx->variants[key] = variant;
  return MSFEM_OK;
}

int msfem_ctx_set_profiling(msfem_ctx *ctx, int on)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (on && !ctx->ev[0])
    for (auto &e : ctx->ev)
      HIP_TRY(ctx, hipEventCreate(&e));
  ctx->profiling = on != 0;
  ctx->ev_recorded = false;
  return MSFEM_OK;
}

int msfem_ctx_mark(msfem_ctx *ctx, int slot)
{
  if (!ctx || slot < 0 || slot >= MSFEM_N_MARKS)
    return fail(ctx, MSFEM_ERR_ARG, "mark slot %d outside [0, %d)", slot, MSFEM_N_MARKS);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  if (!ctx->marks[slot])
    HIP_TRY(ctx, hipEventCreate(&ctx->marks[slot]));
  HIP_TRY(ctx, hipEventRecord(ctx->marks[slot], ctx->stream));
  return MSFEM_OK;
}

int msfem_ctx_elapsed(msfem_ctx *ctx, int from, int to, double *ms)
{
  if (!ctx || !ms || from < 0 || to < 0 || from >= MSFEM_N_MARKS || to >= MSFEM_N_MARKS)
    return fail(ctx, MSFEM_ERR_ARG, "bad mark slots or null output");
  if (!ctx->marks[from] || !ctx->marks[to])
    return fail(ctx, MSFEM_ERR_ARG, "mark slot never recorded");
  HIP_TRY(ctx, hipEventSynchronize(ctx->marks[to]));
  float t = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&t, ctx->marks[from], ctx->marks[to]));
  *ms = t;
  return MSFEM_OK;
}

int msfem_ctx_profile(msfem_ctx *ctx, double *ms_local_solve, double *ms_element)
{
  if (!ctx || !ms_local_solve || !ms_element)
    return fail(ctx, MSFEM_ERR_ARG, "null argument");
  if (!ctx->ev_recorded)
    return fail(ctx, MSFEM_ERR_ARG, "no profiled build recorded");
  HIP_TRY(ctx, hipEventSynchronize(ctx->ev[3]));
  float a = 0, b = 0;
  HIP_TRY(ctx, hipEventElapsedTime(&a, ctx->ev[0], ctx->ev[1]));
  HIP_TRY(ctx, hipEventElapsedTime(&b, ctx->ev[2], ctx->ev[3]));
  *ms_local_solve = a;
  *ms_element = b;
  return MSFEM_OK;
}

int msfem_ctx_destroy(msfem_ctx *ctx)
{
  if (!ctx)
    return MSFEM_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream)
    hipStreamSynchronize(ctx->stream);
  for (auto &e : ctx->ev)
    if (e)
      hipEventDestroy(e);
  for (auto &e : ctx->marks)
    if (e)
      hipEventDestroy(e);
  if (ctx->aux)
    {
      hipStreamSynchronize(ctx->aux);
      hipStreamDestroy(ctx->aux);
    }
  drop_graphs(ctx);
  for (auto &sd : ctx->side)
    if (sd)
      {
        hipStreamSynchronize(sd);
        hipStreamDestroy(sd);
      }
  if (ctx->fork)
    hipEventDestroy(ctx->fork);
  for (auto &e : ctx->join)
    if (e)
      hipEventDestroy(e);
  if (ctx->scratch_done)
    {
      hipEventSynchronize(ctx->scratch_done);
      hipEventDestroy(ctx->scratch_done);
    }
  for (auto &kv : ctx->bufs)
    if (kv.second.p)
      hipFree(kv.second.p);
  if (ctx->stream)
    hipStreamDestroy(ctx->stream);
  delete ctx;
  return MSFEM_OK;
}

int msfem_batch_build_device(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *opts,
                             msfem_batch_out *out, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  return build_device(ctx, in, opts, out, (hipStream_t)stream);
}

int msfem_batch_build(msfem_ctx *ctx, const msfem_batch *in, const msfem_solver_opts *opts, msfem_batch_out *out)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  int rc = validate(ctx, in, opts);
  if (rc)
    return rc;
  if (in->n_cells == 0) // nothing to build or write (the outputs may be NULL)
    return MSFEM_OK;
  if (!out || !out->Ae || !out->Fe)
    return fail(ctx, MSFEM_ERR_ARG, "Ae and Fe outputs are required");
  const size_t nvd = (size_t)(1 << in->dim) * in->dim; // doubles per general cell
  if (in->vertices)
    for (int c = 0; c < in->n_cells; ++c)
      if (in->dim == 2 ? !quad_ok(in->vertices + (size_t)c * nvd) : !hex_ok(in->vertices + (size_t)c * nvd))
        return fail(ctx, MSFEM_ERR_ARG, in->dim == 2 ? "cell %d: the vertices do not form a convex, positively "
                    "oriented quadrilateral (deal.II vertex order)" : "cell %d: the trilinear map of the vertices "
                    "is not certified positively oriented (deal.II vertex order; Jacobian determinant's Bernstein "
                    "coefficients must be > 0)", c);
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int dim = in->dim, nb = 1 << dim, n = in->n_cells;
  long long coef_total = 0, bas_total = 0;
  for (int c = 0; c < n; ++c)
    {
      coef_total += ipow(in->n_fine[c], dim) * coef_stride(dim, in->coef_kind);
      bas_total += nb * ipow(in->n_fine[c] + 1, dim);
    }
  HostStage s{ctx};
  msfem_batch din = *in;
  din.box = s.up("h_box", in->box, (size_t)n * 2 * dim);
  din.vertices = s.up("h_vertices", in->vertices, (size_t)n * nvd);
  din.coef = s.up("h_coef", in->coef, (size_t)coef_total);
  din.x0 = s.up("h_x0", in->x0, (size_t)bas_total);
  msfem_batch_out dout;
  dout.Ae = s.alloc<double>("h_Ae", (size_t)n * nb * nb, true);
  dout.Fe = s.alloc<double>("h_Fe", (size_t)n * nb, true);
  dout.bases = s.alloc<double>("h_bases", (size_t)bas_total, true);
  dout.iters = s.alloc<int32_t>("h_iters", (size_t)n * nb, true);
  dout.status = s.alloc<int32_t>("h_status", (size_t)n * nb, true);
  dout.resid = s.alloc<double>("h_resid", (size_t)n * nb, out->resid != nullptr);
  dout.energy_bound = s.alloc<double>("h_ebound", (size_t)n * nb, out->energy_bound != nullptr);
  if (s.rc)
    return s.rc;
  rc = build_device(ctx, &din, opts, &dout, ctx->stream);
  if (rc)
    return rc;
  std::vector<int32_t> status((size_t)n * nb);
  s.down(out->Ae, dout.Ae, (size_t)n * nb * nb);
  s.down(out->Fe, dout.Fe, (size_t)n * nb);
  s.down(out->bases, dout.bases, (size_t)bas_total);
  s.down(out->iters, dout.iters, (size_t)n * nb);
  s.down(status.data(), dout.status, (size_t)n * nb);
  s.down(out->resid, dout.resid, (size_t)n * nb);
  s.down(out->energy_bound, dout.energy_bound, (size_t)n * nb);
  if (s.rc)
    return s.rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (out->status)
    memcpy(out->status, status.data(), status.size() * sizeof(int32_t));
  int bad = 0;
  for (int32_t v : status)
    bad += v != 0;
  if (bad)
    return fail(ctx, MSFEM_ERR_NOCONV, "%d basis solve(s) did not converge", bad);
  return MSFEM_OK;
}

int msfem_energy_error_device(msfem_ctx *ctx, const msfem_batch *in, const double *bases, const double *bases_ref,
                              double *err, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  msfem_solver_opts o;
  msfem_default_opts(&o);
  int rc = validate(ctx, in, &o);
  if (rc)
    return rc;
  if (in->n_cells == 0)
    return MSFEM_OK;
  if (!bases || !bases_ref || !err)
    return fail(ctx, MSFEM_ERR_ARG, "bases, bases_ref and err are required");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = (hipStream_t)stream;
  bool capturing = false;
  rc = order_after_scratch(ctx, st, &capturing);
  if (rc == MSFEM_OK)
    rc = make_plan(ctx, in);
  if (rc)
    return rc;
  Plan &p = ctx->plan;
  const size_t nae = (size_t)in->n_cells * p.nb * p.nb;
  double *diff, *Ad, *Ar, *Fs;
  rc = scratch(ctx, "ee_diff", p.bas_total * sizeof(double), (void **)&diff);
  if (rc == MSFEM_OK)
    rc = scratch(ctx, "ee_Ae", 2 * nae * sizeof(double), (void **)&Ad);
  if (rc == MSFEM_OK)
    rc = scratch(ctx, "ee_Fe", (size_t)in->n_cells * p.nb * sizeof(double), (void **)&Fs);
  if (rc)
    return rc;
  Ar = Ad + nae;
  HIP_TRY(ctx, launch_sub((size_t)p.bas_total, bases, bases_ref, diff, st));
  for (size_t k = 0; k < p.class_nf.size(); ++k)
    for (int pass = 0; pass < 2; ++pass)
      {
        msfem_batch_out eo = {};
        eo.Ae = pass ? Ar : Ad;
        eo.Fe = Fs;
        rc = launch_class_element(ctx, in, &eo, k, pass ? (double *)bases_ref : diff, nullptr, st);
        if (rc)
          return rc;
      }
  HIP_TRY(ctx, launch_energy_ratio(in->n_cells, p.nb, Ad, Ar, err, st));
  return retire_scratch(ctx, st, capturing);
}

int msfem_coarse_assemble_device(msfem_ctx *ctx, int dim, int nc, const double *Ae, const double *Fe, int bc_kind,
                                 double *Kc, double *Fc, uint8_t *cmask, double *cval, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || nc < 1 || (bc_kind != 0 && bc_kind != 1) || !Ae || !Fe || !Kc || !Fc || !cmask ||
      !cval)
    return fail(ctx, MSFEM_ERR_ARG, "invalid coarse assembly arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  HIP_TRY(ctx, launch_coarse_assemble(dim, nc, Ae, Fe, bc_kind, Kc, Fc, cmask, cval, (hipStream_t)stream));
  return MSFEM_OK;
}

int msfem_coarse_assemble(msfem_ctx *ctx, int dim, int nc, const double *Ae, const double *Fe, int bc_kind,
                          double *Kc, double *Fc, uint8_t *cmask, double *cval)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || nc < 1)
    return fail(ctx, MSFEM_ERR_ARG, "invalid coarse assembly arguments");
  const int nb = 1 << dim;
  const size_t ncell = ipow(nc, dim), nn = ipow(nc + 1, dim), ns3 = ipow(3, dim);
  HostStage s{ctx};
  const double *dA = s.up("c_Ae", Ae, ncell * nb * nb);
  const double *dF = s.up("c_Fe", Fe, ncell * nb);
  double *dK = s.alloc<double>("c_Kc", nn * ns3, true);
  double *dFc = s.alloc<double>("c_Fc", nn, true);
  uint8_t *dm = s.alloc<uint8_t>("c_cm", nn, true);
  double *dv = s.alloc<double>("c_cv", nn, true);
  if (s.rc)
    return s.rc;
  int rc = msfem_coarse_assemble_device(ctx, dim, nc, dA, dF, bc_kind, dK, dFc, dm, dv, ctx->stream);
  if (rc)
    return rc;
  s.down(Kc, dK, nn * ns3);
  s.down(Fc, dFc, nn);
  s.down(cmask, dm, nn);
  s.down(cval, dv, nn);
  if (s.rc)
    return s.rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  return MSFEM_OK;
}

int msfem_coarse_solve_device(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc,
                              const uint8_t *cmask, const double *cval, double rel_tol, int max_iter, double *u,
                              int32_t *iters, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || nc < 1 || !(rel_tol > 0) || max_iter < 1 || !Kc || !Fc || !cmask || !cval || !u ||
      !iters)
    return fail(ctx, MSFEM_ERR_ARG, "invalid coarse solve arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  double *work;
  int rc = scratch(ctx, "coarse_work", coarse_solve_work_doubles(dim, nc) * sizeof(double), (void **)&work);
  if (rc)
    return rc;
  bool capturing = false;
  rc = order_after_scratch(ctx, (hipStream_t)stream, &capturing);
  if (rc)
    return rc;
  HIP_TRY(ctx, launch_coarse_solve(dim, nc, Kc, Fc, cmask, cval, rel_tol, max_iter, u, iters, work,
                                   (hipStream_t)stream));
  return retire_scratch(ctx, (hipStream_t)stream, capturing);
}

int msfem_coarse_solve_device_noalloc(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc,
                                      const uint8_t *cmask, const double *cval, double rel_tol, int max_iter,
                                      double *u, int32_t *iters, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || nc < 1 || !(rel_tol > 0) || max_iter < 1 || !Kc || !Fc || !cmask || !cval || !u ||
      !iters)
    return fail(ctx, MSFEM_ERR_ARG, "invalid coarse solve arguments");
  if ((dim == 2 && nc > 4096) || (dim == 3 && nc > 256))
    return fail(ctx, MSFEM_ERR_ARG, "coarse system of %d^%d cells exceeds the fixed scratch of %zu bytes", nc, dim,
                noalloc_scratch_bytes());
  const size_t need = coarse_solve_work_doubles(dim, nc) * sizeof(double);
  if (need > noalloc_scratch_bytes())
    return fail(ctx, MSFEM_ERR_ARG,
                "coarse system of %d^%d cells needs %zu bytes of scratch, more than the fixed %zu of the "
                "allocation-free entry point (msfem_coarse_solve_device has no such limit)",
                nc, dim, need, noalloc_scratch_bytes());
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  double *w = noalloc_scratch(ctx->device);
  if (!w)
    {
      HIP_TRY(ctx, noalloc_register_device(ctx->device));
      w = noalloc_scratch(ctx->device);
    }
  if (!w)
    return fail(ctx, MSFEM_ERR_HIP, "fixed scratch of device %d unavailable", ctx->device);
  HIP_TRY(ctx, launch_coarse_solve(dim, nc, Kc, Fc, cmask, cval, rel_tol, max_iter, u, iters, w, (hipStream_t)stream));
  return MSFEM_OK;
}

int msfem_coarse_solve(msfem_ctx *ctx, int dim, int nc, const double *Kc, const double *Fc, const uint8_t *cmask,
                       const double *cval, double rel_tol, int max_iter, double *u, int32_t *iters)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || nc < 1 || !iters)
    return fail(ctx, MSFEM_ERR_ARG, "invalid coarse solve arguments");
  const size_t nn = ipow(nc + 1, dim), ns3 = ipow(3, dim);
  HostStage s{ctx};
  const double *dK = s.up("s_Kc", Kc, nn * ns3);
  const double *dF = s.up("s_Fc", Fc, nn);
  const uint8_t *dm = s.up("s_cm", cmask, nn);
  const double *dv = s.up("s_cv", cval, nn);
  double *du = s.alloc<double>("s_u", nn, true);
  int32_t *dit = s.alloc<int32_t>("s_it", 1, true);
  if (s.rc)
    return s.rc;
  int rc = msfem_coarse_solve_device(ctx, dim, nc, dK, dF, dm, dv, rel_tol, max_iter, du, dit, ctx->stream);
  if (rc)
    return rc;
  s.down(u, du, nn);
  s.down(iters, dit, 1);
  if (s.rc)
    return s.rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
  if (*iters >= max_iter)
    return fail(ctx, MSFEM_ERR_NOCONV, "coarse CG reached %d iterations", max_iter);
  return MSFEM_OK;
}

int msfem_gather_weights(int dim, int nc, const double *u, double *w)
{
  if ((dim != 2 && dim != 3) || nc < 1 || !u || !w)
    return MSFEM_ERR_ARG;
  const int nb = 1 << dim, np = nc + 1;
  const long long ncell = ipow(nc, dim);
  for (long long c = 0; c < ncell; ++c)
    {
      const int ci[3] = {(int)(c % nc), (int)((c / nc) % nc), (int)(c / ((long long)nc * nc))};
      for (int v = 0; v < nb; ++v)
        {
          const long long X = ci[0] + (v & 1), Y = ci[1] + ((v >> 1) & 1), Z = dim == 3 ? ci[2] + ((v >> 2) & 1) : 0;
          w[c * nb + v] = u[X + np * (Y + (long long)np * Z)];
        }
    }
  return MSFEM_OK;
}

int msfem_reconstruct_device(msfem_ctx *ctx, int dim, int n_cells, const int32_t *n_fine, const double *bases,
                             const double *w, double *u, void *stream)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if ((dim != 2 && dim != 3) || n_cells < 0 || (n_cells > 0 && (!n_fine || !bases || !w || !u)))
    return fail(ctx, MSFEM_ERR_ARG, "invalid reconstruction arguments");
  if (n_cells == 0)
    return MSFEM_OK;
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  const int nb = 1 << dim;
  std::vector<long long> boff(n_cells), ooff(n_cells);
  std::vector<int> nn(n_cells);
  long long b = 0, o = 0;
  for (int c = 0; c < n_cells; ++c)
    {
      nn[c] = (int)ipow(n_fine[c] + 1, dim);
      boff[c] = b;
      ooff[c] = o;
      b += (long long)nb * nn[c];
      o += nn[c];
    }
  // an earlier reconstruction on another stream may still read the metadata
  int rc = wait_scratch_idle(ctx);
  if (rc)
    return rc;
  HostStage s{ctx};
  const long long *db = s.up("r_boff", boff.data(), n_cells);
  const long long *dq = s.up("r_ooff", ooff.data(), n_cells);
  const int *dn = s.up("r_nn", nn.data(), n_cells);
  if (s.rc)
    return s.rc;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream)); // metadata in place before the caller's stream reads it
  HIP_TRY(ctx, launch_reconstruct(n_cells, nb, db, dq, dn, bases, w, u, (hipStream_t)stream));
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  HIP_TRY(ctx, hipStreamIsCapturing((hipStream_t)stream, &cap));
  return retire_scratch(ctx, (hipStream_t)stream, cap != hipStreamCaptureStatusNone);
}

int msfem_fine_solve(msfem_ctx *ctx, int dim, int n, int coef_kind, const double *coef, int rhs_kind,
                     double rhs_value, int bc_kind, double rel_tol, int max_iter, double *u, int32_t *iters)
{
  if (!ctx)
    return fail(nullptr, MSFEM_ERR_ARG, "ctx is NULL");
  if (dim != 2)
    return fail(ctx, MSFEM_ERR_ARG, "the fine-scale solve is built for dim = 2");
  if (n < 1 || n > 8192 || !coef || !u || coef_stride(2, coef_kind) < 0 ||
      (rhs_kind != MSFEM_RHS_CONST && rhs_kind != MSFEM_RHS_SINSIN) ||
      (bc_kind != MSFEM_BC_ZERO_DIRICHLET && bc_kind != MSFEM_BC_REFERENCE) || !(rel_tol > 0) || max_iter < 0)
    return fail(ctx, MSFEM_ERR_ARG, "invalid fine-solve arguments");
  HIP_TRY(ctx, hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const size_t ncell = (size_t)n * n, np = n + 1, nn = np * np;
  HostStage s{ctx};
  const double *dcoef = s.up("f_coef", coef, ncell * coef_stride(2, coef_kind));
  double *Ae = s.alloc<double>("f_Ae", ncell * 16, true), *Fe = s.alloc<double>("f_Fe", ncell * 4, true);
  double *Kc = s.alloc<double>("f_Kc", nn * 9, true), *Fc = s.alloc<double>("f_Fc", nn, true);
  unsigned char *cm = s.alloc<unsigned char>("f_cm", nn, true);
  double *cv = s.alloc<double>("f_cv", nn, true);
  double *vec = s.alloc<double>("f_vec", 5 * nn, true);
  double *part = s.alloc<double>("f_part", 2 * (size_t)fine_cg_blocks((int)nn), true);
  FineCgState *state = s.alloc<FineCgState>("f_state", 1, true);
  if (s.rc)
    return s.rc;
  HIP_TRY(ctx, launch_fine_elements_2d(n, coef_kind, dcoef, rhs_kind, rhs_value, Ae, Fe, st));
  HIP_TRY(ctx, launch_coarse_assemble(2, n, Ae, Fe, bc_kind, Kc, Fc, cm, cv, st));
  FineCgState init = {};
  init.rel_tol = rel_tol;
  init.max_iter = max_iter;
  HIP_TRY(ctx, hipMemcpyAsync(state, &init, sizeof init, hipMemcpyHostToDevice, st));
  FineCgArgs a;
  a.dim = 2;
  a.np = (int)np;
  a.nn = (int)nn;
  a.Kc = Kc;
  a.Fc = Fc;
  a.cmask = cm;
  a.cval = cv;
  a.x = vec;
  a.r = vec + nn;
  a.p = vec + 2 * nn;
  a.q = vec + 3 * nn;
  a.dinv = vec + 4 * nn;
  a.part = part;
  a.state = state;
  HIP_TRY(ctx, launch_fine_cg_start(a, st));
  // CG steps in blocks of kSteps, captured once as a graph and replayed until
  // the device-side state says done (finished steps turn into no-ops)
  constexpr int kSteps = 32;
  hipGraph_t g = nullptr;
  hipGraphExec_t x = nullptr;
  HIP_TRY(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const hipError_t el = launch_fine_cg_steps(a, kSteps, st);
  const hipError_t ec = hipStreamEndCapture(st, &g);
  if (el != hipSuccess || ec != hipSuccess)
    {
      if (g)
        hipGraphDestroy(g);
      return fail(ctx, MSFEM_ERR_HIP, "fine CG capture: %s", hipGetErrorString(el != hipSuccess ? el : ec));
    }
  hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess)
    return fail(ctx, MSFEM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  FineCgState host = {};
  for (int done_steps = 0;; done_steps += kSteps)
    {
      e = hipMemcpyAsync(&host, state, sizeof host, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess)
        e = hipStreamSynchronize(st);
      if (e != hipSuccess || host.done || done_steps > max_iter)
        break;
      e = hipGraphLaunch(x, st);
      if (e != hipSuccess)
        break;
    }
  hipGraphExecDestroy(x);
  if (e != hipSuccess)
    return fail(ctx, MSFEM_ERR_HIP, "fine CG: %s", hipGetErrorString(e));
  HIP_TRY(ctx, hipMemcpyAsync(u, a.x, nn * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(ctx, hipStreamSynchronize(st));
  if (iters)
    *iters = host.iters;
  if (!host.converged)
    return fail(ctx, MSFEM_ERR_NOCONV, "fine CG: %d iterations, residual %g > %g", host.iters, host.rnorm, host.tol);
  return MSFEM_OK;
}

} // extern "C"
