// 2D local basis solves: for every (coarse cell, basis i) solve
//     K_II x = -K_IB g_i,   g_i = Q1 nodal function N_i on the cell boundary,
// i.e. the constrained problems of DiffusionProblemBasis<2>::run()
// (include/base/diffusion_problem_basis.tpp:450-465, solver :295-317), to the
// same converged values, with a matrix-free Q1 operator on the GPU.
//
// One lane group owns one problem.  Lanes map to node columns (CPL columns per
// lane), SW stacked waves split the node rows (R rows per lane); a workgroup of
// SW waves holds PPW = 64 / lane-group-width problems side by side.  The CG
// vectors of the owned nodes live in registers, the preconditioned residual z
// of the whole cell in a zero-padded LDS grid (the only values a lane needs
// from its neighbours), the fine-cell coefficients in registers
// (kModeScalar / kModeTensor) or global memory (kModeGlobal).
//
// Inner solver: Jacobi-preconditioned CG in the single-reduction form of
// Chronopoulos and Gear (J. Comput. Appl. Math. 25, 1989): per iteration one
// stencil apply w = K z and ONE fused reduction of (r.z, w.z, r.r), so two
// workgroup barriers per iteration (z exchange, reduction).  s = K p is
// carried by recurrence.
//
// Two-level variant (CS, 64^2 tiles, 8 waves x 8 rows): the preconditioner is
// M^-1 = D^-1 + P A_c^-1 P^T with P the bilinear hats of the tile's 8 x 8
// coarse grid (49 interior coarse nodes) and A_c = P^T K P, inverted once per
// cell by coarse_setup_2d and kept in LDS.  Restriction: 8-lane DPP sums into
// fixed LDS slots (each wave covers coarse rows w and w+1, each 8-lane group
// coarse columns g and g+1); coarse product: 8 lanes per coarse row;
// prolongation: bilinear interpolation of the 4 coarse values around the
// lane.  Two more barriers per iteration; about 2.4x fewer iterations on
// lognormal tiles.  The converged values are those of any preconditioner.
//
// Precision: an FP64 outer loop computes the true residual r = -(K phi)_I from
// phi in global memory (FP64 coefficients) and checks
//     ||r|| <= max(abs_tol, rel_tol ||b||);
// the inner CG in type T solves K d = r (T = double: to the final tolerance,
// normally one cycle; T = float: to inner_rel_tol, refined in FP64 until
// converged).
#include <utility>

#include "msfem_device.hpp"
#include "msfem_kernels.hpp"

namespace msfem {

namespace {

// 6 conductances of a fine cell: c01, c02, c03, c12, c13, c23.
template <typename U>
struct Cond6 {
  U c[6];
};

// (K p)_C from the four cells around node C.
//   NE cell: node is vertex 0, NW: vertex 1, SE: vertex 2, SW: vertex 3.
template <typename U>
__device__ __forceinline__ U gather6(U pC, U pE, U pW, U pN, U pS, U pNE, U pNW, U pSE, U pSW, const Cond6<U> &ne,
                                     const Cond6<U> &nw, const Cond6<U> &se, const Cond6<U> &sw)
{
  const U d = (ne.c[0] + ne.c[1] + ne.c[2]) + (nw.c[0] + nw.c[4] + nw.c[3]) + (se.c[5] + se.c[1] + se.c[3]) +
              (sw.c[5] + sw.c[4] + sw.c[2]);
  U y = d * pC;
  y -= (ne.c[0] + se.c[5]) * pE;
  y -= (nw.c[0] + sw.c[5]) * pW;
  y -= (ne.c[1] + nw.c[4]) * pN;
  y -= (se.c[1] + sw.c[4]) * pS;
  y -= ne.c[2] * pNE;
  y -= nw.c[3] * pNW;
  y -= se.c[3] * pSE;
  y -= sw.c[2] * pSW;
  return y;
}

template <typename U>
__device__ __forceinline__ Cond6<U> from4(const Cond4 &c)
{
  return {{(U)c.ch, (U)c.cv, (U)c.cd, (U)c.ca, (U)c.cv, (U)c.ch}};
}

// FP64 conductances of fine cell (cx, cy) straight from global memory.
template <int MODE, int NF>
__device__ __forceinline__ Cond6<double> cell_cond64(const double *cf, const double *cg, int cx, int cy, double rho)
{
  const int f = cy * NF + cx;
  if constexpr (MODE == kModeScalar)
    return from4<double>(cond_scalar(cf[f], rho));
  else if constexpr (MODE == kModeTensor)
    return from4<double>(cond_tensor(cf[3 * f], cf[3 * f + 1], cf[3 * f + 2], rho));
  else
    {
      Cond6<double> r;
#pragma unroll
      for (int k = 0; k < 6; ++k)
        r.c[k] = cg[6 * f + k];
      return r;
    }
}

// Diagonal of the fine stiffness at node (X, Y) from the node-coupling planes:
// the sum of its 8 couplings (aE, aN, aNE, aNW of the node and of its W, S,
// SW, SE neighbours).
template <int NF>
__device__ __forceinline__ double node_diag64(const double *cg, int X, int Y)
{
  constexpr int NP = NF + 1, PSN = coupling_plane_stride(NF);
  const double *q = cg + Y * NP + X;
  return (q[0] + q[-1]) + (q[PSN] + q[PSN - NP]) + (q[2 * PSN] + q[2 * PSN - NP - 1]) +
         (q[3 * PSN] + q[3 * PSN - NP + 1]);
}

} // namespace

template <typename T, int NF, int CPL, int SW, int R, int MODE, int WPE, bool CS>
__global__ void __launch_bounds__(kWave *SW, WPE) local_solve_2d(LocalArgs a)
{
  constexpr int NP = NF + 1;
  constexpr int LW = NF / CPL;
  static_assert(LW == 16 || LW == 32 || LW == 64, "lane group must be 16, 32 or 64 lanes");
  constexpr int PPW = kWave / LW;
  static_assert(!CS || ((NF == 64 || NF == 128) && PPW == 1 && MODE != kModeGlobal),
                "the two-level form: 64^2 and 128^2 tiles, one problem per workgroup");
  static_assert(SW * R >= NF - 1, "rows not covered");
  // LDS grid of one problem: pitch GP, GR rows; the padding column NF+1 and
  // rows beyond NF stay zero so that every window read is in bounds.
  constexpr int GP = NF + 2;
  constexpr int GR = (SW * R + 2 > NP) ? SW * R + 2 : NP;
  constexpr int GSZ = GP * GR;
  constexpr int NCACHE = MODE == kModeScalar ? 1 : (MODE == kModeTensor ? 4 : 1);
  // kModeStream keeps q = K p of the own nodes in LDS (thread-private slots):
  // FP32, and FP64 in the 64^2 two-level form (the FP64 128^2 grid leaves no
  // LDS for it: q goes to a global work array there)
  constexpr bool QLDS = MODE == kModeStream && (sizeof(T) == 4 || (NF == 64 && CS));
  // two-level streamed form: once r is updated, q is dead and the same slots
  // keep z = M^-1 r from the (r, z) sums to the direction update instead of
  // forming the preconditioner's terms twice per iteration
  constexpr bool ZLDS = QLDS && CS;
  constexpr int NT = kWave * SW;
  constexpr int PSN = coupling_plane_stride(NF); // node-coupling plane stride (kModeStream)
  // kModeStream, 64^2 (8 nodes per lane, 256 registers): the lane's window
  // of couplings of the first NHOIST planes (FP32: all four; FP64: E and N)
  // stays in registers across the whole solve, the others are streamed
  constexpr bool HOISTC = MODE == kModeStream && NF == 64 && WPE <= 2;
  constexpr int NHOIST = HOISTC ? (sizeof(T) == 4 ? 4 : 2) : 0;

  __shared__ T sgrid[PPW * GSZ];
  __shared__ T sred[problem_sums_slots<SW, PPW, 3>()];       // inner-CG partial sums (type T)
  __shared__ double sred64[problem_sums_slots<SW, PPW, 1>()]; // FP64 residual norms
  __shared__ T sq[QLDS ? R * CPL * NT : 1];
  __shared__ T sred2[QLDS ? problem_sums_slots<SW, PPW, 1>() : 1]; // second reduction of the two-reduction PCG
  // two-level: inverse coarse matrix, restriction partials [wave][2][group][2], coarse correction
  // inverse coarse matrix, rows padded to kSciStride entries: the coarse product's
  // wave reads 8 rows x 8 consecutive entries, and a stride of 56 (= 24 mod 32
  // words for FP32, 16 mod 32 for FP64) puts those rows on distinct LDS banks
  // (the FP64 128^2 instance has no LDS left for the padding)
  constexpr int kSciStride = (sizeof(T) == 8 && NF == 128) ? kCoarseDofs : 56;
  static_assert(kSciStride >= kCoarseDofs, "padded coarse rows");
  __shared__ T sci[CS ? kCoarseDofs * kSciStride : 1];
  __shared__ T spart[CS ? SW * 2 * 8 * 2 : 1];  // [wave][coarse row 0/1][group][column 0/1]
  __shared__ T suc[CS ? kCoarseDofs : 1];
  // per-wave coarse correction (the FP64 128^2 instance has no LDS left for it)
  constexpr bool WAVE_CS = CS && !(sizeof(T) == 8 && NF == 128);
  __shared__ T swrc[WAVE_CS ? SW * kWave : 1];         // [wave][coarse dof]: P^T r, wave-private
  // FP32 two-level box tiles keep the inner iterate x in thread-private LDS
  // slots ([chunk of 4 nodes][thread][4]: 16-byte accesses, consecutive lanes
  // on consecutive banks) instead of registers: x is only ever updated
  // (x += alpha p) and read once at the end, and the freed registers let two
  // workgroups share a CU
  constexpr bool XLDS = MODE != kModeStream && CS && sizeof(T) == 4 && (R * CPL) % 4 == 0;
  __shared__ float4 sx[XLDS ? R * CPL / 4 * kWave * SW : 1];
  // with four waves per SIMD the recurrence s = K p goes there too
  constexpr bool SLDS = XLDS && WPE >= 4;
  __shared__ float4 ss[SLDS ? R * CPL / 4 * kWave * SW : 1];
  // Intermediate hat levels of the 64^2 two-level tiles (a.mlinv set): M^-1
  // also gains P_2 D_2^-1 P_2^T + P_4 D_4^-1 P_4^T, P_s the bilinear hats of
  // spacing s, D_s = diag(P_s^T K P_s).  Each wave restricts its 8 rows with DPP
  // sums along the wave (x) and fixed row weights (y); the coarse rows it shares
  // with the waves below and above are completed from their published partial
  // rows after the restriction barrier; prolongation by DPP from the coarse
  // values masked to their lanes.
  constexpr bool ML = CS && NF == 64 && CPL == 1 && SW == 8 && R == 8;
  __shared__ T smlp[ML ? 4 * SW * kWave : 1]; // partial rows [B2, T2, B4, T4][wave][lane]
  // The FP32 128^2 streamed form (two node columns per lane) takes the levels
  // of spacing 4 and 8: each wave restricts its rows to the three spacing-4
  // coarse rows it touches (B4 and T4 partial, F4 whole) and publishes them
  // (the odd lanes hold the spacing-4 columns: 32 values per row); the
  // spacing-8 values follow from the spacing-4 ones (P_8 = P_4 P_48), which
  // takes the neighbours' F4 rows too.
  constexpr bool ML8 = CS && NF == 128 && CPL == 2 && SW == 16 && R == 8 && MODE == kModeStream && sizeof(T) == 4;
  __shared__ T smlq[ML8 ? 3 * SW * 32 : 1];           // [B4, F4, T4][wave][lane / 2]
  __shared__ float smld[(ML || ML8) ? kMlFloats : 1]; // D_s1^-1 [31][31], then D_s2^-1 [15][15]

  const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
  const int slot = lane / LW, pl = lane % LW;
  const int prob = xcd_block(blockIdx.x, gridDim.x) * PPW + slot;
  const bool pvalid = prob < a.n_problems;
  const bool mlon = (ML || ML8) && a.mlinv != nullptr; // (uniform)
  int cell = 0, basis = 0;
  if (pvalid)
    {
      cell = a.cells[prob / a.nb_solve];
      basis = prob % a.nb_solve;
    }
  T *grid = sgrid + slot * GSZ;
  const int X0 = 1 + pl * CPL, Y0 = 1 + wave * R;
  T *gown = grid + Y0 * GP + X0;           // own node (row 0, column 0)
  T *gwin = grid + (Y0 - 1) * GP + X0 - 1; // window origin

  double rho = 1.0;
  const double *cf = a.coef;
  const double *cg = a.cond;
  const float *c32 = a.cond32;
  double *phi = a.bases;
  if (pvalid)
    {
      if constexpr (MODE == kModeScalar || MODE == kModeTensor) // boxes only
        {
          const double *bx = a.box + (size_t)cell * 4;
          rho = (bx[3] - bx[2]) / (bx[1] - bx[0]);
        }
      cf = a.coef + a.coef_off[cell];
      if (MODE == kModeGlobal || MODE == kModeStream)
        cg = a.cond + a.cond_off[cell];
      if (MODE == kModeStream) // FP32 copies of the planes follow the FP64 ones
        c32 = (const float *)(cg + 4 * PSN);
      phi = a.bases + a.bas_off[cell] + (size_t)basis * NP * NP;
    }

  // ---- fine-cell data cached in registers ----------------------------------
  // CDPP (FP32 scalar two-level tiles at four waves per SIMD, one node column
  // per lane): only the cells right of the lane's column are cached; the cells
  // left of it are the left neighbour lane's, read with a one-lane rotation of
  // the wave (DPP wave_ror:1).  Lane 63 (column 64, a boundary column whose
  // results are masked) caches cell column 0 for lane 0.
  constexpr bool CDPP = MODE == kModeScalar && CPL == 1 && LW == kWave && sizeof(T) == 4 && WPE >= 4 && CS;
  T cc[R + 1][CPL + 1][NCACHE];
#pragma unroll
  for (int i = 0; i <= R; ++i)
#pragma unroll
    for (int j = 0; j <= CPL; ++j)
      {
        const int cx = (CDPP && pl == LW - 1) ? 0 : X0 - 1 + j, cy = Y0 - 1 + i;
        const bool ok = pvalid && cx < NF && cy < NF && !(CDPP && j == 0);
        if constexpr (MODE == kModeScalar)
          cc[i][j][0] = ok ? (T)cf[cy * NF + cx] : (T)0;
        else if constexpr (MODE == kModeTensor)
          {
            Cond4 c4 = {0, 0, 0, 0};
            if (ok)
              {
                const int f = cy * NF + cx;
                c4 = cond_tensor(cf[3 * f], cf[3 * f + 1], cf[3 * f + 2], rho);
              }
            cc[i][j][0] = (T)c4.ch;
            cc[i][j][1] = (T)c4.cv;
            cc[i][j][2] = (T)c4.cd;
            cc[i][j][3] = (T)c4.ca;
          }
        else
          cc[i][j][0] = 0;
      }
  // scalar coefficient of cached cell (i, j)
  auto ccs = [&](int i, int j) -> T {
    if constexpr (CDPP)
      if (j == 0)
        return dpp_f32<0x13C>(cc[i][1][0]); // wave_ror:1: lane l reads lane l - 1 (lane 0: lane 63)
    return cc[i][j][0];
  };
  const Cond4 k4 = cond_scalar(1.0, rho); // scalar-mode unit conductances
  const T kh = (T)k4.ch, kv = (T)k4.cv, kd = (T)k4.cd, kc = (T)(k4.ch + k4.cv + k4.cd);

  auto cond_T = [&](int i, int j) -> Cond6<T> {
    if constexpr (MODE == kModeScalar)
      {
        const T s = ccs(i, j);
        return {{s * kh, s * kv, s * kd, s * kd, s * kv, s * kh}};
      }
    else if constexpr (MODE == kModeTensor)
      return {{cc[i][j][0], cc[i][j][1], cc[i][j][2], cc[i][j][3], cc[i][j][1], cc[i][j][0]}};
    else
      {
        const int cx = X0 - 1 + j, cy = Y0 - 1 + i;
        Cond6<T> r = {{0, 0, 0, 0, 0, 0}};
        if (pvalid && cx < NF && cy < NF)
          {
            const double *q = cg + 6 * (cy * NF + cx);
#pragma unroll
            for (int k = 0; k < 6; ++k)
              r.c[k] = (T)q[k];
          }
        return r;
      }
  };

  // Jacobi diagonal (FP32 suffices for a preconditioner); 0 on slots that are
  // not interior nodes of a valid problem, which keeps every vector of those
  // slots at exactly 0.
  float dinv[R][CPL];
#pragma unroll
  for (int i = 0; i < R; ++i)
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      {
        const bool act = pvalid && (X0 + c <= NF - 1) && (Y0 + i <= NF - 1);
        T d;
        if constexpr (MODE == kModeStream)
          d = (T)node_diag64<NF>(cg, X0 + c, Y0 + i);
        else
          {
            const Cond6<T> ne = cond_T(i + 1, c + 1), nw = cond_T(i + 1, c), se = cond_T(i, c + 1),
                           sw = cond_T(i, c);
            d = (ne.c[0] + ne.c[1] + ne.c[2]) + (nw.c[0] + nw.c[4] + nw.c[3]) + (se.c[5] + se.c[1] + se.c[3]) +
                (sw.c[5] + sw.c[4] + sw.c[2]);
          }
        dinv[i][c] = act ? (float)((T)1 / d) : 0.0f;
      }
  // kModeStream keeps the diagonal as packed BF16 pairs (8 registers instead
  // of 16); any positive diagonal is a valid Jacobi preconditioner, and a
  // nonzero value stays nonzero, so the active-node mask is unchanged.
  constexpr bool DPK = MODE == kModeStream || (XLDS && WPE >= 4);
  uint32_t dpk[DPK ? (R * CPL + 1) / 2 : 1];
  if constexpr (DPK)
    {
#pragma unroll
      for (int k = 0; k < (R * CPL + 1) / 2; ++k)
        {
          uint32_t lo = 0, hi = 0;
          const int k0 = 2 * k, k1 = 2 * k + 1;
          {
            const uint32_t u = __float_as_uint(dinv[k0 / CPL][k0 % CPL]);
            lo = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
          }
          if (k1 < R * CPL)
            {
              const uint32_t u = __float_as_uint(dinv[k1 / CPL][k1 % CPL]);
              hi = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
            }
          dpk[k] = lo | hi;
        }
    }
  auto dget = [&](int i, int c) -> float {
    if constexpr (DPK)
      {
        const int k = i * CPL + c;
        const uint32_t w = dpk[k >> 1];
        return __uint_as_float((k & 1) ? (w & 0xffff0000u) : (w << 16));
      }
    else
      return dinv[i][c];
  };

  // ---- initial state: LDS grid zero, phi = g on the boundary, x0 inside ----
  for (int k = tid; k < PPW * GSZ; k += blockDim.x)
    sgrid[k] = 0;
  if constexpr (CS)
    if (pvalid)
      {
        const double *ci = a.cinv + (size_t)(prob / a.nb_solve) * (kCoarseDofs * kCoarseDofs);
        for (int k = tid; k < kCoarseDofs * kCoarseDofs; k += blockDim.x)
          sci[(k / kCoarseDofs) * kSciStride + k % kCoarseDofs] = (T)ci[k];
        if (mlon)
          {
            const float *mi = a.mlinv + (size_t)(prob / a.nb_solve) * kMlFloats;
            for (int k = tid; k < kMlFloats; k += blockDim.x)
              smld[k] = mi[k];
          }
      }
  if (pvalid)
    for (int k = pl + LW * wave; k < NP * NP; k += LW * SW)
      {
        const int X = k % NP, Y = k / NP;
        const bool bnd = X == 0 || X == NF || Y == 0 || Y == NF;
        if (bnd && a.verts) // general quadrilateral: boundary data already in place
          continue;
        double v = 0.0;
        if (bnd)
          v = nodal_q1(basis, 2, X, Y, 0, NF);
        else if (a.x0)
          {
            // a non-finite initial value (e.g. a failed prediction) starts from 0
            v = a.x0[a.bas_off[cell] + (size_t)basis * NP * NP + k];
            if (!isfinite(v))
              v = 0.0;
          }
        else if (a.q1_start && !a.verts)
          v = nodal_q1(basis, 2, X, Y, 0, NF); // the Q1 interpolant of the boundary data
        else if (a.q1_start)
          {
            // general cell: transfinite (Coons) interpolant of the boundary values
            // over the parametric node grid (boundary written by quad_dirichlet_2d)
            const double sx = (double)X / NF, ty = (double)Y / NF;
            const double *p = phi;
            v = (1.0 - sx) * p[Y * NP] + sx * p[Y * NP + NF] + (1.0 - ty) * p[X] + ty * p[NF * NP + X] -
                ((1.0 - sx) * (1.0 - ty) * p[0] + sx * (1.0 - ty) * p[NF] + (1.0 - sx) * ty * p[NF * NP] +
                 sx * ty * p[NF * NP + NF]);
          }
        phi[k] = v;
      }
  __syncthreads();

  // (K phi)_C in FP64 from global memory; bonly: interior values taken as 0.
  auto apply64 = [&](int X, int Y, bool bonly) -> double {
    auto v = [&](int XX, int YY) -> double {
      const bool bnd = XX == 0 || XX == NF || YY == 0 || YY == NF;
      return (bonly && !bnd) ? 0.0 : phi[YY * NP + XX];
    };
    if constexpr (MODE == kModeStream)
      {
        const double pC = v(X, Y);
        const double *q = cg + Y * NP + X;
        return q[0] * (pC - v(X + 1, Y)) + q[-1] * (pC - v(X - 1, Y)) + q[PSN] * (pC - v(X, Y + 1)) +
               q[PSN - NP] * (pC - v(X, Y - 1)) + q[2 * PSN] * (pC - v(X + 1, Y + 1)) +
               q[2 * PSN - NP - 1] * (pC - v(X - 1, Y - 1)) + q[3 * PSN] * (pC - v(X - 1, Y + 1)) +
               q[3 * PSN - NP + 1] * (pC - v(X + 1, Y - 1));
      }
    else
      {
        const Cond6<double> ne = cell_cond64<MODE, NF>(cf, cg, X, Y, rho);
        const Cond6<double> nw = cell_cond64<MODE, NF>(cf, cg, X - 1, Y, rho);
        const Cond6<double> se = cell_cond64<MODE, NF>(cf, cg, X, Y - 1, rho);
        const Cond6<double> sw = cell_cond64<MODE, NF>(cf, cg, X - 1, Y - 1, rho);
        return gather6<double>(v(X, Y), v(X + 1, Y), v(X - 1, Y), v(X, Y + 1), v(X, Y - 1), v(X + 1, Y + 1),
                               v(X - 1, Y + 1), v(X + 1, Y - 1), v(X - 1, Y - 1), ne, nw, se, sw);
      }
  };

  // with a given start the tolerance is still relative to ||b|| = ||K_IB g||
  // (the residual of the zero start), computed here from the boundary values
  const bool given_start = a.x0 || a.q1_start;
  double tol_eff = 0.0, bnorm = 0.0;
  if (given_start)
    {
      double s[1] = {0};
#pragma unroll 1
      for (int k = 0; k < R * CPL; ++k)
        {
          const int X = X0 + k % CPL, Y = Y0 + k / CPL;
          if (pvalid && X <= NF - 1 && Y <= NF - 1)
            {
              const double b = apply64(X, Y, true);
              s[0] += b * b;
            }
        }
      problem_sums<double, LW, SW, PPW, 1>(s, false, sred64, wave, slot, pl);
      bnorm = sqrt(s[0]);
      tol_eff = a.ptol ? a.ptol[(size_t)cell * a.nb + basis] : fmax(a.abs_tol, a.rel_tol * bnorm);
      __syncthreads();
    }

  // ---- two-level pieces -------------------------------------------------------------
  // Coarse spacing CH = NF / 8 fine cells.  An 8-lane group g holds the fine
  // columns CH g + 1 .. CH g + CH (8 lanes x CPL columns), i.e. it feeds coarse
  // columns g and g + 1; the rows of wave w lie in coarse band w / WB (WB = CH / R
  // waves per band) and feed coarse rows w / WB and w / WB + 1.
  constexpr int CH = NF / 8, WB = CH / R > 0 ? CH / R : 1;
  static_assert(!CS || (CH * 8 == NF && 8 * CPL == CH && WB * R == CH && SW == 8 * WB),
                "two-level layout: 8-lane groups of CH columns, whole waves per coarse band");
  const int cg8 = pl >> 3, cband = wave / WB;
  const int yy0 = 1 + (wave % WB) * R; // row i: yy = yy0 + i in 1..CH above the band's lower coarse row
  // Interpolation weights: coarse columns g, g + 1 for the lane's column c,
  // coarse rows of the band for row i.  128^2 tiles (16 nodes per lane, 128
  // registers) recompute them where used from opaque copies of the lane and
  // row offsets (two ALU operations each): hoisted out of the iteration loop
  // they would be kept live and spilled.  64^2 tiles have the registers and
  // let the compiler keep them.
  // With four waves per SIMD (128 registers) the lane-dependent LDS and global
  // addresses used once per iteration or cycle are recomputed where used from
  // opaque copies of the lane index: hoisted out of the loops they would be
  // kept live and spilled.
  constexpr bool OPAQUE_CA = WPE >= 4;
  constexpr bool OPAQUE_W = NF == 128;
  static_assert(!OPAQUE_W || LW == kWave, "128^2 tiles: one problem per wave");
  // 128^2 tiles (128 registers, all in use): the lane index is read from the
  // hardware (v_mbcnt) where used instead of kept in a register
  auto lane_here = [&]() -> int {
    if constexpr (OPAQUE_W)
      {
        int v;
        asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(v));
        return v;
      }
    else
      return pl;
  };
  // (and the wave-dependent LDS slots are formed from an opaque copy of the
  // wave index where used)
  auto wave_here = [&]() -> int {
    int w = wave;
    if constexpr (OPAQUE_W)
      asm volatile("" : "+v"(w));
    return w;
  };
  auto lane_off = [&]() {
    int v = (lane_here() & 7) * CPL + 1;
    if constexpr (OPAQUE_W)
      asm volatile("" : "+v"(v));
    return v;
  };
  auto row_off = [&]() {
    int v = yy0;
    if constexpr (OPAQUE_W)
      {
        v = __builtin_amdgcn_readfirstlane(v);
        asm volatile("" : "+s"(v));
      }
    return v;
  };
  auto wxa = [&](int c) -> T { return (T)(CH - lane_off() - c) * (T)(1.0 / CH); };
  auto wxb = [&](int c) -> T { return (T)(lane_off() + c) * (T)(1.0 / CH); };
  auto wya = [&](int i) -> T { return (T)(CH - row_off() - i) * (T)(1.0 / CH); };
  auto wyb = [&](int i) -> T { return (T)(row_off() + i) * (T)(1.0 / CH); };
  auto spart_at = [](int w, int rs, int g, int ab) { return ((w * 2 + rs) * 8 + g) * 2 + ab; };
  // restriction partials of r (8-lane DPP sums; lane 0 of each group writes)
  auto restrict_r = [&](T (&rv)[R][CPL]) {
    if constexpr (CS)
      {
        T v4[4] = {0, 0, 0, 0};
#pragma unroll
        for (int c = 0; c < CPL; ++c)
          {
            T s0 = 0, s1 = 0;
#pragma unroll
            for (int i = 0; i < R; ++i)
              {
                s0 += rv[i][c] * wya(i);
                s1 += rv[i][c] * wyb(i);
              }
            v4[0] += s0 * wxa(c);
            v4[1] += s0 * wxb(c);
            v4[2] += s1 * wxa(c);
            v4[3] += s1 * wxb(c);
          }
#pragma unroll
        for (int k = 0; k < 4; ++k)
          v4[k] = oct_sum_any<T>(v4[k]);
        if ((pl & 7) == 0)
          {
            int g = lane_here() >> 3; // (OPAQUE_CA: the slot address is formed here, not hoisted and spilled)
            if constexpr (OPAQUE_CA)
              asm volatile("" : "+v"(g));
            const int w = wave_here();
            spart[spart_at(w, 0, g, 0)] = v4[0];
            spart[spart_at(w, 0, g, 1)] = v4[1];
            spart[spart_at(w, 1, g, 0)] = v4[2];
            spart[spart_at(w, 1, g, 1)] = v4[3];
          }
      }
  };
  // u_c = A_c^-1 (P^T r): coarse row i on lanes 8 i .. 8 i + 7, fixed order
  auto coarse_solve = [&]() {
    if constexpr (CS)
      if (tid < kCoarseDofs * 8)
        {
          const int i = tid >> 3, c = tid & 7;
          T acc = 0;
#pragma unroll
          for (int k = 0; k < 7; ++k)
            {
              const int j = c + 8 * k;
              if (j < kCoarseDofs)
                {
                  const int jx = j % kCoarseSide + 1, jy = j / kCoarseSide + 1;
                  T rc = 0;
#pragma unroll
                  for (int b = 0; b < WB; ++b)
                    {
                      const int wl = jy * WB + b, wu = (jy - 1) * WB + b; // band jy (lower row), band jy-1 (upper)
                      rc += (spart[spart_at(wl, 0, jx, 0)] + spart[spart_at(wl, 0, jx - 1, 1)]) +
                            (spart[spart_at(wu, 1, jx, 0)] + spart[spart_at(wu, 1, jx - 1, 1)]);
                    }
                  acc += sci[i * kSciStride + j] * rc;
                }
            }
          acc = oct_sum_any<T>(acc);
          if (c == 0)
            suc[i] = acc;
        }
  };
  // coarse value at coarse node (jx, jy); 0 on the tile boundary
  auto ucoarse = [&](int jx, int jy) -> T {
    return (jx >= 1 && jx <= kCoarseSide && jy >= 1 && jy <= kCoarseSide) ? suc[(jy - 1) * kCoarseSide + jx - 1]
                                                                          : (T)0;
  };
  // Per-wave coarse correction: wave w interpolates from coarse rows cband and
  // cband + 1 only, so it forms those <= 14 values of u_c = A_c^-1 (P^T r)
  // itself and needs no barrier between the coarse product and the
  // interpolation.  Lanes j < 49 form (P^T r)_j into the wave's private slots;
  // lane 4 q + h sums row q (coarse row cband + q / 7, column q % 7 + 1) over
  // the entries 13 h .. 13 h + 12, a 4-lane butterfly completes it (every lane
  // of the quad holds the row's value).  Every wave evaluates a shared coarse
  // row in the same order, so the waves agree bitwise.
  auto coarse_wave_acc = [&]() -> T {
    T acc = 0;
    if constexpr (WAVE_CS)
      {
        int lp = lane_here();
        if constexpr (OPAQUE_CA)
          asm volatile("" : "+v"(lp));
        T *wrc = swrc + wave_here() * kWave;
        if (lp < kCoarseDofs)
          {
            const int jx = lp % kCoarseSide + 1, jy = lp / kCoarseSide + 1;
            T rc = 0;
#pragma unroll
            for (int b = 0; b < WB; ++b)
              {
                const int wl = jy * WB + b, wu = (jy - 1) * WB + b; // band jy (lower row), band jy-1 (upper)
                rc += (spart[spart_at(wl, 0, jx, 0)] + spart[spart_at(wl, 0, jx - 1, 1)]) +
                      (spart[spart_at(wu, 1, jx, 0)] + spart[spart_at(wu, 1, jx - 1, 1)]);
              }
            wrc[lp] = rc;
          }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int q = lp >> 2, h = lp & 3;
        const int jyq = cband + (q >= kCoarseSide ? 1 : 0), jxq = q % kCoarseSide + 1;
        if (q < 2 * kCoarseSide && jyq >= 1 && jyq <= kCoarseSide)
          {
            const T *srow = sci + ((jyq - 1) * kCoarseSide + jxq - 1) * kSciStride;
#pragma unroll
            for (int k = 0; k < 13; ++k)
              {
                const int j = 13 * h + k;
                if (j < kCoarseDofs)
                  acc += srow[j] * wrc[j];
              }
          }
        acc += dpp_any<T, 0xB1>(acc);
        acc += dpp_any<T, 0x4E>(acc);
      }
    return acc;
  };
  // the lane's interpolation columns from the wave's coarse values (acc as above)
  auto coarse_wave_interp = [&](T acc, T (&ci0)[CPL], T (&ci1)[CPL]) {
    if constexpr (WAVE_CS)
      {
        // u at coarse column jx (0..8) of coarse row cband + r: lane 4 (7 r + jx - 1)
        auto uq = [&](int jx, int r) -> T {
          const int src = 4 * (r * kCoarseSide + (jx >= 1 && jx <= kCoarseSide ? jx - 1 : 0));
          const T v = __shfl(acc, src, kWave);
          return (jx >= 1 && jx <= kCoarseSide) ? v : (T)0;
        };
        const int g8 = lane_here() >> 3;
        const T u00 = uq(g8, 0), u10 = uq(g8 + 1, 0), u01 = uq(g8, 1), u11 = uq(g8 + 1, 1);
#pragma unroll
        for (int c = 0; c < CPL; ++c)
          {
            ci0[c] = u00 * wxa(c) + u10 * wxb(c);
            ci1[c] = u01 * wxa(c) + u11 * wxb(c);
          }
      }
  };
  auto coarse_wave = [&](T (&ci0)[CPL], T (&ci1)[CPL]) { coarse_wave_interp(coarse_wave_acc(), ci0, ci1); };
  // the lane's interpolation rows: z_c(i, c) = ci0[c] wya(i) + ci1[c] wyb(i)
  auto interp_cols = [&](T (&ci0)[CPL], T (&ci1)[CPL]) {
    const T u00 = ucoarse(cg8, cband), u10 = ucoarse(cg8 + 1, cband);
    const T u01 = ucoarse(cg8, cband + 1), u11 = ucoarse(cg8 + 1, cband + 1);
#pragma unroll
    for (int c = 0; c < CPL; ++c)
      {
        ci0[c] = u00 * wxa(c) + u10 * wxb(c);
        ci1[c] = u01 * wxa(c) + u11 * wxb(c);
      }
  };

  // ---- intermediate levels (ML) ------------------------------------------------------
  auto ror1 = [](T v) -> T { return dpp_any<T, 0x13C>(v); }; // wave_ror:1: lane l reads lane l - 1 (lane 0: 63)
  auto rol1 = [](T v) -> T { return dpp_any<T, 0x134>(v); }; // wave_rol:1: lane l reads lane l + 1 (lane 63: 0)
  auto xw2 = [&](T v) -> T { return v + (T)0.5 * (ror1(v) + rol1(v)); }; // hat-2 weights along the wave
  auto xw4 = [&](T v) -> T {                                              // hat-4 weights along the wave
    const T u = xw2(v);
    return u + (T)0.5 * (ror1(ror1(u)) + rol1(rol1(u)));
  };
  // the partial rows this wave shares with its neighbours: B2 = row 4w (from
  // row y = 8w + 1), T2 = row 4w + 4 (rows 8w + 7, 8w + 8), B4 = row 2w (rows
  // 8w + 1 .. 8w + 3), T4 = row 2w + 2 (rows 8w + 5 .. 8w + 8); lane l = column l + 1
  // (the lane-dependent addresses and masks below are formed from an opaque
  // copy of the lane index where used: hoisted out of the loops they would be
  // kept live and spilled)
  auto ml_lane = [&]() {
    int v = lane;
    if constexpr (OPAQUE_W)
      v = lane_here();
    asm volatile("" : "+v"(v));
    return v;
  };
  // (the wave index as a scalar: the row conditions below are then uniform branches)
  const int mwave = __builtin_amdgcn_readfirstlane(wave);
  auto restrict_ml = [&](T (&rv)[R][CPL]) {
    if constexpr (ML)
      if (mlon)
        {
          T *pp = smlp + mwave * kWave + ml_lane();
          pp[0] = xw2((T)0.5 * rv[0][0]);
          pp[SW * kWave] = xw2(rv[7][0] + (T)0.5 * rv[6][0]);
          pp[2 * SW * kWave] = xw4(((T)0.75 * rv[0][0] + (T)0.5 * rv[1][0]) + (T)0.25 * rv[2][0]);
          pp[3 * SW * kWave] = xw4((((T)0.25 * rv[4][0] + (T)0.5 * rv[5][0]) + (T)0.75 * rv[6][0]) + rv[7][0]);
        }
  };
  // the intermediate levels' share of z on the lane's rows (after the
  // restriction barrier): coarse values y_s = D_s^-1 (P_s^T r) at lanes
  // x = s K (zero elsewhere, and on the tile boundary), spread to the wave's
  // columns with the hat weights, then to its rows
  auto ml_terms = [&](T (&rv)[R][CPL], T (&ml)[R]) {
#pragma unroll
    for (int i = 0; i < R; ++i)
      ml[i] = 0;
    if constexpr (ML)
      if (mlon)
        {
          const int xl = ml_lane(), x = xl + 1;
          const T *pp = smlp + xl;
          const int wlo = mwave > 0 ? mwave - 1 : 0, whi = mwave < SW - 1 ? mwave + 1 : SW - 1;
          const T r0 = rv[0][0], r1 = rv[1][0], r2 = rv[2][0], r3 = rv[3][0], r4 = rv[4][0], r5 = rv[5][0],
                  r6 = rv[6][0];
          const bool lane2 = (x & 1) == 0 && x <= 62, lane4 = (x & 3) == 0 && x <= 60;
          const int i2 = (4 * mwave - 1) * 31 + (x >> 1) - 1, i4 = kMl2 + (2 * mwave - 1) * 15 + (x >> 2) - 1;
          // level 4, coarse rows 2w .. 2w + 2: y4 at lanes x = 4 K, spread to
          // the level-2 lanes (x = 2 K; zero at odd x)
          T s4[3];
          {
            T q[3];
            q[0] = mwave > 0 ? pp[(3 * SW + wlo) * kWave] + pp[(2 * SW + mwave) * kWave] : (T)0;
            q[1] = xw4((((((T)0.25 * r0 + (T)0.5 * r1) + (T)0.75 * r2) + r3) + (T)0.75 * r4) + (T)0.5 * r5 +
                       (T)0.25 * r6);
            q[2] = mwave < SW - 1 ? pp[(3 * SW + mwave) * kWave] + pp[(2 * SW + whi) * kWave] : (T)0;
#pragma unroll
            for (int m = 0; m < 3; ++m)
              {
                const int L = 2 * mwave + m;
                const bool ok = lane4 && L >= 1 && L <= 15;
                const T y = ok ? q[m] * (T)smld[ok ? i4 + 15 * m : 0] : (T)0;
                s4[m] = y + (T)0.5 * (ror1(ror1(y)) + rol1(rol1(y)));
              }
          }
          // level 2, coarse rows 4w .. 4w + 4, plus the level-4 values
          // interpolated to them (P_4 = P_2 P_24: one prolongation from level 2)
          T c2[5];
          {
            T q[5];
            q[0] = mwave > 0 ? pp[(SW + wlo) * kWave] + pp[mwave * kWave] : (T)0;
            q[1] = xw2(r1 + (T)0.5 * (r0 + r2));
            q[2] = xw2(r3 + (T)0.5 * (r2 + r4));
            q[3] = xw2(r5 + (T)0.5 * (r4 + r6));
            q[4] = mwave < SW - 1 ? pp[(SW + mwave) * kWave] + pp[whi * kWave] : (T)0;
            const T s4h[5] = {s4[0], (T)0.5 * (s4[0] + s4[1]), s4[1], (T)0.5 * (s4[1] + s4[2]), s4[2]};
#pragma unroll
            for (int m = 0; m < 5; ++m)
              {
                const int L = 4 * mwave + m;
                const bool ok = lane2 && L >= 1 && L <= 31;
                const T u = (ok ? q[m] * (T)smld[ok ? i2 + 31 * m : 0] : (T)0) + s4h[m];
                c2[m] = u + (T)0.5 * (ror1(u) + rol1(u));
              }
          }
          ml[0] = (T)0.5 * (c2[0] + c2[1]);
          ml[1] = c2[1];
          ml[2] = (T)0.5 * (c2[1] + c2[2]);
          ml[3] = c2[2];
          ml[4] = (T)0.5 * (c2[2] + c2[3]);
          ml[5] = c2[3];
          ml[6] = (T)0.5 * (c2[3] + c2[4]);
          ml[7] = c2[4];
        }
  };

  // ---- intermediate levels of the 128^2 streamed form (ML8) ---------------------------
  // Lane l holds columns x = 2 l + 1 (c = 0) and 2 l + 2 (c = 1): the spacing-2
  // sums v2 = r(2 l + 2) + (r(2 l + 1) + r(2 l + 3)) / 2 sit on every lane, and
  // the hat-2 weights over them give the spacing-4 restriction on the odd
  // lanes (x = 4 K, lane 2 K - 1).
  auto xres4 = [&](T c0, T c1) -> T {
    const T v2 = c1 + (T)0.5 * (c0 + rol1(c0));
    return xw2(v2);
  };
  auto restrict_ml8 = [&](T (&rv)[R][CPL]) {
    if constexpr (ML8)
      if (mlon)
        {
          // rows y = 8 w + 1 + i: B4 = coarse row 2 w (i = 0..2), F4 = row 2 w + 1
          // (i = 0..6), T4 = row 2 w + 2 (i = 4..7)
          T b[2], f[2], t[2];
#pragma unroll
          for (int c = 0; c < 2; ++c)
            {
              b[c] = ((T)0.75 * rv[0][c] + (T)0.5 * rv[1][c]) + (T)0.25 * rv[2][c];
              f[c] = (((((T)0.25 * rv[0][c] + (T)0.5 * rv[1][c]) + (T)0.75 * rv[2][c]) + rv[3][c]) +
                      (T)0.75 * rv[4][c]) +
                     (T)0.5 * rv[5][c] + (T)0.25 * rv[6][c];
              t[c] = (((T)0.25 * rv[4][c] + (T)0.5 * rv[5][c]) + (T)0.75 * rv[6][c]) + rv[7][c];
            }
          const T vb = xres4(b[0], b[1]), vf = xres4(f[0], f[1]), vt = xres4(t[0], t[1]);
          const int xl = ml_lane();
          if (xl & 1)
            {
              T *pp = smlq + mwave * 32 + (xl >> 1);
              pp[0] = vb;
              pp[SW * 32] = vf;
              pp[2 * SW * 32] = vt;
            }
        }
  };
  // After the restriction barrier: the spacing-4 values y_4 = D_4^-1 (P_4^T r)
  // of coarse rows 2 w .. 2 w + 2 on the odd lanes, plus the spacing-8 values
  // y_8 = D_8^-1 (P_8^T r) of coarse rows w, w + 1 interpolated to them
  // (P_8 = P_4 P_48; P_8^T r = P_48^T (P_4^T r)).  The spacing-8 rows shared by
  // two waves are formed by both from the same values in the same order.
  auto ml8_coarse = [&](T (&u4)[3]) {
#pragma unroll
    for (int m = 0; m < 3; ++m)
      u4[m] = 0;
    if constexpr (ML8)
      if (mlon)
        {
          const int xl = ml_lane();
          const bool oddl = (xl & 1) != 0;
          const T *pp = smlq + (xl >> 1);
          const int wlo = mwave > 0 ? mwave - 1 : 0, whi = mwave < SW - 1 ? mwave + 1 : SW - 1;
          auto rd = [&](int k, int w) -> T {
            const T v = pp[(k * SW + w) * 32];
            return oddl ? v : (T)0;
          };
          const T F = rd(1, mwave);
          T q0 = 0, q2 = 0, Fm = 0, Fp = 0;
          if (mwave > 0)
            {
              q0 = rd(2, wlo) + rd(0, mwave);
              Fm = rd(1, wlo);
            }
          if (mwave < SW - 1)
            {
              q2 = rd(2, mwave) + rd(0, whi);
              Fp = rd(1, whi);
            }
          // spacing 8: rows w, w + 1 on the lanes 4 K - 1, spread back to the odd lanes
          const T r8[2] = {((T)0.5 * Fm + q0) + (T)0.5 * F, ((T)0.5 * F + q2) + (T)0.5 * Fp};
          T s8[2];
#pragma unroll
          for (int m = 0; m < 2; ++m)
            {
              const int L = mwave + m;
              const bool ok = (xl & 3) == 3 && xl <= 59 && L >= 1 && L <= 15;
              const T x8 = r8[m] + (T)0.5 * (ror1(ror1(r8[m])) + rol1(rol1(r8[m])));
              const T y = ok ? x8 * (T)smld[ok ? kMl2 + (L - 1) * 15 + ((xl + 1) >> 2) - 1 : 0] : (T)0;
              s8[m] = y + (T)0.5 * (ror1(ror1(y)) + rol1(rol1(y)));
            }
          const T q[3] = {q0, F, q2};
          const T s8h[3] = {s8[0], (T)0.5 * (s8[0] + s8[1]), s8[1]};
#pragma unroll
          for (int m = 0; m < 3; ++m)
            {
              const int L = 2 * mwave + m;
              const bool ok = oddl && xl <= 61 && L >= 1 && L <= 31;
              u4[m] = (ok ? q[m] * (T)smld[ok ? (L - 1) * 31 + ((xl + 1) >> 1) - 1 : 0] : (T)0) + s8h[m];
            }
        }
  };
  // the levels' share of z on the lane's two columns of coarse row m (spread
  // from the odd lanes to every lane's column 1, then to the column 0 midpoints)
  auto ml8_cols = [&](const T (&u4)[3], T (&v)[3][2]) {
    if constexpr (ML8)
#pragma unroll
      for (int m = 0; m < 3; ++m)
        {
          const T c1 = u4[m] + (T)0.5 * (ror1(u4[m]) + rol1(u4[m]));
          v[m][0] = (T)0.5 * (ror1(c1) + c1);
          v[m][1] = c1;
        }
  };
  // row i of the wave between coarse rows 2 w (y = 8 w), 2 w + 1 (8 w + 4), 2 w + 2 (8 w + 8)
  auto ml8_at = [&](const T (&v)[3][2], int i, int c) -> T {
    if constexpr (ML8)
      {
        if (i < 3)
          return (T)(0.25 * (3 - i)) * v[0][c] + (T)(0.25 * (i + 1)) * v[1][c];
        if (i == 3)
          return v[1][c];
        if (i < 7)
          return (T)(0.25 * (7 - i)) * v[1][c] + (T)(0.25 * (i - 3)) * v[2][c];
        return v[2][c];
      }
    return (T)0;
  };

  T xr[R][CPL], rr[R][CPL], pr[R][CPL], sr[R][CPL];
  int it_total = 0;
  int state = pvalid ? 0 : 1; // 0 running, 1 done
  int conv = 0;
  double rnorm = 0.0;

  for (int cyc = 0;; ++cyc)
    {
      // ---- true FP64 residual of the current phi ----------------------------
      // (rolled loop; the values pass through this thread's own LDS slots so
      // the FP64 gathers do not hold registers across rows)
      double s[1] = {0};
#pragma unroll 1
      for (int k = 0; k < R * CPL; ++k)
        {
          const int X = X0 + k % CPL, Y = Y0 + k / CPL;
          if (pvalid && X <= NF - 1 && Y <= NF - 1)
            {
              const double r64 = -apply64(X, Y, false);
              grid[Y * GP + X] = (T)r64;
              s[0] += r64 * r64;
            }
        }
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int c = 0; c < CPL; ++c)
          rr[i][c] = dget(i, c) != 0.0f ? gown[i * GP + c] : (T)0;
      problem_sums<double, LW, SW, PPW, 1>(s, false, sred64, wave, slot, pl);
      rnorm = sqrt(s[0]);
      if (cyc == 0 && !given_start)
        {
          bnorm = rnorm;
          tol_eff = a.ptol ? a.ptol[(size_t)cell * a.nb + basis] : fmax(a.abs_tol, a.rel_tol * bnorm);
        }
      if (state == 0)
        {
          if (rnorm <= tol_eff)
            {
              state = 1;
              conv = 1;
            }
          else if (it_total >= a.max_iter || cyc >= a.max_refine)
            state = 1;
        }
      if (!__syncthreads_or(state == 0))
        break;

      // ---- inner CG (single reduction) on K d = r --------------------------------
      const bool run = state == 0;
      double inner_tol = 0.5 * tol_eff;
      if (sizeof(T) < sizeof(double))
        inner_tol = fmax(a.inner_rel_tol * rnorm, inner_tol);
      const T inner_tol2 = (T)(inner_tol * inner_tol);
#pragma unroll
      for (int i = 0; i < R; ++i)
#pragma unroll
        for (int c = 0; c < CPL; ++c)
          {
            xr[i][c] = 0;
            pr[i][c] = 0;
            sr[i][c] = 0;
          }
      if constexpr (XLDS)
#pragma unroll
        for (int k = 0; k < R * CPL / 4; ++k)
          {
            sx[k * NT + tid] = make_float4(0.f, 0.f, 0.f, 0.f);
            if constexpr (SLDS)
              ss[k * NT + tid] = make_float4(0.f, 0.f, 0.f, 0.f);
          }
      bool irun = run, first = true;
      T inv_gamma_prev = 1, inv_alpha_prev = 1; // 1 / gamma, 1 / alpha of the previous iteration
      if constexpr (MODE == kModeStream)
        {
          // Standard Jacobi PCG with two reductions per iteration.  p lives in
          // the LDS grid (the stencil's input, read back from the own slots for
          // the updates); the node couplings of the node row being processed and
          // of the row below are streamed from global memory as the stencil
          // slides upward: (K p)_C = sum over the 8 neighbours k of a_Ck (p_C - p_k).
          // FP32: x, r and the packed diagonal in registers, q in thread-private
          // LDS slots.  FP64: the FP64 grid leaves no LDS for q, so q goes to this
          // problem's slots of a global work array (thread-private, coalesced)
          // and x is added straight into the FP64 solution phi.
          static_assert(PPW == 1, "kModeStream runs one problem per workgroup");
          // two-level: coarse interpolation rows of the current z (zero without it)
          T ci0[CPL], ci1[CPL];
#pragma unroll
          for (int c = 0; c < CPL; ++c)
            ci0[c] = ci1[c] = 0;
          T cacc = 0; // per-wave coarse sums of the current z (kept for the p update)
          T u4[3];    // ML8: the intermediate levels' coarse values of the current z (idem)
          T mv[3][2];
          T mlv[R]; // ML (64^2): the intermediate levels' share of z on the lane's rows
          if constexpr (CS) // irun is uniform (one problem per workgroup): true here
            {
              restrict_r(rr);
              restrict_ml(rr);
              restrict_ml8(rr);
              __syncthreads(); // restriction partials complete
              if constexpr (WAVE_CS)
                {
                  cacc = coarse_wave_acc();
                  coarse_wave_interp(cacc, ci0, ci1);
                }
              else
                {
                  coarse_solve();
                  __syncthreads(); // coarse correction complete
                  interp_cols(ci0, ci1);
                }
              ml8_coarse(u4);
              ml8_cols(u4, mv);
              ml_terms(rr, mlv);
            }
          T v2[2] = {0, 0}; // (r, z), (r, r)
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < CPL; ++c)
              {
                T z = rr[i][c] * (T)dget(i, c);
                if constexpr (CS)
                  z += ci0[c] * wya(i) + ci1[c] * wyb(i);
                if constexpr (ML8)
                  z += ml8_at(mv, i, c);
                if constexpr (ML)
                  z += mlv[i];
                gown[i * GP + c] = irun ? z : (T)0;
                v2[0] += rr[i][c] * z;
                v2[1] += rr[i][c] * rr[i][c];
              }
          bool any = problem_sums<T, LW, SW, PPW, 2>(v2, irun, sred, wave, slot, pl);
          if (irun && (v2[1] <= inner_tol2 || it_total >= a.max_iter))
            irun = false;
          T rz = v2[0];
          // HOISTC: the couplings of the lane's rows Y0 - 1 .. Y0 + R - 1 and
          // columns X0 - 1 .. X0 + CPL, loaded once (every index lies inside the
          // four padded planes; the unused entries are dead)
          T creg[HOISTC ? R + 1 : 1][HOISTC ? NHOIST : 1][CPL + 2];
          if constexpr (HOISTC)
            {
#pragma unroll
              for (int r = 0; r <= R; ++r)
#pragma unroll
                for (int pn = 0; pn < NHOIST; ++pn)
#pragma unroll
                  for (int k = 0; k < CPL + 2; ++k)
                    {
                      const int q = pn * PSN + (Y0 - 1 + r) * NP + X0 - 1 + k;
                      creg[r][pn][k] = !pvalid ? (T)0 : sizeof(T) == 8 ? (T)cg[q] : (T)c32[q];
                    }
            }
          typedef T __attribute__((address_space(1))) * gtptr;
          typedef const T __attribute__((address_space(1))) * gcptr;
          while (any)
            {
              __syncthreads(); // p complete in the grid
              // opaque per-iteration copies of the (wave-uniform) addresses and of
              // the packed diagonal: keeps loads and conversions inside the loop
              // (hoisted, they would need far more than the 128 registers)
              const T *cbase;
              if constexpr (sizeof(T) == 8)
                cbase = (const T *)cg;
              else
                cbase = (const T *)c32;
              static_assert(R * CPL * NT <= stream_q_slots(NF), "q slots of the FP64 streamed form");
              T *qb = sizeof(T) == 8 ? (T *)(a.work + (size_t)prob * stream_q_slots(NF)) : nullptr;
              double *pb = phi;
              asm volatile("" : "+s"(cbase), "+s"(qb), "+s"(pb));
#pragma unroll
              for (int k = 0; k < (R * CPL + 1) / 2; ++k)
                asm volatile("" : "+v"(dpk[k]));
              const gcptr cp = (gcptr)cbase;
              const gtptr qg = (gtptr)qb;
              typedef double __attribute__((address_space(1))) * gdptr;
              const gdptr ph = (gdptr)pb;
              // 128^2: the planes are read as one buffer (raw buffer loads): the
              // lane's window origin is the only vector offset, formed once per
              // iteration from the hardware lane index and an opaque wave index;
              // plane, row and column go into the constant offsets
              constexpr bool CBUF = OPAQUE_W;
              __amdgpu_buffer_rsrc_t crs;
              int cvo = 0;
              if constexpr (CBUF)
                {
                  crs = __builtin_amdgcn_make_buffer_rsrc((void *)cbase, (short)0, 4 * PSN * (int)sizeof(T),
                                                          0x00020000);
                  cvo = (wave_here() * (R * NP) + lane_here() * CPL) * (int)sizeof(T);
                }
              auto bload = [&](int off) -> T {
                if constexpr (sizeof(T) == 4)
                  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(crs, cvo, off, 0));
                else
                  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(crs, cvo, off, 0));
              };
              // couplings at window column k (node X0 - 1 + k) of node row Y
              // (every index lies inside the four padded planes)
              auto cpl = [&](int plane, int Y, int k) -> T {
                if (plane < NHOIST) // (plane and Y - Y0 are compile-time constants here)
                  return creg[Y - Y0 + 1][HOISTC ? plane : 0][k];
                if constexpr (CBUF)
                  return bload((plane * PSN + (Y - Y0 + 1) * NP + k) * (int)sizeof(T));
                return cp[(unsigned)(plane * PSN + Y * NP + X0 - 1 + k)];
              };
              // CBUF: dst[k] = coupling at window column k for k0 <= k <= k1, else
              // 0 (FP32, k1 - k0 = 1 or 2: one two- or three-dword load)
              auto crow = [&](int plane, int Y, int k0, int k1, T (&dst)[CPL + 2]) {
                if constexpr (CBUF && sizeof(T) == 8)
                  {
#pragma unroll
                    for (int k = 0; k < CPL + 2; ++k)
                      dst[k] = (k >= k0 && k <= k1) ? bload((plane * PSN + (Y - Y0 + 1) * NP + k) * 8) : (T)0;
                  }
                else if constexpr (CBUF)
                  {
                    static_assert(CPL == 2, "three-dword rows");
#pragma unroll
                    for (int k = 0; k < CPL + 2; ++k)
                      dst[k] = 0;
                    const int off = (plane * PSN + (Y - Y0 + 1) * NP + k0) * (int)sizeof(float);
                    if (k1 - k0 == 2)
                      {
                        const auto v = __builtin_amdgcn_raw_buffer_load_b96(crs, cvo, off, 0);
                        dst[k0] = __uint_as_float(v[0]);
                        dst[k0 + 1] = __uint_as_float(v[1]);
                        dst[k0 + 2] = __uint_as_float(v[2]);
                      }
                    else
                      {
                        const auto v = __builtin_amdgcn_raw_buffer_load_b64(crs, cvo, off, 0);
                        dst[k0] = __uint_as_float(v[0]);
                        dst[k0 + 1] = __uint_as_float(v[1]);
                      }
                  }
              };
              // couplings of the node row (E, N, NE, NW) and of the row below
              // (pN, pNE, pNW) in a sliding register window (in FP64 this spills
              // a few values, still faster than reloading each node's 8 couplings)
              T w0[CPL + 2], w1[CPL + 2], w2[CPL + 2];
              T pN[CPL + 2], pNE[CPL + 2], pNW[CPL + 2];
              if constexpr (CBUF)
                {
#pragma unroll
                  for (int k = 0; k < CPL + 2; ++k)
                    {
                      w0[k] = gwin[k];
                      w1[k] = gwin[GP + k];
                    }
                  crow(1, Y0 - 1, 1, CPL, pN);
                  crow(2, Y0 - 1, 0, CPL - 1, pNE);
                  crow(3, Y0 - 1, 2, CPL + 1, pNW);
                }
              else
#pragma unroll
                for (int k = 0; k < CPL + 2; ++k)
                  {
                    w0[k] = gwin[k];
                    w1[k] = gwin[GP + k];
                    pN[k] = (k >= 1 && k <= CPL) ? cpl(1, Y0 - 1, k) : (T)0;
                    pNE[k] = (k <= CPL - 1) ? cpl(2, Y0 - 1, k) : (T)0;
                    pNW[k] = (k >= 2) ? cpl(3, Y0 - 1, k) : (T)0;
                  }
              T pq[1] = {0};
#pragma unroll
              for (int i = 0; i < R; ++i)
                {
                  const int Y = Y0 + i;
                  T E[CPL + 2], N[CPL + 2], NE[CPL + 2], NW[CPL + 2];
                  if constexpr (CBUF)
                    {
                      crow(0, Y, 0, CPL, E);
                      crow(1, Y, 1, CPL, N);
                      crow(2, Y, 0, CPL, NE);
                      crow(3, Y, 1, CPL + 1, NW);
#pragma unroll
                      for (int k = 0; k < CPL + 2; ++k)
                        w2[k] = gwin[(i + 2) * GP + k];
                    }
                  else
#pragma unroll
                    for (int k = 0; k < CPL + 2; ++k)
                      {
                        E[k] = k <= CPL ? cpl(0, Y, k) : (T)0;
                        N[k] = (k >= 1 && k <= CPL) ? cpl(1, Y, k) : (T)0;
                        NE[k] = k <= CPL ? cpl(2, Y, k) : (T)0;
                        NW[k] = k >= 1 ? cpl(3, Y, k) : (T)0;
                        w2[k] = gwin[(i + 2) * GP + k];
                      }
#pragma unroll
                  for (int c = 0; c < CPL; ++c)
                    {
                      const T pC = w1[c + 1];
                      T y = E[c + 1] * (pC - w1[c + 2]);
                      y += E[c] * (pC - w1[c]);
                      y += N[c + 1] * (pC - w2[c + 1]);
                      y += pN[c + 1] * (pC - w0[c + 1]);
                      y += NE[c + 1] * (pC - w2[c + 2]);
                      y += pNE[c] * (pC - w0[c]);
                      y += NW[c + 1] * (pC - w2[c]);
                      y += pNW[c + 2] * (pC - w0[c + 2]);
                      const T q = dget(i, c) != 0.0f ? y : (T)0;
                      if constexpr (QLDS)
                        sq[(i * CPL + c) * NT + tid] = q;
                      else
                        qg[(unsigned)((i * CPL + c) * NT + tid)] = q;
                      pq[0] += pC * q;
                    }
#pragma unroll
                  for (int k = 0; k < CPL + 2; ++k)
                    {
                      w0[k] = w1[k];
                      w1[k] = w2[k];
                      pN[k] = N[k];
                      pNE[k] = NE[k];
                      pNW[k] = NW[k];
                    }
                  __builtin_amdgcn_sched_barrier(0);
                }
              problem_sums<T, LW, SW, PPW, 1>(pq, irun, sred2, wave_here(), slot, lane_here());
              v2[0] = 0;
              v2[1] = 0;
              if (irun)
                {
                  const T alpha = quotient<T>(rz, pq[0]);
#pragma unroll
                  for (int i = 0; i < R; ++i)
#pragma unroll
                    for (int c = 0; c < CPL; ++c)
                      {
                        T q;
                        if constexpr (QLDS)
                          q = sq[(i * CPL + c) * NT + tid];
                        else
                          q = qg[(unsigned)((i * CPL + c) * NT + tid)];
                        if constexpr (sizeof(T) == 4)
                          xr[i][c] += alpha * gown[i * GP + c];
                        else if (dget(i, c) != 0.0f)
                          ph[(unsigned)((Y0 + i) * NP + X0 + c)] += alpha * gown[i * GP + c];
                        rr[i][c] -= alpha * q;
                        if constexpr (!CS)
                          {
                            const T z = rr[i][c] * (T)dget(i, c);
                            v2[0] += rr[i][c] * z;
                            v2[1] += rr[i][c] * rr[i][c];
                          }
                      }
                  if constexpr (CS) // z = D^-1 r + P A_c^-1 P^T r of the new r (uniform branch)
                    {
                      restrict_r(rr);
                      restrict_ml(rr);
                      restrict_ml8(rr);
                      __syncthreads(); // restriction partials complete
                      if constexpr (WAVE_CS)
                        {
                          cacc = coarse_wave_acc();
                          coarse_wave_interp(cacc, ci0, ci1);
                        }
                      else
                        {
                          coarse_solve();
                          __syncthreads(); // coarse correction complete
                          interp_cols(ci0, ci1);
                        }
                      ml8_coarse(u4);
                      ml8_cols(u4, mv);
                      ml_terms(rr, mlv);
#pragma unroll
                      for (int i = 0; i < R; ++i)
#pragma unroll
                        for (int c = 0; c < CPL; ++c)
                          {
                            T z = rr[i][c] * (T)dget(i, c) + (ci0[c] * wya(i) + ci1[c] * wyb(i));
                            if constexpr (ML8)
                              z += ml8_at(mv, i, c);
                            if constexpr (ML)
                              z += mlv[i];
                            if constexpr (ZLDS) // q of this thread's slot is consumed: keep z there
                              sq[(i * CPL + c) * NT + tid] = z;
                            v2[0] += rr[i][c] * z;
                            v2[1] += rr[i][c] * rr[i][c];
                          }
                    }
                  ++it_total;
                }
              any = problem_sums<T, LW, SW, PPW, 2>(v2, irun, sred, wave_here(), slot, lane_here());
              if (irun && (v2[1] <= inner_tol2 || it_total >= a.max_iter))
                irun = false;
              // every lane is past this iteration's stencil (two barriers ago):
              // the own slots of the grid take the next direction in place
              const T beta = irun ? quotient<T>(v2[0], rz) : (T)0;
              rz = v2[0];
              if constexpr (ZLDS) // z of the new r from this thread's slots (irun: it was formed above)
                {
#pragma unroll
                  for (int i = 0; i < R; ++i)
#pragma unroll
                    for (int c = 0; c < CPL; ++c)
                      gown[i * GP + c] = irun ? sq[(i * CPL + c) * NT + tid] + beta * gown[i * GP + c] : (T)0;
                  continue;
                }
              if constexpr (WAVE_CS) // from the kept per-wave coarse sums
                coarse_wave_interp(cacc, ci0, ci1);
              else if constexpr (CS) // the coarse values are still in LDS (next write: next iteration)
                interp_cols(ci0, ci1);
              if constexpr (ML8)
                ml8_cols(u4, mv);
#pragma unroll
              for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                  {
                    T z = rr[i][c] * (T)dget(i, c);
                    if constexpr (CS)
                      z += ci0[c] * wya(i) + ci1[c] * wyb(i);
                    if constexpr (ML8)
                      z += ml8_at(mv, i, c);
                    gown[i * GP + c] = irun ? z + beta * gown[i * GP + c] : (T)0;
                  }
            }
        }
      else
      {
      if constexpr (CS)
        {
          restrict_r(rr);
          restrict_ml(rr);
        }
      while (true)
        {
          // one problem per workgroup: irun is uniform, so the loop ends with
          // it (no idle pass) and the writes below need no select
          if constexpr (PPW == 1)
            if (!irun)
              break;
          if constexpr (CS)
            {
              __syncthreads(); // restriction partials of r complete
              // z = D^-1 r + P u_c (the interpolation vanishes on the tile boundary)
              T ci0[CPL], ci1[CPL];
              coarse_wave(ci0, ci1);
              // the row weights sum to one: ci0 wya + ci1 wyb = ci0 + (ci1 - ci0) wyb
#pragma unroll
              for (int c = 0; c < CPL; ++c)
                ci1[c] -= ci0[c];
              T ml[R];
              ml_terms(rr, ml);
#pragma unroll
              for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                  gown[i * GP + c] = rr[i][c] * (T)dget(i, c) + (ci0[c] + ci1[c] * wyb(i)) + ml[i];
            }
          else
            {
              // z = M r into the LDS grid (zero on padding/boundary slots)
#pragma unroll
              for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                  gown[i * GP + c] = (PPW == 1 || irun) ? rr[i][c] * (T)dget(i, c) : (T)0;
            }
          __syncthreads();
          // w = K z, rows slid through a 3-row window; fused partial sums
          T wv[R][CPL];
          T w0[CPL + 2], w1[CPL + 2], w2[CPL + 2];
#pragma unroll
          for (int k = 0; k < CPL + 2; ++k)
            {
              w0[k] = gwin[k];
              w1[k] = gwin[GP + k];
            }
          T sums[3] = {0, 0, 0}; // (r, z), (w, z), (r, r)
#pragma unroll
          for (int i = 0; i < R; ++i)
            {
#pragma unroll
              for (int k = 0; k < CPL + 2; ++k)
                w2[k] = gwin[(i + 2) * GP + k];
#pragma unroll
              for (int c = 0; c < CPL; ++c)
                {
                  T y;
                  if constexpr (MODE == kModeScalar)
                    {
                      const T aNE = ccs(i + 1, c + 1), aNW = ccs(i + 1, c), aSE = ccs(i, c + 1), aSW = ccs(i, c);
                      y = kc * ((aNE + aNW) + (aSE + aSW)) * w1[c + 1];
                      y -= kh * ((aNE + aSE) * w1[c + 2] + (aNW + aSW) * w1[c]);
                      y -= kv * ((aNE + aNW) * w2[c + 1] + (aSE + aSW) * w0[c + 1]);
                      y -= kd * (aNE * w2[c + 2] + aNW * w2[c] + aSE * w0[c + 2] + aSW * w0[c]);
                    }
                  else
                    y = gather6<T>(w1[c + 1], w1[c + 2], w1[c], w2[c + 1], w0[c + 1], w2[c + 2], w2[c], w0[c + 2],
                                   w0[c], cond_T(i + 1, c + 1), cond_T(i + 1, c), cond_T(i, c + 1), cond_T(i, c));
                  const T z = w1[c + 1];
                  wv[i][c] = dget(i, c) != 0.0f ? y : (T)0;
                  sums[0] += rr[i][c] * z;
                  sums[1] += wv[i][c] * z;
                  sums[2] += rr[i][c] * rr[i][c];
                }
#pragma unroll
              for (int k = 0; k < CPL + 2; ++k)
                {
                  w0[k] = w1[k];
                  w1[k] = w2[k];
                }
            }
          // Converged (or capped) problems stop here with x the current iterate.
          // The flag sent with the partials is the state before this test, so
          // the loop makes one more (idle) pass after the last problem stops.
          T v3[3] = {sums[0], sums[1], sums[2]};
          const bool any = problem_sums<T, LW, SW, PPW, 3>(v3, irun, sred, wave, slot, pl);
          if (irun && (v3[2] <= inner_tol2 || it_total >= a.max_iter))
            irun = false;
          if (irun)
            {
              const T gamma = v3[0], delta = v3[1];
              T beta = 0, alpha;
              // Chronopoulos-Gear: beta = gamma / gamma_prev, alpha = gamma /
              // (delta - beta gamma / alpha_prev), with 1 / gamma and 1 / alpha
              // carried to the next iteration (two reciprocals, no divisions)
              const T inv_gamma = reciprocal<T>(gamma);
              T denom = delta;
              if (!first)
                {
                  beta = gamma * inv_gamma_prev;
                  denom = delta - beta * gamma * inv_alpha_prev;
                }
              alpha = gamma * reciprocal<T>(denom);
              const T tb = beta, ta = alpha;
#pragma unroll
              for (int i = 0; i < R; ++i)
#pragma unroll
                for (int c = 0; c < CPL; ++c)
                  {
                    // z = M r of this iteration: still in the own grid slot (FP64:
                    // one LDS read instead of a conversion and a multiply)
                    const T z = (CS || sizeof(T) == 8) ? gown[i * GP + c] : rr[i][c] * (T)dget(i, c);
                    pr[i][c] = z + tb * pr[i][c];
                    if constexpr (!SLDS)
                      {
                        sr[i][c] = wv[i][c] + tb * sr[i][c];
                        rr[i][c] -= ta * sr[i][c];
                      }
                    if constexpr (!XLDS)
                      xr[i][c] += ta * pr[i][c];
                  }
              if constexpr (SLDS)
#pragma unroll
                for (int k = 0; k < R * CPL / 4; ++k)
                  {
                    float4 v = ss[k * NT + tid];
                    float *vs = &v.x;
#pragma unroll
                    for (int m = 0; m < 4; ++m)
                      {
                        const int n = 4 * k + m, i = n / CPL, c = n % CPL;
                        vs[m] = wv[i][c] + tb * vs[m];
                        rr[i][c] -= ta * vs[m];
                      }
                    ss[k * NT + tid] = v;
                  }
              if constexpr (XLDS)
#pragma unroll
                for (int k = 0; k < R * CPL / 4; ++k)
                  {
                    float4 v = sx[k * NT + tid];
                    const int n = 4 * k;
                    v.x += ta * pr[n / CPL][n % CPL];
                    v.y += ta * pr[(n + 1) / CPL][(n + 1) % CPL];
                    v.z += ta * pr[(n + 2) / CPL][(n + 2) % CPL];
                    v.w += ta * pr[(n + 3) / CPL][(n + 3) % CPL];
                    sx[k * NT + tid] = v;
                  }
              inv_gamma_prev = inv_gamma;
              inv_alpha_prev = denom * inv_gamma;
              first = false;
              ++it_total;
              if constexpr (CS)
                {
                  restrict_r(rr);
                  restrict_ml(rr);
                }
            }
          if (!any)
            break;
        }
      }
      // ---- phi += d in FP64 (the FP64 streamed form has already added it) ---------
      if (run && !(MODE == kModeStream && sizeof(T) == 8))
        {
          // (OPAQUE_CA: the lane's addresses are formed here, not hoisted out of
          // the cycle loop and spilled)
          double *phl = phi + X0;
          if constexpr (OPAQUE_CA)
            asm volatile("" : "+v"(phl));
#pragma unroll
          for (int i = 0; i < R; ++i)
#pragma unroll
            for (int c = 0; c < CPL; ++c)
              if (dget(i, c) != 0.0f)
                {
                  T xv = xr[i][c];
                  if constexpr (XLDS)
                    {
                      const int n = i * CPL + c;
                      const float4 v = sx[(n / 4) * NT + tid];
                      xv = (n & 3) == 0 ? v.x : (n & 3) == 1 ? v.y : (n & 3) == 2 ? v.z : v.w;
                    }
                  phl[(Y0 + i) * NP + c] += (double)xv;
                }
        }
      __syncthreads();
    }

  if (pvalid && pl == 0 && wave == 0)
    {
      const size_t o = (size_t)cell * a.nb + basis;
      if (a.iters)
        a.iters[o] = it_total;
      if (a.status)
        a.status[o] = conv ? 0 : 1;
      if (a.resid)
        a.resid[o] = rnorm;
    }
}

// ---- coefficient preparation -----------------------------------------------------
// Element matrix of fine cell f of a box cell (bx) or a general quadrilateral (vq).
__device__ inline void fine_element_2d(int coef_kind, const double *cf, int f, int nf, const double *bx,
                                       const double *vq, double (&K)[4][4])
{
  if (vq)
    qg_element_quad(coef_kind, cf, f, nf, QuadMap(vq), K, 0, 0.0, nullptr);
  else
    qg_element_2d(coef_kind, cf, f, (bx[1] - bx[0]) / nf, (bx[3] - bx[2]) / nf, K);
}

// kModeGlobal: 6 FP64 conductances (c01, c02, c03, c12, c13, c23 = -K_uv) per fine cell.
__global__ void prep_cond_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                             const double *verts, const double *coef, const long long *coef_off, double *cond,
                             const long long *cond_off)
{
  const int ncell = nf * nf;
  const long long total = (long long)n_cells_class * ncell;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x)
    {
      const int cell = cells[t / ncell];
      const int f = (int)(t % ncell);
      double K[4][4];
      fine_element_2d(coef_kind, coef + coef_off[cell], f, nf, verts ? nullptr : box + (size_t)cell * 4,
                      verts ? verts + (size_t)cell * 8 : nullptr, K);
      double *o = cond + cond_off[cell] + 6 * (size_t)f;
      o[0] = -K[0][1];
      o[1] = -K[0][2];
      o[2] = -K[0][3];
      o[3] = -K[1][2];
      o[4] = -K[1][3];
      o[5] = -K[2][3];
    }
}

// kModeStream: node couplings of every node (X, Y) of the tile, summed from the
// (up to two) fine cells that share each edge; FP64 planes at cond + cond_off,
// and (want32) FP32 copies right after them.  Plane entries of padding are
// never written: the kernels read them only for non-interior nodes.
__global__ void prep_couplings_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                                  const double *verts, const double *coef, const long long *coef_off, double *cond,
                                  const long long *cond_off, int want32)
{
  const int np = nf + 1, nnode = np * np, ps = coupling_plane_stride(nf);
  const long long total = (long long)n_cells_class * nnode;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x)
    {
      const int cell = cells[t / nnode];
      const int n = (int)(t % nnode), X = n % np, Y = n / np;
      const double *bx = verts ? nullptr : box + (size_t)cell * 4;
      const double *vq = verts ? verts + (size_t)cell * 8 : nullptr;
      const double *cf = coef + coef_off[cell];
      double aE = 0, aN = 0, aNE = 0, aNW = 0, K[4][4];
      if (X < nf && Y < nf) // cell (X, Y): the node is its vertex 0
        {
          fine_element_2d(coef_kind, cf, Y * nf + X, nf, bx, vq, K);
          aE -= K[0][1];
          aN -= K[0][2];
          aNE -= K[0][3];
        }
      if (X < nf && Y > 0) // cell (X, Y-1): vertex 2
        {
          fine_element_2d(coef_kind, cf, (Y - 1) * nf + X, nf, bx, vq, K);
          aE -= K[2][3];
        }
      if (X > 0 && Y < nf) // cell (X-1, Y): vertex 1
        {
          fine_element_2d(coef_kind, cf, Y * nf + X - 1, nf, bx, vq, K);
          aN -= K[1][3];
          aNW -= K[1][2];
        }
      double *o = cond + cond_off[cell] + n;
      o[0] = aE;
      o[ps] = aN;
      o[2 * ps] = aNE;
      o[3 * ps] = aNW;
      if (want32)
        {
          float *o32 = (float *)(cond + cond_off[cell] + 4 * (size_t)ps) + n;
          o32[0] = (float)aE;
          o32[ps] = (float)aN;
          o32[2 * ps] = (float)aNE;
          o32[3 * ps] = (float)aNW;
        }
    }
}

// Dirichlet data of general quadrilaterals: BasisQ1<2> (basis_q1.tpp:26-46,
// :86-96) is N_i(x) = sum_m C[m][i] mono_m(x), mono = (1, x, y, xy), with
// C = M^{-1}, M[j][m] = mono_m(v_j).  One thread per (cell, boundary node):
// Gauss-Jordan inversion of M with partial pivoting (as deal.II's
// FullMatrix::invert), then N_i at the node's mapped position F(X/nf, Y/nf),
// written into the node's entry of every basis block.
__global__ void quad_dirichlet_2d(int n_cells_class, const int *cells, int nf, const double *verts, double *bases,
                                  const long long *bas_off)
{
  const int nbnd = 4 * nf, np = nf + 1;
  const long long total = (long long)n_cells_class * nbnd;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total; t += (long long)gridDim.x * blockDim.x)
    {
      const int cell = cells[t / nbnd];
      const int k = (int)(t % nbnd), side = k / nf, j = k % nf;
      // boundary walk: y = 0 (X = j), x = nf (Y = j), y = nf (X = nf - j), x = 0 (Y = nf - j)
      const int X = side == 0 ? j : side == 1 ? nf : side == 2 ? nf - j : 0;
      const int Y = side == 0 ? 0 : side == 1 ? j : side == 2 ? nf : nf - j;
      const double *v = verts + (size_t)cell * 8;
      double a[4][8];
      for (int r = 0; r < 4; ++r)
        {
          const double x = v[2 * r], y = v[2 * r + 1];
          a[r][0] = 1;
          a[r][1] = x;
          a[r][2] = y;
          a[r][3] = x * y;
          for (int c = 0; c < 4; ++c)
            a[r][4 + c] = r == c ? 1.0 : 0.0;
        }
      // fully unrolled with the row swap as selects: every index is static, so
      // the 4 x 8 system stays in registers
#pragma unroll
      for (int c = 0; c < 4; ++c)
        {
#pragma unroll
          for (int r = c + 1; r < 4; ++r)
            if (fabs(a[r][c]) > fabs(a[c][c]))
#pragma unroll
              for (int m = 0; m < 8; ++m)
                {
                  const double tmp = a[c][m];
                  a[c][m] = a[r][m];
                  a[r][m] = tmp;
                }
          const double d = 1.0 / a[c][c];
#pragma unroll
          for (int m = 0; m < 8; ++m)
            a[c][m] *= d;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (r != c)
              {
                const double fct = a[r][c];
#pragma unroll
                for (int m = 0; m < 8; ++m)
                  a[r][m] -= fct * a[c][m];
              }
        }
      double x, y;
      QuadMap(v).point((double)X / nf, (double)Y / nf, x, y);
      const double mono[4] = {1.0, x, y, x * y};
      double *phi = bases + bas_off[cell] + (size_t)Y * np + X;
      for (int i = 0; i < 4; ++i)
        {
          double val = 0;
          for (int m = 0; m < 4; ++m)
            val += a[m][4 + i] * mono[m];
          phi[(size_t)i * np * np] = val;
        }
    }
}

// Two-level setup: A_c = P^T K P over the 49 interior nodes of the 8 x 8
// coarse grid of an NF^2 tile (coarse spacing NF / 8), one workgroup of
// kSetupThreads per cell, then A_c^-1 in place (coarse_invert_2d).  K_c of every fine cell is its QGauss(2) element matrix (box cells or
// general quadrilaterals); each fine cell lies in one coarse block, whose 4 x 4
// block matrix sum_c W^T K_c W (W: bilinear weights of the fine cell's
// vertices to the block corners, in the tile's index space) is summed by 4
// threads (a quarter of the block's fine rows each) and combined in a fixed
// order.  cinv: [cells in class][49][49].
constexpr int kSetupThreads = 256;

// Element matrix of fine cell f for the setup kernels: per-fine-cell
// coefficients on a box take the exact Q1 element matrix a rho Kxx + c / rho
// Kyy + b Kxy (rho = hy / hx; unit-square Kxx, Kyy and Kxy = (s t^T + t s^T) /
// 4, s_u = +-1 by x side, t_u by y side), every other kind fine_element_2d.
__device__ inline void setup_element_2d(int coef_kind, const double *cf, int f, int nf, const double *bx,
                                        const double *vq, double (&K)[4][4])
{
  if (!vq && coef_kind <= 1)
    {
      const double a = coef_kind == 0 ? cf[f] : cf[3 * f], b = coef_kind == 0 ? 0.0 : cf[3 * f + 1];
      const double c = coef_kind == 0 ? cf[f] : cf[3 * f + 2];
      const double rho = (bx[3] - bx[2]) / (bx[1] - bx[0]);
      const double ka = a * rho * (1.0 / 6.0), kc = c / rho * (1.0 / 6.0), kb = 0.25 * b;
      for (int u = 0; u < 4; ++u)
        for (int v = 0; v < 4; ++v)
          {
            const double sx = ((u ^ v) & 1) ? -1.0 : 1.0, sy = ((u ^ v) & 2) ? -1.0 : 1.0;
            const double ny = ((u ^ v) & 2) ? 1.0 : 2.0, nx = ((u ^ v) & 1) ? 1.0 : 2.0;
            const double su = (u & 1) ? 1.0 : -1.0, tu = (u & 2) ? 1.0 : -1.0;
            const double sv = (v & 1) ? 1.0 : -1.0, tv = (v & 2) ? 1.0 : -1.0;
            K[u][v] = ka * sx * ny + kc * sy * nx + kb * (su * tv + tu * sv);
          }
    }
  else
    fine_element_2d(coef_kind, cf, f, nf, bx, vq, K);
}

// Intermediate levels of the two-level preconditioner: the diagonals of
// P_s^T K P_s for the bilinear hats of spacing s = s1 = NF / 32 (31 x 31
// interior coarse nodes) and s = s2 = NF / 16 (15 x 15), i.e. 2 and 4 for the
// 64^2 tiles, 4 and 8 for 128^2, inverted and stored as FP32, mlinv [cells in
// class][kMlFloats] = D_s1^-1 [L - 1][K - 1], then D_s2^-1 [L - 1][K - 1] (node
// (K, L) at fine (s K, s L)).  Each s x s block of fine cells contributes the
// four diagonal entries w_m^T K_e w_m of its corners m (w_m = the corner's hat
// at the fine vertices), summed over the block's cells in a fixed order; a
// node's entry is the sum over its four blocks, in a fixed order.
template <int NF>
__global__ void __launch_bounds__(kSetupThreads) ml_diag_2d(int n_cells_class, const int *cells, int coef_kind,
                                                            const double *box, const double *verts, const double *coef,
                                                            const long long *coef_off, float *mlinv)
{
  constexpr int S1 = NF / 32, S2 = NF / 16, NB2 = 32, NB4 = 16, NBLK = NB2 * NB2 + NB4 * NB4;
  __shared__ double sdiag[NBLK][4];
  // box cells with per-fine-cell scalar or tensor coefficients: K_e = ka Kxx +
  // kc Kyy + kb Kxy (setup_element_2d), so w_m^T K_e w_m = ka gxx + kc gyy + kb
  // gxy with g the same for every cell: tabulated per level, fine position in
  // the block and corner, [level][component][ly S + lx][m]
  __shared__ double sg[2][3][S2 * S2][4];
  const int cc = blockIdx.x, t = threadIdx.x;
  if (cc >= n_cells_class)
    return;
  const int cell = cells[cc];
  const double *bx = verts ? nullptr : box + (size_t)cell * 4;
  const double *vq = verts ? verts + (size_t)cell * 8 : nullptr;
  const double *cf = coef + coef_off[cell];
  const bool fast = !vq && coef_kind <= 1; // (uniform)
  if (fast)
    {
      for (int e = t; e < 2 * 3 * S2 * S2 * 4; e += kSetupThreads)
        {
          const int m = e % 4, pos = e / 4 % (S2 * S2), comp = e / (4 * S2 * S2) % 3, lev = e / (12 * S2 * S2);
          const int S = lev ? S2 : S1, lx = pos % S2, ly = pos / S2;
          double g = 0;
          if (lx < S && ly < S)
            {
              double w[4];
              for (int v = 0; v < 4; ++v)
                {
                  const double px = (double)(lx + (v & 1)) / S, py = (double)(ly + (v >> 1)) / S;
                  w[v] = ((m & 1) ? px : 1.0 - px) * ((m & 2) ? py : 1.0 - py);
                }
              for (int u = 0; u < 4; ++u)
                for (int v = 0; v < 4; ++v)
                  {
                    const double sx = ((u ^ v) & 1) ? -1.0 : 1.0, sy = ((u ^ v) & 2) ? -1.0 : 1.0;
                    const double ny = ((u ^ v) & 2) ? 1.0 : 2.0, nx = ((u ^ v) & 1) ? 1.0 : 2.0;
                    const double su = (u & 1) ? 1.0 : -1.0, tu = (u & 2) ? 1.0 : -1.0;
                    const double sv = (v & 1) ? 1.0 : -1.0, tv = (v & 2) ? 1.0 : -1.0;
                    const double k = comp == 0 ? sx * ny : comp == 1 ? sy * nx : su * tv + tu * sv;
                    g += w[u] * k * w[v];
                  }
            }
          sg[lev][comp][ly * S2 + lx][m] = g;
        }
      __syncthreads();
    }
  const double rho = fast ? (bx[3] - bx[2]) / (bx[1] - bx[0]) : 1.0;
  for (int b = t; b < NBLK; b += kSetupThreads)
    {
      const bool l2 = b < NB2 * NB2;
      const int S = l2 ? S1 : S2, nb = l2 ? NB2 : NB4, bb = l2 ? b : b - NB2 * NB2;
      const int bxi = bb % nb, byi = bb / nb;
      double acc[4] = {0, 0, 0, 0};
      if (fast)
        {
          const int lev = l2 ? 0 : 1;
          for (int ly = 0; ly < S; ++ly)
            for (int lx = 0; lx < S; ++lx)
              {
                const int f = (byi * S + ly) * NF + bxi * S + lx;
                const double a = coef_kind == 0 ? cf[f] : cf[3 * f], bq = coef_kind == 0 ? 0.0 : cf[3 * f + 1];
                const double c = coef_kind == 0 ? cf[f] : cf[3 * f + 2];
                const double ka = a * rho * (1.0 / 6.0), kc = c / rho * (1.0 / 6.0), kb = 0.25 * bq;
                const double *g0 = sg[lev][0][ly * S2 + lx], *g1 = sg[lev][1][ly * S2 + lx],
                             *g2 = sg[lev][2][ly * S2 + lx];
                for (int m = 0; m < 4; ++m)
                  acc[m] += (ka * g0[m] + kc * g1[m]) + kb * g2[m];
              }
        }
      else
        {
        for (int ly = 0; ly < S; ++ly)
          for (int lx = 0; lx < S; ++lx)
            {
              double K[4][4];
              setup_element_2d(coef_kind, cf, (byi * S + ly) * NF + bxi * S + lx, NF, bx, vq, K);
              for (int m = 0; m < 4; ++m)
                {
                  double w[4]; // corner m's hat at the fine cell's vertices
                  for (int v = 0; v < 4; ++v)
                    {
                      const double px = (double)(lx + (v & 1)) / S, py = (double)(ly + (v >> 1)) / S;
                      w[v] = ((m & 1) ? px : 1.0 - px) * ((m & 2) ? py : 1.0 - py);
                    }
                  double q = 0;
                  for (int u = 0; u < 4; ++u)
                    q += w[u] * (((K[u][0] * w[0] + K[u][1] * w[1]) + K[u][2] * w[2]) + K[u][3] * w[3]);
                  acc[m] += q;
                }
            }
        }
      for (int m = 0; m < 4; ++m)
        sdiag[b][m] = acc[m];
    }
  __syncthreads();
  float *out = mlinv + (size_t)cc * kMlFloats;
  for (int e = t; e < kMl2 + kMl4; e += kSetupThreads)
    {
      const bool l2 = e < kMl2;
      const int side = l2 ? NB2 - 1 : NB4 - 1, nb = l2 ? NB2 : NB4, off = l2 ? 0 : NB2 * NB2;
      const int k = l2 ? e : e - kMl2, K = k % side + 1, L = k / side + 1;
      // blocks (K-1, L-1) corner 3, (K, L-1) corner 2, (K-1, L) corner 1, (K, L) corner 0
      const double d = ((sdiag[off + (L - 1) * nb + K - 1][3] + sdiag[off + (L - 1) * nb + K][2]) +
                        sdiag[off + L * nb + K - 1][1]) +
                       sdiag[off + L * nb + K][0];
      out[e] = (float)(1.0 / d);
    }
}

// One Gauss-Jordan pivot step of the register-resident inversion in
// coarse_setup_2d (lane i holds row i; K is a compile-time constant, so every
// register index is): the owner of row K publishes it in rb, every lane reads
// it back, then a_ij -= (a_iK p) a_Kj, a_iK = -a_iK p (i != K), a_Kj *= p,
// a_KK = p with p = 1 / a_KK (to within an ulp).
template <int K, int N>
__device__ __forceinline__ void gj_pivot(double (&a)[N], int i, double (*srow)[N + 1])
{
  double *rb = srow[K & 1];
  if (i == K)
    {
#pragma unroll
      for (int j = 0; j + 1 < N; j += 2)
        *(double2 *)&rb[j] = make_double2(a[j], a[j + 1]);
      if (N % 2)
        rb[N - 1] = a[N - 1];
    }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // p = 1 / a_KK: the hardware estimate refined by two Newton steps (a_KK > 0
  // and normal: A_c is SPD), shorter than an IEEE division on the pivot chain
  const double p = reciprocal<double>(rb[K]);
  // a_ij <- sc a_ij - f a_Kj with (sc, f) = (1, a_iK p), and (p, 0) on row K
  // itself (where a_Kj is the lane's own value): one multiply and one FMA per
  // entry, no select
  const bool piv = i == K;
  const double f = piv ? 0.0 : a[K] * p, sc = piv ? p : 1.0;
#pragma unroll
  for (int j = 0; j + 1 < N; j += 2)
    {
      const double2 v = *(const double2 *)&rb[j];
      a[j] = fma(-f, v.x, sc * a[j]);
      a[j + 1] = fma(-f, v.y, sc * a[j + 1]);
    }
  if (N % 2)
    a[N - 1] = fma(-f, rb[N - 1], sc * a[N - 1]);
  a[K] = piv ? p : -f;
}

template <int N, int... K>
__device__ __forceinline__ void gj_all(double (&a)[N], int i, double (*srow)[N + 1], std::integer_sequence<int, K...>)
{
  (gj_pivot<K, N>(a, i, srow), ...);
}

template <int NF>
__global__ void __launch_bounds__(kSetupThreads) coarse_setup_2d(int n_cells_class, const int *cells, int coef_kind,
                                                                const double *box, const double *verts,
                                                                const double *coef, const long long *coef_off,
                                                                double *cinv)
{
  constexpr int H = NF / 8, NB = kCoarseSide + 1, N = kCoarseDofs;
  static_assert(H % 4 == 0 && NB * H == NF, "four threads per coarse block");
  constexpr int SHN = NB * NB * 4 * 10 > N * N ? NB * NB * 4 * 10 : N * N;
  __shared__ double sh[SHN];           // block partials [64][4][10] (upper triangles), then A_c
  __shared__ double kb[NB * NB * 16];  // block matrices [64][4][4]
  // 64^2 box cells with per-fine-cell scalar or tensor coefficients: K_e = ka
  // Kxx + kc Kyy + kb Kxy (setup_element_2d), so W^T K_e W = ka Gxx + kc Gyy +
  // kb Gxy with G the same for every cell, tabulated per fine position in the
  // block: [component][ly H + lx][upper-triangle entry]
  constexpr bool GT = NF == 64;
  __shared__ double sgt[GT ? 3 : 1][GT ? H * H : 1][10];
  const int cc = blockIdx.x, t = threadIdx.x;
  if (cc >= n_cells_class)
    return;
  const int cell = cells[cc];
  const double *bx = verts ? nullptr : box + (size_t)cell * 4;
  const double *vq = verts ? verts + (size_t)cell * 8 : nullptr;
  const double *cf = coef + coef_off[cell];
  // upper-triangle entry e of (m, n), m <= n: 0..3 row 0, 4..6 row 1, 7..8 row 2, 9
  auto tri = [](int m, int n) { return m * 4 - m * (m - 1) / 2 + (n - m); };
  const bool fast = GT && !vq && coef_kind <= 1; // (uniform)
  if (fast)
    {
#pragma unroll 1
      for (int e = t; e < 3 * H * H * 10; e += kSetupThreads)
        {
          const int ent = e % 10, pos = e / 10 % (H * H), comp = e / (10 * H * H);
          const int m = ent < 4 ? 0 : ent < 7 ? 1 : ent < 9 ? 2 : 3, n = m + ent - tri(m, m);
          const int lx = pos % H, ly = pos / H;
          // block corner c's hat at vertex v of the fine cell
          auto wt = [&](int v, int c) -> double {
            const double px = (double)(lx + (v & 1)) / H, py = (double)(ly + (v >> 1)) / H;
            return ((c & 1) ? px : 1.0 - px) * ((c & 2) ? py : 1.0 - py);
          };
          double g = 0;
#pragma unroll
          for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v)
              {
                const double sx = ((u ^ v) & 1) ? -1.0 : 1.0, sy = ((u ^ v) & 2) ? -1.0 : 1.0;
                const double ny = ((u ^ v) & 2) ? 1.0 : 2.0, nx = ((u ^ v) & 1) ? 1.0 : 2.0;
                const double su = (u & 1) ? 1.0 : -1.0, tu = (u & 2) ? 1.0 : -1.0;
                const double sv = (v & 1) ? 1.0 : -1.0, tv = (v & 2) ? 1.0 : -1.0;
                const double k = comp == 0 ? sx * ny : comp == 1 ? sy * nx : su * tv + tu * sv;
                g += wt(u, m) * k * wt(v, n);
              }
          sgt[comp][pos][ent] = g;
        }
      __syncthreads();
    }
  {
    const int b = t >> 2, part = t & 3, bxi = b % NB, byi = b / NB;
    double acc[4][4] = {};
    if (fast)
      {
        const double rho = (bx[3] - bx[2]) / (bx[1] - bx[0]);
        double at[10] = {};
#pragma unroll 1
        for (int ly = part * (H / 4); ly < (part + 1) * (H / 4); ++ly)
#pragma unroll 2
          for (int lx = 0; lx < H; ++lx)
            {
              const int f = (byi * H + ly) * NF + bxi * H + lx;
              const double a = coef_kind == 0 ? cf[f] : cf[3 * f], bq = coef_kind == 0 ? 0.0 : cf[3 * f + 1];
              const double c = coef_kind == 0 ? cf[f] : cf[3 * f + 2];
              const double ka = a * rho * (1.0 / 6.0), kc = c / rho * (1.0 / 6.0), kbq = 0.25 * bq;
              const double *g0 = sgt[0][ly * H + lx], *g1 = sgt[GT ? 1 : 0][ly * H + lx],
                           *g2 = sgt[GT ? 2 : 0][ly * H + lx];
#pragma unroll
              for (int e = 0; e < 10; ++e)
                at[e] += (ka * g0[e] + kc * g1[e]) + kbq * g2[e];
            }
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = m; n < 4; ++n)
            acc[m][n] = at[tri(m, n)];
      }
    else
    for (int ly = part * (H / 4); ly < (part + 1) * (H / 4); ++ly)
      for (int lx = 0; lx < H; ++lx)
        {
          double K[4][4];
          setup_element_2d(coef_kind, cf, (byi * H + ly) * NF + bxi * H + lx, NF, bx, vq, K);
          double Wt[4][4]; // [vertex a][block corner m]
          for (int v = 0; v < 4; ++v)
            {
              const double px = (double)(lx + (v & 1)) / H, py = (double)(ly + (v >> 1)) / H;
              for (int m = 0; m < 4; ++m)
                Wt[v][m] = ((m & 1) ? px : 1.0 - px) * ((m & 2) ? py : 1.0 - py);
            }
          double Q[4][4]; // K W
          for (int u = 0; u < 4; ++u)
            for (int n = 0; n < 4; ++n)
              Q[u][n] = ((K[u][0] * Wt[0][n] + K[u][1] * Wt[1][n]) + K[u][2] * Wt[2][n]) + K[u][3] * Wt[3][n];
          for (int m = 0; m < 4; ++m)
            for (int n = 0; n < 4; ++n)
              acc[m][n] += ((Wt[0][m] * Q[0][n] + Wt[1][m] * Q[1][n]) + Wt[2][m] * Q[2][n]) + Wt[3][m] * Q[3][n];
        }
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int n = m; n < 4; ++n)
        sh[(b * 4 + part) * 10 + tri(m, n)] = acc[m][n];
  }
  __syncthreads();
  // block matrices, symmetric: the upper triangle's sums, mirrored
  for (int e = t; e < NB * NB * 16; e += kSetupThreads)
    {
      const int b = e >> 4, m = (e >> 2) & 3, n = e & 3;
      const int k = m <= n ? tri(m, n) : tri(n, m);
      kb[e] = ((sh[(b * 4 + 0) * 10 + k] + sh[(b * 4 + 1) * 10 + k]) + sh[(b * 4 + 2) * 10 + k]) +
              sh[(b * 4 + 3) * 10 + k];
    }
  __syncthreads();
  // A_c(I, J) = sum over the blocks holding both coarse nodes
  for (int e = t; e < N * N; e += kSetupThreads)
    {
      const int i = e / N, col = e % N;
      const int ix = i % kCoarseSide + 1, iy = i / kCoarseSide + 1;
      const int jx = col % kCoarseSide + 1, jy = col / kCoarseSide + 1;
      double v = 0;
      for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx)
          {
            const int bxi = ix - 1 + dx, byi = iy - 1 + dy; // block with corner I
            const int mi = (ix - bxi) + 2 * (iy - byi), ox = jx - bxi, oy = jy - byi;
            if (ox >= 0 && ox <= 1 && oy >= 0 && oy <= 1)
              v += kb[(byi * NB + bxi) * 16 + mi * 4 + ox + 2 * oy];
          }
      sh[e] = v;
    }
  __syncthreads();
  double *out = cinv + (size_t)cc * N * N;
  for (int e = t; e < N * N; e += kSetupThreads)
    out[e] = sh[e];
}

// In-place inversion of the coarse matrices written by coarse_setup_2d, one
// wave per cell: Gauss-Jordan without pivoting (A_c is SPD) with p = 1 / a_kk,
// a_ij -= (a_ik p) a_kj (i, j != k), a_kj *= p, a_ik *= -p, a_kk = p.  Lane i
// keeps row i in registers (every index a compile-time constant: the pivots
// are unrolled by gj_all); per pivot the owner of row k publishes it in the
// wave's LDS slots and every lane reads it back (the wave's LDS operations run
// in order, so a wave-level fence suffices).  No workgroup barrier, and small
// enough (registers, LDS) for several waves per SIMD to overlap their pivot
// chains.
__global__ void __launch_bounds__(256) coarse_invert_2d(int n_cells_class, double *cinv)
{
  constexpr int N = kCoarseDofs, WPB = 256 / kWave;
  __shared__ __align__(16) double srow[WPB][2][N + 1];
  const int wave = threadIdx.x / kWave, i = threadIdx.x % kWave;
  const int cc = blockIdx.x * WPB + wave;
  if (cc >= n_cells_class)
    return;
  // lane i takes column i of A_c as its row (A_c is symmetric, up to the
  // rounding of its assembly; the loads and stores are then coalesced) and
  // stores its row of the inverse as column i: the result is the transpose of
  // the inverse of A_c^T, i.e. A_c^-1 to rounding
  double *m = cinv + (size_t)cc * N * N;
  double a[N];
#pragma unroll
  for (int j = 0; j < N; ++j)
    a[j] = i < N ? m[j * N + i] : 0.0;
  gj_all(a, i, srow[wave], std::make_integer_sequence<int, N>{});
  if (i < N)
#pragma unroll
    for (int j = 0; j < N; ++j)
      m[j * N + i] = a[j];
}

// ---- launch table -----------------------------------------------------------------
// Variant ids per fine resolution (the autotuner times every id and keeps the
// fastest per class; id 0 is the built-in default):
//   16:  1 = (SW 2, R 8, 2 waves/EU)   2 = (SW 1, R 15, 2)    3 = (SW 4, R 4, 4)
//   32:  1 = (SW 4, R 8, 2)            2 = (SW 2, R 16, 1)    3 = (SW 8, R 4, 4)
//   64:  1 = (SW 8, R 8, 2)            2 = (SW 4, R 16, 1)    3 = (SW 16, R 4, 4)   4 = (SW 8, R 8, 4)
//        5 = (SW 4, R 16, 3)
//   128: 1 = (SW 16, R 8, CPL 2, 4)
// The two-level 64^2 box classes run shape 4 for id 4 and shape 1 otherwise,
// except that id 0 gives FP32 shape 4 (two workgroups per CU).
// kModeStream (one problem per workgroup, 64 lanes across):
//   64:  1 = (SW 8, R 8, 2)            2 = (SW 4, R 16, 1)    3 = (SW 16, R 4, 4)
//   128: 1 = (SW 16, R 8, CPL 2, 4)
namespace {

template <typename T, int NF, int CPL, int SW, int R, int MODE, int WPE, bool CS = false>
hipError_t launch_one(const LocalArgs &a, hipStream_t st)
{
  constexpr int PPW = kWave / (NF / CPL);
  // a multiple of 8 blocks (xcd_block is a bijection on it); blocks past the
  // last problem find no valid problem and leave at the first residual test
  const int grid = ((a.n_problems + PPW - 1) / PPW + 7) / 8 * 8;
  if (a.n_problems == 0)
    return hipSuccess;
  hipLaunchKernelGGL((local_solve_2d<T, NF, CPL, SW, R, MODE, WPE, CS>), dim3(grid), dim3(kWave * SW), 0, st, a);
  return hipGetLastError();
}

template <typename T, int MODE>
hipError_t launch_mode(int nf, int variant, const LocalArgs &a, hipStream_t st)
{
  switch (nf)
    {
      case 16:
        switch (variant)
          {
            case 2: return launch_one<T, 16, 1, 1, 15, MODE, 2>(a, st);
            case 3: return launch_one<T, 16, 1, 4, 4, MODE, 4>(a, st);
            default: return launch_one<T, 16, 1, 2, 8, MODE, 2>(a, st);
          }
      case 32:
        switch (variant)
          {
            case 2: return launch_one<T, 32, 1, 2, 16, MODE, 1>(a, st);
            case 3: return launch_one<T, 32, 1, 8, 4, MODE, 4>(a, st);
            default: return launch_one<T, 32, 1, 4, 8, MODE, 2>(a, st);
          }
      case 64:
        // the (8 waves, 8 rows) shapes also come in the two-level form (a.cinv set)
        if constexpr (MODE == kModeScalar || MODE == kModeTensor)
          if (a.cinv)
            return (variant == 4 || (variant == 0 && sizeof(T) == 4))
                     ? launch_one<T, 64, 1, 8, 8, MODE, 4, true>(a, st)
                     : launch_one<T, 64, 1, 8, 8, MODE, 2, true>(a, st);
        switch (variant)
          {
            case 2: return launch_one<T, 64, 1, 4, 16, MODE, 1>(a, st);
            case 3: return launch_one<T, 64, 1, 16, 4, MODE, 4>(a, st);
            case 4: return launch_one<T, 64, 1, 8, 8, MODE, 4>(a, st);
            case 5: return launch_one<T, 64, 1, 4, 16, MODE, 3>(a, st);
            default: return launch_one<T, 64, 1, 8, 8, MODE, 2>(a, st);
          }
      case 128:
        return launch_one<T, 128, 2, 16, 8, MODE, 4>(a, st);
      default:
        return hipErrorInvalidValue;
    }
}

} // namespace

template <typename T>
hipError_t launch_stream(int nf, int variant, const LocalArgs &a, hipStream_t st)
{
  if (a.cinv) // two-level form of the default shapes
    {
      if (nf == 64)
        return launch_one<T, 64, 1, 8, 8, kModeStream, 2, true>(a, st);
      if (nf == 128)
        return launch_one<T, 128, 2, 16, 8, kModeStream, 4, true>(a, st);
      return hipErrorInvalidValue;
    }
  switch (nf)
    {
      case 64:
        switch (variant)
          {
            case 2: return launch_one<T, 64, 1, 4, 16, kModeStream, 1>(a, st);
            case 3: return launch_one<T, 64, 1, 16, 4, kModeStream, 4>(a, st);
            default: return launch_one<T, 64, 1, 8, 8, kModeStream, 2>(a, st);
          }
      case 128:
        return launch_one<T, 128, 2, 16, 8, kModeStream, 4>(a, st);
      default:
        return hipErrorInvalidValue;
    }
}

bool local2d_coarse_space(int nf, int mode, int variant)
{
  if (mode == kModeStream)
    return nf == 128 || (nf == 64 && (variant == 0 || variant == 1 || variant > 3));
  return nf == 64 && (mode == kModeScalar || mode == kModeTensor) && (variant == 0 || variant == 1 || variant == 4);
}

int local2d_variants(int nf)
{
  switch (nf)
    {
      case 16: return 3;
      case 32: return 3;
      case 64: return 5;
      case 128: return 1;
      default: return 0;
    }
}

hipError_t launch_local_solve_2d(bool fp32, int mode, int nf, int variant, const LocalArgs &a, hipStream_t st)
{
  if (mode == kModeStream)
    {
      // variant ids beyond the streamed table of a resolution take its default
      if (fp32)
        return launch_stream<float>(nf, variant, a, st);
      return launch_stream<double>(nf, variant, a, st);
    }
  if (fp32)
    {
      if (mode == kModeScalar)
        return launch_mode<float, kModeScalar>(nf, variant, a, st);
      if (mode == kModeTensor)
        return launch_mode<float, kModeTensor>(nf, variant, a, st);
      return launch_mode<float, kModeGlobal>(nf, variant, a, st);
    }
  if (mode == kModeScalar)
    return launch_mode<double, kModeScalar>(nf, variant, a, st);
  if (mode == kModeTensor)
    return launch_mode<double, kModeTensor>(nf, variant, a, st);
  return launch_mode<double, kModeGlobal>(nf, variant, a, st);
}

hipError_t launch_prep_couplings_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                                   const double *verts, const double *coef, const long long *coef_off,
                                   double *cond, const long long *cond_off, bool want32, hipStream_t st)
{
  const long long total = (long long)n_cells_class * (nf + 1) * (nf + 1);
  if (total == 0)
    return hipSuccess;
  const int block = 256;
  long long grid = (total + block - 1) / block;
  if (grid > 65536)
    grid = 65536;
  hipLaunchKernelGGL(prep_couplings_2d, dim3((unsigned)grid), dim3(block), 0, st, n_cells_class, cells, nf,
                     coef_kind, box, verts, coef, coef_off, cond, cond_off, want32 ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_prep_cond_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                               const double *verts, const double *coef, const long long *coef_off, double *cond,
                               const long long *cond_off, hipStream_t st)
{
  const long long total = (long long)n_cells_class * nf * nf;
  if (total == 0)
    return hipSuccess;
  const int block = 256;
  long long grid = (total + block - 1) / block;
  if (grid > 65536)
    grid = 65536;
  hipLaunchKernelGGL(prep_cond_2d, dim3((unsigned)grid), dim3(block), 0, st, n_cells_class, cells, nf, coef_kind,
                     box, verts, coef, coef_off, cond, cond_off);
  return hipGetLastError();
}

hipError_t launch_coarse_setup_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                                  const double *verts, const double *coef, const long long *coef_off, double *cinv,
                                  hipStream_t st)
{
  if (n_cells_class == 0)
    return hipSuccess;
  if (nf == 64)
    hipLaunchKernelGGL(coarse_setup_2d<64>, dim3(n_cells_class), dim3(kSetupThreads), 0, st, n_cells_class, cells,
                       coef_kind, box, verts, coef, coef_off, cinv);
  else if (nf == 128)
    hipLaunchKernelGGL(coarse_setup_2d<128>, dim3(n_cells_class), dim3(kSetupThreads), 0, st, n_cells_class, cells,
                       coef_kind, box, verts, coef, coef_off, cinv);
  else
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(coarse_invert_2d, dim3((n_cells_class + 3) / 4), dim3(256), 0, st, n_cells_class, cinv);
  return hipGetLastError();
}

hipError_t launch_ml_diag_2d(int n_cells_class, const int *cells, int nf, int coef_kind, const double *box,
                             const double *verts, const double *coef, const long long *coef_off, float *mlinv,
                             hipStream_t st)
{
  if (n_cells_class == 0)
    return hipSuccess;
  if (nf == 64)
    hipLaunchKernelGGL(ml_diag_2d<64>, dim3(n_cells_class), dim3(kSetupThreads), 0, st, n_cells_class, cells,
                       coef_kind, box, verts, coef, coef_off, mlinv);
  else if (nf == 128)
    hipLaunchKernelGGL(ml_diag_2d<128>, dim3(n_cells_class), dim3(kSetupThreads), 0, st, n_cells_class, cells,
                       coef_kind, box, verts, coef, coef_off, mlinv);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_quad_dirichlet_2d(int n_cells_class, const int *cells, int nf, const double *verts, double *bases,
                                    const long long *bas_off, hipStream_t st)
{
  const long long total = (long long)n_cells_class * 4 * nf;
  if (total == 0)
    return hipSuccess;
  const int block = 256;
  long long grid = (total + block - 1) / block;
  if (grid > 65536)
    grid = 65536;
  hipLaunchKernelGGL(quad_dirichlet_2d, dim3((unsigned)grid), dim3(block), 0, st, n_cells_class, cells, nf, verts,
                     bases, bas_off);
  return hipGetLastError();
}

} // namespace msfem
