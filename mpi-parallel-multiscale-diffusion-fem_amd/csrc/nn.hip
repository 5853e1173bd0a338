// Neural-network primitives of the learned basis surrogate: BF16 MFMA GEMM
// with FP32 accumulation, im2col/col2im for small convolutions, bias+ReLU,
// 2x2 average pooling, nearest upsampling, masked MSE, Adam.  Activations are
// FP32 in CNHW layout ([channel][sample][y][x]); weights are FP32 masters with
// BF16 copies [Cout][Cin*kh*kw] used by the matrix cores.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "msfem.h"
#include "msfem_device.hpp"

namespace msfem {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) // NaN stays NaN
    return (uint16_t)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u); // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// C[M][N] (+)= op(A)[M][K] * op(B)[K][N]; A is [M][K] (TA: [K][M]), B is [K][N]
// (TB: [N][K]), all row-major BF16; C FP32 row-major.  Tile BM x 128 per
// workgroup (BM = 32, 64 or 128 after M), four waves of 32x32
// v_mfma_f32_32x32x8_bf16 blocks; K staged through LDS 32 at a time (k
// contiguous per row, so each lane's MFMA operand is one 8-byte LDS read),
// with the next K tile's global loads in flight during the current MFMAs.
// VEC: 16-byte global loads (leading dimensions multiples of 8, 16-byte
// aligned bases); otherwise element loads.  Zero fill beyond the edges.
// Split K (gridDim.z > 1): block z covers K range [z kchunk, (z+1) kchunk) and
// writes its partial tile to work[z][M][N] for a fixed-order reduction.
constexpr int kGBN = 128, kGBK = 32, kGPad = 4;

typedef uint16_t bf16x8_t __attribute__((ext_vector_type(8)));
typedef uint16_t bf16x4_t __attribute__((ext_vector_type(4)));

// 8 bf16 into an LDS row at an 8-byte aligned offset (rows are 72 bytes)
__device__ __forceinline__ void store8(uint16_t *dst, bf16x8_t v)
{
  *(bf16x4_t *)dst = v.lo;
  *(bf16x4_t *)(dst + 4) = v.hi;
}

// Geometry of an implicit convolution operand: B[k][n] = bf16(in[ci][b][y + ky
// - pad][x + kx - pad]) with k = (ci kh + ky) kw + kx, n = (b Ho + y) Wo + x
// (the im2col matrix, never stored).  Wo must be a multiple of 8.
struct ConvGeo {
  const float *in;
  int B, H, W, kh, kw, pad, Ho, Wo;
  const float *bias; // fused epilogue without split K: out = act(acc + bias[row]); nullptr: none
  int relu;
};

template <int BM, bool TA, bool TB, bool VEC, bool CONV = false>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const uint16_t *A, int lda,
                                                        const uint16_t *B, int ldb, float *C, int ldc, int accumulate,
                                                        int kchunk, float *work, ConvGeo cg = ConvGeo())
{
  static_assert(!CONV || !TB, "implicit convolution: B is [k][n]");
  static_assert(TB || (kGBN / 8 == 16 && kGBK / 2 == 16), "B[k][n] tile: 16 column octets x 16 k pairs");
  // LDS row (bf16): 72 bytes = 18 banks.  The MFMA operand reads (32 rows, 8
  // bytes each) then cover the 64 banks once, and so do the k-pair writes of
  // a B[k][n] tile (rows 8 apart, 144 = 16 mod 64 banks, plus the pair index).
  constexpr int LDR = kGBK + kGPad;
  __shared__ __align__(16) uint16_t sA[BM][LDR];
  __shared__ __align__(16) uint16_t sB[kGBN][LDR];
  constexpr int WM = BM == 32 ? 1 : 2, WN = 4 / WM;          // waves along M, N
  constexpr int TM = BM / WM / 32, TN = kGBN / WN / 32;      // 32x32 blocks per wave
  constexpr int ACH = BM * kGBK / 8, BCH = kGBN * kGBK / 8;  // 8-element chunks per tile
  constexpr int APT = (ACH + 255) / 256, BPT = BCH / 256;     // chunks per thread

  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * kGBN;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / WN) * TM * 32, wn = (wave % WN) * TN * 32;

  // one chunk: 8 consecutive elements of a row of the stored matrix
  auto load8 = [&](const uint16_t *base, int ld, int row, int rows, int col, int cols) -> bf16x8_t {
    bf16x8_t v;
    if (VEC && row < rows && col + 7 < cols)
      v = *(const bf16x8_t *)(base + (size_t)row * ld + col);
    else
      {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          v[e] = (row < rows && col + e < cols) ? base[(size_t)row * ld + col + e] : (uint16_t)0;
      }
    return v;
  };
  bf16x8_t ra[APT], rb[BPT];
  // B[k][n] tile: thread tid holds column octet tid / 16 at rows k = 2 (tid %
  // 16) + t, t = 0, 1, so it writes LDS as 32-bit k pairs.  CONV: the octet is
  // fixed per thread, so its (b, y, x0) are decoded once
  const int boct = tid / 16, bk2 = 2 * (tid % 16);
  int cb = 0, cy = 0, cx0 = 0;
  bool cvalid = false;
  if constexpr (CONV)
    {
      const int nn = n0 + boct * 8;
      cvalid = nn < N;
      cx0 = nn % cg.Wo;
      const int rest = nn / cg.Wo;
      cy = rest % cg.Ho;
      cb = rest / cg.Ho;
    }
  auto conv8 = [&](int k) -> bf16x8_t {
    bf16x8_t v;
    const int taps = cg.kh * cg.kw, ci = k / taps, r = k - ci * taps, ky = r / cg.kw, kx = r - ky * cg.kw;
    const int yy = cy + ky - cg.pad;
    const bool rowok = cvalid && k < ke && yy >= 0 && yy < cg.H;
    const float *src = cg.in + ((size_t)(ci * cg.B + cb) * cg.H + (rowok ? yy : 0)) * cg.W;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      {
        const int xx = cx0 + e + kx - cg.pad;
        v[e] = (rowok && xx >= 0 && xx < cg.W) ? f32_to_bf16(src[xx]) : (uint16_t)0;
      }
    return v;
  };
  auto fetch = [&](int k0) {
#pragma unroll
    for (int t = 0; t < APT; ++t)
      {
        const int c = tid + 256 * t;
        if (c < ACH)
          {
            if (!TA) // A[m][k]: chunk along k
              ra[t] = load8(A, lda, m0 + c / 4, M, k0 + (c % 4) * 8, ke);
            else // A stored [k][m]: chunk along m
              ra[t] = load8(A, lda, k0 + c / (BM / 8), ke, m0 + (c % (BM / 8)) * 8, M);
          }
      }
#pragma unroll
    for (int t = 0; t < BPT; ++t)
      {
        const int c = tid + 256 * t;
        if constexpr (CONV) // gathered straight from the activations
          rb[t] = conv8(k0 + bk2 + t);
        else if (TB) // B stored [n][k]: chunk along k
          rb[t] = load8(B, ldb, n0 + c / 4, N, k0 + (c % 4) * 8, ke);
        else // B[k][n]: chunk along n
          rb[t] = load8(B, ldb, k0 + bk2 + t, ke, n0 + boct * 8, N);
      }
  };
  auto stash = [&]() {
#pragma unroll
    for (int t = 0; t < APT; ++t)
      {
        const int c = tid + 256 * t;
        if (c < ACH)
          {
            if (!TA)
              store8(&sA[c / 4][(c % 4) * 8], ra[t]);
            else
              {
                const int k = c / (BM / 8), mq = (c % (BM / 8)) * 8;
#pragma unroll
                for (int e = 0; e < 8; ++e)
                  sA[mq + e][k] = ra[t][e];
              }
          }
      }
    if constexpr (TB)
      {
#pragma unroll
        for (int t = 0; t < BPT; ++t)
          {
            const int c = tid + 256 * t;
            store8(&sB[c / 4][(c % 4) * 8], rb[t]);
          }
      }
    else
      {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          *(uint32_t *)&sB[boct * 8 + e][bk2] = (uint32_t)rb[0][e] | ((uint32_t)rb[1][e] << 16);
      }
  };

  f32x16 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e)
        acc[i][j][e] = 0.0f;

  if (kb < ke)
    {
      fetch(kb);
      stash();
    }
  __syncthreads();
  for (int k0 = kb; k0 < ke; k0 += kGBK)
    {
      const bool more = k0 + kGBK < ke;
      if (more)
        fetch(k0 + kGBK); // in flight during the MFMAs below
#pragma unroll
      for (int ks = 0; ks < kGBK; ks += 8)
        {
          const int kq = ks + (lane / 32) * 4;
          bf16x4 av[TM], bv[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i)
            av[i] = *(const bf16x4 *)&sA[wm + i * 32 + lane % 32][kq];
#pragma unroll
          for (int j = 0; j < TN; ++j)
            bv[j] = *(const bf16x4 *)&sB[wn + j * 32 + lane % 32][kq];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av[i], bv[j], acc[i][j], 0, 0, 0);
        }
      __syncthreads();
      if (more)
        {
          stash();
          __syncthreads();
        }
    }
  float *out = C;
  int ldo = ldc;
  bool acc_in = accumulate != 0;
  if (gridDim.z > 1)
    {
      out = work + (size_t)blockIdx.z * M * N;
      ldo = N;
      acc_in = false;
    }
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e)
        {
          const int row = m0 + wm + i * 32 + 8 * (e / 4) + 4 * (lane / 32) + (e % 4);
          const int col = n0 + wn + j * 32 + lane % 32;
          if (row < M && col < N)
            {
              float *c = out + (size_t)row * ldo + col;
              float v = acc_in ? *c + acc[i][j][e] : acc[i][j][e];
              if constexpr (CONV)
                if (cg.bias && gridDim.z == 1)
                  {
                    v += cg.bias[row];
                    if (cg.relu && v < 0.0f) // as bias_act_kernel (NaN passes)
                      v = 0.0f;
                  }
              *c = v;
            }
        }
}

typedef __bf16 bf16x8m_t __attribute__((ext_vector_type(8))); // v_mfma_f32_32x32x16_bf16 operand

// 2x2 average pooling of one finished output tile -- BM channels x kGBN pixels
// of whole rows (W >= 16) or whole W x W images (W <= 8), staged as floats
// s[channel][pixel] in LDS -- into pool[c][b][y/2][x/2] (plane/4 values per
// channel), in pool2_fwd_kernel's order of operations, so the pooled values
// are bitwise those of msfem_nn_pool2 on the stored output.  The tile's pooled
// pixels are contiguous from p0 / 4 (the tile starts on an even row).
template <int BM, int W>
__device__ __forceinline__ void pool_tile(const float *s, int m0, size_t plane, int p0, float *pool)
{
  constexpr int Q = kGBN / 4; // pooled pixels per tile
  static_assert(W >= 16 ? (kGBN / W) % 2 == 0 : W % 2 == 0, "tiles of whole 2x2 windows");
  for (int idx = threadIdx.x; idx < BM * Q; idx += blockDim.x)
    {
      const int r = idx / Q, q = idx % Q;
      int p00;
      if constexpr (W >= 16)
        p00 = 2 * (q / (W / 2)) * W + 2 * (q % (W / 2));
      else
        {
          constexpr int PER = W * W / 4;
          const int qq = q % PER;
          p00 = (q / PER) * W * W + 2 * (qq / (W / 2)) * W + 2 * (qq % (W / 2));
        }
      const float *t = s + r * kGBN + p00;
      pool[(size_t)(m0 + r) * (plane / 4) + (size_t)p0 / 4 + q] = 0.25f * ((t[0] + t[1]) + (t[W] + t[W + 1]));
    }
}

// Direct 3x3 convolution, padding 1, stride 1, on the matrix cores:
//   out[co][b][y][x] = act(bias[co] + sum_{ci,ky,kx} w[co][ci][ky][kx] in[ci][b][y + ky - 1][x + kx - 1])
// with BF16 operands and FP32 accumulation.  A workgroup owns BM output
// channels x 128 output pixels = TR = 128 / W whole rows of one image (W <= 8:
// TI = 128 / W^2 whole W x W images).  Per block of 32 input channels it stages
// the TR + 2 input rows around them (zero halo) of each image in LDS with the
// channels innermost (80-byte pixels) and the
// block's weights as [tap][co][ci]; the 9 taps then take their
// v_mfma_f32_32x32x16_bf16 operands from LDS, so each input value is read from
// memory once per channel block (the implicit GEMM reads it once per tap).
// The next block's input and weights are loaded into registers during the
// current block's MFMAs.  wp: weights packed by conv3x3_pack_kernel,
// wp[co][blk][tap][c] = w[co][32 blk + c][tap].  The sum over k runs tap by
// tap within a channel block, so results differ from the implicit GEMM's by
// FP32 rounding only.
// IB: in holds BF16 values (they enter the staging as they are, i.e. as the
// FP32 values they were rounded from would); OB: out receives the BF16
// rounding of each FP32 result (pool and add stay FP32).
template <int BM, int W, bool UP, bool IB, bool OB>
__global__ void __launch_bounds__(256) conv3x3_bf16_kernel(int Cin, int B, int H, const void *in_, const uint16_t *wp,
                                                           void *out_, const float *bias, int relu, const float *add,
                                                           float *pool)
{
  constexpr int TI = W >= 16 ? 1 : kGBN / (W * W), TR = kGBN / W / TI; // images and rows per tile (W <= 8: H = W)
  constexpr int IR = TR + 2, IC = W + 2, NPIX = TI * IR * IC;
  constexpr int CP = 40;                                  // LDS pitch (bf16) of a pixel and of a weight row
  constexpr int NG = W / 4, IIT = TI * IR * NG * 4;      // input items (staged row, 4-column group, channel octet)
  static_assert(IIT <= 256, "one input item per thread");
  constexpr int WIT = BM * 36, WPT = (WIT + 255) / 256;   // weight items (co, tap, channel octet)
  constexpr int WM = BM == 32 ? 1 : 2, WN = 4 / WM;       // waves along co, pixels
  constexpr int TM = BM / WM / 32, TN = kGBN / WN / 32;   // 32x32 blocks per wave
  static_assert(kGBN % W == 0 && TM == 1, "tile of whole rows, one 32-row block per wave");
  __shared__ __align__(16) uint16_t sIn[NPIX * CP];
  __shared__ __align__(16) uint16_t sW[9 * BM * CP];

  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / WN) * 32, wn = (wave % WN) * TN * 32;
  const int HW = H * W;
  const size_t plane = (size_t)B * HW; // channel stride of in and out
  const int p0 = blockIdx.x * kGBN;    // first output pixel, flattened (b, y, x)
  const int b = p0 / HW, y0 = (p0 - b * HW) / W;
  const int m0 = blockIdx.y * BM;
  const int K = Cin * 9, nblk = Cin / 32;

  // Input staging: item (channel octet, 4-column group, staged row), octet
  // fastest (the 16-byte LDS stores of a quarter wave then hit distinct
  // banks); each loads 8 channel planes x 4 columns as float4 and is kept raw
  // in registers during the MFMAs, converted to BF16 at the LDS store.  The
  // halo columns x = -1 and x = W are outside every image (tiles are whole
  // rows) and are zeroed once.
  // UP: in is the half-resolution tensor [Cin][B][H/2][W/2], read at (y/2,
  // x/2) (nearest upsampling): 2 source columns (float2) per 4-column group.
  float4 ri[UP ? 1 : 8];
  float2 rh[UP ? 8 : 1];
  bf16x8_t rw[WPT];
  const int ioct = tid % 4, ig = tid / 4 % NG, irr = tid / 4 / NG; // irr = img IR + row
  const int iy = y0 - 1 + irr % IR;
  const bool iact = tid < IIT, iok = iact && iy >= 0 && iy < H;
  constexpr bool ib = IB; // BF16 input: the raw pairs sit in the .x/.y registers
  const size_t iplane = UP ? plane / 4 : plane; // channel stride of in
  const size_t ioff = iok ? (UP ? (size_t)(b + irr / IR) * (HW / 4) + (size_t)(iy / 2) * (W / 2) + 2 * ig
                                : (size_t)(b + irr / IR) * HW + (size_t)iy * W + 4 * ig)
                          : 0;
  auto fetch = [&](int blk) {
    if (iok)
      {
        const size_t so = ioff + (size_t)(blk * 32 + ioct * 8) * iplane;
        if constexpr (IB)
          {
            const uint16_t *src = (const uint16_t *)in_ + so;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if constexpr (UP)
                rh[e].x = __uint_as_float(*(const uint32_t *)(src + e * iplane));
              else
                {
                  const uint2 q = *(const uint2 *)(src + e * iplane);
                  ri[e].x = __uint_as_float(q.x);
                  ri[e].y = __uint_as_float(q.y);
                }
          }
        else
          {
            const float *src = (const float *)in_ + so;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if constexpr (UP)
                rh[e] = *(const float2 *)(src + e * iplane);
              else
                ri[e] = *(const float4 *)(src + e * iplane);
          }
      }
#pragma unroll
    for (int t = 0; t < WPT; ++t)
      {
        const int u = tid + 256 * t;
        if (u < WIT)
          rw[t] = *(const bf16x8_t *)(wp + (size_t)(m0 + u / 36) * K + blk * 288 + (u % 36) * 8);
      }
  };
  auto stash = [&]() {
    if (iact)
      {
        uint16_t *dst = &sIn[(irr * IC + 1 + 4 * ig) * CP + ioct * 8];
#pragma unroll
        for (int px = 0; px < 4; ++px)
          {
            bf16x8_t v;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              {
                uint16_t h;
                if constexpr (IB && UP) // (no ternary over rh and ri: one of them has a single element)
                  {
                    const uint32_t q = __float_as_uint(rh[e].x);
                    h = px < 2 ? (uint16_t)q : (uint16_t)(q >> 16);
                  }
                else if constexpr (IB)
                  {
                    const uint32_t q = __float_as_uint(px < 2 ? ri[e].x : ri[e].y);
                    h = px % 2 == 0 ? (uint16_t)q : (uint16_t)(q >> 16);
                  }
                else
                  {
                    float f;
                    if constexpr (UP)
                      f = px < 2 ? rh[e].x : rh[e].y;
                    else
                      f = px == 0 ? ri[e].x : px == 1 ? ri[e].y : px == 2 ? ri[e].z : ri[e].w;
                    h = f32_to_bf16(f);
                  }
                v[e] = iok ? h : (uint16_t)0;
              }
            *(bf16x8_t *)(dst + px * CP) = v;
          }
      }
#pragma unroll
    for (int t = 0; t < WPT; ++t)
      {
        const int u = tid + 256 * t;
        if (u < WIT)
          {
            const int m = u / 36, j = u % 36; // j = tap * 4 + channel octet
            *(bf16x8_t *)&sW[((j / 4) * BM + m) * CP + (j % 4) * 8] = rw[t];
          }
      }
  };
  for (int h = tid; h < TI * IR * 2 * 4; h += 256) // halo columns: (staged row, side, octet)
    {
      const int rr = h / 8, side = h / 4 % 2, oct = h % 4;
      *(bf16x8_t *)&sIn[(rr * IC + side * (IC - 1)) * CP + oct * 8] = bf16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
    }

  f32x16 acc[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e)
      acc[j][e] = 0.0f;
  // this lane's output pixels n = wn + 32 j + lane % 32 (image n / (TR W), row, column)
  // sit at staged (img IR + row, column) for tap (0, 0); its 8 channels of a
  // 16-channel step start at kh
  int pb[TN];
#pragma unroll
  for (int j = 0; j < TN; ++j)
    {
      const int n = wn + j * 32 + lane % 32, img = n / (TR * W), r = n % (TR * W) / W;
      pb[j] = ((img * IR + r) * IC + n % W) * CP;
    }
  const int kh = 8 * (lane / 32);
  const int arow = (wm + lane % 32) * CP + kh;

  // the epilogue's bias and skip values are loaded now, so that they arrive
  // during the MFMAs instead of costing a memory round trip at the end
  float eb[16]; // bias of the lane's 16 output rows (the same for every j)
#pragma unroll
  for (int e = 0; e < 16; ++e)
    eb[e] = bias ? bias[m0 + wm + 8 * (e / 4) + 4 * (lane / 32) + (e % 4)] : 0.0f;
  float ea[TN][16];
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e)
      {
        const int row = m0 + wm + 8 * (e / 4) + 4 * (lane / 32) + (e % 4);
        ea[j][e] = add ? add[(size_t)row * plane + p0 + wn + j * 32 + lane % 32] : 0.0f;
      }

  fetch(0);
  stash();
  __syncthreads();
  for (int blk = 0; blk < nblk; ++blk)
    {
      const bool more = blk + 1 < nblk;
      if (more)
        fetch(blk + 1); // in flight during the MFMAs below
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
        {
          const int toff = ((tap / 3) * IC + tap % 3) * CP;
#pragma unroll
          for (int ks = 0; ks < 32; ks += 16)
            {
              const bf16x8m_t av = *(const bf16x8m_t *)&sW[tap * BM * CP + arow + ks];
#pragma unroll
              for (int j = 0; j < TN; ++j)
                {
                  const bf16x8m_t bv = *(const bf16x8m_t *)&sIn[pb[j] + toff + kh + ks];
                  acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
                }
            }
        }
      __syncthreads();
      if (more)
        {
          stash();
          __syncthreads();
        }
    }
#pragma unroll
  for (int j = 0; j < TN; ++j)
#pragma unroll
    for (int e = 0; e < 16; ++e)
      {
        const int row = m0 + wm + 8 * (e / 4) + 4 * (lane / 32) + (e % 4);
        float v = acc[j][e];
        if (bias)
          {
            v += eb[e];
            if (relu && v < 0.0f) // as bias_act_kernel (NaN passes)
              v = 0.0f;
          }
        const size_t o = (size_t)row * plane + p0 + wn + j * 32 + lane % 32;
        if (add) // skip connection: act(acc + bias) + add, one FP32 addition
          v += ea[j][e];
        if constexpr (OB)
          ((uint16_t *)out_)[o] = f32_to_bf16(v);
        else
          ((float *)out_)[o] = v;
        if (pool) // the block loop ended with a barrier: sW is free
          ((float *)sW)[(row - m0) * kGBN + wn + j * 32 + lane % 32] = v;
      }
  if (pool) // (uniform)
    {
      static_assert(sizeof(sW) >= BM * kGBN * sizeof(float), "the staged tile fits the weight buffer");
      __syncthreads();
      pool_tile<BM, W>((const float *)sW, m0, plane, p0, pool);
    }
}

// Persistent form of conv3x3_bf16_kernel for 32 output channels and NB <= 2
// blocks of 32 input channels (the full-resolution layers of an
// encoder-decoder, whose workgroups would otherwise restage the same weights
// for every tile and find no work to overlap with their first loads; the
// host takes it for NB = 1, where it measured faster than the tile form): a
// workgroup stages the weights of every block once, then walks the steps
// (tile, block), tiles blockIdx.x + k gridDim.x; the input of the next step is
// loaded into registers during the MFMAs and the epilogue of the current one.
// The sums run in the same order as in conv3x3_bf16_kernel (blocks, taps,
// 16-channel steps), so the results are bitwise those of the tile form.
// FIRST (W = 64, one block): the layer's input is itself computed in the
// staging from a one-channel image, as a network's first layer (1 -> 32
// channels, 3x3, padding 1) produces it: in is then [B][H][W] FP32, w0 that
// layer's BF16 weights [32][9] and b0 its bias.  The staged rows are act(conv(in)
// + b0) formed as the implicit GEMM (gemm_bf16_kernel with CONV) forms them:
// the same v_mfma_f32_32x32x8 steps over k = 0..31 (zero beyond 9) on the
// same operands, then bias and ReLU, so the result is bitwise that of the
// two layers run one after the other, without the 32-channel intermediate in
// memory.
template <int W, bool UP, int NB, bool FIRST>
__global__ void __launch_bounds__(256) conv3x3_tiles_kernel(int B, int H, int ntiles, const float *in,
                                                            const uint16_t *wp, float *out, const float *bias,
                                                            int relu, const float *add, float *pool,
                                                            const uint16_t *w0, const float *b0, int relu0)
{
  constexpr int BM = 32, K = NB * 32 * 9;
  constexpr int TI = W >= 16 ? 1 : kGBN / (W * W), TR = kGBN / W / TI;
  constexpr int IR = TR + 2, IC = W + 2, NPIX = TI * IR * IC;
  constexpr int CP = 40;
  constexpr int NG = W / 4, IIT = TI * IR * NG * 4;
  static_assert(IIT <= 256 && kGBN % W == 0 && NB >= 1 && NB <= 2, "shape outside the persistent form");
  static_assert(!FIRST || (W == 64 && !UP && NB == 1), "the fused first layer: 64-wide rows, one block");
  constexpr int XR = IR + 2;                     // FIRST: one-channel rows staged (the layer's rows and their halo)
  constexpr int XB = IR * W / 32;                // FIRST: 32-pixel blocks of the first layer's staged output
  static_assert(!FIRST || (XR * NG <= 256 && XB % 4 == 0), "one one-channel item per thread, whole blocks per wave");
  __shared__ __align__(16) uint16_t sIn[NPIX * CP];
  __shared__ __align__(16) uint16_t sW[NB * 9 * BM * CP];
  __shared__ __align__(16) uint16_t sX[FIRST ? XR * IC : 1]; // [row][column], zero halo columns

  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wn = wave * 32; // four waves along the pixels, one 32 x 32 block each
  const int HW = H * W;
  const size_t plane = (size_t)B * HW;
  const size_t iplane = UP ? plane / 4 : plane;

  // the bias of the lane's 16 output rows, once (every tile has the same rows)
  float eb[16];
#pragma unroll
  for (int e = 0; e < 16; ++e)
    eb[e] = bias ? bias[8 * (e / 4) + 4 * (lane / 32) + (e % 4)] : 0.0f;
  // the weights of every block, [blk][tap][co][ci], once
  for (int u = tid; u < NB * BM * 36; u += 256)
    {
      const int blk = u / (BM * 36), r = u % (BM * 36), m = r / 36, j = r % 36;
      *(bf16x8_t *)&sW[((blk * 9 + j / 4) * BM + m) * CP + (j % 4) * 8] =
          *(const bf16x8_t *)(wp + (size_t)m * K + blk * 288 + j * 8);
    }
  auto zero_halo = [&]() {
    for (int h = tid; h < TI * IR * 2 * 4; h += 256) // halo columns: (staged row, side, octet)
      {
        const int rr = h / 8, side = h / 4 % 2, oct = h % 4;
        *(bf16x8_t *)&sIn[(rr * IC + side * (IC - 1)) * CP + oct * 8] = bf16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
      }
  };
  zero_halo();

  // FIRST: the first layer's weights as this lane's MFMA A operands (row
  // lane % 32, k = 8 q + 4 (lane / 32) + i, zero beyond k = 8) and the bias of
  // its 16 output rows
  bf16x4 a0[4];
  float eb0[16];
  if constexpr (FIRST)
    {
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          {
            const int k = 8 * q + 4 * (lane / 32) + i;
            a0[q][i] = k < 9 ? (short)w0[(lane % 32) * 9 + k] : (short)0;
          }
#pragma unroll
      for (int e = 0; e < 16; ++e)
        eb0[e] = b0 ? b0[8 * (e / 4) + 4 * (lane / 32) + (e % 4)] : 0.0f;
      for (int h = tid; h < XR * 2; h += 256) // the one-channel rows' halo columns
        sX[(h / 2) * IC + (h % 2) * (IC - 1)] = 0;
    }

  // input staging as in conv3x3_bf16_kernel; iok describes the step whose
  // values are in the registers
  float4 ri[UP ? 1 : 8];
  float2 rh[UP ? 8 : 1];
  const int ioct = tid % 4, ig = tid / 4 % NG, irr = tid / 4 / NG;
  const bool iact = tid < IIT;
  bool iok = false;
  auto fetch = [&](int tile, int blk) {
    const int p0 = tile * kGBN, b = p0 / HW, y0 = (p0 - b * HW) / W;
    const int iy = y0 - 1 + irr % IR;
    iok = iact && iy >= 0 && iy < H;
    if (iok)
      {
        const float *src = in + (UP ? (size_t)(b + irr / IR) * (HW / 4) + (size_t)(iy / 2) * (W / 2) + 2 * ig
                                    : (size_t)(b + irr / IR) * HW + (size_t)iy * W + 4 * ig) +
                           (size_t)(blk * 32 + ioct * 8) * iplane;
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if constexpr (UP)
            rh[e] = *(const float2 *)(src + e * iplane);
          else
            ri[e] = *(const float4 *)(src + e * iplane);
      }
  };
  auto stash = [&]() {
    if (iact)
      {
        uint16_t *dst = &sIn[(irr * IC + 1 + 4 * ig) * CP + ioct * 8];
#pragma unroll
        for (int px = 0; px < 4; ++px)
          {
            bf16x8_t v;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              {
                float f;
                if constexpr (UP)
                  f = px < 2 ? rh[e].x : rh[e].y;
                else
                  f = px == 0 ? ri[e].x : px == 1 ? ri[e].y : px == 2 ? ri[e].z : ri[e].w;
                v[e] = iok ? f32_to_bf16(f) : (uint16_t)0;
              }
            *(bf16x8_t *)(dst + px * CP) = v;
          }
      }
  };
  // FIRST: item (one-channel row, 4-column group) of the rows y0 - 2 .. y0 + IR
  // - 1 around the tile whose input is fetched (xt); the first layer's output
  // rows y0 - 1 .. y0 + IR - 2 then go to sIn as act(acc + b0), zero outside
  // the image
  float4 rx = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  bool xok = false;
  int xt = 0;
  const int xr = tid / NG, xg = tid % NG;
  auto fetch_first = [&](int tile) __attribute__((always_inline)) {
    xt = tile;
    const int p0 = tile * kGBN, b = p0 / HW, y0 = (p0 - b * HW) / W;
    const int iy = y0 - 2 + xr;
    xok = tid < XR * NG && iy >= 0 && iy < H;
    if (xok)
      rx = *(const float4 *)(in + (size_t)b * HW + (size_t)iy * W + 4 * xg);
  };
  auto stash_first = [&]() __attribute__((always_inline)) {
    if (tid < XR * NG)
      {
        uint16_t *d = &sX[xr * IC + 1 + 4 * xg];
        d[0] = xok ? f32_to_bf16(rx.x) : (uint16_t)0;
        d[1] = xok ? f32_to_bf16(rx.y) : (uint16_t)0;
        d[2] = xok ? f32_to_bf16(rx.z) : (uint16_t)0;
        d[3] = xok ? f32_to_bf16(rx.w) : (uint16_t)0;
      }
    __syncthreads();
    const int p0 = xt * kGBN, b = p0 / HW, y0 = (p0 - b * HW) / W;
#pragma unroll
    for (int jj = 0; jj < XB / 4; ++jj)
      {
        const int p = 32 * (wave + 4 * jj) + lane % 32, pr = p / W, pc = p % W;
        f32x16 a;
#pragma unroll
        for (int e = 0; e < 16; ++e)
          a[e] = 0.0f;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          {
            bf16x4 bv;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              {
                const int k = 8 * q + 4 * (lane / 32) + i; // tap (k / 3, k % 3)
                bv[i] = k < 9 ? (short)sX[(pr + k / 3) * IC + pc + k % 3] : (short)0;
              }
            a = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a0[q], bv, a, 0, 0, 0);
          }
        const int y = y0 - 1 + pr;
        const bool rok = y >= 0 && y < H;
#pragma unroll
        for (int g = 0; g < 4; ++g)
          {
            bf16x4_t v;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              {
                float f = a[4 * g + i];
                if (b0)
                  {
                    f += eb0[4 * g + i];
                    if (relu0 && f < 0.0f) // as the implicit GEMM's epilogue (NaN passes)
                      f = 0.0f;
                  }
                v[i] = rok ? f32_to_bf16(f) : (uint16_t)0;
              }
            *(bf16x4_t *)&sIn[(pr * IC + 1 + pc) * CP + 8 * g + 4 * (lane / 32)] = v;
          }
      }
  };
  // (always_inline: a lambda called twice and left out of line keeps its captures in scratch)
  auto fetch_step = [&](int tile, int blk) __attribute__((always_inline)) {
    if constexpr (FIRST)
      fetch_first(tile);
    else
      fetch(tile, blk);
  };
  auto stash_step = [&]() __attribute__((always_inline)) {
    if constexpr (FIRST)
      stash_first();
    else
      stash();
  };

  const int n = wn + lane % 32, img = n / (TR * W), r = n % (TR * W) / W;
  const int pb = ((img * IR + r) * IC + n % W) * CP;
  const int kh = 8 * (lane / 32);
  const int arow = (lane % 32) * CP + kh;
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; ++e)
    acc[e] = 0.0f;

  int tile = blockIdx.x, blk = 0;
  if (tile >= ntiles) // (the host sizes the grid to at most ntiles)
    return;
  fetch_step(tile, 0);
  stash_step();
  __syncthreads();
  for (;;)
    {
      int ntile = tile, nblk = blk + 1;
      if (nblk == NB)
        {
          nblk = 0;
          ntile = tile + gridDim.x;
        }
      const bool more = ntile < ntiles;
      if (more)
        fetch_step(ntile, nblk); // in flight during the MFMAs and the epilogue below
#pragma unroll
      for (int tap = 0; tap < 9; ++tap)
        {
          const int toff = ((tap / 3) * IC + tap % 3) * CP;
#pragma unroll
          for (int ks = 0; ks < 32; ks += 16)
            {
              const bf16x8m_t av = *(const bf16x8m_t *)&sW[(blk * 9 + tap) * BM * CP + arow + ks];
              const bf16x8m_t bv = *(const bf16x8m_t *)&sIn[pb + toff + kh + ks];
              acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
            }
        }
      float pv[16]; // (pool: the tile's values, staged in sIn once every wave is done with it)
      if (blk == NB - 1) // the tile is complete: epilogue, then a fresh accumulator
        {
#pragma unroll
          for (int e = 0; e < 16; ++e)
            {
              const int row = 8 * (e / 4) + 4 * (lane / 32) + (e % 4);
              float v = acc[e];
              if (bias)
                {
                  v += eb[e];
                  if (relu && v < 0.0f) // as bias_act_kernel (NaN passes)
                    v = 0.0f;
                }
              const size_t o = (size_t)row * plane + (size_t)tile * kGBN + n;
              if (add)
                v += add[o];
              out[o] = v;
              pv[e] = v;
              acc[e] = 0.0f;
            }
        }
      __syncthreads(); // every wave is done with sIn
      if constexpr (W == 64) // (the host asks for pooling at W = 64 only: sIn holds the tile there)
        if (pool && blk == NB - 1) // (uniform)
          {
            static_assert(sizeof(sIn) >= BM * kGBN * sizeof(float), "the staged tile fits the input buffer");
            float *so = (float *)sIn;
#pragma unroll
            for (int e = 0; e < 16; ++e)
              so[(8 * (e / 4) + 4 * (lane / 32) + (e % 4)) * kGBN + n] = pv[e];
            __syncthreads();
            pool_tile<BM, W>(so, 0, plane, tile * kGBN, pool);
            __syncthreads();
            zero_halo(); // the staged tile covered the halo columns; the next stash writes only the interior
          }
      if (!more)
        break;
      stash_step();
      __syncthreads();
      tile = ntile;
      blk = nblk;
    }
}

// The sums of conv_small_kernel at output pixel (x, y) of image b: acc[c] =
// sum_k bf16(in) w[k][c], k = (ci KS + ky) KS + kx in this order, KS x KS
// windows of CU input channels loaded at once.  IB: in holds BF16 values
// (the same sums as from the FP32 values they were rounded from).
template <int CO, int KS, bool IB>
__device__ __forceinline__ void small_conv_sums(int Cin, int B, int H, int W, int pad, int x, int y, int b,
                                                const void *in, const float *swk, float (&acc)[CO])
{
  constexpr int CU = 8, T = KS * KS; // input channels loaded at once
  const size_t plane = (size_t)B * H * W, base = (size_t)b * H * W;
#pragma unroll
  for (int c = 0; c < CO; ++c)
    acc[c] = 0.0f;
  for (int ci0 = 0; ci0 < Cin; ci0 += CU)
    {
      float v[CU][T];
#pragma unroll
      for (int u = 0; u < CU; ++u)
#pragma unroll
        for (int t = 0; t < T; ++t)
          {
            const int yy = y + t / KS - pad, xx = x + t % KS - pad;
            const bool ok = ci0 + u < Cin && yy >= 0 && yy < H && xx >= 0 && xx < W;
            const size_t o = ok ? base + (ci0 + u) * plane + yy * W + xx : 0;
            if constexpr (IB)
              v[u][t] = ok ? bf16_to_f32(((const uint16_t *)in)[o]) : 0.0f;
            else
              v[u][t] = ok ? ((const float *)in)[o] : 0.0f;
          }
#pragma unroll
      for (int u = 0; u < CU; ++u)
        if (ci0 + u < Cin)
#pragma unroll
          for (int t = 0; t < T; ++t)
            {
              const float f = IB ? v[u][t] : bf16_to_f32(f32_to_bf16(v[u][t]));
              const float *wk = swk + ((ci0 + u) * T + t) * CO;
#pragma unroll
              for (int c = 0; c < CO; ++c)
                acc[c] = fmaf(f, wk[c], acc[c]);
            }
    }
}

// Direct convolution with few output channels (a network's head) on the
// VALU: one thread per output pixel computes CO channels (CO >= Cout; weights
// of the channels beyond Cout are zero, their sums are not stored) from the
// BF16-rounded input and weights (small_conv_sums).  Weights as FP32 in LDS,
// [k][CO].  Each BF16 x BF16 product is exact in FP32, as in the MFMA path.
template <int CO, int KS>
__global__ void __launch_bounds__(256) conv_small_kernel(int Cin, int B, int H, int W, int pad, int Cout,
                                                         const float *in, const uint16_t *w, float *out,
                                                         const float *bias, int relu)
{
  extern __shared__ float swk[];
  const int K = Cin * KS * KS;
  for (int t = threadIdx.x; t < K * CO; t += blockDim.x)
    swk[t] = t % CO < Cout ? bf16_to_f32(w[(size_t)(t % CO) * K + t / CO]) : 0.0f;
  __syncthreads();
  const int Ho = H + 2 * pad - KS + 1, Wo = W + 2 * pad - KS + 1;
  const size_t N = (size_t)B * Ho * Wo;
  for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < N; n += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(n % Wo), y = (int)(n / Wo % Ho), b = (int)(n / ((size_t)Wo * Ho));
      float acc[CO];
      small_conv_sums<CO, KS, false>(Cin, B, H, W, pad, x, y, b, in, swk, acc);
#pragma unroll
      for (int c = 0; c < CO; ++c)
        if (c < Cout)
          {
            float r = acc[c];
            if (bias)
              {
                r += bias[c];
                if (relu && r < 0.0f) // as bias_act_kernel (NaN passes)
                  r = 0.0f;
              }
            out[c * N + n] = r;
          }
    }
}

// The serving head of the basis surrogate, written as bases: for node (X, Y)
// of cell b's (nf + 1)^2 nodes and basis c < 4, bases[b][c][Y (nf + 1) + X] =
// N_c(X, Y) + (double)(scale * h_c) at the interior nodes, with h_c the 2 x 2
// valid convolution + bias (+ ReLU) at pixel (X - 1, Y - 1) of the nf x nf
// activations, summed as conv_small_kernel sums (k = 4 ci + 2 ky + kx
// ascending, fmaf), and N_c on the boundary: bitwise conv_small_kernel,
// axpby_kernel (a = scale, b = 0) and delta_to_bases_kernel in one pass.  One
// thread per node; an interior node's window lies inside the image, so the
// loads of 8 channels go out unconditionally, back to back.
template <bool IB>
__global__ void __launch_bounds__(256) head_bases_kernel(int Cin, int B, int nf, const void *in, const uint16_t *w,
                                                         const float *bias, int relu, float scale, double *bases)
{
  constexpr int CO = 4, CU = 8;
  extern __shared__ float swk[]; // [k][CO]
  const int K = Cin * 4;
  for (int t = threadIdx.x; t < K * CO; t += blockDim.x)
    swk[t] = bf16_to_f32(w[(size_t)(t % CO) * K + t / CO]);
  __syncthreads();
  const int np = nf + 1;
  const size_t per = (size_t)np * np, total = (size_t)B * per, plane = (size_t)B * nf * nf;
  auto value = [&](size_t o) -> float { // the BF16-rounded activation, as a float
    if constexpr (IB)
      return bf16_to_f32(((const uint16_t *)in)[o]);
    else
      return bf16_to_f32(f32_to_bf16(((const float *)in)[o]));
  };
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int b = (int)(t / per), k = (int)(t % per), X = k % np, Y = k / np;
      const bool inner = X > 0 && X < nf && Y > 0 && Y < nf;
      float acc[CO] = {0.0f, 0.0f, 0.0f, 0.0f};
      if (inner)
        {
          const size_t o0 = (size_t)b * nf * nf + (size_t)(Y - 1) * nf + (X - 1);
          int ci0 = 0;
          for (; ci0 + CU <= Cin; ci0 += CU)
            {
              float v[CU][4];
#pragma unroll
              for (int u = 0; u < CU; ++u)
                {
                  const size_t o = o0 + (ci0 + u) * plane;
                  v[u][0] = value(o);
                  v[u][1] = value(o + 1);
                  v[u][2] = value(o + nf);
                  v[u][3] = value(o + nf + 1);
                }
#pragma unroll
              for (int u = 0; u < CU; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                  {
                    const float *wk = swk + ((ci0 + u) * 4 + q) * CO;
#pragma unroll
                    for (int c = 0; c < CO; ++c)
                      acc[c] = fmaf(v[u][q], wk[c], acc[c]);
                  }
            }
          for (; ci0 < Cin; ++ci0) // (Cin not a multiple of CU)
            {
              const size_t o = o0 + ci0 * plane;
              const float v[4] = {value(o), value(o + 1), value(o + nf), value(o + nf + 1)};
#pragma unroll
              for (int q = 0; q < 4; ++q)
                {
                  const float *wk = swk + (ci0 * 4 + q) * CO;
#pragma unroll
                  for (int c = 0; c < CO; ++c)
                    acc[c] = fmaf(v[q], wk[c], acc[c]);
                }
            }
        }
#pragma unroll
      for (int c = 0; c < CO; ++c)
        {
          double v = nodal_q1(c, 2, X, Y, 0, nf);
          if (inner)
            {
              float r = acc[c];
              if (bias)
                {
                  r += bias[c];
                  if (relu && r < 0.0f)
                    r = 0.0f;
                }
              v += (double)(scale * r);
            }
          bases[((size_t)b * CO + c) * per + k] = v;
        }
    }
}

// The same head from BF16 activations, four nodes per thread: nodes X = 4 p + 1
// .. 4 p + 4 of row Y (the thread of p = 0 also writes the boundary node X =
// 0).  Their windows span the pixel columns 4 p .. 4 p + 4 of rows Y - 1 and
// Y: per channel one 8-byte load (columns 4 p .. 4 p + 3) and one 2-byte load
// (column 4 p + 4) per row, shared by the four nodes, and one LDS read of the
// four outputs' weights per tap.  Every node's sums are formed in the order of
// head_bases_kernel (channel, then tap, ascending; fmaf), so the values are
// bitwise the same.  Needs nf % 4 == 0, Cin % CU == 0 and 8-byte aligned
// activations.
template <int CU>
__global__ void __launch_bounds__(256) head_bases4_kernel(int Cin, int B, int nf, const uint16_t *in,
                                                          const uint16_t *w, const float *bias, int relu,
                                                          float scale, double *bases)
{
  constexpr int CO = 4;
  extern __shared__ float swk[]; // [k][CO]
  const int K = Cin * 4;
  for (int t = threadIdx.x; t < K * CO; t += blockDim.x)
    swk[t] = bf16_to_f32(w[(size_t)(t % CO) * K + t / CO]);
  __syncthreads();
  const int np = nf + 1, nq = nf / 4;
  const size_t per = (size_t)np * np, total = (size_t)B * np * nq, plane = (size_t)B * nf * nf;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int p = (int)(t % nq);
      const size_t row = t / nq;
      const int Y = (int)(row % np), b = (int)(row / np), X0 = 4 * p + 1;
      const bool yin = Y > 0 && Y < nf, last = p == nq - 1; // last: node X0 + 3 = nf is on the boundary
      float acc[4][CO];
#pragma unroll
      for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int c = 0; c < CO; ++c)
          acc[n][c] = 0.0f;
      if (yin)
        {
          const size_t o0 = (size_t)b * nf * nf + (size_t)(Y - 1) * nf + 4 * p; // pixel (Y - 1, 4 p)
          for (int ci0 = 0; ci0 < Cin; ci0 += CU)
            {
              uint2 lo[CU][2];
              uint16_t hi[CU][2];
#pragma unroll
              for (int u = 0; u < CU; ++u)
#pragma unroll
                for (int ry = 0; ry < 2; ++ry)
                  {
                    const uint16_t *src = in + o0 + (size_t)(ci0 + u) * plane + (size_t)ry * nf;
                    lo[u][ry] = *(const uint2 *)src;
                    hi[u][ry] = last ? (uint16_t)0 : src[4]; // (column nf does not exist)
                  }
#pragma unroll
              for (int u = 0; u < CU; ++u)
                {
                  float v[2][5]; // [row Y - 1 + ry][column 4 p + j]
#pragma unroll
                  for (int ry = 0; ry < 2; ++ry)
                    {
                      v[ry][0] = __uint_as_float(lo[u][ry].x << 16);
                      v[ry][1] = __uint_as_float(lo[u][ry].x & 0xffff0000u);
                      v[ry][2] = __uint_as_float(lo[u][ry].y << 16);
                      v[ry][3] = __uint_as_float(lo[u][ry].y & 0xffff0000u);
                      v[ry][4] = bf16_to_f32(hi[u][ry]);
                    }
#pragma unroll
                  for (int q = 0; q < 4; ++q)
                    {
                      const float4 wk = *(const float4 *)(swk + ((ci0 + u) * 4 + q) * CO);
#pragma unroll
                      for (int n = 0; n < 4; ++n)
                        {
                          const float val = v[q >> 1][n + (q & 1)];
                          acc[n][0] = fmaf(val, wk.x, acc[n][0]);
                          acc[n][1] = fmaf(val, wk.y, acc[n][1]);
                          acc[n][2] = fmaf(val, wk.z, acc[n][2]);
                          acc[n][3] = fmaf(val, wk.w, acc[n][3]);
                        }
                    }
                }
            }
        }
#pragma unroll
      for (int n = 0; n < 4; ++n)
        {
          const int X = X0 + n;
          const bool inner = yin && X < nf;
#pragma unroll
          for (int c = 0; c < CO; ++c)
            {
              double v = nodal_q1(c, 2, X, Y, 0, nf);
              if (inner)
                {
                  float r = acc[n][c];
                  if (bias)
                    {
                      r += bias[c];
                      if (relu && r < 0.0f)
                        r = 0.0f;
                    }
                  v += (double)(scale * r);
                }
              bases[((size_t)b * CO + c) * per + (size_t)Y * np + X] = v;
            }
        }
      if (p == 0)
#pragma unroll
        for (int c = 0; c < CO; ++c)
          bases[((size_t)b * CO + c) * per + (size_t)Y * np] = nodal_q1(c, 2, 0, Y, 0, nf);
    }
}

// wp[co][blk][tap][c] = w[co][(32 blk + c) 9 + tap]: the weights of a 3x3
// convolution regrouped per 32-channel block and tap (conv3x3_bf16_kernel).
__global__ void conv3x3_pack_kernel(int Cout, int Cin, const uint16_t *w, uint16_t *wp)
{
  const int K = Cin * 9;
  const size_t total = (size_t)Cout * K;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int co = (int)(t / K), k = (int)(t % K);
      const int blk = k / 288, r = k % 288, tap = r / 32, c = r % 32;
      wp[t] = w[(size_t)co * K + (blk * 32 + c) * 9 + tap];
    }
}

// C (+)= sum over z of work[z] in fixed order (the split-K partials).
__global__ void splitk_reduce_kernel(int M, int N, int splits, const float *work, float *C, int ldc, int accumulate)
{
  const size_t total = (size_t)M * N;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      float s = 0.0f;
      for (int z = 0; z < splits; ++z)
        s += work[(size_t)z * total + t];
      float *c = C + (t / N) * (size_t)ldc + t % N;
      *c = accumulate ? *c + s : s;
    }
}

// col[(ci*kh + ky)*kw + kx][n = (b*H + y)*W + x] = bf16(in[ci][b][y + ky - pad][x + kx - pad]),
// output spatial size Ho x Wo = (H + 2 pad - kh + 1) x (W + 2 pad - kw + 1).
__global__ void im2col_kernel(int C, int B, int H, int W, int kh, int kw, int pad, const float *in, uint16_t *col)
{
  const int Ho = H + 2 * pad - kh + 1, Wo = W + 2 * pad - kw + 1;
  const size_t N = (size_t)B * Ho * Wo, total = (size_t)C * kh * kw * N;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const size_t n = t % N;
      const int r = (int)(t / N);
      const int ci = r / (kh * kw), ky = (r / kw) % kh, kx = r % kw;
      const int x = (int)(n % Wo), y = (int)((n / Wo) % Ho), b = (int)(n / ((size_t)Wo * Ho));
      const int yy = y + ky - pad, xx = x + kx - pad;
      float v = 0.0f;
      if (yy >= 0 && yy < H && xx >= 0 && xx < W)
        v = in[(((size_t)ci * B + b) * H + yy) * W + xx];
      col[t] = f32_to_bf16(v);
    }
}

// The same, 8 consecutive columns n per thread (one 16-byte store) with 32-bit
// indexing: block row blockIdx.y = col row r = (ci, ky, kx).  When Wo is a
// multiple of 8 the 8 columns share (b, y) and read one input row segment.
// Requires N = B Ho Wo to be a multiple of 8 (the host checks).
__global__ void im2col8_kernel(int B, int H, int W, int kh, int kw, int pad, int Ho, int Wo, const float *in,
                               uint16_t *col)
{
  const int r = blockIdx.y;
  const int ci = r / (kh * kw), ky = (r / kw) % kh, kx = r % kw;
  const int N = B * Ho * Wo;
  const int n0 = 8 * (blockIdx.x * blockDim.x + threadIdx.x);
  if (n0 >= N)
    return;
  bf16x8_t v;
  if (Wo % 8 == 0)
    {
      const int x0 = n0 % Wo, rest = n0 / Wo, y = rest % Ho, b = rest / Ho;
      const int yy = y + ky - pad;
      const bool rowok = yy >= 0 && yy < H;
      const float *src = in + ((size_t)(ci * B + b) * H + (rowok ? yy : 0)) * W;
#pragma unroll
      for (int e = 0; e < 8; ++e)
        {
          const int xx = x0 + e + kx - pad;
          v[e] = (rowok && xx >= 0 && xx < W) ? f32_to_bf16(src[xx]) : (uint16_t)0;
        }
    }
  else
    {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        {
          const int n = n0 + e, x = n % Wo, rest = n / Wo, y = rest % Ho, b = rest / Ho;
          const int yy = y + ky - pad, xx = x + kx - pad;
          v[e] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? f32_to_bf16(in[((size_t)(ci * B + b) * H + yy) * W + xx])
                                                         : (uint16_t)0;
        }
    }
  *(bf16x8_t *)(col + (size_t)r * N + n0) = v;
}

// dx[ci][b][y][x] = sum over (ky, kx) of dcol[(ci*kh+ky)*kw+kx][n(y - ky + pad, x - kx + pad)] (gather: deterministic)
__global__ void col2im_kernel(int C, int B, int H, int W, int kh, int kw, int pad, const float *dcol, float *dx)
{
  const int Ho = H + 2 * pad - kh + 1, Wo = W + 2 * pad - kw + 1;
  const size_t N = (size_t)B * Ho * Wo, total = (size_t)C * B * H * W;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(t % W), y = (int)((t / W) % H), b = (int)((t / ((size_t)W * H)) % B);
      const int ci = (int)(t / ((size_t)W * H * B));
      float s = 0.0f;
      for (int ky = 0; ky < kh; ++ky)
        for (int kx = 0; kx < kw; ++kx)
          {
            const int yo = y - ky + pad, xo = x - kx + pad;
            if (yo < 0 || yo >= Ho || xo < 0 || xo >= Wo)
              continue;
            s += dcol[(size_t)((ci * kh + ky) * kw + kx) * N + ((size_t)b * Ho + yo) * Wo + xo];
          }
      dx[t] = s;
    }
}

__global__ void bias_act_kernel(int C, size_t N, const float *bias, float *y, int relu)
{
  const size_t total = (size_t)C * N;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      float v = y[t] + bias[t / N];
      y[t] = (relu && v < 0.0f) ? 0.0f : v;
    }
}

// dpre = dy * (y > 0) for ReLU layers (in place), bf16 copy for the GEMMs
__global__ void act_bwd_kernel(size_t total, const float *y, float *dy, uint16_t *dy_bf, int relu)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      float g = dy[t];
      if (relu && !(y[t] > 0.0f))
        g = 0.0f;
      dy[t] = g;
      dy_bf[t] = f32_to_bf16(g);
    }
}

// db[c] = sum_n d[c][n], one workgroup per channel, fixed-order reduction
__global__ void __launch_bounds__(256) bias_grad_kernel(size_t N, const float *d, float *db)
{
  __shared__ float red[4];
  const int c = blockIdx.x;
  float s = 0.0f;
  for (size_t n = threadIdx.x; n < N; n += 256)
    s += d[(size_t)c * N + n];
  for (int m = 32; m >= 1; m >>= 1)
    s += __shfl_xor(s, m, 64);
  if (threadIdx.x % 64 == 0)
    red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    db[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void pool2_fwd_kernel(size_t total_out, int Ho, int Wo, const float *in, float *out)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total_out; t += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(t % Wo), y = (int)((t / Wo) % Ho);
      const size_t plane = t / ((size_t)Wo * Ho);
      const float *p = in + plane * (size_t)(4 * Ho * Wo) + (size_t)(2 * y) * (2 * Wo) + 2 * x;
      out[t] = 0.25f * ((p[0] + p[1]) + (p[2 * Wo] + p[2 * Wo + 1]));
    }
}

__global__ void pool2_bwd_kernel(size_t total_in, int Hi, int Wi, const float *dout, float *din)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total_in; t += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(t % Wi), y = (int)((t / Wi) % Hi);
      const size_t plane = t / ((size_t)Wi * Hi);
      din[t] = 0.25f * dout[plane * (size_t)(Hi / 2) * (Wi / 2) + (size_t)(y / 2) * (Wi / 2) + x / 2];
    }
}

__global__ void up2_fwd_kernel(size_t total_out, int Ho, int Wo, const float *in, float *out)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total_out; t += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(t % Wo), y = (int)((t / Wo) % Ho);
      const size_t plane = t / ((size_t)Wo * Ho);
      out[t] = in[plane * (size_t)(Ho / 2) * (Wo / 2) + (size_t)(y / 2) * (Wo / 2) + x / 2];
    }
}

__global__ void up2_bwd_kernel(size_t total_in, int Hi, int Wi, const float *dout, float *din)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total_in; t += (size_t)gridDim.x * blockDim.x)
    {
      const int x = (int)(t % Wi), y = (int)((t / Wi) % Hi);
      const size_t plane = t / ((size_t)Wi * Hi);
      const float *p = dout + plane * (size_t)(4 * Hi * Wi) + (size_t)(2 * y) * (2 * Wi) + 2 * x;
      din[t] = (p[0] + p[1]) + (p[2 * Wi] + p[2 * Wi + 1]);
    }
}

__global__ void axpy_kernel(size_t n, float a, const float *x, float *y)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    y[t] += a * x[t];
}

// out = a x + b y (y is not read when b == 0; out may alias x or y)
__global__ void axpby_kernel(size_t n, float a, const float *x, float b, const float *y, float *out)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    out[t] = b != 0.0f ? a * x[t] + b * y[t] : a * x[t];
}

__global__ void to_bf16_kernel(size_t n, const float *x, uint16_t *y)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    y[t] = f32_to_bf16(x[t]);
}

// loss partials: grad = 2 (pred - target) / count; per-block sum of squares
__global__ void __launch_bounds__(256) mse_kernel(size_t n, const float *pred, const float *target, float inv_count,
                                                  float *grad, float *partial)
{
  __shared__ float red[4];
  float s = 0.0f;
  for (size_t t = blockIdx.x * (size_t)256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256)
    {
      const float d = pred[t] - target[t];
      grad[t] = 2.0f * d * inv_count;
      s += d * d;
    }
  for (int m = 32; m >= 1; m >>= 1)
    s += __shfl_xor(s, m, 64);
  if (threadIdx.x % 64 == 0)
    red[threadIdx.x / 64] = s;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void adam_kernel(size_t n, float *w, const float *g, float *m, float *v, uint16_t *w_bf, float lr,
                            float b1, float b2, float eps, float c1, float c2, float wd)
{
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    {
      const float gt = g[t] + wd * w[t];
      const float mt = b1 * m[t] + (1.0f - b1) * gt;
      const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
      m[t] = mt;
      v[t] = vt;
      const float wt = w[t] - lr * (mt * c1) / (sqrtf(vt * c2) + eps);
      w[t] = wt;
      if (w_bf)
        w_bf[t] = f32_to_bf16(wt);
    }
}

// Bases from the surrogate's interior corrections: phi_i = N_i + delta_i inside,
// N_i on the boundary; delta [4][cells][(nf-1)^2] FP32 -> bases [cell][4][(nf+1)^2] FP64.
__global__ void delta_to_bases_kernel(int n_cells, int nf, const float *delta, double *bases)
{
  const int np = nf + 1, ni = nf - 1;
  const size_t per = (size_t)4 * np * np, total = (size_t)n_cells * per;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int c = (int)(t / per), i = (int)((t / ((size_t)np * np)) % 4), k = (int)(t % ((size_t)np * np));
      const int X = k % np, Y = k / np;
      double v = nodal_q1(i, 2, X, Y, 0, nf);
      if (X > 0 && X < nf && Y > 0 && Y < nf)
        v += (double)delta[((size_t)i * n_cells + c) * ni * ni + (size_t)(Y - 1) * ni + (X - 1)];
      bases[t] = v;
    }
}

// Training targets: delta_i = phi_i - N_i at the interior nodes, FP32 [4][cells][(nf-1)^2].
__global__ void bases_to_delta_kernel(int n_cells, int nf, const double *bases, float *delta)
{
  const int np = nf + 1, ni = nf - 1;
  const size_t total = (size_t)4 * n_cells * ni * ni;
  for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x)
    {
      const int X = 1 + (int)(t % ni), Y = 1 + (int)((t / ni) % ni);
      const int c = (int)((t / ((size_t)ni * ni)) % n_cells), i = (int)(t / ((size_t)ni * ni * n_cells));
      const double phi = bases[((size_t)c * 4 + i) * np * np + (size_t)Y * np + X];
      delta[t] = (float)(phi - nodal_q1(i, 2, X, Y, 0, nf));
    }
}

} // namespace msfem

using namespace msfem;

namespace {

unsigned grid_for(size_t n)
{
  size_t g = (n + 255) / 256;
  if (g > 65535)
    g = 65535;
  return (unsigned)(g ? g : 1);
}

int check_launch()
{
  return hipGetLastError() == hipSuccess ? MSFEM_OK : MSFEM_ERR_HIP;
}

// Launches a persistent kernel of 256 threads on min(ntiles, resident
// workgroups of the whole device) workgroups.
template <typename... KArgs, typename... Args>
hipError_t launch_tiles(void (*kernel)(KArgs...), int ntiles, hipStream_t st, Args... args)
{
  int dev = 0, ncu = 0, per_cu = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0);
  if (e != hipSuccess)
    return e;
  const long long slots = (long long)(ncu > 0 ? ncu : 1) * (per_cu > 0 ? per_cu : 1);
  const int grid = (int)(ntiles < slots ? ntiles : slots);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(256), 0, st, args...);
  return hipGetLastError();
}

} // namespace

extern "C" {

int msfem_nn_gemm_bf16(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                       float *C, int ldc, int accumulate, void *stream)
{
  return msfem_nn_gemm_bf16_splitk(transA, transB, M, N, K, A, lda, B, ldb, C, ldc, accumulate, 1, nullptr, stream);
}

int msfem_nn_gemm_bf16_splitk(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B,
                              int ldb, float *C, int ldc, int accumulate, int splits, float *work, void *stream)
{
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C || splits < 1 || (splits > 1 && !work))
    return MSFEM_ERR_ARG;
  if (lda < (transA ? M : K) || ldb < (transB ? K : N) || ldc < N)
    return MSFEM_ERR_ARG;
  int kchunk = K;
  if (splits > 1)
    {
      kchunk = ((K + splits - 1) / splits + kGBK - 1) / kGBK * kGBK;
      splits = (K + kchunk - 1) / kchunk;
    }
  const int BM = M <= 32 ? 32 : (M <= 64 ? 64 : 128);
  dim3 grid((N + kGBN - 1) / kGBN, (M + BM - 1) / BM, splits), block(256);
  if (grid.y > 65535 || splits > 65535)
    return MSFEM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *a = (const uint16_t *)A, *b = (const uint16_t *)B;
  const bool vec = lda % 8 == 0 && ldb % 8 == 0 && ((uintptr_t)a % 16) == 0 && ((uintptr_t)b % 16) == 0;
  float *w = splits > 1 ? work : nullptr;
  const int sel = (transA ? 1 : 0) | (transB ? 2 : 0) | (vec ? 4 : 0);
#define MSFEM_GEMM_CASE(BMv)                                                                                           \
  switch (sel)                                                                                                       \
    {                                                                                                                \
      case 0: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, false, false>), grid, block, 0, st, M, N, K, a, lda, b, \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 1: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, true, false, false>), grid, block, 0, st, M, N, K, a, lda, b,  \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 2: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, true, false>), grid, block, 0, st, M, N, K, a, lda, b,  \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 3: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, true, true, false>), grid, block, 0, st, M, N, K, a, lda, b,   \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 4: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, false, true>), grid, block, 0, st, M, N, K, a, lda, b,  \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 5: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, true, false, true>), grid, block, 0, st, M, N, K, a, lda, b,   \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      case 6: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, true, true>), grid, block, 0, st, M, N, K, a, lda, b,   \
                                 ldb, C, ldc, accumulate, kchunk, w); break;                                         \
      default: hipLaunchKernelGGL((gemm_bf16_kernel<BMv, true, true, true>), grid, block, 0, st, M, N, K, a, lda, b,   \
                                  ldb, C, ldc, accumulate, kchunk, w); break;                                        \
    }
  if (BM == 32)
    {
      MSFEM_GEMM_CASE(32)
    }
  else if (BM == 64)
    {
      MSFEM_GEMM_CASE(64)
    }
  else
    {
      MSFEM_GEMM_CASE(128)
    }
#undef MSFEM_GEMM_CASE
  if (splits > 1)
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for((size_t)M * N)), dim3(256), 0, st, M, N, splits, work, C, ldc,
                       accumulate);
  return check_launch();
}

int msfem_nn_im2col(int C, int B, int H, int W, int kh, int kw, int pad, const float *in, void *col, void *stream)
{
  if (C <= 0 || B <= 0 || H <= 0 || W <= 0 || kh <= 0 || kw <= 0 || pad < 0 || !in || !col)
    return MSFEM_ERR_ARG;
  const int Ho = H + 2 * pad - kh + 1, Wo = W + 2 * pad - kw + 1;
  if (Ho <= 0 || Wo <= 0)
    return MSFEM_ERR_ARG;
  const size_t N = (size_t)B * Ho * Wo, rows = (size_t)C * kh * kw;
  if (N % 8 == 0 && N < (1u << 30) && rows <= 65535 && ((uintptr_t)col % 16) == 0)
    {
      const dim3 grid((unsigned)((N / 8 + 255) / 256), (unsigned)rows);
      hipLaunchKernelGGL(im2col8_kernel, grid, dim3(256), 0, (hipStream_t)stream, B, H, W, kh, kw, pad, Ho, Wo, in,
                         (uint16_t *)col);
    }
  else
    hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(rows * N)), dim3(256), 0, (hipStream_t)stream, C, B, H, W, kh, kw,
                       pad, in, (uint16_t *)col);
  return check_launch();
}

int msfem_nn_conv_bf16(int Cin, int B, int H, int W, int kh, int kw, int pad, const void *w, int Cout, const float *in,
                       float *out, int splits, float *work, const float *bias, int relu, void *stream)
{
  if (Cin <= 0 || B <= 0 || H <= 0 || W <= 0 || kh <= 0 || kw <= 0 || pad < 0 || Cout <= 0 || !w || !in || !out ||
      splits < 1 || (splits > 1 && !work))
    return MSFEM_ERR_ARG;
  const int Ho = H + 2 * pad - kh + 1, Wo = W + 2 * pad - kw + 1;
  const long long Nl = (long long)B * Ho * Wo, Kl = (long long)Cin * kh * kw;
  if (Ho <= 0 || Wo <= 0 || Wo % 8 != 0 || Nl >= (1ll << 31) || Kl >= (1ll << 31) ||
      (long long)Cin * B * H * W >= (1ll << 31))
    return MSFEM_ERR_ARG;
  const int N = (int)Nl, K = (int)Kl;
  int kchunk = K;
  if (splits > 1)
    {
      kchunk = ((K + splits - 1) / splits + kGBK - 1) / kGBK * kGBK;
      splits = (K + kchunk - 1) / kchunk;
    }
  const int BM = Cout <= 32 ? 32 : (Cout <= 64 ? 64 : 128);
  dim3 grid((N + kGBN - 1) / kGBN, (Cout + BM - 1) / BM, splits), block(256);
  if (grid.y > 65535 || splits > 65535)
    return MSFEM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *a = (const uint16_t *)w;
  const bool vec = K % 8 == 0 && ((uintptr_t)a % 16) == 0;
  float *wk = splits > 1 ? work : nullptr;
  const ConvGeo cg{in, B, H, W, kh, kw, pad, Ho, Wo, bias, relu};
#define MSFEM_CONV_CASE(BMv)                                                                                           \
  if (vec)                                                                                                           \
    hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, false, true, true>), grid, block, 0, st, Cout, N, K, a, K,       \
                       (const uint16_t *)nullptr, N, out, N, 0, kchunk, wk, cg);                                      \
  else                                                                                                               \
    hipLaunchKernelGGL((gemm_bf16_kernel<BMv, false, false, false, true>), grid, block, 0, st, Cout, N, K, a, K,      \
                       (const uint16_t *)nullptr, N, out, N, 0, kchunk, wk, cg);
  if (BM == 32)
    {
      MSFEM_CONV_CASE(32)
    }
  else if (BM == 64)
    {
      MSFEM_CONV_CASE(64)
    }
  else
    {
      MSFEM_CONV_CASE(128)
    }
#undef MSFEM_CONV_CASE
  if (splits > 1)
    {
      hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for((size_t)Cout * N)), dim3(256), 0, st, Cout, N, splits,
                         work, out, N, 0);
      if (bias)
        hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for((size_t)Cout * N)), dim3(256), 0, st, Cout, (uint64_t)N, bias,
                           out, relu);
    }
  return check_launch();
}

int msfem_nn_conv3x3_pack(int Cin, int Cout, const void *w, void *wp, void *stream)
{
  if (Cin <= 0 || Cin % 32 != 0 || Cout <= 0 || !w || !wp || (long long)Cout * Cin * 9 >= (1ll << 31))
    return MSFEM_ERR_ARG;
  const size_t n = (size_t)Cout * Cin * 9;
  hipLaunchKernelGGL(conv3x3_pack_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, Cout, Cin,
                     (const uint16_t *)w, (uint16_t *)wp);
  return check_launch();
}

int msfem_nn_conv3x3_bf16(int Cin, int B, int H, int W, const void *wp, int Cout, const float *in, int upsample,
                          float *out, const float *bias, int relu, const float *add, void *stream)
{
  return msfem_nn_conv3x3_bf16_ex(Cin, B, H, W, wp, Cout, in, upsample, out, bias, relu, add, nullptr, 0, stream);
}

int msfem_nn_conv3x3_bf16_ex(int Cin, int B, int H, int W, const void *wp, int Cout, const void *in, int upsample,
                             void *out, const float *bias, int relu, const float *add, float *pool, int io,
                             void *stream)
{
  const int ialign = (upsample ? 8 : 16) / ((io & 1) ? 2 : 1); // one staging load: 2 or 4 columns
  if (Cin <= 0 || Cin % 32 != 0 || B <= 0 || H <= 0 || Cout <= 0 || Cout % 32 != 0 || !wp || !in || !out ||
      (W != 4 && W != 8 && W != 16 && W != 32 && W != 64) || (W <= 8 ? H != W : H % (kGBN / W) != 0) ||
      ((uintptr_t)wp % 16) != 0 || ((uintptr_t)in % ialign) != 0 || (upsample && H % 2 != 0) || io < 0 || io > 3)
    return MSFEM_ERR_ARG;
  const long long plane = (long long)B * H * W;
  if (plane % kGBN != 0 || (long long)Cin * plane >= (1ll << 31) || (long long)Cout * plane >= (1ll << 31) ||
      (long long)Cout * Cin * 9 >= (1ll << 31))
    return MSFEM_ERR_ARG;
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *w = (const uint16_t *)wp;
  const int ntiles = (int)(plane / kGBN);
  if (pool && (H % 2 != 0 || (long long)Cout * (plane / 4) >= (1ll << 31)))
    return MSFEM_ERR_ARG;
  if (Cout == 32 && Cin == 32 && (!pool || W == 64) && io == 0) // persistent form: weights staged once per workgroup
    {
      // (with two input blocks the resident weights cost a third of the
      // occupancy, and the tile form is faster)
      const int nb = 1;
      hipError_t e = hipSuccess;
#define MSFEM_C3P_CASE(Wv, UPv, NBv)                                                                                  \
  if (W == Wv && (upsample != 0) == UPv && nb == NBv)                                                                \
    e = launch_tiles(conv3x3_tiles_kernel<Wv, UPv, NBv, false>, ntiles, st, B, H, ntiles, (const float *)in, w,    \
                     (float *)out, bias, relu, add, pool, (const uint16_t *)nullptr, (const float *)nullptr, 0);
#define MSFEM_C3P_W(Wv)                                                                                               \
  MSFEM_C3P_CASE(Wv, false, 1)                                                                                       \
  MSFEM_C3P_CASE(Wv, true, 1)
      MSFEM_C3P_W(4)
      MSFEM_C3P_W(8)
      MSFEM_C3P_W(16)
      MSFEM_C3P_W(32)
      MSFEM_C3P_W(64)
#undef MSFEM_C3P_W
#undef MSFEM_C3P_CASE
      return e == hipSuccess ? check_launch() : MSFEM_ERR_HIP;
    }
  const int BM = Cout % 64 == 0 ? 64 : 32;
  const dim3 grid((unsigned)ntiles, (unsigned)(Cout / BM)), block(256);
#define MSFEM_C3_LAUNCH(BMv, Wv, UPv, IBv, OBv)                                                                       \
  hipLaunchKernelGGL((conv3x3_bf16_kernel<BMv, Wv, UPv, IBv, OBv>), grid, block, 0, st, Cin, B, H, in, w, out, bias,  \
                     relu, add, pool)
#define MSFEM_C3_UP(BMv, Wv, UPv)                                                                                     \
  switch (io)                                                                                                        \
    {                                                                                                                \
    case 0:                                                                                                          \
      MSFEM_C3_LAUNCH(BMv, Wv, UPv, false, false);                                                                   \
      break;                                                                                                         \
    case 1:                                                                                                          \
      MSFEM_C3_LAUNCH(BMv, Wv, UPv, true, false);                                                                    \
      break;                                                                                                         \
    case 2:                                                                                                          \
      MSFEM_C3_LAUNCH(BMv, Wv, UPv, false, true);                                                                    \
      break;                                                                                                         \
    default:                                                                                                         \
      MSFEM_C3_LAUNCH(BMv, Wv, UPv, true, true);                                                                     \
    }
#define MSFEM_C3_CASE(BMv, Wv)                                                                                         \
  if (BM == BMv && W == Wv)                                                                                          \
    {                                                                                                                \
      if (upsample)                                                                                                  \
        {                                                                                                            \
          MSFEM_C3_UP(BMv, Wv, true)                                                                                 \
        }                                                                                                            \
      else                                                                                                           \
        {                                                                                                            \
          MSFEM_C3_UP(BMv, Wv, false)                                                                                \
        }                                                                                                            \
    }
  MSFEM_C3_CASE(32, 4)
  MSFEM_C3_CASE(32, 8)
  MSFEM_C3_CASE(32, 16)
  MSFEM_C3_CASE(32, 32)
  MSFEM_C3_CASE(32, 64)
  MSFEM_C3_CASE(64, 4)
  MSFEM_C3_CASE(64, 8)
  MSFEM_C3_CASE(64, 16)
  MSFEM_C3_CASE(64, 32)
  MSFEM_C3_CASE(64, 64)
#undef MSFEM_C3_CASE
#undef MSFEM_C3_UP
#undef MSFEM_C3_LAUNCH
  return check_launch();
}

int msfem_nn_conv3x3_first_bf16(int B, int H, int W, const void *w0, const float *b0, int relu0, const float *in,
                                const void *wp, int Cout, float *out, const float *bias, int relu, float *pool,
                                void *stream)
{
  if (B <= 0 || H <= 0 || W != 64 || H % 2 != 0 || Cout != 32 || !w0 || !in || !wp || !out ||
      ((uintptr_t)in % 16) != 0 || ((uintptr_t)wp % 16) != 0)
    return MSFEM_ERR_ARG;
  const long long plane = (long long)B * H * W;
  if ((long long)Cout * plane >= (1ll << 31))
    return MSFEM_ERR_ARG;
  const int ntiles = (int)(plane / kGBN);
  const hipError_t e = launch_tiles(conv3x3_tiles_kernel<64, false, 1, true>, ntiles, (hipStream_t)stream, B, H,
                                    ntiles, in, (const uint16_t *)wp, out, bias, relu, (const float *)nullptr, pool,
                                    (const uint16_t *)w0, b0, relu0);
  return e == hipSuccess ? check_launch() : MSFEM_ERR_HIP;
}

int msfem_nn_conv_small(int Cin, int B, int H, int W, int kh, int kw, int pad, const void *w, int Cout,
                        const float *in, float *out, const float *bias, int relu, void *stream)
{
  if (Cin <= 0 || B <= 0 || H <= 0 || W <= 0 || kh != kw || kh < 1 || kh > 3 || pad < 0 || Cout <= 0 ||
      Cout > 8 || !w || !in || !out)
    return MSFEM_ERR_ARG;
  const int Ho = H + 2 * pad - kh + 1, Wo = W + 2 * pad - kw + 1;
  const int CO = Cout <= 4 ? 4 : 8;
  const long long K = (long long)Cin * kh * kw;
  if (Ho <= 0 || Wo <= 0 || K * CO * 4 > 65536 || (long long)Cin * B * H * W >= (1ll << 31) ||
      (long long)Cout * B * Ho * Wo >= (1ll << 31))
    return MSFEM_ERR_ARG;
  const size_t N = (size_t)B * Ho * Wo, lds = (size_t)K * CO * sizeof(float);
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *wb = (const uint16_t *)w;
#define MSFEM_CS_CASE(c, k)                                                                                           \
  if (CO == c && kh == k)                                                                                            \
    hipLaunchKernelGGL((conv_small_kernel<c, k>), dim3(grid_for(N)), dim3(256), lds, st, Cin, B, H, W, pad, Cout, in,  \
                       wb, out, bias, relu);
  MSFEM_CS_CASE(4, 1)
  MSFEM_CS_CASE(4, 2)
  MSFEM_CS_CASE(4, 3)
  MSFEM_CS_CASE(8, 1)
  MSFEM_CS_CASE(8, 2)
  MSFEM_CS_CASE(8, 3)
#undef MSFEM_CS_CASE
  return check_launch();
}

int msfem_nn_head_bases(int Cin, int B, int nf, const void *w, const void *in, int in_bf16, const float *bias,
                        int relu, float scale, double *bases, void *stream)
{
  if (Cin <= 0 || B <= 0 || nf < 2 || !w || !in || !bases || (long long)Cin * 4 * 4 * 4 > 65536 ||
      (long long)Cin * B * nf * nf >= (1ll << 31) || (long long)B * 4 * (nf + 1) * (nf + 1) >= (1ll << 40))
    return MSFEM_ERR_ARG;
  const size_t total = (size_t)B * (nf + 1) * (nf + 1), lds = (size_t)Cin * 4 * 4 * sizeof(float);
  hipStream_t st = (hipStream_t)stream;
  // (4 channels per load batch: 2048 tiles of 32 channels at 64^2 in 325 us
  // against 370 with 8 and 485 with 16, whose registers cut the occupancy)
  if (in_bf16 && nf % 4 == 0 && Cin % 4 == 0 && ((uintptr_t)in & 7) == 0)
    hipLaunchKernelGGL(head_bases4_kernel<4>, dim3(grid_for((size_t)B * (nf + 1) * (nf / 4))), dim3(256), lds, st,
                       Cin, B, nf, (const uint16_t *)in, (const uint16_t *)w, bias, relu, scale, bases);
  else if (in_bf16)
    hipLaunchKernelGGL(head_bases_kernel<true>, dim3(grid_for(total)), dim3(256), lds, st, Cin, B, nf, in,
                       (const uint16_t *)w, bias, relu, scale, bases);
  else
    hipLaunchKernelGGL(head_bases_kernel<false>, dim3(grid_for(total)), dim3(256), lds, st, Cin, B, nf, in,
                       (const uint16_t *)w, bias, relu, scale, bases);
  return check_launch();
}

int msfem_nn_col2im(int C, int B, int H, int W, int kh, int kw, int pad, const float *dcol, float *dx, void *stream)
{
  if (C <= 0 || B <= 0 || H <= 0 || W <= 0 || !dcol || !dx)
    return MSFEM_ERR_ARG;
  const size_t n = (size_t)C * B * H * W;
  hipLaunchKernelGGL(col2im_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, C, B, H, W, kh, kw, pad,
                     dcol, dx);
  return check_launch();
}

int msfem_nn_bias_act(int C, uint64_t N, const float *bias, float *y, int relu, void *stream)
{
  if (C <= 0 || !bias || !y)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(bias_act_kernel, dim3(grid_for((size_t)C * N)), dim3(256), 0, (hipStream_t)stream, C, (size_t)N,
                     bias, y, relu);
  return check_launch();
}

int msfem_nn_act_bwd(uint64_t n, const float *y, float *dy, void *dy_bf16, int relu, void *stream)
{
  if (!y || !dy || !dy_bf16)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(act_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, y, dy,
                     (uint16_t *)dy_bf16, relu);
  return check_launch();
}

int msfem_nn_bias_grad(int C, uint64_t N, const float *d, float *db, void *stream)
{
  if (C <= 0 || !d || !db)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(bias_grad_kernel, dim3(C), dim3(256), 0, (hipStream_t)stream, (size_t)N, d, db);
  return check_launch();
}

int msfem_nn_pool2(int planes, int H, int W, const float *in, float *out, int backward, void *stream)
{
  if (planes <= 0 || H <= 0 || W <= 0 || (H | W) & 1 || !in || !out)
    return MSFEM_ERR_ARG;
  if (!backward)
    {
      const size_t n = (size_t)planes * (H / 2) * (W / 2);
      hipLaunchKernelGGL(pool2_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, H / 2, W / 2, in,
                         out);
    }
  else // in = d(out) at H/2 x W/2, out = d(in) at H x W
    {
      const size_t n = (size_t)planes * H * W;
      hipLaunchKernelGGL(pool2_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, H, W, in, out);
    }
  return check_launch();
}

int msfem_nn_up2(int planes, int H, int W, const float *in, float *out, int backward, void *stream)
{
  if (planes <= 0 || H <= 0 || W <= 0 || !in || !out)
    return MSFEM_ERR_ARG;
  if (!backward) // in H x W -> out 2H x 2W
    {
      const size_t n = (size_t)planes * 4 * H * W;
      hipLaunchKernelGGL(up2_fwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, 2 * H, 2 * W, in,
                         out);
    }
  else // in = d(out) at 2H x 2W, out = d(in) at H x W
    {
      const size_t n = (size_t)planes * H * W;
      hipLaunchKernelGGL(up2_bwd_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n, H, W, in, out);
    }
  return check_launch();
}

int msfem_nn_axpy(uint64_t n, float a, const float *x, float *y, void *stream)
{
  if (!x || !y)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, a, x, y);
  return check_launch();
}

int msfem_nn_axpby(uint64_t n, float a, const float *x, float b, const float *y, float *out, void *stream)
{
  if (!x || !out || (b != 0.0f && !y))
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, a, x, b, y, out);
  return check_launch();
}

int msfem_nn_to_bf16(uint64_t n, const float *x, void *y, void *stream)
{
  if (!x || !y)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, x,
                     (uint16_t *)y);
  return check_launch();
}

int msfem_nn_mse(uint64_t n, const float *pred, const float *target, float *grad, float *partial, int n_partial,
                 void *stream)
{
  if (!pred || !target || !grad || !partial || n_partial <= 0 || n == 0)
    return MSFEM_ERR_ARG;
  hipLaunchKernelGGL(mse_kernel, dim3(n_partial), dim3(256), 0, (hipStream_t)stream, (size_t)n, pred, target,
                     1.0f / (float)n, grad, partial);
  return check_launch();
}

int msfem_nn_adam(uint64_t n, float *w, const float *g, float *m, float *v, void *w_bf16, float lr, float beta1,
                  float beta2, float eps, int step, float weight_decay, void *stream)
{
  if (!w || !g || !m || !v || step < 1)
    return MSFEM_ERR_ARG;
  const float c1 = 1.0f / (1.0f - powf(beta1, (float)step)), c2 = 1.0f / (1.0f - powf(beta2, (float)step));
  hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, (size_t)n, w, g, m, v,
                     (uint16_t *)w_bf16, lr, beta1, beta2, eps, c1, c2, weight_decay);
  return check_launch();
}

int msfem_nn_delta_to_bases(int n_cells, int nf, const float *delta, double *bases, void *stream)
{
  if (n_cells <= 0 || nf < 2 || !delta || !bases)
    return MSFEM_ERR_ARG;
  const size_t n = (size_t)n_cells * 4 * (nf + 1) * (nf + 1);
  hipLaunchKernelGGL(delta_to_bases_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n_cells, nf, delta,
                     bases);
  return check_launch();
}

int msfem_nn_bases_to_delta(int n_cells, int nf, const double *bases, float *delta, void *stream)
{
  if (n_cells <= 0 || nf < 2 || !delta || !bases)
    return MSFEM_ERR_ARG;
  const size_t n = (size_t)n_cells * 4 * (nf - 1) * (nf - 1);
  hipLaunchKernelGGL(bases_to_delta_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, n_cells, nf, bases,
                     delta);
  return check_launch();
}

} // extern "C"
