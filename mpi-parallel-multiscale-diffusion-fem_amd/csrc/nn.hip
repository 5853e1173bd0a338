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
// (TB: [N][K]), all row-major BF16; C FP32 row-major.  64x64 tile per
// workgroup, four waves of one 32x32 v_mfma_f32_32x32x8_bf16 block each;
// K staged through LDS 32 at a time, zero-filled beyond the matrix edges.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_bf16_kernel(int M, int N, int K, const uint16_t *A, int lda,
                                                        const uint16_t *B, int ldb, float *C, int ldc, int accumulate)
{
  __shared__ uint16_t sA[64][40]; // [m][k], padded rows
  __shared__ uint16_t sB[64][40]; // [n][k]
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  f32x16 acc;
#pragma unroll
  for (int j = 0; j < 16; ++j)
    acc[j] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += 32)
    {
      for (int e = tid; e < 64 * 32; e += 256)
        {
          int mm, kk;
          if (TA)
            {
              kk = e / 64;
              mm = e % 64;
            }
          else
            {
              mm = e / 32;
              kk = e % 32;
            }
          const int gm = m0 + mm, gk = k0 + kk;
          uint16_t v = 0;
          if (gm < M && gk < K)
            v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
          sA[mm][kk] = v;
        }
      for (int e = tid; e < 64 * 32; e += 256)
        {
          int nn, kk;
          if (TB)
            {
              nn = e / 32;
              kk = e % 32;
            }
          else
            {
              kk = e / 64;
              nn = e % 64;
            }
          const int gn = n0 + nn, gk = k0 + kk;
          uint16_t v = 0;
          if (gn < N && gk < K)
            v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
          sB[nn][kk] = v;
        }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < 32; ks += 8)
        {
          bf16x4 a, b;
          const int kq = ks + (lane / 32) * 4;
#pragma unroll
          for (int v = 0; v < 4; ++v)
            {
              a[v] = (short)sA[wm + lane % 32][kq + v];
              b[v] = (short)sB[wn + lane % 32][kq + v];
            }
          acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
      __syncthreads();
    }
#pragma unroll
  for (int j = 0; j < 16; ++j)
    {
      const int row = wm + 8 * (j / 4) + 4 * (lane / 32) + (j % 4), col = wn + lane % 32;
      const int gm = m0 + row, gn = n0 + col;
      if (gm < M && gn < N)
        {
          float *c = C + (size_t)gm * ldc + gn;
          *c = accumulate ? *c + acc[j] : acc[j];
        }
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

} // namespace

extern "C" {

int msfem_nn_gemm_bf16(int transA, int transB, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                       float *C, int ldc, int accumulate, void *stream)
{
  if (M <= 0 || N <= 0 || K <= 0 || !A || !B || !C)
    return MSFEM_ERR_ARG;
  dim3 grid((N + 63) / 64, (M + 63) / 64), block(256);
  hipStream_t st = (hipStream_t)stream;
  const uint16_t *a = (const uint16_t *)A, *b = (const uint16_t *)B;
  if (!transA && !transB)
    hipLaunchKernelGGL((gemm_bf16_kernel<false, false>), grid, block, 0, st, M, N, K, a, lda, b, ldb, C, ldc, accumulate);
  else if (!transA && transB)
    hipLaunchKernelGGL((gemm_bf16_kernel<false, true>), grid, block, 0, st, M, N, K, a, lda, b, ldb, C, ldc, accumulate);
  else if (transA && !transB)
    hipLaunchKernelGGL((gemm_bf16_kernel<true, false>), grid, block, 0, st, M, N, K, a, lda, b, ldb, C, ldc, accumulate);
  else
    hipLaunchKernelGGL((gemm_bf16_kernel<true, true>), grid, block, 0, st, M, N, K, a, lda, b, ldb, C, ldc, accumulate);
  return check_launch();
}

int msfem_nn_im2col(int C, int B, int H, int W, int kh, int kw, int pad, const float *in, void *col, void *stream)
{
  if (C <= 0 || B <= 0 || H <= 0 || W <= 0 || kh <= 0 || kw <= 0 || pad < 0 || !in || !col)
    return MSFEM_ERR_ARG;
  const size_t n = (size_t)C * kh * kw * B * (H + 2 * pad - kh + 1) * (W + 2 * pad - kw + 1);
  hipLaunchKernelGGL(im2col_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, C, B, H, W, kh, kw, pad, in,
                     (uint16_t *)col);
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
