// Band-proposal network kernels for gfx950: a BF16 MFMA GEMM with FP32
// accumulation (v_mfma_f32_16x16x32_bf16) under im2col Conv1d layers in
// channels-last layout, bias + GELU epilogues, the per-bin sigmoid head
// with a weighted binary cross-entropy, and AdamW.  Every reduction runs in
// a fixed order, so a training step is deterministic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cnn_kernel.h"

namespace fbls {
namespace cnn {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;

__device__ __forceinline__ float bf2f(bf16 x) { return __uint_as_float((uint32_t)x << 16); }
__device__ __forceinline__ bf16 f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16)(u >> 16);
}

// 256 threads = 4 waves; the block computes a 64 x 64 tile of C, each wave
// a 32 x 32 quarter as 2 x 2 MFMA 16x16x32 blocks.  Operand layout of the
// MFMA: A lane l holds row l % 16 and eight k of group l / 16, B lane l holds
// column l % 16 and the same eight k; D lane l holds column l % 16 and rows
// 4 (l / 16) .. + 3.
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[kBM][kBK + kPad];
  __shared__ __attribute__((aligned(16))) bf16 Bs[kBN][kBK + kPad];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * kBM, n0 = blockIdx.y * kBN;
  const int kb = blockIdx.z * g.kc;
  const int ke = min(g.K, kb + g.kc);
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  const bool a_m_fast = g.sam == 1, b_n_fast = g.sbn == 1;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    for (int e = tid; e < kBM * kBK; e += 256) {
      const int r = a_m_fast ? e % kBM : e / kBK;
      const int kk = a_m_fast ? e / kBM : e % kBK;
      const int m = m0 + r, k = k0 + kk;
      As[r][kk] = (m < g.M && k < ke) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : (bf16)0;
    }
    for (int e = tid; e < kBN * kBK; e += 256) {
      const int c = b_n_fast ? e % kBN : e / kBK;
      const int kk = b_n_fast ? e / kBN : e % kBK;
      const int n = n0 + c, k = k0 + kk;
      Bs[c][kk] = (n < g.N && k < ke) ? g.B[(long long)k * g.sbk + (long long)n * g.sbn] : (bf16)0;
    }
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    bf16x8 a[2], b[2];
    for (int i = 0; i < 2; i++) a[i] = *(const bf16x8*)&As[wm + 16 * i + (lane & 15)][kq];
    for (int j = 0; j < 2; j++) b[j] = *(const bf16x8*)&Bs[wn + 16 * j + (lane & 15)][kq];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  float* C = g.C + (long long)blockIdx.z * g.M * g.N;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int r = 0; r < 4; r++) {
        const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = n0 + wn + 16 * j + (lane & 15);
        if (m < g.M && n < g.N) C[(long long)m * g.N + n] = acc[i][j][r];
      }
}

__device__ __forceinline__ float gelu(float x) {
  return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// Conv1d + bias + GELU for inference, as one implicit GEMM: a workgroup
// owns 128 consecutive bins of one sample and all 64 output channels.  The
// input rows the 128 bins need (with the kw - 1 halo, zero outside the
// sample) and the layer's bf16 weights W[n][k] (k = tap * CIN + c) are
// staged in LDS; the MFMA A operand of output row r and k group kq..kq+7 is
// X[r + tap][c..c+7] (eight channels of one tap, CIN % 8 == 0), so no
// im2col matrix is formed.  The k steps, operand values and MFMA order are
// those of gemm_kernel over the im2col matrix, so H is bit-identical to the
// im2col + gemm + bias_gelu path.
constexpr int kConvRows = 128, kConvN = 64;

template <int CIN>
__global__ __launch_bounds__(256) void conv_gelu_kernel(ConvArgs a) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  constexpr int XS = CIN + 8;  // padded LDS row of the input, in bf16
  const int kw = a.kw, half = kw / 2, K = kw * CIN, KP = (K + 31) / 32 * 32;
  const int WS = KP + 8;       // padded LDS row of the weights, in bf16
  bf16* Xs = (bf16*)smem;                                   // [kConvRows + kw - 1][XS]
  bf16* Ws = Xs + (size_t)(kConvRows + kw - 1) * XS;        // [kConvN][WS]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tiles = (a.L + kConvRows - 1) / kConvRows;
  const int b = blockIdx.x / tiles, l0 = (blockIdx.x % tiles) * kConvRows;
  // weights: 16-byte pieces, zero beyond K
  for (int e = tid; e < kConvN * (KP / 8); e += 256) {
    const int n = e / (KP / 8), k8 = (e % (KP / 8)) * 8;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (k8 < K) v = *(const uint4*)(a.W + (size_t)n * K + k8);
    *(uint4*)(Ws + (size_t)n * WS + k8) = v;
  }
  // input rows l0 - half .. l0 + kConvRows + half - 1 of sample b
  const int prow = kConvRows + kw - 1;
  for (int e = tid; e < prow * (CIN / 8); e += 256) {
    const int p = e / (CIN / 8), c8 = (e % (CIN / 8)) * 8;
    const int l = l0 - half + p;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (l >= 0 && l < a.L) v = *(const uint4*)(a.X + ((size_t)b * a.L + l) * CIN + c8);
    *(uint4*)(Xs + (size_t)p * XS + c8) = v;
  }
  __syncthreads();
  const int wm = wave * 32;  // this wave's 32 rows, all 64 columns
  f32x4 acc[2][4];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kq = 8 * (lane >> 4), r16 = lane & 15;
  for (int k0 = 0; k0 < KP; k0 += 32) {
    const int kk = k0 + kq;
    bf16x8 av[2], bv[4];
    if (kk < K) {
      const int tap = kk / CIN, c = kk % CIN;
      for (int i = 0; i < 2; i++)
        av[i] = *(const bf16x8*)(Xs + (size_t)(wm + 16 * i + r16 + tap) * XS + c);
    } else {
      for (int i = 0; i < 2; i++) {
        const uint4 z = make_uint4(0, 0, 0, 0);
        av[i] = *(const bf16x8*)&z;
      }
    }
    for (int j = 0; j < 4; j++) bv[j] = *(const bf16x8*)(Ws + (size_t)(16 * j + r16) * WS + kk);
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 4; j++)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
  }
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 4; j++)
      for (int r = 0; r < 4; r++) {
        const int l = l0 + wm + 16 * i + 4 * (lane >> 4) + r;
        const int n = 16 * j + r16;
        if (l < a.L) {
          const float y = acc[i][j][r] + a.bias[n];
          a.H[((size_t)b * a.L + l) * kConvN + n] = f2bf(gelu(y));
        }
      }
}

__global__ void reduce_kernel(const float* part, float* out, int splits, long long mn) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float s = 0.f;
  for (int z = 0; z < splits; z++) s += part[(long long)z * mn + i];
  out[i] = s;
}

__global__ void im2col_kernel(const bf16* X, bf16* Xcol, int B, int L, int C, int kw) {
  const long long kc = (long long)kw * C;
  const long long total = (long long)B * L * kc;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long row = i / kc;
  const int j = (int)(i % kc), k = j / C, c = j % C;
  const int b = (int)(row / L), l = (int)(row % L);
  const int src = l + k - kw / 2;
  Xcol[i] = (src >= 0 && src < L) ? X[((long long)b * L + src) * C + c] : (bf16)0;
}

__global__ void col2im_kernel(const float* dXcol, float* dX, int B, int L, int C, int kw) {
  const long long total = (long long)B * L * C;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long long row = i / C;
  const int c = (int)(i % C);
  const int b = (int)(row / L), l = (int)(row % L);
  const long long kc = (long long)kw * C;
  float s = 0.f;
  for (int k = 0; k < kw; k++) {
    const int r = l - k + kw / 2;
    if (r >= 0 && r < L) s += dXcol[((long long)b * L + r) * kc + (long long)k * C + c];
  }
  dX[i] = s;
}

__global__ void bias_gelu_kernel(float* Y, const float* bias, bf16* H, long long M, int N) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * N) return;
  const float y = Y[i] + bias[i % N];
  Y[i] = y;
  H[i] = f2bf(gelu(y));
}

__global__ void gelu_back_kernel(const float* dH, const float* Y, float* dY, bf16* dYb,
                                 long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = dH[i] * gelu_grad(Y[i]);
  dY[i] = d;
  dYb[i] = f2bf(d);
}

// fixed-order block reduction of 256 values
__device__ float block_reduce(float x, float* red) {
  red[threadIdx.x] = x;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(256) void col_sums_kernel(const float* dY, float* db, long long M,
                                                       int N) {
  __shared__ float red[256];
  const int n = blockIdx.x;
  float s = 0.f;
  for (long long m = threadIdx.x; m < M; m += 256) s += dY[m * N + n];
  const float t = block_reduce(s, red);
  if (threadIdx.x == 0) db[n] = t;
}

__global__ void head_forward_kernel(const bf16* H, const float* w, const float* b, float* z,
                                    long long M, int C) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  float s = 0.f;
  for (int c = 0; c < C; c++) s += bf2f(H[m * C + c]) * w[c];
  z[m] = s + b[0];
}

__device__ __forceinline__ float softplus(float x) {
  return fmaxf(x, 0.f) + log1pf(__expf(-fabsf(x)));
}
__device__ __forceinline__ float sigmoid(float x) { return 1.0f / (1.0f + __expf(-x)); }

__global__ void head_loss_kernel(const float* z, const float* y, float pw, float* terms,
                                 float* dz, long long M) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const float zz = z[m], yy = y[m];
  terms[m] = pw * yy * softplus(-zz) + (1.0f - yy) * softplus(zz);
  dz[m] = (-pw * yy * sigmoid(-zz) + (1.0f - yy) * sigmoid(zz)) / (float)M;
}

__global__ void head_dh_kernel(const float* w, const float* dz, float* dH, long long M, int C) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * C) return;
  dH[i] = dz[i / C] * w[i % C];
}

__global__ __launch_bounds__(256) void head_dw_kernel(const bf16* H, const float* dz, float* dw,
                                                      float* db, long long M, int C) {
  __shared__ float red[256];
  const int c = blockIdx.x;  // c == C: the bias
  float s = 0.f;
  for (long long m = threadIdx.x; m < M; m += 256)
    s += c < C ? dz[m] * bf2f(H[m * C + c]) : dz[m];
  const float t = block_reduce(s, red);
  if (threadIdx.x == 0) {
    if (c < C)
      dw[c] = t;
    else
      db[0] = t;
  }
}

__global__ void adamw_kernel(float* p, bf16* pb, const float* g, float* m, float* v,
                             long long n, float lr, float b1, float b2, float eps, float wd,
                             float c1, float c2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / c1, vh = vi / c2;
  const float pi = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
  p[i] = pi;
  if (pb) pb[i] = f2bf(pi);
}

__global__ void to_bf16_kernel(const float* x, bf16* y, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

__global__ __launch_bounds__(256) void sum_kernel(const float* x, float* out, long long n) {
  __shared__ float red[256];
  float s = 0.f;
  for (long long i = threadIdx.x; i < n; i += 256) s += x[i];
  const float t = block_reduce(s, red);
  if (threadIdx.x == 0) out[0] = t;
}

inline unsigned blocks(long long n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t gemm(const GemmArgs& g, int splits, hipStream_t s) {
  dim3 grid((g.M + kBM - 1) / kBM, (g.N + kBN - 1) / kBN, splits);
  hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t reduce_splits(const float* part, float* out, int splits, long long mn, hipStream_t s) {
  hipLaunchKernelGGL(reduce_kernel, dim3(blocks(mn)), dim3(256), 0, s, part, out, splits, mn);
  return hipGetLastError();
}

size_t conv_gelu_lds(int cin, int kw) {
  const int K = kw * cin, KP = (K + 31) / 32 * 32;
  return ((size_t)(kConvRows + kw - 1) * (cin + 8) + (size_t)kConvN * (KP + 8)) * 2;
}

hipError_t conv_gelu(const ConvArgs& a, hipStream_t s) {
  const int tiles = (a.L + kConvRows - 1) / kConvRows;
  const size_t lds = conv_gelu_lds(a.cin, a.kw);
  const void* fn = a.cin == 8 ? (const void*)conv_gelu_kernel<8>
                              : (const void*)conv_gelu_kernel<64>;
  if (a.cin != 8 && a.cin != 64) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (a.cin == 8)
    hipLaunchKernelGGL(conv_gelu_kernel<8>, dim3(a.B * tiles), dim3(256), lds, s, a);
  else
    hipLaunchKernelGGL(conv_gelu_kernel<64>, dim3(a.B * tiles), dim3(256), lds, s, a);
  return hipGetLastError();
}

hipError_t im2col(const bf16* X, bf16* Xcol, int B, int L, int C, int kw, hipStream_t s) {
  const long long n = (long long)B * L * kw * C;
  hipLaunchKernelGGL(im2col_kernel, dim3(blocks(n)), dim3(256), 0, s, X, Xcol, B, L, C, kw);
  return hipGetLastError();
}

hipError_t col2im(const float* dXcol, float* dX, int B, int L, int C, int kw, hipStream_t s) {
  const long long n = (long long)B * L * C;
  hipLaunchKernelGGL(col2im_kernel, dim3(blocks(n)), dim3(256), 0, s, dXcol, dX, B, L, C, kw);
  return hipGetLastError();
}

hipError_t bias_gelu(float* Y, const float* bias, bf16* H, long long M, int N, hipStream_t s) {
  hipLaunchKernelGGL(bias_gelu_kernel, dim3(blocks(M * N)), dim3(256), 0, s, Y, bias, H, M, N);
  return hipGetLastError();
}

hipError_t gelu_back(const float* dH, const float* Y, float* dY, bf16* dYb, long long n,
                     hipStream_t s) {
  hipLaunchKernelGGL(gelu_back_kernel, dim3(blocks(n)), dim3(256), 0, s, dH, Y, dY, dYb, n);
  return hipGetLastError();
}

hipError_t col_sums(const float* dY, float* db, long long M, int N, hipStream_t s) {
  hipLaunchKernelGGL(col_sums_kernel, dim3(N), dim3(256), 0, s, dY, db, M, N);
  return hipGetLastError();
}

hipError_t head_forward(const bf16* H, const float* w, const float* b, float* z, long long M,
                        int C, hipStream_t s) {
  hipLaunchKernelGGL(head_forward_kernel, dim3(blocks(M)), dim3(256), 0, s, H, w, b, z, M, C);
  return hipGetLastError();
}

hipError_t head_loss(const float* z, const float* y, float pos_weight, float* loss_terms,
                     float* dz, long long M, hipStream_t s) {
  hipLaunchKernelGGL(head_loss_kernel, dim3(blocks(M)), dim3(256), 0, s, z, y, pos_weight,
                     loss_terms, dz, M);
  return hipGetLastError();
}

hipError_t head_backward(const bf16* H, const float* w, const float* dz, float* dH, float* dw,
                         float* db, long long M, int C, hipStream_t s) {
  hipLaunchKernelGGL(head_dh_kernel, dim3(blocks(M * C)), dim3(256), 0, s, w, dz, dH, M, C);
  hipLaunchKernelGGL(head_dw_kernel, dim3(C + 1), dim3(256), 0, s, H, dz, dw, db, M, C);
  return hipGetLastError();
}

hipError_t adamw(float* p, bf16* pb, const float* g, float* m, float* v, long long n, float lr,
                 float beta1, float beta2, float eps, float wd, int step, hipStream_t s) {
  const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(adamw_kernel, dim3(blocks(n)), dim3(256), 0, s, p, pb, g, m, v, n, lr,
                     beta1, beta2, eps, wd, c1, c2);
  return hipGetLastError();
}

hipError_t to_bf16(const float* x, bf16* y, long long n, hipStream_t s) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks(n)), dim3(256), 0, s, x, y, n);
  return hipGetLastError();
}

hipError_t sum_f32(const float* x, float* out, long long n, hipStream_t s) {
  hipLaunchKernelGGL(sum_kernel, dim3(1), dim3(256), 0, s, x, out, n);
  return hipGetLastError();
}

}  // namespace cnn
}  // namespace fbls
