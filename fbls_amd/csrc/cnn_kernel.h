// Launch interface of the band-proposal network kernels (cnn_kernel.hip):
// BF16 MFMA GEMM with FP32 accumulation, im2col / col2im for Conv1d in
// channels-last layout, bias + GELU epilogues, the sigmoid head with a
// weighted binary cross-entropy, and AdamW.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbls {
namespace cnn {

typedef uint16_t bf16;  // storage of bfloat16 (upper half of an fp32)

// C[z][m][n] = sum over k in chunk z of A(m, k) * B(k, n) with
// A(m, k) = A[m * sam + k * sak] and B(k, n) = B[k * sbk + n * sbn];
// chunks of kc (a multiple of 32) split K over gridDim.z.
struct GemmArgs {
  const bf16* A;
  const bf16* B;
  float* C;  // splits x M x N
  int M, N, K, kc;
  long long sam, sak, sbk, sbn;
};
hipError_t gemm(const GemmArgs& g, int splits, hipStream_t s);
// out[m][n] = sum over z of part[z][m][n], in z order (deterministic)
hipError_t reduce_splits(const float* part, float* out, int splits, long long mn,
                         hipStream_t s);

// Inference Conv1d (channels-last, zero padding, 64 output channels) with
// bias and GELU fused: H[b][l][n] = bf16(GELU(bias[n] + sum over (k, c) of
// W[n][k cin + c] X[b][l + k - kw/2][c])), bit-identical to im2col + gemm +
// bias_gelu.  cin is 8 or 64.
struct ConvArgs {
  const bf16* X;     // B x L x cin
  const bf16* W;     // 64 x (kw cin)
  const float* bias; // 64
  bf16* H;           // B x L x 64
  int B, L, cin, kw;
};
size_t conv_gelu_lds(int cin, int kw);
hipError_t conv_gelu(const ConvArgs& a, hipStream_t s);

// Xcol[(b L + l)][k C + c] = X[b][l + k - kw/2][c] (0 outside), bf16
hipError_t im2col(const bf16* X, bf16* Xcol, int B, int L, int C, int kw, hipStream_t s);
// dX[b][l][c] = sum over k of dXcol[(b L + l - k + kw/2)][k C + c] (fp32)
hipError_t col2im(const float* dXcol, float* dX, int B, int L, int C, int kw, hipStream_t s);

// Y = Ypre + bias (in place, fp32), H = bf16(GELU(Y))
hipError_t bias_gelu(float* Y, const float* bias, bf16* H, long long M, int N, hipStream_t s);
// dY = dH * GELU'(Y) (fp32), dYb = bf16(dY)
hipError_t gelu_back(const float* dH, const float* Y, float* dY, bf16* dYb, long long n,
                     hipStream_t s);
// column sums of dY (M x N) in row order: db[n] = sum_m dY[m][n]
hipError_t col_sums(const float* dY, float* db, long long M, int N, hipStream_t s);

// z[m] = sum_c H[m][c] w[c] + b; loss terms and dz of the weighted BCE:
// loss_m = pw y softplus(-z) + (1 - y) softplus(z), dz = d loss_m / dz / M
hipError_t head_forward(const bf16* H, const float* w, const float* b, float* z,
                        long long M, int C, hipStream_t s);
hipError_t head_loss(const float* z, const float* y, float pos_weight, float* loss_terms,
                     float* dz, long long M, hipStream_t s);
// dH[m][c] = dz[m] w[c]; dw[c] = sum_m dz[m] H[m][c]; db = sum_m dz[m]
hipError_t head_backward(const bf16* H, const float* w, const float* dz, float* dH,
                         float* dw, float* db, long long M, int C, hipStream_t s);

// AdamW on fp32 master weights p (and their bf16 copy pb, may be null)
hipError_t adamw(float* p, bf16* pb, const float* g, float* m, float* v, long long n,
                 float lr, float beta1, float beta2, float eps, float wd, int step,
                 hipStream_t s);
hipError_t to_bf16(const float* x, bf16* y, long long n, hipStream_t s);
// deterministic sum of n floats into *out (one workgroup)
hipError_t sum_f32(const float* x, float* out, long long n, hipStream_t s);

}  // namespace cnn
}  // namespace fbls
