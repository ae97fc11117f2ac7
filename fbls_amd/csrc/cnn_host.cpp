// Host driver of the band-proposal network: parameters, activations and the
// forward / backward / AdamW sequence over the kernels of cnn_kernel.hip.
//
// Network: `layers` Conv1d layers (kernel `kernel`, `width` channels, zero
// padding, GELU) over a periodogram of L period bins x in_channels duration
// groups, then a per-bin linear head (width -> 1) whose sigmoid is the
// probability that the full scan finds its detection in that bin.  Loss:
// binary cross-entropy with positive weight pos_weight, mean over bins.
// Parameters are FP32 masters with BF16 copies for the MFMA GEMMs;
// activations enter the GEMMs as BF16 and accumulate in FP32.
//
// Flat parameter order: for each layer l, W_l[cout][k][cin] then b_l[cout];
// then the head w[width] and b[1].
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/fbls.h"
#include "cnn_kernel.h"
#include "host_internal.h"

using namespace fbls::host;
namespace cn = fbls::cnn;

struct fbls_cnn {
  fbls_ctx* c = nullptr;
  fbls_cnn_cfg cfg{};
  std::vector<size_t> w_off, b_off, k_in;  // per layer
  size_t head_w = 0, head_b = 0, count = 0;
  DevBuf P, PB, G, Mo, Vo;                 // params, bf16 copy, grads, Adam moments
  int step = 0;
  // activations for the current B x L
  long long rows = 0;
  std::vector<DevBuf> Xcol, Y, H;
  DevBuf X0, z, y, terms, dz, dHa, dHb, dY, dYb, dXcol, part, scalar;
  DevBuf Ha, Hb;  // inference activations (ping-pong)
  int fused = 1;  // inference through conv_gelu when the shapes allow it
  ~fbls_cnn() {
    for (DevBuf* b : {&P, &PB, &G, &Mo, &Vo, &X0, &z, &y, &terms, &dz, &dHa, &dHb, &dY, &dYb,
                      &dXcol, &part, &scalar, &Ha, &Hb})
      b->release();
    for (auto* v : {&Xcol, &Y, &H})
      for (DevBuf& b : *v) b.release();
  }
};

namespace {

uint64_t next(uint64_t* s) {
  uint64_t x = *s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  *s = x;
  return x;
}
double uniform(uint64_t* s) { return (double)(next(s) >> 11) * (1.0 / 9007199254740992.0); }

int ensure_acts(fbls_cnn* n, long long rows) {
  if (rows <= n->rows) return FBLS_OK;
  const int Lr = (int)n->cfg.layers, W = (int)n->cfg.width;
  n->Xcol.resize(Lr);
  n->Y.resize(Lr);
  n->H.resize(Lr);
  size_t kmax = 0;
  for (int l = 0; l < Lr; l++) kmax = std::max(kmax, n->k_in[l]);
  int st;
  for (int l = 0; l < Lr; l++) {
    if ((st = n->Xcol[l].ensure(rows * n->k_in[l] * 2)) || (st = n->Y[l].ensure(rows * W * 4)) ||
        (st = n->H[l].ensure(rows * W * 2)))
      return st;
  }
  if ((st = n->X0.ensure(rows * n->cfg.in_channels * 4 + rows * n->cfg.in_channels * 2)) ||
      (st = n->z.ensure(rows * 4)) || (st = n->y.ensure(rows * 4)) ||
      (st = n->terms.ensure(rows * 4)) || (st = n->dz.ensure(rows * 4)) ||
      (st = n->dHa.ensure(rows * W * 4)) || (st = n->dHb.ensure(rows * W * 4)) ||
      (st = n->dY.ensure(rows * W * 4)) || (st = n->dYb.ensure(rows * W * 2)) ||
      (st = n->dXcol.ensure(rows * kmax * 4)) || (st = n->scalar.ensure(64)))
    return st;
  n->rows = rows;
  return FBLS_OK;
}

constexpr int kSplitChunk = 2048;  // rows of K per split of the weight-gradient GEMM

// Inference: every layer as one fused implicit-GEMM conv + bias + GELU
// launch (no im2col matrix, no fp32 pre-activations), then the head.  x is
// host input, or null when the input is already staged in X0.
int infer(fbls_cnn* n, const float* x, uint32_t B, uint32_t L) {
  const long long rows = (long long)B * L;
  const int Cin = (int)n->cfg.in_channels, W = (int)n->cfg.width, kw = (int)n->cfg.kernel;
  int st;
  if ((st = n->X0.ensure(rows * Cin * 6)) || (st = n->z.ensure(rows * 4)) ||
      (st = n->Ha.ensure(rows * W * 2)) || (st = n->Hb.ensure(rows * W * 2)))
    return st;
  hipStream_t s = n->c->stream;
  float* x32 = (float*)n->X0.ptr;
  cn::bf16* x16 = (cn::bf16*)((unsigned char*)n->X0.ptr + rows * Cin * 4);
  if (x) HIP_TRY(hipMemcpyAsync(x32, x, rows * Cin * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(cn::to_bf16(x32, x16, rows * Cin, s));
  const float* P = (const float*)n->P.ptr;
  const cn::bf16* PB = (const cn::bf16*)n->PB.ptr;
  const cn::bf16* in = x16;
  int C = Cin;
  for (uint32_t l = 0; l < n->cfg.layers; l++) {
    cn::ConvArgs a;
    a.X = in;
    a.W = PB + n->w_off[l];
    a.bias = P + n->b_off[l];
    a.H = (cn::bf16*)((l & 1) ? n->Hb.ptr : n->Ha.ptr);
    a.B = (int)B;
    a.L = (int)L;
    a.cin = C;
    a.kw = kw;
    HIP_TRY(cn::conv_gelu(a, s));
    in = a.H;
    C = W;
  }
  HIP_TRY(cn::head_forward(in, P + n->head_w, P + n->head_b, (float*)n->z.ptr, rows, W, s));
  return FBLS_OK;
}

// The fused path needs 64 channels, 8 or 64 inputs per layer and aligned
// weight rows.
bool fusable(const fbls_cnn* n) {
  if (n->cfg.width != 64) return false;
  if (n->cfg.in_channels != 8 && n->cfg.in_channels != 64) return false;
  for (uint32_t l = 0; l < n->cfg.layers; l++)
    if (n->w_off[l] % 8) return false;
  return true;
}

// forward of B x L bins; x host fp32 (B, L, in_channels)
int forward(fbls_cnn* n, const float* x, uint32_t B, uint32_t L) {
  const long long rows = (long long)B * L;
  if (int st = ensure_acts(n, rows)) return st;
  hipStream_t s = n->c->stream;
  const int Cin = (int)n->cfg.in_channels, W = (int)n->cfg.width, kw = (int)n->cfg.kernel;
  float* x32 = (float*)n->X0.ptr;
  cn::bf16* x16 = (cn::bf16*)((unsigned char*)n->X0.ptr + rows * Cin * 4);
  HIP_TRY(hipMemcpyAsync(x32, x, rows * Cin * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(cn::to_bf16(x32, x16, rows * Cin, s));
  const float* P = (const float*)n->P.ptr;
  const cn::bf16* PB = (const cn::bf16*)n->PB.ptr;
  const cn::bf16* in = x16;
  int C = Cin;
  for (uint32_t l = 0; l < n->cfg.layers; l++) {
    cn::bf16* xcol = (cn::bf16*)n->Xcol[l].ptr;
    HIP_TRY(cn::im2col(in, xcol, (int)B, (int)L, C, kw, s));
    cn::GemmArgs g;
    g.A = xcol;
    g.B = PB + n->w_off[l];
    g.C = (float*)n->Y[l].ptr;
    g.M = (int)rows;
    g.N = W;
    g.K = (int)n->k_in[l];
    g.kc = ((g.K + 31) / 32) * 32;
    g.sam = g.K;
    g.sak = 1;
    g.sbk = 1;
    g.sbn = g.K;
    HIP_TRY(cn::gemm(g, 1, s));
    HIP_TRY(cn::bias_gelu((float*)n->Y[l].ptr, P + n->b_off[l], (cn::bf16*)n->H[l].ptr, rows, W, s));
    in = (const cn::bf16*)n->H[l].ptr;
    C = W;
  }
  HIP_TRY(cn::head_forward(in, P + n->head_w, P + n->head_b, (float*)n->z.ptr, rows, W, s));
  return FBLS_OK;
}

// loss and gradients (into n->G) of the last forward
int backward(fbls_cnn* n, const float* y, uint32_t B, uint32_t L, float* loss) {
  const long long rows = (long long)B * L;
  hipStream_t s = n->c->stream;
  const int Cin = (int)n->cfg.in_channels, W = (int)n->cfg.width, kw = (int)n->cfg.kernel;
  const int Lr = (int)n->cfg.layers;
  float* G = (float*)n->G.ptr;
  const float* P = (const float*)n->P.ptr;
  const cn::bf16* PB = (const cn::bf16*)n->PB.ptr;
  HIP_TRY(hipMemcpyAsync(n->y.ptr, y, rows * 4, hipMemcpyHostToDevice, s));
  HIP_TRY(cn::head_loss((const float*)n->z.ptr, (const float*)n->y.ptr, n->cfg.pos_weight,
                        (float*)n->terms.ptr, (float*)n->dz.ptr, rows, s));
  HIP_TRY(cn::sum_f32((const float*)n->terms.ptr, (float*)n->scalar.ptr, rows, s));
  float* dH = (float*)n->dHa.ptr;
  float* dHnext = (float*)n->dHb.ptr;
  HIP_TRY(cn::head_backward((const cn::bf16*)n->H[Lr - 1].ptr, P + n->head_w,
                            (const float*)n->dz.ptr, dH, G + n->head_w, G + n->head_b, rows, W, s));
  const int splits = (int)((rows + kSplitChunk - 1) / kSplitChunk);
  for (int l = Lr - 1; l >= 0; l--) {
    const int K = (int)n->k_in[l];
    const int C = l == 0 ? Cin : W;
    HIP_TRY(cn::gelu_back(dH, (const float*)n->Y[l].ptr, (float*)n->dY.ptr,
                          (cn::bf16*)n->dYb.ptr, rows * W, s));
    HIP_TRY(cn::col_sums((const float*)n->dY.ptr, G + n->b_off[l], rows, W, s));
    // dW = dY^T Xcol: (W x rows) (rows x K), split over rows
    if (int st = n->part.ensure((size_t)splits * W * K * 4)) return st;
    cn::GemmArgs g;
    g.A = (const cn::bf16*)n->dYb.ptr;
    g.B = (const cn::bf16*)n->Xcol[l].ptr;
    g.C = (float*)n->part.ptr;
    g.M = W;
    g.N = K;
    g.K = (int)rows;
    g.kc = kSplitChunk;
    g.sam = 1;
    g.sak = W;
    g.sbk = K;
    g.sbn = 1;
    HIP_TRY(cn::gemm(g, splits, s));
    HIP_TRY(cn::reduce_splits((const float*)n->part.ptr, G + n->w_off[l], splits,
                              (long long)W * K, s));
    if (l > 0) {
      // dXcol = dY W: (rows x W) (W x K)
      cn::GemmArgs h;
      h.A = (const cn::bf16*)n->dYb.ptr;
      h.B = PB + n->w_off[l];
      h.C = (float*)n->dXcol.ptr;
      h.M = (int)rows;
      h.N = K;
      h.K = W;
      h.kc = ((W + 31) / 32) * 32;
      h.sam = W;
      h.sak = 1;
      h.sbk = K;
      h.sbn = 1;
      HIP_TRY(cn::gemm(h, 1, s));
      HIP_TRY(cn::col2im((const float*)n->dXcol.ptr, dHnext, (int)B, (int)L, C, kw, s));
      std::swap(dH, dHnext);
    }
  }
  float total = 0.f;
  HIP_TRY(hipMemcpyAsync(&total, n->scalar.ptr, 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (loss) *loss = total / (float)rows;
  return FBLS_OK;
}

int check_batch(const fbls_cnn* n, const float* x, uint32_t B, uint32_t L) {
  if (!n || !x) return fail(FBLS_E_ARG, "null argument");
  if (B == 0 || L == 0 || (uint64_t)B * L > (1ull << 24))
    return fail(FBLS_E_ARG, "batch of %u x %u bins out of range", B, L);
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_cnn_create(fbls_ctx* c, const fbls_cnn_cfg* cfg, uint64_t seed,
                               fbls_cnn** out) {
  return fbls::host::guard([&]() -> int {
    if (!c || !cfg || !out) return fail(FBLS_E_ARG, "null argument");
    *out = nullptr;
    if (cfg->in_channels < 1 || cfg->in_channels > 256 || cfg->width < 1 || cfg->width > 256 ||
        cfg->layers < 1 || cfg->layers > 32 || cfg->kernel < 1 || cfg->kernel % 2 == 0 ||
        cfg->kernel > 31)
      return fail(FBLS_E_ARG, "network shape out of range");
    HIP_TRY(hipSetDevice(c->device));
    fbls_cnn* n = new fbls_cnn();
    n->c = c;
    n->cfg = *cfg;
    size_t off = 0;
    for (uint32_t l = 0; l < cfg->layers; l++) {
      const size_t cin = l == 0 ? cfg->in_channels : cfg->width;
      n->k_in.push_back(cfg->kernel * cin);
      n->w_off.push_back(off);
      off += (size_t)cfg->width * cfg->kernel * cin;
      n->b_off.push_back(off);
      off += cfg->width;
    }
    n->head_w = off;
    off += cfg->width;
    n->head_b = off;
    off += 1;
    n->count = off;
    std::vector<float> p(off, 0.f);
    uint64_t s = seed ? seed : 0x9e3779b97f4a7c15ull;
    for (uint32_t l = 0; l < cfg->layers; l++) {
      const double bound = std::sqrt(6.0 / (double)n->k_in[l]);
      for (size_t i = 0; i < (size_t)cfg->width * n->k_in[l]; i++)
        p[n->w_off[l] + i] = (float)((2.0 * uniform(&s) - 1.0) * bound);
    }
    const double hb = std::sqrt(3.0 / (double)cfg->width);
    for (uint32_t i = 0; i < cfg->width; i++) p[n->head_w + i] = (float)((2.0 * uniform(&s) - 1.0) * hb);
    p[n->head_b] = -3.0f;  // sparse positives: start near sigmoid(-3)
    int st;
    if ((st = n->P.ensure(off * 4)) || (st = n->PB.ensure(off * 2)) || (st = n->G.ensure(off * 4)) ||
        (st = n->Mo.ensure(off * 4)) || (st = n->Vo.ensure(off * 4))) {
      delete n;
      return st;
    }
    if (fbls_cnn_set_params(n, p.data()) != FBLS_OK) {
      delete n;
      return FBLS_E_HIP;
    }
    *out = n;
    return FBLS_OK;
  });
}

extern "C" void fbls_cnn_destroy(fbls_cnn* n) {
  if (!n) return;
  (void)hipSetDevice(n->c->device);
  delete n;
}

extern "C" size_t fbls_cnn_param_count(const fbls_cnn* n) { return n ? n->count : 0; }

extern "C" int fbls_cnn_get_params(fbls_cnn* n, float* out) {
  return fbls::host::guard([&]() -> int {
    if (!n || !out) return fail(FBLS_E_ARG, "null argument");
    HIP_TRY(hipMemcpy(out, n->P.ptr, n->count * 4, hipMemcpyDeviceToHost));
    return FBLS_OK;
  });
}

extern "C" int fbls_cnn_set_params(fbls_cnn* n, const float* in) {
  return fbls::host::guard([&]() -> int {
    if (!n || !in) return fail(FBLS_E_ARG, "null argument");
    hipStream_t s = n->c->stream;
    HIP_TRY(hipMemcpyAsync(n->P.ptr, in, n->count * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(cn::to_bf16((const float*)n->P.ptr, (cn::bf16*)n->PB.ptr, (long long)n->count, s));
    std::vector<float> zero(n->count, 0.f);
    HIP_TRY(hipMemcpyAsync(n->Mo.ptr, zero.data(), n->count * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(n->Vo.ptr, zero.data(), n->count * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    n->step = 0;
    return FBLS_OK;
  });
}

extern "C" int fbls_cnn_forward(fbls_cnn* n, const float* x, uint32_t B, uint32_t L,
                                float* logits) {
  return fbls::host::guard([&]() -> int {
    if (int st = check_batch(n, x, B, L)) return st;
    if (!logits) return fail(FBLS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(n->c->device));
    if (int st = (n->fused && fusable(n)) ? infer(n, x, B, L) : forward(n, x, B, L)) return st;
    HIP_TRY(hipMemcpyAsync(logits, n->z.ptr, (size_t)B * L * 4, hipMemcpyDeviceToHost, n->c->stream));
    HIP_TRY(hipStreamSynchronize(n->c->stream));
    return FBLS_OK;
  });
}

extern "C" int fbls_cnn_loss_grad(fbls_cnn* n, const float* x, const float* y, uint32_t B,
                                  uint32_t L, float* loss, float* grads) {
  return fbls::host::guard([&]() -> int {
    if (int st = check_batch(n, x, B, L)) return st;
    if (!y) return fail(FBLS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(n->c->device));
    if (int st = forward(n, x, B, L)) return st;
    if (int st = backward(n, y, B, L, loss)) return st;
    if (grads) HIP_TRY(hipMemcpy(grads, n->G.ptr, n->count * 4, hipMemcpyDeviceToHost));
    return FBLS_OK;
  });
}

extern "C" int fbls_cnn_train_step(fbls_cnn* n, const float* x, const float* y, uint32_t B,
                                   uint32_t L, float lr, float weight_decay, float* loss) {
  return fbls::host::guard([&]() -> int {
    if (int st = check_batch(n, x, B, L)) return st;
    if (!y) return fail(FBLS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(n->c->device));
    if (int st = forward(n, x, B, L)) return st;
    if (int st = backward(n, y, B, L, loss)) return st;
    n->step++;
    hipStream_t s = n->c->stream;
    float* P = (float*)n->P.ptr;
    cn::bf16* PB = (cn::bf16*)n->PB.ptr;
    const float* G = (const float*)n->G.ptr;
    float* M = (float*)n->Mo.ptr;
    float* V = (float*)n->Vo.ptr;
    // decoupled weight decay on the weights, none on the biases
    for (uint32_t l = 0; l < n->cfg.layers; l++) {
      const size_t wo = n->w_off[l], wn = n->b_off[l] - wo, bo = n->b_off[l];
      HIP_TRY(cn::adamw(P + wo, PB + wo, G + wo, M + wo, V + wo, wn, lr, 0.9f, 0.999f, 1e-8f,
                        weight_decay, n->step, s));
      HIP_TRY(cn::adamw(P + bo, PB + bo, G + bo, M + bo, V + bo, n->cfg.width, lr, 0.9f, 0.999f,
                        1e-8f, 0.0f, n->step, s));
    }
    HIP_TRY(cn::adamw(P + n->head_w, PB + n->head_w, G + n->head_w, M + n->head_w, V + n->head_w,
                      n->cfg.width, lr, 0.9f, 0.999f, 1e-8f, weight_decay, n->step, s));
    HIP_TRY(cn::adamw(P + n->head_b, PB + n->head_b, G + n->head_b, M + n->head_b, V + n->head_b,
                      1, lr, 0.9f, 0.999f, 1e-8f, 0.0f, n->step, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FBLS_OK;
  });
}

extern "C" int fbls_cnn_set_fused(fbls_cnn* n, int enable) {
  return fbls::host::guard([&]() -> int {
    if (!n) return fail(FBLS_E_ARG, "null argument");
    n->fused = enable ? 1 : 0;
    return FBLS_OK;
  });
}

namespace fbls {
namespace host {

int cnn_stage_input(fbls_cnn* n, uint32_t B, uint32_t L, float** device_input) {
  const long long rows = (long long)B * L;
  if (B == 0 || L == 0 || (uint64_t)rows > (1ull << 24))
    return fail(FBLS_E_ARG, "batch of %u x %u bins out of range", B, L);
  HIP_TRY(hipSetDevice(n->c->device));
  if (int st = n->X0.ensure(rows * n->cfg.in_channels * 6)) return st;
  *device_input = (float*)n->X0.ptr;
  return FBLS_OK;
}

int cnn_run_staged(fbls_cnn* n, uint32_t B, uint32_t L, float* logits) {
  if (!(n->fused && fusable(n)))
    return fail(FBLS_E_ARG, "staged inference needs the fused network shape");
  if (int st = infer(n, nullptr, B, L)) return st;
  HIP_TRY(hipMemcpyAsync(logits, n->z.ptr, (size_t)B * L * 4, hipMemcpyDeviceToHost,
                         n->c->stream));
  HIP_TRY(hipStreamSynchronize(n->c->stream));
  return FBLS_OK;
}

fbls_ctx* cnn_context(const fbls_cnn* n) { return n->c; }
uint32_t cnn_in_channels(const fbls_cnn* n) { return n->cfg.in_channels; }

}  // namespace host
}  // namespace fbls
