// Internal host-side types shared by the C ABI sources (not installed).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/fbls.h"

namespace fbls {
namespace host {

// Records the thread's last error message and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Runs an entry point's body: a C++ exception (host memory exhausted, a
// container's size limit) becomes an error status instead of unwinding into
// the C caller's frames.
template <class Body>
int guard(Body&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(FBLS_E_OOM, "host memory allocation failed");
  } catch (const std::exception& e) {
    return fail(FBLS_E_ARG, "%s", e.what());
  }
}
double now_s();
// work(a, e) over [0, m) split into contiguous ranges on up to 16 host
// threads (about `per` items per thread at least); ranges whose thread
// cannot start run on the caller's thread.  work must not throw.
void parallel_ranges(size_t m, size_t per, const std::function<void(size_t, size_t)>& work);
// The network's device input buffer for a B x L batch (channels-last
// float32, in_channels per bin), and the fused inference from it into host
// logits (cnn_host.cpp).
int cnn_stage_input(fbls_cnn* net, uint32_t B, uint32_t L, float** device_input);
int cnn_run_staged(fbls_cnn* net, uint32_t B, uint32_t L, float* logits);
fbls_ctx* cnn_context(const fbls_cnn* net);
uint32_t cnn_in_channels(const fbls_cnn* net);
// The biweight detrend of sorted times on the context's device
// (raw_host.cpp): out[i] = (1 + rel[i]) / location_i - 1.
int biweight_gpu(fbls_ctx* c, const std::vector<double>& t, const std::vector<double>& rel,
                 double window, double gap, double cst, int iterations, std::vector<double>* out);

// Growable device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return FBLS_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 4096);
    if (hipMalloc(&ptr, want) != hipSuccess)
      return fail(FBLS_E_OOM, "device allocation of %zu bytes failed", want);
    cap = want;
    return FBLS_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// Growable pinned host buffer (staging of the per-launch tables/results).
struct HostBuf {
  void* ptr = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return FBLS_OK;
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes + bytes / 2, 1 << 16);
    if (hipHostMalloc(&ptr, want, hipHostMallocDefault) != hipSuccess)
      return fail(FBLS_E_OOM, "pinned host allocation of %zu bytes failed", want);
    cap = want;
    return FBLS_OK;
  }
  void release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    cap = 0;
  }
};

// Device buffers of released uploads, kept for the next ones (a hipFree
// waits for the whole device, which would stall a pipelined batch).  Shared
// by the context and every arena it made; the last holder frees them.
class ArenaPool {
 public:
  explicit ArenaPool(int device) : device_(device) {}
  ~ArenaPool() {
    (void)hipSetDevice(device_);
    for (DevBuf& b : free_) b.release();
  }
  // A buffer of at least `bytes` (the smallest kept one that fits, else a
  // new allocation).
  int take(size_t bytes, DevBuf* out) {
    {
      std::lock_guard<std::mutex> hold(mu_);
      size_t best = free_.size();
      for (size_t k = 0; k < free_.size(); k++)
        if (free_[k].cap >= bytes && (best == free_.size() || free_[k].cap < free_[best].cap))
          best = k;
      if (best < free_.size()) {
        *out = free_[best];
        free_.erase(free_.begin() + (ptrdiff_t)best);
        return FBLS_OK;
      }
    }
    (void)hipSetDevice(device_);
    return out->ensure(bytes);
  }
  // Keeps a released buffer; beyond kKeep buffers the smallest is freed.
  void give(DevBuf b) {
    if (!b.ptr) return;
    std::lock_guard<std::mutex> hold(mu_);
    free_.push_back(b);
    if (free_.size() > kKeep) {
      size_t small = 0;
      for (size_t k = 1; k < free_.size(); k++)
        if (free_[k].cap < free_[small].cap) small = k;
      (void)hipSetDevice(device_);
      free_[small].release();
      free_.erase(free_.begin() + (ptrdiff_t)small);
    }
  }

 private:
  static constexpr size_t kKeep = 4;
  int device_;
  std::mutex mu_;
  std::vector<DevBuf> free_;
};

// Device memory shared by the families of one upload (one family, or a
// packed batch of stars); returned to the context's pool with the last
// family that uses it.
struct DevArena {
  std::shared_ptr<ArenaPool> pool;
  DevBuf buf;
  ~DevArena() {
    if (pool) pool->give(buf);
    else buf.release();
  }
};

constexpr size_t kMaxLds = 160 * 1024;  // gfx950 LDS per CU
constexpr uint32_t kObsAlign = 512;     // observations per wave block

}  // namespace host
}  // namespace fbls

#define HIP_TRY(expr)                                                     \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess)                                                 \
      return fbls::host::fail(FBLS_E_HIP, "%s failed: %s", #expr,         \
                              hipGetErrorString(_e));                     \
  } while (0)

struct fbls_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, evm = nullptr, ev1 = nullptr;
  uint32_t cus = 256;       // compute units (persistent launch grids)
  size_t rows_hint = 4096;  // rows copied back with a search's counters
  uint32_t lds_override = 0, block_override = 0;
  size_t max_lds = fbls::host::kMaxLds;
  uint32_t run_fold = 1;
  uint32_t count_from_weights = 1;  // equal weights: counts = W / w (no count atomics)
  fbls_launch_totals totals = {0, 0, 0, 0.0};
  int use_graphs = 1;                  // evidence calls replay a hipGraph
  hipGraphExec_t ev_exec = nullptr;    // the evidence call's executable graph
  uint64_t graph_launches = 0;
  fbls::host::DevBuf dev_in, dev_out, dev_ev;
  fbls::host::HostBuf pin_in, pin_out;
  // family uploads: their own stream (they overlap the scans of the main
  // stream) and pinned staging, one upload at a time
  hipStream_t up_stream = nullptr;
  fbls::host::HostBuf pin_up;
  std::mutex up_mutex;
  std::shared_ptr<fbls::host::ArenaPool> arenas;  // released upload buffers
};

struct fbls_family {
  fbls_ctx* ctx = nullptr;
  std::shared_ptr<fbls::host::DevArena> arena;
  const double* t_dev = nullptr;    // n_pad origin-relative times (pad 0)
  const int32_t* vw_dev = nullptr;  // n_pad x {v, w} fixed point (pad 0)
  uint32_t n = 0, n_pad = 0;
  double C = 0.0, origin = 0.0, t_max = 0.0, t_last = 0.0;
  int ev = 0, ew = 0;
  uint32_t fx = 0;       // bit 0: v, bit 1: w in block mode (fixed_value)
  bool sorted = false;   // times ascend
  int32_t w_uniform = 0; // > 0: every fixed-point weight equals it
  std::vector<int32_t> vq, wq;
};
