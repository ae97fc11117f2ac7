// Host side of the MI355X FBLS hot path: C ABI (include/fbls.h), device
// context, family upload, and the scan drivers that restate the reference's
// src/search/scan.rs:144-473 around the fused HIP kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include <sys/resource.h>
#include <sys/syscall.h>
#include <unistd.h>

#include "../../include/fbls.h"
#include "prep.h"
#include "band_kernel.h"
#include "scan_host.h"
#include "scan_kernel.h"

#include "host_internal.h"

using namespace fbls::host;

namespace {
thread_local std::string g_last_error;
}  // namespace

int fbls::host::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

double fbls::host::now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

extern "C" const char* fbls_last_error(void) { return g_last_error.c_str(); }

namespace {

// Host worker threads shared by every call of the library: a call splits its
// range into pieces, queues all but the first, runs the first itself and
// then, until its own pieces are done, runs queued pieces (its own or
// another call's) or waits.  Nothing blocks on a piece that no thread will
// run, whatever the number of concurrent callers or nesting.
class Pool {
 public:
  static Pool& get() {
    static Pool* pool = new Pool(false);  // never destroyed: workers outlive exit handlers
    return *pool;
  }
  // The preparation threads (fbls_family_state_many): a pool of their own at
  // a lower CPU priority (nice + 5) and four threads short of the cores, so
  // that a batch being prepared does not delay the threads that plan
  // searches and feed the device (measured on the single-sector pruned batch
  // on 16 cores: 12 preparation threads 53-55 ms per 512 stars, 16 threads
  // 55-57, 10 threads 57-58, 8 threads 60-61).
  static Pool& background() {
    static Pool* pool = new Pool(true);
    return *pool;
  }
  size_t workers() const { return threads_.size(); }

  void run(size_t m, size_t pieces, const std::function<void(size_t, size_t)>& work) {
    Group g;
    g.left = pieces;
    {
      std::lock_guard<std::mutex> hold(mu_);
      for (size_t k = 1; k < pieces; k++)
        queue_.push_back(Task{&g, &work, m * k / pieces, m * (k + 1) / pieces});
    }
    cv_.notify_all();
    work(0, m / pieces);
    finish(&g);
    std::unique_lock<std::mutex> lock(mu_);
    while (g.left > 0) {
      if (!queue_.empty()) {
        Task t = queue_.front();
        queue_.pop_front();
        lock.unlock();
        (*t.work)(t.a, t.e);
        finish(t.g);
        lock.lock();
      } else {
        done_.wait(lock);
      }
    }
  }

 private:
  struct Group {
    size_t left = 0;
  };
  struct Task {
    Group* g;
    const std::function<void(size_t, size_t)>* work;
    size_t a, e;
  };

  explicit Pool(bool background) {
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t cap = std::min<size_t>(hw, 16);
    const size_t n = background ? std::max<size_t>(1, cap > 4 ? cap - 4 : 1) : cap - 1;
    for (size_t k = 0; k < n; k++) {
      try {
        threads_.emplace_back([this, background]() {
          if (background)  // this thread only; refused or not, it still works
            (void)setpriority(PRIO_PROCESS, (id_t)syscall(SYS_gettid), 5);
          loop();
        });
      } catch (const std::system_error&) {
        break;  // fewer workers: the callers run more pieces themselves
      }
    }
    for (std::thread& t : threads_) t.detach();
  }
  void finish(Group* g) {
    std::lock_guard<std::mutex> hold(mu_);
    if (--g->left == 0) done_.notify_all();
  }
  void loop() {
    std::unique_lock<std::mutex> lock(mu_);
    for (;;) {
      cv_.wait(lock, [this]() { return !queue_.empty(); });
      Task t = queue_.front();
      queue_.pop_front();
      lock.unlock();
      (*t.work)(t.a, t.e);
      finish(t.g);
      lock.lock();
    }
  }

  std::mutex mu_;
  std::condition_variable cv_, done_;
  std::deque<Task> queue_;
  std::vector<std::thread> threads_;
};

}  // namespace

void fbls::host::parallel_ranges(size_t m, size_t per,
                                 const std::function<void(size_t, size_t)>& work) {
  if (m == 0) return;
  Pool& pool = Pool::get();
  const size_t nt = std::min<size_t>(pool.workers() + 1, (m + per - 1) / std::max<size_t>(per, 1));
  if (nt <= 1) {
    work(0, m);
    return;
  }
  pool.run(m, nt, work);
}

// ---------------------------------------------------------------- context

extern "C" int fbls_ctx_create(int device, fbls_ctx** out) {
  return fbls::host::guard([&]() -> int {
    if (!out) return fail(FBLS_E_ARG, "null output pointer");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
      return fail(FBLS_E_ARG, "device %d out of range (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    fbls_ctx* c = new fbls_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess &&
        prop.maxSharedMemoryPerMultiProcessor > 0)
      c->max_lds = std::min<size_t>(kMaxLds, prop.maxSharedMemoryPerMultiProcessor);
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      c->cus = (uint32_t)prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->up_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess ||
        hipEventCreate(&c->evm) != hipSuccess ||
        hipEventCreate(&c->ev1) != hipSuccess) {
      delete c;
      return fail(FBLS_E_HIP, "stream/event creation failed");
    }
    hipError_t e = fbls::scan_configure(c->max_lds);
    if (e != hipSuccess) {
      fbls_ctx_destroy(c);
      return fail(FBLS_E_HIP, "kernel LDS configuration failed: %s",
                  hipGetErrorString(e));
    }
    fbls::host::noalloc_setup(device);
    c->arenas = std::make_shared<ArenaPool>(device);
    *out = c;
    return FBLS_OK;
  });
}

extern "C" void fbls_ctx_destroy(fbls_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->up_stream) (void)hipStreamSynchronize(c->up_stream);
  if (c->ev_exec) (void)hipGraphExecDestroy(c->ev_exec);
  c->dev_in.release();
  c->dev_out.release();
  c->dev_ev.release();
  c->pin_in.release();
  c->pin_out.release();
  c->pin_up.release();
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->evm) (void)hipEventDestroy(c->evm);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->up_stream) (void)hipStreamDestroy(c->up_stream);
  delete c;
}

extern "C" int fbls_mem_info(fbls_ctx* c, size_t* free_bytes, size_t* total) {
  return fbls::host::guard([&]() -> int {
    if (!c || !free_bytes || !total) return fail(FBLS_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemGetInfo(free_bytes, total));
    return FBLS_OK;
  });
}

extern "C" int fbls_ctx_set_launch(fbls_ctx* c, uint32_t lds_bytes,
                                   uint32_t block, uint32_t fold_mode) {
  return fbls::host::guard([&]() -> int {
    if (!c) return fail(FBLS_E_ARG, "null context");
    if (fold_mode > 2) return fail(FBLS_E_ARG, "fold_mode must be 0, 1 or 2");
    if (block != 0 && block != 256 && block != 512 && block != 1024)
      return fail(FBLS_E_ARG, "block must be 0, 256, 512 or 1024");
    if (lds_bytes > c->max_lds)
      return fail(FBLS_E_ARG, "lds_bytes %u exceeds %zu", lds_bytes, c->max_lds);
    c->lds_override = lds_bytes;
    c->block_override = block;
    c->run_fold = fold_mode == 2 ? 0u : 1u;
    c->count_from_weights = fold_mode == 0 ? 1u : 0u;
    return FBLS_OK;
  });
}

extern "C" int fbls_ctx_set_graphs(fbls_ctx* c, int enable, uint64_t* graph_launches) {
  return fbls::host::guard([&]() -> int {
    if (!c) return fail(FBLS_E_ARG, "null context");
    if (enable >= 0) c->use_graphs = enable ? 1 : 0;
    if (graph_launches) *graph_launches = c->graph_launches;
    return FBLS_OK;
  });
}

extern "C" int fbls_ctx_launch_totals(fbls_ctx* c, fbls_launch_totals* out,
                                      int reset) {
  return fbls::host::guard([&]() -> int {
    if (!c || !out) return fail(FBLS_E_ARG, "null argument");
    *out = c->totals;
    if (reset) c->totals = fbls_launch_totals{0, 0, 0, 0.0};
    return FBLS_OK;
  });
}

// ----------------------------------------------------------------- family

extern "C" int fbls_quantize(const double* x, size_t n, int32_t* q, int* block) {
  return fbls::host::guard([&]() -> int {
    const fbls::Fixed f = fbls::quantize(x, n, q);
    if (block) *block = f.block ? 1 : 0;
    return f.e;
  });
}

namespace {

// Validates one family and fills its host-side fields (fixed point, sizes);
// the device pointers are set by the caller.
int prepare_family(fbls_ctx* c, const double* t, const double* v,
                   const double* w, size_t n, double baseline_norm,
                   double origin, fbls_family* f) {
  if (!t || !v || !w) return fail(FBLS_E_ARG, "null argument");
  if (n < 6) return fail(FBLS_E_DATA, "at least six observations are required");
  if (n > (1u << 22))
    return fail(FBLS_E_ARG, "at most 2^22 observations per family (exact sums)");
  if (!(std::isfinite(baseline_norm) && baseline_norm > 0.0))
    return fail(FBLS_E_DATA, "invalid baseline precision");
  double tmax = 0.0;
  for (size_t i = 0; i < n; i++) {
    if (!(t[i] >= 0.0) || !std::isfinite(t[i]))
      return fail(FBLS_E_DATA, "times must be finite and origin-relative (>= 0)");
    if (!std::isfinite(v[i]) || !std::isfinite(w[i]))
      return fail(FBLS_E_DATA, "fields must be finite");
    tmax = std::max(tmax, t[i]);
  }
  f->ctx = c;
  f->n = (uint32_t)n;
  f->n_pad = (uint32_t)((n + kObsAlign - 1) / kObsAlign * kObsAlign);
  f->C = baseline_norm;
  f->origin = origin;
  f->t_max = tmax;
  f->t_last = t[n - 1];
  f->vq.resize(n);
  f->wq.resize(n);
  const fbls::Fixed fv = fbls::quantize(v, n, f->vq.data());
  const fbls::Fixed fw = fbls::quantize(w, n, f->wq.data());
  f->ev = fv.e;
  f->ew = fw.e;
  f->fx = (fv.block ? 1u : 0u) | (fw.block ? 2u : 0u);
  f->sorted = true;
  for (size_t i = 1; i < n; i++) f->sorted = f->sorted && t[i] >= t[i - 1];
  // equal weights are never in block mode (max = mean)
  f->w_uniform = f->wq[0] > 0 && !fw.block ? f->wq[0] : 0;
  for (size_t i = 1; i < n && f->w_uniform; i++)
    if (f->wq[i] != f->w_uniform) f->w_uniform = 0;
  return FBLS_OK;
}

// Device image of one family: t[n_pad] then vw[2 n_pad] (16 B per point).
size_t family_bytes(const fbls_family* f) { return (size_t)f->n_pad * 16; }

void write_family_image(const fbls_family* f, const double* t, unsigned char* dst) {
  double* tp = (double*)dst;
  int32_t* vw = (int32_t*)(dst + (size_t)f->n_pad * 8);
  std::memcpy(tp, t, (size_t)f->n * 8);
  std::fill(tp + f->n, tp + f->n_pad, 0.0);
  for (uint32_t i = 0; i < f->n; i++) {
    vw[2 * i] = f->vq[i];
    vw[2 * i + 1] = f->wq[i];
  }
  std::fill(vw + 2 * (size_t)f->n, vw + 2 * (size_t)f->n_pad, 0);
}

// Uploads the families (already prepared) into one arena: the device image
// is written into pinned staging by up to 16 threads and copied on the
// context's upload stream.
int upload_arena(fbls_ctx* c, std::vector<fbls_family*>& fams,
                 const std::vector<const double*>& times) {
  std::vector<size_t> at(fams.size() + 1, 0);
  for (size_t k = 0; k < fams.size(); k++) at[k + 1] = at[k] + family_bytes(fams[k]);
  const size_t total = at.back();
  auto arena = std::make_shared<DevArena>();
  arena->pool = c->arenas;
  HIP_TRY(hipSetDevice(c->device));
  if (int st = c->arenas->take(total, &arena->buf)) return st;
  {
    // in pieces of at most 256 MiB of staging (a larger family alone)
    constexpr size_t kPiece = (size_t)256 << 20;
    std::lock_guard<std::mutex> hold(c->up_mutex);
    for (size_t k0 = 0; k0 < fams.size();) {
      size_t k1 = k0 + 1;
      while (k1 < fams.size() && at[k1 + 1] - at[k0] <= kPiece) k1++;
      const size_t bytes = at[k1] - at[k0];
      if (int st = c->pin_up.ensure(bytes)) return st;
      unsigned char* img = (unsigned char*)c->pin_up.ptr;
      fbls::host::parallel_ranges(k1 - k0, 4, [&](size_t a, size_t e) {
        for (size_t k = k0 + a; k < k0 + e; k++)
          write_family_image(fams[k], times[k], img + (at[k] - at[k0]));
      });
      HIP_TRY(hipMemcpyAsync((unsigned char*)arena->buf.ptr + at[k0], img, bytes,
                             hipMemcpyHostToDevice, c->up_stream));
      HIP_TRY(hipStreamSynchronize(c->up_stream));
      k0 = k1;
    }
  }
  for (size_t k = 0; k < fams.size(); k++) {
    unsigned char* base = (unsigned char*)arena->buf.ptr + at[k];
    fams[k]->arena = arena;
    fams[k]->t_dev = (const double*)base;
    fams[k]->vw_dev = (const int32_t*)(base + (size_t)fams[k]->n_pad * 8);
  }
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_family_upload(fbls_ctx* c, const double* t,
                                  const double* v, const double* w, size_t n,
                                  double baseline_norm, double origin,
                                  fbls_family** out) {
  return fbls::host::guard([&]() -> int {
    if (!c || !out) return fail(FBLS_E_ARG, "null argument");
    *out = nullptr;
    std::unique_ptr<fbls_family> f(new fbls_family());
    int st = prepare_family(c, t, v, w, n, baseline_norm, origin, f.get());
    if (st) return st;
    std::vector<fbls_family*> one{f.get()};
    if ((st = upload_arena(c, one, {t}))) return st;
    *out = f.release();
    return FBLS_OK;
  });
}

namespace {

// Many families into one arena: family k holds n[k] observations at t[k],
// v[k], w[k].  Checks and fixed point on up to 16 host threads, then one
// upload (upload_arena); every handle or none is written to out.
int upload_many(fbls_ctx* c, size_t nfam, const double* const* t, const double* const* v,
                const double* const* w, const size_t* n, const double* baseline_norm,
                const double* origin, fbls_family** out) {
  for (size_t k = 0; k < nfam; k++) out[k] = nullptr;
  // the families are owned here until every one is uploaded
  std::vector<std::unique_ptr<fbls_family>> owned;
  std::vector<fbls_family*> fams;
  std::vector<const double*> times(t, t + nfam);
  owned.reserve(nfam);
  fams.reserve(nfam);
  for (size_t k = 0; k < nfam; k++) {
    owned.emplace_back(new fbls_family());
    fams.push_back(owned.back().get());
  }
  std::vector<int> status(nfam, FBLS_OK);
  std::vector<std::string> message(nfam);
  fbls::host::parallel_ranges(nfam, 8, [&](size_t lo, size_t hi) {
    for (size_t k = lo; k < hi; k++) {
      try {
        status[k] = prepare_family(c, t[k], v[k], w[k], n[k], baseline_norm[k], origin[k],
                                   fams[k]);
        if (status[k]) message[k] = fbls_last_error();
      } catch (const std::bad_alloc&) {
        status[k] = FBLS_E_OOM;
        message[k] = "host memory allocation failed";
      } catch (const std::exception& ex) {
        status[k] = FBLS_E_ARG;
        try {
          message[k] = ex.what();
        } catch (...) {
        }
      }
      if (status[k]) return;
    }
  });
  for (size_t k = 0; k < nfam; k++)
    if (status[k]) return fail(status[k], "family %zu: %s", k, message[k].c_str());
  if (!fams.empty())
    if (int st = upload_arena(c, fams, times)) return st;
  for (size_t k = 0; k < nfam; k++) out[k] = owned[k].release();
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_family_prepare_upload_many(
    fbls_ctx* c, size_t count, const double* const* time, const double* const* flux,
    const double* const* sigma, const size_t* n, double flux_offset, const double* durations,
    size_t nd, int correlated, size_t* out_n, double* baseline_norm, double* origin,
    double* rho, double* inflation, size_t* segments, double* residual_energy, double* t_last,
    int* status, fbls_family** out) {
  return fbls::host::guard([&]() -> int {
    if (!c || !out) return fail(FBLS_E_ARG, "null argument");
    for (size_t k = 0; k < count; k++) out[k] = nullptr;
    if (count == 0) return FBLS_OK;
    if (!time || !flux || !sigma || !n || !out_n || !baseline_norm || !origin ||
        !inflation || !segments || (nd && (!durations || !rho)))
      return fail(FBLS_E_ARG, "null argument");
    // staging slots sized by the input counts (a star keeps at most n[k])
    std::vector<size_t> at(count + 1, 0);
    for (size_t k = 0; k < count; k++)
      at[k + 1] = at[k] + (n[k] + kObsAlign - 1) / kObsAlign * kObsAlign * 16;
    std::vector<std::unique_ptr<fbls_family>> owned(count);
    for (auto& f : owned) f.reset(new fbls_family());
    std::vector<int> codes(count, FBLS_OK);
    std::vector<std::string> msgs(count);
    auto arena = std::make_shared<DevArena>();
    arena->pool = c->arenas;
    HIP_TRY(hipSetDevice(c->device));
    if (int st = c->arenas->take(at[count], &arena->buf)) return st;
    {
      // in pieces of at most 256 MiB of staging (a larger star alone)
      constexpr size_t kPiece = (size_t)256 << 20;
      std::lock_guard<std::mutex> hold(c->up_mutex);
      for (size_t k0 = 0; k0 < count;) {
        size_t k1 = k0 + 1;
        while (k1 < count && at[k1 + 1] - at[k0] <= kPiece) k1++;
        if (int st = c->pin_up.ensure(at[k1] - at[k0])) return st;
        unsigned char* img = (unsigned char*)c->pin_up.ptr;
        Pool::background().run(k1 - k0, k1 - k0, [&](size_t a, size_t e) {
          struct Out {
            std::vector<double> t, v, w;
          };
          thread_local Out o;
          for (size_t k = k0 + a; k < k0 + e; k++) {
            try {
              o.t.resize(n[k]);
              o.v.resize(n[k]);
              o.w.resize(n[k]);
              double energy = 0.0;
              codes[k] = fbls::family_state(time[k], flux[k], sigma[k], n[k], durations, nd,
                                            correlated, o.t.data(), o.v.data(), o.w.data(),
                                            out_n + k, baseline_norm + k, origin + k,
                                            rho + k * nd, inflation + k, segments + k,
                                            &energy, &msgs[k], flux_offset);
              if (residual_energy) residual_energy[k] = energy;
              if (codes[k]) continue;
              codes[k] = prepare_family(c, o.t.data(), o.v.data(), o.w.data(), out_n[k],
                                        baseline_norm[k], origin[k], owned[k].get());
              if (codes[k]) {
                msgs[k] = fbls_last_error();
                continue;
              }
              write_family_image(owned[k].get(), o.t.data(), img + (at[k] - at[k0]));
              if (o.t.capacity() > ((size_t)4 << 20)) o = Out();  // an unusually large star
            } catch (const std::bad_alloc&) {
              codes[k] = FBLS_E_OOM;
              msgs[k] = "host memory allocation failed";
            } catch (const std::exception& x) {
              codes[k] = FBLS_E_ARG;
              msgs[k] = x.what();
            }
          }
        });
        HIP_TRY(hipMemcpyAsync((unsigned char*)arena->buf.ptr + at[k0], img, at[k1] - at[k0],
                               hipMemcpyHostToDevice, c->up_stream));
        HIP_TRY(hipStreamSynchronize(c->up_stream));
        k0 = k1;
      }
    }
    size_t first = count;
    for (size_t k = 0; k < count; k++) {
      if (status) status[k] = codes[k];
      if (codes[k] && first == count) first = k;
    }
    if (first < count) return fail(codes[first], "star %zu: %s", first, msgs[first].c_str());
    for (size_t k = 0; k < count; k++) {
      fbls_family* f = owned[k].get();
      if (t_last) t_last[k] = f->t_last;
      unsigned char* base = (unsigned char*)arena->buf.ptr + at[k];
      f->arena = arena;
      f->t_dev = (const double*)base;
      f->vw_dev = (const int32_t*)(base + (size_t)f->n_pad * 8);
    }
    for (size_t k = 0; k < count; k++) out[k] = owned[k].release();
    return FBLS_OK;
  });
}

extern "C" int fbls_family_upload_packed(fbls_ctx* c, size_t nfam,
                                         const size_t* offsets, const double* t,
                                         const double* v, const double* w,
                                         const double* baseline_norm,
                                         const double* origin,
                                         fbls_family** out) {
  return fbls::host::guard([&]() -> int {
    if (!c || !offsets || !t || !v || !w || !baseline_norm || !origin || !out)
      return fail(FBLS_E_ARG, "null argument");
    for (size_t k = 0; k < nfam; k++) out[k] = nullptr;
    std::vector<const double*> tp(nfam), vp(nfam), wp(nfam);
    std::vector<size_t> n(nfam);
    for (size_t k = 0; k < nfam; k++) {
      if (offsets[k + 1] < offsets[k]) return fail(FBLS_E_ARG, "offsets must be nondecreasing");
      tp[k] = t + offsets[k];
      vp[k] = v + offsets[k];
      wp[k] = w + offsets[k];
      n[k] = offsets[k + 1] - offsets[k];
    }
    return upload_many(c, nfam, tp.data(), vp.data(), wp.data(), n.data(), baseline_norm,
                       origin, out);
  });
}

extern "C" int fbls_family_upload_list(fbls_ctx* c, size_t nfam, const double* const* t,
                                       const double* const* v, const double* const* w,
                                       const size_t* n, const double* baseline_norm,
                                       const double* origin, fbls_family** out) {
  return fbls::host::guard([&]() -> int {
    if (!c || !out || (nfam && (!t || !v || !w || !n || !baseline_norm || !origin)))
      return fail(FBLS_E_ARG, "null argument");
    for (size_t k = 0; k < nfam; k++) out[k] = nullptr;
    for (size_t k = 0; k < nfam; k++)
      if (!t[k] || !v[k] || !w[k]) return fail(FBLS_E_ARG, "null array of family %zu", k);
    return upload_many(c, nfam, t, v, w, n, baseline_norm, origin, out);
  });
}

extern "C" void fbls_family_free(fbls_family* f) {
  if (!f) return;
  delete f;  // the arena goes with its last family
}

extern "C" int fbls_family_fields(const fbls_family* f, double* v_out,
                                  double* w_out, int* ev, int* ew) {
  return fbls::host::guard([&]() -> int {
    if (!f) return fail(FBLS_E_ARG, "null family");
    for (uint32_t i = 0; i < f->n; i++) {
      if (v_out) v_out[i] = std::ldexp(fbls::fixed_value(f->vq[i], f->fx & 1), -f->ev);
      if (w_out) w_out[i] = std::ldexp(fbls::fixed_value(f->wq[i], f->fx & 2), -f->ew);
    }
    if (ev) *ev = f->ev;
    if (ew) *ew = f->ew;
    return FBLS_OK;
  });
}

// ------------------------------------------------------------- scan_bank

int fbls::host::check_bank(const fbls_bank* b) {
  if (!b) return fail(FBLS_E_ARG, "null bank");
  if (b->nd == 0 || b->nd > 32)
    return fail(FBLS_E_ARG, "1..32 durations per band are supported, got %zu", b->nd);
  if (b->ns == 0 || b->ns > 8) return fail(FBLS_E_ARG, "1..8 shapes supported");
  if (b->layers < 1 || b->layers > 16) return fail(FBLS_E_ARG, "1..16 layers supported");
  for (size_t i = 0; i < b->ns; i++)
    if (!(b->ingress[i] >= 0.0 && b->ingress[i] < 0.5))
      return fail(FBLS_E_ARG, "ingress fractions must be in [0, 0.5)");
  for (size_t i = 0; i < b->nd; i++)
    if (!(b->durations[i] > 0.0 && std::isfinite(b->durations[i])) ||
        !(b->rho[i] > 0.0 && std::isfinite(b->rho[i])))
      return fail(FBLS_E_ARG, "durations and rho must be positive and finite");
  return FBLS_OK;
}

uint64_t fbls::host::bank_batch(uint32_t nb, uint32_t period_batch, double max_fold_mib) {
  double per_period =
      (double)(3 * (4 * (uint64_t)nb + 2) * 8 + 3 * (2 * (uint64_t)nb + 1) * 8);
  return (uint64_t)std::floor((max_fold_mib * 1048576.0) / per_period);
}

fbls::host::Shared fbls::host::shared_of(const fbls_bank* b) {
  Shared sh;
  sh.ns = (uint32_t)b->ns;
  sh.layers = b->layers;
  sh.ingress = b->ingress;
  sh.gate = b->shape_gate;
  return sh;
}

bool fbls::host::needs_split(const fbls_ctx* c, const Work& x, const Shared& sh) {
  return fbls::scan_lds_bytes(x.nb, 1, x.nd, sh.ns, sh.layers, nullptr) > c->max_lds;
}

int fbls::host::check_periods(const Work& x, double* smax) {
  double s = 0.0;
  for (size_t k = 0; k < x.np; k++) {
    const double P = x.periods[k];
    if (!(P > 0.0) || !std::isfinite(P))
      return fail(FBLS_E_ARG, "periods must be positive and finite");
    s = std::max(s, (double)x.nb / P);  // scan.rs:204
  }
  if (x.fam->t_max * s >= 4294967295.0) return fail(FBLS_E_ARG, "t * nb / P exceeds 2^32");
  *smax = s;
  return FBLS_OK;
}

namespace {

size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Longest template of a segment, in bins: max over its periods and
// durations of min(max(1, floor(D nb / P + 0.5)), nb / 2) (kernels.rs:679-680).
uint32_t longest_template(const Work& x) {
  double pmin = INFINITY, dmax = 0.0;
  for (size_t k = 0; k < x.np; k++) pmin = std::min(pmin, x.periods[k]);
  for (uint32_t k = 0; k < x.nd; k++) dmax = std::max(dmax, x.dur[k]);
  const double rk = std::floor(dmax * ((double)x.nb / pmin) + 0.5);
  const uint32_t raw = rk < 1.0 ? 1u : (rk > 1e9 ? 1000000000u : (uint32_t)rk);
  return std::min(raw, x.nb / 2);
}

}  // namespace

int fbls::host::segment_shapes(const fbls_ctx* c, const Work* w, size_t nw, const Shared& sh,
                               bool split, SegShape* shapes, size_t* lds_max,
                               uint32_t* groups, bool cap_by_total) {
  const uint32_t ns = sh.ns, layers = sh.layers;
  size_t total = 0;
  for (size_t i = 0; i < nw; i++) total += w[i].np;
  // the launch's LDS budget: the configured one, raised to the largest
  // single-period need of its segments; by default half the CU's LDS (two
  // 512-thread workgroups per CU), or all of it when one period needs more
  // than half (one 1024-thread workgroup per CU)
  size_t lds = c->lds_override ? c->lds_override : c->max_lds / 2;
  for (size_t i = 0; i < nw; i++) {
    shapes[i] = SegShape();
    if (w[i].np == 0) continue;
    if (w[i].nb < 2 || w[i].nb >= 65536)
      return fail(FBLS_E_ARG, "phase bins must be in [2, 65535], got %u", w[i].nb);
    if (split) {
      if (sh.ingress[0] != 0.0)
        return fail(FBLS_E_ARG, "phase bins beyond the LDS need ingress[0] = 0");
      const uint32_t halo = longest_template(w[i]);
      // widest multiple of 128 bins whose window fits the budget: by default
      // the whole LDS (one workgroup per CU).  Measured on MI355X at
      // 9000-27000 bins: windows of half the LDS (two workgroups per CU)
      // are 0-26 % slower with 512 threads, the extra ranges costing more
      // than the overlap gains.
      const size_t budget = c->lds_override ? c->lds_override : c->max_lds;
      uint32_t B = 128;
      while (fbls::split_lds_bytes(w[i].nb, B + 128, halo, w[i].nd, ns, layers) <= budget &&
             B < w[i].nb)
        B += 128;
      // a window of B + halo bins holds each bin at most once (the fold
      // places an observation in one slot): B <= nb - halo
      const uint32_t cap = (w[i].nb > halo ? (w[i].nb - halo) / 128 * 128 : 0);
      B = std::min(B, std::max<uint32_t>(cap, 128));
      if (B + halo > w[i].nb)
        return fail(FBLS_E_ARG, "templates of %u bins leave no range of %u bins", halo,
                    w[i].nb);
      if (fbls::split_lds_bytes(w[i].nb, B, halo, w[i].nd, ns, layers) > c->max_lds)
        return fail(FBLS_E_ARG, "templates of %u bins do not fit the LDS", halo);
      shapes[i].halo = halo;
      shapes[i].range = B;
      shapes[i].lds = fbls::split_lds_bytes(w[i].nb, B, halo, w[i].nd, ns, layers);
    } else {
      shapes[i].lds = fbls::scan_lds_bytes(w[i].nb, 1, w[i].nd, ns, layers, nullptr);
      if (shapes[i].lds > c->max_lds)
        return fail(FBLS_E_ARG, "%u phase bins need %zu bytes of LDS per period (limit %zu)",
                    w[i].nb, shapes[i].lds, c->max_lds);
    }
    lds = std::max(lds, shapes[i].lds);
  }
  if (!c->lds_override && !split && 2 * lds > c->max_lds) lds = c->max_lds;
  // at most total / 1024 periods per workgroup, so the grid fills 256 CUs
  // (without the cap: the largest G the LDS allows, at most 16)
  const uint32_t g_cap =
      cap_by_total ? (uint32_t)std::max<size_t>(1, total / 1024) : 16u;
  *lds_max = 0;
  *groups = 0;
  for (size_t i = 0; i < nw; i++) {
    const Work& x = w[i];
    if (x.np == 0) continue;
    SegShape& g = shapes[i];
    if (split) {
      g.G = 1;
      *lds_max = std::max(*lds_max, g.lds);
    } else {
      uint32_t G = (uint32_t)std::max<size_t>(1, lds / g.lds);
      G = std::min<uint32_t>(std::min<uint32_t>(G, 16), g_cap);
      while (G > 1 && fbls::scan_lds_bytes(x.nb, G, x.nd, ns, layers, nullptr) > lds) G--;
      g.G = G;
      *lds_max = std::max(*lds_max, fbls::scan_lds_bytes(x.nb, G, x.nd, ns, layers, &g.pstride));
    }
    *groups += (uint32_t)((x.np + g.G - 1) / g.G);
  }
  return FBLS_OK;
}

uint32_t fbls::host::launch_block(const fbls_ctx* c, bool split, size_t lds) {
  // The range-split kernel: 512 threads (its 155 VGPRs allow three waves
  // per SIMD; 1024 threads would spill), or 256 when asked.
  if (split) return c->block_override == 256 ? 256 : 512;
  if (c->block_override) return c->block_override;
  // 16 waves per CU (four per SIMD, the occupancy of the kernel's 128
  // VGPRs): two 512-thread workgroups when two fit the LDS, else one of
  // 1024.  Measured on MI355X over stars of 76k-494k observations: against
  // 256-thread workgroups of 32 KiB (four per CU), 2-3 % faster at 384 phase
  // bins, up to 1.3x at 1536 (two periods per workgroup instead of one) and
  // 1.3-1.9x at 3072-6144, where one period fills more than 40 KiB and
  // 256-thread workgroups leave 4-8 waves per CU.
  return 2 * lds > c->max_lds ? 1024 : 512;
}

void fbls::host::segment_fields(const fbls_ctx* c, const Work& x, const SegShape& g,
                                bool split, double smax, fbls::Segment* S) {
  const fbls_family* f = x.fam;
  std::memset(S, 0, sizeof(*S));
  S->t = f->t_dev;
  S->vw = f->vw_dev;
  S->sv = std::ldexp(1.0, -f->ev);
  S->sw = std::ldexp(1.0, -f->ew);
  S->fx = f->fx;
  S->C = f->C;
  S->invC = 1.0 / f->C;
  S->n = f->n;
  S->n_pad = f->n_pad;
  S->np = (uint32_t)x.np;
  S->nb = x.nb;
  S->nd = x.nd;
  S->G = g.G;
  S->pstride = g.pstride;
  if (split) {
    S->range_bins = g.range;
    S->halo_bins = g.halo;
  }
  S->wide_cells = (f->t_max * smax >= 16777215.0) ? 1u : 0u;
  S->sorted = f->sorted ? 1u : 0u;
  S->uniform_w = c->count_from_weights ? f->w_uniform : 0;
}

void fbls::host::bank_factors(const Shared& sh, double* factors) {
  for (uint32_t s = 0; s < sh.ns; s++)
    for (uint32_t l = 0; l < sh.layers; l++)  // scan.rs:157-171
      factors[s * sh.layers + l] =
          1.0 - 2.0 * sh.ingress[s] *
                    (sh.layers > 1 ? (double)l / (double)(sh.layers - 1) : 0.0);
}

void fbls::host::account(fbls_scan_stats* st, const Work& x, const fbls_bank* b,
                         uint64_t budget, double seconds, const unsigned long long* cycles) {
  st->phase_bins = x.nb;
  uint64_t batch = std::min<uint64_t>(std::min<uint64_t>(b->period_batch, budget), x.np);
  if (batch) st->batches += (x.np + batch - 1) / batch;
  st->templates_scanned += (uint64_t)x.np * x.nd * b->ns * x.nb;  // scan.rs:352
  st->exact_tile_retries += x.retries;
  st->seconds_device += seconds;
  // FBLS_PROFILE (scan.rs:152,222-229): the kernel's shader cycles per stage
  // share out the work's device time; unprofiled, the stages stay 0 as in
  // the reference
  if (cycles) {
    const double tot = (double)cycles[0] + (double)cycles[1] + (double)cycles[2];
    if (tot > 0.0) {
      st->seconds_fold += seconds * (double)cycles[0] / tot;
      st->seconds_cumulative += seconds * (double)cycles[1] / tot;
      st->seconds_scan += seconds * (double)cycles[2] / tot;
    }
  }
}

bool fbls::host::profiling() { return std::getenv("FBLS_PROFILE") != nullptr; }

namespace {

// One kernel launch over the Works: the LDS-resident kernel, or (split) the
// range-split kernel for periods whose histogram exceeds the LDS.  Writes
// each Work's host outputs and retry count; returns the launch's device time
// through *seconds.
int run_launch(fbls_ctx* c, Work* w, size_t nw, const Shared& sh, bool split,
               double* seconds) {
  *seconds = 0.0;
  size_t total = 0;
  for (size_t i = 0; i < nw; i++) total += w[i].np;
  if (total == 0) return FBLS_OK;
  if (total > 0x7fffffffu) return fail(FBLS_E_ARG, "too many periods in one launch");
  HIP_TRY(hipSetDevice(c->device));
  const uint32_t ns = sh.ns, layers = sh.layers;
  std::vector<SegShape> shapes(nw);
  size_t lds_max = 0;
  uint32_t groups = 0;
  if (int st = segment_shapes(c, w, nw, sh, split, shapes.data(), &lds_max, &groups)) return st;
  std::vector<fbls::Segment> segs;
  std::vector<size_t> seg_work;
  segs.reserve(nw);
  // staging layout (doubles): per segment periods | scales | dur | rho, then
  // ingress | factors, then the segment table and zeroed counters
  size_t in_doubles = ns + ns * layers;
  for (size_t i = 0; i < nw; i++)
    if (w[i].np) in_doubles += 2 * w[i].np + 2 * w[i].nd;
  size_t nseg = 0;
  for (size_t i = 0; i < nw; i++) nseg += w[i].np ? 1 : 0;
  const size_t seg_off = align16(in_doubles * 8);
  const size_t cnt_off = align16(seg_off + nseg * sizeof(fbls::Segment));
  // counters: per segment two u32, then (profiled) three u64 stage cycles
  const bool prof = profiling();
  const size_t cnt_bytes = nseg * 8 + (prof ? nseg * 24 : 0);
  const size_t in_bytes = cnt_off + align16(cnt_bytes);
  const size_t out_bytes = total * (9 + 2) * 8 + total * 16;
  int st;
  if ((st = c->dev_in.ensure(in_bytes)) || (st = c->dev_out.ensure(out_bytes)) ||
      (st = c->pin_in.ensure(in_bytes)) || (st = c->pin_out.ensure(out_bytes)))
    return st;
  unsigned char* hin = (unsigned char*)c->pin_in.ptr;
  unsigned char* din = (unsigned char*)c->dev_in.ptr;
  unsigned char* dout = (unsigned char*)c->dev_out.ptr;
  double* hd = (double*)hin;
  size_t cur = 0;
  const size_t ing_at = cur;
  for (uint32_t s = 0; s < ns; s++) hd[cur++] = sh.ingress[s];
  const size_t fac_at = cur;
  bank_factors(sh, hd + cur);
  cur += ns * layers;
  size_t out_at = 0;  // in periods
  uint32_t group0 = 0;
  const size_t rec_base = 0, score_base = total * 9 * 8,
               depth_base = total * 10 * 8, win_base = total * 11 * 8;
  for (size_t i = 0; i < nw; i++) {
    const Work& x = w[i];
    if (x.np == 0) continue;
    double smax = 0.0;
    if (int e = check_periods(x, &smax)) return e;
    const size_t per_at = cur;
    for (size_t k = 0; k < x.np; k++) {
      hd[cur + k] = x.periods[k];
      hd[cur + x.np + k] = (double)x.nb / x.periods[k];  // scan.rs:204
    }
    cur += 2 * x.np;
    const size_t dur_at = cur;
    for (uint32_t k = 0; k < x.nd; k++) {
      hd[cur + k] = x.dur[k];
      hd[cur + x.nd + k] = x.rho[k];
    }
    cur += 2 * x.nd;
    fbls::Segment S;
    segment_fields(c, x, shapes[i], split, smax, &S);
    S.periods = (const double*)(din + per_at * 8);
    S.scales = S.periods + x.np;
    S.dur = (const double*)(din + dur_at * 8);
    S.rho = S.dur + x.nd;
    S.records = (double*)(dout + rec_base) + out_at * 9;
    S.score = (double*)(dout + score_base) + out_at;
    S.depth = (double*)(dout + depth_base) + out_at;
    S.winner = x.winner ? (uint32_t*)(dout + win_base) + out_at * 4 : nullptr;
    S.group0 = group0;
    S.index = (uint32_t)segs.size();
    group0 += (uint32_t)((x.np + S.G - 1) / S.G);
    segs.push_back(S);
    seg_work.push_back(i);
    out_at += x.np;
  }
  std::memcpy(hin + seg_off, segs.data(), segs.size() * sizeof(fbls::Segment));
  std::memset(hin + cnt_off, 0, align16(cnt_bytes));
  fbls::LaunchArgs A;
  std::memset(&A, 0, sizeof(A));
  A.segs = (const fbls::Segment*)(din + seg_off);
  A.nseg = (uint32_t)segs.size();
  A.ns = ns;
  A.layers = layers;
  A.run_fold = c->run_fold;
  A.ingress = (const double*)(din + ing_at * 8);
  A.factors = (const double*)(din + fac_at * 8);
  A.gate = sh.gate;
  A.counters = (unsigned int*)(din + cnt_off);
  if (prof) A.stages = (unsigned long long*)(din + cnt_off + nseg * 8);
  const uint32_t block = launch_block(c, split, lds_max);
  HIP_TRY(hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  HIP_TRY(fbls::scan_launch(A, groups, block, lds_max, split, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  bool any_winner = false, all_score = true;
  for (size_t i = 0; i < nw; i++) {
    any_winner = any_winner || (w[i].winner && w[i].np);
    all_score = all_score && (w[i].score_only || !w[i].np);
  }
  if (all_score) {  // a prescan launch: only the scores come back (8 of 88 B per period)
    HIP_TRY(hipMemcpyAsync((unsigned char*)c->pin_out.ptr + score_base, dout + score_base,
                           total * 8, hipMemcpyDeviceToHost, c->stream));
  } else {
    const size_t copy_bytes = any_winner ? out_bytes : win_base;
    HIP_TRY(hipMemcpyAsync(c->pin_out.ptr, dout, copy_bytes, hipMemcpyDeviceToHost, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(hin + cnt_off, din + cnt_off, cnt_bytes, hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  float ms = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  *seconds = ms * 1e-3;
  c->totals.launches += 1;
  c->totals.seconds += *seconds;
  for (size_t i = 0; i < nw; i++) {
    c->totals.periods += w[i].np;
    c->totals.fold_pairs += (uint64_t)w[i].np * w[i].fam->n;
  }
  const unsigned char* hout = (const unsigned char*)c->pin_out.ptr;
  const unsigned int* cnt = (const unsigned int*)(hin + cnt_off);
  bool bad = false;
  out_at = 0;
  for (size_t s = 0; s < segs.size(); s++) {
    Work& x = w[seg_work[s]];
    std::memcpy(x.score, (const double*)(hout + score_base) + out_at, x.np * 8);
    if (!x.score_only) {  // (a launch with any such work copies everything back)
      std::memcpy(x.records, (const double*)(hout + rec_base) + out_at * 9, x.np * 72);
      std::memcpy(x.depth, (const double*)(hout + depth_base) + out_at, x.np * 8);
    }
    if (x.winner)
      std::memcpy(x.winner, (const uint32_t*)(hout + win_base) + out_at * 4, x.np * 16);
    bad = bad || cnt[2 * s] != 0;
    x.retries = cnt[2 * s + 1];
    x.profiled = prof;
    if (prof) {
      const unsigned long long* cyc =
          (const unsigned long long*)(hin + cnt_off + nseg * 8) + 3 * s;
      for (int k = 0; k < 3; k++) x.cycles[k] = cyc[k];
    }
    out_at += x.np;
  }
  if (bad)
    return fail(FBLS_E_NONPOSITIVE_PRECISION,
                "nonpositive template precision; inspect noise model");
  return FBLS_OK;
}

// Runs every Work: the LDS-resident kernel for those whose histogram fits,
// then the range-split kernel for the rest (same stream, two launches).
int run_work(fbls_ctx* c, Work* w, size_t nw, const Shared& sh,
             double* seconds) {
  std::vector<Work> lo, hi;
  std::vector<size_t> lo_at, hi_at;
  for (size_t i = 0; i < nw; i++) {
    if (needs_split(c, w[i], sh)) {
      hi.push_back(w[i]);
      hi_at.push_back(i);
    } else {
      lo.push_back(w[i]);
      lo_at.push_back(i);
    }
  }
  double s1 = 0.0, s2 = 0.0;
  int st = run_launch(c, lo.data(), lo.size(), sh, false, &s1);
  if (st) return st;
  st = run_launch(c, hi.data(), hi.size(), sh, true, &s2);
  if (st) return st;
  for (size_t k = 0; k < lo.size(); k++) w[lo_at[k]] = lo[k];
  for (size_t k = 0; k < hi.size(); k++) w[hi_at[k]] = hi[k];
  *seconds = s1 + s2;
  return FBLS_OK;
}

// Work estimate of a segment: fold pairs plus box templates.
double work_cost(const Work& x) {
  return (double)x.np * ((double)x.fam->n + 3.0 * x.nb * x.nd);
}

}  // namespace

extern "C" int fbls_scan_bank(fbls_ctx* c, const fbls_family* f,
                              const double* periods, size_t np,
                              const fbls_bank* b, uint32_t nb, double* records,
                              double* score, double* depth, uint32_t* winner,
                              fbls_scan_stats* stats) {
  return fbls::host::guard([&]() -> int {
    if (!c || !f || (!periods && np) || !records || !score || !depth)
      return fail(FBLS_E_ARG, "null argument");
    int st = check_bank(b);
    if (st) return st;
    const uint64_t budget = bank_batch(nb, b->period_batch, b->max_fold_mib);
    if (budget == 0) return fail(FBLS_E_FOLD_BUDGET, "one folded period exceeds max_fold_mib");
    Work x;
    x.fam = f;
    x.periods = periods;
    x.np = np;
    x.nb = nb;
    x.nd = (uint32_t)b->nd;
    x.dur = b->durations;
    x.rho = b->rho;
    x.records = records;
    x.score = score;
    x.depth = depth;
    x.winner = winner;
    double secs = 0.0;
    st = run_work(c, &x, 1, shared_of(b), &secs);
    if (st) return st;
    if (stats) account(stats, x, b, budget, secs, x.profiled ? x.cycles : nullptr);
    return FBLS_OK;
  });
}

namespace {

// One band (one reference scan_periods call): its periods (host), phase bins
// of the full scan and of the prescan (0: none), bank durations and rho.
struct BandRun {
  const fbls_family* fam;
  const double* periods;
  size_t np;
  uint32_t nb, nb_coarse, nd;
  const double* dur;
  const double* rho;
  fbls_scan_stats stats;
};

// The launch of one segment table: the Works (their periods on the host),
// shapes, LDS bytes, workgroup size and workgroups (an upper bound for a
// persistent table).
struct Table {
  std::vector<Work> w;
  std::vector<size_t> band;  // the active band of each Work
  std::vector<SegShape> shape;
  size_t lds = 0;
  uint32_t groups = 0, block = 0;
  bool split = false;
};

size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// scan.rs:363-408 for any number of bands of one or more families, as one
// stream of launches: every band's first scan (coarse, or full when there is
// no prescan) in one launch per kernel, the survivors compacted and the full
// scan's segment tables completed on the device, the full scan of every
// band's survivors in one persistent launch per kernel, the rows extracted
// on the device; the host waits once, at the end.  box_first (searches
// restricted to proposed period bins, where nearly every period passes the
// prescan): each prescan period scans its boxes first and skips its
// staircase templates when its best box already reaches the cut (the
// survivors are the same; where few periods survive the extra box pass
// would cost more than it saves).
int scan_bands(fbls_ctx* c, std::vector<BandRun>& bands, const fbls_bank* b,
               fbls_row* rows, size_t* nrows, size_t* row_counts, bool box_first) {
  const double t_start = now_s();
  const Shared sh = shared_of(b);
  const uint32_t ns = sh.ns, layers = sh.layers;
  *nrows = 0;
  std::vector<size_t> act;
  for (size_t i = 0; i < bands.size(); i++) {
    BandRun& r = bands[i];
    std::memset(&r.stats, 0, sizeof(r.stats));
    r.stats.periods = r.np;
    if (row_counts) row_counts[i] = 0;
    if (r.np == 0 || r.nd == 0) continue;
    const uint32_t nb0 = r.nb_coarse ? r.nb_coarse : r.nb;
    if (bank_batch(nb0, b->period_batch, b->max_fold_mib) == 0 ||
        bank_batch(r.nb, b->period_batch, b->max_fold_mib) == 0)
      return fail(FBLS_E_FOLD_BUDGET, "one folded period exceeds max_fold_mib");
    if (r.np > 0x7fffffffu) return fail(FBLS_E_ARG, "too many periods in one band");
    act.push_back(i);
  }
  const size_t na = act.size();
  if (na == 0) return FBLS_OK;
  HIP_TRY(hipSetDevice(c->device));
  // tables: [0] first scan LDS-resident, [1] first scan range-split,
  // [2] full scan LDS-resident, [3] full scan range-split
  Table T[4];
  T[1].split = T[3].split = true;
  size_t np_total = 0, nprescan = 0;
  for (size_t k = 0; k < na; k++) {
    const BandRun& r = bands[act[k]];
    Work x;
    x.fam = r.fam;
    x.periods = r.periods;  // the host's copy: checks, scales, halo
    x.np = r.np;
    x.nb = r.nb_coarse ? r.nb_coarse : r.nb;
    x.nd = r.nd;
    x.dur = r.dur;
    x.rho = r.rho;
    x.records = x.score = x.depth = nullptr;
    x.winner = nullptr;
    x.cost = work_cost(x);
    Table& t1 = T[needs_split(c, x, sh) ? 1 : 0];
    t1.w.push_back(x);
    t1.band.push_back(k);
    np_total += r.np;
    if (r.nb_coarse) {
      // the full scan's bound: every period of the band survives
      x.nb = r.nb;
      Table& t2 = T[needs_split(c, x, sh) ? 3 : 2];
      t2.w.push_back(x);
      t2.band.push_back(k);
      nprescan++;
    }
  }
  if (np_total > 0x7fffffffu) return fail(FBLS_E_ARG, "too many periods in one launch");
  for (int t = 0; t < 4; t++) {
    Table& tb = T[t];
    if (tb.w.empty()) continue;
    tb.shape.resize(tb.w.size());
    if (int st = segment_shapes(c, tb.w.data(), tb.w.size(), sh, tb.split, tb.shape.data(),
                                &tb.lds, &tb.groups, /*cap_by_total=*/t < 2))
      return st;
    tb.block = launch_block(c, tb.split, tb.lds);
    // the full pass: persistent (at most the workgroups the chip holds at
    // once), or, where nearly every period survives (box_first), one group
    // per workgroup of a grid sized for every period surviving at the host's
    // G (the loop-free kernel spills fewer registers; its idle tail is short)
    if (t >= 2 && !box_first) {
      const uint32_t per_cu = (tb.split || tb.block == 1024) ? 1u : 2u;
      tb.groups = std::min<uint32_t>(tb.groups, std::max<uint32_t>(1, c->cus) * per_cu);
    }
  }
  // ---- staging (one host-to-device copy): bank tables, per-band periods,
  // scales, durations and rho, the four segment tables, the band table and
  // the zeroed counters
  std::vector<size_t> per_at(na), dur_at(na);
  size_t cur = 0;  // doubles
  const size_t ing_at = cur;
  cur += ns;
  const size_t fac_at = cur;
  cur += (size_t)ns * layers;
  for (size_t k = 0; k < na; k++) {
    const BandRun& r = bands[act[k]];
    per_at[k] = cur;
    cur += 2 * r.np;  // periods | first-scan scales
    dur_at[k] = cur;
    cur += 2 * (size_t)r.nd;
  }
  size_t nseg = 0;
  for (const Table& tb : T) nseg += tb.w.size();
  const size_t seg_off = align16(cur * 8);
  const size_t band_off = align16(seg_off + nseg * sizeof(fbls::Segment));
  const size_t cnt_off = align16(band_off + na * sizeof(fbls::BandDesc));
  // counters (u32): counts[2 na + 1] | vgroups[2] | tickets[2] | per segment [2],
  // then (profiled) per segment three u64 stage-cycle sums
  const bool prof = profiling();
  const size_t ncnt = 2 * na + 1 + 4 + 2 * nseg;
  const size_t stage_at = ((ncnt * 4 + 7) & ~(size_t)7);
  const size_t cnt_bytes = stage_at + (prof ? nseg * 24 : 0);
  const size_t in_bytes = cnt_off + align16(cnt_bytes);
  // device outputs, per band: first scan records | score | depth, then for a
  // band with prescan survivor periods | scales | full-scan records | score |
  // depth, then the band's rows; after the bands every row in band order and
  // the row offsets
  std::vector<size_t> out_at(na);
  size_t out_bytes = 0;
  for (size_t k = 0; k < na; k++) {
    const BandRun& r = bands[act[k]];
    out_at[k] = out_bytes;
    out_bytes += round256(r.np * 88);
    if (r.nb_coarse) out_bytes += round256(r.np * 104);
    out_bytes += round256(r.np * sizeof(fbls::RowOut));
  }
  const size_t rows_at = out_bytes;
  out_bytes += round256(np_total * sizeof(fbls::RowOut));
  const size_t off_at = out_bytes;
  out_bytes += round256(na * 4);
  const size_t guess = std::min<size_t>(np_total, std::max<size_t>(c->rows_hint, 1024));
  const size_t back_bytes = align16(cnt_bytes) + guess * sizeof(fbls::RowOut);
  int st;
  if ((st = c->dev_in.ensure(in_bytes)) || (st = c->dev_out.ensure(out_bytes)) ||
      (st = c->pin_in.ensure(in_bytes)) || (st = c->pin_out.ensure(back_bytes)))
    return st;
  unsigned char* hin = (unsigned char*)c->pin_in.ptr;
  unsigned char* din = (unsigned char*)c->dev_in.ptr;
  unsigned char* dout = (unsigned char*)c->dev_out.ptr;
  double* hd = (double*)hin;
  for (uint32_t s = 0; s < ns; s++) hd[ing_at + s] = sh.ingress[s];
  bank_factors(sh, hd + fac_at);
  // periods and scales in pieces of at most 16384 periods on the host threads
  struct Piece {
    size_t k, a, e;
  };
  std::vector<Piece> pieces;
  for (size_t k = 0; k < na; k++) {
    const BandRun& r = bands[act[k]];
    for (size_t a = 0; a < r.np; a += 16384) pieces.push_back({k, a, std::min(r.np, a + 16384)});
    double* d = hd + dur_at[k];
    for (uint32_t i = 0; i < r.nd; i++) {
      d[i] = r.dur[i];
      d[r.nd + i] = r.rho[i];
    }
  }
  parallel_ranges(pieces.size(), 2, [&](size_t lo, size_t hi) {
    for (size_t q = lo; q < hi; q++) {
      const BandRun& r = bands[act[pieces[q].k]];
      const uint32_t nb0 = r.nb_coarse ? r.nb_coarse : r.nb;
      double* p = hd + per_at[pieces[q].k];
      for (size_t i = pieces[q].a; i < pieces[q].e; i++) {
        p[i] = r.periods[i];
        p[r.np + i] = (double)nb0 / r.periods[i];  // scan.rs:204
      }
    }
  });
  auto dev_d = [&](size_t at_doubles) { return (double*)(din + at_doubles * 8); };
  uint32_t* dcnt = (uint32_t*)(din + cnt_off);
  uint32_t* d_counts = dcnt;
  uint32_t* d_vgroups = dcnt + 2 * na + 1;
  unsigned int* d_tickets = d_vgroups + 2;
  unsigned int* d_segcnt = d_tickets + 2;
  std::vector<fbls::BandDesc> bd(na);
  for (size_t k = 0; k < na; k++) {
    const BandRun& r = bands[act[k]];
    unsigned char* o = dout + out_at[k];
    fbls::BandDesc& B = bd[k];
    std::memset(&B, 0, sizeof(B));
    B.periods = dev_d(per_at[k]);
    B.rec1 = (const double*)o;
    B.score1 = (const double*)(o + r.np * 72);
    B.depth1 = (const double*)(o + r.np * 80);
    o += round256(r.np * 88);
    if (r.nb_coarse) {
      B.surv_p = (double*)o;
      B.surv_s = (double*)(o + r.np * 8);
      B.rec2 = (const double*)(o + r.np * 16);
      B.score2 = (const double*)(o + r.np * 88);
      B.depth2 = (const double*)(o + r.np * 96);
      o += round256(r.np * 104);
    }
    B.rows = (fbls::RowOut*)o;
    B.origin = r.fam->origin;
    B.nb = (double)r.nb;
    B.np = (uint32_t)r.np;
    B.prescan = r.nb_coarse ? 1u : 0u;
  }
  // segment tables
  const double cut = b->prescan_fraction * b->threshold;
  fbls::Segment* hseg = (fbls::Segment*)(hin + seg_off);
  const fbls::Segment* dseg = (const fbls::Segment*)(din + seg_off);
  size_t seg_at[4];
  size_t sbase = 0;
  for (int t = 0; t < 4; t++) {
    Table& tb = T[t];
    seg_at[t] = sbase;
    uint32_t group0 = 0;
    for (size_t j = 0; j < tb.w.size(); j++) {
      const Work& x = tb.w[j];
      const size_t k = tb.band[j];
      double smax = 0.0;
      if (int e = check_periods(x, &smax)) return e;
      fbls::Segment S;
      segment_fields(c, x, tb.shape[j], tb.split, smax, &S);
      const fbls::BandDesc& B = bd[k];
      S.dur = dev_d(dur_at[k]);
      S.rho = S.dur + x.nd;
      S.winner = nullptr;
      S.index = (uint32_t)j;
      if (t < 2) {
        S.periods = B.periods;
        S.scales = B.periods + x.np;
        S.records = const_cast<double*>(B.rec1);
        S.score = const_cast<double*>(B.score1);
        S.depth = const_cast<double*>(B.depth1);
        // a prescan only decides survival (score >= cut)
        if (box_first && B.prescan && cut > 0.0) S.survive = cut;
        S.group0 = group0;
        group0 += (uint32_t)((x.np + S.G - 1) / S.G);
      } else {
        // np, G (capped) and group0 are completed on the device
        S.periods = B.surv_p;
        S.scales = B.surv_s;
        S.records = const_cast<double*>(B.rec2);
        S.score = const_cast<double*>(B.score2);
        S.depth = const_cast<double*>(B.depth2);
        S.np = 0;
        S.group0 = 0;
        bd[k].fine_table = t - 2;
        bd[k].fine_seg = (uint32_t)j;
      }
      hseg[sbase + j] = S;
    }
    sbase += tb.w.size();
  }
  std::memcpy(hin + band_off, bd.data(), na * sizeof(fbls::BandDesc));
  std::memset(hin + cnt_off, 0, align16(cnt_bytes));
  unsigned long long* d_stages = (unsigned long long*)(din + cnt_off + stage_at);
  auto args_of = [&](int t) {
    fbls::LaunchArgs A;
    std::memset(&A, 0, sizeof(A));
    A.segs = dseg + seg_at[t];
    A.nseg = (uint32_t)T[t].w.size();
    A.ns = ns;
    A.layers = layers;
    A.run_fold = c->run_fold;
    A.ingress = dev_d(ing_at);
    A.factors = dev_d(fac_at);
    A.gate = sh.gate;
    A.counters = d_segcnt + 2 * seg_at[t];
    if (prof) A.stages = d_stages + 3 * seg_at[t];
    A.box_first = (box_first && t < 2) ? 1u : 0u;
    if (t >= 2) {
      A.vgroups = d_vgroups + (t - 2);
      A.ticket = box_first ? nullptr : d_tickets + (t - 2);
    }
    return A;
  };
  fbls::BandArgs BA;
  std::memset(&BA, 0, sizeof(BA));
  BA.bands = (fbls::BandDesc*)(din + band_off);
  BA.nbands = (uint32_t)na;
  BA.cut = cut;
  BA.threshold = b->threshold;
  BA.counts = d_counts;
  BA.row_off = (uint32_t*)(dout + off_at);
  BA.out = (fbls::RowOut*)(dout + rows_at);
  for (int t = 0; t < 2; t++) {
    BA.fine[t].segs = (fbls::Segment*)(din + seg_off) + seg_at[2 + t];
    BA.fine[t].nseg = (uint32_t)T[2 + t].w.size();
    // the device keeps the host's G for a bounded grid
    BA.fine[t].cap_by_total = (t == 0 && !box_first) ? 1u : 0u;
    BA.fine[t].vgroups = d_vgroups + t;
    BA.fine[t].ticket = d_tickets + t;
  }
  // ---- the stream
  HIP_TRY(hipMemcpyAsync(din, hin, in_bytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int t = 0; t < 2; t++)
    if (!T[t].w.empty())
      HIP_TRY(fbls::scan_launch(args_of(t), T[t].groups, T[t].block, T[t].lds, T[t].split,
                                c->stream));
  HIP_TRY(hipEventRecord(c->evm, c->stream));
  if (nprescan) {
    HIP_TRY(fbls::band_survivors(BA, c->stream));
    HIP_TRY(fbls::band_plan(BA, c->stream));
    for (int t = 2; t < 4; t++)
      if (!T[t].w.empty())
        HIP_TRY(fbls::scan_launch(args_of(t), T[t].groups, T[t].block, T[t].lds, T[t].split,
                                  c->stream));
  }
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(fbls::band_rows(BA, c->stream));
  unsigned char* hback = (unsigned char*)c->pin_out.ptr;
  const size_t rows_back = align16(cnt_bytes);
  HIP_TRY(hipMemcpyAsync(hback, din + cnt_off, cnt_bytes, hipMemcpyDeviceToHost, c->stream));
  if (guess)
    HIP_TRY(hipMemcpyAsync(hback + rows_back, dout + rows_at, guess * sizeof(fbls::RowOut),
                           hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  const uint32_t* hc = (const uint32_t*)hback;
  const size_t total_rows = hc[2 * na];
  if (total_rows > np_total)
    return fail(FBLS_E_HIP, "row count %zu exceeds the periods scanned", total_rows);
  const fbls_row* src = (const fbls_row*)(hback + rows_back);
  std::vector<fbls_row> more;
  if (total_rows > guess) {
    more.resize(total_rows);
    HIP_TRY(hipMemcpyAsync(more.data(), dout + rows_at, total_rows * sizeof(fbls::RowOut),
                           hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    src = more.data();
    c->rows_hint = total_rows + total_rows / 4;
  }
  float ms1 = 0.0f, ms2 = 0.0f;
  HIP_TRY(hipEventElapsedTime(&ms1, c->ev0, c->evm));
  HIP_TRY(hipEventElapsedTime(&ms2, c->evm, c->ev1));
  const double s1 = ms1 * 1e-3, s2 = ms2 * 1e-3;
  // ---- statistics (scan.rs:352 and the prescan survivors), device time
  // apportioned by each pass's work estimate
  bool bad = false;
  double c1 = 0.0, c2 = 0.0;
  std::vector<uint32_t> surv(na, 0);
  for (size_t k = 0; k < na; k++) surv[k] = bands[act[k]].nb_coarse ? hc[2 * k] : 0u;
  for (int t = 0; t < 4; t++)
    for (size_t j = 0; j < T[t].w.size(); j++) {
      Work& x = T[t].w[j];
      if (t >= 2) {
        x.np = surv[T[t].band[j]];
        x.cost = work_cost(x);
      }
      (t < 2 ? c1 : c2) += x.cost;
    }
  uint64_t launches = 0;
  for (int t = 0; t < 4; t++) {
    if (T[t].w.empty() || (t >= 2 && !nprescan)) continue;
    launches++;
    for (size_t j = 0; j < T[t].w.size(); j++) {
      Work& x = T[t].w[j];
      const uint32_t* cnt = hc + 2 * na + 1 + 4 + 2 * (seg_at[t] + j);
      bad = bad || cnt[0] != 0;
      x.retries = cnt[1];
      // a band without survivors has no full scan (its phase_bins stay the
      // prescan's, scan.rs:397-398)
      if (t >= 2 && x.np == 0) continue;
      BandRun& r = bands[act[T[t].band[j]]];
      const double secs = t < 2 ? (c1 > 0 ? s1 * x.cost / c1 : 0.0)
                                : (c2 > 0 ? s2 * x.cost / c2 : 0.0);
      const unsigned long long* cyc =
          prof ? (const unsigned long long*)(hback + stage_at) + 3 * (seg_at[t] + j) : nullptr;
      account(&r.stats, x, b, bank_batch(x.nb, b->period_batch, b->max_fold_mib), secs, cyc);
      c->totals.periods += x.np;
      c->totals.fold_pairs += (uint64_t)x.np * x.fam->n;
    }
  }
  c->totals.launches += launches;
  c->totals.seconds += s1 + s2;
  if (bad)
    return fail(FBLS_E_NONPOSITIVE_PRECISION,
                "nonpositive template precision; inspect noise model");
  for (size_t k = 0; k < na; k++) {
    BandRun& r = bands[act[k]];
    if (r.nb_coarse) r.stats.prescan_survivors = surv[k];
    if (row_counts) row_counts[act[k]] = hc[2 * k + 1];
  }
  std::memcpy(rows, src, total_rows * sizeof(fbls_row));
  *nrows = total_rows;
  // wall time of the call, apportioned like the device time
  const double wall = now_s() - t_start;
  const double sdev = s1 + s2;
  for (BandRun& r : bands)
    r.stats.seconds = sdev > 0 ? wall * r.stats.seconds_device / sdev : 0.0;
  return FBLS_OK;
}

uint32_t band_nb(const double* periods, size_t np, const fbls_bank* b) {
  double dmin = INFINITY;
  for (size_t i = 0; i < b->nd; i++) dmin = std::min(dmin, b->durations[i]);
  const double nbf = std::ceil((double)b->phase_oversample * periods[np - 1] / dmin);
  return nbf < 8.0 ? 8u : (nbf > 4e9 ? 0xffffffffu : (uint32_t)nbf);
}

}  // namespace

// scan.rs:363-408
extern "C" int fbls_scan_periods(fbls_ctx* c, const fbls_family* f,
                                 const double* periods, size_t np,
                                 const fbls_bank* b, fbls_row* rows,
                                 size_t* nrows, fbls_scan_stats* stats) {
  return fbls::host::guard([&]() -> int {
    if (!c || !f || !nrows || (np && (!periods || !rows)))
      return fail(FBLS_E_ARG, "null argument");
    *nrows = 0;
    if (np == 0 || b == nullptr || b->nd == 0) {
      if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->periods = np;
      }
      return FBLS_OK;
    }
    int st = check_bank(b);
    if (st) return st;
    std::vector<BandRun> bands(1);
    BandRun& r = bands[0];
    r.fam = f;
    r.periods = periods;
    r.np = np;
    r.nb = band_nb(periods, np, b);
    r.nb_coarse = (b->prescan_factor > 1 && r.nb / b->prescan_factor >= 8)
                      ? r.nb / b->prescan_factor
                      : 0u;
    r.nd = (uint32_t)b->nd;
    r.dur = b->durations;
    r.rho = b->rho;
    st = scan_bands(c, bands, b, rows, nrows, nullptr, false);
    if (st) return st;
    if (stats) *stats = r.stats;
    return FBLS_OK;
  });
}

// ---------------------------------------------------------- signal_search

extern "C" size_t fbls_period_grid_count(double minimum, double maximum,
                                         double baseline, double dmin,
                                         double oversample) {
  return fbls::period_grid_count(minimum, maximum, baseline, dmin, oversample);
}

extern "C" int fbls_period_grid(double minimum, double maximum,
                                double baseline, double dmin,
                                double oversample, double* out, size_t count) {
  return fbls::host::guard([&]() -> int {
    if (!out || count == 0) return fail(FBLS_E_ARG, "empty grid");
    parallel_ranges(count, 8192, [&](size_t a, size_t e) {
      fbls::period_grid_range(minimum, maximum, out, count, a, e);
    });
    return FBLS_OK;
  });
}

extern "C" size_t fbls_durations_days(const double* hours, size_t n,
                                      double* out) {
  return fbls::durations_days(hours, n, out);
}

extern "C" int fbls_plan_bands(const double* grid, size_t count,
                               const double* sorted_durations, size_t nd,
                               uint32_t phase_oversample,
                               uint32_t prescan_factor, fbls_band* bands,
                               size_t bands_cap, size_t* nbands) {
  return fbls::host::guard([&]() -> int {
    if (!grid || !sorted_durations || !nbands) return fail(FBLS_E_ARG, "null argument");
    std::vector<fbls_band> plan = fbls::plan_bands(grid, count, sorted_durations, nd,
                                                   phase_oversample, prescan_factor);
    *nbands = plan.size();
    if (bands) {
      if (plan.size() > bands_cap)
        return fail(FBLS_E_ARG, "band capacity %zu < %zu", bands_cap, plan.size());
      std::copy(plan.begin(), plan.end(), bands);
    }
    return FBLS_OK;
  });
}

extern "C" uint64_t fbls_bank_batches(size_t np, uint32_t nb,
                                      uint32_t period_batch,
                                      double max_fold_mib) {
  uint64_t budget = bank_batch(nb, period_batch, max_fold_mib);
  if (budget == 0 || np == 0) return 0;
  uint64_t batch = std::min<uint64_t>(std::min<uint64_t>(period_batch, budget), np);
  return (np + batch - 1) / batch;
}

namespace {

// One family's search plan (scan.rs:413-471): sorted bank, grid, bands.
// With a keep mask the grid is not materialised: the bands are found by
// binary search on the grid formula and only kept periods are evaluated.
struct SearchPlan {
  const fbls_family* fam = nullptr;
  std::vector<double> sd, sr, grid;
  std::vector<fbls_band> bands;
  std::vector<std::vector<double>> kept;  // per band, when a keep mask applies
  fbls::GridFormula formula;
  size_t count = 0;                       // grid periods
};

// Restricts each band to the grid periods whose keep flag is set; each kept
// period is scanned exactly as in the full search (same nb, same prescan).
void apply_keep(SearchPlan& p, const uint8_t* keep) {
  p.kept.assign(p.bands.size(), {});
  for (size_t b = 0; b < p.bands.size(); b++)
    for (size_t i = 0; i < p.bands[b].count; i++) {
      const size_t g = p.bands[b].first + i;
      if (keep[g]) p.kept[b].push_back(p.grid.empty() ? p.formula(g) : p.grid[g]);
    }
}

// Restricts each band to the grid periods inside the kept bins (edges[k] <=
// P < edges[k + 1]); index ranges by binary search on the grid formula.
void apply_bins(SearchPlan& p, const uint8_t* bin_keep, size_t nbins, const double* edges) {
  p.kept.assign(p.bands.size(), {});
  auto first_at_least = [&](size_t lo, size_t hi, double bound) {
    while (lo < hi) {
      const size_t mid = lo + (hi - lo) / 2;
      if (p.formula(mid) < bound)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  for (size_t b = 0; b < p.bands.size(); b++) {
    const size_t first = p.bands[b].first, end = first + p.bands[b].count;
    for (size_t k = 0; k < nbins; k++) {
      if (!bin_keep[k]) continue;
      const size_t i0 = first_at_least(first, end, edges[k]);
      const size_t i1 = first_at_least(i0, end, edges[k + 1]);
      for (size_t i = i0; i < i1; i++) p.kept[b].push_back(p.formula(i));
    }
  }
}

// SearchConfig's period range and grid density (config.rs:122-133): an
// oversample that is not finite and positive, or a range outside
// 0 < period_min < period_max, would make the grid count meaningless.
int check_search_cfg(const fbls_search_cfg* cfg) {
  if (!(std::isfinite(cfg->period_oversample) && cfg->period_oversample > 0.0))
    return fail(FBLS_E_ARG, "period_oversample must be finite and positive");
  if (!(std::isfinite(cfg->period_min) && std::isfinite(cfg->period_max) &&
        0.0 < cfg->period_min && cfg->period_min < cfg->period_max))
    return fail(FBLS_E_ARG, "require 0 < period_min < period_max");
  return FBLS_OK;
}

// Grid count with its size checked: at most 2^31 periods per family.
int grid_count_checked(double minimum, double maximum, double baseline, double dmin,
                       double oversample, size_t* count) {
  const double log_step = std::log1p(dmin / (oversample * baseline));  // grid.rs:12-13
  const double periods = std::ceil(std::log(maximum / minimum) / log_step) + 1.0;
  if (!(periods >= 1.0 && periods <= 2147483648.0))
    return fail(FBLS_E_ARG, "period grid of %g periods (at most 2^31 per family)", periods);
  *count = fbls::period_grid_count(minimum, maximum, baseline, dmin, oversample);
  return FBLS_OK;
}

int plan_search(const fbls_family* f, const double* durations, const double* rho,
                size_t nd, const fbls_bank* bank, const fbls_search_cfg* cfg,
                SearchPlan* p, bool lazy = false) {
  p->fam = f;
  p->grid.clear();
  p->bands.clear();
  if (nd == 0) return FBLS_OK;
  std::vector<std::pair<double, double>> ds(nd);
  for (size_t i = 0; i < nd; i++) {
    if (!(durations[i] > 0.0 && std::isfinite(durations[i])) ||
        !(rho[i] > 0.0 && std::isfinite(rho[i])))
      return fail(FBLS_E_ARG, "durations and rho must be positive and finite");
    ds[i] = {durations[i], rho[i]};
  }
  std::stable_sort(ds.begin(), ds.end(),
                   [](const std::pair<double, double>& a,
                      const std::pair<double, double>& b) { return a.first < b.first; });
  p->sd.resize(nd);
  p->sr.resize(nd);
  for (size_t i = 0; i < nd; i++) {
    p->sd[i] = ds[i].first;
    p->sr[i] = ds[i].second;
  }
  const double baseline = f->t_last;  // last relative time (scan.rs:425)
  const double maximum = std::min(cfg->period_max, baseline);
  const double minimum = std::max(cfg->period_min, 4.0 * p->sd[0] * (1.0 + 1e-9));
  if (maximum <= minimum) return FBLS_OK;
  size_t count = 0;
  if (int st = grid_count_checked(minimum, maximum, baseline, p->sd[0], cfg->period_oversample,
                                  &count))
    return st;
  p->count = count;
  p->formula = fbls::GridFormula(minimum, maximum, count);
  if (lazy) {
    p->bands = fbls::plan_bands(p->formula, p->sd.data(), nd, bank->phase_oversample,
                                bank->prescan_factor);
    return FBLS_OK;
  }
  p->grid.resize(count);
  double* g = p->grid.data();
  parallel_ranges(count, 8192, [&](size_t a, size_t e) {
    fbls::period_grid_range(minimum, maximum, g, count, a, e);
  });
  p->bands = fbls::plan_bands(p->grid.data(), count, p->sd.data(), nd,
                              bank->phase_oversample, bank->prescan_factor);
  return FBLS_OK;
}

// A restriction of the batch: a per-period mask (keep, offsets keep_at per
// star) or per-star bins (bin_keep, nbins, edges); both null for the full
// search.
struct Restriction {
  const uint8_t* keep = nullptr;
  const uint8_t* bin_keep = nullptr;
  size_t nbins = 0;
  const double* edges = nullptr;
  // ragged bins: star s has bins [bin_offsets[s], bin_offsets[s + 1]) of
  // bin_keep and its own edges at edges + bin_offsets[s] + s
  const size_t* bin_offsets = nullptr;
  bool any() const { return keep || bin_keep; }
};

// Plans stars [s0, s1) on up to 16 host threads under the restriction
// (keep_at[s - s0] is star s's offset in a per-period mask).
int plan_stars(const fbls_family* const* fams, size_t s0, size_t s1, const double* durations,
               const double* rho, size_t nd, const fbls_bank* bank, const fbls_search_cfg* cfg,
               const Restriction& rs, const std::vector<size_t>& keep_at,
               std::vector<SearchPlan>* plans) {
  const size_t m = s1 - s0;
  plans->assign(m, SearchPlan());
  std::vector<int> status(m, FBLS_OK);
  std::vector<std::string> message(m);
  // a worker records its stars' errors, C++ exceptions included (a worker
  // thread has no guard of its own: an escaping exception would terminate)
  auto plan_one = [&](size_t k) {
    const size_t s = s0 + k;
    if (!fams[s]) {
      status[k] = FBLS_E_ARG;
      message[k] = "null family " + std::to_string(s);
      return false;
    }
    status[k] = plan_search(fams[s], durations, rho + s * nd, nd, bank, cfg, &(*plans)[k],
                            rs.any());
    if (status[k]) {
      message[k] = fbls_last_error();
      return false;
    }
    if (rs.keep) apply_keep((*plans)[k], rs.keep + keep_at[k]);
    if (rs.bin_keep && rs.bin_offsets) {
      const size_t b0 = rs.bin_offsets[s];
      apply_bins((*plans)[k], rs.bin_keep + b0, rs.bin_offsets[s + 1] - b0, rs.edges + b0 + s);
    } else if (rs.bin_keep) {
      apply_bins((*plans)[k], rs.bin_keep + s * rs.nbins, rs.nbins, rs.edges);
    }
    return true;
  };
  auto work = [&](size_t a, size_t e) noexcept {
    for (size_t k = a; k < e; k++) {
      try {
        if (!plan_one(k)) return;
      } catch (const std::bad_alloc&) {
        status[k] = FBLS_E_OOM;
        message[k] = "host memory allocation failed";
        return;
      } catch (const std::exception& ex) {
        status[k] = FBLS_E_ARG;
        try {
          message[k] = ex.what();
        } catch (...) {
        }
        return;
      }
    }
  };
  const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), (m + 31) / 32);
  if (nt <= 1) {
    work(0, m);
  } else {
    // the threads that did start are joined before any error returns; the
    // ranges of threads that could not start are planned on this thread
    std::vector<std::thread> pool;
    pool.reserve(nt);
    size_t started = 0;
    try {
      for (; started < nt; started++)
        pool.emplace_back(work, m * started / nt, m * (started + 1) / nt);
    } catch (const std::system_error&) {
    }
    for (std::thread& th : pool) th.join();
    if (started < nt) work(m * started / nt, m);
  }
  for (size_t k = 0; k < m; k++)
    if (status[k]) return fail(status[k], "%s", message[k].c_str());
  return FBLS_OK;
}

void append_runs(const SearchPlan& p, std::vector<BandRun>& runs) {
  for (size_t b = 0; b < p.bands.size(); b++) {
    const fbls_band& band = p.bands[b];
    BandRun r;
    r.fam = p.fam;
    if (p.kept.empty()) {
      r.periods = p.grid.data() + band.first;
      r.np = band.count;
    } else {
      r.periods = p.kept[b].data();
      r.np = p.kept[b].size();
    }
    r.nb = band.nb;
    r.nb_coarse = band.nb_coarse;
    r.nd = (uint32_t)band.nd;
    r.dur = p.sd.data();
    r.rho = p.sr.data();
    runs.push_back(std::move(r));
  }
}

// Gpu::fold_budget_mib (scan.rs:93-98)
fbls_bank budgeted(const fbls_bank* bank) {
  fbls_bank b = *bank;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) == hipSuccess)
    b.max_fold_mib = std::min(b.max_fold_mib, (double)fr / 1048576.0 / 2.0);
  return b;
}

}  // namespace

// scan.rs:413-473
extern "C" int fbls_signal_search(fbls_ctx* c, const fbls_family* f,
                                  const double* durations, const double* rho,
                                  size_t nd, const fbls_bank* bank,
                                  const fbls_search_cfg* cfg, fbls_row* rows,
                                  size_t rows_cap, size_t* nrows,
                                  fbls_scan_stats* stats, size_t stats_cap,
                                  size_t* nbands, double* grid_out,
                                  size_t grid_cap, size_t* grid_count) {
  return fbls::host::guard([&]() -> int {
    if (!c || !f || !bank || !cfg || !nrows || !nbands || !grid_count)
      return fail(FBLS_E_ARG, "null argument");
    *nrows = 0;
    *nbands = 0;
    *grid_count = 0;
    if (int st = check_search_cfg(cfg)) return st;
    if (nd == 0) return FBLS_OK;
    if (!durations || !rho) return fail(FBLS_E_ARG, "null argument");
    SearchPlan p;
    int st = plan_search(f, durations, rho, nd, bank, cfg, &p);
    if (st) return st;
    const size_t count = p.grid.size();
    if (count == 0) return FBLS_OK;
    *grid_count = count;
    if (grid_out) {
      if (grid_cap < count) return fail(FBLS_E_ARG, "grid capacity %zu < %zu", grid_cap, count);
      std::memcpy(grid_out, p.grid.data(), count * sizeof(double));
    }
    if (rows_cap < count) return fail(FBLS_E_ARG, "rows capacity %zu < grid %zu", rows_cap, count);
    if (stats && p.bands.size() > stats_cap)
      return fail(FBLS_E_ARG, "band statistics capacity %zu < %zu bands", stats_cap,
                  p.bands.size());
    fbls_bank b = budgeted(bank);
    b.durations = p.sd.data();
    b.rho = p.sr.data();
    b.nd = nd;
    st = check_bank(&b);
    if (st) return st;
    std::vector<BandRun> runs;
    append_runs(p, runs);
    st = scan_bands(c, runs, &b, rows, nrows, nullptr, false);
    if (st) return st;
    for (size_t i = 0; i < runs.size(); i++)
      if (stats) stats[i] = runs[i].stats;
    *nbands = runs.size();
    return FBLS_OK;
  });
}

// Many stars (one family each) through the same two-launch band loop: the
// coarse passes of every band of every star in a chunk share one launch,
// their fine passes a second.  Chunks hold at most kBatchPeriods periods.
namespace {

int batch_search(fbls_ctx* c, const fbls_family* const* fams, size_t nfam,
                 const double* durations, const double* rho, size_t nd, const fbls_bank* bank,
                 const fbls_search_cfg* cfg, fbls_row* rows, size_t rows_cap,
                 size_t* row_offsets, fbls_scan_stats* star_stats, const Restriction& rs) {
  if (!bank || !cfg || !row_offsets) return fail(FBLS_E_ARG, "null argument");
  constexpr size_t kBatchPeriods = size_t(1) << 21;
  for (size_t k = 0; k <= nfam; k++) row_offsets[k] = 0;
  if (star_stats) std::memset(star_stats, 0, nfam * sizeof(fbls_scan_stats));
  // the configuration is checked before anything else touches the device
  if (int st = check_search_cfg(cfg)) return st;
  if (!c || (nfam && !fams) || (nd && (!durations || !rho)))
    return fail(FBLS_E_ARG, "null argument");
  if (nd == 0 || nfam == 0) return FBLS_OK;
  fbls_bank b = budgeted(bank);
  size_t written = 0, needed = 0;
  size_t s0 = 0, keep_at = 0;
  // grid sizes first: they set the chunks and each star's keep offset
  std::vector<size_t> grid_size(nfam, 0);
  {
    double dmin = INFINITY;
    for (size_t i = 0; i < nd; i++) dmin = std::min(dmin, durations[i]);
    for (size_t s = 0; s < nfam; s++) {
      if (!fams[s]) return fail(FBLS_E_ARG, "null family %zu", s);
      const double baseline = fams[s]->t_last;
      const double maximum = std::min(cfg->period_max, baseline);
      const double minimum = std::max(cfg->period_min, 4.0 * dmin * (1.0 + 1e-9));
      if (maximum > minimum)
        if (int st = grid_count_checked(minimum, maximum, baseline, dmin,
                                        cfg->period_oversample, &grid_size[s]))
          return st;
    }
  }
  // a restricted search plans every star first (its grids are never
  // materialised) and chunks by the periods actually scanned, so a pruned
  // batch needs few launches; the full search plans chunk by chunk
  std::vector<SearchPlan> all;
  std::vector<size_t> scanned = grid_size;
  if (rs.any()) {
    std::vector<size_t> off(nfam, 0);
    for (size_t s = 1; s < nfam; s++) off[s] = off[s - 1] + grid_size[s - 1];
    if (int st = plan_stars(fams, 0, nfam, durations, rho, nd, bank, cfg, rs, off, &all))
      return st;
    for (size_t s = 0; s < nfam; s++) {
      scanned[s] = 0;
      for (const auto& k : all[s].kept) scanned[s] += k.size();
    }
  }
  while (s0 < nfam) {
    // one chunk of stars
    std::vector<SearchPlan> plans;
    std::vector<BandRun> runs;
    std::vector<size_t> run_star;
    size_t periods = 0, s1 = s0;
    std::vector<size_t> keep_off;
    while (s1 < nfam && (s1 == s0 || periods < kBatchPeriods)) {
      keep_off.push_back(keep_at);
      keep_at += grid_size[s1];
      periods += scanned[s1];
      s1++;
    }
    if (rs.any()) {
      for (size_t k = s0; k < s1; k++) plans.push_back(std::move(all[k]));
    } else if (int st = plan_stars(fams, s0, s1, durations, rho, nd, bank, cfg, rs, keep_off,
                                   &plans)) {
      return st;
    }
    for (size_t k = 0; k < plans.size(); k++)
      if (plans[k].count != grid_size[s0 + k])
        return fail(FBLS_E_ARG, "grid size of star %zu changed while planning", s0 + k);
    for (size_t k = 0; k < plans.size(); k++) {
      const size_t before = runs.size();
      append_runs(plans[k], runs);
      for (size_t i = before; i < runs.size(); i++) run_star.push_back(s0 + k);
    }
    // every plan shares the sorted bank durations (rho differs per star)
    if (!plans.empty()) {
      b.durations = plans[0].sd.data();
      b.rho = plans[0].sr.data();
      b.nd = nd;
      int st = check_bank(&b);
      if (st) return st;
    }
    std::unique_ptr<fbls_row[]> chunk_rows(new fbls_row[periods ? periods : 1]);
    std::vector<size_t> counts(runs.size(), 0);
    size_t nchunk = 0;
    if (!runs.empty()) {
      int st = scan_bands(c, runs, &b, chunk_rows.get(), &nchunk, counts.data(), rs.any());
      if (st) return st;
    }
    size_t at = 0;
    for (size_t i = 0; i < runs.size(); i++) {
      const size_t s = run_star[i];
      row_offsets[s + 1] += counts[i];
      for (size_t k = 0; k < counts[i]; k++, at++) {
        if (written < rows_cap && rows) rows[written] = chunk_rows[at];
        written++;
      }
      if (star_stats) {
        fbls_scan_stats& a = star_stats[s];
        const fbls_scan_stats& r = runs[i].stats;
        a.periods += r.periods;
        a.prescan_survivors += r.prescan_survivors;
        a.phase_bins = std::max(a.phase_bins, r.phase_bins);
        a.batches += r.batches;
        a.templates_scanned += r.templates_scanned;
        a.exact_tile_retries += r.exact_tile_retries;
        a.seconds += r.seconds;
        a.seconds_fold += r.seconds_fold;
        a.seconds_cumulative += r.seconds_cumulative;
        a.seconds_scan += r.seconds_scan;
        a.seconds_device += r.seconds_device;
      }
    }
    needed += nchunk;
    s0 = s1;
  }
  for (size_t k = 0; k < nfam; k++) row_offsets[k + 1] += row_offsets[k];
  if (needed > rows_cap)
    return fail(FBLS_E_ARG, "rows capacity %zu < %zu rows (offsets filled)", rows_cap, needed);
  return FBLS_OK;
}

}  // namespace

extern "C" int fbls_signal_search_batch(fbls_ctx* c, const fbls_family* const* fams,
                                        size_t nfam, const double* durations,
                                        const double* rho, size_t nd,
                                        const fbls_bank* bank,
                                        const fbls_search_cfg* cfg, fbls_row* rows,
                                        size_t rows_cap, size_t* row_offsets,
                                        fbls_scan_stats* star_stats,
                                        const uint8_t* keep) {
  return fbls::host::guard([&]() -> int {
    Restriction rs;
    rs.keep = keep;
    return batch_search(c, fams, nfam, durations, rho, nd, bank, cfg, rows, rows_cap, row_offsets,
                        star_stats, rs);
  });
}

extern "C" int fbls_signal_search_bins(fbls_ctx* c, const fbls_family* const* fams, size_t nfam,
                                       const double* durations, const double* rho, size_t nd,
                                       const fbls_bank* bank, const fbls_search_cfg* cfg,
                                       fbls_row* rows, size_t rows_cap, size_t* row_offsets,
                                       fbls_scan_stats* star_stats, const uint8_t* bin_keep,
                                       size_t nbins, const double* edges) {
  return fbls::host::guard([&]() -> int {
    if (!bin_keep || !edges || nbins == 0) return fail(FBLS_E_ARG, "null argument");
    for (size_t k = 0; k < nbins; k++)
      if (!(edges[k + 1] > edges[k])) return fail(FBLS_E_ARG, "bin edges must ascend");
    Restriction rs;
    rs.bin_keep = bin_keep;
    rs.nbins = nbins;
    rs.edges = edges;
    return batch_search(c, fams, nfam, durations, rho, nd, bank, cfg, rows, rows_cap, row_offsets,
                        star_stats, rs);
  });
}

extern "C" int fbls_signal_search_bins_ragged(
    fbls_ctx* c, const fbls_family* const* fams, size_t nfam, const double* durations,
    const double* rho, size_t nd, const fbls_bank* bank, const fbls_search_cfg* cfg,
    fbls_row* rows, size_t rows_cap, size_t* row_offsets, fbls_scan_stats* star_stats,
    const uint8_t* bin_keep, const size_t* bin_offsets, const double* edges) {
  return fbls::host::guard([&]() -> int {
    if (!bin_keep || !bin_offsets || !edges) return fail(FBLS_E_ARG, "null argument");
    if (bin_offsets[0] != 0) return fail(FBLS_E_ARG, "bin offsets must start at 0");
    for (size_t s = 0; s < nfam; s++) {
      const size_t b0 = bin_offsets[s], b1 = bin_offsets[s + 1];
      if (b1 <= b0) return fail(FBLS_E_ARG, "star %zu has no bins", s);
      const double* e = edges + b0 + s;
      for (size_t k = 0; k < b1 - b0; k++)
        if (!(e[k + 1] > e[k])) return fail(FBLS_E_ARG, "bin edges of star %zu must ascend", s);
    }
    Restriction rs;
    rs.bin_keep = bin_keep;
    rs.edges = edges;
    rs.bin_offsets = bin_offsets;
    return batch_search(c, fams, nfam, durations, rho, nd, bank, cfg, rows, rows_cap, row_offsets,
                        star_stats, rs);
  });
}

// ------------------------------------------------------------ host helpers

extern "C" int fbls_family_state(const double* time, const double* flux,
                                 const double* sigma, size_t n,
                                 const double* durations, size_t nd,
                                 int correlated, double* t, double* v,
                                 double* w, size_t* out_n,
                                 double* baseline_norm, double* origin,
                                 double* rho, double* inflation,
                                 size_t* segments, double* residual_energy) {
  return fbls::host::guard([&]() -> int {
    std::string msg;
    double energy = 0.0;
    int st = fbls::family_state(time, flux, sigma, n, durations, nd, correlated,
                                t, v, w, out_n, baseline_norm, origin, rho,
                                inflation, segments, &energy, &msg);
    if (residual_energy) *residual_energy = energy;
    if (st) return fail(st, "%s", msg.c_str());
    return FBLS_OK;
  });
}

extern "C" int fbls_family_state_many(
    size_t count, const double* const* time, const double* const* flux,
    const double* const* sigma, const size_t* n, double flux_offset,
    const double* durations, size_t nd, int correlated, double* const* t, double* const* v,
    double* const* w, size_t* out_n, double* baseline_norm, double* origin, double* rho,
    double* inflation, size_t* segments, double* residual_energy, int* status) {
  return fbls::host::guard([&]() -> int {
    if (count == 0) return FBLS_OK;
    if (!time || !flux || !sigma || !n || !t || !v || !w || !out_n || !baseline_norm ||
        !origin || !inflation || !segments || (nd && (!durations || !rho)))
      return fail(FBLS_E_ARG, "null argument");
    std::vector<int> codes(count, FBLS_OK);
    std::vector<std::string> msgs(count);
    // one star per piece: the threads take the stars in order as they free
    Pool::background().run(count, count, [&](size_t a, size_t e) {
      for (size_t k = a; k < e; k++) {
        try {
          double energy = 0.0;
          codes[k] = fbls::family_state(time[k], flux[k], sigma[k], n[k], durations, nd,
                                        correlated, t[k], v[k], w[k], out_n + k,
                                        baseline_norm + k, origin + k, rho + k * nd,
                                        inflation + k, segments + k, &energy, &msgs[k],
                                        flux_offset);
          if (residual_energy) residual_energy[k] = energy;
        } catch (const std::bad_alloc&) {
          codes[k] = FBLS_E_OOM;
          msgs[k] = "host memory allocation failed";
        } catch (const std::exception& x) {
          codes[k] = FBLS_E_ARG;
          msgs[k] = x.what();
        }
      }
    });
    size_t first = count;
    for (size_t k = 0; k < count; k++) {
      if (status) status[k] = codes[k];
      if (codes[k] && first == count) first = k;
    }
    if (first < count) return fail(codes[first], "star %zu: %s", first, msgs[first].c_str());
    return FBLS_OK;
  });
}

extern "C" int fbls_synth_star(const fbls_synth_cfg* cfg, double* time,
                               double* flux, double* flux_err, size_t cap,
                               size_t* n) {
  return fbls::host::guard([&]() -> int {
    std::string msg;
    int st = fbls::synth_star(cfg, time, flux, flux_err, cap, n, &msg);
    if (st) return fail(st, "%s", msg.c_str());
    return FBLS_OK;
  });
}

// ---------------------------------------------------------------- autotune

// Times fbls_signal_search of one family under every candidate launch shape
// (LDS budget per workgroup, which sets the periods per workgroup, and the
// workgroup size), checks that every shape returns bit-identical rows, and
// keeps the fastest on the context.
extern "C" int fbls_autotune(fbls_ctx* c, const fbls_family* f, const double* durations,
                             const double* rho, size_t nd, const fbls_bank* bank,
                             const fbls_search_cfg* cfg, uint32_t repeats,
                             fbls_tune_entry* table, size_t table_cap, size_t* ntable,
                             fbls_tune_entry* best) {
  return fbls::host::guard([&]() -> int {
    if (!c || !f || !bank || !cfg || !ntable || !best) return fail(FBLS_E_ARG, "null argument");
    *ntable = 0;
    if (int st = check_search_cfg(cfg)) return st;
    const uint32_t lds_opts[] = {16384, 32768, 49152, 65536, 81920, 98304, 163840};
    const uint32_t block_opts[] = {256, 512, 1024};
    const uint32_t saved_lds = c->lds_override, saved_block = c->block_override;
    const double d0 = *std::min_element(durations, durations + nd);
    const double hi = std::min(cfg->period_max, f->t_last);
    const double lo = std::max(cfg->period_min, 4.0 * d0 * (1.0 + 1e-9));
    const size_t cap = hi > lo ? std::max<size_t>(1, fbls::period_grid_count(
                                                         lo, hi, f->t_last, d0,
                                                         cfg->period_oversample))
                               : 1;
    std::vector<fbls_row> rows(cap), ref;
    std::vector<fbls_scan_stats> st(64);
    size_t nrows = 0, nbands = 0, gc = 0;
    best->seconds = INFINITY;
    int status = FBLS_OK;
    for (uint32_t block : block_opts) {
      for (uint32_t lds : lds_opts) {
        if (lds > c->max_lds) continue;
        c->lds_override = lds;
        c->block_override = block;
        double secs = INFINITY;
        for (uint32_t r = 0; r <= std::max<uint32_t>(repeats, 1); r++) {
          fbls_launch_totals before = c->totals;
          status = fbls_signal_search(c, f, durations, rho, nd, bank, cfg, rows.data(), cap,
                                      &nrows, st.data(), st.size(), &nbands, nullptr, 0, &gc);
          if (status) break;
          if (r > 0) secs = std::min(secs, c->totals.seconds - before.seconds);  // r = 0 warms up
        }
        if (status) break;
        if (ref.empty()) {
          ref.assign(rows.begin(), rows.begin() + nrows);
        } else if (nrows != ref.size() ||
                   std::memcmp(ref.data(), rows.data(), nrows * sizeof(fbls_row)) != 0) {
          status = fail(FBLS_E_HIP, "launch shape lds %u block %u changed the rows", lds, block);
          break;
        }
        fbls_tune_entry e{lds, block, secs};
        if (table && *ntable < table_cap) table[*ntable] = e;
        (*ntable)++;
        if (secs < best->seconds) *best = e;
      }
      if (status) break;
    }
    if (status) {
      c->lds_override = saved_lds;
      c->block_override = saved_block;
      return status;
    }
    c->lds_override = best->lds_bytes;
    c->block_override = best->block;
    return FBLS_OK;
  });
}
