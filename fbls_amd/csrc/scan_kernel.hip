// Fused FBLS period scan for gfx950 (MI355X).
//
// A launch carries segments (one family at one phase-bin count over a list
// of periods: the work of one reference scan_bank call).  Each workgroup owns
// G periods of one segment.  It folds every observation of the family into
// per-period LDS histograms (fixed-point values, so the sums are exact and
// order-independent), turns them into exclusive prefix sums in place, scans
// every (phase bin, duration, shape) template with an FP64 surrogate of the
// squared merit, resolves near ties with the exact FP64 merit, and writes one
// record per period.  Nothing but the record touches HBM.
//
// Reference semantics restated here (isakruas/fbls):
//   phase cell           src/search/kernels.rs:314-327
//   fold                 src/search/kernels.rs:464-553
//   doubled prefix       src/search/kernels.rs:237-309 (cum[nb+b+1] = cum[b+1] + total)
//   template scan        src/search/kernels.rs:608-845
//     k = min(max(1, floor(D*scale + 0.5)), nb/2); usable j < nb,
//     raw_k < nb/2, 3 <= N <= n-3; valid usable && numerator > 0
//     shapes only in 128-bin tiles where a valid box merit >= gate
//   staircase sums       src/search/kernels.rs:84-113
//   exact merit          src/search/kernels.rs:427-455 (rank_exact)
//   winner / tie rule    larger merit, then smaller key, then lower lane
//                        (kernels.rs:154-198, 1134-1143)
//   winner recompute     src/search/kernels.rs:800-843
//
// LDS per period: vw[nb+1] {v, w} double pairs (16 B) then n[nb+1] u32;
// slot b+1 accumulates bin b, so after the in-place scan slot x holds the
// exclusive prefix E[x] = sum of bins < x.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scan_kernel.h"

namespace fbls {

namespace {

constexpr uint32_t kTile = 128;       // gate granularity (kernels.rs:26)
constexpr uint32_t kObsPerLane = 8;   // contiguous observations per lane
constexpr uint32_t kObsPerWave = 64 * kObsPerLane;
constexpr double kInf = __builtin_inf();
constexpr double kNear = 1e-10;       // near-tie window of the surrogate

struct Cum {
  const double2* vw;  // exclusive prefix of {v, w}, dequantised
  const uint32_t* n;  // exclusive prefix of counts
  double tv, tw;      // totals E[nb]
  uint32_t tn, nb;
};

// doubled prefix (kernels.rs:7-10): cum[x] for x in [0, 2nb]
__device__ __forceinline__ double2 cum_vw(const Cum& c, uint32_t x) {
  if (x <= c.nb) return c.vw[x];
  double2 e = c.vw[x - c.nb];
  return make_double2(e.x + c.tv, e.y + c.tw);
}

__device__ __forceinline__ uint32_t cum_n(const Cum& c, uint32_t x) {
  return x <= c.nb ? c.n[x] : c.n[x - c.nb] + c.tn;
}

// Reciprocal with two Newton steps: relative error far below kNear.
__device__ __forceinline__ double rcp_nr(double x) {
  double y = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, y, 1.0);
  y = __builtin_fma(y, e, y);
  e = __builtin_fma(-x, y, 1.0);
  return __builtin_fma(y, e, y);
}

// rank_exact's merit (kernels.rs:439-447), IEEE division and square root.
__device__ __forceinline__ double exact_merit(double a, double hp, double norm,
                                              double rho, double c,
                                              bool* bad) {
  double d = hp * hp / c;
  double q = norm - d;
  if (q <= 0.0) *bad = true;
  double cand = (a / rho) / __builtin_sqrt(__builtin_fmax(q, 1e-300));
  return (q > 0.0 && cand < kInf) ? cand : -kInf;
}

// Squared-merit surrogate of a valid template; exact fallback when the
// precision is small relative to the norm or the ratio is out of range.
__device__ __forceinline__ double surrogate(double a, double hp, double norm,
                                            double rho, double rho2, double c,
                                            double inv_c, bool* bad) {
  double q = norm - hp * hp * inv_c;
  if (norm > 0.0 && q > 1e-3 * norm) {
    double s = a * a * rcp_nr(rho2 * q);
    if (s < 1e300) return s;
  }
  double m = exact_merit(a, hp, norm, rho, c, bad);
  return m > -kInf ? m * m : -kInf;
}

struct LaneState {
  double s, second;         // best surrogate and best other candidate
  double a, hp, norm, rho;  // inputs of the best (identical inputs tie exactly)
  uint64_t key;
  uint32_t di;
};

__device__ __forceinline__ void state_init(LaneState& st) {
  st.s = -kInf;
  st.second = -kInf;
  st.a = st.hp = st.norm = st.rho = 0.0;
  st.key = ~0ull;
  st.di = 0;
}

__device__ __forceinline__ void state_update(LaneState& st, double s,
                                             uint64_t key, uint32_t di,
                                             double a, double hp, double norm,
                                             double rho) {
  if (!(s > -kInf)) return;
  if (st.s > -kInf && a == st.a && hp == st.hp && norm == st.norm &&
      rho == st.rho) {
    // bit-identical inputs: the exact merits are equal, smaller key wins
    if (key < st.key) {
      st.key = key;
      st.di = di;
    }
    return;
  }
  if (s > st.s || (s == st.s && key < st.key)) {
    st.second = st.s;
    st.s = s;
    st.key = key;
    st.di = di;
    st.a = a;
    st.hp = hp;
    st.norm = norm;
    st.rho = rho;
  } else if (s > st.second) {
    st.second = s;
  }
}

// Staircase sums (kernels.rs:84-113) with precomputed layer widths/offsets.
__device__ __forceinline__ void shape_sums(const Cum& c, uint32_t jj, double a0,
                                           double w0, uint32_t layers,
                                           const uint32_t* tab, double* pa,
                                           double* php, double* pnorm) {
  double l = (double)layers;
  double a = a0 / l;
  double hp = w0 / l;
  double norm = w0 / (l * l);
  for (uint32_t layer = 1; layer < layers; layer++) {
    uint32_t e = tab[layer - 1];
    uint32_t width = e & 0xffffu;
    uint32_t start = jj + (e >> 16);
    double2 hi = cum_vw(c, start + width);
    double2 lo = cum_vw(c, start);
    double dw = hi.y - lo.y;
    double dv = hi.x - lo.x;
    a += dv / l;
    hp += dw / l;
    norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

// Staircase sums with widths computed from the factors (record recompute).
__device__ void shape_sums_direct(const Cum& c, uint32_t j, uint32_t k,
                                  double a0, double w0, double ingress,
                                  const double* factors, uint32_t si,
                                  uint32_t layers, double* pa, double* php,
                                  double* pnorm) {
  if (ingress == 0.0) {
    *pa = a0;
    *php = w0;
    *pnorm = w0;
    return;
  }
  double l = (double)layers;
  double a = a0 / l;
  double hp = w0 / l;
  double norm = w0 / (l * l);
  for (uint32_t layer = 1; layer < layers; layer++) {
    double factor = factors[si * layers + layer];
    double wf = __builtin_floor((double)k * factor + 0.5);
    uint32_t width = wf < 1.0 ? 1u : (uint32_t)wf;
    uint32_t start = j + (k - width) / 2;
    double2 hi = cum_vw(c, start + width);
    double2 lo = cum_vw(c, start);
    double dw = hi.y - lo.y;
    a += (hi.x - lo.x) / l;
    hp += dw / l;
    norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

__device__ __forceinline__ double wave_max(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = __builtin_fmax(x, __shfl_xor(x, off, 64));
  return x;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    uint64_t y = __shfl_xor(x, off, 64);
    x = y < x ? y : x;
  }
  return x;
}

template <typename T>
__device__ __forceinline__ T wave_excl_scan(T x, uint32_t lane) {
  T incl = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  return incl - x;
}

// In-place inclusive scan of slots [1, nb] of {v, w} pairs by one wave
// (exact: integer-valued doubles below 2^53), then the power-of-two scales.
__device__ void wave_prefix_vw(double2* arr, uint32_t nb, uint32_t lane,
                               double sv, double sw) {
  uint32_t len = (nb + 63) / 64;
  uint32_t lo = 1 + lane * len;
  uint32_t hi = lo + len;
  if (hi > nb + 1) hi = nb + 1;
  double av = 0.0, aw = 0.0;
  for (uint32_t s = lo; s < hi; s++) {
    double2 x = arr[s];
    av += x.x;
    aw += x.y;
    arr[s] = make_double2(av, aw);
  }
  const bool has = lo < hi;
  double ev = wave_excl_scan(has ? av : 0.0, lane);
  double ew = wave_excl_scan(has ? aw : 0.0, lane);
  for (uint32_t s = lo; s < hi; s++) {
    double2 x = arr[s];
    arr[s] = make_double2((x.x + ev) * sv, (x.y + ew) * sw);
  }
}

__device__ void wave_prefix_n(uint32_t* arr, uint32_t nb, uint32_t lane) {
  uint32_t len = (nb + 63) / 64;
  len |= 1u;  // odd stride between lanes spreads LDS banks
  uint32_t lo = 1 + lane * len;
  uint32_t hi = lo + len;
  if (hi > nb + 1) hi = nb + 1;
  uint32_t acc = 0;
  for (uint32_t s = lo; s < hi; s++) {
    acc += arr[s];
    arr[s] = acc;
  }
  uint32_t ex = wave_excl_scan(lo < hi ? acc : 0u, lane);
  for (uint32_t s = lo; s < hi; s++) arr[s] += ex;
}

// floor(fl(t * scale)) mod nb (kernels.rs:314-327).  Narrow: t*scale < 2^24,
// the quotient from a float reciprocal biased low by ~2^-22 is floor(u/nb)
// or one less, so one conditional subtraction (as unsigned min) finishes.
template <bool WIDE>
__device__ __forceinline__ uint32_t phase_bin(double t, double scale,
                                              uint32_t nb, float inv_lo) {
  const uint32_t u = (uint32_t)(t * scale);
  if (WIDE) return u % nb;
  const uint32_t q = (uint32_t)((float)u * inv_lo);
  const uint32_t r = u - __umul24(q, nb);
  return min(r, r - nb);
}

__device__ __forceinline__ void lds_add(double* vw, uint32_t* n, uint32_t b,
                                        double v, double w, uint32_t c) {
  __hip_atomic_fetch_add(vw + 2 * (b + 1), v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(vw + 2 * (b + 1) + 1, w, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(n + b + 1, c, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Fold one lane's 8 consecutive observations into one period's histogram.
// RUN: consecutive observations falling in the same bin are summed in
// registers and flushed once (exact: integer-valued doubles).
template <bool WIDE, bool RUN, bool FULL>
__device__ __forceinline__ void fold_lane(double* vw, uint32_t* n,
                                          const double* tt, const double* vv,
                                          const double* ww, double scale,
                                          uint32_t nb, float inv_lo,
                                          uint32_t base, uint32_t nobs) {
  if (RUN && FULL) {
    uint32_t cb = phase_bin<WIDE>(tt[0], scale, nb, inv_lo);
    double av = vv[0], aw = ww[0];
    uint32_t an = 1;
#pragma unroll
    for (uint32_t r = 1; r < kObsPerLane; r++) {
      const uint32_t b = phase_bin<WIDE>(tt[r], scale, nb, inv_lo);
      const bool flush = b != cb;
      if (flush) lds_add(vw, n, cb, av, aw, an);
      const double sv = av + vv[r], sw = aw + ww[r];
      av = flush ? vv[r] : sv;
      aw = flush ? ww[r] : sw;
      an = flush ? 1u : an + 1u;
      cb = b;
    }
    lds_add(vw, n, cb, av, aw, an);
  } else {
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) {
      const uint32_t b = phase_bin<WIDE>(tt[r], scale, nb, inv_lo);
      if (FULL || base + r < nobs) lds_add(vw, n, b, vv[r], ww[r], 1u);
    }
  }
}

template <bool WIDE, bool RUN>
__device__ void fold_all(const Segment& S, unsigned char* hist,
                         uint32_t p0, uint32_t gcount, uint32_t wave,
                         uint32_t lane, uint32_t nwaves) {
  const uint32_t nb = S.nb;
  const uint32_t nb1 = nb + 1;
  const float inv_lo = (float)((1.0 / (double)nb) * (1.0 - 0x1p-22));
  const uint32_t nblk = S.n_pad / kObsPerWave;
  for (uint32_t blk = wave; blk < nblk; blk += nwaves) {
    const uint32_t base = blk * kObsPerWave + lane * kObsPerLane;
    double tt[kObsPerLane], vv[kObsPerLane], ww[kObsPerLane];
    const double2* tp = (const double2*)(S.t + base);
    const int4* qp = (const int4*)(S.vw + 2 * (size_t)base);
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane / 2; r++) {
      double2 d = tp[r];
      int4 q = qp[r];
      tt[2 * r] = d.x;
      tt[2 * r + 1] = d.y;
      vv[2 * r] = (double)q.x;
      ww[2 * r] = (double)q.y;
      vv[2 * r + 1] = (double)q.z;
      ww[2 * r + 1] = (double)q.w;
    }
    const bool full = blk * kObsPerWave + kObsPerWave <= S.n;
    for (uint32_t g = 0; g < gcount; g++) {
      const double scale = S.scales[p0 + g];
      double* vw = (double*)(hist + (size_t)g * S.pstride);
      uint32_t* n = (uint32_t*)(vw + 2 * nb1);
      if (full)
        fold_lane<WIDE, RUN, true>(vw, n, tt, vv, ww, scale, nb, inv_lo, base, S.n);
      else
        fold_lane<WIDE, RUN, false>(vw, n, tt, vv, ww, scale, nb, inv_lo, base, S.n);
    }
  }
}

__device__ __forceinline__ uint32_t find_segment(const LaunchArgs& A,
                                                 uint32_t group) {
  uint32_t lo = 0, hi = A.nseg;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) / 2;
    if (A.segs[mid].group0 <= group) lo = mid; else hi = mid;
  }
  return lo;
}

template <uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void fused_scan(LaunchArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  constexpr uint32_t kWaves = BLOCK / 64;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const Segment S = A.segs[find_segment(A, blockIdx.x)];
  const uint32_t nb = S.nb;
  const uint32_t nb1 = nb + 1;
  const uint32_t p0 = (blockIdx.x - S.group0) * S.G;
  const uint32_t gcount = min(S.G, S.np - p0);
  const uint32_t nd = S.nd, ns = A.ns, layers = A.layers;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  const uint32_t nsl = (ns - 1) * (layers - 1);  // table entries per duration

  // ---- LDS carve-up ---------------------------------------------------
  unsigned char* p = lds;
  unsigned char* hist = p;
  p += (size_t)S.G * S.pstride;
  uint32_t* ktab = (uint32_t*)p;  // [G][nd]: k | kok << 31
  p += S.G * nd * 4;
  uint32_t* stab = (uint32_t*)p;  // [G][nd][nsl]: width | offset << 16
  p += S.G * nd * nsl * 4;
  p = (unsigned char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  double* tval = (double*)p;  // [G][tiles]
  p += S.G * tiles * 8;
  uint64_t* tkey = (uint64_t*)p;
  p += S.G * tiles * 8;
  uint32_t* tdi = (uint32_t*)p;

  // ---- phase 0: clear histograms, duration tables ----------------------
  {
    uint4* h4 = (uint4*)hist;
    const uint32_t words = (uint32_t)((size_t)gcount * S.pstride / 16);
    for (uint32_t i = tid; i < words; i += BLOCK) h4[i] = make_uint4(0, 0, 0, 0);
    for (uint32_t i = tid; i < gcount * nd; i += BLOCK) {
      uint32_t g = i / nd, di = i % nd;
      double scale = S.scales[p0 + g];
      double rk = __builtin_floor(S.dur[di] * scale + 0.5);
      uint32_t raw_k = rk < 1.0 ? 1u : (rk >= 2147483648.0 ? 0x7fffffffu : (uint32_t)rk);
      uint32_t k = raw_k < nb / 2 ? raw_k : nb / 2;
      uint32_t kok = raw_k < nb / 2 ? 1u : 0u;
      ktab[i] = k | (kok << 31);
      for (uint32_t si = 1; si < ns; si++) {
        for (uint32_t l = 1; l < layers; l++) {
          double factor = A.factors[si * layers + l];
          double wf = __builtin_floor((double)k * factor + 0.5);
          uint32_t width = wf < 1.0 ? 1u : (uint32_t)wf;
          uint32_t off = (k - width) / 2;
          stab[(i * (ns - 1) + (si - 1)) * (layers - 1) + (l - 1)] = width | (off << 16);
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 1: fold every observation into the G histograms ----------
  if (S.wide_cells) {
    if (A.run_fold) fold_all<true, true>(S, hist, p0, gcount, wave, lane, kWaves);
    else fold_all<true, false>(S, hist, p0, gcount, wave, lane, kWaves);
  } else {
    if (A.run_fold) fold_all<false, true>(S, hist, p0, gcount, wave, lane, kWaves);
    else fold_all<false, false>(S, hist, p0, gcount, wave, lane, kWaves);
  }
  __syncthreads();

  // ---- phase 2: exclusive prefix sums in place (exact) -----------------
  for (uint32_t item = wave; item < 2 * gcount; item += kWaves) {
    uint32_t g = item >> 1;
    double2* vw = (double2*)(hist + (size_t)g * S.pstride);
    if (item & 1)
      wave_prefix_n((uint32_t*)(vw + nb1), nb, lane);
    else
      wave_prefix_vw(vw, nb, lane, S.sv, S.sw);
  }
  __syncthreads();

  // ---- phase 3: template scan, one 128-bin tile per wave at a time -----
  const double c = S.C, inv_c = S.invC;
  const uint32_t nobs = S.n;
  const double gate = A.gate;
  const double g2 = gate * gate;
  for (uint32_t item = wave; item < gcount * tiles; item += kWaves) {
    const uint32_t g = item / tiles, tile = item % tiles;
    Cum cm;
    cm.vw = (const double2*)(hist + (size_t)g * S.pstride);
    cm.n = (const uint32_t*)(cm.vw + nb1);
    cm.nb = nb;
    {
      double2 t = cm.vw[nb];
      cm.tv = t.x;
      cm.tw = t.y;
      cm.tn = cm.n[nb];
    }
    const uint32_t* kt = ktab + g * nd;
    const uint32_t* st_tab = stab + (size_t)g * nd * nsl;

    uint32_t jx[2], jjx[2];
    double ev[2], ew[2];
    uint32_t en[2];
    LaneState st[2];
#pragma unroll
    for (int x = 0; x < 2; x++) {
      jx[x] = tile * kTile + lane + 64 * x;
      jjx[x] = jx[x] < nb ? jx[x] : nb - 1;
      double2 e = cm.vw[jjx[x]];
      ev[x] = e.x;
      ew[x] = e.y;
      en[x] = cm.n[jjx[x]];
      state_init(st[x]);
    }
    bool bad = false;
    uint32_t gmask = 0;
    for (uint32_t di = 0; di < nd; di++) {
      const uint32_t kte = kt[di];
      const uint32_t k = kte & 0x7fffffffu;
      const bool kok = (kte >> 31) != 0;
      const double rho = S.rho[di];
      const double rho2 = rho * rho;
      double bv[2], bw[2];
      bool busable[2];
      bool pass = false;
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const uint32_t e = jjx[x] + k;
        const uint32_t cnt = cum_n(cm, e) - en[x];
        const double2 hi = cum_vw(cm, e);
        const double w = hi.y - ew[x];
        const double a = hi.x - ev[x];
        const bool usable = jx[x] < nb && kok && cnt >= 3 && cnt + 3 <= nobs;
        const bool valid = usable && a > 0.0;
        bv[x] = a;
        bw[x] = w;
        busable[x] = usable;
        if (valid) {
          const uint64_t key = (uint64_t)k * nb1 + jjx[x];
          const double s = surrogate(a, w, w, rho, rho2, c, inv_c, &bad);
          state_update(st[x], s, key, di, a, w, w, rho);
          if (s > -kInf) {
            if (gate <= 0.0 || s >= g2 * (1.0 + kNear)) {
              pass = true;
            } else if (s >= g2 * (1.0 - kNear)) {
              pass = pass || exact_merit(a, w, w, rho, c, &bad) >= gate;
            }
          }
        }
      }
      if (!__any(pass)) continue;
      gmask |= 1u << di;
      for (uint32_t si = 1; si < ns; si++) {
        const uint32_t* tab = st_tab + (di * (ns - 1) + (si - 1)) * (layers - 1);
        const bool staircase = A.ingress[si] != 0.0;
#pragma unroll
        for (int x = 0; x < 2; x++) {
          double a = bv[x], hp = bw[x], norm = bw[x];
          if (staircase)
            shape_sums(cm, jjx[x], bv[x], bw[x], layers, tab, &a, &hp, &norm);
          if (busable[x] && a > 0.0) {
            const uint64_t key = ((uint64_t)si * nb1 + k) * nb1 + jjx[x];
            const double s = surrogate(a, hp, norm, rho, rho2, c, inv_c, &bad);
            state_update(st[x], s, key, di, a, hp, norm, rho);
          }
        }
      }
    }

    // ---- tile winner: surrogate top, exact value, near-tie rescan -------
    const double m = wave_max(__builtin_fmax(st[0].s, st[1].s));
    double tile_value = -kInf;
    uint64_t tile_key = ~0ull;
    uint32_t tile_di = 0;
    if (m > -kInf) {
      const double thr = m * (1.0 - kNear);
      bool near[2];
      uint64_t kk = ~0ull;
#pragma unroll
      for (int x = 0; x < 2; x++) {
        near[x] = st[x].s >= thr;
        if (st[x].s == m && st[x].key < kk) kk = st[x].key;
      }
      const uint64_t top_key = wave_min_u64(kk);
      const uint64_t owner_mask =
          __ballot(st[0].key == top_key || st[1].key == top_key);
      const int owner = __builtin_ctzll(owner_mask);
      const int ox = __shfl((st[1].key == top_key) ? 1 : 0, owner, 64);
      const double ta = __shfl(ox ? st[1].a : st[0].a, owner, 64);
      const double thp = __shfl(ox ? st[1].hp : st[0].hp, owner, 64);
      const double tno = __shfl(ox ? st[1].norm : st[0].norm, owner, 64);
      const double trh = __shfl(ox ? st[1].rho : st[0].rho, owner, 64);
      const uint32_t tdi0 = __shfl(ox ? st[1].di : st[0].di, owner, 64);
      bool need = false;
#pragma unroll
      for (int x = 0; x < 2; x++) {
        if (near[x]) {
          const bool same = st[x].a == ta && st[x].hp == thp &&
                            st[x].norm == tno && st[x].rho == trh;
          need = need || !same || st[x].second >= st[x].s * (1.0 - kNear);
        }
      }
      if (!__any(need)) {
        tile_key = top_key;
        tile_di = tdi0;
        tile_value = exact_merit(ta, thp, tno, trh, c, &bad);
      } else {
        // exact rescan of the near states (kernels.rs:662-798 retry pass)
        double ebest[2] = {-kInf, -kInf};
        uint64_t ekey[2] = {~0ull, ~0ull};
        uint32_t edi[2] = {0, 0};
#pragma unroll
        for (int x = 0; x < 2; x++) {
          if (!near[x]) continue;
          for (uint32_t di = 0; di < nd; di++) {
            const uint32_t kte = kt[di];
            const uint32_t k = kte & 0x7fffffffu;
            const bool kok = (kte >> 31) != 0;
            const double rho = S.rho[di];
            const uint32_t e = jjx[x] + k;
            const uint32_t cnt = cum_n(cm, e) - en[x];
            const double2 hi = cum_vw(cm, e);
            const double w = hi.y - ew[x];
            const double a = hi.x - ev[x];
            const bool usable = jx[x] < nb && kok && cnt >= 3 && cnt + 3 <= nobs;
            if (usable && a > 0.0) {
              const uint64_t key = (uint64_t)k * nb1 + jjx[x];
              const double mm = exact_merit(a, w, w, rho, c, &bad);
              if (mm > ebest[x] || (mm == ebest[x] && key < ekey[x])) {
                ebest[x] = mm;
                ekey[x] = key;
                edi[x] = di;
              }
            }
            if (!((gmask >> di) & 1u)) continue;
            for (uint32_t si = 1; si < ns; si++) {
              const uint32_t* tab = st_tab + (di * (ns - 1) + (si - 1)) * (layers - 1);
              double sa = a, shp = w, snorm = w;
              if (A.ingress[si] != 0.0)
                shape_sums(cm, jjx[x], a, w, layers, tab, &sa, &shp, &snorm);
              if (usable && sa > 0.0) {
                const uint64_t key = ((uint64_t)si * nb1 + k) * nb1 + jjx[x];
                const double mm = exact_merit(sa, shp, snorm, rho, c, &bad);
                if (mm > ebest[x] || (mm == ebest[x] && key < ekey[x])) {
                  ebest[x] = mm;
                  ekey[x] = key;
                  edi[x] = di;
                }
              }
            }
          }
        }
        // wave argmax by (value desc, key asc)
        int bx = (ebest[1] > ebest[0] || (ebest[1] == ebest[0] && ekey[1] < ekey[0])) ? 1 : 0;
        double bvv = ebest[bx];
        uint64_t bkk = ekey[bx];
        uint32_t bdd = edi[bx];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          double ov = __shfl_xor(bvv, off, 64);
          uint64_t ok = __shfl_xor(bkk, off, 64);
          uint32_t od = __shfl_xor(bdd, off, 64);
          if (ov > bvv || (ov == bvv && ok < bkk)) {
            bvv = ov;
            bkk = ok;
            bdd = od;
          }
        }
        tile_value = bvv;
        tile_key = bkk;
        tile_di = bdd;
        if (lane == 0) atomicAdd(A.counters + 2 * S.index + 1, 1u);
      }
    }
    if (__any(bad) && lane == 0) atomicOr(A.counters + 2 * S.index, 1u);
    if (lane == 0) {
      tval[g * tiles + tile] = tile_value;
      tkey[g * tiles + tile] = tile_key;
      tdi[g * tiles + tile] = tile_di;
    }
  }
  __syncthreads();

  // ---- phase 4: merge tiles (kernels.rs:1113-1155), exact record -------
  for (uint32_t g = tid; g < gcount; g += BLOCK) {
    double best = -kInf;
    uint64_t key = ~0ull;
    uint32_t bdi = 0;
    for (uint32_t t = 0; t < tiles; t++) {
      double v = tval[g * tiles + t];
      uint64_t k = tkey[g * tiles + t];
      if (v > -kInf && (v > best || (v == best && k < key))) {
        best = v;
        key = k;
        bdi = tdi[g * tiles + t];
      }
    }
    if (!(best > -kInf)) {
      key = nb1;  // no usable template: canonical (phase 0, width 1, box)
      bdi = 0;
    }
    Cum cm;
    cm.vw = (const double2*)(hist + (size_t)g * S.pstride);
    cm.n = (const uint32_t*)(cm.vw + nb1);
    cm.nb = nb;
    cm.tv = cm.vw[nb].x;
    cm.tw = cm.vw[nb].y;
    cm.tn = cm.n[nb];
    const uint32_t phase = (uint32_t)(key % nb1);
    const uint32_t width = (uint32_t)((key / nb1) % nb1);
    const uint32_t si = (uint32_t)(key / ((uint64_t)nb1 * nb1));
    const uint32_t sic = si < ns - 1 ? si : ns - 1;
    const double ing = A.ingress[sic];
    const double r = S.rho[bdi];
    const double2 hi = cum_vw(cm, phase + width);
    const double2 lo = cum_vw(cm, phase);
    const double w0 = hi.y - lo.y;
    const double a00 = hi.x - lo.x;
    double a0, hp, norm;
    shape_sums_direct(cm, phase, width, a00, w0, ing, A.factors, sic, layers,
                      &a0, &hp, &norm);
    const double a = a0 / r;
    const double d = hp * hp / c;
    const double q = __builtin_fmax(norm - d, 1e-300);
    const double value = best > -kInf ? a / __builtin_sqrt(q) : -kInf;
    const double depth = a / q * r;
    const uint32_t f = p0 + g;
    double* rec = S.records + (size_t)f * 9;
    rec[0] = S.periods[f];
    rec[1] = S.scales[f];
    rec[2] = (double)phase;
    rec[3] = (double)width;
    rec[4] = a;
    rec[5] = norm;
    rec[6] = d;
    rec[7] = ing;
    rec[8] = r;
    S.score[f] = value;
    S.depth[f] = depth;
    if (S.winner) {
      uint32_t* wn = S.winner + (size_t)f * 4;
      wn[0] = phase;
      wn[1] = width;
      wn[2] = si;
      wn[3] = bdi;
    }
  }
}

}  // namespace

size_t scan_lds_bytes(uint32_t nb, uint32_t G, uint32_t nd, uint32_t ns,
                      uint32_t layers, uint32_t* period_stride) {
  const uint32_t nb1 = nb + 1;
  uint32_t ps = (uint32_t)((20ull * nb1 + 15) & ~15ull);
  if (period_stride) *period_stride = ps;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  size_t bytes = (size_t)G * ps;
  bytes += (size_t)G * nd * 4;
  bytes += (size_t)G * nd * (ns - 1) * (layers - 1) * 4;
  bytes = (bytes + 7) & ~(size_t)7;
  bytes += (size_t)G * tiles * (8 + 8 + 4);
  return bytes;
}

hipError_t scan_configure(size_t max_lds) {
  hipError_t e = hipFuncSetAttribute((const void*)fused_scan<256>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)max_lds);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)fused_scan<512>,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)max_lds);
}

hipError_t scan_launch(const LaunchArgs& args, uint32_t groups, uint32_t block,
                       size_t lds, hipStream_t stream) {
  if (groups == 0) return hipSuccess;
  if (block == 512)
    hipLaunchKernelGGL(fused_scan<512>, dim3(groups), dim3(512), lds, stream, args);
  else
    hipLaunchKernelGGL(fused_scan<256>, dim3(groups), dim3(256), lds, stream, args);
  return hipGetLastError();
}

}  // namespace fbls
