// Fused FBLS period scan for gfx950 (MI355X).
//
// One workgroup owns G periods of one band.  It folds every observation of
// the star into per-period LDS histograms (fixed-point values, so the sums
// are exact and order-independent), turns them into exclusive prefix sums in
// place, scans every (phase bin, duration, shape) template with a cheap FP64
// surrogate of the squared merit, resolves near ties with the exact FP64
// merit, and writes one record per period.  Nothing but the record touches
// HBM.
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

__device__ __forceinline__ double cum_d(const double* e, uint32_t x,
                                        uint32_t nb, double tot) {
  // doubled prefix: cum[x] for x in [0, 2nb]
  return x <= nb ? e[x] : e[x - nb] + tot;
}

__device__ __forceinline__ uint32_t cum_n(const uint32_t* e, uint32_t x,
                                          uint32_t nb, uint32_t tot) {
  return x <= nb ? e[x] : e[x - nb] + tot;
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
  double s, second;       // best surrogate and best other candidate
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
__device__ __forceinline__ void shape_sums(const double* cv, const double* cw,
                                           uint32_t nb, double tv, double tw,
                                           uint32_t jj, double a0, double w0,
                                           uint32_t layers,
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
    double dw = cum_d(cw, start + width, nb, tw) - cum_d(cw, start, nb, tw);
    double dv = cum_d(cv, start + width, nb, tv) - cum_d(cv, start, nb, tv);
    a += dv / l;
    hp += dw / l;
    norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

// Staircase sums with widths computed from the factors (record recompute).
__device__ void shape_sums_direct(const double* cv, const double* cw,
                                  uint32_t nb, double tv, double tw,
                                  uint32_t j, uint32_t k, double a0, double w0,
                                  double ingress, const double* factors,
                                  uint32_t si, uint32_t layers, double* pa,
                                  double* php, double* pnorm) {
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
    double dw = cum_d(cw, start + width, nb, tw) - cum_d(cw, start, nb, tw);
    double dv = cum_d(cv, start + width, nb, tv) - cum_d(cv, start, nb, tv);
    a += dv / l;
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
__device__ __forceinline__ T wave_excl_scan(T x, uint32_t lane, T* total) {
  T incl = x;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  *total = __shfl(incl, 63, 64);
  return incl - x;
}

// In-place inclusive scan of slots [1, nb] by one wave, then scale.
template <typename T>
__device__ void wave_prefix(T* arr, uint32_t nb, uint32_t lane, T scale,
                            bool do_scale) {
  uint32_t len = (nb + 63) / 64;
  len |= 1u;  // odd stride between lanes spreads LDS banks
  uint32_t lo = 1 + lane * len;
  uint32_t hi = lo + len;
  if (hi > nb + 1) hi = nb + 1;
  T acc = 0;
  for (uint32_t s = lo; s < hi; s++) {
    acc += arr[s];
    arr[s] = acc;
  }
  T total;
  T excl = wave_excl_scan(lo < hi ? acc : (T)0, lane, &total);
  for (uint32_t s = lo; s < hi; s++) {
    T x = arr[s] + excl;
    arr[s] = do_scale ? x * scale : x;
  }
}

struct PeriodView {
  const double* cv;
  const double* cw;
  const uint32_t* cn;
  double tv, tw;
  uint32_t tn;
};

template <uint32_t BLOCK>
__global__ __launch_bounds__(BLOCK) void fused_scan(ScanArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  constexpr uint32_t kWaves = BLOCK / 64;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const uint32_t nb = A.nb;
  const uint32_t nb1 = nb + 1;
  const uint32_t p0 = blockIdx.x * A.G;
  const uint32_t gcount = min(A.G, A.np - p0);
  const uint32_t nd = A.nd, ns = A.ns, layers = A.layers;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  const uint32_t nsl = (ns - 1) * (layers - 1);  // table entries per duration

  // ---- LDS carve-up ---------------------------------------------------
  const uint32_t pstride = A.period_stride;  // bytes per period histogram
  unsigned char* p = lds;
  unsigned char* hist = p;
  p += (size_t)A.G * pstride;
  uint32_t* ktab = (uint32_t*)p;  // [G][nd]: k | kok << 31
  p += A.G * nd * 4;
  uint32_t* stab = (uint32_t*)p;  // [G][nd][nsl]: width | offset << 16
  p += A.G * nd * nsl * 4;
  p = (unsigned char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  double* tval = (double*)p;  // [G][tiles]
  p += A.G * tiles * 8;
  uint64_t* tkey = (uint64_t*)p;
  p += A.G * tiles * 8;
  uint32_t* tdi = (uint32_t*)p;

  // ---- phase 0: clear histograms, duration tables ----------------------
  {
    uint4* h4 = (uint4*)hist;
    const uint32_t words = (uint32_t)((size_t)gcount * pstride / 16);
    for (uint32_t i = tid; i < words; i += BLOCK) h4[i] = make_uint4(0, 0, 0, 0);
    for (uint32_t i = tid; i < gcount * nd; i += BLOCK) {
      uint32_t g = i / nd, di = i % nd;
      double scale = A.scales[p0 + g];
      double rk = __builtin_floor(A.dur[di] * scale + 0.5);
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
  {
    const float inv_nb = 1.0f / (float)nb;
    const uint32_t nblk = A.n_pad / kObsPerWave;
    for (uint32_t blk = wave; blk < nblk; blk += kWaves) {
      const uint32_t base = blk * kObsPerWave + lane * kObsPerLane;
      double tt[kObsPerLane], vv[kObsPerLane], ww[kObsPerLane];
      const double2* tp = (const double2*)(A.t + base);
      const int4* qp = (const int4*)(A.vw + 2 * (size_t)base);
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
      const bool full = blk * kObsPerWave + kObsPerWave <= A.n;
      for (uint32_t g = 0; g < gcount; g++) {
        const double scale = A.scales[p0 + g];
        double* cv = (double*)(hist + (size_t)g * pstride);
        double* cw = cv + nb1;
        uint32_t* cn = (uint32_t*)(cw + nb1);
#pragma unroll
        for (uint32_t r = 0; r < kObsPerLane; r++) {
          const uint32_t u = (uint32_t)(tt[r] * scale);
          uint32_t b;
          if (A.wide_cells) {
            b = u % nb;
          } else {
            uint32_t q = (uint32_t)((float)u * inv_nb);
            int32_t rr = (int32_t)(u - q * nb);
            rr += rr < 0 ? (int32_t)nb : 0;
            rr -= rr >= (int32_t)nb ? (int32_t)nb : 0;
            b = (uint32_t)rr;
          }
          if (full || base + r < A.n) {
            __hip_atomic_fetch_add(cv + b + 1, vv[r], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(cw + b + 1, ww[r], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(cn + b + 1, 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- phase 2: exclusive prefix sums in place (exact) -----------------
  for (uint32_t item = wave; item < 3 * gcount; item += kWaves) {
    uint32_t g = item / 3, f = item % 3;
    double* cv = (double*)(hist + (size_t)g * pstride);
    double* cw = cv + nb1;
    uint32_t* cn = (uint32_t*)(cw + nb1);
    if (f == 0)
      wave_prefix<double>(cv, nb, lane, A.sv, true);
    else if (f == 1)
      wave_prefix<double>(cw, nb, lane, A.sw, true);
    else
      wave_prefix<uint32_t>(cn, nb, lane, 1u, false);
  }
  __syncthreads();

  // ---- phase 3: template scan, one 128-bin tile per wave at a time -----
  const double c = A.C, inv_c = A.invC;
  const uint32_t nobs = A.n;
  for (uint32_t item = wave; item < gcount * tiles; item += kWaves) {
    const uint32_t g = item / tiles, tile = item % tiles;
    PeriodView pv;
    pv.cv = (const double*)(hist + (size_t)g * pstride);
    pv.cw = pv.cv + nb1;
    pv.cn = (const uint32_t*)(pv.cw + nb1);
    pv.tv = pv.cv[nb];
    pv.tw = pv.cw[nb];
    pv.tn = pv.cn[nb];
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
      ev[x] = pv.cv[jjx[x]];
      ew[x] = pv.cw[jjx[x]];
      en[x] = pv.cn[jjx[x]];
      state_init(st[x]);
    }
    bool bad = false;
    uint32_t gmask = 0;
    for (uint32_t di = 0; di < nd; di++) {
      const uint32_t kte = kt[di];
      const uint32_t k = kte & 0x7fffffffu;
      const bool kok = (kte >> 31) != 0;
      const double rho = A.rho[di];
      const double rho2 = rho * rho;
      double bv[2], bw[2];
      bool busable[2];
      bool pass = false;
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const uint32_t e = jjx[x] + k;
        const uint32_t cnt = cum_n(pv.cn, e, nb, pv.tn) - en[x];
        const double w = cum_d(pv.cw, e, nb, pv.tw) - ew[x];
        const double a = cum_d(pv.cv, e, nb, pv.tv) - ev[x];
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
            const double g2 = A.gate * A.gate;
            if (A.gate <= 0.0 || s >= g2 * (1.0 + kNear)) {
              pass = true;
            } else if (s >= g2 * (1.0 - kNear)) {
              pass = pass || exact_merit(a, w, w, rho, c, &bad) >= A.gate;
            }
          }
        }
      }
      if (!__any(pass)) continue;
      gmask |= 1u << di;
      for (uint32_t si = 1; si < ns; si++) {
        const uint32_t* tab = st_tab + (di * (ns - 1) + (si - 1)) * (layers - 1);
#pragma unroll
        for (int x = 0; x < 2; x++) {
          double a = bv[x], hp = bw[x], norm = bw[x];
          if (A.ingress[si] != 0.0)
            shape_sums(pv.cv, pv.cw, nb, pv.tv, pv.tw, jjx[x], bv[x], bw[x],
                       layers, tab, &a, &hp, &norm);
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
            const double rho = A.rho[di];
            const uint32_t e = jjx[x] + k;
            const uint32_t cnt = cum_n(pv.cn, e, nb, pv.tn) - en[x];
            const double w = cum_d(pv.cw, e, nb, pv.tw) - ew[x];
            const double a = cum_d(pv.cv, e, nb, pv.tv) - ev[x];
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
                shape_sums(pv.cv, pv.cw, nb, pv.tv, pv.tw, jjx[x], a, w,
                           layers, tab, &sa, &shp, &snorm);
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
        if (lane == 0) atomicAdd(A.counters + 1, 1u);
      }
    }
    if (__any(bad) && lane == 0) atomicOr(A.counters, 1u);
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
    const double* cv = (const double*)(hist + (size_t)g * pstride);
    const double* cw = cv + nb1;
    const double tv = cv[nb], tw = cw[nb];
    const uint32_t phase = (uint32_t)(key % nb1);
    const uint32_t width = (uint32_t)((key / nb1) % nb1);
    const uint32_t si = (uint32_t)(key / ((uint64_t)nb1 * nb1));
    const uint32_t sic = si < ns - 1 ? si : ns - 1;
    const double ing = A.ingress[sic];
    const double r = A.rho[bdi];
    const double w0 = cum_d(cw, phase + width, nb, tw) - cum_d(cw, phase, nb, tw);
    const double a00 = cum_d(cv, phase + width, nb, tv) - cum_d(cv, phase, nb, tv);
    double a0, hp, norm;
    shape_sums_direct(cv, cw, nb, tv, tw, phase, width, a00, w0, ing,
                      A.factors, sic, layers, &a0, &hp, &norm);
    const double a = a0 / r;
    const double d = hp * hp / c;
    const double q = __builtin_fmax(norm - d, 1e-300);
    const double value = best > -kInf ? a / __builtin_sqrt(q) : -kInf;
    const double depth = a / q * r;
    const uint32_t f = p0 + g;
    double* rec = A.records + (size_t)f * 9;
    rec[0] = A.periods[f];
    rec[1] = A.scales[f];
    rec[2] = (double)phase;
    rec[3] = (double)width;
    rec[4] = a;
    rec[5] = norm;
    rec[6] = d;
    rec[7] = ing;
    rec[8] = r;
    A.score[f] = value;
    A.depth[f] = depth;
    if (A.winner) {
      uint32_t* wn = A.winner + (size_t)f * 4;
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

hipError_t scan_launch(const ScanArgs& args, uint32_t block, size_t lds,
                       hipStream_t stream) {
  const uint32_t grid = (args.np + args.G - 1) / args.G;
  if (block == 512)
    hipLaunchKernelGGL(fused_scan<512>, dim3(grid), dim3(512), lds, stream, args);
  else
    hipLaunchKernelGGL(fused_scan<256>, dim3(grid), dim3(256), lds, stream, args);
  return hipGetLastError();
}

}  // namespace fbls
