// Fused FBLS period scan for gfx950 (MI355X).
//
// A launch carries segments (one family at one phase-bin count over a list
// of periods: the work of one reference scan_bank call).  In the LDS-resident
// kernel each workgroup owns G periods of one segment: it folds every
// observation of the family into per-period LDS histograms (fixed-point
// values, so the sums are exact and order-independent), turns them into
// exclusive prefix sums in place, scans every (phase bin, duration, shape)
// template with an FP64 surrogate of the squared merit, resolves near ties
// with the exact FP64 merit, and writes one record per period.  When one
// period's histogram does not fit in LDS (large nb), the range-split kernel
// folds the family once per bin range (plus a halo as wide as the longest
// template), carries the exact prefix across ranges and scans each range's
// tiles in turn.  Nothing but the record touches HBM.
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
// LDS per period: vw[nb+1+kOver] {v, w} double pairs (16 B) then
// n[nb+1+kOver] u32; slot b+1 accumulates bin b (slots past nb are the
// fold's overflow slots, added back into the first bins before the scan),
// so after the in-place scan slot x holds the exclusive prefix E[x] = sum of
// bins < x.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scan_kernel.h"
#include "fixed_point.h"
#include "fold.h"

namespace fbls {

namespace {

using namespace fold;

constexpr uint32_t kTile = 128;       // gate granularity (kernels.rs:26)
constexpr double kInf = __builtin_inf();
constexpr double kNear = 1e-10;       // near-tie window of the surrogate

// ------------------------------------------------------------ prefix views

// The whole period in LDS; the doubled prefix (kernels.rs:7-10) is virtual:
// cum[x] = E[x] for x <= nb, E[x - nb] + total above.
struct WrapCum {
  const double2* vw;
  const uint32_t* n;
  double tv, tw;
  uint32_t tn, nb;
  __device__ __forceinline__ double2 vw_at(uint32_t x) const {
    if (x <= nb) return vw[x];
    double2 e = vw[x - nb];
    return make_double2(e.x + tv, e.y + tw);
  }
  __device__ __forceinline__ uint32_t n_at(uint32_t x) const {
    return x <= nb ? n[x] : n[x - nb] + tn;
  }
};

// A window [x0, x0 + len] of the doubled prefix: the exact base cum[x0] plus
// the window's local prefix (all values are integers below 2^53, so the sum
// is the same double as the full prefix would hold).
struct WindowCum {
  const double2* vw;
  const uint32_t* n;
  double bv, bw;
  uint32_t bn, x0;
  __device__ __forceinline__ double2 vw_at(uint32_t x) const {
    double2 e = vw[x - x0];
    return make_double2(bv + e.x, bw + e.y);
  }
  __device__ __forceinline__ uint32_t n_at(uint32_t x) const {
    return bn + n[x - x0];
  }
};

// ------------------------------------------------------------ merit helpers

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

// Staircase layer count and its reciprocals.  With a power-of-two count n,
// x / n and x / (n n) are exact scalings, equal to x * (1 / n) and
// x * (1 / (n n)) for every x, so the hot loop multiplies instead of
// dividing.
struct Layers {
  uint32_t n;
  bool pow2;
  double l, inv_l, inv_l2;
};

__device__ __forceinline__ Layers layers_of(uint32_t n) {
  Layers y;
  y.n = n;
  y.pow2 = (n & (n - 1)) == 0;
  y.l = (double)n;
  y.inv_l = 1.0 / y.l;
  y.inv_l2 = 1.0 / (y.l * y.l);
  return y;
}

// Staircase sums (kernels.rs:84-113) with precomputed layer widths/offsets.
// POW2 (a power-of-two layer count): the divisions by the count are exact
// scalings, done as multiplications by its reciprocals.
template <bool POW2, class CumT>
__device__ __forceinline__ void shape_sums(const CumT& c, uint32_t jj,
                                           double a0, double w0,
                                           const Layers& ly,
                                           const uint32_t* tab, double* pa,
                                           double* php, double* pnorm) {
  const double l = ly.l;
  double a = POW2 ? a0 * ly.inv_l : a0 / l;
  double hp = POW2 ? w0 * ly.inv_l : w0 / l;
  double norm = POW2 ? w0 * ly.inv_l2 : w0 / (l * l);
  for (uint32_t layer = 1; layer < ly.n; layer++) {
    uint32_t e = tab[layer - 1];
    uint32_t width = e & 0xffffu;
    uint32_t start = jj + (e >> 16);
    double2 hi = c.vw_at(start + width);
    double2 lo = c.vw_at(start);
    double dw = hi.y - lo.y;
    double dv = hi.x - lo.x;
    a += POW2 ? dv * ly.inv_l : dv / l;
    hp += POW2 ? dw * ly.inv_l : dw / l;
    norm += dw * (POW2 ? (2.0 * (double)layer + 1.0) * ly.inv_l2
                       : (2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

// Staircase sums with widths computed from the factors (record recompute).
__device__ void shape_sums_direct(const WrapCum& c, uint32_t j, uint32_t k,
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
    double2 hi = c.vw_at(start + width);
    double2 lo = c.vw_at(start);
    double dw = hi.y - lo.y;
    a += (hi.x - lo.x) / l;
    hp += dw / l;
    norm += dw * ((2.0 * (double)layer + 1.0) / (l * l));
  }
  *pa = a;
  *php = hp;
  *pnorm = norm;
}

// ------------------------------------------------------- wave primitives

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

// In-place inclusive prefix of slots [1, len] of the {v, w} pairs (then the
// power-of-two scales) and of the counts, by the whole workgroup: each
// thread sums a contiguous run of slots, the run totals are scanned across
// the waves, and each run adds its offset.  Exact: every partial sum is an
// integer-valued double below 2^53, so any order gives the same doubles.
// wtot: 3 x BLOCK/64 doubles of LDS.
template <uint32_t BLOCK>
__device__ void block_prefix(double2* vw, uint32_t* n, uint32_t len, double sv,
                             double sw, double* wtot, uint32_t tid) {
  constexpr uint32_t kW = BLOCK / 64;
  const uint32_t lane = tid & 63, wave = tid >> 6;
  const uint32_t per = (len + BLOCK - 1) / BLOCK;
  const uint32_t lo = min(1 + tid * per, len + 1);
  const uint32_t hi = min(lo + per, len + 1);
  double av = 0.0, aw = 0.0;
  uint32_t an = 0;
  for (uint32_t s = lo; s < hi; s++) {
    const double2 x = vw[s];
    av += x.x;
    aw += x.y;
    vw[s] = make_double2(av, aw);
    an += n[s];
    n[s] = an;
  }
  double ev = wave_excl_scan(av, lane), ew = wave_excl_scan(aw, lane);
  uint32_t en = wave_excl_scan(an, lane);
  if (lane == 63) {
    wtot[wave] = ev + av;
    wtot[kW + wave] = ew + aw;
    wtot[2 * kW + wave] = (double)(en + an);
  }
  __syncthreads();
  for (uint32_t w = 0; w < wave; w++) {
    ev += wtot[w];
    ew += wtot[kW + w];
    en += (uint32_t)wtot[2 * kW + w];
  }
  for (uint32_t s = lo; s < hi; s++) {
    const double2 x = vw[s];
    vw[s] = make_double2((x.x + ev) * sv, (x.y + ew) * sw);
    n[s] += en;
  }
}

// ------------------------------------------------------------------ fold

// NOCNT skips the count field (equal weights: the counts are rebuilt from
// the weight sums after the fold).
template <bool NOCNT>
__device__ __forceinline__ void lds_add(double* vw, uint32_t* n, uint32_t b,
                                        double v, double w, uint32_t c) {
  __hip_atomic_fetch_add(vw + 2 * (b + 1), v, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add(vw + 2 * (b + 1) + 1, w, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
  if (!NOCNT)
    __hip_atomic_fetch_add(n + b + 1, c, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Equal fixed-point weights w0: a bin's count is its weight sum / w0, exact
// (the sum is count * w0, an integer below 2^53, and IEEE division of an
// exact multiple returns the integer).
__device__ __forceinline__ uint32_t count_of(double wsum, double w0) {
  return (uint32_t)(wsum / w0);
}

__device__ __forceinline__ void load_obs(const Segment& S, uint32_t base, double* tt,
                                         double* vv, double* ww) {
  fold::load_obs(S.t, S.vw, S.fx, base, tt, vv, ww);
}

template <bool WIDE, bool RUN, bool UNI>
__device__ void fold_all(const Segment& S, unsigned char* lds, uint32_t hist_off,
                         uint32_t p0, uint32_t gcount, uint32_t wave,
                         uint32_t lane, uint32_t nwaves) {
  const uint32_t nb = S.nb;
  const uint32_t nsl = nb + 1 + kOver;
  const uint32_t span = min(kOver, nb);
  const float inv_lo = inv_nb_low(nb);
  const uint32_t nblk = S.n_pad / kObsPerWave;
  // Lane l holds the scale of period p0 + l (gcount <= 16), loaded once; a
  // period's scale is then a v_readlane.  A memory read inside the period
  // loop would wait on the counter the LDS atomics share and drain them.
  const double my_scale = lane < gcount ? ((gdouble*)S.scales)[p0 + lane] : 0.0;
  const uint64_t my_bits = (uint64_t)__double_as_longlong(my_scale);
  const int my_lo = (int)(uint32_t)my_bits, my_hi = (int)(uint32_t)(my_bits >> 32);
  for (uint32_t blk = wave; blk < nblk; blk += nwaves) {
    const uint32_t base = blk * kObsPerWave + lane * kObsPerLane;
    double tt[kObsPerLane], pv[kObsPerLane], pw[kObsPerLane];
    load_obs(S, base, tt, pv, pw);
#pragma unroll
    for (uint32_t r = 1; r < kObsPerLane; r++) {  // inclusive prefix, exact
      pv[r] += pv[r - 1];
      if (!UNI) pw[r] += pw[r - 1];
    }
    const bool full = blk * kObsPerWave + kObsPerWave <= S.n;
    for (uint32_t g = 0; g < gcount; g++) {
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane(my_lo, (int)g);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane(my_hi, (int)g);
      const double scale = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
      HistAt h;
      h.vo = hist_off + g * S.pstride;
      h.no = h.vo + 16u * nsl;
      uint32_t u[kObsPerLane];
#pragma unroll
      for (uint32_t r = 0; r < kObsPerLane; r++) u[r] = phase_cell_u(tt[r], scale);
      if (!full) {
        fold_lane<WIDE, RUN, false, UNI, false>(lds, h, u, pv, pw, nb, inv_lo, base, S.n);
      } else if (RUN && S.sorted && __all(u[kObsPerLane - 1] - u[0] < span)) {
        fold_lane<WIDE, RUN, true, UNI, true>(lds, h, u, pv, pw, nb, inv_lo, base, S.n);
      } else {
        fold_lane<WIDE, RUN, true, UNI, false>(lds, h, u, pv, pw, nb, inv_lo, base, S.n);
      }
    }
  }
}

// Fold into the window [x0, x0 + wlen) of the bins (mod nb) of one period:
// slot rel + 1 accumulates bin (x0 + rel) mod nb for rel < wlen.  With
// ascending times, the 512 observations of a wave's block fall in the cells
// floor(t_first * scale) .. floor(t_last * scale), so a block spanning fewer
// than nb cells whose bins all miss the window is skipped from its two end
// times alone (scalar loads).  Inside the window, runs of equal cells are
// summed in registers and bins advance by cell differences, as in fold_lane.
template <bool WIDE, bool NOCNT>
__device__ void fold_window(const Segment& S, double* vw, uint32_t* n,
                            double scale, uint32_t x0, uint32_t wlen,
                            uint32_t wave, uint32_t lane, uint32_t nwaves) {
  const uint32_t nb = S.nb;
  const float inv_lo = inv_nb_low(nb);
  const uint32_t nblk = S.n_pad / kObsPerWave;
  gdouble* tg = (gdouble*)S.t;
  auto rel_of = [&](uint32_t u) {
    const uint32_t b = phase_bin<WIDE>(u, nb, inv_lo);
    return b >= x0 ? b - x0 : b + nb - x0;
  };
  for (uint32_t blk = wave; blk < nblk; blk += nwaves) {
    const uint32_t first = blk * kObsPerWave;
    const bool full = first + kObsPerWave <= S.n;
    if (S.sorted && full) {
      const uint32_t uf = phase_cell_u(tg[first], scale);
      const uint32_t ul = phase_cell_u(tg[first + kObsPerWave - 1], scale);
      // the block's bins relative to x0: rf, rf + 1, ..., rf + (ul - uf) mod nb
      if (ul - uf < nb) {
        const uint32_t rf = rel_of(uf);
        if (rf >= wlen && rf + (ul - uf) < nb) continue;
      }
    }
    const uint32_t base = first + lane * kObsPerLane;
    double tt[kObsPerLane], vv[kObsPerLane], ww[kObsPerLane];
    load_obs(S, base, tt, vv, ww);
    uint32_t u[kObsPerLane];
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) u[r] = phase_cell_u(tt[r], scale);
    if (full && S.sorted && __all(u[kObsPerLane - 1] - u[0] < nb)) {
      uint32_t cr = rel_of(u[0]);
      double av = vv[0], aw = ww[0];
      uint32_t an = 1;
#pragma unroll
      for (uint32_t r = 1; r < kObsPerLane; r++) {
        const uint32_t d = u[r] - u[r - 1];
        const bool flush = d != 0;
        uint32_t rr = cr + d;
        rr = min(rr, rr - nb);
        if (flush && cr < wlen) lds_add<NOCNT>(vw, n, cr, av, aw, an);
        const double sv = av + vv[r];
        av = flush ? vv[r] : sv;
        const double sw = aw + ww[r];
        aw = flush ? ww[r] : sw;
        if (!NOCNT) an = flush ? 1u : an + 1u;
        cr = rr;
      }
      if (cr < wlen) lds_add<NOCNT>(vw, n, cr, av, aw, an);
    } else {
#pragma unroll
      for (uint32_t r = 0; r < kObsPerLane; r++) {
        const uint32_t rel = rel_of(u[r]);
        if (rel < wlen && base + r < S.n) lds_add<NOCNT>(vw, n, rel, vv[r], ww[r], 1u);
      }
    }
  }
}

// ------------------------------------------------------------ tile scan

struct TileCtx {
  const uint32_t* kt;     // [nd] k | kok << 31
  const uint32_t* stab;   // [nd][ns-1][layers-1] width | offset << 16
  const double* rho;      // [nd]
  const double* ingress;  // [ns]
  double c, inv_c, gate, g2;
  uint32_t nb, nd, ns, layers, nobs, lane;
  Layers ly;
};

struct TileOut {
  double value, a, hp, norm;
  uint64_t key;
  uint32_t di;
};

// The winner of one 128-bin tile (kernels.rs:608-845): every duration's box
// for both of the lane's bins, the tile gate, the staircase shapes, the
// surrogate lane states, the exact value of the winner and an exact rescan
// when the top is near-tied.  Wave-uniform result.
template <bool POW2, class CumT>
__device__ __forceinline__ TileOut scan_tile(const CumT& cm, const TileCtx& k, uint32_t tile,
                             bool* bad, unsigned int* rescans) {
  const uint32_t nb = k.nb, nb1 = nb + 1, nd = k.nd, ns = k.ns;
  const uint32_t layers = k.layers, lane = k.lane;
  const uint32_t nsl = (ns - 1) * (layers - 1);
  uint32_t jx[2], jjx[2];
  double ev[2], ew[2];
  uint32_t en[2];
  LaneState st[2];
#pragma unroll
  for (int x = 0; x < 2; x++) {
    jx[x] = tile * kTile + lane + 64 * x;
    jjx[x] = jx[x] < nb ? jx[x] : nb - 1;
    double2 e = cm.vw_at(jjx[x]);
    ev[x] = e.x;
    ew[x] = e.y;
    en[x] = cm.n_at(jjx[x]);
    state_init(st[x]);
  }
  uint32_t gmask = 0;
  for (uint32_t di = 0; di < nd; di++) {
    const uint32_t kte = k.kt[di];
    const uint32_t kk = kte & 0x7fffffffu;
    const bool kok = (kte >> 31) != 0;
    const double rho = k.rho[di];
    const double rho2 = rho * rho;
    double bv[2], bw[2];
    bool busable[2];
    bool pass = false;
#pragma unroll
    for (int x = 0; x < 2; x++) {
      const uint32_t e = jjx[x] + kk;
      const uint32_t cnt = cm.n_at(e) - en[x];
      const double2 hi = cm.vw_at(e);
      const double w = hi.y - ew[x];
      const double a = hi.x - ev[x];
      const bool usable = jx[x] < nb && kok && cnt >= 3 && cnt + 3 <= k.nobs;
      const bool valid = usable && a > 0.0;
      bv[x] = a;
      bw[x] = w;
      busable[x] = usable;
      if (valid) {
        const uint64_t key = (uint64_t)kk * nb1 + jjx[x];
        const double s = surrogate(a, w, w, rho, rho2, k.c, k.inv_c, bad);
        state_update(st[x], s, key, di, a, w, w, rho);
        if (s > -kInf) {
          if (k.gate <= 0.0 || s >= k.g2 * (1.0 + kNear)) {
            pass = true;
          } else if (s >= k.g2 * (1.0 - kNear)) {
            pass = pass || exact_merit(a, w, w, rho, k.c, bad) >= k.gate;
          }
        }
      }
    }
    if (!__any(pass)) continue;
    gmask |= 1u << di;
    for (uint32_t si = 1; si < ns; si++) {
      const uint32_t* tab = k.stab + (di * (ns - 1) + (si - 1)) * (layers - 1);
      const bool staircase = k.ingress[si] != 0.0;
#pragma unroll
      for (int x = 0; x < 2; x++) {
        double a = bv[x], hp = bw[x], norm = bw[x];
        if (staircase) shape_sums<POW2>(cm, jjx[x], bv[x], bw[x], k.ly, tab, &a, &hp, &norm);
        if (busable[x] && a > 0.0) {
          const uint64_t key = ((uint64_t)si * nb1 + kk) * nb1 + jjx[x];
          const double s = surrogate(a, hp, norm, rho, rho2, k.c, k.inv_c, bad);
          state_update(st[x], s, key, di, a, hp, norm, rho);
        }
      }
    }
  }
  (void)nsl;

  TileOut out;
  out.value = -kInf;
  out.key = ~0ull;
  out.di = 0;
  out.a = out.hp = out.norm = 0.0;
  const double m = wave_max(__builtin_fmax(st[0].s, st[1].s));
  if (!(m > -kInf)) return out;
  const double thr = m * (1.0 - kNear);
  bool near[2];
  uint64_t kmin = ~0ull;
#pragma unroll
  for (int x = 0; x < 2; x++) {
    near[x] = st[x].s >= thr;
    if (st[x].s == m && st[x].key < kmin) kmin = st[x].key;
  }
  const uint64_t top_key = wave_min_u64(kmin);
  const uint64_t owner_mask = __ballot(st[0].key == top_key || st[1].key == top_key);
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
      const bool same = st[x].a == ta && st[x].hp == thp && st[x].norm == tno &&
                        st[x].rho == trh;
      need = need || !same || st[x].second >= st[x].s * (1.0 - kNear);
    }
  }
  if (!__any(need)) {
    out.key = top_key;
    out.di = tdi0;
    out.a = ta;
    out.hp = thp;
    out.norm = tno;
    out.value = exact_merit(ta, thp, tno, trh, k.c, bad);
    return out;
  }
  // exact rescan of the near states (kernels.rs:662-798 retry pass)
  double eb[2] = {-kInf, -kInf};
  uint64_t ek[2] = {~0ull, ~0ull};
  uint32_t ed[2] = {0, 0};
  for (int x = 0; x < 2; x++) {
    if (!near[x]) continue;
    for (uint32_t di = 0; di < nd; di++) {
      const uint32_t kte = k.kt[di];
      const uint32_t kk = kte & 0x7fffffffu;
      const bool kok = (kte >> 31) != 0;
      const double rho = k.rho[di];
      const uint32_t e = jjx[x] + kk;
      const uint32_t cnt = cm.n_at(e) - en[x];
      const double2 hi = cm.vw_at(e);
      const double w = hi.y - ew[x];
      const double a = hi.x - ev[x];
      const bool usable = jx[x] < nb && kok && cnt >= 3 && cnt + 3 <= k.nobs;
      if (usable && a > 0.0) {
        const uint64_t key = (uint64_t)kk * nb1 + jjx[x];
        const double mm = exact_merit(a, w, w, rho, k.c, bad);
        if (mm > eb[x] || (mm == eb[x] && key < ek[x])) {
          eb[x] = mm;
          ek[x] = key;
          ed[x] = di;
        }
      }
      if (!((gmask >> di) & 1u)) continue;
      for (uint32_t si = 1; si < ns; si++) {
        const uint32_t* tab = k.stab + (di * (ns - 1) + (si - 1)) * (layers - 1);
        double sa = a, shp = w, snorm = w;
        if (k.ingress[si] != 0.0)
          shape_sums<POW2>(cm, jjx[x], a, w, k.ly, tab, &sa, &shp, &snorm);
        if (usable && sa > 0.0) {
          const uint64_t key = ((uint64_t)si * nb1 + kk) * nb1 + jjx[x];
          const double mm = exact_merit(sa, shp, snorm, rho, k.c, bad);
          if (mm > eb[x] || (mm == eb[x] && key < ek[x])) {
            eb[x] = mm;
            ek[x] = key;
            ed[x] = di;
          }
        }
      }
    }
  }
  const int bx = (eb[1] > eb[0] || (eb[1] == eb[0] && ek[1] < ek[0])) ? 1 : 0;
  double bv = eb[bx];
  uint64_t bk = ek[bx];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double ov = __shfl_xor(bv, off, 64);
    uint64_t okk = __shfl_xor(bk, off, 64);
    if (ov > bv || (ov == bv && okk < bk)) {
      bv = ov;
      bk = okk;
    }
  }
  const uint64_t wmask = __ballot((ek[0] == bk && eb[0] == bv) || (ek[1] == bk && eb[1] == bv));
  const int wl = __builtin_ctzll(wmask);
  const int wx = __shfl((ek[1] == bk && eb[1] == bv) ? 1 : 0, wl, 64);
  out.value = bv;
  out.key = bk;
  out.di = __shfl(wx ? ed[1] : ed[0], wl, 64);
  // the winner's sums, recomputed once (wave-uniform) from its key
  const uint32_t wj = (uint32_t)(bk % nb1);
  const uint32_t wk = (uint32_t)((bk / nb1) % nb1);
  const uint32_t wsi = (uint32_t)(bk / ((uint64_t)nb1 * nb1));
  const double2 hi = cm.vw_at(wj + wk);
  const double2 lo = cm.vw_at(wj);
  double a = hi.x - lo.x, w = hi.y - lo.y, hp = w, norm = w;
  if (wsi > 0 && k.ingress[wsi] != 0.0)
    shape_sums<POW2>(cm, wj, hi.x - lo.x, w, k.ly,
               k.stab + (out.di * (ns - 1) + (wsi - 1)) * (layers - 1), &a, &hp,
               &norm);
  out.a = a;
  out.hp = hp;
  out.norm = norm;
  if (lane == 0) atomicAdd(rescans, 1u);
  return out;
}

// Per-period duration tables: k and its usability flag, and the staircase
// layer widths/offsets (scan.rs:157-171, kernels.rs:104-106,679-681).
__device__ void duration_tables(const Segment& S, const LaunchArgs& A,
                                uint32_t p, uint32_t* ktab, uint32_t* stab,
                                uint32_t tid, uint32_t block) {
  const uint32_t nb = S.nb, nd = S.nd, ns = A.ns, layers = A.layers;
  for (uint32_t di = tid; di < nd; di += block) {
    double scale = S.scales[p];
    double rk = __builtin_floor(S.dur[di] * scale + 0.5);
    uint32_t raw_k = rk < 1.0 ? 1u : (rk >= 2147483648.0 ? 0x7fffffffu : (uint32_t)rk);
    uint32_t k = raw_k < nb / 2 ? raw_k : nb / 2;
    uint32_t kok = raw_k < nb / 2 ? 1u : 0u;
    ktab[di] = k | (kok << 31);
    for (uint32_t si = 1; si < ns; si++) {
      for (uint32_t l = 1; l < layers; l++) {
        double factor = A.factors[si * layers + l];
        double wf = __builtin_floor((double)k * factor + 0.5);
        uint32_t width = wf < 1.0 ? 1u : (uint32_t)wf;
        uint32_t off = (k - width) / 2;
        stab[(di * (ns - 1) + (si - 1)) * (layers - 1) + (l - 1)] = width | (off << 16);
      }
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

// Record of the period's winner (kernels.rs:800-843, 1144-1154).
__device__ void write_record(const Segment& S, uint32_t f, uint64_t key,
                             uint32_t di, bool valid, uint32_t ns,
                             const double* ingress, double a0, double hp,
                             double norm) {
  const uint32_t nb1 = S.nb + 1;
  const uint32_t phase = (uint32_t)(key % nb1);
  const uint32_t width = (uint32_t)((key / nb1) % nb1);
  const uint32_t si = (uint32_t)(key / ((uint64_t)nb1 * nb1));
  const uint32_t sic = si < ns - 1 ? si : ns - 1;
  const double r = S.rho[di];
  const double a = a0 / r;
  const double d = hp * hp / S.C;
  const double q = __builtin_fmax(norm - d, 1e-300);
  const double value = valid ? a / __builtin_sqrt(q) : -kInf;
  const double depth = a / q * r;
  double* rec = S.records + (size_t)f * 9;
  rec[0] = S.periods[f];
  rec[1] = S.scales[f];
  rec[2] = (double)phase;
  rec[3] = (double)width;
  rec[4] = a;
  rec[5] = norm;
  rec[6] = d;
  rec[7] = ingress[sic];
  rec[8] = r;
  S.score[f] = value;
  S.depth[f] = depth;
  if (S.winner) {
    uint32_t* wn = S.winner + (size_t)f * 4;
    wn[0] = phase;
    wn[1] = width;
    wn[2] = si;
    wn[3] = di;
  }
}

// ------------------------------------------------- LDS-resident kernel

// Stage clock of a profiled launch: thread 0 adds the shader-clock cycles
// since its previous mark to the segment's stage k (0 fold, 1 prefix sums,
// 2 template scan and merge); every mark follows a barrier, so the cycles
// are the workgroup's.
// PROF false: no clock (the kernels without it keep their registers).
template <bool PROF>
struct StageClock {
  unsigned long long* acc;
  uint64_t last;
  __device__ __forceinline__ explicit StageClock(unsigned long long* a) : acc(a), last(0) {
    if (PROF && threadIdx.x == 0) last = clock64();
  }
  __device__ __forceinline__ void mark(uint32_t k) {
    if (PROF && threadIdx.x == 0) {
      const uint64_t now = clock64();
      atomicAdd(acc + k, (unsigned long long)(now - last));
      last = now;
    }
  }
};

// One group of the launch: G periods of one segment, from the fold to the
// records.  LDS: the workgroup's dynamic LDS.
template <uint32_t BLOCK, bool PROF, bool BOX>
__device__ __forceinline__ void scan_group(const LaunchArgs& A, uint32_t group,
                                           unsigned char* lds) {
  constexpr uint32_t kWaves = BLOCK / 64;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const Segment S = A.segs[find_segment(A, group)];
  const uint32_t nb = S.nb;
  const uint32_t nb1 = nb + 1;
  const uint32_t slots = nb1 + kOver;  // {v, w} and count slots per period
  const uint32_t p0 = (group - S.group0) * S.G;
  const uint32_t gcount = min(S.G, S.np - p0);
  const uint32_t nd = S.nd, ns = A.ns, layers = A.layers;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  const uint32_t nsl = (ns - 1) * (layers - 1);
  StageClock<PROF> clk(PROF ? A.stages + 3 * (size_t)S.index : nullptr);

  unsigned char* p = lds;
  unsigned char* hist = p;
  p += (size_t)S.G * S.pstride;
  uint32_t* ktab = (uint32_t*)p;  // [G][nd]
  p += S.G * nd * 4;
  uint32_t* stab = (uint32_t*)p;  // [G][nd][nsl]
  p += S.G * nd * nsl * 4;
  p = (unsigned char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  double* tval = (double*)p;  // [G][tiles]
  p += S.G * tiles * 8;
  uint64_t* tkey = (uint64_t*)p;
  p += S.G * tiles * 8;
  uint32_t* tdi = (uint32_t*)p;
  p += S.G * tiles * 4;
  uint32_t* sflag = (uint32_t*)p;  // [G] prescan survival flags

  // ---- phase 0: clear histograms, duration tables
  {
    uint4* h4 = (uint4*)hist;
    const uint32_t words = (uint32_t)((size_t)gcount * S.pstride / 16);
    for (uint32_t i = tid; i < words; i += BLOCK) h4[i] = make_uint4(0, 0, 0, 0);
    for (uint32_t g = wave; g < gcount; g += kWaves)
      duration_tables(S, A, p0 + g, ktab + g * nd, stab + g * nd * nsl, lane, 64);
  }
  __syncthreads();

  // ---- phase 1: fold every observation into the G histograms
  const bool uni = A.run_fold && S.uniform_w > 0;
  const uint32_t hist_off = (uint32_t)(hist - lds);
  if (S.wide_cells) {
    if (!A.run_fold) fold_all<true, false, false>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
    else if (uni) fold_all<true, true, true>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
    else fold_all<true, true, false>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
  } else {
    if (!A.run_fold) fold_all<false, false, false>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
    else if (uni) fold_all<false, true, true>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
    else fold_all<false, true, false>(S, lds, hist_off, p0, gcount, wave, lane, kWaves);
  }
  __syncthreads();
  fold_finish<BLOCK>(hist, S.pstride, gcount, nb, slots, uni, (double)S.uniform_w);
  clk.mark(0);

  // ---- phase 2: exclusive prefix sums in place (exact)
  for (uint32_t item = wave; item < 2 * gcount; item += kWaves) {
    uint32_t g = item >> 1;
    double2* vw = (double2*)(hist + (size_t)g * S.pstride);
    if (item & 1)
      wave_prefix_n((uint32_t*)(vw + slots), nb, lane);
    else
      wave_prefix_vw(vw, nb, lane, S.sv, S.sw);
  }
  __syncthreads();
  clk.mark(1);

  // ---- phase 3: template scan, one 128-bin tile per wave at a time.  In a
  // box-first launch (BOX) a prescan segment (survival cut > 0) first scans
  // the boxes alone (the
  // shape gate at +inf); a period whose best box reaches the cut survives
  // whatever its staircases give and keeps that score, the others are
  // scanned in full in a second pass.
  TileCtx k;
  k.rho = S.rho;
  k.ingress = A.ingress;
  k.c = S.C;
  k.inv_c = S.invC;
  k.nb = nb;
  k.nd = nd;
  k.ns = ns;
  k.layers = layers;
  k.ly = layers_of(layers);
  k.nobs = S.n;
  k.lane = lane;
  unsigned int* rescans = A.counters + 2 * S.index + 1;
  const bool surv = BOX && S.survive > 0.0;
  for (uint32_t pass = surv ? 0u : 1u; pass < 2; pass++) {
    k.gate = pass == 0 ? kInf : A.gate;
    k.g2 = k.gate * k.gate;
    for (uint32_t item = wave; item < gcount * tiles; item += kWaves) {
      const uint32_t g = item / tiles, tile = item % tiles;
      if (surv && pass == 1 && sflag[g]) continue;  // survives by its boxes
      WrapCum cm;
      cm.vw = (const double2*)(hist + (size_t)g * S.pstride);
      cm.n = (const uint32_t*)(cm.vw + slots);
      cm.nb = nb;
      const double2 t = cm.vw[nb];
      cm.tv = t.x;
      cm.tw = t.y;
      cm.tn = cm.n[nb];
      k.kt = ktab + g * nd;
      k.stab = stab + (size_t)g * nd * nsl;
      bool bad = false;
      const TileOut o = k.ly.pow2 ? scan_tile<true>(cm, k, tile, &bad, rescans)
                                  : scan_tile<false>(cm, k, tile, &bad, rescans);
      if (__any(bad) && lane == 0) atomicOr(A.counters + 2 * S.index, 1u);
      if (lane == 0) {
        tval[g * tiles + tile] = o.value;
        tkey[g * tiles + tile] = o.key;
        tdi[g * tiles + tile] = o.di;
      }
    }
    __syncthreads();
    if (pass == 0) {
      for (uint32_t g = tid; g < gcount; g += BLOCK) {
        double m = -kInf;
        for (uint32_t t = 0; t < tiles; t++) m = fmax(m, tval[g * tiles + t]);
        sflag[g] = m >= S.survive ? 1u : 0u;
      }
      __syncthreads();
    }
  }

  // ---- phase 4: merge tiles (kernels.rs:1113-1155), exact record
  for (uint32_t g = tid; g < gcount; g += BLOCK) {
    double best = -kInf;
    uint64_t key = ~0ull;
    uint32_t bdi = 0;
    for (uint32_t t = 0; t < tiles; t++) {
      double v = tval[g * tiles + t];
      uint64_t kk = tkey[g * tiles + t];
      if (v > -kInf && (v > best || (v == best && kk < key))) {
        best = v;
        key = kk;
        bdi = tdi[g * tiles + t];
      }
    }
    const bool valid = best > -kInf;
    if (!valid) {
      key = nb1;  // no usable template: canonical (phase 0, width 1, box)
      bdi = 0;
    }
    WrapCum cm;
    cm.vw = (const double2*)(hist + (size_t)g * S.pstride);
    cm.n = (const uint32_t*)(cm.vw + slots);
    cm.nb = nb;
    cm.tv = cm.vw[nb].x;
    cm.tw = cm.vw[nb].y;
    cm.tn = cm.n[nb];
    const uint32_t phase = (uint32_t)(key % nb1);
    const uint32_t width = (uint32_t)((key / nb1) % nb1);
    const uint32_t si = (uint32_t)(key / ((uint64_t)nb1 * nb1));
    const uint32_t sic = si < ns - 1 ? si : ns - 1;
    const double2 hi = cm.vw_at(phase + width);
    const double2 lo = cm.vw_at(phase);
    double a0, hp, norm;
    shape_sums_direct(cm, phase, width, hi.x - lo.x, hi.y - lo.y,
                      A.ingress[sic], A.factors, sic, layers, &a0, &hp, &norm);
    write_record(S, p0 + g, key, bdi, valid, ns, A.ingress, a0, hp, norm);
  }
  clk.mark(2);
}

// Runs body(group) for the launch's groups.  PERSIST: a fixed grid whose
// workgroups take every group from the ticket counter (the ticket is passed
// through the first word of the LDS; body may use all of the LDS, so each
// turn ends at a barrier).  Otherwise one group per workgroup.  The two are
// separate kernels: the loop raises register pressure, and the launches
// whose groups the host knows keep the loop-free body.
template <bool PERSIST, class Body>
__device__ __forceinline__ void for_each_group(const LaunchArgs& A, unsigned char* lds,
                                               Body body) {
  if (!PERSIST) {
    // a grid sized for every period surviving: the groups past the
    // device-written total have nothing to do
    if (A.vgroups && blockIdx.x >= *A.vgroups) return;
    body(A, blockIdx.x);
    return;
  }
  const uint32_t total = *A.vgroups;
  volatile uint32_t* slot = (volatile uint32_t*)lds;
  for (;;) {
    if (threadIdx.x == 0) *slot = atomicAdd(A.ticket, 1u);
    __syncthreads();
    const uint32_t group = __builtin_amdgcn_readfirstlane(*slot);
    __syncthreads();
    if (group >= total) break;
    body(A, group);
    __syncthreads();
  }
}

// MODE: 0 one group per workgroup, 1 persistent, 2 one group per workgroup
// with the box-first prescan.
template <uint32_t BLOCK, uint32_t MODE, bool PROF>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4)))
void fused_scan(LaunchArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  for_each_group<MODE == 1>(A, lds, [&](const LaunchArgs& a, uint32_t group) {
    scan_group<BLOCK, PROF, MODE == 2>(a, group, lds);
  });
}

// ----------------------------------------------------- range-split kernel

// One period per workgroup; the bins are processed in ranges of
// S.range_bins with a halo of S.halo_bins (>= the longest template).
template <uint32_t BLOCK, bool PROF>
__device__ __forceinline__ void split_group(const LaunchArgs& A, uint32_t group,
                                            unsigned char* lds) {
  constexpr uint32_t kWaves = BLOCK / 64;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const Segment S = A.segs[find_segment(A, group)];
  const uint32_t nb = S.nb;
  const uint32_t nb1 = nb + 1;
  const uint32_t f = group - S.group0;  // G = 1
  const uint32_t nd = S.nd, ns = A.ns, layers = A.layers;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  const uint32_t nsl = (ns - 1) * (layers - 1);
  const uint32_t B = S.range_bins, H = S.halo_bins;
  const uint32_t slots = B + H + 1;

  unsigned char* p = lds;
  double2* wvw = (double2*)p;  // [slots]
  p += (size_t)slots * 16;
  uint32_t* wn = (uint32_t*)p;  // [slots]
  p += ((size_t)slots * 4 + 15) & ~(size_t)15;
  uint32_t* ktab = (uint32_t*)p;
  p += nd * 4;
  uint32_t* stab = (uint32_t*)p;
  p += nd * nsl * 4;
  p = (unsigned char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  double* tval = (double*)p;  // [tiles] x {value, a, hp, norm}
  p += tiles * 32;
  uint64_t* tkey = (uint64_t*)p;
  p += tiles * 8;
  uint32_t* tdi = (uint32_t*)p;
  p += tiles * 4;
  p = (unsigned char*)(((uintptr_t)p + 7) & ~(uintptr_t)7);
  double* carry = (double*)p;  // {bv, bw, bin0 v, bin0 w}, bn as double
  p += 5 * 8;
  double* wtot = (double*)p;   // [3][BLOCK / 64] wave totals of the prefix

  if (wave == 0) duration_tables(S, A, f, ktab, stab, lane, 64);
  if (tid == 0) {
    carry[0] = 0.0;
    carry[1] = 0.0;
    carry[4] = 0.0;
  }
  TileCtx k;
  k.kt = ktab;
  k.stab = stab;
  k.rho = S.rho;
  k.ingress = A.ingress;
  k.c = S.C;
  k.inv_c = S.invC;
  k.gate = A.gate;
  k.g2 = A.gate * A.gate;
  k.nb = nb;
  k.nd = nd;
  k.ns = ns;
  k.layers = layers;
  k.ly = layers_of(layers);
  k.nobs = S.n;
  k.lane = lane;
  unsigned int* rescans = A.counters + 2 * S.index + 1;
  const double scale = S.scales[f];
  StageClock<PROF> clk(PROF ? A.stages + 3 * (size_t)S.index : nullptr);
  for (uint32_t x0 = 0; x0 < nb; x0 += B) {
    const uint32_t W = min(B, nb - x0);
    const uint32_t wlen = W + H;
    {
      uint4* z = (uint4*)wvw;
      for (uint32_t i = tid; i < slots; i += BLOCK) z[i] = make_uint4(0, 0, 0, 0);
      for (uint32_t i = tid; i < slots; i += BLOCK) wn[i] = 0;
    }
    __syncthreads();
    const bool nocnt = A.run_fold && S.uniform_w > 0;
    if (S.wide_cells) {
      if (nocnt) fold_window<true, true>(S, (double*)wvw, wn, scale, x0, wlen, wave, lane, kWaves);
      else fold_window<true, false>(S, (double*)wvw, wn, scale, x0, wlen, wave, lane, kWaves);
    } else {
      if (nocnt) fold_window<false, true>(S, (double*)wvw, wn, scale, x0, wlen, wave, lane, kWaves);
      else fold_window<false, false>(S, (double*)wvw, wn, scale, x0, wlen, wave, lane, kWaves);
    }
    __syncthreads();
    if (nocnt) {
      const double w0 = (double)S.uniform_w;
      for (uint32_t i = tid; i < wlen; i += BLOCK) wn[i + 1] = count_of(wvw[i + 1].y, w0);
      __syncthreads();
    }
    clk.mark(0);
    block_prefix<BLOCK>(wvw, wn, wlen, S.sv, S.sw, wtot, tid);
    __syncthreads();
    clk.mark(1);
    WindowCum cm;
    cm.vw = wvw;
    cm.n = wn;
    cm.bv = carry[0];
    cm.bw = carry[1];
    cm.bn = (uint32_t)carry[4];
    cm.x0 = x0;
    if (x0 == 0 && tid == 0) {
      carry[2] = wvw[1].x;
      carry[3] = wvw[1].y;
    }
    const uint32_t t0 = x0 / kTile, t1 = (x0 + W + kTile - 1) / kTile;
    for (uint32_t tile = t0 + wave; tile < t1; tile += kWaves) {
      bool bad = false;
      const TileOut o = k.ly.pow2 ? scan_tile<true>(cm, k, tile, &bad, rescans)
                                  : scan_tile<false>(cm, k, tile, &bad, rescans);
      if (__any(bad) && lane == 0) atomicOr(A.counters + 2 * S.index, 1u);
      if (lane == 0) {
        tval[4 * tile] = o.value;
        tval[4 * tile + 1] = o.a;
        tval[4 * tile + 2] = o.hp;
        tval[4 * tile + 3] = o.norm;
        tkey[tile] = o.key;
        tdi[tile] = o.di;
      }
    }
    __syncthreads();
    if (tid == 0) {
      carry[0] = cm.bv + wvw[W].x;
      carry[1] = cm.bw + wvw[W].y;
      carry[4] = (double)(cm.bn + wn[W]);
    }
    __syncthreads();
    clk.mark(2);
  }
  if (tid == 0) {
    double best = -kInf, a0 = 0.0, hp = 0.0, norm = 0.0;
    uint64_t key = ~0ull;
    uint32_t bdi = 0;
    for (uint32_t t = 0; t < tiles; t++) {
      double v = tval[4 * t];
      uint64_t kk = tkey[t];
      if (v > -kInf && (v > best || (v == best && kk < key))) {
        best = v;
        key = kk;
        bdi = tdi[t];
        a0 = tval[4 * t + 1];
        hp = tval[4 * t + 2];
        norm = tval[4 * t + 3];
      }
    }
    const bool valid = best > -kInf;
    if (!valid) {
      // canonical (phase 0, width 1, box): the sums of bin 0
      key = nb1;
      bdi = 0;
      a0 = carry[2];
      hp = carry[3];
      norm = carry[3];
    }
    write_record(S, f, key, bdi, valid, ns, A.ingress, a0, hp, norm);
    clk.mark(2);
  }
}

// MODE: 0 one group per workgroup, 1 persistent.
template <uint32_t BLOCK, uint32_t MODE, bool PROF>
__global__ __launch_bounds__(BLOCK) void fused_scan_split(LaunchArgs A) {
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  for_each_group<MODE == 1>(A, lds, [&](const LaunchArgs& a, uint32_t group) {
    split_group<BLOCK, PROF>(a, group, lds);
  });
}

}  // namespace

size_t scan_lds_bytes(uint32_t nb, uint32_t G, uint32_t nd, uint32_t ns,
                      uint32_t layers, uint32_t* period_stride) {
  const uint32_t nb1 = nb + 1;
  uint32_t ps = (uint32_t)((20ull * (nb1 + kOver) + 15) & ~15ull);
  if (period_stride) *period_stride = ps;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  size_t bytes = (size_t)G * ps;
  bytes += (size_t)G * nd * 4;
  bytes += (size_t)G * nd * (ns - 1) * (layers - 1) * 4;
  bytes = (bytes + 7) & ~(size_t)7;
  bytes += (size_t)G * tiles * (8 + 8 + 4);
  bytes += (size_t)G * 4;  // prescan survival flags
  return bytes;
}

size_t split_lds_bytes(uint32_t nb, uint32_t range_bins, uint32_t halo_bins,
                       uint32_t nd, uint32_t ns, uint32_t layers) {
  const size_t slots = (size_t)range_bins + halo_bins + 1;
  const uint32_t tiles = (nb + kTile - 1) / kTile;
  size_t bytes = slots * 16 + ((slots * 4 + 15) & ~(size_t)15);
  bytes += (size_t)nd * 4 + (size_t)nd * (ns - 1) * (layers - 1) * 4;
  bytes = (bytes + 7) & ~(size_t)7;
  bytes += (size_t)tiles * (32 + 8 + 4);
  bytes = (bytes + 7) & ~(size_t)7;
  bytes += 5 * 8;
  bytes += 3 * 16 * 8;  // wave totals of the prefix (up to 1024 threads)
  return bytes;
}

namespace {

// Every instantiation of a kernel: workgroup size x persistent x profiled.
template <uint32_t BLOCK, uint32_t MODE, bool PROF>
struct ScanK {
  static const void* fn() { return (const void*)fused_scan<BLOCK, MODE, PROF>; }
  static void launch(const LaunchArgs& a, uint32_t groups, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((fused_scan<BLOCK, MODE, PROF>), dim3(groups), dim3(BLOCK), lds, s, a);
  }
};
template <uint32_t BLOCK, uint32_t MODE, bool PROF>
struct SplitK {
  static const void* fn() { return (const void*)fused_scan_split<BLOCK, MODE, PROF>; }
  static void launch(const LaunchArgs& a, uint32_t groups, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL((fused_scan_split<BLOCK, MODE, PROF>), dim3(groups), dim3(BLOCK), lds,
                       s, a);
  }
};

template <template <uint32_t, uint32_t, bool> class K, uint32_t MODE, bool PROF>
hipError_t configure_all(size_t max_lds) {
  const void* fns[] = {K<256, MODE, PROF>::fn(), K<512, MODE, PROF>::fn(),
                       K<1024, MODE, PROF>::fn()};
  for (const void* fn : fns) {
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)max_lds);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <template <uint32_t, uint32_t, bool> class K, uint32_t MODE, bool PROF>
void launch_block_of(const LaunchArgs& a, uint32_t groups, uint32_t block, size_t lds,
                     hipStream_t s) {
  if (block == 1024)
    K<1024, MODE, PROF>::launch(a, groups, lds, s);
  else if (block == 512)
    K<512, MODE, PROF>::launch(a, groups, lds, s);
  else
    K<256, MODE, PROF>::launch(a, groups, lds, s);
}

template <template <uint32_t, uint32_t, bool> class K, uint32_t BOXMODE>
void launch_kind(const LaunchArgs& a, uint32_t groups, uint32_t block, size_t lds,
                 hipStream_t s) {
  const bool persist = a.ticket != nullptr, prof = a.stages != nullptr;
  const bool box = !persist && a.box_first;
  if (persist && prof) launch_block_of<K, 1, true>(a, groups, block, lds, s);
  else if (persist) launch_block_of<K, 1, false>(a, groups, block, lds, s);
  else if (box && prof) launch_block_of<K, BOXMODE, true>(a, groups, block, lds, s);
  else if (box) launch_block_of<K, BOXMODE, false>(a, groups, block, lds, s);
  else if (prof) launch_block_of<K, 0, true>(a, groups, block, lds, s);
  else launch_block_of<K, 0, false>(a, groups, block, lds, s);
}

}  // namespace

hipError_t scan_configure(size_t max_lds) {
  hipError_t e;
  if ((e = configure_all<ScanK, 0, false>(max_lds)) != hipSuccess ||
      (e = configure_all<ScanK, 0, true>(max_lds)) != hipSuccess ||
      (e = configure_all<ScanK, 1, false>(max_lds)) != hipSuccess ||
      (e = configure_all<ScanK, 1, true>(max_lds)) != hipSuccess ||
      (e = configure_all<ScanK, 2, false>(max_lds)) != hipSuccess ||
      (e = configure_all<ScanK, 2, true>(max_lds)) != hipSuccess ||
      (e = configure_all<SplitK, 0, false>(max_lds)) != hipSuccess ||
      (e = configure_all<SplitK, 0, true>(max_lds)) != hipSuccess ||
      (e = configure_all<SplitK, 1, false>(max_lds)) != hipSuccess ||
      (e = configure_all<SplitK, 1, true>(max_lds)) != hipSuccess)
    return e;
  return hipSuccess;
}

hipError_t scan_launch(const LaunchArgs& args, uint32_t groups, uint32_t block,
                       size_t lds, bool split, hipStream_t stream) {
  if (groups == 0) return hipSuccess;
  if (split)  // the range-split kernel has no box-first mode
    launch_kind<SplitK, 0>(args, groups, block, lds, stream);
  else
    launch_kind<ScanK, 2>(args, groups, block, lds, stream);
  return hipGetLastError();
}

}  // namespace fbls
