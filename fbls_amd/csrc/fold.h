// Device-side fold shared by the scan kernels (scan_kernel.hip) and the
// coarse periodogram (feature_kernel.hip): observations into LDS phase
// histograms with exact integer sums held in doubles (fixed-point values,
// so every order of addition gives the same doubles), and the in-place
// exclusive prefix of a histogram.  Internal, not installed.
//
// Reference semantics restated here (isakruas/fbls):
//   phase cell   src/search/kernels.rs:314-327 (floor(fl(t * scale)) mod nb)
//   fold         src/search/kernels.rs:464-553
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fixed_point.h"

namespace fbls {
namespace fold {

constexpr uint32_t kObsPerLane = 8;   // contiguous observations per lane
constexpr uint32_t kObsPerWave = 64 * kObsPerLane;
// Overflow slots after bin nb - 1 of a period's LDS histogram: the
// incremental fold may address up to kOver bins past the end (bins 0 ..
// kOver - 1 wrapped once); they are added back after the fold.
constexpr uint32_t kOver = 16;

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

// In-place inclusive scan of slots [1, len] of {v, w} pairs by one wave
// (exact: integer-valued doubles below 2^53), then the power-of-two scales.
__device__ inline void wave_prefix_vw(double2* arr, uint32_t len_slots, uint32_t lane,
                               double sv, double sw) {
  uint32_t len = (len_slots + 63) / 64;
  uint32_t lo = 1 + lane * len;
  uint32_t hi = lo + len;
  if (hi > len_slots + 1) hi = len_slots + 1;
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

__device__ inline void wave_prefix_n(uint32_t* arr, uint32_t len_slots, uint32_t lane) {
  uint32_t len = (len_slots + 63) / 64;
  len |= 1u;  // odd stride between lanes spreads LDS banks
  uint32_t lo = 1 + lane * len;
  uint32_t hi = lo + len;
  if (hi > len_slots + 1) hi = len_slots + 1;
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
__device__ __forceinline__ uint32_t phase_cell_u(double t, double scale) {
  return (uint32_t)(t * scale);  // floor for t >= 0
}

template <bool WIDE>
__device__ __forceinline__ uint32_t phase_bin(uint32_t u, uint32_t nb,
                                              float inv_lo) {
  if (WIDE) return u % nb;
  const uint32_t q = (uint32_t)((float)u * inv_lo);
  const uint32_t r = u - __umul24(q, nb);
  return min(r, r - nb);
}

// One period's histogram in LDS, addressed in bytes from the LDS base:
// slot s (bin s - 1; slots nb + 1 .. nb + kOver are the overflow slots of
// bins 0 .. kOver - 1) holds {v, w} at vo + 16 s and its count at no + 4 s.
struct HistAt {
  uint32_t vo, no;
};


// One run of observations added to a slot: v (and, unless the weights are
// equal, w) as FP64 atomics and the count as a u32 atomic.  UNI (equal
// fixed-point weights w0): a slot's weight sum is count * w0, rebuilt after
// the fold, so only v and the count are folded.
template <bool UNI>
__device__ __forceinline__ void add_run(unsigned char* lds, uint32_t va, uint32_t na, double v,
                                        double w, uint32_t c) {
  double* p = (double*)(lds + va);
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if (!UNI) __hip_atomic_fetch_add(p + 1, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  __hip_atomic_fetch_add((uint32_t*)(lds + na), c, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The observations R0 .. R0 + LEN - 1 of a lane whose cells lie in
// u[R0] .. u[R0] + S (S wave-uniform and small): one atomic set per cell
// rather than one per observation.  Times ascend, so the observations below
// cell u[R0] + k are a prefix of the span: the prefix sum at the last of
// them is picked by selects, and the run of cell u[R0] + k - 1 is the
// difference of two such prefixes.  A lane whose cells end before a cell
// adds nothing there.  Slot addresses: cell * 16 + cv, cell * 4 + cn.
template <uint32_t S, uint32_t R0, uint32_t LEN, bool UNI>
__device__ __forceinline__ void fold_cells(unsigned char* lds, uint32_t cv, uint32_t cn,
                                           const uint32_t* u, const double* pv,
                                           const double* pw) {
  constexpr uint32_t L = R0 + LEN - 1;
  double lv = R0 ? pv[R0 - 1] : 0.0;
  double lw = (!UNI && R0) ? pw[R0 - 1] : 0.0;
  uint32_t lr = R0;
#pragma unroll
  for (uint32_t k = 1; k <= S; k++) {
    const uint32_t ck = u[R0] + k;
    double sv = pv[R0], sw = UNI ? 0.0 : pw[R0];
    uint32_t sr = R0 + 1;
#pragma unroll
    for (uint32_t r = R0 + 1; r <= L; r++) {
      const bool below = u[r] < ck;
      sv = below ? pv[r] : sv;
      if (!UNI) sw = below ? pw[r] : sw;
      sr = below ? r + 1 : sr;
      // keeps the chain as register selects (the compiler would otherwise
      // turn it into an indexed load from a copy of pv in scratch memory)
      asm("" : "+v"(sv));
      if (!UNI) asm("" : "+v"(sw));
    }
    if (ck - 1 <= u[L])
      add_run<UNI>(lds, (ck - 1) * 16u + cv, (ck - 1) * 4u + cn, sv - lv, sw - lw, sr - lr);
    lv = sv;
    lw = sw;
    lr = sr;
  }
  const uint32_t cl = u[R0] + S;
  if (cl <= u[L])
    add_run<UNI>(lds, cl * 16u + cv, cl * 4u + cn, pv[L] - lv, UNI ? 0.0 : pw[L] - lw,
                 L + 1 - lr);
}

// The same observations with one flush per observation position: a run ends
// wherever the next cell differs.
template <uint32_t R0, uint32_t LEN, bool UNI>
__device__ __forceinline__ void fold_positions(unsigned char* lds, uint32_t cv, uint32_t cn,
                                               const uint32_t* u, const double* pv,
                                               const double* pw) {
  constexpr uint32_t L = R0 + LEN - 1;
  double lv = R0 ? pv[R0 - 1] : 0.0;
  double lw = (!UNI && R0) ? pw[R0 - 1] : 0.0;
  uint32_t lr = R0;
#pragma unroll
  for (uint32_t r = R0; r < L; r++) {
    if (u[r] != u[r + 1]) {
      add_run<UNI>(lds, u[r] * 16u + cv, u[r] * 4u + cn, pv[r] - lv, UNI ? 0.0 : pw[r] - lw,
                   r + 1 - lr);
      lv = pv[r];
      if (!UNI) lw = pw[r];
      lr = r + 1;
    }
  }
  add_run<UNI>(lds, u[L] * 16u + cv, u[L] * 4u + cn, pv[L] - lv, UNI ? 0.0 : pw[L] - lw,
               L + 1 - lr);
}

// A lane's 8 observations in the cheapest form its wave allows: one atomic
// set per cell when every lane's cells span at most 3 cells, else one flush
// per observation position.
template <bool UNI>
__device__ __forceinline__ void fold_runs(unsigned char* lds, uint32_t cv, uint32_t cn,
                                          const uint32_t* u, const double* pv,
                                          const double* pw) {
  const uint32_t span = u[kObsPerLane - 1] - u[0];
  if (__all(span <= 1))
    fold_cells<1, 0, kObsPerLane, UNI>(lds, cv, cn, u, pv, pw);
  else if (__all(span <= 2))
    fold_cells<2, 0, kObsPerLane, UNI>(lds, cv, cn, u, pv, pw);
  else
    fold_positions<0, kObsPerLane, UNI>(lds, cv, cn, u, pv, pw);
}

// Fold one lane's 8 consecutive observations into one period's histogram.
// pv, pw: the lane's inclusive prefix sums of v and w (exact integers), so a
// run of observations r0..r in one bin adds pv[r] - pv[r0 - 1].
// INCR (times ascend, the lane's 8 cells span fewer than min(kOver, nb)):
// cell u lands in slot u - off + 1 with off = u[0] - bin(u[0]), a slot below
// nb + 1 + kOver: the bins past nb - 1 are the overflow slots, added back
// into bins 0 .. kOver - 1 after the fold, so a run's address is one
// shift-add of its cell.  RUN: runs of equal bins flushed once.  Otherwise
// one atomic set per observation.  FULL: all 8 observations are real.
template <bool WIDE, bool RUN, bool FULL, bool UNI, bool INCR>
__device__ __forceinline__ void fold_lane(unsigned char* lds, HistAt h, const uint32_t* u,
                                          const double* pv, const double* pw, uint32_t nb,
                                          float inv_lo, uint32_t base, uint32_t nobs) {
  if (RUN && FULL && INCR) {
    const uint32_t off = u[0] - phase_bin<WIDE>(u[0], nb, inv_lo);
    // slot addresses u * 16 + cv (and u * 4 + cn), modulo 2^32; the empty
    // asm keeps each base in one register (otherwise the compiler expands
    // it again at every add)
    uint32_t cv = h.vo + 16u - off * 16u;
    uint32_t cn = h.no + 4u - off * 4u;
    asm("" : "+v"(cv), "+v"(cn));
    fold_runs<UNI>(lds, cv, cn, u, pv, pw);
  } else if (RUN && FULL) {
    uint32_t b[kObsPerLane];
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) b[r] = phase_bin<WIDE>(u[r], nb, inv_lo);
    double lv = 0.0, lw = 0.0;
    uint32_t lr = 0;
#pragma unroll
    for (uint32_t r = 0; r + 1 < kObsPerLane; r++) {
      if (b[r] != b[r + 1]) {
        add_run<UNI>(lds, h.vo + 16u * (b[r] + 1), h.no + 4u * (b[r] + 1), pv[r] - lv,
                     UNI ? 0.0 : pw[r] - lw, r + 1 - lr);
        lv = pv[r];
        if (!UNI) lw = pw[r];
        lr = r + 1;
      }
    }
    constexpr uint32_t L = kObsPerLane - 1;
    add_run<UNI>(lds, h.vo + 16u * (b[L] + 1), h.no + 4u * (b[L] + 1), pv[L] - lv,
                 UNI ? 0.0 : pw[L] - lw, kObsPerLane - lr);
  } else {
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) {
      const uint32_t bb = phase_bin<WIDE>(u[r], nb, inv_lo) + 1;
      const double v = r ? pv[r] - pv[r - 1] : pv[0];
      const double w = UNI ? 0.0 : (r ? pw[r] - pw[r - 1] : pw[0]);
      if (FULL || base + r < nobs) add_run<UNI>(lds, h.vo + 16u * bb, h.no + 4u * bb, v, w, 1u);
    }
  }
}

// Global-address-space views: the observation loads become global (not flat)
// loads, so waiting for them never drains the LDS atomics in flight.
typedef __attribute__((address_space(1))) const double gdouble;
typedef __attribute__((address_space(1))) const int gint;

// A lane's 8 observations from `base`: times and the values of the v and w
// words (fx: bit 0 v, bit 1 w in block mode).
__device__ __forceinline__ void load_obs(const double* t, const int32_t* vw, uint32_t fx,
                                         uint32_t base, double* tt, double* vv, double* ww) {
  gdouble* tp = (gdouble*)__builtin_assume_aligned(t + base, 64);
  gint* qp = (gint*)__builtin_assume_aligned(vw + 2 * (size_t)base, 64);
  if (fx == 0) {  // plain words (a uniform branch: no cost to the common case)
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) {
      tt[r] = tp[r];
      vv[r] = (double)qp[2 * r];
      ww[r] = (double)qp[2 * r + 1];
    }
  } else {
    const bool bv = fx & 1u, bw = fx & 2u;
#pragma unroll
    for (uint32_t r = 0; r < kObsPerLane; r++) {
      tt[r] = tp[r];
      vv[r] = fixed_dvalue(qp[2 * r], bv);
      ww[r] = fixed_dvalue(qp[2 * r + 1], bw);
    }
  }
}

__device__ __forceinline__ float inv_nb_low(uint32_t nb) {
  return (float)((1.0 / (double)nb) * (1.0 - 0x1p-22));
}

// The fold after the last observation: the overflow slots of each of the
// G histograms (at lds + hist_off + g * pstride, `slots` slots each) added
// back into bins 0 .. min(kOver, nb) - 1, then, for equal weights w0 (uni),
// every weight sum rebuilt as count * w0 (exact).  Ends at a barrier.
template <uint32_t BLOCK>
__device__ __forceinline__ void fold_finish(unsigned char* hist, uint32_t pstride,
                                            uint32_t gcount, uint32_t nb, uint32_t slots,
                                            bool uni, double w0) {
  const uint32_t tid = threadIdx.x;
  const uint32_t nb1 = nb + 1;
  const uint32_t wrap = min(kOver, nb);
  for (uint32_t i = tid; i < gcount * wrap; i += BLOCK) {
    const uint32_t g = i / wrap, x = i - g * wrap;
    double2* vw = (double2*)(hist + (size_t)g * pstride);
    uint32_t* n = (uint32_t*)(vw + slots);
    const double2 o = vw[nb1 + x];
    const double2 e = vw[1 + x];
    vw[1 + x] = make_double2(e.x + o.x, e.y + o.y);
    n[1 + x] += n[nb1 + x];
  }
  __syncthreads();
  if (uni) {
    for (uint32_t i = tid; i < gcount * nb; i += BLOCK) {
      const uint32_t g = i / nb, b = i - g * nb;
      double* vw = (double*)(hist + (size_t)g * pstride);
      const uint32_t* n = (const uint32_t*)(vw + 2 * slots);
      vw[2 * (b + 1) + 1] = (double)n[b + 1] * w0;
    }
    __syncthreads();
  }
}

}  // namespace fold
}  // namespace fbls
