// Launch interface of the coarse-periodogram kernel (feature_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbls {

// Most consecutive periods one workgroup folds from one load of a star.
constexpr uint32_t kFeatureMaxGroup = 16;
// Most durations of one launch.
constexpr uint32_t kFeatureMaxDurations = 16;

struct FeatureStar {
  const double* t;    // n_pad origin-relative times (pad 0)
  const int32_t* vw;  // n_pad x {v, w} fixed point (pad 0)
  uint32_t n, n_pad;  // observations, padded to a multiple of 512
  uint32_t sorted;    // 1: times ascend
  uint32_t fx;        // fixed-point modes: bit 0 v, bit 1 w in block mode
  int32_t uniform_w;  // > 0: every fixed-point weight equals it
  uint32_t pad_;
  double sv, sw;      // fixed point -> value
  double C;           // baseline norm
};

struct FeatureArgs {
  const FeatureStar* stars;
  const double* periods;    // np
  const uint32_t* nb;       // np phase bins per period
  uint32_t np, ng;
  const double* durations;  // ng
  double* out;              // stars x np x ng best SNRs, or null
  float* inputs;            // stars x np x ng network inputs, or null
  float* peaks;             // stars x np largest SNR over the durations, or null
  uint32_t nb_max;          // largest nb
  uint32_t G;               // most periods in one group (<= kFeatureMaxGroup)
  uint32_t narrow;          // 1: every t * nb / P < 2^24 (32-bit cells, float modulo)
  uint32_t ngroups;         // groups of consecutive periods, one workgroup each per star
  const uint32_t* gstart;   // ngroups + 1: group k holds periods [gstart[k], gstart[k+1])
  const uint32_t* gstride;  // ngroups: LDS slots per period of group k (feature_slots of
                            // its largest nb)
  uint32_t lds_bytes;       // LDS of the largest group
  uint32_t uniform;         // 1: every star has equal weights (the fold skips w)
};

// LDS bytes of a group of G periods of at most nb_max bins, and the slots
// of one period's histogram (bins, the prefix's leading zero slot and the
// fold's overflow slots).
size_t feature_lds_bytes(uint32_t nb_max, uint32_t G);
uint32_t feature_slots(uint32_t nb_max);
hipError_t feature_launch(const FeatureArgs& a, uint32_t nstars, hipStream_t s);

}  // namespace fbls
