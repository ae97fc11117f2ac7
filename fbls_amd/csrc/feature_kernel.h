// Launch interface of the coarse-periodogram kernel (feature_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbls {

struct FeatureStar {
  const double* t;    // n origin-relative times
  const int32_t* vw;  // n x {v, w} fixed point
  uint32_t n;
  uint32_t sorted;    // 1: times ascend
  uint32_t fx;        // fixed-point modes: bit 0 v, bit 1 w in block mode
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
  uint32_t G;               // most periods in one group (<= 4)
  uint32_t narrow;          // 1: every t * nb / P < 2^24 (32-bit cells, float modulo)
  uint32_t ngroups;         // groups of consecutive periods, one workgroup each per star
  const uint32_t* gstart;   // ngroups + 1: group k holds periods [gstart[k], gstart[k+1])
  const uint32_t* gstride;  // ngroups: LDS slots per period of group k (its largest nb + 1)
  uint32_t lds_bytes;       // LDS of the largest group
  uint32_t pad_;
};

size_t feature_lds_bytes(uint32_t nb_max, uint32_t G);
hipError_t feature_launch(const FeatureArgs& a, uint32_t nstars, hipStream_t s);

}  // namespace fbls
