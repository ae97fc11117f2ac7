// Launch interface of the peak-selection rescoring kernels (select_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbls {

struct RescoreArgs {
  const double* t;         // n origin-relative times of the family
  const int32_t* vw;       // n x {v, w} fixed point
  const uint8_t* excluded; // n: 1 when the observation is excluded
  uint32_t n;
  uint32_t fx;             // fixed-point modes: bit 0 v, bit 1 w in block mode
  double origin;
  const double* rows;      // nrows x {period, epoch (absolute), duration}
  long long* sums;         // nrows x {sum v, sum w} (fixed point)
};

struct CountArgs {
  const double* time_all;  // n merged absolute times
  const uint8_t* excluded; // n
  uint32_t n;
  const double* rows;      // nrows x {period, epoch, duration}
  uint32_t* counts;        // nrows x {in transit, in transit and not excluded}
};

hipError_t rescore_launch(const RescoreArgs& a, uint32_t nrows, hipStream_t s);
hipError_t count_launch(const CountArgs& a, uint32_t nrows, hipStream_t s);

}  // namespace fbls
