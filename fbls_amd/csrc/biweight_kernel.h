// Launch interface of the biweight detrending kernel (biweight_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbls {

constexpr uint32_t kBiweightBlock = 128;    // points per workgroup
constexpr uint32_t kBiweightWindow = 4096;  // union of their windows staged in LDS

struct BiweightArgs {
  const double* flux;      // n values 1 + relative flux
  const uint32_t* lo;      // n: first index of each point's window
  const uint32_t* hi;      // n: one past the last
  const uint32_t* blocks;  // per workgroup {first point, count, union lo, union hi}
  double c;                // tuning constant of the scale
  uint32_t iterations;
  double* out;             // n: flux / location - 1
};

// Blocks whose union of windows may exceed the staging size (read from
// global memory, counting selection of the medians).
hipError_t biweight_launch(const BiweightArgs& a, uint32_t nblocks, hipStream_t s);
// Blocks whose union of windows holds at most kBiweightWindow values (sorted
// once in LDS).
hipError_t biweight_sorted_launch(const BiweightArgs& a, uint32_t nblocks, hipStream_t s);

}  // namespace fbls
