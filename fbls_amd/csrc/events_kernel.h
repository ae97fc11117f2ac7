// Launch interface of the single-event scan kernels (events_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbls {

constexpr uint32_t kEventPrefixBlock = 1024;  // one workgroup builds the prefix
constexpr uint32_t kEventScanBlock = 256;

struct EventPrefixArgs {
  const double* t;    // n ascending origin-relative times
  const int32_t* vw;  // n x {v, w} fixed point
  uint32_t n, nb;
  uint32_t fx;        // fixed-point modes: bit 0 v, bit 1 w in block mode
  double scale;       // nb / trial period
  double sv, sw;      // 2^-ev, 2^-ew
  double* cv;         // nb + 1: exclusive prefix of v over bins
  double* cw;         // nb + 1
  uint32_t* cn;       // nb + 1
  uint32_t* counter;  // record counter, cleared here
};

struct EventScanArgs {
  const double* cv;
  const double* cw;
  const uint32_t* cn;
  const double* durations;  // nd ascending (days)
  const double* rho;        // nd
  const double* ingress;    // ns
  const double* factors;    // ns x layers staircase width factors
  double scale, C, threshold;
  uint32_t n, nb, nd, ns, layers, capacity;
  uint32_t* counter;
  double* records;          // capacity x {phase, width, score, depth}
  uint64_t* keys;           // capacity: (phase * nd + duration) * ns + shape
};

hipError_t event_prefix_launch(const EventPrefixArgs& a, hipStream_t s);
hipError_t event_scan_launch(const EventScanArgs& a, hipStream_t s);

}  // namespace fbls
