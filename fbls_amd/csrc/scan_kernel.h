// Launch interface of the fused FBLS scan kernel (scan_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbls {

struct ScanArgs {
  const double* t;        // n_pad origin-relative times (pad = 0)
  const int32_t* vw;      // n_pad x {v, w} fixed point (pad = 0)
  uint32_t n, n_pad;      // observations, padded to a multiple of 512
  double sv, sw;          // 2^-ev, 2^-ew: fixed point -> value
  double C, invC;         // baseline norm and its reciprocal
  const double* periods;  // np
  const double* scales;   // np: nb / P, computed on the host in f64
  uint32_t np, nb, G;     // periods, phase bins, periods per workgroup
  uint32_t nd, ns, layers;
  const double* dur;      // nd
  const double* rho;      // nd
  const double* ingress;  // ns
  const double* factors;  // ns x layers staircase width factors
  double gate;            // shape gate on the exact merit
  uint32_t wide_cells;    // 1: t*scale may reach 2^24, use integer modulo
  uint32_t period_stride; // LDS bytes per period histogram
  double* records;        // np x 9
  double* score;          // np
  double* depth;          // np
  uint32_t* winner;       // np x 4 or null
  unsigned int* counters; // [0] nonpositive-precision flag, [1] exact tile rescans
};

// LDS bytes of one workgroup holding G periods of nb bins.
size_t scan_lds_bytes(uint32_t nb, uint32_t G, uint32_t nd, uint32_t ns,
                      uint32_t layers, uint32_t* period_stride);
hipError_t scan_configure(size_t max_lds);
hipError_t scan_launch(const ScanArgs& args, uint32_t block, size_t lds,
                       hipStream_t stream);

}  // namespace fbls
