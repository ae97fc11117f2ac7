// Launch interface of the fused FBLS scan kernel (scan_kernel.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbls {

// One scan_bank's worth of work: one family (star) at one phase-bin count
// over a list of periods.  A launch carries any number of segments.
struct Segment {
  const double* t;        // n_pad origin-relative times (pad = 0)
  const int32_t* vw;      // n_pad x {v, w} fixed point (pad = 0)
  const double* periods;  // np
  const double* scales;   // np: nb / P, computed on the host in f64
  const double* dur;      // nd durations (ascending)
  const double* rho;      // nd correlated-noise factors
  double* records;        // np x 9
  double* score;          // np
  double* depth;          // np
  uint32_t* winner;       // np x 4 or null
  double sv, sw;          // 2^-ev, 2^-ew: fixed point -> value
  double C, invC;         // baseline norm and its reciprocal
  double survive;         // > 0: a prescan segment's survival cut.  Its periods
                          // only need to be known above or below it: in a
                          // box_first launch the boxes are scanned first, and
                          // a period whose best box reaches the cut keeps that
                          // score (>= the cut) without its staircase templates
                          // (the LDS-resident kernel; the range-split kernel
                          // scans everything).  0: every template, the exact
                          // maximum.
  uint32_t n, n_pad;      // observations, padded to a multiple of 512
  uint32_t np, nb, G, nd; // periods, phase bins, periods per group, durations
  uint32_t pstride;       // LDS bytes per period histogram
  uint32_t wide_cells;    // 1: t*scale may reach 2^24, use integer modulo
  uint32_t group0;        // first workgroup of this segment in the launch
  uint32_t index;         // segment index: counters[2 index], [2 index + 1]
  uint32_t range_bins;    // range-split kernel: bins per range (x128)
  uint32_t halo_bins;     // range-split kernel: halo >= the longest template
  uint32_t sorted;        // 1: times ascend (incremental phase bins)
  int32_t uniform_w;      // > 0: every fixed-point weight equals it; the fold
                          // skips the count atomics, each count being W / uniform_w
  uint32_t fx;            // fixed-point modes: bit 0 v, bit 1 w in block mode
};

struct LaunchArgs {
  const Segment* segs;    // nseg
  uint32_t nseg;
  uint32_t ns, layers;    // shapes and staircase layers (shared)
  uint32_t run_fold;      // 1: run-length pre-aggregation in the fold
  const double* ingress;  // ns
  const double* factors;  // ns x layers staircase width factors
  double gate;            // shape gate on the exact merit
  unsigned int* counters; // per segment: nonpositive-precision flag, exact tile rescans
  // Persistent launch (both non-null): a fixed grid whose workgroups take
  // the launch's groups in order from *ticket (zero when the launch starts)
  // until *vgroups are taken; the segment table and *vgroups may have been
  // written on the device by an earlier kernel of the stream.  vgroups
  // alone: one group per workgroup of a grid at least *vgroups long, the
  // workgroups past *vgroups exit at once.  Both null: one group per
  // workgroup of the grid.
  const uint32_t* vgroups;
  unsigned int* ticket;
  // Stage attribution (the reference's FBLS_PROFILE): per segment three
  // shader-clock sums over its workgroups -- fold, prefix sums, template
  // scan and merge.  Null: not recorded.
  unsigned long long* stages;
  // 1: segments with a survival cut scan their boxes first (non-persistent
  // LDS-resident launches; ignored elsewhere).
  uint32_t box_first;
};

// LDS bytes of one workgroup holding G periods of nb bins.
size_t scan_lds_bytes(uint32_t nb, uint32_t G, uint32_t nd, uint32_t ns,
                      uint32_t layers, uint32_t* period_stride);
// LDS bytes of one range-split workgroup (one period).
size_t split_lds_bytes(uint32_t nb, uint32_t range_bins, uint32_t halo_bins,
                       uint32_t nd, uint32_t ns, uint32_t layers);
hipError_t scan_configure(size_t max_lds);
// groups: total workgroups (sum over segments of ceil(np / G)); split: the
// range-split kernel (every segment with G = 1 and range_bins set).
hipError_t scan_launch(const LaunchArgs& args, uint32_t groups, uint32_t block,
                       size_t lds, bool split, hipStream_t stream);

}  // namespace fbls
