// Host-side pieces of the scan driver shared by the entry points that launch
// the fused scan (fbls_host.cpp, noalloc.hip): the unit of scan work, the
// launch geometry and the segment fields.  Internal, not installed.
#pragma once
#include <cstddef>
#include <cstdint>

#include "host_internal.h"
#include "scan_kernel.h"

namespace fbls {
namespace host {

// One unit of scan work: the periods of one reference scan_bank call.
struct Work {
  const fbls_family* fam;
  const double* periods;
  size_t np;
  uint32_t nb, nd;
  const double* dur;  // nd (ascending)
  const double* rho;  // nd
  double* records;    // np x 9 (host)
  double* score;      // np
  double* depth;      // np
  uint32_t* winner;   // np x 4 or null
  uint32_t retries = 0;
  double cost = 0.0;  // work estimate used to apportion launch time
  bool score_only = false;  // the caller reads only score (records, depth stay on the device)
  bool profiled = false;             // cycles hold a profiled launch's stage cycles
  unsigned long long cycles[3] = {0, 0, 0};  // fold, prefix sums, scan and merge
};

// The bank's shape fields, shared by every segment of a launch.
struct Shared {
  uint32_t ns, layers;
  const double* ingress;
  double gate;
};

// Launch geometry of one segment: periods per workgroup (G), the LDS stride
// of one period's histogram, and for the range-split kernel the bins per
// range and the halo; lds = the segment's LDS bytes at G = 1.
struct SegShape {
  uint32_t G = 1, pstride = 0, range = 0, halo = 0;
  size_t lds = 0;
};

// Bank limits: nd 1..32, ns 1..8, layers 1..16, ingress in [0, 0.5),
// durations and rho positive and finite.
int check_bank(const fbls_bank* b);
// scan.rs:177-183: periods per batch under the fold budget (0 = too small).
uint64_t bank_batch(uint32_t nb, uint32_t period_batch, double max_fold_mib);
Shared shared_of(const fbls_bank* b);
// Does one period's histogram exceed the LDS (range-split kernel)?
bool needs_split(const fbls_ctx* c, const Work& x, const Shared& sh);
// Periods positive and finite, t * nb / P below 2^32; *smax = nb / min P.
int check_periods(const Work& x, double* smax);
// The geometry run_launch gives the segments of one launch: shapes[i] for
// w[i] (nw entries), the launch's LDS bytes and workgroup count.
// cap_by_total: at most max(1, total periods / 1024) periods per workgroup
// (off for a table whose periods are counted on the device).
int segment_shapes(const fbls_ctx* c, const Work* w, size_t nw, const Shared& sh, bool split,
                   SegShape* shapes, size_t* lds_max, uint32_t* groups,
                   bool cap_by_total = true);
// Workgroup size of a launch whose workgroups hold `lds` bytes of LDS.
uint32_t launch_block(const fbls_ctx* c, bool split, size_t lds);
// Every Segment field that does not point into the launch's staging
// (periods, scales, dur, rho, outputs, group0 and index are the caller's).
void segment_fields(const fbls_ctx* c, const Work& x, const SegShape& g, bool split,
                    double smax, fbls::Segment* S);
// Staircase width factors (scan.rs:157-171): factors[s * layers + l].
void bank_factors(const Shared& sh, double* factors);
// Band statistics of one Work's scan (scan.rs:352 for templates_scanned);
// cycles (or null): the Work's fold / prefix / scan shader cycles of a
// profiled launch, which share out `seconds` into the stage fields.
void account(fbls_scan_stats* st, const Work& x, const fbls_bank* b, uint64_t budget,
             double seconds, const unsigned long long* cycles = nullptr);
// FBLS_PROFILE set in the environment (scan.rs:152): launches record stage
// cycles.
bool profiling();

// Resolves the device address of the allocation-free entry point's
// scratch array on the current device (once per device); called by
// fbls_ctx_create with `device` current.
void noalloc_setup(int device);

}  // namespace host
}  // namespace fbls
