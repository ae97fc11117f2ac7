// Device-side band pipeline of a search (band_kernel.hip): between the
// first (coarse) scan and the full scan of the prescan survivors, and after
// it, the work the reference does on the host (src/search/scan.rs:381-405)
// runs on the device, so one search -- or one chunk of a batch -- is a
// single stream of launches with one synchronisation at its end.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "scan_kernel.h"

namespace fbls {

// A reported row (scan.rs:16-23, the layout of fbls_row).
struct RowOut {
  double period, epoch, duration, score, depth;
};

// One band (one reference scan_periods call) on the device.
struct BandDesc {
  const double* periods;  // np: the first scan's periods
  const double* score1;   // np: first-scan scores
  const double* rec1;     // np x 9: first-scan records (a band without prescan)
  const double* depth1;   // np
  double* surv_p;         // np capacity: survivor periods (a band with prescan)
  double* surv_s;         // np capacity: their scales nb / P at the full nb
  const double* rec2;     // np capacity x 9: full-scan records of the survivors
  const double* score2;   // np capacity
  const double* depth2;   // np capacity
  RowOut* rows;           // np capacity: the band's rows (score >= threshold)
  double origin;          // the family's time origin (epochs are absolute)
  double nb;              // full-scan phase bins (the survivors' scales)
  uint32_t np;            // periods of the first scan
  uint32_t prescan;       // 1: survivors of the first scan get a full scan
  uint32_t fine_table;    // the full scan's segment table: 0 LDS-resident, 1 range-split
  uint32_t fine_seg;      // the band's segment in that table
};

// The full scan's segment tables, completed on the device: a segment's np is
// its band's survivor count; G (periods per workgroup) the smaller of the
// host's bound (the largest count the launch's LDS holds, or 1 in the
// range-split table) and total / 1024, so the grid fills 256 CUs; group0 the
// running sum of ceil(np / G).
struct FineTable {
  Segment* segs;
  uint32_t nseg;
  uint32_t cap_by_total;  // 1: G <= max(1, total periods / 1024)
  uint32_t* vgroups;      // out: groups of the table (the persistent launch's count)
  unsigned int* ticket;   // out: zeroed for the persistent launch
};

struct BandArgs {
  BandDesc* bands;
  uint32_t nbands;
  double cut;             // prescan_fraction * threshold (scan.rs:390-395)
  double threshold;       // rows: finite score >= threshold (scan.rs:397-405)
  uint32_t* counts;       // per band: {survivors, rows}; then rows in total
  uint32_t* row_off;      // per band: first row in `out` (exclusive prefix)
  RowOut* out;            // every band's rows in band order
  FineTable fine[2];
};

// Survivors of every band with prescan: the periods whose first-scan score
// is finite and >= cut, in period order, and their full-nb scales.
hipError_t band_survivors(const BandArgs& a, hipStream_t stream);
// The full scan's segment tables from the survivor counts (one workgroup).
hipError_t band_plan(const BandArgs& a, hipStream_t stream);
// Each band's rows (from its full scan, or its only scan) in period order,
// then every band's rows gathered in band order into `out`.
hipError_t band_rows(const BandArgs& a, hipStream_t stream);

}  // namespace fbls
