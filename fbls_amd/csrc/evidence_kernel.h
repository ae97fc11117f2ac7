// Launch interface of the evidence (model comparison) kernels
// (evidence_kernel.hip): one trial period of one family.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fbls {

constexpr uint32_t kEvTile = 64;   // bins per tile = one wave64 (evidence.rs:29)
constexpr uint32_t kEvPart = 15;   // values per (tile, selection) (evidence.rs:30)
constexpr uint32_t kEvOrbits = 3;  // secondary phases 0.25, 0.5, 0.75 (evidence.rs:31)
constexpr uint32_t kEvParams = 30; // evidence.rs:32

struct EvArgs {
  const double* t;        // n origin-relative times
  const int32_t* vw;      // n x {v, w} fixed point
  uint32_t n;
  uint32_t fx;            // fixed-point modes: bit 0 v, bit 1 w in block mode
  uint32_t nb, nb2, nd, ns, layers, tiles;
  uint32_t pad_;
  double sv, sw;          // 2^-ev, 2^-ew
  double scale;           // nb / P (host f64)
  double C;               // baseline norm
  double threshold;       // SNR threshold of valid templates
  double screen_threshold;  // threshold * (1 - 1e-12) (evidence.rs:849)
  double margin;          // ln(rtol) - ln((NORB + 1) nb nd ns), or -inf
  const double* dur;      // nd ascending durations
  const double* rho;      // nd
  const double* dlm;      // nd duration log masses
  const double* shapes;   // ns x layers x {fraction, increment, squared coefficient}
  const double* priors;   // ns x {planet, binary} log shape priors
  const double* params;   // 30 model parameters
  const double* orbit;    // nd x 3 x {phase, duration ratio, log w planet, log w binary}
  unsigned long long* hist;  // 2 nb2: v, w fixed-point bin sums (two's complement)
  unsigned int* hcount;      // nb2 bin counts
  double* cum;            // 3 x (nb2 + 1) exclusive prefix sums: v, w, count
  double* screen;         // tiles x 4
  double* parts;          // tiles x 2 x 15
  double* floors;         // {floor_p, floor_b, pilot key, passed}
  // per (template, phase), template-major with a row of tiles x 64 phases;
  // a template is t = duration index x ns + shape index
  double* t_lp;           // planet log evidence of the template
  double* t_lb;           // binary log evidence
  double* t_snr;          // SNR of a valid template, -inf otherwise
  double* t_depth;        // depth
  uint8_t* t_code;        // winning scenario (1-15 planet, 16-22 odd/even, 0 none),
                          // bit 7: the binary diagnostics are not certain
  uint32_t* t_kept;       // tiles x templates: scenario groups evaluated (bits 0-2
                          // secondary phases, bit 3 odd/even)
};

// Bytes of device scratch the kernels need beyond the inputs.
size_t evidence_scratch_bytes(uint32_t nb2, uint32_t tiles, uint32_t templates);
// Carves the scratch regions of `a` out of `base` (evidence_scratch_bytes).
void evidence_bind_scratch(EvArgs* a, void* base);
// Enqueues zero, fold, prefix, screen, pilot and joint kernels on `s`.
hipError_t evidence_launch(const EvArgs& a, hipStream_t s);

}  // namespace fbls
