// device.h — structures shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/nraps.h"
#include "host.h"

namespace nraps {

constexpr int kBlock = 256;        // threads per block of every kernel
constexpr int kTreeWidth = 256;    // canonical k-sum group width
constexpr int kScanTile = 2048;    // elements per tile of the site-count scan
constexpr int kMaxSitesPerCollision = 1000;

// Per-problem generation state, resident in device memory so that one
// captured cycle can be replayed without host involvement.
struct CycleState {
  int32_t cycle;      // generation being run
  int32_t from_bank;  // births read the source bank
  int32_t make_sites; // collisions bank fission sites
  int32_t error;      // bit 0: bank overflow, bit 1: no fission sites, bit 2: done
  double inv_k;       // 1 / k of the previous generation (site production)
  double k_prev;      // k that normalises this generation's flux (k_initial at 0)
  unsigned long long site_count; // append counter of the unordered site buffer
  int64_t n_sites;    // sites banked by the last generation
  int64_t queue_head; // next unclaimed history of this generation
  int32_t active;     // problem takes part in launches
  int32_t pad;
};

enum StatIdx {
  kStCollisions = 0,
  kStVirtual,
  kStCrossings,
  kStAbsorb,
  kStLeak,
  kStCapped,
  kStSteps,
  kNumStats
};

// Everything a kernel needs to know about one problem of the batch.
struct DevProblem {
  const uint8_t *tab;  // packed tables (TabLayout)
  TabLayout lay;
  int32_t nc, G, M, R, nf;
  int32_t xs_every, fission;
  int32_t dstart;       // fission bank started from the diffusion source
  int32_t pass;         // which event-kernel launch of a cycle runs it: 0 with
                        // the tally in LDS (or the only launch), 1 global
  int32_t tally_bins;  // 2*G*nc (direct + difference array)
  int32_t nb;          // cell-lookup buckets
  double inv_h;        // 1 / bucket width
  int64_t N;
  uint64_t seed, problem_id;
  int64_t max_events;
  double boundl, boundr, XL, XR, maj, inv_maj;
  CycleState *cs;
  double *k;           // [generations]
  double *kacc;        // [N] per-history k contribution
  int32_t *counts;     // [N] sites per history (fission)
  int64_t *offsets;    // [N] exclusive scan of counts
  double *tally_glob;  // [tally_bins] generation tally (direct | difference)
  // fission bank
  void *site_x;        // [site_cap] Real, unordered
  int32_t *site_cell;
  uint64_t *site_key;  // (history << 32) | sequence
  int64_t site_cap;
  void *ord_x;         // [site_cap] ordered by (history, sequence)
  int32_t *ord_cell;
  void *src_x;         // [N] source bank of the next generation
  int32_t *src_cell;
  // canonical k tree scratch
  double *tree_a, *tree_b;
  // parity traces
  nraps_trace_event *trace;
  int32_t *trace_len;
  int64_t *draws;
  double *trace_kacc;
  int64_t trace_n;
  int32_t trace_ev, trace_cycle;
  unsigned long long *stats; // [kNumStats]
};


// Finalize-stage per-problem data (double precision, global memory).
struct DevFinal {
  const double *dx;     // [nc] Mesh.delta_x
  const double *width;  // [nc] mesh_right - mesh_left
  const double *nut, *sigf; // [G*M]
  const int32_t *mat;   // [nc]
  double nut_conv, length;
  int32_t generations, skip;
  double *tally_cycle;  // [2*G*nc] reduced tally of this generation
  double *tally_last;   // [G*nc]
  double *flux_acc, *fs_acc, *fsum, *fsq; // accumulators
  double *k_ref;        // [generations]
  int64_t *sites_hist;  // [generations]
  double *kpart;        // [G] per-group share of the reference k (scratch)
  int32_t has_diff;     // tally_cycle carries a difference array
};

}  // namespace nraps
