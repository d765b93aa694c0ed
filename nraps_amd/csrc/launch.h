// launch.h — host-side launch wrappers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace nraps {

void launch_cycle_begin(const DevProblem *d, int nprob, int max_bins, hipStream_t s);
void launch_surface(const DevProblem *d, int nprob, int64_t max_n, hipStream_t s);
// Which instantiation of the event kernel a launch uses.
struct WcVariant {
  bool fp32 = false;      // FP32 tracking/tallies (else FP64)
  bool lds_tally = true;  // tallies privatised in LDS (else global FP64 atomics)
  bool big_block = false; // one 1024-thread block per CU owning the whole LDS
  bool trace = false;     // per-history parity traces compiled in
  bool fission = false;   // fission-site banking compiled in
  bool region = false;    // region-by-region surface tracking (else Woodcock)
  int block_size() const { return big_block ? 1024 : 256; }
};
void launch_event(const WcVariant &v, const DevProblem *d, int nprob, int grid, size_t lds_bytes,
                     int claim, int refill_min, int pass, hipStream_t s);
// Resident blocks per CU of the event kernel for this LDS size.
int event_blocks_per_cu(const WcVariant &v, size_t lds_bytes);
hipError_t event_set_lds_limit(const WcVariant &v, size_t lds_bytes);
void launch_ktree(const DevProblem *d, int nprob, int level, int64_t max_groups, hipStream_t s);
void launch_bank(bool fp32, const DevProblem *d, int nprob, int64_t max_n, int64_t max_tiles,
                 int64_t *tile_sums, const int64_t *tile_base, hipStream_t s);
void launch_finalize(const DevProblem *d, const DevFinal *f, int nprob, int max_groups, int levels,
                     hipStream_t s);
void launch_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n, int draws, uint32_t *out,
                       hipStream_t s);

}  // namespace nraps
