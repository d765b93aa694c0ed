// launch.h — host-side launch wrappers of the kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "device.h"

namespace nraps {

void launch_cycle_begin(const DevProblem *d, int nprob, int max_bins, hipStream_t s);
void launch_surface(const DevProblem *d, int nprob, int64_t max_n, hipStream_t s);
void launch_woodcock(bool fp32, bool lds_tally, const DevProblem *d, int nprob, int grid,
                     size_t lds_bytes, int tab_lds_bytes, int claim, hipStream_t s);
// Resident blocks per CU of the Woodcock kernel for this LDS size.
int woodcock_blocks_per_cu(bool fp32, bool lds_tally, size_t lds_bytes);
hipError_t woodcock_set_lds_limit(bool fp32, bool lds_tally, size_t lds_bytes);
void launch_ktree(const DevProblem *d, int nprob, int level, int64_t max_groups, hipStream_t s);
void launch_bank(bool fp32, const DevProblem *d, int nprob, int64_t max_n, int64_t max_tiles,
                 int64_t *tile_sums, const int64_t *tile_base, hipStream_t s);
void launch_finalize(const DevProblem *d, const DevFinal *f, int nprob, int levels, hipStream_t s);
void launch_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n, int draws, uint32_t *out,
                       hipStream_t s);

}  // namespace nraps
