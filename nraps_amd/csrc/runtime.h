// runtime.h — host runtime objects shared by the solver and the server.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <vector>

#include "device.h"
#include "diffusion.h"
#include "host.h"
#include "launch.h"

namespace nraps {

struct DeviceError : std::runtime_error {
  hipError_t err;
  DeviceError(const char *what, hipError_t e);
};

#define NRAPS_HIP_TRY(x)                                     \
  do {                                                       \
    hipError_t e_ = (x);                                     \
    if (e_ != hipSuccess) throw ::nraps::DeviceError(#x, e_); \
  } while (0)

// Owns device allocations; frees them on destruction.  Bound to a stream,
// it allocates from the device's stream-ordered memory pool on that stream
// (no device-wide synchronisation per allocation or free); the stream must
// outlive the arena.
class DevArena {
 public:
  DevArena() = default;
  DevArena(const DevArena &) = delete;
  DevArena &operator=(const DevArena &) = delete;
  ~DevArena();
  void bind(hipStream_t s) { stream_ = s; pooled_ = true; }
  void *raw(size_t bytes);
  template <class T>
  T *alloc(size_t n) {
    return static_cast<T *>(raw(n * sizeof(T)));
  }
  size_t bytes() const { return bytes_; }

 private:
  std::vector<void *> ptrs_;
  size_t bytes_ = 0;
  hipStream_t stream_ = nullptr;
  bool pooled_ = false;
};

// Keep freed stream-ordered allocations of `device` cached in its default
// memory pool (set once per process and device).
void keep_pool_memory(int device);

int ktree_levels(int64_t n);
int64_t site_capacity(int64_t N);
// Device bytes held for one problem (banks, tallies, accumulators).
int64_t problem_footprint(const nraps_problem *pb, const nraps_run_opts *o);

// One problem's device state: tables, banks, tallies, accumulators.
class ProblemDev {
 public:
  // Host half of the set-up: derived tables and the packed blob (no HIP).
  void prepare(const nraps_problem *pb, const nraps_run_opts &o, bool fp32);
  // Device half: allocations and uploads, asynchronous on `s` when it is
  // not null (host sources are owned by this object or by the caller's
  // problem, which must stay alive until the stream has run them).
  // cs_slot: externally owned CycleState (server slot array) or nullptr.
  void upload(bool event, CycleState *cs_slot, bool want_trace, nraps_result *trace_res,
              hipStream_t s);
  void setup(const nraps_problem *pb, const nraps_run_opts &o, bool fp32, bool event,
             CycleState *cs_slot, bool want_trace, nraps_result *trace_res, hipStream_t s) {
    prepare(pb, o, fp32);
    upload(event, cs_slot, want_trace, trace_res, s);
  }
  // Copy results into r (one synchronisation of the set-up stream); returns
  // the status of the run.
  // `via`: stream for the copies instead of the set-up stream (the caller
  // guarantees the problem's device work has completed).
  int download(nraps_result *r, double ms_total, double ms_device, double ms_transport,
               int64_t launches, hipStream_t via = nullptr);
  // Stop the problem from taking part in further launches.
  void deactivate(hipStream_t s);

  const nraps_problem *pb = nullptr;
  nraps_run_opts o{};
  Prepared prep;
  std::vector<uint8_t> blob;
  TabLayout lay{};
  bool fission = false, fp32 = false;
  int64_t trace_n = 0;
  std::vector<double> width;
  CycleState cs0{};
  hipStream_t stream = nullptr;
  DevProblem dev{};
  DevFinal fin{};
  CycleState *d_cs = nullptr;
  // byte offsets inside the block of small per-problem arrays: [tab .. mat)
  // is uploaded from `stage`, [up_end, end) holds the zeroed accumulators
  struct Blk {
    size_t tab, cs, dx, width, nut, sigf, mat, up_end;
    size_t k, kref, sites, flux, fs, fsum, fsq, tl, stats, kpart, end;
  } blk{};
  uint8_t *d_base = nullptr;
  std::vector<uint8_t> stage;
  DevArena arena;
  // diffusion solve behind a diffusion-started fission bank
  std::unique_ptr<DiffusionSet> diff;
};

// Launch tables and parameters for a set of problems run together.
class Launcher {
 public:
  explicit Launcher(int device);
  ~Launcher();
  // Launch plan for a set of problems; the tables go to the device on `s`
  // (ordered after the work already queued there) when it is not null.
  void build(const std::vector<ProblemDev *> &ps, bool fp32, bool event, hipStream_t s = nullptr);
  // Launches of one generation for every problem of the set.
  void enqueue_cycle(hipStream_t s, hipEvent_t tr0, hipEvent_t tr1);
  // Capture one generation into a graph (replaces any previous capture).
  void capture(hipStream_t s);
  void launch_graph(hipStream_t s);
  bool has_graph() const { return graph_exec_ != nullptr; }

 private:
  void drop_graph();
  int device_;
  int n_ = 0, n_cu_ = 0, lds_limit_ = 0, refill_min_ = 1;
  int max_bins_ = 0, levels_ = 1, max_groups_ = 1;
  int64_t max_n_ = 0, max_tiles_ = 0;
  bool fp32_ = false, event_ = false, fission_any_ = false;
  // event-kernel launches of a cycle: [0] problems with the tally in LDS,
  // [1] problems whose tables + tally do not fit (global FP64 tally)
  struct EventPlan {
    bool used = false;
    WcVariant v{};
    size_t lds_bytes = 0;
    int grid = 1, claim = 256;
  };
  EventPlan plans_[2];
  std::vector<DevProblem> h_probs_;
  std::vector<DevFinal> h_fins_;
  std::vector<int64_t> h_tile_base_;
  DevProblem *d_probs_ = nullptr;
  DevFinal *d_fins_ = nullptr;
  int64_t *d_tile_sums_ = nullptr, *d_tile_base_ = nullptr;
  int64_t cap_probs_ = 0, cap_tiles_ = 0;
  std::vector<void *> owned_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
};

int solve_batch(int n, const nraps_problem *const *pbs, const nraps_run_opts *opts,
                nraps_result *const *results);

}  // namespace nraps
