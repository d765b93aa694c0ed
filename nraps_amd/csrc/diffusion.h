// diffusion.h — multigroup finite-difference diffusion on the device: the
// k-eigenvalue solver of src/discrete.rs:169-356 and the fission-bank start
// that uses its fission source.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <vector>

#include "../../include/nraps.h"
#include "device.h"

namespace nraps {

class DevArena;

// One workgroup of BS threads (256 or 1024) per problem; each thread owns
// L = ceil(n/BS) consecutive cells.  Per-cell arrays are stored thread-major:
// cell i = t*L + j sits at j*BS + t, so a wave's loads are contiguous and LDS
// accesses hit distinct banks.
// Iteration cap of the diffusion solve behind a diffusion-started bank (its
// flux is only a starting guess; the reference's own rule stops far earlier).
constexpr int kDiffusionStartIterations = 20000;

// One problem of k_diffusion (one workgroup each).
struct DiffProblem {
  int32_t n, G, M, L;        // cells, groups, materials, cells per thread
  int32_t bs;                // threads per workgroup the layout is made for
  int32_t in_lds;            // flux and fission rate live in LDS (else gflux/gF)
  int32_t max_it;
  double ftol, ktol;
  int32_t pos0, pos1, pos2;  // thread-major slots of cells 0, n-2, n-1
  int32_t own0, own1, own2;  // threads owning them
  // Thomas factors per group, thread-major [G][L*bs]:
  //   forward  y_i = alpha_i y_{i-1} + m_i rhs_i
  //   backward x_i = gamma_i x_{i+1} + y_i
  const double *alpha, *mdiag, *gamma;
  const int32_t *mat;        // [L*bs] thread-major
  const double *dx;          // [L*bs] thread-major
  const double *dx_nat;      // [n] cell order
  const double *left, *right; // [n] mesh edges (bank start)
  const double *nsf, *chi, *nut, *sigf; // [G*M] ([g][mat])
  const double *scat;        // [M*G*G] [mat][from][to]
  double *gflux, *gF;        // global fallback for flux [G][L*bs] and rate [L*bs]
  double *flux_out;          // [G*n] power-normalised, cell order
  double *fs_out;            // [n] sum_g nut*sigf*flux
  double *scal;              // [5]: k, delta_flux, delta_k, iterations, status
};

// Diffusion problems prepared on the host and staged into one device block.
class DiffusionSet {
 public:
  // Host half: checks, group matrices and their factors (no HIP).  Throws
  // std::invalid_argument on a malformed problem.
  void add(const nraps_problem *pb, const nraps_diffusion_opts &o, int lds_limit);
  // Device half: allocation from `arena` and one upload on `s`.
  void upload(DevArena &arena, hipStream_t s);
  // One launch for every problem of the set.
  void launch(hipStream_t s);
  // Copy every problem's outputs back (one copy; synchronises s), then
  // unpack problem i's results into r.
  void fetch(hipStream_t s);
  int download(int i, nraps_diffusion_result *r);
  const DiffProblem *dev_problem(int i) const { return d_probs_ + slot_[i]; }
  int size() const { return (int)host_.size(); }

 private:
  // Workgroup size for the set, thread-major layouts, LDS sizes.
  void layout();
  struct Host {
    const nraps_problem *pb;
    DiffProblem d;            // sizes and flags (pointers filled at upload)
    std::vector<double> alpha, mdiag, gamma;        // factors, [G][n] cell order
    std::vector<double> p_alpha, p_mdiag, p_gamma;  // the same, thread-major
    std::vector<double> dx, nsf, chi, nut, sigf;
    std::vector<int32_t> mat;
  };
  std::vector<Host> host_;
  // one launch per workgroup size; its problems occupy device slots
  // [first, first + count)
  struct Launch {
    int bs = 256, first = 0, count = 0, max_G = 0, max_L = 0;
    size_t lds_generic = 0;  // LDS of the LDS/global form (tables + flux when it fits)
    size_t tab_bytes = 0;    // tables + scratch (the register form adds its slots)
  };
  std::vector<Launch> launches_;
  std::vector<int> slot_;  // device slot of problem i
  int lds_limit_ = 0;
  DiffProblem *d_probs_ = nullptr;
  std::vector<DiffProblem> h_probs_;
  std::vector<uint8_t> stage_;  // source of the asynchronous upload
  uint8_t *d_base_ = nullptr;
  size_t out_begin_ = 0, out_end_ = 0;           // device byte range of the outputs
  std::vector<std::array<size_t, 3>> out_off_;   // per problem: flux, fs, scal offsets
  std::vector<uint8_t> out_;                     // host copy of the outputs
};

// Fission-bank start of a Monte Carlo problem: N sources at the quantiles
// (p + 1/2)/N of problem dp's fission source, generation-0 state of cs
// (1/k_diffusion for site production).  cdf: n doubles of scratch.
void launch_diffusion_bank(bool fp32, const DiffProblem *dp, int64_t N, void *src_x,
                           int32_t *src_cell, CycleState *cs, double *cdf, hipStream_t s);

}  // namespace nraps
