// runtime.hip — host runtime of the MI355X solver: per-problem device
// state, the per-generation launch sequence (optionally captured once in a
// hipGraph and replayed), and the download of results into the caller's
// arrays.
//
// nraps_mc_solve() replaces the body of monte_carlo (src/mc_code.rs:276-380):
// same inputs, same SolutionResults arrays, same normalisation.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "launch.h"
#include "runtime.h"

namespace nraps {

DeviceError::DeviceError(const char *what, hipError_t e)
    : std::runtime_error(std::string(what) + ": " + hipGetErrorString(e)), err(e) {}

#define HIP_TRY NRAPS_HIP_TRY

DevArena::~DevArena() {
  for (void *p : ptrs_) (void)(pooled_ ? hipFreeAsync(p, stream_) : hipFree(p));
}

void *DevArena::raw(size_t bytes) {
  if (bytes == 0) bytes = 16;
  bytes = (bytes + 255) & ~(size_t)255;
  void *p = nullptr;
  hipError_t e = pooled_ ? hipMallocAsync(&p, bytes, stream_) : hipMalloc(&p, bytes);
  if (e != hipSuccess) throw DeviceError(pooled_ ? "hipMallocAsync" : "hipMalloc", e);
  ptrs_.push_back(p);
  bytes_ += bytes;
  return p;
}

void keep_pool_memory(int device) {
  static std::mutex mu;
  static std::vector<int> done;
  std::lock_guard<std::mutex> g(mu);
  if (std::find(done.begin(), done.end(), device) != done.end()) return;
  hipMemPool_t pool;
  NRAPS_HIP_TRY(hipDeviceGetDefaultMemPool(&pool, device));
  uint64_t keep = UINT64_MAX;
  NRAPS_HIP_TRY(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  done.push_back(device);
}

static int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

static int env_int(const char *name, int dflt) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  char *e = nullptr;
  long x = std::strtol(v, &e, 10);
  return (e && *e == '\0') ? (int)x : dflt;
}

int ktree_levels(int64_t n) {
  int L = 0;
  do {
    n = ceil_div(n, kTreeWidth);
    L++;
  } while (n > 1);
  return L;
}

// Unordered site-buffer slots: up to four sites per history plus the
// unused tails of the waves' 128-slot chunks (a few thousand waves).
int64_t site_capacity(int64_t N) { return 4 * N + (1 << 20); }

int64_t problem_footprint(const nraps_problem *pb, const nraps_run_opts *o) {
  const int64_t N = o->histories, nc = pb->n_cells, G = pb->n_groups;
  const int rs = o->precision == NRAPS_FP32 ? 4 : 8;
  int64_t b = 0;
  b += N * 8;                                  // kacc
  b += 2 * G * nc * 8;                         // generation tally
  b += 4 * G * nc * 8 + 4 * nc * 8;            // accumulators, finalize tables
  b += (ceil_div(N, kTreeWidth) + ceil_div(N, 65536) + 2) * 8;
  b += (int64_t)o->generations * 24;
  if (o->source == NRAPS_SOURCE_FISSION_BANK) {
    int64_t cap = site_capacity(N);
    b += N * (4 + 8 + rs + 4);                 // counts, offsets, source bank
    b += cap * (rs + 4 + 8) + cap * (rs + 4);  // unordered + ordered sites
    b += ceil_div(N, kScanTile) * 8;
  }
  return b;
}

// ---------------------------------------------------------------------------
// ProblemDev
// ---------------------------------------------------------------------------
void ProblemDev::prepare(const nraps_problem *pb_, const nraps_run_opts &o_, bool fp32_) {
  pb = pb_;
  o = o_;
  fp32 = fp32_;
  prep.build(pb);
  blob = prep.pack(fp32, &lay);
  fission = o.source == NRAPS_SOURCE_FISSION_BANK;
  width.resize(prep.nc);
  for (int c = 0; c < prep.nc; c++) width[c] = pb->mesh_right[c] - pb->mesh_left[c];
}

void ProblemDev::upload(bool event, CycleState *cs_slot, bool want_trace, nraps_result *r,
                        hipStream_t s) {
  stream = s;
  if (s) arena.bind(s);
  auto h2d = [&](void *dst, const void *src, size_t n) {
    HIP_TRY(s ? hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, s)
              : hipMemcpy(dst, src, n, hipMemcpyHostToDevice));
  };
  auto zero = [&](void *dst, size_t n) {
    HIP_TRY(s ? hipMemsetAsync(dst, 0, n, s) : hipMemset(dst, 0, n));
  };
  const Prepared &P = prep;
  const int64_t N = o.histories;
  const int rs = fp32 ? 4 : 8;
  const int nc = P.nc, GM = P.G * P.M, Gnc = P.G * P.nc, gens = o.generations;

  // Every small per-problem array lives in one block: the part the host
  // fills (tables, initial cycle state, finalize inputs) goes up in one
  // copy from a staging buffer, the accumulators after it are zeroed by one
  // fill and come back in one copy.
  size_t at = 0;
  auto put = [&](size_t bytes) {
    size_t off = at;
    at = (at + bytes + 255) & ~(size_t)255;
    return off;
  };
  Blk &B = blk;
  B.tab = put(blob.size());
  B.cs = cs_slot ? 0 : put(sizeof(CycleState));
  B.dx = put((size_t)nc * 8);
  B.width = put((size_t)nc * 8);
  B.nut = put((size_t)GM * 8);
  B.sigf = put((size_t)GM * 8);
  B.mat = put((size_t)nc * 4);
  B.up_end = at;
  B.k = put((size_t)gens * 8);
  B.kref = put((size_t)gens * 8);
  B.sites = put((size_t)gens * 8);
  B.flux = put((size_t)Gnc * 8);
  B.fs = put((size_t)nc * 8);
  B.fsum = put((size_t)Gnc * 8);
  B.fsq = put((size_t)Gnc * 8);
  B.tl = put((size_t)Gnc * 8);
  B.stats = put(sizeof(unsigned long long) * kNumStats);
  B.kpart = put((size_t)P.G * 8);
  B.end = at;
  // the large per-history and bank arrays follow in the same allocation:
  // one pool allocation (and one free) per problem
  const int64_t site_cap = fission ? site_capacity(N) : 0;
  const size_t o_kacc = put((size_t)N * 8);
  const size_t o_tally = put((size_t)2 * P.G * P.nc * 8);
  size_t o_counts = 0, o_offsets = 0, o_sx = 0, o_scell = 0, o_skey = 0, o_ox = 0, o_ocell = 0,
         o_srcx = 0, o_srccell = 0;
  if (fission) {
    o_counts = put((size_t)N * 4);
    o_offsets = put((size_t)N * 8);
    o_sx = put((size_t)site_cap * rs);
    o_scell = put((size_t)site_cap * 4);
    o_skey = put((size_t)site_cap * 8);
    o_ox = put((size_t)site_cap * rs);
    o_ocell = put((size_t)site_cap * 4);
    o_srcx = put((size_t)N * rs);
    o_srccell = put((size_t)N * 4);
  }
  const size_t o_tree_a = put((size_t)(ceil_div(N, kTreeWidth) + 1) * 8);
  const size_t o_tree_b = put((size_t)(ceil_div(N, 65536) + 1) * 8);
  const size_t total = at;
  uint8_t *base = arena.alloc<uint8_t>(total);
  d_base = base;

  std::memset(&cs0, 0, sizeof cs0);
  cs0.k_prev = o.k_initial;
  cs0.active = 1;
  const bool dstart = fission && o.bank_init == NRAPS_BANK_INIT_DIFFUSION;
  if (dstart) {
    // generation 0 is born from the diffusion bank and banks sites; its
    // 1/k is written on the device once the diffusion k is known
    cs0.from_bank = 1;
    cs0.make_sites = 1;
  }
  stage.assign(B.up_end, 0);
  std::memcpy(&stage[B.tab], blob.data(), blob.size());
  if (!cs_slot) std::memcpy(&stage[B.cs], &cs0, sizeof cs0);
  std::memcpy(&stage[B.dx], pb->delta_x, (size_t)nc * 8);
  std::memcpy(&stage[B.width], width.data(), (size_t)nc * 8);
  std::memcpy(&stage[B.nut], pb->nut, (size_t)GM * 8);
  std::memcpy(&stage[B.sigf], pb->sigf, (size_t)GM * 8);
  std::memcpy(&stage[B.mat], pb->matid, (size_t)nc * 4);
  h2d(base, stage.data(), B.up_end);
  zero(base + B.up_end, B.end - B.up_end);
  d_cs = cs_slot ? cs_slot : reinterpret_cast<CycleState *>(base + B.cs);
  if (cs_slot) h2d(cs_slot, &cs0, sizeof cs0);

  DevProblem &d = dev;
  std::memset(&d, 0, sizeof d);
  d.tab = base + B.tab;
  d.lay = lay;
  d.nc = P.nc;
  d.G = P.G;
  d.M = P.M;
  d.R = P.R;
  d.nf = P.nf;
  d.xs_every = o.xs_refresh == NRAPS_XS_EVERY_COLLISION;
  d.fission = fission;
  d.dstart = dstart ? 1 : 0;
  d.tally_bins = 2 * P.G * P.nc;
  d.nb = P.nb;
  d.inv_h = P.inv_h;
  d.N = N;
  d.seed = o.seed;
  d.problem_id = o.problem_id;
  d.max_events = o.max_events > 0 ? std::min<int64_t>(o.max_events, 2147483647) : 10000000;
  d.boundl = P.boundl;
  d.boundr = P.boundr;
  d.XL = P.XL;
  d.XR = P.XR;
  d.maj = P.maj;
  d.inv_maj = 1.0 / P.maj;
  d.cs = d_cs;
  d.k = reinterpret_cast<double *>(base + B.k);
  d.stats = reinterpret_cast<unsigned long long *>(base + B.stats);
  d.kacc = reinterpret_cast<double *>(base + o_kacc);
  d.tally_glob = reinterpret_cast<double *>(base + o_tally);
  if (fission) {
    d.counts = reinterpret_cast<int32_t *>(base + o_counts);
    d.offsets = reinterpret_cast<int64_t *>(base + o_offsets);
    d.site_cap = site_cap;
    d.site_x = base + o_sx;
    d.site_cell = reinterpret_cast<int32_t *>(base + o_scell);
    d.site_key = reinterpret_cast<uint64_t *>(base + o_skey);
    d.ord_x = base + o_ox;
    d.ord_cell = reinterpret_cast<int32_t *>(base + o_ocell);
    d.src_x = base + o_srcx;
    d.src_cell = reinterpret_cast<int32_t *>(base + o_srccell);
  }
  d.tree_a = reinterpret_cast<double *>(base + o_tree_a);
  d.tree_b = reinterpret_cast<double *>(base + o_tree_b);
  trace_n = 0;
  if (want_trace && o.trace_histories > 0 && r && r->trace && r->trace_len && r->draws && r->kacc) {
    trace_n = o.trace_histories;
    d.trace = arena.alloc<nraps_trace_event>(trace_n * o.trace_events);
    d.trace_len = arena.alloc<int32_t>(trace_n);
    d.draws = arena.alloc<int64_t>(trace_n);
    d.trace_kacc = arena.alloc<double>(trace_n);
    zero(d.trace, sizeof(nraps_trace_event) * trace_n * o.trace_events);
    zero(d.trace_len, sizeof(int32_t) * trace_n);
    d.trace_n = trace_n;
    d.trace_ev = o.trace_events;
    d.trace_cycle = o.trace_cycle;
  }

  if (dstart) {
    // diffusion k-eigenvalue of the problem, then N generation-0 sources at
    // the quantiles of its fission source (all on the device, in order on s)
    int dev = 0, lds_limit = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    nraps_diffusion_opts dopt;
    nraps_diffusion_opts_default(&dopt);
    dopt.max_iterations = kDiffusionStartIterations;
    diff.reset(new DiffusionSet());
    diff->add(pb, dopt, lds_limit);
    diff->upload(arena, s);
    diff->launch(s);
    double *cdf = arena.alloc<double>(nc);
    launch_diffusion_bank(fp32, diff->dev_problem(0), N, d.src_x, d.src_cell, d_cs, cdf, s);
  }

  DevFinal &f = fin;
  std::memset(&f, 0, sizeof f);
  f.dx = reinterpret_cast<const double *>(base + B.dx);
  f.width = reinterpret_cast<const double *>(base + B.width);
  f.nut = reinterpret_cast<const double *>(base + B.nut);
  f.sigf = reinterpret_cast<const double *>(base + B.sigf);
  f.mat = reinterpret_cast<const int32_t *>(base + B.mat);
  f.nut_conv = P.nut_conv;
  f.length = P.XR;
  f.generations = o.generations;
  f.skip = o.skip;
  f.tally_cycle = d.tally_glob;
  f.tally_last = reinterpret_cast<double *>(base + B.tl);
  f.flux_acc = reinterpret_cast<double *>(base + B.flux);
  f.fs_acc = reinterpret_cast<double *>(base + B.fs);
  f.fsum = reinterpret_cast<double *>(base + B.fsum);
  f.fsq = reinterpret_cast<double *>(base + B.fsq);
  f.k_ref = reinterpret_cast<double *>(base + B.kref);
  f.sites_hist = reinterpret_cast<int64_t *>(base + B.sites);
  f.kpart = reinterpret_cast<double *>(base + B.kpart);
  f.has_diff = event ? 1 : 0;
}

void ProblemDev::deactivate(hipStream_t s) {
  const int32_t zero = 0;
  HIP_TRY(hipMemcpyAsync(&d_cs->active, &zero, sizeof zero, hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));
}

int ProblemDev::download(nraps_result *r, double ms_total, double ms_device, double ms_transport,
                         int64_t launches, hipStream_t via) {
  const Prepared &P = prep;
  const int nc = P.nc, G = P.G, gens = o.generations, Gnc = G * nc;
  // every copy is queued on one stream (by default the set-up stream, after
  // the work of the run) and waited for once
  hipStream_t cs_stream = via ? via : stream;
  auto d2h = [&](void *dst, const void *src, size_t n) {
    HIP_TRY(cs_stream ? hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, cs_stream)
                      : hipMemcpy(dst, src, n, hipMemcpyDeviceToHost));
  };
  CycleState cs;
  std::vector<uint8_t> acc(blk.end - blk.up_end);
  d2h(&cs, d_cs, sizeof cs);
  d2h(acc.data(), d_base + blk.up_end, acc.size());
  if (r && trace_n > 0) {
    d2h(r->trace, dev.trace, sizeof(nraps_trace_event) * trace_n * o.trace_events);
    d2h(r->trace_len, dev.trace_len, 4 * trace_n);
    d2h(r->draws, dev.draws, 8 * trace_n);
    d2h(r->kacc, dev.trace_kacc, 8 * trace_n);
  }
  if (cs_stream) HIP_TRY(hipStreamSynchronize(cs_stream));
  auto take = [&](size_t off, size_t n, void *dst) {
    std::memcpy(dst, &acc[off - blk.up_end], n);
  };
  std::vector<double> k(gens), kref(gens), flux(Gnc), fs(nc), fsum(Gnc), fsq(Gnc), tl(Gnc);
  std::vector<int64_t> sites(gens);
  unsigned long long st[kNumStats];
  take(blk.k, (size_t)gens * 8, k.data());
  take(blk.kref, (size_t)gens * 8, kref.data());
  take(blk.sites, (size_t)gens * 8, sites.data());
  take(blk.flux, (size_t)Gnc * 8, flux.data());
  take(blk.fs, (size_t)nc * 8, fs.data());
  take(blk.fsum, (size_t)Gnc * 8, fsum.data());
  take(blk.fsq, (size_t)Gnc * 8, fsq.data());
  take(blk.tl, (size_t)Gnc * 8, tl.data());
  take(blk.stats, sizeof st, st);
  // generations actually run (the server may stop a problem early)
  const int ran = std::min(gens, (int)cs.cycle);
  const int skip = o.skip;
  if (!r) return NRAPS_OK;
  auto put = [](double *dst, const std::vector<double> &v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * 8);
  };
  put(r->k, k);
  put(r->k_ref, kref);
  put(r->flux, flux);
  put(r->fission_source, fs);
  put(r->tally_last, tl);
  if (r->sites) std::memcpy(r->sites, sites.data(), gens * 8);
  // average_assembly, src/mc_code.rs:259-274
  if (r->assembly_average) {
    std::fill(r->assembly_average, r->assembly_average + Gnc, 0.0);
    const int na = P.numass, per = nc / na;
    for (int g = 0; g < G && per > 0; g++)
      for (int a = 0; a < na; a++) {
        double s = 0.0;
        for (int c = a * per; c < (a + 1) * per; c++) s += flux[g * nc + c];
        double avg = s / (double)per;
        for (int c = a * per; c < (a + 1) * per; c++) r->assembly_average[g * nc + c] = avg;
      }
  }
  // k_fund, src/mc_code.rs:368-376
  if (r->k_fund) {
    std::fill(r->k_fund, r->k_fund + gens, 0.0);
    for (int gg = skip; gg < ran; gg++) {
      double s = 0.0;
      for (int x = skip; x <= gg; x++) s += k[x];
      r->k_fund[gg] = s / (double)(gg - (skip - 1));
    }
  }
  const int nact = std::max(ran - skip, 0);
  for (int j = 0; j < Gnc; j++) {
    double mean = nact > 0 ? fsum[j] / nact : 0.0;
    double var = nact > 1 ? (fsq[j] - fsum[j] * fsum[j] / nact) / (nact - 1) : 0.0;
    if (var < 0) var = 0;
    if (r->flux_mean) r->flux_mean[j] = mean;
    if (r->flux_sigma) r->flux_sigma[j] = nact > 0 ? std::sqrt(var / nact) : 0.0;
  }
  double s = 0, s2 = 0;
  for (int x = skip; x < ran; x++) {
    s += k[x];
    s2 += k[x] * k[x];
  }
  r->k_mean = nact > 0 ? s / nact : 0.0;
  double var = nact > 1 ? (s2 - s * s / nact) / (nact - 1) : 0.0;
  r->k_sigma = var > 0 ? std::sqrt(var / nact) : 0.0;
  r->histories_run = o.histories * (int64_t)ran;
  r->collisions = (int64_t)st[kStCollisions];
  r->virtual_collisions = (int64_t)st[kStVirtual];
  r->crossings = (int64_t)st[kStCrossings];
  r->absorptions = (int64_t)st[kStAbsorb];
  r->leaks = (int64_t)st[kStLeak];
  r->capped = (int64_t)st[kStCapped];
  r->mean_flights =
      r->histories_run > 0 ? (double)(r->collisions + r->leaks) / (double)r->histories_run : 0.0;
  r->ms_total = ms_total;
  r->ms_device = ms_device;
  r->ms_transport = ms_transport;
  r->transport_launches = launches;
  int status = NRAPS_OK;
  if (cs.error & 1)
    status = NRAPS_E_BANK_OVERFLOW;
  else if (cs.error & 2)
    status = NRAPS_E_NO_FISSION;
  else if (r->capped > 0)
    status = NRAPS_E_EVENT_CAP;
  r->status = status;
  if (status == NRAPS_E_BANK_OVERFLOW) set_error("fission sites exceeded the bank capacity");
  if (status == NRAPS_E_NO_FISSION) set_error("a generation banked no fission sites");
  if (status == NRAPS_E_EVENT_CAP) set_error("histories stopped at max_events");
  return status;
}

// ---------------------------------------------------------------------------
// Launcher
// ---------------------------------------------------------------------------
Launcher::Launcher(int device) : device_(device) {
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  n_cu_ = prop.multiProcessorCount;
  HIP_TRY(hipDeviceGetAttribute(&lds_limit_, hipDeviceAttributeMaxSharedMemoryPerBlock, device_));
}

Launcher::~Launcher() {
  drop_graph();
  for (void *p : owned_) (void)hipFree(p);
}

void Launcher::drop_graph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  graph_exec_ = nullptr;
  graph_ = nullptr;
}

void Launcher::build(const std::vector<ProblemDev *> &ps, bool fp32, bool event, hipStream_t s) {
  drop_graph();
  n_ = (int)ps.size();
  fp32_ = fp32;
  event_ = event;
  fission_any_ = false;
  max_n_ = 0;
  max_bins_ = 0;
  max_groups_ = 1;
  levels_ = 1;
  max_tiles_ = 0;
  int tab_max = 0;  // largest tables of one problem
  bool trace_any = false;
  int64_t tiles_total = 0, total = 0;
  std::vector<int64_t> tile_base(std::max(n_, 1));
  for (int i = 0; i < n_; i++) {
    const ProblemDev &p = *ps[i];
    fission_any_ |= p.fission;
    trace_any |= p.trace_n > 0;
    max_n_ = std::max<int64_t>(max_n_, p.o.histories);
    max_bins_ = std::max(max_bins_, p.dev.tally_bins);
    max_groups_ = std::max(max_groups_, p.dev.G);
    levels_ = std::max(levels_, ktree_levels(p.o.histories));
    tile_base[i] = tiles_total;
    const int64_t t = ceil_div(p.o.histories, kScanTile);
    tiles_total += t;
    max_tiles_ = std::max(max_tiles_, t);
    tab_max = std::max(tab_max, p.lay.hot);
    total += p.o.histories;
  }
  std::vector<int32_t> pass(std::max(n_, 1), 0);
  plans_[0] = plans_[1] = EventPlan{};
  if (event_ && n_ > 0) {
    // LDS plan.  The hot tables always go to LDS.  The tally goes there too,
    // right after the tables of the problem a block works on: with
    // 256-thread blocks when the largest tables + tally of any one problem
    // leave room for four blocks per CU, else with one 1024-thread block per
    // CU owning the whole LDS.  Problems whose tables + tally exceed even
    // that are run by a second launch of the cycle that scores into the
    // global FP64 tally, so they do not take the rest of the set with them.
    if (tab_max > lds_limit_)
      throw std::invalid_argument("problem tables (" + std::to_string(tab_max) +
                                  " B) exceed the LDS of one block");
    // tuning knobs (environment): LDS budgets of the two block shapes, and
    // the number of idle lanes that triggers a refill
    const int small_budget = std::min(lds_limit_ / 4, env_int("NRAPS_LDS_SMALL", lds_limit_ / 4));
    const int large_budget = std::min(lds_limit_ - 256, env_int("NRAPS_LDS_BUDGET", lds_limit_ - 256));
    refill_min_ = std::max(1, std::min(64, env_int("NRAPS_REFILL_MIN", 16)));
    int need0 = 0, tab1 = 0;
    int64_t total_pass[2] = {0, 0};
    for (int i = 0; i < n_; i++) {
      const ProblemDev &p = *ps[i];
      const int need = p.lay.hot + p.dev.tally_bins * (fp32 ? 4 : 8);
      pass[i] = need <= large_budget ? 0 : 1;
      if (pass[i] == 0)
        need0 = std::max(need0, need);
      else
        tab1 = std::max(tab1, p.lay.hot);
      total_pass[pass[i]] += p.o.histories;
    }
    for (int k = 0; k < 2; k++) {
      EventPlan &P = plans_[k];
      P.used = total_pass[k] > 0;
      if (!P.used) continue;
      P.v = WcVariant{};
      P.v.lds_tally = k == 0;
      P.v.big_block = k == 0 && need0 > small_budget;
      P.lds_bytes = (size_t)(k == 0 ? need0 : tab1);
      P.v.fp32 = fp32_;
      P.v.fission = fission_any_;
      P.v.trace = trace_any;
      P.v.region = ps[0]->o.method == NRAPS_EVENT_REGION;  // one method per batch
      (void)event_set_lds_limit(P.v, P.lds_bytes);
      P.grid = n_cu_ * event_blocks_per_cu(P.v, P.lds_bytes);
      const int64_t waves = (int64_t)P.grid * (P.v.block_size() / 64);
      int64_t c = total_pass[k] / (waves * 6);
      // about 6 claims per wave and launch (measured best between 4 and 8
      // on the 9-region and 201-region slabs); the cap bounds the tail a
      // last claim leaves at the end of the launch (forced 1024-history
      // claims are 11% slower than automatic ones on a 10^7-history fission
      // run) while large serving mixes (10^8 histories per launch) still get
      // few atomics per wave (a cap of 256 there costs 16% of throughput)
      const int claim_max = std::max(64, env_int("NRAPS_CLAIM_MAX", 1024));
      c = std::max<int64_t>(64, std::min<int64_t>(claim_max, c));
      P.claim = (int)((c + 63) / 64 * 64);
      if (n_ == 1 && ps[0]->o.chunk > 0) P.claim = (int)std::min<int64_t>(ps[0]->o.chunk, 1 << 20);
    }
  }
  if (n_ > cap_probs_) {
    if (d_probs_) (void)hipFree(d_probs_);
    if (d_fins_) (void)hipFree(d_fins_);
    if (d_tile_base_) (void)hipFree(d_tile_base_);
    owned_.erase(std::remove_if(owned_.begin(), owned_.end(),
                                [&](void *q) { return q == d_probs_ || q == d_fins_ || q == d_tile_base_; }),
                 owned_.end());
    cap_probs_ = std::max<int64_t>(n_, 2 * cap_probs_);
    HIP_TRY(hipMalloc(&d_probs_, sizeof(DevProblem) * cap_probs_));
    HIP_TRY(hipMalloc(&d_fins_, sizeof(DevFinal) * cap_probs_));
    HIP_TRY(hipMalloc(&d_tile_base_, sizeof(int64_t) * cap_probs_));
    owned_.push_back(d_probs_);
    owned_.push_back(d_fins_);
    owned_.push_back(d_tile_base_);
  }
  if (tiles_total > cap_tiles_ || !d_tile_sums_) {
    if (d_tile_sums_) {
      (void)hipFree(d_tile_sums_);
      owned_.erase(std::remove(owned_.begin(), owned_.end(), (void *)d_tile_sums_), owned_.end());
    }
    cap_tiles_ = std::max<int64_t>(std::max<int64_t>(tiles_total, 1), 2 * cap_tiles_);
    HIP_TRY(hipMalloc(&d_tile_sums_, sizeof(int64_t) * cap_tiles_));
    owned_.push_back(d_tile_sums_);
  }
  std::vector<DevProblem> dp(n_);
  std::vector<DevFinal> df(n_);
  for (int i = 0; i < n_; i++) {
    dp[i] = ps[i]->dev;
    dp[i].pass = pass[i];
    df[i] = ps[i]->fin;
  }
  if (n_ > 0) {
    // staged copies: the host vectors die with this call, so the async path
    // keeps them in the launcher until the next build
    h_probs_ = std::move(dp);
    h_fins_ = std::move(df);
    h_tile_base_ = std::move(tile_base);
    if (s) {
      HIP_TRY(hipMemcpyAsync(d_probs_, h_probs_.data(), sizeof(DevProblem) * n_, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(d_fins_, h_fins_.data(), sizeof(DevFinal) * n_, hipMemcpyHostToDevice, s));
      HIP_TRY(hipMemcpyAsync(d_tile_base_, h_tile_base_.data(), sizeof(int64_t) * n_,
                             hipMemcpyHostToDevice, s));
    } else {
      HIP_TRY(hipMemcpy(d_probs_, h_probs_.data(), sizeof(DevProblem) * n_, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_fins_, h_fins_.data(), sizeof(DevFinal) * n_, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(d_tile_base_, h_tile_base_.data(), sizeof(int64_t) * n_, hipMemcpyHostToDevice));
    }
  }
}

void Launcher::enqueue_cycle(hipStream_t s, hipEvent_t tr0, hipEvent_t tr1) {
  if (n_ == 0) return;
  launch_cycle_begin(d_probs_, n_, max_bins_, s);
  if (tr0) HIP_TRY(hipEventRecord(tr0, s));
  if (event_) {
    for (int k = 0; k < 2; k++) {
      const EventPlan &P = plans_[k];
      if (P.used)
        launch_event(P.v, d_probs_, n_, P.grid, P.lds_bytes, P.claim, refill_min_, k, s);
    }
  } else
    launch_surface(d_probs_, n_, max_n_, s);
  if (tr1) HIP_TRY(hipEventRecord(tr1, s));
  int64_t groups = max_n_;
  for (int l = 0; l < levels_; l++) {
    groups = ceil_div(groups, kTreeWidth);
    launch_ktree(d_probs_, n_, l, groups, s);
  }
  if (fission_any_) launch_bank(fp32_, d_probs_, n_, max_n_, max_tiles_, d_tile_sums_, d_tile_base_, s);
  launch_finalize(d_probs_, d_fins_, n_, max_groups_, levels_, s);
  HIP_TRY(hipGetLastError());
}

void Launcher::capture(hipStream_t s) {
  drop_graph();
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  enqueue_cycle(s, nullptr, nullptr);
  HIP_TRY(hipStreamEndCapture(s, &graph_));
  HIP_TRY(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
}

void Launcher::launch_graph(hipStream_t s) { HIP_TRY(hipGraphLaunch(graph_exec_, s)); }

// ---------------------------------------------------------------------------
// Solve a batch to completion
// ---------------------------------------------------------------------------
namespace {
struct StreamGuard {
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
  ~StreamGuard() {
    for (auto e : ev) (void)hipEventDestroy(e);
    if (s) (void)hipStreamDestroy(s);
  }
  hipEvent_t make() {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    ev.push_back(e);
    return e;
  }
};
}  // namespace

int solve_batch(int n, const nraps_problem *const *pbs, const nraps_run_opts *opts,
                nraps_result *const *results) {
  auto t0 = std::chrono::steady_clock::now();
  try {
    if (n < 1 || !pbs || !opts) {
      set_error("empty batch");
      return NRAPS_E_INVALID;
    }
    for (int i = 0; i < n; i++) {
      std::string why;
      if (validate(pbs[i], &opts[i], &why)) {
        set_error(n > 1 ? "problem " + std::to_string(i) + ": " + why : why);
        return NRAPS_E_INVALID;
      }
      if (opts[i].precision != opts[0].precision || opts[i].method != opts[0].method) {
        set_error("precision and method must agree across a batch");
        return NRAPS_E_INVALID;
      }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      set_error("no HIP device available");
      return NRAPS_E_DEVICE;
    }
    const int device = opts[0].device;
    if (device < 0 || device >= ndev) {
      set_error("device ordinal out of range");
      return NRAPS_E_DEVICE;
    }
    HIP_TRY(hipSetDevice(device));
    const bool fp32 = opts[0].precision == NRAPS_FP32;
    const bool event = opts[0].method != NRAPS_HISTORY_SURFACE;  // event pipeline
    const bool use_graph = opts[0].use_graph != 0;
    // the stream outlives the problems: their memory is freed on it
    StreamGuard sg;
    HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    std::vector<std::unique_ptr<ProblemDev>> owned(n);
    std::vector<ProblemDev *> ps(n);
    int gens_max = 0;
    for (int i = 0; i < n; i++) {
      owned[i].reset(new ProblemDev());
      owned[i]->setup(pbs[i], opts[i], fp32, event, nullptr, true, results ? results[i] : nullptr,
                      sg.s);
      ps[i] = owned[i].get();
      gens_max = std::max(gens_max, (int)opts[i].generations);
    }
    Launcher L(device);
    L.build(ps, fp32, event, sg.s);
    hipEvent_t e0 = sg.make(), e1 = sg.make();
    std::vector<hipEvent_t> tr;
    HIP_TRY(hipEventRecord(e0, sg.s));
    if (use_graph) {
      L.capture(sg.s);
      for (int c = 0; c < gens_max; c++) L.launch_graph(sg.s);
    } else {
      for (int c = 0; c < 2 * gens_max; c++) tr.push_back(sg.make());
      for (int c = 0; c < gens_max; c++) L.enqueue_cycle(sg.s, tr[2 * c], tr[2 * c + 1]);
    }
    HIP_TRY(hipEventRecord(e1, sg.s));
    HIP_TRY(hipStreamSynchronize(sg.s));
    float ms_dev = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms_dev, e0, e1));
    double ms_tr = 0.0;
    for (size_t c = 0; c + 1 < tr.size(); c += 2) {
      float t = 0.f;
      HIP_TRY(hipEventElapsedTime(&t, tr[c], tr[c + 1]));
      ms_tr += t;
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    int worst = NRAPS_OK;
    for (int i = 0; i < n; i++) {
      int st = ps[i]->download(results ? results[i] : nullptr, ms, ms_dev, ms_tr,
                               use_graph ? 0 : gens_max);
      if (st != NRAPS_OK && worst == NRAPS_OK) worst = st;
    }
    return worst;
  } catch (const DeviceError &e) {
    set_error(e.what());
    return e.err == hipErrorOutOfMemory ? NRAPS_E_NOMEM : NRAPS_E_DEVICE;
  } catch (const std::invalid_argument &e) {
    set_error(e.what());
    return NRAPS_E_INVALID;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return NRAPS_E_NOMEM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_DEVICE;
  }
}

}  // namespace nraps

using namespace nraps;

extern "C" int nraps_mc_solve(const nraps_problem *problem, const nraps_run_opts *opts,
                              nraps_result *result) {
  const nraps_problem *pbs[1] = {problem};
  nraps_result *rs[1] = {result};
  if (!opts) {
    set_error("opts is NULL");
    return NRAPS_E_INVALID;
  }
  int st = solve_batch(1, pbs, opts, rs);
  if (result) result->status = st;
  return st;
}

extern "C" int nraps_mc_solve_batch(int32_t n, const nraps_problem *const *problems,
                                    const nraps_run_opts *opts, nraps_result *const *results) {
  return solve_batch(n, problems, opts, results);
}

extern "C" int nraps_device_info(int32_t device, char *buf, int32_t buflen) {
  if (!buf || buflen < 1) return NRAPS_E_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || device < 0 || device >= ndev) {
    std::snprintf(buf, buflen, "nraps %d no-device", NRAPS_ABI_VERSION);
    set_error("no HIP device available");
    return NRAPS_E_DEVICE;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, device) != hipSuccess) return NRAPS_E_DEVICE;
  int lds = 0;
  (void)hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  std::snprintf(buf, buflen, "nraps %d %s cus=%d lds_per_block=%d hbm_bytes=%zu clock_khz=%d",
                NRAPS_ABI_VERSION, p.gcnArchName, p.multiProcessorCount, lds, (size_t)p.totalGlobalMem,
                p.clockRate);
  return NRAPS_OK;
}

extern "C" int nraps_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n_streams, int32_t draws,
                                uint32_t *out) {
  try {
    if (n_streams < 1 || draws < 1 || !out) {
      set_error("bad arguments");
      return NRAPS_E_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      set_error("no HIP device available");
      return NRAPS_E_DEVICE;
    }
    DevArena a;
    uint32_t *d = a.alloc<uint32_t>((size_t)n_streams * draws);
    launch_pcg32_fill(seed, stream0, n_streams, draws, d, nullptr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpy(out, d, (size_t)n_streams * draws * 4, hipMemcpyDeviceToHost));
    return NRAPS_OK;
  } catch (const DeviceError &e) {
    set_error(e.what());
    return NRAPS_E_DEVICE;
  }
}
