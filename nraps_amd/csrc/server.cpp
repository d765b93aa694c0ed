// server.cpp — device-resident request loop with continuous batching.
//
// One owner thread drives the device.  Between generations it admits queued
// requests while slots and the memory budget (bank + tally bytes) allow,
// then runs one generation of every resident problem in the same launches
// (the persistent event kernel visits all of them), reads back the per-slot
// generation state in one copy, and retires requests that reached their
// generation count or their relative k standard-error target.  A retired request's
// device memory is released at once, so queued work is admitted while the
// rest keeps running.  Every request keeps its own RNG streams (problem_id),
// so its result does not depend on what else is resident.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "runtime.h"

namespace nraps {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  ProblemBuf pb;
  nraps_run_opts o{};
  int64_t footprint = 0;
  enum State { QUEUED, RUNNING, DONE } state = QUEUED;
  bool cancel = false;
  int status = NRAPS_OK;
  std::unique_ptr<ProblemDev> dev;
  int slot = -1;
  Clock::time_point t_submit, t_admit, t_done;
  int cycles_run = 0;
  double ks = 0, ks2 = 0;
  int kn = 0;
  // results
  std::vector<double> flux, aa, fs, k, kf, fmean, fsig, tl, kref;
  std::vector<int64_t> sites;
  nraps_result res{};
};

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace
}  // namespace nraps

using namespace nraps;

struct nraps_server {
  nraps_server_opts opts{};
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<Request *> pending;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::vector<Request *> resident;
  int64_t used = 0;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::string fatal;
  std::thread th;
  // device side (owner thread only)
  CycleState *d_slots = nullptr;
  std::vector<CycleState> h_slots;
  std::vector<int> free_slots;
  std::unique_ptr<Launcher> L;
  hipStream_t s = nullptr;   // set-up, cycles and slot read-back
  hipStream_t s_out = nullptr;  // result read-back of retiring requests

  void loop();
  void retire(Request *r, int status);
};

void nraps_server::retire(Request *r, int status) {
  // owner thread, stream idle: copy results out and release the device memory
  if (r->dev) {
    try {
      if (status != NRAPS_E_CANCELLED) {
        const int nc = r->pb.view.n_cells, G = r->pb.view.n_groups, gens = r->o.generations;
        r->flux.assign((size_t)G * nc, 0.0);
        r->aa.assign((size_t)G * nc, 0.0);
        r->fs.assign(nc, 0.0);
        r->k.assign(gens, 0.0);
        r->kf.assign(gens, 0.0);
        r->fmean.assign((size_t)G * nc, 0.0);
        r->fsig.assign((size_t)G * nc, 0.0);
        r->tl.assign((size_t)G * nc, 0.0);
        r->kref.assign(gens, 0.0);
        r->sites.assign(gens, 0);
        nraps_result &o = r->res;
        std::memset(&o, 0, sizeof o);
        o.flux = r->flux.data();
        o.assembly_average = r->aa.data();
        o.fission_source = r->fs.data();
        o.k = r->k.data();
        o.k_fund = r->kf.data();
        o.flux_mean = r->fmean.data();
        o.flux_sigma = r->fsig.data();
        o.tally_last = r->tl.data();
        o.k_ref = r->kref.data();
        o.sites = r->sites.data();
        double total = ms_between(r->t_submit, Clock::now());
        // the request's device work has completed; its read-back runs on
        // its own stream, beside the cycle already queued for the others
        int st = r->dev->download(&o, total, 0.0, 0.0, r->cycles_run, s_out);
        if (status == NRAPS_OK) status = st;
      }
    } catch (const std::exception &e) {
      status = NRAPS_E_DEVICE;
      set_error(e.what());
    }
    r->dev.reset();  // memory returns to the pool, ordered after the queued cycle
  }
  std::lock_guard<std::mutex> g(m);
  if (r->slot >= 0) free_slots.push_back(r->slot);
  r->slot = -1;
  used -= r->footprint;
  r->status = status;
  r->res.status = status;
  r->t_done = Clock::now();
  r->state = Request::DONE;
  resident.erase(std::remove(resident.begin(), resident.end(), r), resident.end());
  cv_done.notify_all();
}

// One iteration of the owner loop:
//   1. admit queued requests (first fit by memory footprint and free slots)
//      and build their host tables -- while the GPU runs the cycle queued by
//      the previous iteration;
//   2. wait for that cycle, read every slot's generation state, and take
//      the requests that finished (or were cancelled) out of the resident
//      set;
//   3. queue the device set-up of the admitted requests, rebuild the launch
//      plan if membership changed, and queue the next cycle of every
//      resident request together with the read-back of the slot states;
//   4. read back and release the finished requests while that cycle runs.
// Set-up, cycles and slot read-back are ordered on one stream, result
// read-back on a second one; allocations come from the stream-ordered pool,
// so set-up and retirement do not stall the device.
void nraps_server::loop() {
  try {
    NRAPS_HIP_TRY(hipSetDevice(opts.device));
    NRAPS_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    NRAPS_HIP_TRY(hipStreamCreateWithFlags(&s_out, hipStreamNonBlocking));
    keep_pool_memory(opts.device);
    L.reset(new Launcher(opts.device));
    bool dirty = true;
    const bool fp32 = opts.precision == NRAPS_FP32;
    std::vector<Request *> inflight;  // requests of the queued cycle
    for (;;) {
      std::vector<Request *> admit;
      {
        std::unique_lock<std::mutex> lk(m);
        cv_work.wait(lk, [&] {
          return stop || !pending.empty() || !resident.empty() || !inflight.empty();
        });
        if (stop) break;
        for (auto it = pending.begin(); it != pending.end();) {
          Request *r = *it;
          if (r->cancel) {
            r->state = Request::DONE;
            r->status = NRAPS_E_CANCELLED;
            r->t_done = Clock::now();
            it = pending.erase(it);
            cv_done.notify_all();
            continue;
          }
          if ((int)(resident.size() + admit.size()) < opts.max_slots && !free_slots.empty() &&
              used + r->footprint <= opts.budget_bytes) {
            used += r->footprint;
            r->slot = free_slots.back();
            free_slots.pop_back();
            r->state = Request::RUNNING;
            admit.push_back(r);
            it = pending.erase(it);
          } else {
            ++it;
          }
        }
      }
      // host half of the set-up, overlapping the queued cycle
      std::vector<Request *> ready;
      for (Request *r : admit) {
        r->t_admit = Clock::now();
        try {
          r->dev.reset(new ProblemDev());
          r->dev->prepare(&r->pb.view, r->o, fp32);
          ready.push_back(r);
        } catch (const std::exception &e) {
          set_error(e.what());
          r->dev.reset();
          retire(r, NRAPS_E_INVALID);
        }
      }
      // the queued cycle: wait, then read the generation state of its slots
      std::vector<std::pair<Request *, int>> done_list;
      if (!inflight.empty()) {
        NRAPS_HIP_TRY(hipStreamSynchronize(s));
        for (Request *r : inflight) {
          const CycleState &cs = h_slots[r->slot];
          if (cs.cycle > r->cycles_run) {
            const int done_gen = cs.cycle - 1;  // generation just closed
            if (done_gen >= r->o.skip) {
              r->ks += cs.k_prev;
              r->ks2 += cs.k_prev * cs.k_prev;
              r->kn++;
            }
            r->cycles_run = cs.cycle;
          }
          bool cancel;
          {
            std::lock_guard<std::mutex> g(m);
            cancel = r->cancel;
          }
          bool finished = !cs.active;
          if (!finished && opts.sigma_k_target > 0 && r->kn >= std::max(2, opts.min_active)) {
            double n = r->kn, mean = r->ks / n;
            double var = (r->ks2 - n * mean * mean) / (n - 1);
            double sem = var > 0 ? std::sqrt(var / n) : 0.0;
            if (sem <= opts.sigma_k_target * std::fabs(mean)) finished = true;
          }
          if (cancel || finished) {
            // out of the resident set now, so the next launch plan leaves it
            // out; its results are read back once that cycle is queued
            std::lock_guard<std::mutex> g(m);
            resident.erase(std::remove(resident.begin(), resident.end(), r), resident.end());
            done_list.emplace_back(r, cancel ? NRAPS_E_CANCELLED : NRAPS_OK);
            dirty = true;
          }
        }
        inflight.clear();
      }
      // device half of the set-up, queued on the stream
      for (Request *r : ready) {
        try {
          r->dev->upload(true, &d_slots[r->slot], false, nullptr, s);
          std::lock_guard<std::mutex> g(m);
          resident.push_back(r);
          dirty = true;
        } catch (const std::invalid_argument &e) {
          // e.g. a diffusion-started bank whose diffusion matrix is singular
          set_error(e.what());
          NRAPS_HIP_TRY(hipStreamSynchronize(s));
          r->dev.reset();
          retire(r, NRAPS_E_INVALID);
        } catch (const std::exception &e) {
          set_error(e.what());
          NRAPS_HIP_TRY(hipStreamSynchronize(s));
          r->dev.reset();
          retire(r, NRAPS_E_NOMEM);
        }
      }
      std::vector<Request *> run;
      {
        std::lock_guard<std::mutex> g(m);
        run = resident;
      }
      if (run.empty()) {
        for (auto &f : done_list) retire(f.first, f.second);
        continue;
      }
      if (dirty) {
        std::vector<ProblemDev *> ps;
        for (Request *r : run) ps.push_back(r->dev.get());
        L->build(ps, fp32, true, s);
        if (opts.use_graph) L->capture(s);
        dirty = false;
      }
      if (opts.use_graph)
        L->launch_graph(s);
      else
        L->enqueue_cycle(s, nullptr, nullptr);
      NRAPS_HIP_TRY(hipMemcpyAsync(h_slots.data(), d_slots, sizeof(CycleState) * h_slots.size(),
                                   hipMemcpyDeviceToHost, s));
      inflight = run;
      for (auto &f : done_list) retire(f.first, f.second);
    }
  } catch (const std::exception &e) {
    std::lock_guard<std::mutex> g(m);
    fatal = e.what();
    for (auto &kv : reqs) {
      Request *r = kv.second.get();
      if (r->state != Request::DONE) {
        r->state = Request::DONE;
        r->status = NRAPS_E_DEVICE;
        r->t_done = Clock::now();
      }
    }
    pending.clear();
    cv_done.notify_all();
  }
  // release the device memory of every request still holding some (while
  // the stream its pool allocations are ordered on is alive)
  if (s) (void)hipStreamSynchronize(s);
  if (s_out) (void)hipStreamSynchronize(s_out);
  std::vector<Request *> left;
  {
    std::lock_guard<std::mutex> g(m);
    for (auto &kv : reqs)
      if (kv.second->dev) left.push_back(kv.second.get());
  }
  const bool failed = !fatal.empty();
  for (Request *r : left) {
    if (failed) {
      r->dev.reset();
      std::lock_guard<std::mutex> g(m);
      resident.erase(std::remove(resident.begin(), resident.end(), r), resident.end());
    } else {
      retire(r, NRAPS_E_CANCELLED);
    }
  }
  L.reset();
  if (s) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  if (s_out) (void)hipStreamDestroy(s_out);
  s = nullptr;
  s_out = nullptr;
}

extern "C" int nraps_server_create(const nraps_server_opts *o, nraps_server **out) {
  if (!o || !out) {
    set_error("opts/out is NULL");
    return NRAPS_E_INVALID;
  }
  *out = nullptr;
  if (o->max_slots < 1 || o->budget_bytes <= 0 ||
      (o->precision != NRAPS_FP32 && o->precision != NRAPS_FP64) ||
      (o->method != 0 && o->method != NRAPS_EVENT_REGION && o->method != NRAPS_EVENT_WOODCOCK)) {
    set_error("server needs max_slots >= 1, budget_bytes > 0, a valid precision and an event method");
    return NRAPS_E_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1 || o->device < 0 || o->device >= ndev) {
    set_error("no HIP device available");
    return NRAPS_E_DEVICE;
  }
  auto *srv = new nraps_server();
  srv->opts = *o;
  try {
    NRAPS_HIP_TRY(hipSetDevice(o->device));
    NRAPS_HIP_TRY(hipMalloc(&srv->d_slots, sizeof(CycleState) * o->max_slots));
    NRAPS_HIP_TRY(hipMemset(srv->d_slots, 0, sizeof(CycleState) * o->max_slots));
  } catch (const std::exception &e) {
    set_error(e.what());
    delete srv;
    return NRAPS_E_DEVICE;
  }
  srv->h_slots.assign(o->max_slots, CycleState{});
  for (int i = o->max_slots - 1; i >= 0; i--) srv->free_slots.push_back(i);
  srv->th = std::thread([srv] { srv->loop(); });
  *out = srv;
  return NRAPS_OK;
}

extern "C" int nraps_server_submit(nraps_server *srv, const nraps_problem *pb, const nraps_run_opts *o,
                                   uint64_t *ticket) {
  if (!srv || !pb || !o || !ticket) {
    set_error("bad arguments");
    return NRAPS_E_INVALID;
  }
  nraps_run_opts oo = *o;
  oo.method = srv->opts.method == NRAPS_EVENT_WOODCOCK ? NRAPS_EVENT_WOODCOCK : NRAPS_EVENT_REGION;
  oo.precision = srv->opts.precision;
  oo.trace_histories = 0;
  std::string why;
  if (validate(pb, &oo, &why)) {
    set_error(why);
    return NRAPS_E_INVALID;
  }
  auto r = std::make_unique<Request>();
  r->pb.copy_from(pb);
  r->o = oo;
  r->footprint = problem_footprint(pb, &oo);
  if (r->footprint > srv->opts.budget_bytes) {
    set_error("request footprint exceeds the server memory budget");
    return NRAPS_E_OVER_BUDGET;
  }
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> g(srv->m);
  if (!srv->fatal.empty()) {
    set_error(srv->fatal);
    return NRAPS_E_DEVICE;
  }
  r->ticket = srv->next_ticket++;
  *ticket = r->ticket;
  srv->pending.push_back(r.get());
  srv->reqs[r->ticket] = std::move(r);
  srv->cv_work.notify_one();
  return NRAPS_OK;
}

static void copy_out(const Request *r, nraps_result *dst) {
  if (!dst) return;
  const nraps_result &s = r->res;
  auto cp = [](double *d, const std::vector<double> &v) {
    if (d && !v.empty()) std::memcpy(d, v.data(), v.size() * 8);
  };
  cp(dst->flux, r->flux);
  cp(dst->assembly_average, r->aa);
  cp(dst->fission_source, r->fs);
  cp(dst->k, r->k);
  cp(dst->k_fund, r->kf);
  cp(dst->flux_mean, r->fmean);
  cp(dst->flux_sigma, r->fsig);
  cp(dst->tally_last, r->tl);
  cp(dst->k_ref, r->kref);
  if (dst->sites && !r->sites.empty()) std::memcpy(dst->sites, r->sites.data(), r->sites.size() * 8);
  dst->k_mean = s.k_mean;
  dst->k_sigma = s.k_sigma;
  dst->histories_run = s.histories_run;
  dst->collisions = s.collisions;
  dst->virtual_collisions = s.virtual_collisions;
  dst->crossings = s.crossings;
  dst->absorptions = s.absorptions;
  dst->leaks = s.leaks;
  dst->capped = s.capped;
  dst->mean_flights = s.mean_flights;
  dst->ms_total = ms_between(r->t_submit, r->t_done);
  dst->ms_device = 0;
  dst->ms_transport = 0;
  dst->transport_launches = r->cycles_run;
  dst->status = r->status;
}

extern "C" int nraps_server_poll(nraps_server *srv, uint64_t ticket, nraps_result *res) {
  if (!srv) return NRAPS_E_INVALID;
  std::lock_guard<std::mutex> g(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    set_error("unknown ticket");
    return NRAPS_E_NOT_FOUND;
  }
  Request *r = it->second.get();
  if (r->state != Request::DONE) return NRAPS_PENDING;
  copy_out(r, res);
  if (r->status != NRAPS_OK) set_error(r->status == NRAPS_E_CANCELLED ? "request cancelled" : "request failed");
  return r->status;
}

extern "C" int nraps_server_wait(nraps_server *srv, uint64_t ticket, nraps_result *res,
                                 int64_t timeout_ms) {
  if (!srv) return NRAPS_E_INVALID;
  std::unique_lock<std::mutex> lk(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    set_error("unknown ticket");
    return NRAPS_E_NOT_FOUND;
  }
  Request *r = it->second.get();
  auto done = [&] { return r->state == Request::DONE; };
  if (timeout_ms < 0)
    srv->cv_done.wait(lk, done);
  else if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
    return NRAPS_PENDING;
  copy_out(r, res);
  if (r->status != NRAPS_OK) set_error(r->status == NRAPS_E_CANCELLED ? "request cancelled" : "request failed");
  return r->status;
}

extern "C" int nraps_server_cancel(nraps_server *srv, uint64_t ticket) {
  if (!srv) return NRAPS_E_INVALID;
  std::lock_guard<std::mutex> g(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    set_error("unknown ticket");
    return NRAPS_E_NOT_FOUND;
  }
  it->second->cancel = true;
  srv->cv_work.notify_one();
  return NRAPS_OK;
}

extern "C" int nraps_server_latency(nraps_server *srv, uint64_t ticket, double *queued_ms,
                                    double *total_ms, int32_t *cycles_run) {
  if (!srv) return NRAPS_E_INVALID;
  std::lock_guard<std::mutex> g(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    set_error("unknown ticket");
    return NRAPS_E_NOT_FOUND;
  }
  Request *r = it->second.get();
  if (r->state != Request::DONE) return NRAPS_PENDING;
  const bool admitted = r->t_admit.time_since_epoch().count() != 0;
  if (queued_ms) *queued_ms = admitted ? ms_between(r->t_submit, r->t_admit) : ms_between(r->t_submit, r->t_done);
  if (total_ms) *total_ms = ms_between(r->t_submit, r->t_done);
  if (cycles_run) *cycles_run = r->cycles_run;
  return NRAPS_OK;
}

extern "C" int nraps_server_forget(nraps_server *srv, uint64_t ticket) {
  if (!srv) return NRAPS_E_INVALID;
  std::lock_guard<std::mutex> g(srv->m);
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) {
    set_error("unknown ticket");
    return NRAPS_E_NOT_FOUND;
  }
  if (it->second->state != Request::DONE) return NRAPS_PENDING;
  srv->reqs.erase(it);
  return NRAPS_OK;
}

extern "C" void nraps_server_destroy(nraps_server *srv) {
  if (!srv) return;
  {
    std::lock_guard<std::mutex> g(srv->m);
    srv->stop = true;
    for (auto &kv : srv->reqs) kv.second->cancel = true;
  }
  srv->cv_work.notify_all();
  if (srv->th.joinable()) srv->th.join();
  if (srv->d_slots) (void)hipFree(srv->d_slots);
  delete srv;
}
