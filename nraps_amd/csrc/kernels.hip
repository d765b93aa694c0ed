// kernels.hip — HIP kernels of the MI355X Monte Carlo k-eigenvalue solver.
//
// Generation pipeline (every kernel takes the batch's DevProblem table and
// works on any number of problems; blockIdx.y or the chunk list selects the
// problem):
//
//   k_cycle_begin      reset the history queues, zero the generation tallies
//   k_surface          plain FP64 port: one thread per history, cell-by-cell
//                      surface tracking, global FP64 atomics
//                      (src/mc_code.rs:134-257)
//   k_woodcock<Real>   event pipeline: persistent blocks stage a problem's
//                      tables in LDS; each wave claims ranges of histories with
//                      one atomic and refills idle lanes by ballot + mbcnt;
//                      Woodcock flights, collisions and difference-array
//                      track-length tallies in LDS (flushed once per block with
//                      FP64 atomics); fission sites appended by wave ballot
//                      (one global atomic per wave)
//   k_ktree            canonical 256-wide pairwise sum of per-history k
//   k_scan_*, k_place, k_resample
//                      fission bank: sites placed in (history, sequence) order
//                      by a prefix sum of per-history counts, then systematic
//                      resampling to N sources in FP64
//   k_finalize         k, reference normalisation and accumulation
//                      (src/mc_code.rs:340-363), variance sums, next-cycle state
//
// Compiled with -ffp-contract=off: the FP64 histories must reproduce the CPU
// replay bit for bit.
#include <type_traits>

#include "device.h"
#include "launch.h"
#include "nraps_core.h"

namespace nraps {

// ---------------------------------------------------------------------------
// Precision traits
// ---------------------------------------------------------------------------
template <typename Real>
struct Num;
template <>
struct Num<double> {
  __device__ static double u01(uint32_t u) { return nraps_u01_f64(u); }
  __device__ static double flight(uint32_t u) { return nraps_uflight_f64(u); }
  __device__ static double nlog(double v) { return nraps_log(v); }
};
template <>
struct Num<float> {
  __device__ static float u01(uint32_t u) { return nraps_u01_f32(u); }
  __device__ static float flight(uint32_t u) { return nraps_uflight_f32(u); }
  __device__ static float nlog(float v) { return __logf(v); }
};

__device__ inline int lane_id() { return (int)__lane_id(); }
__device__ inline int rank_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Views of the packed tables (in LDS or in global memory).
template <typename Real>
struct Tabs {
  const Real *edges, *dx, *sigw, *invtr, *absr, *invs, *nsf, *chicdf, *cumrow;
  const uint32_t *cellinfo;
  const int32_t *fuel;
  const double *nsfd, *reglo, *kreg;
  __device__ void bind(const uint8_t *base, const TabLayout &L) {
    edges = (const Real *)(base + L.edges);
    cellinfo = (const uint32_t *)(base + L.cellinfo);
    fuel = (const int32_t *)(base + L.fuel);
    dx = (const Real *)(base + L.dx);
    sigw = (const Real *)(base + L.sigw);
    invtr = (const Real *)(base + L.invtr);
    absr = (const Real *)(base + L.absr);
    invs = (const Real *)(base + L.invs);
    nsf = (const Real *)(base + L.nsf);
    chicdf = (const Real *)(base + L.chicdf);
    cumrow = (const Real *)(base + L.cumrow);
    nsfd = (const double *)(base + L.nsfd);
    reglo = (const double *)(base + L.reglo);
    kreg = (const double *)(base + L.kreg);
  }
  __device__ int mat(int c) const { return (int)(cellinfo[c] & 0xffu); }
  __device__ int reg(int c) const { return (int)(cellinfo[c] >> 8); }
};

struct LaneRng {
  nraps_pcg32 r;
  int32_t draws;
  __device__ uint32_t next() {
    draws++;
    return nraps_pcg32_next(&r);
  }
};

struct LocalStats {
  uint32_t v[kNumStats];
};

template <typename Real>
struct Particle {
  Real x, mu;
  int32_t cell, g, gxs, rxs;
  int64_t p, events;
  double kacc;
  int32_t nsites, tr;
  bool traced;
};

// First index whose CDF value is not below u, clamped to the last group:
// slice.partition_point(|&x| x < u).min(n - 1), src/mc_code.rs:30-31, :124-126.
template <typename Real>
__device__ inline int cdf_index(const Real *cdf, int n, Real u) {
  int i = 0;
  while (i < n && cdf[i] < u) i++;
  return i < n - 1 ? i : n - 1;
}

template <typename Real>
__device__ inline void trace_event(const DevProblem &P, Particle<Real> &q, int type, int cell, int g,
                                   int aux, Real x, Real mu) {
  if (!q.traced) return;
  if (q.tr < P.trace_ev) {
    nraps_trace_event e;
    e.type = type;
    e.cell = cell;
    e.group = g;
    e.aux = aux;
    e.x = (double)x;
    e.mu = (double)mu;
    P.trace[q.p * P.trace_ev + q.tr] = e;
  }
  q.tr++;
}

// Birth, src/mc_code.rs:39-53 and :228-230 (uniform fuel cell, position,
// direction, chi group), or a banked fission site (direction, chi group).
template <typename Real>
__device__ inline void birth(const DevProblem &P, const Tabs<Real> &T, bool from_bank, int64_t p,
                             LaneRng &rng, Particle<Real> &q) {
  const int G = P.G;
  if (!from_bank) {
    uint32_t u = rng.next();
    int c = T.fuel[nraps_uindex(u, (uint32_t)P.nf)];
    Real frac = Num<Real>::u01(rng.next());
    q.mu = Real(2) * Num<Real>::u01(rng.next()) - Real(1);
    Real ue = Num<Real>::u01(rng.next());
    q.cell = c;
    Real xx = T.edges[c] + frac * T.dx[c];
    q.x = xx > T.edges[c + 1] ? T.edges[c + 1] : xx;
    q.g = cdf_index(&T.chicdf[T.mat(c) * G], G, ue);
  } else {
    q.mu = Real(2) * Num<Real>::u01(rng.next()) - Real(1);
    Real ue = Num<Real>::u01(rng.next());
    int c = P.src_cell[p];
    q.cell = c;
    q.x = ((const Real *)P.src_x)[p];
    q.g = cdf_index(&T.chicdf[T.mat(c) * G], G, ue);
  }
}

// Fission sites of one collision.  WAVE: append by ballot + one atomic per
// wave (event pipeline); otherwise one atomic per site (plain port).
template <typename Real, bool WAVE>
__device__ inline void emit_sites(const DevProblem &P, int n, Real x, int cell, Particle<Real> &q) {
  if (WAVE) {
    int rem = n;
    for (;;) {
      bool want = rem > 0;
      uint64_t m = __ballot(want);
      if (m == 0) break;
      int cnt = __popcll(m);
      int leader = __ffsll((unsigned long long)m) - 1;
      unsigned long long base = 0;
      if (lane_id() == leader) base = atomicAdd(&P.cs->site_count, (unsigned long long)cnt);
      base = __shfl(base, leader);
      if (want) {
        unsigned long long idx = base + (unsigned long long)rank_below(m);
        if (idx < (unsigned long long)P.site_cap) {
          ((Real *)P.site_x)[idx] = x;
          P.site_cell[idx] = cell;
          P.site_key[idx] = ((uint64_t)q.p << 32) | (uint32_t)q.nsites;
        }
        q.nsites++;
        rem--;
      }
    }
  } else {
    for (int i = 0; i < n; i++) {
      unsigned long long idx = atomicAdd(&P.cs->site_count, 1ull);
      if (idx < (unsigned long long)P.site_cap) {
        ((Real *)P.site_x)[idx] = x;
        P.site_cell[idx] = cell;
        P.site_key[idx] = ((uint64_t)q.p << 32) | (uint32_t)q.nsites;
      }
      q.nsites++;
    }
  }
}

// Collision, src/mc_code.rs:113-132 (draw order U_int, U_mu, U_scat; the
// scatter CDF uses the row of the current group and the normaliser of the
// possibly stale XS index), then fission sites when banking.  Returns true
// if the particle scattered.
template <typename Real, bool WAVE>
__device__ inline bool collide(const DevProblem &P, const Tabs<Real> &T, const CycleState &cs,
                               Real inv_k, int xs, LaneRng &rng, Particle<Real> &q, LocalStats &st) {
  const int G = P.G, M = P.M, m = T.mat(q.cell);
  Real u_int = Num<Real>::u01(rng.next());
  Real mu_new = Real(2) * Num<Real>::u01(rng.next()) - Real(1);
  Real u_scat = Num<Real>::u01(rng.next());
  int nsite = 0;
  if (cs.make_sites) {
    Real nsf = T.nsf[m + M * q.g];
    if (nsf > Real(0)) {
      Real y = nsf * T.invtr[xs] * inv_k + Num<Real>::u01(rng.next());
      nsite = y < Real(kMaxSitesPerCollision) ? (int)y : kMaxSitesPerCollision;
    }
  }
  if (WAVE) {
    if (cs.make_sites) emit_sites<Real, true>(P, nsite, q.x, q.cell, q);
  } else if (nsite) {
    emit_sites<Real, false>(P, nsite, q.x, q.cell, q);
  }
  st.v[kStCollisions]++;
  if (u_int < T.absr[xs]) {
    st.v[kStAbsorb]++;
    trace_event(P, q, NRAPS_EV_ABSORB, q.cell, q.g, q.g, q.x, q.mu);
    return false;
  }
  const Real *row = &T.cumrow[(m * G + q.g) * G];
  const Real is = T.invs[xs];
  int gn = G - 1;
  for (int j = 0; j < G; j++) {
    if (!(row[j] * is < u_scat)) {
      gn = j;
      break;
    }
  }
  trace_event(P, q, NRAPS_EV_SCATTER, q.cell, q.g, gn, q.x, q.mu);
  q.g = gn;
  q.mu = mu_new;
  return true;
}

template <typename Real>
__device__ inline void finish(const DevProblem &P, const CycleState &cs, Particle<Real> &q,
                              const LaneRng &rng) {
  P.kacc[q.p] = q.kacc;
  if (P.fission) P.counts[q.p] = q.nsites;
  if (q.traced) {
    P.trace_len[q.p] = q.tr;
    P.draws[q.p] = rng.draws;
    P.trace_kacc[q.p] = q.kacc;
  }
}

template <typename Real>
__device__ inline void init_particle(const DevProblem &P, const CycleState &cs, int64_t p,
                                     LaneRng &rng, Particle<Real> &q) {
  nraps_pcg32_seed(&rng.r, P.seed, nraps_stream_id(P.problem_id, (uint64_t)cs.cycle, (uint64_t)p));
  rng.draws = 0;
  q.p = p;
  q.events = 0;
  q.kacc = 0.0;
  q.nsites = 0;
  q.tr = 0;
  q.traced = P.trace != nullptr && cs.cycle == P.trace_cycle && p < P.trace_n;
}

__device__ inline void flush_stats(const DevProblem &P, LocalStats &st) {
  for (int i = 0; i < kNumStats; i++) {
    unsigned long long v = st.v[i];
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    if (lane_id() == 0 && v) atomicAdd(&P.stats[i], v);
    st.v[i] = 0;
  }
}

// ---------------------------------------------------------------------------
// Cycle start
// ---------------------------------------------------------------------------
__global__ void k_cycle_begin(const DevProblem *probs) {
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) P.cs->queue_head = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.tally_bins; i += gridDim.x * blockDim.x)
    P.tally_glob[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Plain FP64 port: one thread per history, surface tracking cell by cell,
// src/mc_code.rs:134-257.  Tables are read from global memory and every
// segment is scored with a global FP64 atomic.
// ---------------------------------------------------------------------------
__device__ inline double seg_len(double a, double b, double mu) {
  return a == b ? 0.0 : fabs((a - b) / mu);
}

__device__ void history_surface(const DevProblem &P, const Tabs<double> &T, const CycleState &cs,
                                int64_t p, LocalStats &st) {
  const int nc = P.nc, M = P.M;
  const bool xs_every = P.xs_every;
  double *tally = P.tally_glob;
  LaneRng rng;
  Particle<double> q;
  init_particle(P, cs, p, rng, q);
  birth(P, T, cs.from_bank, p, rng, q);
  trace_event(P, q, NRAPS_EV_BIRTH, q.cell, q.g, 0, q.x, q.mu);
  const double inv_k = cs.inv_k;
  for (;;) {
    int xs = T.mat(q.cell) + M * q.g;
    double delta_s = q.mu * -nraps_log(nraps_uflight_f64(rng.next())) * T.invtr[xs];
    bool same = true;
    while (same) {
      if (++q.events > P.max_events) {
        st.v[kStCapped]++;
        trace_event(P, q, NRAPS_EV_CAP, q.cell, q.g, 0, q.x, q.mu);
        finish(P, cs, q, rng);
        return;
      }
      const int cell = q.cell;
      const double x = q.x, mu = q.mu;
      double end_x = x + delta_s;
      double mesh_end = mu >= 0.0 ? T.edges[cell + 1] : T.edges[cell];
      const int mc = T.mat(cell);
      if ((mu < 0.0 && mesh_end > end_x && cell == 0) ||
          (mu >= 0.0 && end_x > mesh_end && cell == nc - 1)) {
        double tl = seg_len(x, mesh_end, mu);
        atomicAdd(&tally[q.g * nc + cell], tl);
        q.kacc += T.nsfd[mc + M * q.g] * tl;
        double bound = mu >= 0.0 ? P.boundr : P.boundl;
        st.v[kStCrossings]++;
        if (bound > 0.0) {
          q.mu = mu * (-bound);
          delta_s = (delta_s + (x - mesh_end)) * (-bound);
          q.x = mesh_end;
        } else {
          st.v[kStLeak]++;
          trace_event(P, q, NRAPS_EV_LEAK, cell, q.g, 0, mesh_end, mu);
          finish(P, cs, q, rng);
          return;
        }
      } else if (fabs(end_x - x) > fabs(mesh_end - x)) {
        double tl = seg_len(x, mesh_end, mu);
        atomicAdd(&tally[q.g * nc + cell], tl);
        q.kacc += T.nsfd[mc + M * q.g] * tl;
        int ni = mu >= 0.0 ? cell + 1 : cell - 1;
        delta_s = delta_s + (x - mesh_end);
        q.x = mesh_end;
        q.cell = ni;
        st.v[kStCrossings]++;
        if (mc != T.mat(ni)) same = false;
      } else {
        double tl = seg_len(x, end_x, mu);
        atomicAdd(&tally[q.g * nc + cell], tl);
        q.kacc += T.nsfd[mc + M * q.g] * tl;
        q.x = end_x;
        if (!collide<double, false>(P, T, cs, inv_k, xs, rng, q, st)) {
          finish(P, cs, q, rng);
          return;
        }
        if (xs_every) xs = T.mat(q.cell) + M * q.g;
        delta_s = q.mu * -nraps_log(nraps_uflight_f64(rng.next())) * T.invtr[xs];
      }
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_surface(const DevProblem *probs, int nprob) {
  const DevProblem &P = probs[blockIdx.y];
  const CycleState cs = *P.cs;
  LocalStats st;
  for (int i = 0; i < kNumStats; i++) st.v[i] = 0;
  if (cs.active) {
    Tabs<double> T;
    T.bind(P.tab, P.lay);
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < P.N;
         p += (int64_t)gridDim.x * blockDim.x)
      history_surface(P, T, cs, p, st);
  }
  flush_stats(P, st);
}

// ---------------------------------------------------------------------------
// Event pipeline with Woodcock delta tracking
// ---------------------------------------------------------------------------
// Cell holding the end of a move: moving right (mu >= 0) a point on an edge
// belongs to the cell on its left, moving left to the cell on its right
// (the strict crossing test of src/mc_code.rs:171).  Fixed-step binary
// search over the cell edges.
template <typename Real>
__device__ inline int locate(const Real *edges, int nc, int top, Real xe, bool right) {
  int c = 0;
  for (int step = top; step > 0; step >>= 1) {
    int t = c + step;
    if (t < nc) {
      Real e = edges[t];
      if (right ? (e < xe) : (e <= xe)) c = t;
    }
  }
  return c;
}

// Track-length tally of a straight piece from (xa, ca) to (xb, cb) with
// weight w = 1/|mu|: the partial end cells are scored directly, the cells
// in between by two updates of a difference array.  The generation tally
// is direct + width * prefix_sum(difference) (done in k_finalize).
template <typename Real, typename Acc>
__device__ inline void tally_piece(Acc *direct, Acc *diff, const Real *edges, int base, Real xa,
                                   int ca, Real xb, int cb, Real w) {
  if (xa == xb) return;
  if (ca == cb) {
    atomicAdd(&direct[base + ca], (Acc)(fabs(xb - xa) * w));
  } else if (ca < cb) {
    atomicAdd(&direct[base + ca], (Acc)((edges[ca + 1] - xa) * w));
    atomicAdd(&direct[base + cb], (Acc)((xb - edges[cb]) * w));
    if (cb > ca + 1) {
      atomicAdd(&diff[base + ca + 1], (Acc)w);
      atomicAdd(&diff[base + cb], (Acc)(-w));
    }
  } else {
    atomicAdd(&direct[base + ca], (Acc)((xa - edges[ca]) * w));
    atomicAdd(&direct[base + cb], (Acc)((edges[cb + 1] - xb) * w));
    if (ca > cb + 1) {
      atomicAdd(&diff[base + cb + 1], (Acc)w);
      atomicAdd(&diff[base + ca], (Acc)(-w));
    }
  }
}

// Integral of nu*Sigma_f from the slab's left edge to x (group g), in FP64:
// the k contribution of a piece is the difference at its ends times 1/|mu|.
template <typename Real>
__device__ inline double kint(const Tabs<Real> &T, int M, int R, int g, Real x, int c) {
  int r = T.reg(c);
  return T.kreg[g * (R + 1) + r] + T.nsfd[T.mat(c) + M * g] * ((double)x - T.reglo[r]);
}

// One Woodcock step: flight to a tentative site (folding reflections),
// track-length tally, XS refresh when the path left the region of the last
// refresh, acceptance test, and the collision if real.  Returns false when
// the particle leaks, is absorbed or hits the event cap.
template <typename Real, typename Acc>
__device__ inline bool woodcock_step(const DevProblem &P, const Tabs<Real> &T, const CycleState &cs,
                                     Real inv_k, Real maj, Real inv_maj, Real XL, Real XR,
                                     Real bl, Real br, int top, Acc *direct, Acc *diff,
                                     LaneRng &rng, Particle<Real> &q, LocalStats &st) {
  const int nc = P.nc, M = P.M, R = P.R;
  if (++q.events > P.max_events) {
    st.v[kStCapped]++;
    trace_event(P, q, NRAPS_EV_CAP, q.cell, q.g, 0, q.x, q.mu);
    return false;
  }
  st.v[kStSteps]++;
  Real mu = q.mu, x = q.x;
  int cell = q.cell;
  const int g = q.g, gbase = g * nc;
  Real disp = mu * -Num<Real>::nlog(Num<Real>::flight(rng.next())) * inv_maj;
  bool left_region = false;
  Real wgt = Real(1) / fabs(mu);
  for (;;) {
    Real end = x + disp;
    if (mu >= Real(0) && end > XR) {
      tally_piece<Real, Acc>(direct, diff, T.edges, gbase, x, cell, XR, nc - 1, wgt);
      if (x != XR) q.kacc += fabs(kint(T, M, R, g, XR, nc - 1) - kint(T, M, R, g, x, cell)) * (double)wgt;
      if (T.reg(nc - 1) != q.rxs) left_region = true;
      st.v[kStCrossings]++;
      if (br > Real(0)) {
        Real m2 = mu * (-br);
        disp = (disp + (x - XR)) * (-br);
        x = XR;
        mu = m2;
        cell = nc - 1;
        wgt = Real(1) / fabs(mu);
        continue;
      }
      st.v[kStLeak]++;
      trace_event(P, q, NRAPS_EV_LEAK, nc - 1, g, 0, XR, mu);
      return false;
    } else if (mu < Real(0) && XL > end) {
      tally_piece<Real, Acc>(direct, diff, T.edges, gbase, x, cell, XL, 0, wgt);
      if (x != XL) q.kacc += fabs(kint(T, M, R, g, XL, 0) - kint(T, M, R, g, x, cell)) * (double)wgt;
      if (T.reg(0) != q.rxs) left_region = true;
      st.v[kStCrossings]++;
      if (bl > Real(0)) {
        Real m2 = mu * (-bl);
        disp = (disp + (x - XL)) * (-bl);
        x = XL;
        mu = m2;
        cell = 0;
        wgt = Real(1) / fabs(mu);
        continue;
      }
      st.v[kStLeak]++;
      trace_event(P, q, NRAPS_EV_LEAK, 0, g, 0, XL, mu);
      return false;
    } else {
      int ce = end == x ? cell : locate(T.edges, nc, top, end, mu >= Real(0));
      tally_piece<Real, Acc>(direct, diff, T.edges, gbase, x, cell, end, ce, wgt);
      if (end != x) q.kacc += fabs(kint(T, M, R, g, end, ce) - kint(T, M, R, g, x, cell)) * (double)wgt;
      x = end;
      cell = ce;
      break;
    }
  }
  q.x = x;
  q.mu = mu;
  q.cell = cell;
  if (!P.xs_every && (left_region || T.reg(cell) != q.rxs)) {
    q.gxs = g;
    q.rxs = T.reg(cell);
  }
  const int xs = T.mat(cell) + M * q.gxs;
  Real u_acc = Num<Real>::u01(rng.next());
  if (u_acc * maj < T.sigw[xs]) {
    if (!collide<Real, true>(P, T, cs, inv_k, xs, rng, q, st)) return false;
    if (P.xs_every) {
      q.gxs = q.g;
      q.rxs = T.reg(cell);
    }
  } else {
    st.v[kStVirtual]++;
  }
  return true;
}

// Persistent event kernel.  A block visits the problems of the batch in
// cyclic order starting at its own index; on a problem with histories left
// it stages the tables in LDS, zeroes its LDS tally, and its waves run
// histories until the problem's queue is empty: each wave claims `claim`
// histories with one global atomic on the queue head and refills idle lanes
// from that range by ballot + mbcnt, so lanes stay busy without block
// barriers.  The LDS tally is then flushed with FP64 atomics.  Tallies that
// do not fit in LDS are scored straight into the global FP64 tally.
// Dynamic LDS: [tables][tally direct|diff].
template <typename Real, bool LDS_TALLY>
__global__ __launch_bounds__(kBlock) void k_woodcock(const DevProblem *probs, int nprob,
                                                     int tab_lds_bytes, int claim) {
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int s_has_work;
  typedef typename std::conditional<LDS_TALLY, Real, double>::type Acc;
  Real *ltally = (Real *)(smem + tab_lds_bytes);
  int loaded = -1;
  for (int k = 0; k < nprob; k++) {
    const int pi = (int)((blockIdx.x + (unsigned)k) % (unsigned)nprob);
    const DevProblem &P = probs[pi];
    if (threadIdx.x == 0) {
      const CycleState *c = P.cs;
      s_has_work = c->active && __hip_atomic_load(&P.cs->queue_head, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT) < P.N;
    }
    __syncthreads();
    if (!s_has_work) {
      __syncthreads();
      continue;
    }
    const CycleState cs = *P.cs;
    if (pi != loaded) {
      const uint32_t *src = (const uint32_t *)P.tab;
      uint32_t *dst = (uint32_t *)smem;
      for (int i = threadIdx.x; i < P.lay.bytes / 4; i += blockDim.x) dst[i] = src[i];
      loaded = pi;
    }
    const int TB = P.tally_bins;
    if (LDS_TALLY)
      for (int i = threadIdx.x; i < TB; i += blockDim.x) ltally[i] = Real(0);
    __syncthreads();

    Tabs<Real> T;
    T.bind(smem, P.lay);
    const int nc = P.nc;
    Acc *direct = LDS_TALLY ? (Acc *)ltally : (Acc *)P.tally_glob;
    Acc *diff = direct + P.G * nc;
    const Real maj = (Real)P.maj, inv_maj = (Real)P.inv_maj;
    const Real XL = (Real)P.XL, XR = (Real)P.XR, bl = (Real)P.boundl, br = (Real)P.boundr;
    const Real inv_k = (Real)cs.inv_k;
    const int64_t N = P.N;
    int top = 1;
    while (top * 2 < nc) top *= 2;
    LocalStats st;
    for (int i = 0; i < kNumStats; i++) st.v[i] = 0;

    Particle<Real> q;
    LaneRng rng;
    bool active = false, queue_done = false;
    int64_t r_cur = 0, r_end = 0;  // wave-uniform claimed range
    for (;;) {
      uint64_t need = __ballot(!active);
      while (need && !queue_done) {
        if (r_cur >= r_end) {
          unsigned long long b = 0;
          if (lane_id() == 0) b = atomicAdd((unsigned long long *)&P.cs->queue_head,
                                            (unsigned long long)claim);
          b = __shfl(b, 0);
          if ((int64_t)b >= N) {
            queue_done = true;
            break;
          }
          r_cur = (int64_t)b;
          r_end = r_cur + claim < N ? r_cur + claim : N;
        }
        const int cnt = __popcll(need);
        const int64_t avail = r_end - r_cur;
        const int take = avail < cnt ? (int)avail : cnt;
        const int rk = rank_below(need);
        if (!active && rk < take) {
          const int64_t p = r_cur + rk;
          init_particle(P, cs, p, rng, q);
          birth(P, T, cs.from_bank, p, rng, q);
          q.gxs = q.g;
          q.rxs = T.reg(q.cell);
          trace_event(P, q, NRAPS_EV_BIRTH, q.cell, q.g, 0, q.x, q.mu);
          active = true;
        }
        r_cur += take;
        need = __ballot(!active);
      }
      if (!__any(active)) break;
      if (active) {
        if (!woodcock_step<Real, Acc>(P, T, cs, inv_k, maj, inv_maj, XL, XR, bl, br, top, direct,
                                      diff, rng, q, st)) {
          finish(P, cs, q, rng);
          active = false;
        }
      }
    }
    flush_stats(P, st);
    __syncthreads();
    if (LDS_TALLY) {
      for (int i = threadIdx.x; i < TB; i += blockDim.x) {
        Real v = ltally[i];
        if (v != Real(0)) atomicAdd(&P.tally_glob[i], (double)v);
      }
    }
    __syncthreads();
  }
}

template __global__ void k_woodcock<float, true>(const DevProblem *, int, int, int);
template __global__ void k_woodcock<float, false>(const DevProblem *, int, int, int);
template __global__ void k_woodcock<double, true>(const DevProblem *, int, int, int);
template __global__ void k_woodcock<double, false>(const DevProblem *, int, int, int);


// ---------------------------------------------------------------------------
// Canonical k sum: groups of 256 consecutive values, pairwise tree with
// strides 128..1, zero padded; `levels` passes (the result of the last one
// is the total).  Level l reads kacc (l == 0) or the previous level's output.
// ---------------------------------------------------------------------------
__device__ inline int64_t tree_len(int64_t n, int level) {
  for (int l = 0; l < level; l++) n = (n + kTreeWidth - 1) / kTreeWidth;
  return n;
}

__global__ __launch_bounds__(kTreeWidth) void k_ktree(const DevProblem *probs, int level) {
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active) return;
  __shared__ double s[kTreeWidth];
  const int64_t n_in = tree_len(P.N, level);
  const int64_t n_out = (n_in + kTreeWidth - 1) / kTreeWidth;
  const double *in = level == 0 ? P.kacc : ((level - 1) % 2 == 0 ? P.tree_a : P.tree_b);
  double *out = level % 2 == 0 ? P.tree_a : P.tree_b;
  for (int64_t gi = blockIdx.x; gi < n_out; gi += gridDim.x) {
    int64_t j = gi * kTreeWidth + threadIdx.x;
    s[threadIdx.x] = j < n_in ? in[j] : 0.0;
    __syncthreads();
    for (int stride = kTreeWidth / 2; stride >= 1; stride >>= 1) {
      if ((int)threadIdx.x < stride) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + stride];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[gi] = s[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Fission bank: exclusive scan of per-history site counts, placement of the
// unordered sites in (history, sequence) order, systematic resampling.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const DevProblem *probs, int64_t *tile_sums,
                                                       const int64_t *tile_base) {
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active || !P.cs->make_sites) return;
  const int64_t ntiles = (P.N + kScanTile - 1) / kScanTile;
  __shared__ int64_t red[kBlock];
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int64_t s = 0;
    for (int i = threadIdx.x; i < kScanTile; i += blockDim.x) {
      int64_t j = t * kScanTile + i;
      if (j < P.N) s += P.counts[j];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int stride = kBlock / 2; stride >= 1; stride >>= 1) {
      if ((int)threadIdx.x < stride) red[threadIdx.x] += red[threadIdx.x + stride];
      __syncthreads();
    }
    if (threadIdx.x == 0) tile_sums[tile_base[blockIdx.y] + t] = red[0];
    __syncthreads();
  }
}

// One block per problem: exclusive scan over the tile sums (each thread a
// contiguous run), total -> CycleState.n_sites, overflow / empty flags.
__global__ __launch_bounds__(kBlock) void k_scan_tilesums(const DevProblem *probs, int64_t *tile_sums,
                                                          const int64_t *tile_base) {
  const DevProblem &P = probs[blockIdx.x];
  if (!P.cs->active || !P.cs->make_sites) return;
  const int64_t ntiles = (P.N + kScanTile - 1) / kScanTile;
  int64_t *ts = tile_sums + tile_base[blockIdx.x];
  __shared__ int64_t part[kBlock];
  const int64_t per = (ntiles + kBlock - 1) / kBlock;
  const int64_t lo = threadIdx.x * per, hi = lo + per < ntiles ? lo + per : ntiles;
  int64_t s = 0;
  for (int64_t t = lo; t < hi; t++) s += ts[t];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t run = 0;
    for (int i = 0; i < kBlock; i++) {
      int64_t v = part[i];
      part[i] = run;
      run += v;
    }
    CycleState *cs = P.cs;
    cs->n_sites = run;
    if (run == 0) cs->error |= 2;
    if ((int64_t)cs->site_count > P.site_cap || (int64_t)cs->site_count != run) cs->error |= 1;
  }
  __syncthreads();
  int64_t run = part[threadIdx.x];
  for (int64_t t = lo; t < hi; t++) {
    int64_t v = ts[t];
    ts[t] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const DevProblem *probs, const int64_t *tile_sums,
                                                       const int64_t *tile_base) {
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active || !P.cs->make_sites) return;
  const int64_t ntiles = (P.N + kScanTile - 1) / kScanTile;
  constexpr int per = kScanTile / kBlock;
  __shared__ int64_t part[kBlock];
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t b0 = t * kScanTile + (int64_t)threadIdx.x * per;
    int64_t s = 0;
    for (int i = 0; i < per; i++) {
      int64_t j = b0 + i;
      if (j < P.N) s += P.counts[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t run = tile_sums[tile_base[blockIdx.y] + t];
      for (int i = 0; i < kBlock; i++) {
        int64_t v = part[i];
        part[i] = run;
        run += v;
      }
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for (int i = 0; i < per; i++) {
      int64_t j = b0 + i;
      if (j < P.N) {
        P.offsets[j] = run;
        run += P.counts[j];
      }
    }
    __syncthreads();
  }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_place(const DevProblem *probs) {
  const DevProblem &P = probs[blockIdx.y];
  const CycleState *cs = P.cs;
  if (!cs->active || !cs->make_sites || (cs->error & 1)) return;
  const int64_t n = cs->n_sites;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key = P.site_key[i];
    int64_t dst = P.offsets[key >> 32] + (int64_t)(key & 0xffffffffu);
    ((Real *)P.ord_x)[dst] = ((const Real *)P.site_x)[i];
    P.ord_cell[dst] = P.site_cell[i];
  }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_resample(const DevProblem *probs) {
  const DevProblem &P = probs[blockIdx.y];
  const CycleState *cs = P.cs;
  if (!cs->active || !cs->make_sites || (cs->error & 3)) return;
  const int64_t M = cs->n_sites;
  nraps_pcg32 r;
  nraps_pcg32_seed(&r, P.seed, nraps_stream_id(P.problem_id, (uint64_t)cs->cycle, NRAPS_SERVICE_PARTICLE));
  const double u0 = nraps_u01_f64(nraps_pcg32_next(&r));
  const double ratio = (double)M / (double)P.N;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < P.N;
       j += (int64_t)gridDim.x * blockDim.x) {
    int64_t idx = (int64_t)(((double)j + u0) * ratio);
    if (idx > M - 1) idx = M - 1;
    ((Real *)P.src_x)[j] = ((const Real *)P.ord_x)[idx];
    P.src_cell[j] = P.ord_cell[idx];
  }
}
template __global__ void k_place<float>(const DevProblem *);
template __global__ void k_place<double>(const DevProblem *);
template __global__ void k_resample<float>(const DevProblem *);
template __global__ void k_resample<double>(const DevProblem *);

// ---------------------------------------------------------------------------
// Generation close: k, the reference normalisation/accumulation
// (src/mc_code.rs:340-363), variance sums, next-generation state.
// One block per problem.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_finalize(const DevProblem *probs, const DevFinal *fins,
                                                     int levels) {
  const DevProblem &P = probs[blockIdx.x];
  const DevFinal &F = fins[blockIdx.x];
  CycleState *csp = P.cs;
  if (!csp->active) return;
  const int nc = P.nc, G = P.G, M = P.M;
  const int cyc = csp->cycle;
  const double kprev = csp->k_prev;
  const double N = (double)P.N;
  const double *tot = (levels - 1) % 2 == 0 ? P.tree_a : P.tree_b;
  const double knew = tot[0] / N;
  double *T = F.tally_cycle;  // direct part becomes the cell tally
  __shared__ double part[kBlock];
  // cell tally = direct + width * inclusive_prefix(diff), per group
  if (F.has_diff) {
    const int per = (nc + kBlock - 1) / kBlock;
    const int lo = threadIdx.x * per, hi = lo + per < nc ? lo + per : nc;
    for (int g = 0; g < G; g++) {
      const double *d = T + G * nc + g * nc;
      double s = 0.0;
      for (int c = lo; c < hi; c++) s += d[c];
      part[threadIdx.x] = s;
      __syncthreads();
      if (threadIdx.x == 0) {
        double run = 0.0;
        for (int i = 0; i < kBlock; i++) {
          double v = part[i];
          part[i] = run;
          run += v;
        }
      }
      __syncthreads();
      double run = part[threadIdx.x];
      for (int c = lo; c < hi; c++) {
        run += d[c];
        T[g * nc + c] += F.width[c] * run;
      }
      __syncthreads();
    }
  }
  const bool active_gen = cyc >= F.skip;
  const double fund = 1.0 / (double)(F.generations - (F.skip - 1));
  const double conversion =
      (3565e6 * kprev * 36.2) / (200e6 * 1.602176634e-19 * F.nut_conv * F.length);
  double kref = 0.0;
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    const double dxc = F.dx[c];
    const int mc = F.mat[c];
    double fs_c = 0.0;
    for (int g = 0; g < G; g++) {
      const int xi = mc + M * g;
      const double t = T[g * nc + c];
      double flux = t / (kprev * N * dxc);
      double fsrc = F.nut[xi] * F.sigf[xi] * flux;
      kref += kprev * dxc * fsrc;
      if (active_gen) {
        F.flux_acc[g * nc + c] += flux * conversion * fund;
        fs_c += fsrc * fund;
        double fv = t / (N * dxc);
        F.fsum[g * nc + c] += fv;
        F.fsq[g * nc + c] += fv * fv;
      }
      if (cyc == F.generations - 1 && F.tally_last) F.tally_last[g * nc + c] = t;
    }
    if (active_gen) F.fs_acc[c] += fs_c;
  }
  part[threadIdx.x] = kref;
  __syncthreads();
  for (int stride = kBlock / 2; stride >= 1; stride >>= 1) {
    if ((int)threadIdx.x < stride) part[threadIdx.x] += part[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    P.k[cyc] = knew;
    F.k_ref[cyc] = part[0];
    F.sites_hist[cyc] = csp->make_sites ? csp->n_sites : 0;
    const int next = cyc + 1;
    csp->k_prev = knew;
    csp->inv_k = 1.0 / knew;
    csp->cycle = next;
    csp->from_bank = P.fission && next >= 2;
    csp->make_sites = P.fission && next >= 1;
    csp->site_count = 0;
    if (P.fission && !(knew > 0.0)) csp->error |= 2;
    if (next >= F.generations) {
      csp->active = 0;
      csp->error |= 4;
    }
  }
}

// PCG32 streams for the generator known-answer test.
__global__ void k_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n, int draws, uint32_t *out) {
  int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  nraps_pcg32 r;
  nraps_pcg32_seed(&r, seed, stream0 + (uint64_t)s);
  for (int i = 0; i < draws; i++) out[s * draws + i] = nraps_pcg32_next(&r);
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
static inline unsigned clamp_grid(int64_t n, int64_t cap) {
  if (n < 1) n = 1;
  return (unsigned)(n < cap ? n : cap);
}

void launch_cycle_begin(const DevProblem *d, int nprob, int max_bins, hipStream_t s) {
  dim3 grid(clamp_grid((max_bins + kBlock - 1) / kBlock, 64), (unsigned)nprob);
  hipLaunchKernelGGL(k_cycle_begin, grid, dim3(kBlock), 0, s, d);
}

void launch_surface(const DevProblem *d, int nprob, int64_t max_n, hipStream_t s) {
  dim3 grid(clamp_grid((max_n + kBlock - 1) / kBlock, 1 << 20), (unsigned)nprob);
  hipLaunchKernelGGL(k_surface, grid, dim3(kBlock), 0, s, d, nprob);
}

template <typename Real, bool L>
static void launch_wc(const DevProblem *d, int nprob, int grid, size_t lds, int tab, int claim,
                      hipStream_t s) {
  hipLaunchKernelGGL((k_woodcock<Real, L>), dim3((unsigned)grid), dim3(kBlock), lds, s, d, nprob, tab,
                     claim);
}

void launch_woodcock(bool fp32, bool lds_tally, const DevProblem *d, int nprob, int grid,
                     size_t lds_bytes, int tab_lds_bytes, int claim, hipStream_t s) {
  if (fp32 && lds_tally) launch_wc<float, true>(d, nprob, grid, lds_bytes, tab_lds_bytes, claim, s);
  if (fp32 && !lds_tally) launch_wc<float, false>(d, nprob, grid, lds_bytes, tab_lds_bytes, claim, s);
  if (!fp32 && lds_tally) launch_wc<double, true>(d, nprob, grid, lds_bytes, tab_lds_bytes, claim, s);
  if (!fp32 && !lds_tally) launch_wc<double, false>(d, nprob, grid, lds_bytes, tab_lds_bytes, claim, s);
}

template <typename Real, bool L>
static const void *wc_fn() {
  return reinterpret_cast<const void *>(&k_woodcock<Real, L>);
}
static const void *wc_select(bool fp32, bool lds_tally) {
  if (fp32) return lds_tally ? wc_fn<float, true>() : wc_fn<float, false>();
  return lds_tally ? wc_fn<double, true>() : wc_fn<double, false>();
}

hipError_t woodcock_set_lds_limit(bool fp32, bool lds_tally, size_t lds_bytes) {
  return hipFuncSetAttribute(wc_select(fp32, lds_tally), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)lds_bytes);
}

int woodcock_blocks_per_cu(bool fp32, bool lds_tally, size_t lds_bytes) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, wc_select(fp32, lds_tally), kBlock, lds_bytes) !=
      hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

void launch_ktree(const DevProblem *d, int nprob, int level, int64_t max_groups, hipStream_t s) {
  dim3 grid(clamp_grid(max_groups, 65535), (unsigned)nprob);
  hipLaunchKernelGGL(k_ktree, grid, dim3(kTreeWidth), 0, s, d, level);
}

void launch_bank(bool fp32, const DevProblem *d, int nprob, int64_t max_n, int64_t max_tiles,
                 int64_t *tile_sums, const int64_t *tile_base, hipStream_t s) {
  dim3 gt(clamp_grid(max_tiles, 65535), (unsigned)nprob);
  hipLaunchKernelGGL(k_scan_tiles, gt, dim3(kBlock), 0, s, d, tile_sums, tile_base);
  hipLaunchKernelGGL(k_scan_tilesums, dim3((unsigned)nprob), dim3(kBlock), 0, s, d, tile_sums, tile_base);
  hipLaunchKernelGGL(k_scan_apply, gt, dim3(kBlock), 0, s, d, (const int64_t *)tile_sums, tile_base);
  dim3 gp(clamp_grid((4 * max_n + 65536 + kBlock - 1) / kBlock, 8192), (unsigned)nprob);
  dim3 gr(clamp_grid((max_n + kBlock - 1) / kBlock, 8192), (unsigned)nprob);
  if (fp32) {
    hipLaunchKernelGGL(k_place<float>, gp, dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_resample<float>, gr, dim3(kBlock), 0, s, d);
  } else {
    hipLaunchKernelGGL(k_place<double>, gp, dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_resample<double>, gr, dim3(kBlock), 0, s, d);
  }
}

void launch_finalize(const DevProblem *d, const DevFinal *f, int nprob, int levels, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3((unsigned)nprob), dim3(kBlock), 0, s, d, f, levels);
}

void launch_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n, int draws, uint32_t *out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_pcg32_fill, dim3(clamp_grid((n + kBlock - 1) / kBlock, 1 << 30)), dim3(kBlock), 0,
                     s, seed, stream0, n, draws, out);
}

}  // namespace nraps
