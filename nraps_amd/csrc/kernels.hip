// kernels.hip — HIP kernels of the MI355X Monte Carlo k-eigenvalue solver.
//
// Generation pipeline (every kernel takes the batch's DevProblem table and
// works on any number of problems; blockIdx.y or the block's problem walk
// selects the problem):
//
//   k_cycle_begin      reset the history queues, zero the generation tallies
//   k_surface          plain FP64 port: one thread per history, cell-by-cell
//                      surface tracking, global FP64 atomics
//                      (src/mc_code.rs:134-257)
//   k_event<Real>      event pipeline: persistent blocks stage a problem's
//                      tables in LDS; each wave claims ranges of histories with
//                      one atomic and refills idle lanes by ballot + mbcnt;
//                      one event per lane and step (region surface tracking:
//                      a collision, a region edge or a face; Woodcock: a
//                      tentative flight), difference-array
//                      track-length tallies in LDS (flushed once per block with
//                      FP64 atomics); fission sites appended by wave ballot
//                      (one global atomic per wave)
//   k_ktree            canonical 256-wide pairwise sum of per-history k
//   k_scan_*, k_place, k_resample
//                      fission bank: sites placed in (history, sequence) order
//                      by a prefix sum of per-history counts, then systematic
//                      resampling to N sources in FP64
//   k_finalize_rows, k_finalize_close
//                      k, reference normalisation and accumulation
//                      (src/mc_code.rs:340-363), variance sums, next-cycle state
//
// Compiled with -ffp-contract=off: the FP64 histories must reproduce the CPU
// replay bit for bit.
#include <algorithm>
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
  __device__ static double rcp(double v) { return 1.0 / v; }  // exact: CPU replay
};
template <>
struct Num<float> {
  __device__ static float u01(uint32_t u) { return nraps_u01_f32(u); }
  __device__ static float flight(uint32_t u) { return nraps_uflight_f32(u); }
  // v_log_f32 (log2) times ln 2: flight arguments are >= 2^-32, never
  // denormal, so no range reduction is needed (1.6% of the event kernel
  // against __logf)
  __device__ static float nlog(float v) { return __builtin_amdgcn_logf(v) * 0.69314718f; }
  __device__ static float rcp(float v) { return __builtin_amdgcn_rcpf(v); }
};

__device__ inline int lane_id() { return (int)__lane_id(); }
// Lane 0's value, read into scalar registers (the caller knows it is the
// same in every lane): wave-uniform counters then take no vector registers.
__device__ inline uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}
// Lane 0's copy of a wave-uniform value.  Fields read from the problem
// descriptor arrive through vector loads (the kernel writes global memory,
// so the compiler cannot use scalar loads for them); passing them through
// readfirstlane moves them to scalar registers for the rest of the visit.
__device__ inline int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline int64_t uni(int64_t v) { return (int64_t)uniform_u64((uint64_t)v); }
__device__ inline uint64_t uni(uint64_t v) { return uniform_u64(v); }
__device__ inline float uni(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ inline double uni(double v) {
  return __longlong_as_double((long long)uniform_u64((uint64_t)__double_as_longlong(v)));
}
template <class T>
__device__ inline T *uni(T *ptr) {  // global-memory pointers only (not LDS)
  return reinterpret_cast<T *>(uniform_u64(reinterpret_cast<uint64_t>(ptr)));
}
// uni() where U holds: a kernel whose vector registers are scarce moves the
// descriptor to scalar registers; one whose scalar registers are the tighter
// budget keeps it in vector registers.
template <bool U, class T>
__device__ inline T uni_if(T v) {
  if constexpr (U) return uni(v);
  else return v;
}

__device__ inline int rank_below(uint64_t mask) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Exclusive prefix sum over a 256-thread block (every thread calls it):
// inclusive scans inside each wave by shuffles, then the wave totals.
// *total receives the block's sum.  The order of the additions differs from
// a serial scan, which matters for floating point only.
template <typename T>
__device__ inline T block_exclusive_scan(T v, T *total) {
  __shared__ T wsum[kBlock / 64];
  const int lane = lane_id(), w = (int)(threadIdx.x >> 6);
  T inc = v;
  for (int off = 1; off < 64; off <<= 1) {
    const T u = __shfl_up(inc, off);
    if (lane >= off) inc = inc + u;
  }
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  T base = T(0), tot = T(0);
  for (int i = 0; i < kBlock / 64; i++) {
    if (i < w) base = base + wsum[i];
    tot = tot + wsum[i];
  }
  __syncthreads();  // wsum may be reused by the next call
  *total = tot;
  return base + (inc - v);
}

// Views of the packed tables: the hot part (in LDS or global memory) and the
// birth-only part (global memory).
template <typename Real>
struct Tabs {
  const Real *edges, *dx, *sigw, *invtr, *absr, *invs, *nsf, *chicdf, *cumrow;
  const uint8_t *mat8;
  const uint16_t *reg16;
  const int32_t *fuel;
  const double *nsfd, *kreg;
  const uint16_t *bucket;
  const double *regedge;  // region edges in FP64: the k integral's origins too
  const uint16_t *regcell;
  // The offsets are wave-uniform: with U, uni() keeps them in scalar
  // registers (the base keeps its address space, so LDS tables stay LDS
  // accesses).
  template <bool U = true>
  __device__ void bind(const uint8_t *base, const uint8_t *global_base, const TabLayout &L) {
    edges = (const Real *)(base + uni_if<U>(L.edges));
    mat8 = base + uni_if<U>(L.mat8);
    reg16 = (const uint16_t *)(base + uni_if<U>(L.reg16));
    // birth tables: staged in LDS with the hot part when placed before it
    fuel = (const int32_t *)((L.fuel < L.hot ? base : global_base) + uni_if<U>(L.fuel));
    dx = (const Real *)((L.dx < L.hot ? base : global_base) + uni_if<U>(L.dx));
    sigw = (const Real *)(base + uni_if<U>(L.sigw));
    invtr = (const Real *)(base + uni_if<U>(L.invtr));
    absr = (const Real *)(base + uni_if<U>(L.absr));
    invs = (const Real *)(base + uni_if<U>(L.invs));
    nsf = (const Real *)(base + uni_if<U>(L.nsf));
    chicdf = (const Real *)(base + uni_if<U>(L.chicdf));
    cumrow = (const Real *)(base + uni_if<U>(L.cumrow));
    nsfd = (const double *)(base + uni_if<U>(L.nsfd));
    kreg = (const double *)(base + uni_if<U>(L.kreg));
    bucket = (const uint16_t *)(base + uni_if<U>(L.bucket));
    regedge = (const double *)(base + uni_if<U>(L.regedge));
    regcell = (const uint16_t *)(base + uni_if<U>(L.regcell));
  }
  __device__ int mat(int c) const { return (int)mat8[c]; }
  __device__ int reg(int c) const { return (int)reg16[c]; }
};

// Uniform values of one problem, read once per problem visit so the hot
// loop keeps them in scalar registers.
template <typename Real>
struct Ctx {
  Tabs<Real> T;
  int nc, G, M, R, nf, nb;
  int xs_every, fission, from_bank, make_sites, cycle;
  int32_t max_events;  // per-history cap, clamped to the range of the event counter
  int64_t N;
  uint64_t seed, problem_id;
  Real maj, inv_maj, XL, XR, bl, br, inv_k, inv_h;
  double *kacc;
  int32_t *counts;
  const Real *src_x;
  const int32_t *src_cell;
  Real *site_x;
  int32_t *site_cell;
  uint64_t *site_key;
  unsigned long long *site_count;
  int64_t site_cap;
  nraps_trace_event *trace;
  int32_t *trace_len;
  int64_t *draws;
  double *trace_kacc;
  int64_t trace_n;
  int trace_ev, trace_cycle;

  // U: descriptor fields through uni(); UT: the table offsets.
  template <bool U = true, bool UT = U>
  __device__ void load(const DevProblem &P, const CycleState &cs, const uint8_t *tab) {
    T.template bind<UT>(tab, uni_if<UT>(P.tab), P.lay);
    nc = uni_if<U>(P.nc);
    G = uni_if<U>(P.G);
    M = uni_if<U>(P.M);
    R = uni_if<U>(P.R);
    nf = uni_if<U>(P.nf);
    nb = uni_if<U>(P.nb);
    xs_every = uni_if<U>(P.xs_every);
    fission = uni_if<U>(P.fission);
    from_bank = uni_if<U>(cs.from_bank);
    make_sites = uni_if<U>(cs.make_sites);
    cycle = uni_if<U>(cs.cycle);
    max_events = uni_if<U>((int32_t)(P.max_events < INT32_MAX - 1 ? P.max_events : INT32_MAX - 1));
    N = uni_if<U>(P.N);
    seed = uni_if<U>(P.seed);
    problem_id = uni_if<U>(P.problem_id);
    maj = uni_if<U>((Real)P.maj);
    inv_maj = uni_if<U>((Real)P.inv_maj);
    XL = uni_if<U>((Real)P.XL);
    XR = uni_if<U>((Real)P.XR);
    bl = uni_if<U>((Real)P.boundl);
    br = uni_if<U>((Real)P.boundr);
    inv_k = uni_if<U>((Real)cs.inv_k);
    inv_h = uni_if<U>((Real)P.inv_h);
    kacc = uni_if<U>(P.kacc);
    counts = uni_if<U>(P.counts);
    src_x = uni_if<U>((const Real *)P.src_x);
    src_cell = uni_if<U>(P.src_cell);
    site_x = uni_if<U>((Real *)P.site_x);
    site_cell = uni_if<U>(P.site_cell);
    site_key = uni_if<U>(P.site_key);
    site_count = &uni_if<U>(P.cs)->site_count;
    site_cap = uni_if<U>(P.site_cap);
    trace = uni_if<U>(P.trace);
    trace_len = uni_if<U>(P.trace_len);
    draws = uni_if<U>(P.draws);
    trace_kacc = uni_if<U>(P.trace_kacc);
    trace_n = uni_if<U>(P.trace_n);
    trace_ev = uni_if<U>(P.trace_ev);
    trace_cycle = uni_if<U>(P.trace_cycle);
  }
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

// Compile-time features of a history kernel: per-history parity traces and
// fission-site banking.  Production launches without them carry neither the
// code nor the registers.
template <bool TRACE_, bool FISSION_>
struct Feat {
  static constexpr bool TRACE = TRACE_;
  static constexpr bool FISSION = FISSION_;
};
using FeatAll = Feat<true, true>;

// Outcome bits of one event-kernel step.
enum : uint32_t {
  kEvStep = 1u,     // a flight was sampled
  kEvReal = 2u,     // real collision
  kEvAbsorb = 4u,   // absorbed
  kEvLeak = 8u,     // leaked through a vacuum face
  kEvCap = 16u,     // stopped by the event cap
  kEvDead = kEvAbsorb | kEvLeak | kEvCap,
};

template <typename Real>
struct Particle {
  Real x, mu;
  int32_t cell, g, gxs, rxs;
  int32_t p, events;
  int32_t nreal;  // real collisions of this history
  double kacc;
  // The event kernel scores a straight segment once, when it ends (real
  // collision, boundary, cap): its start point and the k integral there
  Real sx;
  int32_t scell;
  double ks;
  // XS of the particle while it stays in region rxs: index (mat + M*gxs) and
  // the Woodcock acceptance XS there
  int32_t xs;
  Real sw;
  // region tracking: displacement left of the current flight
  Real ds;
  int32_t nsites, tr;
  bool traced;
};

// First index whose CDF value is not below u, clamped to the last group:
// slice.partition_point(|&x| x < u).min(n - 1), src/mc_code.rs:30-31, :124-126.
// A cumulative table is non-decreasing, so that index is the number of the
// first n - 1 entries below u: counted with a fixed trip count (n is uniform
// across the wave), no divergent early exit.
template <typename Real>
__device__ inline int cdf_index(const Real *cdf, int n, Real u) {
  int i = 0;
  for (int j = 0; j < n - 1; j++) i += cdf[j] < u ? 1 : 0;
  return i;
}

template <class F, typename Real>
__device__ inline void trace_event(const Ctx<Real> &C, Particle<Real> &q, int type, int cell, int g,
                                   int aux, Real x, Real mu) {
  if (!F::TRACE || !q.traced) return;
  if (q.tr < C.trace_ev) {
    nraps_trace_event e;
    e.type = type;
    e.cell = cell;
    e.group = g;
    e.aux = aux;
    e.x = (double)x;
    e.mu = (double)mu;
    C.trace[(int64_t)q.p * C.trace_ev + q.tr] = e;
  }
  q.tr++;
}

template <class F, typename Real>
__device__ inline void init_particle(const Ctx<Real> &C, int64_t p, LaneRng &rng, Particle<Real> &q) {
  nraps_history_rng(&rng.r, C.seed, nraps_stream_id(C.problem_id, (uint64_t)C.cycle, (uint64_t)p));
  rng.draws = 0;
  q.p = (int32_t)p;
  q.events = 0;
  q.nreal = 0;
  q.kacc = 0.0;
  q.nsites = 0;
  q.tr = 0;
  q.traced = F::TRACE && C.trace != nullptr && C.cycle == C.trace_cycle && p < C.trace_n;
}

// Birth, src/mc_code.rs:39-53 and :228-230 (uniform fuel cell, position,
// direction, chi group), or a banked fission site (direction, chi group).
template <typename Real>
__device__ inline void birth(const Ctx<Real> &C, int64_t p, LaneRng &rng, Particle<Real> &q) {
  const Tabs<Real> &T = C.T;
  const int G = C.G;
  if (!C.from_bank) {
    uint32_t u = rng.next();
    int c = T.fuel[nraps_uindex(u, (uint32_t)C.nf)];
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
    int c = C.src_cell[p];
    q.cell = c;
    q.x = C.src_x[p];
    q.g = cdf_index(&T.chicdf[T.mat(c) * G], G, ue);
  }
}

template <typename Real>
__device__ inline void put_site(const Ctx<Real> &C, unsigned long long idx, Particle<Real> &q) {
  if (idx < (unsigned long long)C.site_cap) {
    C.site_x[idx] = q.x;
    C.site_cell[idx] = q.cell;
    C.site_key[idx] = ((uint64_t)(uint32_t)q.p << 32) | (uint32_t)q.nsites;
  }
  q.nsites++;
}

// Fission sites of the plain port: one atomic per site.
template <typename Real>
__device__ inline void emit_sites_plain(const Ctx<Real> &C, int n, Particle<Real> &q) {
  for (int i = 0; i < n; i++) put_site(C, atomicAdd(C.site_count, 1ull), q);
}

// Site-slot cursor of one wave in the event pipeline.  Slots of the
// unordered site buffer are reserved kSiteChunk at a time with one atomic;
// lanes fill them by ballot + mbcnt.  The unused tail of a chunk is marked
// with an invalid key, which the placement kernel skips.  Must be used in
// wave-uniform control flow (every lane holds the same cursor).
constexpr int kSiteChunk = 128;
constexpr uint64_t kNoSite = ~0ull;

struct SiteCursor {
  unsigned long long next = 0, end = 0;
  template <typename Real>
  __device__ void close(const Ctx<Real> &C) {
    for (unsigned long long i = next + (unsigned long long)lane_id(); i < end; i += 64)
      if (i < (unsigned long long)C.site_cap) C.site_key[i] = kNoSite;
    next = end;
  }
  // Bank `n` sites of this lane's particle at its position (n may be 0).
  template <typename Real>
  __device__ void emit(const Ctx<Real> &C, int n, Particle<Real> &q) {
    int rem = n;
    for (;;) {
      const bool want = rem > 0;
      const uint64_t m = __ballot(want);
      if (m == 0) break;
      const unsigned long long cnt = (unsigned long long)__popcll(m);
      if (next + cnt > end) {
        close(C);
        unsigned long long b = 0;
        if (lane_id() == 0) b = atomicAdd(C.site_count, (unsigned long long)kSiteChunk);
        next = uniform_u64(b);
        end = next + kSiteChunk;
      }
      if (want) {
        put_site(C, next + (unsigned long long)rank_below(m), q);
        rem--;
      }
      next = uniform_u64(next + cnt);
    }
  }
};

// Collision at (q.x, q.cell) in material m, src/mc_code.rs:113-132: draws U_int, U_mu,
// U_scat; the scatter CDF uses the row of the current group and the
// normaliser of the (possibly stale) XS index; the number of fission sites
// is drawn after the reference's three draws and returned in `nsite` (the
// caller banks them at the collision site).  Returns true if the particle
// scattered.
template <class F, typename Real>
__device__ inline bool collide(const Ctx<Real> &C, int xs, int m, LaneRng &rng, Particle<Real> &q,
                               int &nsite) {
  const Tabs<Real> &T = C.T;
  const int G = C.G;
  q.nreal++;
  Real u_int = Num<Real>::u01(rng.next());
  Real mu_new = Real(2) * Num<Real>::u01(rng.next()) - Real(1);
  Real u_scat = Num<Real>::u01(rng.next());
  nsite = 0;
  if (F::FISSION && C.make_sites) {
    Real nsf = T.nsf[m + C.M * q.g];
    if (nsf > Real(0)) {
      Real y = nsf * T.invtr[xs] * C.inv_k + Num<Real>::u01(rng.next());
      nsite = y < Real(kMaxSitesPerCollision) ? (int)y : kMaxSitesPerCollision;
    }
  }
  if (u_int < T.absr[xs]) {
    trace_event<F>(C, q, NRAPS_EV_ABSORB, q.cell, q.g, q.g, q.x, q.mu);
    return false;
  }
  // first j with row[j] * (1/sigma_s) >= U, clamped to G - 1 (the row is
  // cumulative, so this is a count over the first G - 1 entries)
  const Real *row = &T.cumrow[(m * G + q.g) * G];
  const Real is = T.invs[xs];
  int gn = 0;
  for (int j = 0; j < G - 1; j++) gn += row[j] * is < u_scat ? 1 : 0;
  trace_event<F>(C, q, NRAPS_EV_SCATTER, q.cell, q.g, gn, q.x, q.mu);
  q.g = gn;
  q.mu = mu_new;
  return true;
}

template <class F, typename Real>
__device__ inline void finish(const Ctx<Real> &C, Particle<Real> &q, const LaneRng &rng) {
  C.kacc[q.p] = q.kacc;
  if (F::FISSION && C.fission) C.counts[q.p] = q.nsites;
  if (F::TRACE && q.traced) {
    C.trace_len[q.p] = q.tr;
    C.draws[q.p] = rng.draws;
    C.trace_kacc[q.p] = q.kacc;
  }
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

__device__ void history_surface(const Ctx<double> &C, const DevProblem &P, int64_t p, LocalStats &st) {
  const Tabs<double> &T = C.T;
  const int nc = C.nc, M = C.M;
  const bool xs_every = C.xs_every;
  double *tally = P.tally_glob;
  LaneRng rng;
  Particle<double> q;
  init_particle<FeatAll>(C, p, rng, q);
  birth(C, p, rng, q);
  trace_event<FeatAll>(C, q, NRAPS_EV_BIRTH, q.cell, q.g, 0, q.x, q.mu);
  for (;;) {
    int xs = T.mat(q.cell) + M * q.g;
    double delta_s = q.mu * -nraps_log(nraps_uflight_f64(rng.next())) * T.invtr[xs];
    bool same = true;
    while (same) {
      if (++q.events > C.max_events) {
        st.v[kStCapped]++;
        trace_event<FeatAll>(C, q, NRAPS_EV_CAP, q.cell, q.g, 0, q.x, q.mu);
        finish<FeatAll>(C, q, rng);
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
          trace_event<FeatAll>(C, q, NRAPS_EV_LEAK, cell, q.g, 0, mesh_end, mu);
          finish<FeatAll>(C, q, rng);
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
        st.v[kStCollisions]++;
        int nsite;
        const bool scattered = collide<FeatAll, double>(C, xs, mc, rng, q, nsite);
        if (nsite) emit_sites_plain(C, nsite, q);
        if (!scattered) {
          st.v[kStAbsorb]++;
          finish<FeatAll>(C, q, rng);
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
    Ctx<double> C;
    C.load(P, cs, P.tab);
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < C.N;
         p += (int64_t)gridDim.x * blockDim.x)
      history_surface(C, P, p, st);
  }
  flush_stats(P, st);
}

// ---------------------------------------------------------------------------
// Event pipeline: region surface tracking and Woodcock delta tracking
// ---------------------------------------------------------------------------
// Cell holding the end of a move: moving right (mu >= 0) a point on an edge
// belongs to the cell on its left, moving left to the cell on its right
// (the strict crossing test of src/mc_code.rs:171).  The bucket table gives
// a starting cell; short walks make the answer exact whatever the rounding
// of the bucket index.
template <typename Real>
__device__ inline int locate(const Ctx<Real> &C, Real xe, bool right) {
  const Real *e = C.T.edges;
  int b = (int)((xe - C.XL) * C.inv_h);
  b = b < 0 ? 0 : (b >= C.nb ? C.nb - 1 : b);
  int c = C.T.bucket[b];
  // edge e below the point: e < xe moving right, e <= xe moving left; the
  // conditions are combined without short-circuit branches (every index
  // read is in 0..nc, edges has nc + 1 entries)
  const bool left = !right;
  while ((c > 0) & !((e[c] < xe) | (left & (e[c] == xe)))) c--;
  while ((c + 1 < C.nc) & ((e[c + 1] < xe) | (left & (e[c + 1] == xe)))) c++;
  return c;
}

// Track-length tally of a straight piece from (xa, ca) to (xb, cb) with
// weight w = 1/|mu|: the partial end cells are scored directly, the cells
// in between by two updates of a difference array.  The generation tally
// is direct + width * prefix_sum(difference) (done in k_finalize_rows).
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
__device__ inline double kint(const Ctx<Real> &C, int g, Real x, int c) {
  const Tabs<Real> &T = C.T;
  int r = T.reg(c);
  return T.kreg[g * (C.R + 1) + r] + T.nsfd[T.mat(c) + C.M * g] * ((double)x - T.regedge[r]);
}

// The same integral in region r of material m (what a region-tracked
// particle holds): no cell-to-region or cell-to-material lookups.
template <typename Real>
__device__ inline double kint_rm(const Ctx<Real> &C, int g, Real x, int r, int m) {
  const Tabs<Real> &T = C.T;
  return T.kreg[g * (C.R + 1) + r] + T.nsfd[m + C.M * g] * ((double)x - T.regedge[r]);
}

// XS refresh: the particle's XS group gxs is taken in region r (material of
// `cell`).  Staleness follows src/mc_code.rs:147: with the reference rule the
// caller sets gxs = g only on entering another region.
template <typename Real>
__device__ inline void set_xs(const Ctx<Real> &C, Particle<Real> &q, int r, int cell) {
  q.rxs = r;
  q.xs = C.T.mat(cell) + C.M * q.gxs;
  q.sw = C.T.sigw[q.xs];
}

// Close the straight segment that started at (q.sx, q.scell) at (xe, ce):
// its track-length tally and its k contribution, both with weight 1/|mu|.
// The next segment starts there with k integral `kb`.
// RM: the end lies in the particle's XS region q.rxs (region tracking), so
// the k integral takes that region and its material directly.
#ifndef NRAPS_RS_KINT_RM
#define NRAPS_RS_KINT_RM 1
#endif
template <typename Real, typename Acc, bool RM = false>
__device__ inline void score_segment(const Ctx<Real> &C, Acc *direct, Acc *diff, Particle<Real> &q,
                                     Real xe, int ce) {
  const Real w = Num<Real>::rcp(fabs(q.mu));
  tally_piece<Real, Acc>(direct, diff, C.T.edges, q.g * C.nc, q.sx, q.scell, xe, ce, w);
  const double kb = (RM && NRAPS_RS_KINT_RM) ? kint_rm(C, q.g, xe, q.rxs, q.xs - C.M * q.gxs)
                                             : kint(C, q.g, xe, ce);
  if (q.sx != xe) q.kacc += fabs(kb - q.ks) * (double)w;
  q.ks = kb;
  q.sx = xe;
  q.scell = ce;
}

// One Woodcock step: flight to a tentative site (folding reflections), XS
// refresh when the path left the region of the last refresh, acceptance
// test, and the collision if real.  Virtual collisions change neither the
// direction nor the group, so the path between real events is one straight
// segment: it is scored (tally and k) only where it ends, at a boundary, a
// real collision or the event cap.  Returns kEv* bits; boundary crossings
// are added to `ncross`, fission sites to bank at the particle's position
// are returned in `nsite`.
template <class F, typename Real, typename Acc>
__device__ inline uint32_t woodcock_step(const Ctx<Real> &C, Acc *direct, Acc *diff, LaneRng &rng,
                                         Particle<Real> &q, uint32_t &ncross, int &nsite) {
  const Tabs<Real> &T = C.T;
  const int nc = C.nc;
  if (++q.events > C.max_events) {
    score_segment<Real, Acc>(C, direct, diff, q, q.x, q.cell);
    trace_event<F>(C, q, NRAPS_EV_CAP, q.cell, q.g, 0, q.x, q.mu);
    return kEvCap;
  }
  Real x = q.x;
  int cell = q.cell;
  Real disp = q.mu * -Num<Real>::nlog(Num<Real>::flight(rng.next())) * C.inv_maj;
  bool left_region = false;
  for (;;) {
    const Real mu = q.mu;
    Real end = x + disp;
    if (mu >= Real(0) && end > C.XR) {
      score_segment<Real, Acc>(C, direct, diff, q, C.XR, nc - 1);
      if (T.reg(nc - 1) != q.rxs) left_region = true;
      ncross++;
      if (C.br > Real(0)) {
        q.mu = mu * (-C.br);
        disp = (disp + (x - C.XR)) * (-C.br);
        x = C.XR;
        cell = nc - 1;
        continue;
      }
      trace_event<F>(C, q, NRAPS_EV_LEAK, nc - 1, q.g, 0, C.XR, mu);
      return kEvStep | kEvLeak;
    } else if (mu < Real(0) && C.XL > end) {
      score_segment<Real, Acc>(C, direct, diff, q, C.XL, 0);
      if (T.reg(0) != q.rxs) left_region = true;
      ncross++;
      if (C.bl > Real(0)) {
        q.mu = mu * (-C.bl);
        disp = (disp + (x - C.XL)) * (-C.bl);
        x = C.XL;
        cell = 0;
        continue;
      }
      trace_event<F>(C, q, NRAPS_EV_LEAK, 0, q.g, 0, C.XL, mu);
      return kEvStep | kEvLeak;
    } else {
      if (end != x) {
        cell = locate(C, end, mu >= Real(0));
        x = end;
      }
      break;
    }
  }
  q.x = x;
  q.cell = cell;
  const int r = T.reg(cell);
  if (left_region || r != q.rxs) {
    if (!C.xs_every) q.gxs = q.g;
    set_xs(C, q, r, cell);
  }
  Real u_acc = Num<Real>::u01(rng.next());
  if (!(u_acc * C.maj < q.sw)) return kEvStep;
  score_segment<Real, Acc>(C, direct, diff, q, x, cell);
  const int g = q.g;
  if (!collide<F, Real>(C, q.xs, q.xs - C.M * q.gxs, rng, q, nsite))
    return kEvStep | kEvReal | kEvAbsorb;
  if (q.g != g) {
    q.ks = kint(C, q.g, x, cell);
    if (C.xs_every) {
      q.gxs = q.g;
      set_xs(C, q, r, cell);
    }
  }
  return kEvStep | kEvReal;
}

// One event of surface tracking taken region by region: the reference's
// particle_travel (src/mc_code.rs:134-213) with the cell crossings inside
// one material region merged, since they change neither the flight nor the
// XS.  The flight left in q.ds either ends inside the region (a collision;
// the cell is looked up there), or reaches the region's edge in the
// direction of travel: a face of the slab (reflection keeping the remaining
// displacement, src/mc_code.rs:55-62, or leak) or the edge of the next
// region (XS index refreshed to the current group, src/mc_code.rs:147, and a
// new flight).  A new flight is also sampled after a scatter, with the XS
// index the particle holds (stale after a group change, as in the
// reference).  The straight path is scored where it changes direction, as
// in Woodcock tracking.  Returns kEv* bits; face and region crossings are
// added to `ncross`.
template <class F, typename Real, typename Acc>
__device__ inline uint32_t region_step(const Ctx<Real> &C, Acc *direct, Acc *diff, LaneRng &rng,
                                       Particle<Real> &q, uint32_t &ncross, int &nsite) {
  const Tabs<Real> &T = C.T;
  if (++q.events > C.max_events) {
    score_segment<Real, Acc, true>(C, direct, diff, q, q.x, q.cell);
    trace_event<F>(C, q, NRAPS_EV_CAP, q.cell, q.g, 0, q.x, q.mu);
    return kEvCap;
  }
  const Real x = q.x, mu = q.mu;
  const Real end = x + q.ds;
  const bool fwd = mu >= Real(0);
  const int r = q.rxs;
  const Real edge = (Real)T.regedge[fwd ? r + 1 : r];
  uint32_t ev = kEvStep;
  // both comparisons, combined without short-circuit branches
  const bool past = (fwd & (end > edge)) | (!fwd & (edge > end));
  if (past) {
    ncross++;
    if ((fwd & (r == C.R - 1)) | (!fwd & (r == 0))) {
      const int fc = fwd ? C.nc - 1 : 0;
      score_segment<Real, Acc, true>(C, direct, diff, q, edge, fc);
      const Real b = fwd ? C.br : C.bl;
      if (!(b > Real(0))) {
        trace_event<F>(C, q, NRAPS_EV_LEAK, fc, q.g, 0, edge, mu);
        return kEvStep | kEvLeak;
      }
      q.mu = mu * (-b);
      q.ds = (q.ds + (x - edge)) * (-b);
      q.x = edge;
      q.cell = fc;
      return kEvStep;
    }
    const int rn = fwd ? r + 1 : r - 1;
    q.x = edge;
    // first cell of region r + 1 moving right, last cell of region r - 1
    // moving left: one load at a computed index, no branch
    q.cell = (int)T.regcell[r + (int)fwd] - (int)!fwd;
    q.gxs = q.g;
    set_xs(C, q, rn, q.cell);
  } else {
    if (end != x) {
      q.cell = locate(C, end, fwd);
      q.x = end;
    }
    score_segment<Real, Acc, true>(C, direct, diff, q, q.x, q.cell);
    const int g = q.g;
    if (!collide<F, Real>(C, q.xs, q.xs - C.M * q.gxs, rng, q, nsite))
      return kEvStep | kEvReal | kEvAbsorb;
    if (q.g != g) {
      q.ks = NRAPS_RS_KINT_RM ? kint_rm(C, q.g, q.x, r, q.xs - C.M * q.gxs)
                              : kint(C, q.g, q.x, q.cell);
      if (C.xs_every) {
        q.gxs = q.g;
        set_xs(C, q, r, q.cell);
      }
    }
    ev |= kEvReal;
  }
  q.ds = q.mu * -Num<Real>::nlog(Num<Real>::flight(rng.next())) * T.invtr[q.xs];
  return ev;
}

// Per-wave counters in scalar registers: every update is the population
// count of a ballot taken in wave-uniform control flow.
struct WaveStats {
  uint64_t steps = 0, real = 0, absorb = 0, leak = 0, cap = 0;
  template <typename Real>
  __device__ void add_history(const Particle<Real> &, uint32_t) {}
  __device__ void add(uint32_t ev) {
    steps += __popcll(__ballot(ev & kEvStep));
    real += __popcll(__ballot(ev & kEvReal));
    absorb += __popcll(__ballot(ev & kEvAbsorb));
    leak += __popcll(__ballot(ev & kEvLeak));
    cap += __popcll(__ballot(ev & kEvCap));
  }
  // `region`: region tracking has no virtual collisions
  __device__ void flush(const DevProblem &P, uint32_t ncross, bool region) {
    unsigned long long x = ncross;
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    if (lane_id() == 0) {
      if (real) atomicAdd(&P.stats[kStCollisions], (unsigned long long)real);
      if (steps) atomicAdd(&P.stats[kStSteps], (unsigned long long)steps);
      if (absorb) atomicAdd(&P.stats[kStAbsorb], (unsigned long long)absorb);
      if (leak) atomicAdd(&P.stats[kStLeak], (unsigned long long)leak);
      if (cap) atomicAdd(&P.stats[kStCapped], (unsigned long long)cap);
      if (x) atomicAdd(&P.stats[kStCrossings], x);
      // virtual collisions: flights that were neither real nor a leak
      const uint64_t virt = region ? 0 : steps - real - leak;
      if (virt) atomicAdd(&P.stats[kStVirtual], (unsigned long long)virt);
    }
  }
};

// Per-lane counters, updated only when a history ends (from the history's
// own event and collision counts) and summed over the wave once per
// problem visit: the steps of a history are its events less the capped one.
#ifndef NRAPS_LANE_STATS
#define NRAPS_LANE_STATS 1
#endif
struct LaneStats {
  uint64_t steps = 0, real = 0;
  uint32_t absorb = 0, leak = 0, cap = 0;
  template <typename Real>
  __device__ void add_history(const Particle<Real> &q, uint32_t ev) {
    const uint32_t c = (ev & kEvCap) ? 1u : 0u;
    steps += (uint64_t)(uint32_t)q.events - c;
    real += (uint64_t)(uint32_t)q.nreal;
    absorb += (ev & kEvAbsorb) ? 1u : 0u;
    leak += (ev & kEvLeak) ? 1u : 0u;
    cap += c;
  }
  __device__ void add(uint32_t) {}
  __device__ static unsigned long long wave_sum(unsigned long long x) {
    for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off);
    return x;
  }
  // `region`: region tracking has no virtual collisions
  __device__ void flush(const DevProblem &P, uint32_t ncross, bool region) {
    const unsigned long long st = wave_sum(steps), re = wave_sum(real), ab = wave_sum(absorb),
                             le = wave_sum(leak), ca = wave_sum(cap), cr = wave_sum(ncross);
    if (lane_id() == 0) {
      if (re) atomicAdd(&P.stats[kStCollisions], re);
      if (st) atomicAdd(&P.stats[kStSteps], st);
      if (ab) atomicAdd(&P.stats[kStAbsorb], ab);
      if (le) atomicAdd(&P.stats[kStLeak], le);
      if (ca) atomicAdd(&P.stats[kStCapped], ca);
      if (cr) atomicAdd(&P.stats[kStCrossings], cr);
      const unsigned long long virt = region ? 0 : st - re - le;
      if (virt) atomicAdd(&P.stats[kStVirtual], virt);
    }
  }
};

// Minimum resident waves per SIMD requested from the register allocator
// (measured on one MI355X):
//   FP32, 256-thread blocks, no fission: 5 waves (96 VGPRs, spills only in
//     the refill and boundary paths), 5.6% faster than 4 waves with 127
//     VGPRs and no spills; 6 waves is slower;
//   FP32, 256-thread blocks, fission sites: 4 waves (94 VGPRs with the
//     descriptor in scalar registers), 0.6% faster than 5 waves; with the
//     descriptor in vector registers it needed 128 VGPRs and spilled;
//   1024-thread blocks run 4 waves per SIMD whatever the count; the FP64
//     and trace variants keep the allocator's choice (at least 3 waves);
//   region tracking, FP32, 256-thread blocks: 3, 4 and 5 waves within 0.5%
//     of each other (5 kept); 6 waves is 35% slower; with fission sites 3,
//     4 and 5 waves within 0.5% (4 kept).
#ifndef NRAPS_WC_WAVES_PER_EU
#define NRAPS_WC_WAVES_PER_EU 3
#endif
#ifndef NRAPS_WC_HOT_UNI_TABLES
#define NRAPS_WC_HOT_UNI_TABLES 0
#endif
#ifndef NRAPS_RS_HOT_VGPR
#define NRAPS_RS_HOT_VGPR 0
#endif
#ifndef NRAPS_RS_FP32_WAVES
#define NRAPS_RS_FP32_WAVES NRAPS_WC_FP32_WAVES
#endif
#ifndef NRAPS_RS_FP32_FISSION_WAVES
#define NRAPS_RS_FP32_FISSION_WAVES NRAPS_WC_FP32_FISSION_WAVES
#endif
#ifndef NRAPS_WC_FP32_WAVES
#define NRAPS_WC_FP32_WAVES 5
#endif
#ifndef NRAPS_WC_FP32_FISSION_WAVES
#define NRAPS_WC_FP32_FISSION_WAVES 4
#endif
template <typename Real, bool TRACE, bool FISSION, int BS, bool RS>
constexpr int wc_waves_per_eu() {
  return (sizeof(Real) == 4 && !TRACE && BS == 256)
             ? (RS ? (FISSION ? NRAPS_RS_FP32_FISSION_WAVES : NRAPS_RS_FP32_WAVES)
                   : (FISSION ? NRAPS_WC_FP32_FISSION_WAVES : NRAPS_WC_FP32_WAVES))
             : NRAPS_WC_WAVES_PER_EU;
}

// Persistent event kernel (RS: region surface tracking, else Woodcock
// delta tracking).  A block visits the problems of the batch in
// cyclic order starting at its own index; on a problem with histories left
// it stages the tables in LDS, zeroes its LDS tally, and its waves run
// histories until the problem's queue is empty: each wave claims `claim`
// histories with one global atomic on the queue head and refills idle lanes
// from that range by ballot + mbcnt once at least `refill_min` lanes are
// idle (or none is busy), so lanes stay busy without block barriers.  The
// LDS tally is then flushed with FP64 atomics.  Tallies that do not fit in
// LDS are scored straight into the global FP64 tally.  BS is the block
// size: 256 for small LDS footprints (several blocks per CU), 1024 when one
// block per CU uses the whole LDS.
// Dynamic LDS: [hot tables of the problem][its tally direct|diff].
template <typename Real, bool LDS_TALLY, bool TRACE, bool FISSION, int BS, bool RS>
__global__ __launch_bounds__(BS)
__attribute__((amdgpu_waves_per_eu(wc_waves_per_eu<Real, TRACE, FISSION, BS, RS>())))
void k_event(const DevProblem *probs, int nprob, int claim, int refill_min, int pass) {
  using F = Feat<TRACE, FISSION>;
  extern __shared__ __align__(16) uint8_t smem[];
  __shared__ int s_has_work;
  typedef typename std::conditional<LDS_TALLY, Real, double>::type Acc;
  int loaded = -1;
  for (int k = 0; k < nprob; k++) {
    const int pi = (int)((blockIdx.x + (unsigned)k) % (unsigned)nprob);
    const DevProblem &P = probs[pi];
    if (threadIdx.x == 0) {
      const CycleState *c = P.cs;
      s_has_work = P.pass == pass && c->active &&
                   __hip_atomic_load(&P.cs->queue_head, __ATOMIC_RELAXED,
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
      for (int i = threadIdx.x; i < P.lay.hot / 4; i += BS) dst[i] = src[i];
      loaded = pi;
    }
    const int TB = P.tally_bins;
    // the tally follows this problem's tables (the dynamic LDS holds the
    // largest tables + tally of any one problem of the batch)
    Real *ltally = (Real *)(smem + uni(P.lay.hot));
    if (LDS_TALLY)
      for (int i = threadIdx.x; i < TB; i += BS) ltally[i] = Real(0);
    __syncthreads();

    Ctx<Real> C;
    // descriptor in scalar registers for every variant but the 256-thread
    // FP32 Woodcock one without fission sites: that one measured 0.8% faster
    // with it in vector registers (96 VGPRs, a few spills) than in scalar
    // registers (83 VGPRs), while the fission variant gained 6% from scalar
    // registers and the region-tracking variant 7% (1.81 vs 1.94 ms per
    // launch of 10^7 histories on the 9-region slab)
    constexpr bool kHot =
        sizeof(Real) == 4 && !FISSION && !TRACE && BS == 256 && (!RS || NRAPS_RS_HOT_VGPR);
    C.template load<!kHot, !kHot || NRAPS_WC_HOT_UNI_TABLES>(P, cs, smem);
    Acc *direct = LDS_TALLY ? (Acc *)ltally : (Acc *)P.tally_glob;
    Acc *diff = direct + C.G * C.nc;
    unsigned long long *queue = (unsigned long long *)&P.cs->queue_head;
    // region tracking: per-lane counts (c1 -1.1%, c2 -1.8% against ballots);
    // Woodcock keeps the per-step ballots (per-lane counts measured +1%)
    typename std::conditional<RS && NRAPS_LANE_STATS != 0, LaneStats, WaveStats>::type ws;
    SiteCursor sites;
    uint32_t ncross = 0;

    Particle<Real> q;
    LaneRng rng;
    bool active = false, queue_done = false;
    int64_t r_cur = 0, r_end = 0;  // wave-uniform claimed range
    for (;;) {
      uint64_t need = __ballot(!active);
      if (need && !queue_done && (__popcll(need) >= refill_min || need == ~0ull)) {
        while (need && !queue_done) {
          if (r_cur >= r_end) {
            unsigned long long b = 0;
            if (lane_id() == 0) b = atomicAdd(queue, (unsigned long long)claim);
            b = uniform_u64(b);
            if ((int64_t)b >= C.N) {
              queue_done = true;
              break;
            }
            r_cur = (int64_t)b;
            r_end = r_cur + claim < C.N ? r_cur + claim : C.N;
          }
          const int cnt = __popcll(need);
          const int64_t avail = r_end - r_cur;
          const int take = avail < cnt ? (int)avail : cnt;
          const int rk = rank_below(need);
          if (!active && rk < take) {
            const int64_t p = r_cur + rk;
            init_particle<F>(C, p, rng, q);
            birth(C, p, rng, q);
            q.gxs = q.g;
            set_xs(C, q, C.T.reg(q.cell), q.cell);
            q.sx = q.x;
            q.scell = q.cell;
            q.ks = kint(C, q.g, q.x, q.cell);
            trace_event<F>(C, q, NRAPS_EV_BIRTH, q.cell, q.g, 0, q.x, q.mu);
            if (RS) q.ds = q.mu * -Num<Real>::nlog(Num<Real>::flight(rng.next())) * C.T.invtr[q.xs];
            active = true;
          }
          r_cur += take;
          need = __ballot(!active);
        }
      }
      if (!__any(active)) {
        if (queue_done) break;
        continue;
      }
      uint32_t ev = 0;
      int nsite = 0;
      if (active) {
        if constexpr (RS)
          ev = region_step<F, Real, Acc>(C, direct, diff, rng, q, ncross, nsite);
        else
          ev = woodcock_step<F, Real, Acc>(C, direct, diff, rng, q, ncross, nsite);
      }
      if (F::FISSION && C.make_sites) sites.emit(C, nsite, q);  // wave-uniform
      if (active && (ev & kEvDead)) {
        finish<F>(C, q, rng);
        ws.add_history(q, ev);
        active = false;
      }
      ws.add(ev);
    }
    if (F::FISSION && C.make_sites) sites.close(C);
    ws.flush(P, ncross, RS);
    __syncthreads();
    if (LDS_TALLY) {
      for (int i = threadIdx.x; i < TB; i += BS) {
        Real v = ltally[i];
        if (v != Real(0)) atomicAdd(&P.tally_glob[i], (double)v);
      }
    }
    __syncthreads();
  }
}

#define NRAPS_WC_INST1(R, L, T, Fi, BS)                                                     \
  template __global__ void k_event<R, L, T, Fi, BS, false>(const DevProblem *, int, int, int, \
                                                            int);                             \
  template __global__ void k_event<R, L, T, Fi, BS, true>(const DevProblem *, int, int, int, int);
#define NRAPS_WC_INST(R, L, BS)            \
  NRAPS_WC_INST1(R, L, false, false, BS) \
  NRAPS_WC_INST1(R, L, false, true, BS)  \
  NRAPS_WC_INST1(R, L, true, true, BS)
NRAPS_WC_INST(float, true, 256)
NRAPS_WC_INST(float, false, 256)
NRAPS_WC_INST(float, true, 1024)
NRAPS_WC_INST(double, true, 256)
NRAPS_WC_INST(double, false, 256)
NRAPS_WC_INST(double, true, 1024)
#undef NRAPS_WC_INST
#undef NRAPS_WC_INST1


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
  static_assert(kTreeWidth == 256, "the tree below is written for 256-wide groups");
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active) return;
  __shared__ double s[kTreeWidth / 2];
  const int t = threadIdx.x;
  const int64_t n_in = tree_len(P.N, level);
  const int64_t n_out = (n_in + kTreeWidth - 1) / kTreeWidth;
  const double *in = level == 0 ? P.kacc : ((level - 1) % 2 == 0 ? P.tree_a : P.tree_b);
  double *out = level % 2 == 0 ? P.tree_a : P.tree_b;
  for (int64_t gi = blockIdx.x; gi < n_out; gi += gridDim.x) {
    const int64_t j = gi * kTreeWidth + t;
    double v = j < n_in ? in[j] : 0.0;
    // strides 128 and 64 through LDS: element i takes element i + stride
    if (t >= 128) s[t - 128] = v;
    __syncthreads();
    if (t < 128) v = v + s[t];
    __syncthreads();
    if (t >= 64 && t < 128) s[t - 64] = v;
    __syncthreads();
    if (t < 64) {
      v = v + s[t];
      // strides 32 .. 1 inside the first wave: lane i adds lane i + stride
      for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_down(v, off);
      if (t == 0) out[gi] = v;
    }
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
  int64_t run;
  part[threadIdx.x] = block_exclusive_scan<int64_t>(s, &run);
  if (threadIdx.x == 0) {
    CycleState *cs = P.cs;
    cs->n_sites = run;
    if (run == 0) cs->error |= 2;
    // site_count counts reserved slots (chunk tails included): it bounds
    // the buffer scan of k_place; overflow if it exceeds the capacity
    if ((int64_t)cs->site_count > P.site_cap || (int64_t)cs->site_count < run) cs->error |= 1;
  }
  __syncthreads();
  int64_t r = part[threadIdx.x];
  for (int64_t t = lo; t < hi; t++) {
    int64_t v = ts[t];
    ts[t] = r;
    r += v;
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_apply(const DevProblem *probs, const int64_t *tile_sums,
                                                       const int64_t *tile_base) {
  const DevProblem &P = probs[blockIdx.y];
  if (!P.cs->active || !P.cs->make_sites) return;
  const int64_t ntiles = (P.N + kScanTile - 1) / kScanTile;
  constexpr int per = kScanTile / kBlock;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t b0 = t * kScanTile + (int64_t)threadIdx.x * per;
    int64_t c[per];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < per; i++) {
      const int64_t j = b0 + i;
      c[i] = j < P.N ? P.counts[j] : 0;
      s += c[i];
    }
    int64_t tot;
    int64_t run = tile_sums[tile_base[blockIdx.y] + t] + block_exclusive_scan<int64_t>(s, &tot);
#pragma unroll
    for (int i = 0; i < per; i++) {
      const int64_t j = b0 + i;
      if (j < P.N) P.offsets[j] = run;
      run += c[i];
    }
  }
}

template <typename Real>
__global__ __launch_bounds__(kBlock) void k_place(const DevProblem *probs) {
  const DevProblem &P = probs[blockIdx.y];
  const CycleState *cs = P.cs;
  if (!cs->active || !cs->make_sites || (cs->error & 1)) return;
  const int64_t n = (int64_t)cs->site_count;  // reserved slots, <= site_cap
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t key = P.site_key[i];
    if (key == ~0ull) continue;  // unused tail of a wave's slot chunk
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
  nraps_history_rng(&r, P.seed, nraps_stream_id(P.problem_id, (uint64_t)cs->cycle, NRAPS_SERVICE_PARTICLE));
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
// k_finalize_rows: one block per (group, problem) rebuilds the cell tally of
// its group row from the difference array and updates that row's
// accumulators; the per-cell fission-source terms go to the consumed
// difference row and the row's share of the reference k to kpart[g].
// k_finalize_close: one block per problem sums those over groups in group
// order and advances the generation state.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_finalize_rows(const DevProblem *probs,
                                                          const DevFinal *fins) {
  const DevProblem &P = probs[blockIdx.y];
  const DevFinal &F = fins[blockIdx.y];
  const int g = (int)blockIdx.x;
  if (!P.cs->active || g >= P.G) return;
  const int nc = P.nc, G = P.G, M = P.M;
  const int cyc = P.cs->cycle;
  const double kprev = P.cs->k_prev;
  const double N = (double)P.N;
  double *row = F.tally_cycle + g * nc;             // direct part -> cell tally
  double *drow = F.tally_cycle + G * nc + g * nc;   // difference row -> scratch
  if (F.has_diff) {
    // cell tally = direct + width * inclusive_prefix(diff)
    const int per = (nc + kBlock - 1) / kBlock;
    const int lo = threadIdx.x * per, hi = lo + per < nc ? lo + per : nc;
    double s = 0.0;
    for (int c = lo; c < hi; c++) s += drow[c];
    double total;
    double run = block_exclusive_scan<double>(s, &total);
    for (int c = lo; c < hi; c++) {
      run += drow[c];
      row[c] += F.width[c] * run;
    }
    __syncthreads();  // the difference row is reused below
  }
  const bool active_gen = cyc >= F.skip;
  const double fund = 1.0 / (double)(F.generations - (F.skip - 1));
  const double conversion =
      (3565e6 * kprev * 36.2) / (200e6 * 1.602176634e-19 * F.nut_conv * F.length);
  const bool last = cyc == F.generations - 1 && F.tally_last;
  double kref = 0.0;
  for (int c = threadIdx.x; c < nc; c += blockDim.x) {
    const double dxc = F.dx[c];
    const int xi = F.mat[c] + M * g;
    const double t = row[c];
    const double flux = t / (kprev * N * dxc);
    const double fsrc = F.nut[xi] * F.sigf[xi] * flux;
    kref += kprev * dxc * fsrc;
    if (active_gen) {
      F.flux_acc[g * nc + c] += flux * conversion * fund;
      drow[c] = fsrc * fund;
      const double fv = t / (N * dxc);
      F.fsum[g * nc + c] += fv;
      F.fsq[g * nc + c] += fv * fv;
    }
    if (last) F.tally_last[g * nc + c] = t;
  }
  __shared__ double part[kBlock];
  part[threadIdx.x] = kref;
  __syncthreads();
  for (int stride = kBlock / 2; stride >= 1; stride >>= 1) {
    if ((int)threadIdx.x < stride) part[threadIdx.x] += part[threadIdx.x + stride];
    __syncthreads();
  }
  if (threadIdx.x == 0) F.kpart[g] = part[0];
}

__global__ __launch_bounds__(kBlock) void k_finalize_close(const DevProblem *probs,
                                                           const DevFinal *fins, int levels) {
  const DevProblem &P = probs[blockIdx.x];
  const DevFinal &F = fins[blockIdx.x];
  CycleState *csp = P.cs;
  if (!csp->active) return;
  const int nc = P.nc, G = P.G;
  const int cyc = csp->cycle;
  const double *tot = (levels - 1) % 2 == 0 ? P.tree_a : P.tree_b;
  const double knew = tot[0] / (double)P.N;
  if (cyc >= F.skip) {
    const double *fsr = F.tally_cycle + G * nc;  // fission-source terms, [G][nc]
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      double fs_c = 0.0;
      for (int g = 0; g < G; g++) fs_c += fsr[g * nc + c];
      F.fs_acc[c] += fs_c;
    }
  }
  if (threadIdx.x == 0) {
    double kref = 0.0;
    for (int g = 0; g < G; g++) kref += F.kpart[g];
    P.k[cyc] = knew;
    F.k_ref[cyc] = kref;
    F.sites_hist[cyc] = csp->make_sites ? csp->n_sites : 0;
    const int next = cyc + 1;
    csp->k_prev = knew;
    csp->inv_k = 1.0 / knew;
    csp->cycle = next;
    csp->from_bank = P.fission && (P.dstart || next >= 2);
    csp->make_sites = P.fission && (P.dstart || next >= 1);
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

// Blocks per problem for the grid-stride kernels of a batch (blockIdx.y is
// the problem): enough for the largest problem alone, but with many
// problems of uneven size a cap of ~64k blocks in the whole launch, since
// every block of a small problem past its end still costs a dispatch.
static inline unsigned batch_grid(int64_t n, int64_t cap, int nprob) {
  const int64_t share = std::max<int64_t>(256, (int64_t)65536 / std::max(nprob, 1));
  return clamp_grid(n, std::min(cap, share));
}

void launch_cycle_begin(const DevProblem *d, int nprob, int max_bins, hipStream_t s) {
  dim3 grid(clamp_grid((max_bins + kBlock - 1) / kBlock, 64), (unsigned)nprob);
  hipLaunchKernelGGL(k_cycle_begin, grid, dim3(kBlock), 0, s, d);
}

void launch_surface(const DevProblem *d, int nprob, int64_t max_n, hipStream_t s) {
  dim3 grid(clamp_grid((max_n + kBlock - 1) / kBlock, 1 << 20), (unsigned)nprob);
  hipLaunchKernelGGL(k_surface, grid, dim3(kBlock), 0, s, d, nprob);
}

// Variant of the event kernel: precision, tally placement, and the
// compile-time features (traces imply the fission code path).
template <typename Real, bool L, bool T, bool Fi, int BS>
static const void *wc_fn(bool region) {
  return region ? reinterpret_cast<const void *>(&k_event<Real, L, T, Fi, BS, true>)
                : reinterpret_cast<const void *>(&k_event<Real, L, T, Fi, BS, false>);
}
template <typename Real, bool L, int BS>
static const void *wc_pick(bool trace, bool fission, bool region) {
  if (trace) return wc_fn<Real, L, true, true, BS>(region);
  return fission ? wc_fn<Real, L, false, true, BS>(region) : wc_fn<Real, L, false, false, BS>(region);
}
template <typename Real>
static const void *wc_pick_mode(const WcVariant &v) {
  if (v.big_block) return wc_pick<Real, true, 1024>(v.trace, v.fission, v.region);
  return v.lds_tally ? wc_pick<Real, true, 256>(v.trace, v.fission, v.region)
                     : wc_pick<Real, false, 256>(v.trace, v.fission, v.region);
}
static const void *wc_select(const WcVariant &v) {
  return v.fp32 ? wc_pick_mode<float>(v) : wc_pick_mode<double>(v);
}

void launch_event(const WcVariant &v, const DevProblem *d, int nprob, int grid, size_t lds_bytes,
                     int claim, int refill_min, int pass, hipStream_t s) {
  // in the order of k_event's parameters
  void *args[] = {(void *)&d, (void *)&nprob, (void *)&claim, (void *)&refill_min,
                  (void *)&pass};
  hipError_t e = hipLaunchKernel(wc_select(v), dim3((unsigned)grid), dim3(v.block_size()), args,
                                 lds_bytes, s);
  (void)e;  // reported by the caller's hipGetLastError
}

hipError_t event_set_lds_limit(const WcVariant &v, size_t lds_bytes) {
  return hipFuncSetAttribute(wc_select(v), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
}

int event_blocks_per_cu(const WcVariant &v, size_t lds_bytes) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, wc_select(v), v.block_size(), lds_bytes) !=
      hipSuccess)
    return 1;
  return nb > 0 ? nb : 1;
}

void launch_ktree(const DevProblem *d, int nprob, int level, int64_t max_groups, hipStream_t s) {
  dim3 grid(batch_grid(max_groups, 65535, nprob), (unsigned)nprob);
  hipLaunchKernelGGL(k_ktree, grid, dim3(kTreeWidth), 0, s, d, level);
}

void launch_bank(bool fp32, const DevProblem *d, int nprob, int64_t max_n, int64_t max_tiles,
                 int64_t *tile_sums, const int64_t *tile_base, hipStream_t s) {
  dim3 gt(batch_grid(max_tiles, 65535, nprob), (unsigned)nprob);
  hipLaunchKernelGGL(k_scan_tiles, gt, dim3(kBlock), 0, s, d, tile_sums, tile_base);
  hipLaunchKernelGGL(k_scan_tilesums, dim3((unsigned)nprob), dim3(kBlock), 0, s, d, tile_sums, tile_base);
  hipLaunchKernelGGL(k_scan_apply, gt, dim3(kBlock), 0, s, d, (const int64_t *)tile_sums, tile_base);
  dim3 gp(batch_grid((4 * max_n + 65536 + kBlock - 1) / kBlock, 8192, nprob), (unsigned)nprob);
  dim3 gr(batch_grid((max_n + kBlock - 1) / kBlock, 8192, nprob), (unsigned)nprob);
  if (fp32) {
    hipLaunchKernelGGL(k_place<float>, gp, dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_resample<float>, gr, dim3(kBlock), 0, s, d);
  } else {
    hipLaunchKernelGGL(k_place<double>, gp, dim3(kBlock), 0, s, d);
    hipLaunchKernelGGL(k_resample<double>, gr, dim3(kBlock), 0, s, d);
  }
}

void launch_finalize(const DevProblem *d, const DevFinal *f, int nprob, int max_groups, int levels,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_finalize_rows, dim3((unsigned)max_groups, (unsigned)nprob), dim3(kBlock), 0, s,
                     d, f);
  hipLaunchKernelGGL(k_finalize_close, dim3((unsigned)nprob), dim3(kBlock), 0, s, d, f, levels);
}

void launch_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n, int draws, uint32_t *out,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_pcg32_fill, dim3(clamp_grid((n + kBlock - 1) / kBlock, 1 << 30)), dim3(kBlock), 0,
                     s, seed, stream0, n, draws, out);
}

}  // namespace nraps
