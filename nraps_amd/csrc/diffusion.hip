// diffusion.hip — finite-difference multigroup diffusion k-eigenvalue on the
// device (nalgebra_method, src/discrete.rs:169-356), and the fission-bank
// start that draws generation-0 sources from its fission source.
//
// The reference builds each group's n x n matrix (matrix_gen,
// src/discrete.rs:4-101), inverts it densely once (:186-205) and then, per
// power iteration, multiplies every row of the inverse with the right-hand
// side q/k + in-scatter (:217-284).  The matrix is tridiagonal.  Here the host
// factors it once (Gaussian elimination without pivoting, the Thomas
// algorithm), and the device solves each group in O(n): the forward and back
// substitutions are first-order linear recurrences, which one workgroup of 256
// or 1024 threads evaluates as a parallel scan of affine maps (per-thread runs
// of L cells, DPP moves inside a wave, one LDS step across the waves).  The whole power
// iteration, convergence test included, runs in one launch with no host round
// trip; a batch puts one problem on each workgroup.
//
// Kept as the reference has them: the matrix entries (harmonic-mean
// couplings, the beta boundary rule, the last row's coupling of cells 0 and 1),
// the group sweep in order with the newest flux of the other groups, the flux
// and k change measures and the loop condition (:207, :252-289), and the
// power normalisation (:293-343).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>

#include "diffusion.h"
#include "host.h"
#include "runtime.h"

namespace nraps {

namespace {

// Doubles of LDS scratch after the tables: scan totals and reductions for up
// to 16 waves, and the change-measure slots.
constexpr int kScratch = 128;

#define HIP_TRY NRAPS_HIP_TRY

// y -> a*y + b.  compose(p, q) applies p first, then q.
struct Aff {
  double a, b;
};
__device__ __forceinline__ Aff compose(Aff p, Aff q) { return Aff{q.a * p.a, q.a * p.b + q.b}; }

// DPP controls (GFX9 encoding): shift right within a row of 16 lanes, lane
// 15 of a row to the next row, lane 31 to rows 2 and 3, whole wave right by 1.
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

// A double moved across lanes by DPP; lanes without a source (or outside
// ROW_MASK) get `old`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v, double old) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROW_MASK,
                                             0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROW_MASK,
                                             0xf, false);
  return __hiloint2double(hi, lo);
}
// The map of the lane CTRL names, or the identity where there is none.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ Aff dpp_take(Aff v) {
  return Aff{dpp_f64<CTRL, ROW_MASK>(v.a, 1.0), dpp_f64<CTRL, ROW_MASK>(v.b, 0.0)};
}
__device__ __forceinline__ Aff readlane(Aff v, int l) {
  auto rl = [&](double x) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), l),
                            __builtin_amdgcn_readlane(__double2loint(x), l));
  };
  return Aff{rl(v.a), rl(v.b)};
}
__device__ __forceinline__ Aff lane_from(Aff v, int src) {
  return Aff{__shfl(v.a, src), __shfl(v.b, src)};
}

// Inclusive scan over lanes 0..63 in lane order: rows of 16 by shifts of
// 1, 2, 4, 8, then the row totals by two broadcasts (all DPP moves).
__device__ __forceinline__ Aff wave_inscan(Aff v) {
  v = compose(dpp_take<kDppRowShr1>(v), v);
  v = compose(dpp_take<kDppRowShr2>(v), v);
  v = compose(dpp_take<kDppRowShr4>(v), v);
  v = compose(dpp_take<kDppRowShr8>(v), v);
  v = compose(dpp_take<kDppRowBcast15, 0xa>(v), v);
  v = compose(dpp_take<kDppRowBcast31, 0xc>(v), v);
  return v;
}

// Exclusive scan of one affine map per thread of a BS-thread workgroup, in
// thread order (FWD) or in reverse thread order; wtot: 2*BS/64 doubles of LDS.
// The reverse order reverses the lanes (one shuffle each way) and scans
// forward.  The wave totals are combined by a scan over the first BS/64
// lanes of every wave.
template <bool FWD, bool SYNC_AFTER, int BS>
__device__ __forceinline__ Aff block_exscan(Aff v, double *wtot) {
  constexpr int kW = BS / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const Aff in = FWD ? v : lane_from(v, 63 - lane);
  const Aff inc = wave_inscan(in);
  const Aff ex_s = dpp_take<kDppWaveShr1>(inc);  // exclusive, in scan order
  const Aff ex = FWD ? ex_s : lane_from(ex_s, 63 - lane);
  const Aff tot = readlane(inc, 63);
  if (lane == 0) {
    wtot[2 * w] = tot.a;
    wtot[2 * w + 1] = tot.b;
  }
  __syncthreads();
  // scan of the wave totals in scan order: lane r holds wave r (FWD) or
  // wave kW-1-r
  Aff t{1.0, 0.0};
  if (lane < kW) {
    const int u = FWD ? lane : kW - 1 - lane;
    t = Aff{wtot[2 * u], wtot[2 * u + 1]};
  }
  t = compose(dpp_take<kDppRowShr1>(t), t);
  if (kW > 2) t = compose(dpp_take<kDppRowShr2>(t), t);
  if (kW > 4) t = compose(dpp_take<kDppRowShr4>(t), t);
  if (kW > 8) t = compose(dpp_take<kDppRowShr8>(t), t);
  const int r = FWD ? w : kW - 1 - w;  // this wave's place in scan order
  const Aff pre = r > 0 ? readlane(t, r - 1) : Aff{1.0, 0.0};
  if (SYNC_AFTER) __syncthreads();  // wtot is written again by the next scan
  return compose(pre, ex);
}

// Sum over the workgroup in a fixed order; red: BS/64 doubles of LDS.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double *red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int u = 0; u < BS / 64; u++) s += red[u];
  __syncthreads();
  return s;
}

struct Tabs {
  const double *nsf, *chi, *nut, *sigf, *scat;
};

// q_gen (src/discrete.rs:103-126): fission rate F_i = sum_x nut sigf flux
// per cell, and the sum over groups and cells of q[g][i] = F_i dx_i chi_g.
template <int BS>
__device__ __forceinline__ double qgen(const DiffProblem &P, const Tabs &T, const double *fl,
                                       double *F, double *red) {
  const int t = threadIdx.x, n = P.n, G = P.G, M = P.M, L = P.L, LT = L * BS;
  double part = 0.0;
  for (int j = 0; j < L; j++) {
    if (t * L + j >= n) break;
    const int idx = j * BS + t;
    const int m = P.mat[idx];
    double f = 0.0;
    for (int x = 0; x < G; x++) f += T.nsf[m + M * x] * fl[x * LT + idx];
    F[idx] = f;
    const double fd = f * P.dx[idx];
    for (int g = 0; g < G; g++) part += fd * T.chi[m + M * g];
  }
  return block_sum<BS>(part, red);
}

template <int BS>
__device__ __forceinline__ void diffusion_run(const DiffProblem &P, const Tabs &T, double *scr,
                                              double *fl, double *F) {
  constexpr int kW = BS / 64;
  const int t = threadIdx.x, n = P.n, G = P.G, M = P.M, L = P.L, LT = L * BS;
  double *wtot = scr, *red = scr + 2 * kW, *sc = scr + 3 * kW;
  for (int j = 0; j < L; j++) {
    if (t * L + j >= n) break;
    for (int g = 0; g < G; g++) fl[g * LT + j * BS + t] = 1.0;
  }
  double sumq = qgen<BS>(P, T, fl, F, red);
  double k = 1.0, dflux = 1.0, dk = 1.0;
  int it = 0, status = NRAPS_OK;
  while (dflux >= P.ftol && dk >= P.ktol) {
    if (it >= P.max_it) {
      status = NRAPS_E_NOT_CONVERGED;
      break;
    }
    it++;
    const double sq_old = sumq, invk = 1.0 / k;
    for (int g = 0; g < G; g++) {
      double *fg = fl + g * LT;
      const double *al = P.alpha + (size_t)g * LT, *mg = P.mdiag + (size_t)g * LT,
                   *ga = P.gamma + (size_t)g * LT;
      // old flux at the cells the change measure reads
      if (t == P.own0) sc[0] = fg[P.pos0];
      if (t == P.own1) sc[1] = fg[P.pos1];
      if (t == P.own2) sc[2] = fg[P.pos2];
      // right-hand side q/k + in-scatter (scat_calc, src/discrete.rs:128-149)
      // and the thread's composed forward map
      Aff acc{1.0, 0.0};
      for (int j = 0; j < L; j++) {
        if (t * L + j >= n) break;
        const int idx = j * BS + t;
        const int m = P.mat[idx];
        const double d = P.dx[idx];
        double s = 0.0;
        for (int e = 0; e < G; e++)
          if (e != g) s += T.scat[G * G * m + G * e + g] * fl[e * LT + idx] * d;
        const double rhs = F[idx] * d * T.chi[m + M * g] * invk + s;
        const double beta = rhs * mg[idx];
        fg[idx] = beta;
        acc = compose(acc, Aff{al[idx], beta});
      }
      double y = block_exscan<true, true, BS>(acc, wtot).b;
      for (int j = 0; j < L; j++) {
        if (t * L + j >= n) break;
        const int idx = j * BS + t;
        y = al[idx] * y + fg[idx];
        fg[idx] = y;
      }
      acc = Aff{1.0, 0.0};
      for (int j = L - 1; j >= 0; j--) {
        if (t * L + j >= n) continue;
        const int idx = j * BS + t;
        acc = compose(acc, Aff{ga[idx], fg[idx]});
      }
      double x = block_exscan<false, true, BS>(acc, wtot).b;
      for (int j = L - 1; j >= 0; j--) {
        if (t * L + j >= n) continue;
        const int idx = j * BS + t;
        x = ga[idx] * x + fg[idx];
        fg[idx] = x;
      }
      __syncthreads();
      if (t == 0) {
        // src/discrete.rs:252-283
        const double new0 = fg[P.pos0];
        double dfl = fmax(fabs((sc[0] - new0) / sc[0]), dflux);
        if (n >= 3) dfl = fabs((sc[1] - fg[P.pos1]) / new0);
        dfl = fmax(fabs((sc[2] - fg[P.pos2]) / new0), dfl);
        sc[3] = dfl;
      }
      __syncthreads();
      dflux = sc[3];
    }
    sumq = qgen<BS>(P, T, fl, F, red);
    const double kold = k;
    k = kold * (sumq / sq_old);
    dk = fabs((k - kold) / kold);
  }
  // power normalisation (src/discrete.rs:304-343): the group-major array of
  // nut sigf flux is summed in chunks of G, chunk j weighted by dx_j
  double part = 0.0;
  for (int j = 0; j < L; j++) {
    const int i = t * L + j;
    if (i >= n) break;
    const int idx = j * BS + t;
    const int m = P.mat[idx];
    for (int g = 0; g < G; g++)
      part += fl[g * LT + idx] * T.nut[m + M * g] * T.sigf[m + M * g] * P.dx_nat[(g * n + i) / G];
  }
  const double ptot = block_sum<BS>(part, red);
  const double pc = 3565e6 / (1.6022e-13 * 200.0 * ptot);
  for (int j = 0; j < L; j++) {
    const int i = t * L + j;
    if (i >= n) break;
    const int idx = j * BS + t;
    const int m = P.mat[idx];
    double fs = 0.0;
    for (int g = 0; g < G; g++) {
      const double v = fl[g * LT + idx] * pc;
      P.flux_out[(size_t)g * n + i] = v;
      fs += T.nsf[m + M * g] * v;
    }
    P.fs_out[i] = fs;
  }
  if (t == 0) {
    P.scal[0] = k;
    P.scal[1] = dflux;
    P.scal[2] = dk;
    P.scal[3] = (double)it;
    P.scal[4] = (double)status;
  }
}

// Register-resident form for G <= GM groups and L <= LM cells per thread:
// each thread keeps its cells' flux, fission rate, material and width in
// registers for the whole solve, so LDS carries only the tables and the
// scan/reduction scratch.  Same arithmetic, in the same order, as
// diffusion_run.  Per group: loads of the group's factors (independent),
// the right-hand side, two scans and one barrier for the change measure;
// the scan totals and the change-measure slots alternate between two
// buffers so no barrier is needed to recycle them.
template <int GM, int LM, int BS>
__device__ __forceinline__ void diffusion_regs(const DiffProblem &P, const Tabs &T, double *scr) {
  constexpr int kW = BS / 64;
  const int t = threadIdx.x, n = P.n, G = P.G, M = P.M, L = P.L, LT = L * BS;
  double *wtA = scr, *wtB = scr + 2 * kW, *redA = scr + 4 * kW, *redB = scr + 5 * kW;
  double *scs = scr + 6 * kW;  // [2][6]: old and new flux at cells 0, n-2, n-1
  // the thread's cell widths, fission rates and materials: LDS slots
  // [LM][256] of this thread only (registers stay free for the flux)
  double *dx = scr + kScratch, *F = dx + LM * BS;
  int *mat = reinterpret_cast<int *>(F + LM * BS);
  const int jn = max(0, min(L, n - t * L));
  const int j0 = P.pos0 / BS, j1 = P.pos1 / BS, j2 = P.pos2 / BS;
  double phi[GM][LM];
#pragma unroll
  for (int j = 0; j < LM; j++) {
    const bool on = j < jn;
    mat[j * BS + t] = on ? P.mat[j * BS + t] : 0;
    dx[j * BS + t] = on ? P.dx[j * BS + t] : 0.0;
#pragma unroll
    for (int g = 0; g < GM; g++) phi[g][j] = 1.0;
  }
  auto qgen = [&]() {
    double part = 0.0;
#pragma unroll
    for (int j = 0; j < LM; j++) {
      if (j < jn) {
        const int m = mat[j * BS + t];
        double f = 0.0;
#pragma unroll
        for (int x = 0; x < GM; x++)
          if (x < G) f += T.nsf[m + M * x] * phi[x][j];
        F[j * BS + t] = f;
        const double fd = f * dx[j * BS + t];
#pragma unroll
        for (int g = 0; g < GM; g++)
          if (g < G) part += fd * T.chi[m + M * g];
      }
    }
    return block_sum<BS>(part, redA);
  };
  double sumq = qgen();
  double k = 1.0, dflux = 1.0, dk = 1.0;
  int it = 0, status = NRAPS_OK;
  while (dflux >= P.ftol && dk >= P.ktol) {
    if (it >= P.max_it) {
      status = NRAPS_E_NOT_CONVERGED;
      break;
    }
    it++;
    const double sq_old = sumq, invk = 1.0 / k;
#pragma unroll
    for (int g = 0; g < GM; g++) {
      if (g >= G) break;
      const size_t go = (size_t)g * LT;
      double fa[LM], fm[LM], fc[LM], v[LM];
#pragma unroll
      for (int j = 0; j < LM; j++) {
        if (j < jn) {
          fa[j] = P.alpha[go + j * BS + t];
          fm[j] = P.mdiag[go + j * BS + t];
          fc[j] = P.gamma[go + j * BS + t];
        }
      }
      double *sc = scs + 6 * (g & 1);
#pragma unroll
      for (int j = 0; j < LM; j++) {
        if (t == P.own0 && j == j0) sc[0] = phi[g][j];
        if (t == P.own1 && j == j1) sc[1] = phi[g][j];
        if (t == P.own2 && j == j2) sc[2] = phi[g][j];
      }
      Aff acc{1.0, 0.0};
#pragma unroll
      for (int j = 0; j < LM; j++) {
        if (j < jn) {
          const int m = mat[j * BS + t];
          const double d = dx[j * BS + t];
          double sct = 0.0;
#pragma unroll
          for (int e = 0; e < GM; e++)
            if (e < G && e != g) sct += T.scat[G * G * m + G * e + g] * phi[e][j] * d;
          const double rhs = F[j * BS + t] * d * T.chi[m + M * g] * invk + sct;
          v[j] = rhs * fm[j];
          acc = compose(acc, Aff{fa[j], v[j]});
        }
      }
      double y = block_exscan<true, false, BS>(acc, (g & 1) ? wtB : wtA).b;
#pragma unroll
      for (int j = 0; j < LM; j++) {
        if (j < jn) {
          y = fa[j] * y + v[j];
          v[j] = y;
        }
      }
      acc = Aff{1.0, 0.0};
#pragma unroll
      for (int j = LM - 1; j >= 0; j--)
        if (j < jn) acc = compose(acc, Aff{fc[j], v[j]});
      double x = block_exscan<false, false, BS>(acc, (g & 1) ? wtA : wtB).b;
#pragma unroll
      for (int j = LM - 1; j >= 0; j--) {
        if (j < jn) {
          x = fc[j] * x + v[j];
          phi[g][j] = x;
        }
      }
#pragma unroll
      for (int j = 0; j < LM; j++) {
        if (t == P.own0 && j == j0) sc[3] = phi[g][j];
        if (t == P.own1 && j == j1) sc[4] = phi[g][j];
        if (t == P.own2 && j == j2) sc[5] = phi[g][j];
      }
      __syncthreads();
      // src/discrete.rs:252-283, evaluated by every thread
      const double new0 = sc[3];
      double dfl = fmax(fabs((sc[0] - new0) / sc[0]), dflux);
      if (n >= 3) dfl = fabs((sc[1] - sc[4]) / new0);
      dflux = fmax(fabs((sc[2] - sc[5]) / new0), dfl);
    }
    sumq = qgen();
    const double kold = k;
    k = kold * (sumq / sq_old);
    dk = fabs((k - kold) / kold);
  }
  double part = 0.0;
#pragma unroll
  for (int j = 0; j < LM; j++) {
    if (j < jn) {
      const int i = t * L + j, m = mat[j * BS + t];
#pragma unroll
      for (int g = 0; g < GM; g++)
        if (g < G) part += phi[g][j] * T.nut[m + M * g] * T.sigf[m + M * g] * P.dx_nat[(g * n + i) / G];
    }
  }
  const double ptot = block_sum<BS>(part, redB);
  const double pc = 3565e6 / (1.6022e-13 * 200.0 * ptot);
#pragma unroll
  for (int j = 0; j < LM; j++) {
    if (j < jn) {
      const int i = t * L + j, m = mat[j * BS + t];
      double fs = 0.0;
#pragma unroll
      for (int g = 0; g < GM; g++) {
        if (g < G) {
          const double val = phi[g][j] * pc;
          P.flux_out[(size_t)g * n + i] = val;
          fs += T.nsf[m + M * g] * val;
        }
      }
      P.fs_out[i] = fs;
    }
  }
  if (t == 0) {
    P.scal[0] = k;
    P.scal[1] = dflux;
    P.scal[2] = dk;
    P.scal[3] = (double)it;
    P.scal[4] = (double)status;
  }
}

// Tables to LDS; returns the scratch that follows them.
template <int BS>
__device__ __forceinline__ double *load_tables(const DiffProblem &P, double *sm, Tabs &T) {
  const int G = P.G, M = P.M, GM = G * M, GGM = G * G * M;
  for (int i = threadIdx.x; i < GM; i += BS) {
    sm[i] = P.nsf[i];
    sm[GM + i] = P.chi[i];
    sm[2 * GM + i] = P.nut[i];
    sm[3 * GM + i] = P.sigf[i];
  }
  for (int i = threadIdx.x; i < GGM; i += BS) sm[4 * GM + i] = P.scat[i];
  T = Tabs{sm, sm + GM, sm + 2 * GM, sm + 3 * GM, sm + 4 * GM};
  __syncthreads();
  return sm + 4 * GM + GGM;
}

template <int GM, int LM, int BS>
__global__ __launch_bounds__(BS) void k_diffusion_regs(const DiffProblem *probs) {
  extern __shared__ double sm[];
  const DiffProblem &P = probs[blockIdx.x];
  Tabs T;
  double *scr = load_tables<BS>(P, sm, T);
  diffusion_regs<GM, LM, BS>(P, T, scr);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_diffusion(const DiffProblem *probs) {
  extern __shared__ double sm[];
  const DiffProblem &P = probs[blockIdx.x];
  const int G = P.G;
  Tabs T;
  double *scr = load_tables<BS>(P, sm, T);
  if (P.in_lds) {
    double *fl = scr + kScratch;
    diffusion_run<BS>(P, T, scr, fl, fl + (size_t)G * P.L * BS);
  } else {
    diffusion_run<BS>(P, T, scr, P.gflux, P.gF);
  }
}

// Fission-bank start, part 1 (one thread): cumulative weights s_c dx_c in
// cell order, and generation 0's state.
__global__ void k_dbank_cdf(const DiffProblem *dp, double *cdf, CycleState *cs) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const DiffProblem &P = *dp;
  double acc = 0.0;
  for (int c = 0; c < P.n; c++) {
    const double s = P.fs_out[c];
    acc = acc + (s > 0.0 ? s * P.dx_nat[c] : 0.0);
    cdf[c] = acc;
  }
  const double k = P.scal[0];
  if (!(acc > 0.0) || !(k > 0.0)) {
    cs->error |= 2;
    cs->active = 0;
    return;
  }
  cs->inv_k = 1.0 / k;
}

// Part 2: source p at the quantile (p + 1/2)/N of the weights, placed
// linearly inside its cell and clamped to the cell's right edge.
template <typename Real>
__global__ void k_dbank_sites(const DiffProblem *dp, const double *cdf, int64_t N, Real *src_x,
                              int32_t *src_cell) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const DiffProblem &P = *dp;
  const int n = P.n;
  const double acc = cdf[n - 1];
  if (!(acc > 0.0)) return;
  const double t = (((double)p + 0.5) / (double)N) * acc;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cdf[mid] > t) hi = mid;
    else lo = mid + 1;
  }
  const int c = lo;
  const double below = c > 0 ? cdf[c - 1] : 0.0;
  const double w = cdf[c] - below;
  const double frac = w > 0.0 ? (t - below) / w : 0.5;
  const double xx = P.left[c] + frac * P.dx_nat[c];
  src_x[p] = (Real)(xx > P.right[c] ? P.right[c] : xx);
  src_cell[p] = c;
}

double d_over_dx(const nraps_problem *pb, const std::vector<double> &ist, int c, int g) {
  return (1.0 / 3.0) * ist[pb->matid[c] + pb->n_mats * g] / pb->delta_x[c];
}

double coupling(double a, double b) { return (2.0 * a * b) / (a + b); }

// Boundary beta, src/discrete.rs:26-33 and :81-88.
double beta_of(double bound, double d_next, double d_curr) {
  if (bound == 1.0) return 1.0;
  if (bound == 0.0) return 0.25;
  const double r = (1.0 - bound) / (1.0 + bound);
  return (1.0 - 0.25 * (r * (1.0 / d_next))) / (1.0 + 0.25 * (r * (1.0 / d_curr)));
}

}  // namespace

void DiffusionSet::add(const nraps_problem *pb, const nraps_diffusion_opts &o, int lds_limit) {
  if (!pb) throw std::invalid_argument("problem is NULL");
  const int n = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  if (n < 2 || n > 1000000) throw std::invalid_argument("diffusion needs n_cells in [2, 10^6]");
  if (G < 1 || G > 255 || M < 1 || M > 255)
    throw std::invalid_argument("n_groups and n_mats must be in [1, 255]");
  if (!pb->matid || !pb->delta_x || !pb->mesh_left || !pb->mesh_right || !pb->sigt || !pb->nut ||
      !pb->sigf || !pb->chit || !pb->scat_matrix || (!pb->inv_sigtr && (!pb->mu || !pb->sigs)))
    throw std::invalid_argument("diffusion needs the mesh, sigt, nut, sigf, chit, scat_matrix "
                                "and inv_sigtr (or mu and sigs)");
  for (int c = 0; c < n; c++) {
    if (pb->matid[c] < 0 || pb->matid[c] >= M)
      throw std::invalid_argument("matid out of range at cell " + std::to_string(c));
    if (!(pb->delta_x[c] > 0.0) || !std::isfinite(pb->delta_x[c]))
      throw std::invalid_argument("delta_x must be positive at cell " + std::to_string(c));
  }
  if ((int64_t)n * G > (1 << 30)) throw std::invalid_argument("problem too large");
  const int GM = G * M;
  std::vector<double> ist(GM);
  for (int i = 0; i < GM; i++)
    ist[i] = pb->inv_sigtr ? pb->inv_sigtr[i] : 1.0 / (pb->sigt[i] - pb->mu[i] * pb->sigs[i]);

  Host h;
  h.pb = pb;
  DiffProblem &d = h.d;
  std::memset(&d, 0, sizeof d);
  d.n = n;
  d.G = G;
  d.M = M;
  d.max_it = o.max_iterations > 0 ? o.max_iterations : 1000000;
  d.ftol = o.flux_tol > 0 ? o.flux_tol : 1e-5;
  d.ktol = o.k_tol > 0 ? o.k_tol : 1e-6;

  // group matrices (matrix_gen, src/discrete.rs:4-101) and their factors,
  // in cell order (the thread-major layout is made at upload)
  h.alpha.assign((size_t)G * n, 0.0);
  h.mdiag.assign((size_t)G * n, 0.0);
  h.gamma.assign((size_t)G * n, 0.0);
  std::vector<double> sub(n), diag(n), sup(n);
  for (int g = 0; g < G; g++) {
    const double d0 = d_over_dx(pb, ist, 0, g), d1 = d_over_dx(pb, ist, 1, g);
    const double h01 = coupling(d0, d1);
    for (int c = 0; c < n; c++) {
      const int m = pb->matid[c];
      const double rem =
          pb->delta_x[c] * (pb->sigt[m + M * g] - pb->scat_matrix[G * G * m + (G + 1) * g]);
      const double dc = d_over_dx(pb, ist, c, g);
      const double hl = c > 0 ? coupling(dc, d_over_dx(pb, ist, c - 1, g)) : 0.0;
      const double hr = c < n - 1 ? coupling(dc, d_over_dx(pb, ist, c + 1, g)) : 0.0;
      sub[c] = -hl;
      sup[c] = -hr;
      if (c == 0)
        diag[c] = 2.0 * dc * (1.0 - beta_of(pb->boundl, d1, dc)) + rem + h01;
      else if (c == n - 1)
        diag[c] = 2.0 * dc * (1.0 - beta_of(pb->boundr, d1, dc)) + rem + h01;
      else
        diag[c] = hl + rem + hr;
    }
    double wprev = 0.0;
    for (int c = 0; c < n; c++) {
      const double m = 1.0 / (diag[c] - sub[c] * wprev);
      const double w = sup[c] * m;
      if (!std::isfinite(m) || !std::isfinite(w))
        throw std::invalid_argument("diffusion matrix of group " + std::to_string(g) +
                                    " is singular");
      const size_t at = (size_t)g * n + c;
      h.alpha[at] = -sub[c] * m;
      h.mdiag[at] = m;
      h.gamma[at] = -w;
      wprev = w;
    }
  }
  h.nsf.resize(GM);
  h.chi.assign(pb->chit, pb->chit + GM);
  h.nut.assign(pb->nut, pb->nut + GM);
  h.sigf.assign(pb->sigf, pb->sigf + GM);
  for (int i = 0; i < GM; i++) h.nsf[i] = pb->nut[i] * pb->sigf[i];

  const size_t tab_bytes = ((size_t)4 * GM + (size_t)G * G * M + kScratch) * sizeof(double);
  if (tab_bytes > (size_t)lds_limit) throw std::invalid_argument("diffusion tables exceed LDS");
  lds_limit_ = lds_limit;
  host_.push_back(std::move(h));
}

// Workgroup size of one problem: 1024 threads (4 waves per SIMD, 1-4 cells
// per thread) for meshes of 513..4096 cells with at most 8 groups, else 256.
// It depends on the problem alone, so a result does not depend on the batch
// it is solved in.
static int pick_threads(int n, int G) { return (G <= 8 && n > 512 && n <= 4096) ? 1024 : 256; }

void DiffusionSet::layout() {
  const int np = (int)host_.size();
  launches_.clear();
  slot_.assign(np, 0);
  for (int bs : {256, 1024}) {
    Launch ln;
    ln.bs = bs;
    ln.first = launches_.empty() ? 0 : launches_.back().first + launches_.back().count;
    for (int i = 0; i < np; i++) {
      Host &h = host_[i];
      DiffProblem &d = h.d;
      if (pick_threads(d.n, d.G) != bs) continue;
      slot_[i] = ln.first + ln.count++;
      const int n = d.n, G = d.G;
      const int L = (n + bs - 1) / bs, LT = L * bs;
      auto pos = [&](int c) { return (c % L) * bs + c / L; };
      d.L = L;
      d.bs = bs;
      d.pos0 = pos(0);
      d.pos1 = pos(n - 2);
      d.pos2 = pos(n - 1);
      d.own0 = 0;
      d.own1 = (n - 2) / L;
      d.own2 = (n - 1) / L;
      h.p_alpha.assign((size_t)G * LT, 0.0);
      h.p_mdiag.assign((size_t)G * LT, 0.0);
      h.p_gamma.assign((size_t)G * LT, 0.0);
      for (int g = 0; g < G; g++)
        for (int c = 0; c < n; c++) {
          h.p_alpha[(size_t)g * LT + pos(c)] = h.alpha[(size_t)g * n + c];
          h.p_mdiag[(size_t)g * LT + pos(c)] = h.mdiag[(size_t)g * n + c];
          h.p_gamma[(size_t)g * LT + pos(c)] = h.gamma[(size_t)g * n + c];
        }
      h.mat.assign(LT, 0);
      h.dx.assign(LT, 0.0);
      for (int c = 0; c < n; c++) {
        h.mat[pos(c)] = h.pb->matid[c];
        h.dx[pos(c)] = h.pb->delta_x[c];
      }
      const size_t tab_bytes =
          ((size_t)4 * G * d.M + (size_t)G * G * d.M + kScratch) * sizeof(double);
      const size_t dyn_bytes = ((size_t)G + 1) * LT * sizeof(double);
      d.in_lds = tab_bytes + dyn_bytes <= (size_t)lds_limit_;
      ln.lds_generic = std::max(ln.lds_generic, tab_bytes + (d.in_lds ? dyn_bytes : 0));
      ln.tab_bytes = std::max(ln.tab_bytes, tab_bytes);
      ln.max_G = std::max(ln.max_G, G);
      ln.max_L = std::max(ln.max_L, L);
    }
    launches_.push_back(ln);
  }
}

void DiffusionSet::upload(DevArena &arena, hipStream_t s) {
  layout();
  const int np = (int)host_.size();
  // byte layout: [DiffProblem x np][per problem: uploaded arrays][outputs, fallbacks]
  size_t at = 0;
  auto put = [&](size_t bytes) {
    size_t o = at;
    at += (bytes + 255) & ~(size_t)255;
    return o;
  };
  struct Off {
    size_t alpha, mdiag, gamma, mat, dx, dxn, left, right, nsf, chi, nut, sigf, scat;
    size_t gflux, gF, flux, fs, scal;
  };
  std::vector<Off> off(np);
  auto off_outputs = [](const std::vector<Off> &o) {
    std::vector<std::array<size_t, 3>> v(o.size());
    for (size_t i = 0; i < o.size(); i++) v[i] = {o[i].flux, o[i].fs, o[i].scal};
    return v;
  };
  const size_t probs = put(sizeof(DiffProblem) * np);
  for (int i = 0; i < np; i++) {
    const Host &h = host_[i];
    const size_t n = h.d.n, G = h.d.G, GM = G * h.d.M, LT = (size_t)h.d.L * h.d.bs;
    Off &f = off[i];
    f.alpha = put(G * LT * 8);
    f.mdiag = put(G * LT * 8);
    f.gamma = put(G * LT * 8);
    f.mat = put(LT * 4);
    f.dx = put(LT * 8);
    f.dxn = put(n * 8);
    f.left = put(n * 8);
    f.right = put(n * 8);
    f.nsf = put(GM * 8);
    f.chi = put(GM * 8);
    f.nut = put(GM * 8);
    f.sigf = put(GM * 8);
    f.scat = put(GM * G * 8);
  }
  const size_t up_end = at;
  for (int i = 0; i < np; i++) {
    const Host &h = host_[i];
    const size_t G = h.d.G, LT = (size_t)h.d.L * h.d.bs;
    Off &f = off[i];
    f.gflux = h.d.in_lds ? 0 : put(G * LT * 8);
    f.gF = h.d.in_lds ? 0 : put(LT * 8);
  }
  // outputs of every problem in one range: one copy back for the set
  out_begin_ = at;
  for (int i = 0; i < np; i++) {
    const size_t n = host_[i].d.n, G = host_[i].d.G;
    Off &f = off[i];
    f.flux = put(G * n * 8);
    f.fs = put(n * 8);
    f.scal = put(8 * 8);
  }
  out_end_ = at;
  out_off_ = off_outputs(off);
  uint8_t *base = arena.alloc<uint8_t>(at);
  d_base_ = base;
  std::vector<uint8_t> stage(up_end, 0);
  h_probs_.resize(np);
  for (int i = 0; i < np; i++) {
    const Host &h = host_[i];
    const nraps_problem *pb = h.pb;
    const size_t n = h.d.n, G = h.d.G, GM = G * h.d.M;
    const Off &f = off[i];
    auto cp = [&](size_t o, const void *src, size_t bytes) { std::memcpy(&stage[o], src, bytes); };
    cp(f.alpha, h.p_alpha.data(), h.p_alpha.size() * 8);
    cp(f.mdiag, h.p_mdiag.data(), h.p_mdiag.size() * 8);
    cp(f.gamma, h.p_gamma.data(), h.p_gamma.size() * 8);
    cp(f.mat, h.mat.data(), h.mat.size() * 4);
    cp(f.dx, h.dx.data(), h.dx.size() * 8);
    cp(f.dxn, pb->delta_x, n * 8);
    cp(f.left, pb->mesh_left, n * 8);
    cp(f.right, pb->mesh_right, n * 8);
    cp(f.nsf, h.nsf.data(), GM * 8);
    cp(f.chi, h.chi.data(), GM * 8);
    cp(f.nut, h.nut.data(), GM * 8);
    cp(f.sigf, h.sigf.data(), GM * 8);
    cp(f.scat, pb->scat_matrix, GM * G * 8);
    DiffProblem d = h.d;
    d.alpha = reinterpret_cast<const double *>(base + f.alpha);
    d.mdiag = reinterpret_cast<const double *>(base + f.mdiag);
    d.gamma = reinterpret_cast<const double *>(base + f.gamma);
    d.mat = reinterpret_cast<const int32_t *>(base + f.mat);
    d.dx = reinterpret_cast<const double *>(base + f.dx);
    d.dx_nat = reinterpret_cast<const double *>(base + f.dxn);
    d.left = reinterpret_cast<const double *>(base + f.left);
    d.right = reinterpret_cast<const double *>(base + f.right);
    d.nsf = reinterpret_cast<const double *>(base + f.nsf);
    d.chi = reinterpret_cast<const double *>(base + f.chi);
    d.nut = reinterpret_cast<const double *>(base + f.nut);
    d.sigf = reinterpret_cast<const double *>(base + f.sigf);
    d.scat = reinterpret_cast<const double *>(base + f.scat);
    d.gflux = h.d.in_lds ? nullptr : reinterpret_cast<double *>(base + f.gflux);
    d.gF = h.d.in_lds ? nullptr : reinterpret_cast<double *>(base + f.gF);
    d.flux_out = reinterpret_cast<double *>(base + f.flux);
    d.fs_out = reinterpret_cast<double *>(base + f.fs);
    d.scal = reinterpret_cast<double *>(base + f.scal);
    h_probs_[i] = d;
  }
  // device slots are grouped by workgroup size (one launch per size)
  for (int i = 0; i < np; i++)
    std::memcpy(&stage[probs + sizeof(DiffProblem) * slot_[i]], &h_probs_[i], sizeof(DiffProblem));
  d_probs_ = reinterpret_cast<DiffProblem *>(base + probs);
  if (s) {
    // the staging vector must outlive the copy: keep it until the set dies
    stage_.swap(stage);
    HIP_TRY(hipMemcpyAsync(base, stage_.data(), up_end, hipMemcpyHostToDevice, s));
  } else {
    HIP_TRY(hipMemcpy(base, stage.data(), up_end, hipMemcpyHostToDevice));
  }
}

namespace {
using DiffKernel = void (*)(const DiffProblem *);

int regs_lm(int L) { return L <= 2 ? 2 : L <= 4 ? 4 : 8; }

// Register-resident instantiation for the set's workgroup size, largest G
// (rounded up to 2, 4 or 8) and cells per thread, or nullptr.
DiffKernel regs_kernel(int bs, int G, int L) {
  if (G > 8 || (bs == 256 && L > 2) || (bs == 1024 && L > 4)) return nullptr;
  const int gm = G <= 2 ? 2 : G <= 4 ? 4 : 8, lm = regs_lm(L);
#define NRAPS_DK(a, b, c) \
  if (gm == a && lm == b && bs == c) return &k_diffusion_regs<a, b, c>;
  NRAPS_DK(2, 2, 256) NRAPS_DK(4, 2, 256) NRAPS_DK(8, 2, 256)
  NRAPS_DK(2, 2, 1024) NRAPS_DK(4, 2, 1024) NRAPS_DK(8, 2, 1024)
  NRAPS_DK(2, 4, 1024) NRAPS_DK(4, 4, 1024) NRAPS_DK(8, 4, 1024)
#undef NRAPS_DK
  return nullptr;
}
}  // namespace

void DiffusionSet::launch(hipStream_t s) {
  // NRAPS_DIFFUSION_GENERIC=1 forces the LDS/global-memory form (A/B checks)
  const char *env = std::getenv("NRAPS_DIFFUSION_GENERIC");
  const bool generic = env && std::atoi(env) != 0;
  for (const Launch &ln : launches_) {
    if (ln.count == 0) continue;
    DiffKernel kr = generic ? nullptr : regs_kernel(ln.bs, ln.max_G, ln.max_L);
    DiffKernel kern = kr ? kr : ln.bs == 1024 ? &k_diffusion<1024> : &k_diffusion<256>;
    const size_t lds = kr ? ln.tab_bytes + (size_t)regs_lm(ln.max_L) * ln.bs * 20 : ln.lds_generic;
    if (lds > 65536)
      HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(ln.count), dim3(ln.bs), lds, s, d_probs_ + ln.first);
    HIP_TRY(hipGetLastError());
  }
}

void DiffusionSet::fetch(hipStream_t s) {
  out_.resize(out_end_ - out_begin_);
  HIP_TRY(hipMemcpyAsync(out_.data(), d_base_ + out_begin_, out_.size(), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
}

int DiffusionSet::download(int i, nraps_diffusion_result *r) {
  const DiffProblem &d = h_probs_[i];
  const size_t n = d.n, G = d.G;
  auto at = [&](size_t off) { return reinterpret_cast<const double *>(&out_[off - out_begin_]); };
  const double *flux = at(out_off_[i][0]), *fs = at(out_off_[i][1]), *scal = at(out_off_[i][2]);
  const int status = (int)scal[4];
  if (!r) return status;
  if (r->flux) std::memcpy(r->flux, flux, G * n * 8);
  if (r->fission_source) std::memcpy(r->fission_source, fs, n * 8);
  if (r->assembly_average) {
    // average_assembly, src/discrete.rs:151-167
    std::fill(r->assembly_average, r->assembly_average + G * n, 0.0);
    const nraps_problem *pb = host_[i].pb;
    const int na = pb->numass > 0 ? pb->numass : 1;
    const size_t per = n / na;
    for (size_t g = 0; g < G && per > 0; g++)
      for (int a = 0; a < na; a++) {
        double sum = 0.0;
        for (size_t c = a * per; c < (a + 1) * per; c++) sum += flux[g * n + c];
        const double avg = sum / (double)per;
        for (size_t c = a * per; c < (a + 1) * per; c++) r->assembly_average[g * n + c] = avg;
      }
  }
  r->k = scal[0];
  r->delta_flux = scal[1];
  r->delta_k = scal[2];
  r->iterations = (int32_t)scal[3];
  r->status = status;
  return status;
}

void launch_diffusion_bank(bool fp32, const DiffProblem *dp, int64_t N, void *src_x,
                           int32_t *src_cell, CycleState *cs, double *cdf, hipStream_t s) {
  hipLaunchKernelGGL(k_dbank_cdf, dim3(1), dim3(64), 0, s, dp, cdf, cs);
  HIP_TRY(hipGetLastError());
  const int bs = 256;
  const int64_t grid = (N + bs - 1) / bs;
  if (fp32)
    hipLaunchKernelGGL(k_dbank_sites<float>, dim3((unsigned)grid), dim3(bs), 0, s, dp, cdf, N,
                       static_cast<float *>(src_x), src_cell);
  else
    hipLaunchKernelGGL(k_dbank_sites<double>, dim3((unsigned)grid), dim3(bs), 0, s, dp, cdf, N,
                       static_cast<double *>(src_x), src_cell);
  HIP_TRY(hipGetLastError());
}

namespace {

int diffusion_batch(int n, const nraps_problem *const *pbs, const nraps_diffusion_opts *opts,
                    nraps_diffusion_result *const *results) {
  const auto t0 = std::chrono::steady_clock::now();
  try {
    if (n < 1 || !pbs) {
      set_error("empty batch");
      return NRAPS_E_INVALID;
    }
    nraps_diffusion_opts o;
    nraps_diffusion_opts_default(&o);
    if (opts) o = *opts;
    if (o.max_iterations < 0 || o.flux_tol < 0 || o.k_tol < 0) {
      set_error("max_iterations and tolerances must be >= 0");
      return NRAPS_E_INVALID;
    }
    // problems are checked before the device, so bad input is reported
    // as such on any machine
    int ndev = 0;
    const bool have = hipGetDeviceCount(&ndev) == hipSuccess && ndev >= 1;
    const bool in_range = have && o.device >= 0 && o.device < ndev;
    int lds_limit = 160 * 1024;
    if (in_range) {
      HIP_TRY(hipSetDevice(o.device));
      HIP_TRY(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, o.device));
    }
    DiffusionSet set;
    for (int i = 0; i < n; i++) {
      try {
        set.add(pbs[i], o, lds_limit);
      } catch (const std::invalid_argument &e) {
        set_error(n > 1 ? "problem " + std::to_string(i) + ": " + e.what() : std::string(e.what()));
        return NRAPS_E_INVALID;
      }
    }
    if (!have) {
      set_error("no HIP device available");
      return NRAPS_E_DEVICE;
    }
    if (!in_range) {
      set_error("device ordinal out of range");
      return NRAPS_E_DEVICE;
    }
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct SG {
      hipStream_t s;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      ~SG() {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamDestroy(s);
      }
    } sg{s};
    HIP_TRY(hipEventCreate(&sg.e0));
    HIP_TRY(hipEventCreate(&sg.e1));
    int worst = NRAPS_OK;
    {
      DevArena arena;
      set.upload(arena, s);
      HIP_TRY(hipEventRecord(sg.e0, s));
      set.launch(s);
      HIP_TRY(hipEventRecord(sg.e1, s));
      HIP_TRY(hipStreamSynchronize(s));
      float ms_dev = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms_dev, sg.e0, sg.e1));
      set.fetch(s);
      for (int i = 0; i < n; i++) {
        nraps_diffusion_result *r = results ? results[i] : nullptr;
        const int st = set.download(i, r);
        if (st != NRAPS_OK && worst == NRAPS_OK) worst = st;
        if (r) r->ms_device = ms_dev;
      }
    }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; i++)
      if (results && results[i]) results[i]->ms_total = ms;
    if (worst == NRAPS_E_NOT_CONVERGED) set_error("iteration limit reached before convergence");
    return worst;
  } catch (const DeviceError &e) {
    set_error(e.what());
    return e.err == hipErrorOutOfMemory ? NRAPS_E_NOMEM : NRAPS_E_DEVICE;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed");
    return NRAPS_E_NOMEM;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_DEVICE;
  }
}

}  // namespace
}  // namespace nraps

using namespace nraps;

extern "C" void nraps_diffusion_opts_default(nraps_diffusion_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->max_iterations = 1000000;
  o->flux_tol = 1e-5;
  o->k_tol = 1e-6;
}

extern "C" int nraps_diffusion_solve(const nraps_problem *problem, const nraps_diffusion_opts *opts,
                                     nraps_diffusion_result *result) {
  const nraps_problem *pbs[1] = {problem};
  nraps_diffusion_result *rs[1] = {result};
  const int st = diffusion_batch(1, pbs, opts, rs);
  if (result) result->status = st;
  return st;
}

extern "C" int nraps_diffusion_solve_batch(int32_t n, const nraps_problem *const *problems,
                                           const nraps_diffusion_opts *opts,
                                           nraps_diffusion_result *const *results) {
  return diffusion_batch(n, problems, opts, results);
}
