// host.cpp — validation, derived tables, synthetic workloads, the input-deck
// reader, mesh expansion and the CSV writers of the MI355X solver.
#include "host.h"

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <memory>
#include <stdexcept>
#include <sstream>

#include "nraps_core.h"

namespace nraps {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const std::string &last_error() { return g_err; }

// ---------------------------------------------------------------------------
// Validation
// ---------------------------------------------------------------------------
int validate(const nraps_problem *pb, const nraps_run_opts *o, std::string *why) {
  auto bad = [&](const std::string &m) {
    if (why) *why = m;
    return NRAPS_E_INVALID;
  };
  if (!pb) return bad("problem is NULL");
  const int nc = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  if (nc < 1 || nc > kMaxCells) return bad("n_cells must be in [1, 65535]");
  if (G < 1 || G > kMaxGroups) return bad("n_groups must be in [1, 32]");
  if (M < 1 || M > kMaxMats) return bad("n_mats must be in [1, 255]");
  if (!pb->matid || !pb->delta_x || !pb->mesh_left || !pb->mesh_right)
    return bad("mesh arrays must not be NULL");
  if (!pb->sigt || !pb->sigs || !pb->mu || !pb->siga || !pb->sigf || !pb->nut || !pb->chit ||
      !pb->scat_matrix)
    return bad("cross-section arrays must not be NULL");
  if (!(pb->boundl <= 1.0) || !(pb->boundr <= 1.0))
    return bad("boundary albedo must be <= 1 (> 0 reflects, <= 0 leaks)");
  for (int c = 0; c < nc; c++) {
    if (pb->matid[c] < 0 || pb->matid[c] >= M) return bad("matid out of range at cell " + std::to_string(c));
    if (!(pb->mesh_right[c] > pb->mesh_left[c]) || !std::isfinite(pb->mesh_right[c]) ||
        !std::isfinite(pb->mesh_left[c]))
      return bad("cell " + std::to_string(c) + " has non-positive width");
    if (!(pb->delta_x[c] > 0.0) || !std::isfinite(pb->delta_x[c]))
      return bad("delta_x must be positive at cell " + std::to_string(c));
    if (c + 1 < nc && pb->mesh_right[c] != pb->mesh_left[c + 1])
      return bad("mesh_right[c] must equal mesh_left[c+1] at cell " + std::to_string(c));
  }
  std::vector<char> present(M, 0);
  for (int c = 0; c < nc; c++) present[pb->matid[c]] = 1;
  for (int m = 0; m < M; m++) {
    if (!present[m]) continue;
    for (int g = 0; g < G; g++) {
      int i = m + M * g;
      double itr = pb->inv_sigtr ? pb->inv_sigtr[i] : 1.0 / (pb->sigt[i] - pb->mu[i] * pb->sigs[i]);
      if (!(itr > 0.0) || !std::isfinite(itr))
        return bad("transport cross section must be positive (mat " + std::to_string(m) + ", group " +
                   std::to_string(g) + ")");
      if (!(pb->sigt[i] > 0.0)) return bad("sigt must be positive for materials in the mesh");
      if (!(pb->sigs[i] >= 0.0) || !(pb->siga[i] >= 0.0) || !(pb->nut[i] >= 0.0) ||
          !(pb->sigf[i] >= 0.0))
        return bad("negative cross section");
    }
  }
  if (pb->fuel_indices) {
    if (pb->n_fuel < 1) return bad("n_fuel must be >= 1");
    for (int i = 0; i < pb->n_fuel; i++)
      if (pb->fuel_indices[i] < 0 || pb->fuel_indices[i] >= nc) return bad("fuel index out of range");
  } else {
    int nf = 0;
    for (int c = 0; c < nc; c++) nf += (pb->matid[c] == 0 || pb->matid[c] == 1);
    if (nf < 1) return bad("no fuel cells (matid 0 or 1) for the source");
  }
  if (!o) return NRAPS_OK;
  if (o->histories < 1 || o->histories > NRAPS_MAX_HISTORIES)
    return bad("histories must be in [1, 2^28 - 2]");
  if (o->generations < 2 || o->generations > 65535) return bad("generations must be in [2, 65535]");
  if (o->skip < 1 || o->skip >= o->generations)
    return bad("skip must satisfy 1 <= skip < generations");
  if (o->precision != NRAPS_FP64 && o->precision != NRAPS_FP32) return bad("unknown precision");
  if (o->method != NRAPS_HISTORY_SURFACE && o->method != NRAPS_EVENT_WOODCOCK &&
      o->method != NRAPS_EVENT_REGION)
    return bad("unknown method");
  if (o->method == NRAPS_HISTORY_SURFACE && o->precision != NRAPS_FP64)
    return bad("the plain surface-tracking port runs in FP64 only");
  if (o->source != NRAPS_SOURCE_UNIFORM_FUEL && o->source != NRAPS_SOURCE_FISSION_BANK)
    return bad("unknown source mode");
  if (o->xs_refresh != NRAPS_XS_ON_MATERIAL_CHANGE && o->xs_refresh != NRAPS_XS_EVERY_COLLISION)
    return bad("unknown xs_refresh");
  if (o->bank_init != NRAPS_BANK_INIT_UNIFORM_FUEL && o->bank_init != NRAPS_BANK_INIT_DIFFUSION)
    return bad("unknown bank_init");
  if (o->bank_init == NRAPS_BANK_INIT_DIFFUSION && o->source != NRAPS_SOURCE_FISSION_BANK)
    return bad("bank_init = diffusion needs the fission-bank source");
  if (o->bank_init == NRAPS_BANK_INIT_DIFFUSION && nc < 2)
    return bad("bank_init = diffusion needs at least 2 cells");
  if (o->source == NRAPS_SOURCE_FISSION_BANK && o->bank_init == NRAPS_BANK_INIT_UNIFORM_FUEL &&
      (o->skip < 2 || o->generations < 3))
    return bad("fission-bank mode needs skip >= 2 (two uniform-source start-up generations)");
  if (o->problem_id >= (1ull << 19)) return bad("problem_id must be < 2^19");
  if (!(o->k_initial > 0.0) || !std::isfinite(o->k_initial)) return bad("k_initial must be positive");
  if (o->max_events < 0) return bad("max_events must be >= 0");
  if (o->trace_histories < 0 || (o->trace_histories > 0 && o->trace_events < 1))
    return bad("trace_events must be >= 1 when tracing");
  if (o->trace_histories > o->histories) return bad("trace_histories must be <= histories");
  if (o->trace_histories > 0 && (o->trace_cycle < 0 || o->trace_cycle >= o->generations))
    return bad("trace_cycle out of range");
  if (o->chunk < 0) return bad("chunk must be >= 0");
  if (o->source == NRAPS_SOURCE_FISSION_BANK) {
    bool any = false;
    for (int c = 0; c < nc && !any; c++)
      for (int g = 0; g < G; g++) {
        int i = pb->matid[c] + M * g;
        if (pb->nut[i] * pb->sigf[i] > 0.0) any = true;
      }
    if (!any) return bad("fission-bank mode needs a fissile material in the mesh");
  }
  return NRAPS_OK;
}

// ---------------------------------------------------------------------------
// Derived tables (same formulas as the reference: src/process_input.rs:152-156,
// src/mc_code.rs:6-32, :81-111)
// ---------------------------------------------------------------------------
void Prepared::build(const nraps_problem *pb) {
  nc = pb->n_cells;
  G = pb->n_groups;
  M = pb->n_mats;
  numass = pb->numass > 0 ? pb->numass : 1;
  boundl = pb->boundl;
  boundr = pb->boundr;
  XL = pb->mesh_left[0];
  XR = pb->mesh_right[nc - 1];
  edges.resize(nc + 1);
  for (int c = 0; c < nc; c++) edges[c] = pb->mesh_left[c];
  edges[nc] = XR;
  dx.assign(pb->delta_x, pb->delta_x + nc);
  mat.assign(pb->matid, pb->matid + nc);
  fuel.clear();
  if (pb->fuel_indices)
    fuel.assign(pb->fuel_indices, pb->fuel_indices + pb->n_fuel);
  else
    for (int c = 0; c < nc; c++)
      if (mat[c] == 0 || mat[c] == 1) fuel.push_back(c);
  nf = (int)fuel.size();
  const int GM = G * M;
  inv_sigtr.resize(GM);
  sigw.resize(GM);
  abs_ratio.resize(GM);
  inv_sigs.resize(GM);
  nusigf.resize(GM);
  for (int i = 0; i < GM; i++) {
    inv_sigtr[i] = pb->inv_sigtr ? pb->inv_sigtr[i] : 1.0 / (pb->sigt[i] - pb->mu[i] * pb->sigs[i]);
    sigw[i] = 1.0 / inv_sigtr[i];
    abs_ratio[i] = pb->siga[i] / pb->sigt[i];
    inv_sigs[i] = 1.0 / pb->sigs[i];
    nusigf[i] = pb->nut[i] * pb->sigf[i];
  }
  chicdf.resize((size_t)M * G);
  cumrow.resize((size_t)M * G * G);
  for (int m = 0; m < M; m++) {
    double cum = 0.0;
    for (int g = 0; g < G; g++) {
      cum += pb->chit[m + M * g];
      chicdf[(size_t)m * G + g] = cum;
    }
    for (int g = 0; g < G; g++) {
      double c2 = 0.0;
      for (int j = 0; j < G; j++) {
        c2 += pb->scat_matrix[(size_t)G * G * m + (size_t)G * g + j];
        cumrow[((size_t)m * G + g) * G + j] = c2;
      }
    }
  }
  maj = 0.0;
  for (int c = 0; c < nc; c++)
    for (int g = 0; g < G; g++) maj = std::max(maj, sigw[mat[c] + M * g]);
  region.resize(nc);
  R = 0;
  for (int c = 0; c < nc; c++) {
    if (c > 0 && mat[c] != mat[c - 1]) R++;
    region[c] = R;
  }
  R++;
  reg_lo.assign(R, 0.0);
  reg_mat.assign(R, 0);
  reg_cell.assign(R + 1, nc);
  std::vector<double> reg_hi(R, 0.0);
  for (int c = 0; c < nc; c++) {
    int r = region[c];
    if (c == 0 || region[c - 1] != r) {
      reg_cell[r] = c;
      reg_lo[r] = pb->mesh_left[c];
      reg_mat[r] = mat[c];
    }
    reg_hi[r] = pb->mesh_right[c];
  }
  kreg.assign((size_t)G * (R + 1), 0.0);
  for (int g = 0; g < G; g++) {
    double acc = 0.0;
    for (int r = 0; r < R; r++) {
      acc = acc + nusigf[reg_mat[r] + M * g] * (reg_hi[r] - reg_lo[r]);
      kreg[(size_t)g * (R + 1) + r + 1] = acc;
    }
  }
  nut_conv = pb->nut[0 + M * (G > 1 ? 1 : 0)];
  // cell lookup buckets of width h (twice the narrowest cell, at least
  // L / kMaxBuckets): bucket[b] = largest c with edges[c] < XL + b*h, so a
  // point x is in cell bucket[floor((x - XL) / h)] or a few cells to its right
  double hmin = XR - XL;
  for (int c = 0; c < nc; c++) hmin = std::min(hmin, edges[c + 1] - edges[c]);
  double h = std::max(2.0 * hmin, (XR - XL) / (kMaxBuckets - 2));
  nb = (int)std::ceil((XR - XL) / h) + 2;
  nb = std::min(nb, kMaxBuckets);
  inv_h = 1.0 / h;
  bucket.assign(nb, 0);
  int c = 0;
  for (int b = 0; b < nb; b++) {
    double xb = XL + b * h;
    while (c + 1 < nc && edges[c + 1] < xb) c++;
    bucket[b] = (uint16_t)c;
  }
}

TabLayout Prepared::layout(bool fp32) const {
  const int rs = fp32 ? 4 : 8;
  TabLayout L{};
  int32_t at = 0;
  auto put = [&](int32_t bytes) {
    int32_t o = at;
    at += (bytes + 15) & ~15;
    return o;
  };
  const int GM = G * M;
  L.edges = put(rs * (nc + 1));
  L.mat8 = put(nc);
  L.reg16 = put(2 * nc);
  L.sigw = put(rs * GM);
  L.invtr = put(rs * GM);
  L.absr = put(rs * GM);
  L.invs = put(rs * GM);
  L.nsf = put(rs * GM);
  L.chicdf = put(rs * GM);
  L.cumrow = put(rs * M * G * G);
  L.nsfd = put(8 * GM);
  L.kreg = put(8 * G * (R + 1));
  L.bucket = put(2 * nb);
  L.regedge = put(8 * (R + 1));
  L.regcell = put(2 * (R + 1));
  // the birth tables join the LDS-staged part when they still leave the
  // tables + tally within a quarter of the LDS (several 256-thread blocks
  // per CU); larger problems read them from global memory
  const int64_t birth = ((4 * (int64_t)nf + 15) & ~15) + ((rs * (int64_t)nc + 15) & ~15);
  const int64_t tally = 2 * (int64_t)G * nc * rs;
  if (kBirthInLds && at + birth + tally <= kLdsBytes / 4) {
    L.fuel = put(4 * nf);
    L.dx = put(rs * nc);
    L.hot = at;
  } else {
    L.hot = at;
    L.fuel = put(4 * nf);
    L.dx = put(rs * nc);
  }
  L.bytes = at;
  return L;
}

std::vector<uint8_t> Prepared::pack(bool fp32, TabLayout *lay) const {
  TabLayout L = layout(fp32);
  std::vector<uint8_t> blob(L.bytes, 0);
  auto putr = [&](int32_t off, const std::vector<double> &v) {
    if (fp32) {
      for (size_t i = 0; i < v.size(); i++) {
        float f = (float)v[i];
        std::memcpy(&blob[off + 4 * i], &f, 4);
      }
    } else {
      std::memcpy(&blob[off], v.data(), 8 * v.size());
    }
  };
  auto putd = [&](int32_t off, const std::vector<double> &v) {
    std::memcpy(&blob[off], v.data(), 8 * v.size());
  };
  putr(L.edges, edges);
  for (int c = 0; c < nc; c++) {
    blob[L.mat8 + c] = (uint8_t)mat[c];
    uint16_t r = (uint16_t)region[c];
    std::memcpy(&blob[L.reg16 + 2 * c], &r, 2);
  }
  std::memcpy(&blob[L.fuel], fuel.data(), 4 * fuel.size());
  putr(L.dx, dx);
  putr(L.sigw, sigw);
  putr(L.invtr, inv_sigtr);
  putr(L.absr, abs_ratio);
  putr(L.invs, inv_sigs);
  putr(L.nsf, nusigf);
  putr(L.chicdf, chicdf);
  putr(L.cumrow, cumrow);
  putd(L.nsfd, nusigf);
  putd(L.kreg, kreg);
  std::memcpy(&blob[L.bucket], bucket.data(), 2 * bucket.size());
  std::vector<double> regedge(R + 1);
  for (int r = 0; r <= R; r++) {
    regedge[r] = edges[reg_cell[r]];
    uint16_t c = (uint16_t)reg_cell[r];
    std::memcpy(&blob[L.regcell + 2 * r], &c, 2);
  }
  putd(L.regedge, regedge);
  if (lay) *lay = L;
  return blob;
}

// ---------------------------------------------------------------------------
// Owned problem storage
// ---------------------------------------------------------------------------
void ProblemBuf::refresh_view() {
  view.matid = matid.data();
  view.delta_x = delta_x.data();
  view.mesh_left = left.data();
  view.mesh_right = right.data();
  view.n_fuel = (int32_t)fuel.size();
  view.fuel_indices = fuel.empty() ? nullptr : fuel.data();
  view.sigt = sigt.data();
  view.sigs = sigs.data();
  view.mu = mu.data();
  view.siga = siga.data();
  view.sigf = sigf.data();
  view.nut = nut.data();
  view.chit = chit.data();
  view.scat_matrix = scat.data();
  view.inv_sigtr = invtr.empty() ? nullptr : invtr.data();
}

void ProblemBuf::copy_from(const nraps_problem *pb) {
  const int nc = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  view = *pb;
  matid.assign(pb->matid, pb->matid + nc);
  delta_x.assign(pb->delta_x, pb->delta_x + nc);
  left.assign(pb->mesh_left, pb->mesh_left + nc);
  right.assign(pb->mesh_right, pb->mesh_right + nc);
  fuel.clear();
  if (pb->fuel_indices) fuel.assign(pb->fuel_indices, pb->fuel_indices + pb->n_fuel);
  auto cp = [&](std::vector<double> &d, const double *s, size_t n) { d.assign(s, s + n); };
  cp(sigt, pb->sigt, G * M);
  cp(sigs, pb->sigs, G * M);
  cp(mu, pb->mu, G * M);
  cp(siga, pb->siga, G * M);
  cp(sigf, pb->sigf, G * M);
  cp(nut, pb->nut, G * M);
  cp(chit, pb->chit, G * M);
  cp(scat, pb->scat_matrix, (size_t)M * G * G);
  invtr.clear();
  if (pb->inv_sigtr) cp(invtr, pb->inv_sigtr, G * M);
  refresh_view();
}

// ---------------------------------------------------------------------------
// Rust Display for f64
// ---------------------------------------------------------------------------
std::string rust_display(double v) {
  if (std::isnan(v)) return "NaN";
  if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
  if (v == 0.0) return std::signbit(v) ? "-0" : "0";
  char buf[64];
  auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::scientific);
  std::string sci(buf, r.ptr);  // shortest round-trip: d.ddddde[+-]xx
  bool neg = sci[0] == '-';
  if (neg) sci.erase(0, 1);
  size_t epos = sci.find('e');
  int exp10 = std::atoi(sci.c_str() + epos + 1);
  std::string digits;
  for (size_t i = 0; i < epos; i++)
    if (sci[i] != '.') digits.push_back(sci[i]);
  // value = 0.d1d2d3... * 10^(exp10+1)
  int point = exp10 + 1;  // digits before the decimal point
  std::string out;
  if (point <= 0) {
    out = "0." + std::string(-point, '0') + digits;
  } else if ((size_t)point >= digits.size()) {
    out = digits + std::string(point - digits.size(), '0');
  } else {
    out = digits.substr(0, point) + "." + digits.substr(point);
  }
  return neg ? "-" + out : out;
}

}  // namespace nraps

using namespace nraps;

// ---------------------------------------------------------------------------
// Synthetic workloads
// ---------------------------------------------------------------------------
namespace {

struct GenRng {
  nraps_pcg32 r;
  GenRng(uint64_t seed, uint64_t stream) { nraps_pcg32_seed(&r, seed, stream); }
  double u() { return nraps_u01_f64(nraps_pcg32_next(&r)); }
  double range(double a, double b) { return a + (b - a) * u(); }
  uint32_t index(uint32_t n) { return nraps_uindex(nraps_pcg32_next(&r), n); }
};

// Generator stream of (config, index): the top problem id is reserved.
uint64_t gen_stream(int config, int index) {
  return nraps_stream_id((1u << 19) - 1u, (uint64_t)config, (uint64_t)index);
}

// Cross sections of one material kind for G groups, appended at material
// slot m (reference layout).  Draw order per group: sigma_t, c, [nu*Sigma_f
// for fuel in the lower half of the groups], [upscatter fraction for the
// two lowest groups when upscatter is on], then G-g split weights.
void gen_material(GenRng &rng, ProblemBuf &b, int m, int M, int G, bool fuel, bool upscatter) {
  static const double chi0[2] = {0.75, 0.25};
  for (int g = 0; g < G; g++) {
    const int i = m + M * g;
    double st = fuel ? rng.range(0.2, 0.6) : rng.range(0.3, 2.0);
    double c = fuel ? rng.range(0.5, 0.9) : rng.range(0.95, 0.99);
    double nsf = (fuel && g >= G / 2) ? rng.range(0.005, 0.1) : 0.0;
    double up = (upscatter && g >= 1 && g >= G - 2) ? rng.range(0.0, 0.05) : 0.0;
    b.sigt[i] = st;
    b.sigs[i] = c * st;
    b.siga[i] = st - c * st;
    b.mu[i] = 0.0;
    b.nut[i] = fuel ? 2.5 : 0.0;
    b.sigf[i] = fuel ? nsf / 2.5 : 0.0;
    b.chit[i] = (fuel && g < 2) ? chi0[g] : 0.0;
    std::vector<double> w(G - g);
    double ws = 0.0;
    for (int j = 0; j < G - g; j++) {
      w[j] = rng.u();
      ws += w[j];
    }
    double *row = &b.scat[(size_t)G * G * m + (size_t)G * g];
    for (int j = 0; j < G; j++) row[j] = 0.0;
    if (up > 0.0) row[g - 1] = up * b.sigs[i];
    for (int j = 0; j < G - g; j++) row[g + j] = (1.0 - up) * b.sigs[i] * (w[j] / ws);
  }
}

void copy_material(ProblemBuf &b, int dst, int src, int M, int G) {
  for (int g = 0; g < G; g++) {
    int i = dst + M * g, j = src + M * g;
    b.sigt[i] = b.sigt[j];
    b.sigs[i] = b.sigs[j];
    b.mu[i] = b.mu[j];
    b.siga[i] = b.siga[j];
    b.sigf[i] = b.sigf[j];
    b.nut[i] = b.nut[j];
    b.chit[i] = b.chit[j];
  }
  for (int k = 0; k < G * G; k++) b.scat[(size_t)G * G * dst + k] = b.scat[(size_t)G * G * src + k];
}

void build_mesh(ProblemBuf &b, const std::vector<double> &widths, const std::vector<int> &mats,
                const std::vector<int> &cells) {
  b.matid.clear();
  b.delta_x.clear();
  b.left.clear();
  b.right.clear();
  double x = 0.0;
  for (size_t r = 0; r < widths.size(); r++) {
    double d = widths[r] / cells[r];
    for (int k = 0; k < cells[r]; k++) {
      b.matid.push_back(mats[r]);
      b.delta_x.push_back(d);
      b.left.push_back(x);
      x = x + d;
      b.right.push_back(x);
    }
  }
}

// Largest-remainder apportionment of `total` cells to regions by width,
// at least one cell each.
std::vector<int> apportion(const std::vector<double> &w, int total) {
  double W = 0.0;
  for (double v : w) W += v;
  std::vector<int> n(w.size());
  std::vector<std::pair<double, int>> rem;
  int used = 0;
  for (size_t i = 0; i < w.size(); i++) {
    double q = total * w[i] / W;
    n[i] = std::max(1, (int)std::floor(q));
    used += n[i];
    rem.push_back({q - std::floor(q), (int)i});
  }
  std::stable_sort(rem.begin(), rem.end(),
                   [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                     return a.first > b.first;
                   });
  for (size_t k = 0; used < total && k < rem.size(); k++, used++) n[rem[k].second]++;
  return n;
}

void alloc_xs(ProblemBuf &b, int G, int M) {
  b.sigt.assign(G * M, 0.0);
  b.sigs.assign(G * M, 0.0);
  b.mu.assign(G * M, 0.0);
  b.siga.assign(G * M, 0.0);
  b.sigf.assign(G * M, 0.0);
  b.nut.assign(G * M, 0.0);
  b.chit.assign(G * M, 0.0);
  b.scat.assign((size_t)M * G * G, 0.0);
}

}  // namespace

extern "C" int nraps_problem_generate(int32_t config, int32_t index, uint64_t seed,
                                      nraps_problem_buf **out, nraps_run_opts *opts) {
  try {
    if (!out) {
      set_error("out is NULL");
      return NRAPS_E_INVALID;
    }
    *out = nullptr;
    if (config < 0 || config > 4) {
      set_error("config must be 0..4");
      return NRAPS_E_NOT_FOUND;
    }
    if ((config >= 3) && (index < 0 || index >= (1 << 19) - 1)) {
      set_error("batch index out of range");
      return NRAPS_E_NOT_FOUND;
    }
    auto *b = new nraps_problem_buf();
    const int M = 3;  // {0 fuel, 1 copy of fuel, 2 water}
    nraps_run_opts o;
    nraps_run_opts_default(&o);
    o.seed = seed;
    if (config <= 1) {
      const int G = 4;
      GenRng rng(seed, gen_stream(0, 0));
      alloc_xs(*b, G, M);
      gen_material(rng, *b, 0, M, G, true, false);
      gen_material(rng, *b, 2, M, G, false, false);
      copy_material(*b, 1, 0, M, G);
      std::vector<double> w;
      std::vector<int> mt, n;
      for (int r = 0; r < 9; r++) {
        w.push_back(r % 2 == 0 ? 1.0 : 0.5);
        mt.push_back(r % 2 == 0 ? 0 : 2);
        n.push_back(20);
      }
      build_mesh(*b, w, mt, n);
      b->view.n_groups = G;
      if (config == 0) {
        o.histories = 10000;
        o.generations = 150;
        o.skip = 50;
      } else {
        o.histories = 10000000;
        o.generations = 120;
        o.skip = 20;
        o.method = NRAPS_EVENT_REGION;
        o.precision = NRAPS_FP32;
      }
    } else {
      int G, regions, cells_total = 0;
      int64_t N;
      GenRng rng(seed, gen_stream(config >= 3 ? 3 : 2, config >= 3 ? index : 0));
      if (config == 2) {
        G = 8;
        regions = 201;
        N = 100000000;
      } else {
        regions = 3 + (int)rng.index(199);
        static const int gs[3] = {2, 4, 8};
        G = gs[rng.index(3)];
        N = (int64_t)std::llround(std::pow(10.0, 4.0 + 2.0 * rng.u()));
      }
      std::vector<double> w(regions);
      std::vector<int> mt(regions);
      for (int r = 0; r < regions; r++) {
        w[r] = rng.range(0.4, 1.2);
        mt[r] = r % 2 == 0 ? 0 : 2;
      }
      std::vector<int> n;
      if (config == 2) {
        n = apportion(w, 2000);
      } else {
        n.assign(regions, 10);
      }
      for (int v : n) cells_total += v;
      (void)cells_total;
      alloc_xs(*b, G, M);
      gen_material(rng, *b, 0, M, G, true, G >= 4);
      gen_material(rng, *b, 2, M, G, false, G >= 4);
      copy_material(*b, 1, 0, M, G);
      build_mesh(*b, w, mt, n);
      b->view.n_groups = G;
      o.histories = N;
      o.generations = 120;
      o.skip = 20;
      o.method = NRAPS_EVENT_REGION;
      o.precision = NRAPS_FP32;
      o.source = NRAPS_SOURCE_FISSION_BANK;
      o.problem_id = config >= 3 ? (uint64_t)index : 0;
    }
    b->view.n_cells = (int32_t)b->matid.size();
    b->view.n_mats = M;
    b->view.numass = 1;
    b->view.boundl = 1.0;
    b->view.boundr = 0.0;
    b->refresh_view();
    *out = b;
    if (opts) *opts = o;
    return NRAPS_OK;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_NOMEM;
  }
}

extern "C" const nraps_problem *nraps_problem_view(const nraps_problem_buf *buf) {
  return buf ? &buf->view : nullptr;
}

extern "C" void nraps_problem_free(nraps_problem_buf *buf) { delete buf; }

extern "C" int nraps_validate(const nraps_problem *pb, const nraps_run_opts *o) {
  std::string why;
  int st = validate(pb, o, &why);
  if (st) set_error(why);
  return st;
}

extern "C" void nraps_run_opts_default(nraps_run_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof *o);
  o->histories = 10000;
  o->generations = 150;
  o->skip = 50;
  o->seed = 42;
  o->problem_id = 0;
  o->precision = NRAPS_FP64;
  o->method = NRAPS_HISTORY_SURFACE;
  o->source = NRAPS_SOURCE_UNIFORM_FUEL;
  o->xs_refresh = NRAPS_XS_ON_MATERIAL_CHANGE;
  o->k_initial = 1.0;
  o->max_events = 0;
  o->device = 0;
  o->use_graph = 1;
  o->chunk = 0;
}

extern "C" const char *nraps_last_error(void) { return last_error().c_str(); }

// ---------------------------------------------------------------------------
// Input deck (src/process_input.rs:1-175) and mesh expansion (src/main.rs:85-143)
// ---------------------------------------------------------------------------
namespace {

// Key bucket by (last two characters, length), src/process_input.rs:13-42.
int key_index(const std::string &k2, size_t len) {
  struct E {
    const char *k;
    size_t n;
    int i;
  };
  static const E tab[] = {{"lk", 5, 0},   {"es", 8, 1},   {"ps", 12, 2},  {"ns", 11, 3},
                          {"es", 9, 4},   {"ip", 4, 5},   {"ss", 6, 6},   {"ds", 7, 7},
                          {"ia", 6, 8},   {"ch", 8, 9},   {"fr", 4, 10},  {"wr", 4, 11},
                          {"dl", 6, 12},  {"dr", 6, 13},  {"gt", 4, 14},  {"gs", 4, 15},
                          {"mu", 2, 16},  {"ga", 4, 17},  {"gf", 4, 18},  {"ut", 3, 19},
                          {"it", 4, 20},  {"at", 4, 21},  {"id", 5, 22},  {"on", 8, 23},
                          {"er", 6, 24}};
  for (const E &e : tab)
    if (len == e.n && k2 == e.k) return e.i;
  return 25;
}

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) a++;
  while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
  return s.substr(a, b - a);
}

// Byte scanner of src/process_input.rs:44-83, including its handling of
// '#': the rest of the line is skipped and the first byte of the next line
// is stepped over.
void scan_deck(const std::string &buf, std::string temp[26]) {
  const size_t end = buf.size();
  size_t pos = 0, line_start = 0, name_end = 0, val_start = 0;
  while (pos < end) {
    char ch = buf[pos];
    if (ch == '#') {
      while (pos < end && buf[pos] != '\n') pos++;
      pos += 1;
      line_start = pos;
    } else if (ch == '=') {
      name_end = pos - 1;
      val_start = pos + 1;
    } else if (ch == '\n') {
      if (name_end > line_start) {
        std::string key = buf.substr(line_start, name_end - line_start);
        key = trim(key);
        for (auto &c : key) c = (char)std::tolower((unsigned char)c);
        std::string value = trim(buf.substr(val_start, pos - val_start));
        size_t n = key.size();
        if (n >= 2) temp[key_index(key.substr(n - 2), n)] += " " + value;
      }
      line_start = pos + 1;
    }
    pos += 1;
  }
}

bool parse_list(const std::string &s, std::vector<double> &out) {
  std::istringstream is(s);
  std::string tok;
  out.clear();
  while (is >> tok) {
    char *e = nullptr;
    errno = 0;
    double v = std::strtod(tok.c_str(), &e);
    if (e == tok.c_str() || *e != '\0' || errno) return false;
    out.push_back(v);
  }
  return true;
}

bool parse_int(const std::string &s, int64_t &v) {
  std::string t = trim(s);
  if (t.empty()) return false;
  char *e = nullptr;
  errno = 0;
  long long x = std::strtoll(t.c_str(), &e, 10);
  if (*e != '\0' || errno) return false;
  v = x;
  return true;
}

bool parse_real(const std::string &s, double &v) {
  std::vector<double> l;
  if (!parse_list(s, l) || l.size() != 1) return false;
  v = l[0];
  return true;
}

}  // namespace

extern "C" int nraps_deck_load(const char *path, nraps_problem_buf **out, nraps_deck_info *info) {
  try {
    if (!path || !out) {
      set_error("path/out is NULL");
      return NRAPS_E_INVALID;
    }
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f) {
      set_error(std::string("cannot read ") + path);
      return NRAPS_E_IO;
    }
    std::string buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    std::string t[26];
    scan_deck(buf, t);
    nraps_deck_info d{};
    int64_t iv;
    double rv;
    auto need_int = [&](int idx, const char *name, int64_t &dst) {
      if (!parse_int(t[idx], iv)) throw std::runtime_error(std::string("deck: bad or missing ") + name);
      dst = iv;
    };
    auto need_real = [&](int idx, const char *name, double &dst) {
      if (!parse_real(t[idx], rv)) throw std::runtime_error(std::string("deck: bad or missing ") + name);
      dst = rv;
    };
    int64_t a;
    need_int(0, "Analk", a);
    d.analk = (int32_t)a;
    need_int(1, "MatTypes", a);
    d.mattypes = (int32_t)a;
    need_int(2, "EnergyGroups", a);
    d.energygroups = (int32_t)a;
    need_int(3, "Generations", d.generations);
    need_int(4, "Histories", d.histories);
    need_int(5, "Skip", d.skip);
    need_int(6, "NumAss", a);
    d.numass = (int32_t)a;
    need_int(7, "NumRods", a);
    d.numrods = (int32_t)a;
    need_real(8, "RodDia", d.roddia);
    double pitch;
    need_real(9, "RodPitch", pitch);
    d.rodpitch = pitch - d.roddia;
    need_int(10, "MPFR", d.mpfr);
    need_int(11, "MPWR", d.mpwr);
    need_real(12, "BoundL", d.boundl);
    need_real(13, "BoundR", d.boundr);
    if (d.mpfr < 1 || d.mpwr < 0 || d.numass < 1) throw std::runtime_error("deck: bad mesh counts");
    d.delta_fuel = d.roddia / (double)d.mpfr;
    d.delta_water = d.mpwr > 0 ? d.rodpitch / (double)d.mpwr : 0.0;
    int64_t sol = 0, slv = 0;
    d.solution_mc = parse_int(t[23], sol) ? (sol == 1) : 0;
    d.solver = parse_int(t[24], slv) ? (int32_t)slv : 0;
    const int G = d.energygroups, M = d.mattypes;
    if (G < 1 || M < 1 || G > kMaxGroups || M > kMaxMats) throw std::runtime_error("deck: bad sizes");
    auto *b = new nraps_problem_buf();
    std::unique_ptr<nraps_problem_buf> guard(b);
    const char *names[8] = {"SigT", "SigS", "mu", "SigA", "SigF", "nuT", "ChiT", "Scat"};
    std::vector<double> *dst[8] = {&b->sigt, &b->sigs, &b->mu, &b->siga,
                                   &b->sigf, &b->nut,  &b->chit, &b->scat};
    for (int k = 0; k < 8; k++) {
      if (!parse_list(t[14 + k], *dst[k])) throw std::runtime_error(std::string("deck: bad ") + names[k]);
      size_t want = k < 7 ? (size_t)G * M : (size_t)M * G * G;
      if (dst[k]->size() != want)
        throw std::runtime_error(std::string("deck: ") + names[k] + " has " +
                                 std::to_string(dst[k]->size()) + " values, expected " +
                                 std::to_string(want));
    }
    std::vector<double> mid;
    if (!parse_list(t[22], mid) || mid.empty()) throw std::runtime_error("deck: bad MatID");
    // mesh_gen, src/main.rs:85-143
    std::vector<int> temp;
    for (double v : mid) {
      int m = (int)v;
      int64_t reps = (m == 0 || m == 1) ? d.mpfr : d.mpwr;
      for (int64_t k = 0; k < reps; k++) temp.push_back(m);
    }
    for (int i1 = 1; i1 < d.numass; i1++)
      for (int64_t k = 0; k < d.mpwr; k++) {
        size_t at = ((size_t)i1 * temp.size()) / (size_t)d.numass;
        if (at >= temp.size()) throw std::runtime_error("deck: mesh too small");
        temp.erase(temp.begin() + at);
      }
    size_t half = (size_t)(d.mpwr / 2);
    if (temp.size() < 2 * half + 1) throw std::runtime_error("deck: mesh too small");
    temp.erase(temp.begin(), temp.begin() + half);
    temp.resize(temp.size() - half);
    b->fuel.clear();
    double left = 0.0;
    for (size_t c = 0; c < temp.size(); c++) {
      int m = temp[c];
      bool fuel = (m == 0 || m == 1);
      double w = fuel ? d.delta_fuel : d.delta_water;
      b->matid.push_back(m);
      b->delta_x.push_back(w);
      b->left.push_back(left);
      b->right.push_back(left + w);
      left += w;
      if (fuel) b->fuel.push_back((int32_t)c);
    }
    b->view.n_cells = (int32_t)temp.size();
    b->view.n_groups = G;
    b->view.n_mats = M;
    b->view.numass = d.numass;
    b->view.boundl = d.boundl;
    b->view.boundr = d.boundr;
    b->refresh_view();
    if (info) *info = d;
    *out = guard.release();
    return NRAPS_OK;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_INVALID;
  }
}

// ---------------------------------------------------------------------------
// Output files, src/plot_solution.rs:14-58
// ---------------------------------------------------------------------------
extern "C" int nraps_write_outputs(const char *dir, const nraps_problem *pb, const nraps_run_opts *o,
                                   const nraps_result *r) {
  try {
    if (!dir || !pb || !o || !r || !r->flux || !r->assembly_average || !r->fission_source || !r->k ||
        !r->k_fund) {
      set_error("nraps_write_outputs: missing argument or result array");
      return NRAPS_E_INVALID;
    }
    const int nc = pb->n_cells, G = pb->n_groups, gens = o->generations;
    auto row = [](const double *v, int n) {
      std::string s;
      for (int i = 0; i < n; i++) {
        if (i) s += ',';
        s += rust_display(v[i]);
      }
      s += '\n';
      return s;
    };
    auto write = [&](const char *name, const std::string &body) {
      std::string path = std::string(dir) + "/" + name;
      FILE *fp = std::fopen(path.c_str(), "wb");
      if (!fp) throw std::runtime_error("cannot write " + path);
      std::fwrite(body.data(), 1, body.size(), fp);
      std::fclose(fp);
    };
    std::string vars = rust_display(pb->mesh_right[nc - 1]) + "\n" + std::to_string(nc) + "\n" +
                       std::to_string(gens) + "\n";
    write("vars.csv", vars);
    std::string iface;
    for (int g = 0; g < G; g++) iface += row(r->flux + (size_t)g * nc, nc);
    for (int g = 0; g < G; g++) iface += row(r->assembly_average + (size_t)g * nc, nc);
    iface += row(r->fission_source, nc);
    write("interface.csv", iface);
    write("k_eff.csv", row(r->k, gens) + row(r->k_fund, gens));
    return NRAPS_OK;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_IO;
  }
}

extern "C" int nraps_write_diffusion_outputs(const char *dir, const nraps_problem *pb,
                                             int64_t generations,
                                             const nraps_diffusion_result *r) {
  try {
    if (!dir || !pb || !r || !r->flux || !r->assembly_average || !r->fission_source ||
        pb->n_cells < 1) {
      set_error("nraps_write_diffusion_outputs: missing argument or result array");
      return NRAPS_E_INVALID;
    }
    const int nc = pb->n_cells, G = pb->n_groups;
    std::string iface;
    auto row = [&](const double *v) {
      for (int i = 0; i < nc; i++) {
        if (i) iface += ',';
        iface += rust_display(v[i]);
      }
      iface += '\n';
    };
    for (int g = 0; g < G; g++) row(r->flux + (size_t)g * nc);
    for (int g = 0; g < G; g++) row(r->assembly_average + (size_t)g * nc);
    row(r->fission_source);
    auto write = [&](const char *name, const std::string &body) {
      std::string path = std::string(dir) + "/" + name;
      FILE *fp = std::fopen(path.c_str(), "wb");
      if (!fp) throw std::runtime_error("cannot write " + path);
      std::fwrite(body.data(), 1, body.size(), fp);
      std::fclose(fp);
    };
    write("vars.csv", rust_display(pb->mesh_right[nc - 1]) + "\n" + std::to_string(nc) + "\n" +
                          std::to_string(generations) + "\n");
    write("interface.csv", iface);
    write("k_eff.csv", rust_display(r->k) + "\n");
    return NRAPS_OK;
  } catch (const std::exception &e) {
    set_error(e.what());
    return NRAPS_E_IO;
  }
}
