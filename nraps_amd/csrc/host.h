// host.h — host-side problem model of the MI355X solver (internal).
//
// Validation of the C-ABI inputs, the derived tables the kernels read
// (transport XS, absorption ratios, cumulative scatter rows, material
// regions, the Woodcock majorant, region prefix integrals of nu*Sigma_f),
// and their packing into one device blob.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nraps.h"

namespace nraps {

// Thread-local error message behind nraps_last_error().
void set_error(const std::string &msg);
const std::string &last_error();

constexpr int kMaxGroups = 32;
constexpr int kMaxMats = 255;

int validate(const nraps_problem *pb, const nraps_run_opts *o, std::string *why);

// Byte offsets of each table inside the packed blob (all 16-byte aligned).
// The first `hot` bytes are what the event kernel stages in LDS; the
// birth-only tables after them stay in global memory.
struct TabLayout {
  int32_t edges;    // Real[nc+1]   cell edges, edges[c] = mesh_left[c]
  int32_t mat8;     // u8[nc]       material of each cell
  int32_t reg16;    // u16[nc]      material region of each cell
  int32_t sigw;     // Real[G*M]    1/inv_sigtr (Woodcock site XS)
  int32_t invtr;    // Real[G*M]    inv_sigtr
  int32_t absr;     // Real[G*M]    siga/sigt
  int32_t invs;     // Real[G*M]    1/sigs
  int32_t nsf;      // Real[G*M]    nut*sigf (site production)
  int32_t chicdf;   // Real[M*G]    cumulative chi per material
  int32_t cumrow;   // Real[M*G*G]  cumulative scatter rows
  int32_t nsfd;     // double[G*M]  nut*sigf (k accumulation)
  int32_t kreg;     // double[G*(R+1)] region prefix integrals of nu*Sigma_f
  int32_t bucket;   // u16[nb]      cell lookup: largest c with edge[c] < XL + b*h
  int32_t regedge;  // double[R+1]  region edges, regedge[r] = edges[regcell[r]]
  int32_t regcell;  // u16[R+1]     first cell of each region, regcell[R] = nc
  int32_t hot;      // bytes of the LDS-staged part (all of the above, and the
                    // birth tables when they are placed before `hot`)
  int32_t fuel;     // i32[nf]      source cells (birth only)
  int32_t dx;       // Real[nc]     Mesh.delta_x (birth only)
  int32_t bytes;    // total (multiple of 16)
};

constexpr int kMaxBuckets = 8192;
constexpr int64_t kLdsBytes = 160 * 1024;  // LDS of one gfx950 workgroup
#ifndef NRAPS_BIRTH_IN_LDS
#define NRAPS_BIRTH_IN_LDS 1
#endif
constexpr bool kBirthInLds = NRAPS_BIRTH_IN_LDS != 0;
constexpr int kMaxCells = 65535;

struct Prepared {
  int nc = 0, G = 0, M = 0, R = 0, nf = 0, numass = 1, nb = 1;
  double boundl = 0, boundr = 0, XL = 0, XR = 0, maj = 0, nut_conv = 0, inv_h = 0;
  std::vector<double> edges, dx, sigw, inv_sigtr, abs_ratio, inv_sigs, nusigf, chicdf,
      cumrow, reg_lo, kreg;
  std::vector<int32_t> mat, region, fuel, reg_mat, reg_cell;
  std::vector<uint16_t> bucket;

  void build(const nraps_problem *pb);
  TabLayout layout(bool fp32) const;
  std::vector<uint8_t> pack(bool fp32, TabLayout *lay) const;
};

// Owned problem storage behind nraps_problem_buf.
struct ProblemBuf {
  nraps_problem view{};
  std::vector<int32_t> matid, fuel;
  std::vector<double> delta_x, left, right, sigt, sigs, mu, siga, sigf, nut, chit, scat, invtr;
  void refresh_view();
  void copy_from(const nraps_problem *pb);
};

// Rust `Display` for f64: shortest round-trip digits, no exponent.
std::string rust_display(double v);

}  // namespace nraps

struct nraps_problem_buf : nraps::ProblemBuf {};
