// cli.cpp — command-line driver: input deck -> mesh -> Monte Carlo solve ->
// reference CSV outputs, the run that src/main.rs performs for the Monte Carlo
// solution (process_input, mesh_gen, monte_carlo, plot_solution; src/main.rs:
// 85-143, :337, :358-364), on the GPU through the C ABI.
//
//   nraps_run DECK [OUTDIR] [--fp32|--fp64] [--event|--plain] [--bank]
//             [--histories N] [--generations N] [--skip N] [--seed S]
//             [--every-collision]
//
// Defaults: the deck's histories/generations/skip, FP32 event-based
// Woodcock tracking, the reference source and XS rules, seed 42.  Writes
// vars.csv, interface.csv and k_eff.csv into OUTDIR (default ".").
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nraps.h"

namespace {

int usage() {
  std::fprintf(stderr,
               "usage: nraps_run DECK [OUTDIR] [--fp32|--fp64] [--event|--plain] [--bank]\n"
               "                 [--histories N] [--generations N] [--skip N] [--seed S]\n"
               "                 [--every-collision]\n");
  return 2;
}

bool parse_i64(const char *s, long long *v) {
  errno = 0;
  char *e = nullptr;
  long long x = std::strtoll(s, &e, 10);
  if (errno || !e || *e) return false;
  *v = x;
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return usage();
  std::string deck, outdir = ".";
  nraps_run_opts o;
  nraps_run_opts_default(&o);
  o.precision = NRAPS_FP32;
  o.method = NRAPS_EVENT_WOODCOCK;
  long long hist = -1, gens = -1, skip = -1, seed = -1;
  std::vector<std::string> pos;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](long long *v) {
      if (i + 1 >= argc || !parse_i64(argv[i + 1], v)) return false;
      i++;
      return true;
    };
    if (a == "--fp32") o.precision = NRAPS_FP32;
    else if (a == "--fp64") o.precision = NRAPS_FP64;
    else if (a == "--event") o.method = NRAPS_EVENT_WOODCOCK;
    else if (a == "--plain") o.method = NRAPS_HISTORY_SURFACE;
    else if (a == "--bank") o.source = NRAPS_SOURCE_FISSION_BANK;
    else if (a == "--every-collision") o.xs_refresh = NRAPS_XS_EVERY_COLLISION;
    else if (a == "--histories") { if (!next(&hist)) return usage(); }
    else if (a == "--generations") { if (!next(&gens)) return usage(); }
    else if (a == "--skip") { if (!next(&skip)) return usage(); }
    else if (a == "--seed") { if (!next(&seed)) return usage(); }
    else if (a.size() > 1 && a[0] == '-') return usage();
    else pos.push_back(a);
  }
  if (pos.empty() || pos.size() > 2) return usage();
  deck = pos[0];
  if (pos.size() == 2) outdir = pos[1];
  if (o.method == NRAPS_HISTORY_SURFACE) o.precision = NRAPS_FP64;  // the plain port is FP64

  nraps_problem_buf *buf = nullptr;
  nraps_deck_info info;
  int st = nraps_deck_load(deck.c_str(), &buf, &info);
  if (st != NRAPS_OK) {
    std::fprintf(stderr, "nraps_run: %s\n", nraps_last_error());
    return 1;
  }
  const nraps_problem *p = nraps_problem_view(buf);
  o.histories = hist > 0 ? hist : info.histories;
  o.generations = (int32_t)(gens > 0 ? gens : info.generations);
  o.skip = (int32_t)(skip > 0 ? skip : info.skip);
  if (seed >= 0) o.seed = (uint64_t)seed;

  const size_t G = (size_t)p->n_groups, nc = (size_t)p->n_cells, ng = (size_t)o.generations;
  std::vector<double> flux(G * nc), aa(G * nc), fs(nc), k(ng), kf(ng), fm(G * nc), fsg(G * nc);
  nraps_result r;
  std::memset(&r, 0, sizeof r);
  r.flux = flux.data();
  r.assembly_average = aa.data();
  r.fission_source = fs.data();
  r.k = k.data();
  r.k_fund = kf.data();
  r.flux_mean = fm.data();
  r.flux_sigma = fsg.data();
  std::fprintf(stderr, "nraps_run: %s, %d cells, %d groups, %lld histories x %d generations (skip %d)\n",
               deck.c_str(), p->n_cells, p->n_groups, (long long)o.histories, o.generations, o.skip);
  st = nraps_mc_solve(p, &o, &r);
  if (st != NRAPS_OK && st != NRAPS_E_EVENT_CAP) {
    std::fprintf(stderr, "nraps_run: solve failed (%d): %s\n", st, nraps_last_error());
    nraps_problem_free(buf);
    return 1;
  }
  std::printf("k = %.8f +- %.8f  (%.3g histories/s, %.1f ms)\n", r.k_mean, r.k_sigma,
              r.ms_total > 0 ? (double)r.histories_run / (r.ms_total * 1e-3) : 0.0, r.ms_total);
  st = nraps_write_outputs(outdir.c_str(), p, &o, &r);
  nraps_problem_free(buf);
  if (st != NRAPS_OK) {
    std::fprintf(stderr, "nraps_run: %s\n", nraps_last_error());
    return 1;
  }
  return 0;
}
