// cli.cpp — command-line driver: input deck -> mesh -> solve -> reference CSV
// outputs, the run of src/main.rs (process_input, mesh_gen, the solution
// dispatch and plot_solution; src/main.rs:85-143, :335-364), on the GPU
// through the C ABI.
//
//   nraps_run DECK [OUTDIR] [--mc|--diffusion] [--fp32|--fp64] [--event|--region|--plain]
//             [--bank] [--diffusion-start] [--histories N] [--generations N]
//             [--skip N] [--seed S] [--every-collision]
//
// The deck's `Solution` picks the solver as the reference's dispatcher does
// (src/main.rs:336-353): 1 runs the Monte Carlo solve, otherwise the
// finite-difference diffusion solve (nalgebra_method) for `solver` 0; the
// reference implements no other solver.  --mc / --diffusion override it.
// Monte Carlo defaults: the deck's histories/generations/skip, the FP32
// event kernel with region-by-region surface tracking (--event: Woodcock),
// the reference source and XS rules, seed 42.
// Writes vars.csv, interface.csv and k_eff.csv into OUTDIR (default ".").
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
               "usage: nraps_run DECK [OUTDIR] [--mc|--diffusion] [--fp32|--fp64] [--event|--region|--plain]\n"
               "                 [--bank] [--diffusion-start] [--histories N] [--generations N]\n"
               "                 [--skip N] [--seed S] [--every-collision]\n");
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
  o.method = NRAPS_EVENT_REGION;
  long long hist = -1, gens = -1, skip = -1, seed = -1;
  int solver = -1;  // -1: the deck's Solution, 0: Monte Carlo, 1: diffusion
  std::vector<std::string> pos;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&](long long *v) {
      if (i + 1 >= argc || !parse_i64(argv[i + 1], v)) return false;
      i++;
      return true;
    };
    if (a == "--mc") solver = 0;
    else if (a == "--diffusion") solver = 1;
    else if (a == "--diffusion-start") o.bank_init = NRAPS_BANK_INIT_DIFFUSION;
    else if (a == "--fp32") o.precision = NRAPS_FP32;
    else if (a == "--fp64") o.precision = NRAPS_FP64;
    else if (a == "--event") o.method = NRAPS_EVENT_WOODCOCK;
    else if (a == "--region") o.method = NRAPS_EVENT_REGION;
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
  if (solver < 0) solver = info.solution_mc ? 0 : 1;
  if (solver == 1) {
    if (info.solver >= 1 && info.solver <= 3) {
      std::fprintf(stderr, "nraps_run: solver %d is not implemented by the reference (only 0, "
                           "the direct solve); use --mc for the Monte Carlo solution\n", info.solver);
      nraps_problem_free(buf);
      return 1;
    }
    const size_t G = (size_t)p->n_groups, nc = (size_t)p->n_cells;
    std::vector<double> flux(G * nc), aa(G * nc), fs(nc);
    nraps_diffusion_result d;
    std::memset(&d, 0, sizeof d);
    d.flux = flux.data();
    d.assembly_average = aa.data();
    d.fission_source = fs.data();
    nraps_diffusion_opts dopt;
    nraps_diffusion_opts_default(&dopt);
    std::fprintf(stderr, "nraps_run: %s, %d cells, %d groups, diffusion\n", deck.c_str(), p->n_cells,
                 p->n_groups);
    st = nraps_diffusion_solve(p, &dopt, &d);
    if (st != NRAPS_OK && st != NRAPS_E_NOT_CONVERGED) {
      std::fprintf(stderr, "nraps_run: solve failed (%d): %s\n", st, nraps_last_error());
      nraps_problem_free(buf);
      return 1;
    }
    // src/discrete.rs:345 prints k with 10 decimals
    std::printf("%.10f\nk = %.10f  (%d iterations, %.2f ms device, %.1f ms)\n", d.k, d.k,
                d.iterations, d.ms_device, d.ms_total);
    st = nraps_write_diffusion_outputs(outdir.c_str(), p, info.generations, &d);
    nraps_problem_free(buf);
    if (st != NRAPS_OK) {
      std::fprintf(stderr, "nraps_run: %s\n", nraps_last_error());
      return 1;
    }
    return 0;
  }
  if (o.bank_init == NRAPS_BANK_INIT_DIFFUSION) o.source = NRAPS_SOURCE_FISSION_BANK;
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
