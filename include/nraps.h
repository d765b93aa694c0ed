/*
 * nraps.h — C ABI of the MI355X Monte Carlo k-eigenvalue solver.
 *
 * This is the drop-in boundary for the NRAPS particle history loop.  The
 * reference has no FFI: its path is the crate-internal Rust function
 *
 *   pub fn monte_carlo(variables: &Variables, xsdata: &XSData, delta_x: &DeltaX,
 *                      meshid: &Vec<Mesh>, fuel_indices: &Vec<usize>,
 *                      k_new: f32) -> SolutionResults          (src/mc_code.rs:276-283)
 *
 * called from the (commented) dispatcher at src/main.rs:337.  nraps_mc_solve()
 * replaces the body of that function: nraps_problem carries the flattened
 * Variables/XSData/Mesh/fuel_indices inputs (src/main.rs:22-75), nraps_run_opts
 * the run controls, and nraps_result the caller-allocated SolutionResults
 * arrays (src/main.rs:77-83).
 *
 * Conventions
 *   - Plain C types only; every array is a caller-owned pointer plus a size
 *     derivable from the problem.  The library copies inputs to the device
 *     and never frees caller memory.
 *   - Every entry point returns an int status (NRAPS_OK == 0, negative on
 *     error) and never lets a C++ exception cross the ABI.  The message for
 *     the last failure on the calling thread is nraps_last_error().
 *   - Cross-section arrays use the reference layout: [g][mat] flattened as
 *     mat + n_mats*g (src/process_input.rs:114); the scatter matrix is
 *     [mat][from][to] = n_groups*n_groups*mat + n_groups*from + to
 *     (src/mc_code.rs:89-90).
 *   - Values cross the boundary in double precision; FP32 is internal to the
 *     precision form.
 */
#ifndef NRAPS_H
#define NRAPS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NRAPS_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------ */
#define NRAPS_OK 0
#define NRAPS_PENDING 1            /* server: request not finished yet */
#define NRAPS_E_INVALID -1         /* malformed problem or options */
#define NRAPS_E_NOMEM -2           /* host or device allocation failed */
#define NRAPS_E_DEVICE -3          /* HIP runtime error / no device */
#define NRAPS_E_OVER_BUDGET -4     /* server: request exceeds the memory budget */
#define NRAPS_E_CANCELLED -5       /* server: request was cancelled */
#define NRAPS_E_EVENT_CAP -6       /* a history hit max_events (results kept) */
#define NRAPS_E_NO_FISSION -7      /* fission-bank mode produced no sites */
#define NRAPS_E_BANK_OVERFLOW -8   /* fission sites exceeded the bank capacity */
#define NRAPS_E_NOT_FOUND -9       /* unknown ticket / config */
#define NRAPS_E_IO -10             /* file could not be read or written */
#define NRAPS_E_NOT_CONVERGED -11  /* diffusion: iteration limit reached (results kept) */

/* ---- run option enumerations ------------------------------------------- */
#define NRAPS_FP64 0 /* double tracking and tallies */
#define NRAPS_FP32 1 /* float tracking/tallies, double k, bank and variance */

#define NRAPS_HISTORY_SURFACE 0 /* plain port: one thread per history, cell-by-cell
                                   surface tracking, src/mc_code.rs:134-213 */
#define NRAPS_EVENT_WOODCOCK 1  /* event pipeline with Woodcock delta tracking */
#define NRAPS_EVENT_REGION 2    /* event pipeline with the reference's surface tracking
                                   taken region by region: a flight is sampled at
                                   birth, after a scatter and on entry to another
                                   material region (src/mc_code.rs:142-149), cells
                                   are crossed without events */

#define NRAPS_SOURCE_UNIFORM_FUEL 0 /* reference source: uniform fuel cell each
                                       generation, src/mc_code.rs:39-53 */
#define NRAPS_SOURCE_FISSION_BANK 1 /* fission-site bank carried between cycles */

#define NRAPS_XS_ON_MATERIAL_CHANGE 0 /* reference: XS index refreshed only on
                                         entry to a new material, src/mc_code.rs:147 */
#define NRAPS_XS_EVERY_COLLISION 1    /* XS follow the current group */

#define NRAPS_BANK_INIT_UNIFORM_FUEL 0 /* fission bank: generations 0 and 1 start
                                          from the reference source */
#define NRAPS_BANK_INIT_DIFFUSION 1    /* fission bank: generation 0 starts from
                                          the diffusion fission source */

/* ---- problem ----------------------------------------------------------- */
typedef struct nraps_problem {
  int32_t n_cells;  /* meshid.len() */
  int32_t n_groups; /* Variables.energygroups */
  int32_t n_mats;   /* Variables.mattypes */
  int32_t numass;   /* Variables.numass (assembly averaging, src/mc_code.rs:259) */
  double boundl;    /* Variables.boundl: >0 reflect (albedo), <=0 vacuum */
  double boundr;    /* Variables.boundr */
  /* mesh, struct-of-arrays form of Vec<Mesh> (src/main.rs:68-75) */
  const int32_t *matid;     /* [n_cells] material of each cell */
  const double *delta_x;    /* [n_cells] Mesh.delta_x (normalisation + source) */
  const double *mesh_left;  /* [n_cells] */
  const double *mesh_right; /* [n_cells]; mesh_right[c] == mesh_left[c+1] */
  /* fuel_indices (src/main.rs:110-115); NULL => cells with matid 0 or 1 */
  int32_t n_fuel;
  const int32_t *fuel_indices;
  /* XSData (src/main.rs:46-56), each [n_groups*n_mats] */
  const double *sigt, *sigs, *mu, *siga, *sigf, *nut, *chit;
  const double *scat_matrix; /* [n_mats*n_groups*n_groups] */
  const double *inv_sigtr;   /* optional; NULL => 1/(sigt - mu*sigs)
                                (src/process_input.rs:152-156) */
} nraps_problem;

/* ---- run options ------------------------------------------------------- */
typedef struct nraps_run_opts {
  int64_t histories;  /* Variables.histories: particles per generation */
  int32_t generations;
  int32_t skip;       /* first active generation (1 <= skip < generations) */
  uint64_t seed;      /* PCG32 seed (the reference seeds from the clock) */
  uint64_t problem_id; /* high bits of every particle stream id (< 2^19) */
  int32_t precision;  /* NRAPS_FP64 / NRAPS_FP32 */
  int32_t method;     /* NRAPS_HISTORY_SURFACE / NRAPS_EVENT_WOODCOCK / NRAPS_EVENT_REGION */
  int32_t source;     /* NRAPS_SOURCE_* */
  int32_t xs_refresh; /* NRAPS_XS_* */
  double k_initial;   /* the k_new argument of monte_carlo (reference: 1.0) */
  int64_t max_events; /* per-history event cap; 0 => 10^7; caps above 2^31 - 2
                         act as 2^31 - 2 (the width of the event counter) */
  int32_t device;     /* HIP device ordinal */
  int32_t use_graph;  /* capture one cycle in a hipGraph and replay it */
  int64_t chunk;      /* particles per work chunk; 0 => automatic */
  /* debugging/parity: per-history records for the first trace_histories
     particles of generation trace_cycle (needs the trace arrays in result) */
  int64_t trace_histories;
  int32_t trace_events;
  int32_t trace_cycle;
  /* fission-bank start (NRAPS_BANK_INIT_*).  With NRAPS_BANK_INIT_DIFFUSION
     the solver first runs nraps_diffusion_solve on the device and places the
     N generation-0 sources at the quantiles (p + 1/2)/N of its fission source
     (piecewise constant per cell); generation 0 then banks sites with
     1/k_diffusion.  Only with NRAPS_SOURCE_FISSION_BANK. */
  int32_t bank_init;
  int32_t reserved[7];
} nraps_run_opts;

/* One recorded history event (parity traces). */
#define NRAPS_EV_BIRTH 0
#define NRAPS_EV_ABSORB 1
#define NRAPS_EV_SCATTER 2
#define NRAPS_EV_LEAK 3
#define NRAPS_EV_CAP 4
typedef struct nraps_trace_event {
  int32_t type;  /* NRAPS_EV_* */
  int32_t cell;  /* cell index at the event */
  int32_t group; /* group before the event */
  int32_t aux;   /* scatter: new group; birth: sites-mode flag */
  double x;      /* position at the event */
  double mu;     /* direction before the event */
} nraps_trace_event;

/* ---- results ----------------------------------------------------------- */
typedef struct nraps_result {
  /* caller-allocated outputs; any pointer may be NULL to skip it */
  double *flux;             /* [G*nc] SolutionResults.flux, reference
                               normalisation (src/mc_code.rs:340-360) */
  double *assembly_average; /* [G*nc] (src/mc_code.rs:259-274) */
  double *fission_source;   /* [nc] */
  double *k;                /* [generations] k per generation */
  double *k_fund;           /* [generations] running active mean (src/mc_code.rs:368-376) */
  double *flux_mean;        /* [G*nc] unbiased mean over active generations */
  double *flux_sigma;       /* [G*nc] standard error of flux_mean */
  double *tally_last;       /* [G*nc] raw track-length tally of the last generation */
  double *k_ref;            /* [generations] k recomputed from the tally with
                               the reference formula (src/mc_code.rs:342-351) */
  int64_t *sites;           /* [generations] fission sites banked per generation */
  /* per-history parity records (generation trace_cycle, p < trace_histories) */
  nraps_trace_event *trace; /* [trace_histories * trace_events] */
  int32_t *trace_len;       /* [trace_histories] events recorded (may exceed cap) */
  int64_t *draws;           /* [trace_histories] PCG32 draws consumed */
  double *kacc;             /* [trace_histories] per-history k contribution */
  /* scalar outputs */
  double k_mean;            /* mean of k over active generations */
  double k_sigma;           /* standard error of k_mean */
  int64_t histories_run;
  int64_t collisions;       /* real collisions */
  int64_t virtual_collisions;
  int64_t crossings;        /* surface/boundary crossings (surface tracking) */
  int64_t absorptions;
  int64_t leaks;
  int64_t capped;           /* histories stopped by max_events */
  double mean_flights;      /* (collisions + leaks) / histories */
  double ms_total;          /* wall time of the solve, upload to download */
  double ms_device;         /* device time of the generation loop */
  double ms_transport;      /* summed device time of the transport launches */
  int64_t transport_launches;
  int32_t status;           /* same value the call returned */
  int32_t reserved[7];
} nraps_result;

/* Fill opts with the reference defaults (FP64, surface, uniform fuel source,
   XS refreshed on material change, k_initial 1.0, seed 42). */
void nraps_run_opts_default(nraps_run_opts *opts);

/* Run the k-eigenvalue calculation.  Replaces mc_code::monte_carlo
   (src/mc_code.rs:276-380). */
int nraps_mc_solve(const nraps_problem *problem, const nraps_run_opts *opts,
                   nraps_result *result);

/* Run n independent problems in one device pipeline (shared persistent
   launches, one hipGraph per cycle).  Each problem keeps its own RNG streams
   (opts[i].problem_id), tallies and fission bank.  Precision and method must
   agree across the batch; histories, generations and mesh size may differ. */
int nraps_mc_solve_batch(int32_t n, const nraps_problem *const *problems,
                         const nraps_run_opts *opts, nraps_result *const *results);

/* Message for the last failure on this thread ("" if none). */
const char *nraps_last_error(void);

/* Library build / device description, e.g. "nraps 1 gfx950 ...". */
int nraps_device_info(int32_t device, char *buf, int32_t buflen);

/* ---- multigroup diffusion (src/discrete.rs) ----------------------------- */
/* Finite-difference multigroup diffusion k-eigenvalue by power iteration,
   the reference's nalgebra_method (src/discrete.rs:169-356), on the device.
   The reference inverts each group's n x n matrix densely
   (src/discrete.rs:186-205) and multiplies the inverse with the right-hand
   side every iteration.  The matrix is tridiagonal: the library factors it
   once and solves each group in O(n), which gives the same A^-1 rhs.  The
   matrix rules, the group sweep, the convergence test and the power
   normalisation of the flux are the reference's. */
typedef struct nraps_diffusion_opts {
  int32_t device;
  int32_t max_iterations; /* 0 => 1000000; hitting it => NRAPS_E_NOT_CONVERGED */
  double flux_tol;        /* iterate while flux change >= flux_tol (0 => 1e-5) */
  double k_tol;           /*   and k change >= k_tol (0 => 1e-6), src/discrete.rs:207 */
  int32_t reserved[8];
} nraps_diffusion_opts;

typedef struct nraps_diffusion_result {
  /* caller-allocated outputs; any pointer may be NULL to skip it */
  double *flux;             /* [G*nc] power-normalised flux (src/discrete.rs:304-343) */
  double *assembly_average; /* [G*nc] (src/discrete.rs:151-167) */
  double *fission_source;   /* [nc] sum over groups of nut*sigf*flux */
  /* scalar outputs */
  double k;
  double delta_flux;        /* convergence measures at exit (src/discrete.rs:252-289) */
  double delta_k;
  int32_t iterations;
  int32_t status;           /* same value the call returned */
  double ms_total;          /* wall time, upload to download */
  double ms_device;         /* device time of the iteration kernel */
  int32_t reserved[8];
} nraps_diffusion_result;

void nraps_diffusion_opts_default(nraps_diffusion_opts *opts);

/* Solve one problem.  Uses n_cells >= 2, the XS of the problem (inv_sigtr,
   sigt, scat_matrix, nut, sigf, chit), boundl/boundr and numass. */
int nraps_diffusion_solve(const nraps_problem *problem, const nraps_diffusion_opts *opts,
                          nraps_diffusion_result *result);

/* Solve n independent problems in one launch (one workgroup per problem). */
int nraps_diffusion_solve_batch(int32_t n, const nraps_problem *const *problems,
                                const nraps_diffusion_opts *opts,
                                nraps_diffusion_result *const *results);

/* ---- PCG32 (src/rand.rs:36-107) ---------------------------------------- */
/* out[s*draws + i] = i-th next_u32 of stream (stream0 + s) seeded like
   Prn32::new with an explicit seed (src/rand.rs:49-71).  Runs on the device. */
int nraps_pcg32_fill(uint64_t seed, uint64_t stream0, int64_t n_streams,
                     int32_t draws, uint32_t *out);

/* ---- host-side problem tools (the steps either side of the path) -------- */
typedef struct nraps_problem_buf nraps_problem_buf; /* opaque, library-owned */

/* Reference run controls read from an input deck (src/main.rs:22-39). */
typedef struct nraps_deck_info {
  int32_t analk, mattypes, energygroups;
  int64_t generations, histories, skip;
  int32_t numass, numrods;
  double roddia, rodpitch; /* rodpitch is stored as pitch - diameter, like
                              src/process_input.rs:101 */
  int64_t mpfr, mpwr;
  double boundl, boundr;
  double delta_fuel, delta_water; /* DeltaX (src/process_input.rs:108-111) */
  int32_t solution_mc;            /* Solution == 1 */
  int32_t solver;
} nraps_deck_info;

/* Parse a reference input deck (src/process_input.rs:85-175) and expand the
   mesh (src/main.rs:85-143).  The returned buffer owns every array. */
int nraps_deck_load(const char *path, nraps_problem_buf **out,
                    nraps_deck_info *info);

/* Build one of the seeded synthetic workloads:
   config 0/1: 9-region 4-group slab; 2: 201-region 8-group upscatter slab;
   3/4: the index-th problem of the uneven batch.  Histories/cycles for the
   config are written to opts (may be NULL). */
int nraps_problem_generate(int32_t config, int32_t index, uint64_t seed,
                           nraps_problem_buf **out, nraps_run_opts *opts);

/* Borrow the problem view of a buffer (valid until nraps_problem_free). */
const nraps_problem *nraps_problem_view(const nraps_problem_buf *buf);
void nraps_problem_free(nraps_problem_buf *buf);

/* Check a problem/options pair without touching the device. */
int nraps_validate(const nraps_problem *problem, const nraps_run_opts *opts);

/* Write vars.csv, interface.csv and k_eff.csv into dir in the reference's
   layout and number format (src/plot_solution.rs:14-58). */
int nraps_write_outputs(const char *dir, const nraps_problem *problem,
                        const nraps_run_opts *opts, const nraps_result *result);

/* Write the diffusion solution in the same files: vars.csv (generations is
   the deck's, as src/main.rs passes it), interface.csv (G flux rows, G
   assembly rows and the fission-source row) and k_eff.csv (one row, k).
   The reference hands plot_solution an empty fission source and k_fund for
   this path (src/discrete.rs:347-355); here the rows are filled. */
int nraps_write_diffusion_outputs(const char *dir, const nraps_problem *problem,
                                  int64_t generations, const nraps_diffusion_result *result);

/* ---- serving: continuous batching of solve requests --------------------- */
typedef struct nraps_server nraps_server;

typedef struct nraps_server_opts {
  int32_t device;
  int32_t precision;        /* NRAPS_FP32 or NRAPS_FP64 for every request */
  int64_t budget_bytes;     /* admission limit on bank + tally bytes */
  int32_t max_slots;        /* problems resident at once */
  int32_t min_active;       /* active cycles before the sigma stop applies */
  double sigma_k_target;    /* stop a request once its relative k standard error,
                               SEM(k) / |mean k|, is <= this (50 pcm = 5e-4;
                               0 => off) */
  int32_t use_graph;
  int32_t method;           /* event method of every request: 0 or NRAPS_EVENT_REGION
                               (region surface tracking), NRAPS_EVENT_WOODCOCK */
  int32_t reserved[6];
} nraps_server_opts;

int nraps_server_create(const nraps_server_opts *opts, nraps_server **out);
/* Queue a request; the problem and options are copied. */
int nraps_server_submit(nraps_server *srv, const nraps_problem *problem,
                        const nraps_run_opts *opts, uint64_t *ticket);
/* NRAPS_OK and result filled when done, NRAPS_PENDING otherwise. */
int nraps_server_poll(nraps_server *srv, uint64_t ticket, nraps_result *result);
/* Block until the request finishes (timeout_ms < 0: forever). */
int nraps_server_wait(nraps_server *srv, uint64_t ticket, nraps_result *result,
                      int64_t timeout_ms);
int nraps_server_cancel(nraps_server *srv, uint64_t ticket);
/* Latency of a finished request: submit->admit and submit->done, ms. */
int nraps_server_latency(nraps_server *srv, uint64_t ticket, double *queued_ms,
                         double *total_ms, int32_t *cycles_run);
/* Drop a finished request's record and results (the server keeps them until
   then, so a long-running server should forget what it has read).  Returns
   NRAPS_PENDING for a request that is not finished. */
int nraps_server_forget(nraps_server *srv, uint64_t ticket);
void nraps_server_destroy(nraps_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* NRAPS_H */
