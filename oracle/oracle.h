/*
 * oracle.h — CPU FP64 restatement of the NRAPS Monte Carlo history loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP product:
 * only the tests, the smoke check and the CPU-baseline leg of the bench may
 * load it.  The product never links or calls it.
 *
 * It restates src/mc_code.rs:1-380 and src/rand.rs:36-107 of the reference
 * (and, in diffusion.c, the diffusion solver of src/discrete.rs:1-356)
 * with the deviations the product documents (single-draw CDF inversion,
 * exactly N histories, guarded edge draws, 32-bit indices), in plain C and
 * double precision, history by history.  It shares exactly one routine with
 * the product: the ln() of nraps_core.h, so FP64 histories can be compared
 * bit for bit; that routine is pinned separately against a 60-digit log.
 * The problem/options/result structs are the ones of include/nraps.h.
 */
#ifndef NRAPS_ORACLE_H
#define NRAPS_ORACLE_H

#include <stdint.h>

#include "../include/nraps.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Full k-eigenvalue run on `threads` CPU threads (0 => nproc - 1, as
   src/mc_code.rs:302).  Results do not depend on the thread count. */
int oracle_solve(const nraps_problem *pb, const nraps_run_opts *opts,
                 int threads, nraps_result *res);

/* Per-cycle bank access for parity tests: after oracle_solve with
   keep_banks=1, the source bank used by generation c (x, cell). */
int oracle_solve_banks(const nraps_problem *pb, const nraps_run_opts *opts,
                       int threads, nraps_result *res, int32_t bank_cycle,
                       double *bank_x, int32_t *bank_cell, int64_t *bank_n);

/* Fission-bank run whose generation 0 is born from the given N sources
   (opts->bank_init == NRAPS_BANK_INIT_DIFFUSION) and banks sites with
   1/init_k. */
int oracle_solve_init(const nraps_problem *pb, const nraps_run_opts *opts, int threads,
                      nraps_result *res, const double *init_x, const int32_t *init_cell,
                      double init_k);

/* Multigroup diffusion k-eigenvalue, src/discrete.rs:169-356 (diffusion.c). */
int oracle_diffusion(const nraps_problem *pb, const nraps_diffusion_opts *opts,
                     nraps_diffusion_result *res);
/* The tridiagonal matrix of group g (src/discrete.rs:4-101): sub[c], diag[c],
   sup[c] are the entries (c, c-1), (c, c), (c, c+1). */
void oracle_diffusion_matrix(const nraps_problem *pb, int32_t g, double *sub, double *diag,
                             double *sup);
/* N generation-0 sources at the quantiles (p + 1/2)/N of a per-cell fission
   source (weights s_c dx_c). */
int oracle_diffusion_bank(const nraps_problem *pb, const double *fission_source, int64_t N,
                          double *x, int32_t *cell);

/* Reference helper functions, restated (for the reference's own KATs). */
void oracle_hit_boundary(double mu, double start_x, double delta_s, double bound,
                         double mesh_end, double out[3]);     /* src/mc_code.rs:55-62 */
void oracle_cross_mesh(int64_t mesh_index, double mu, double start_x,
                       double mesh_end, double delta_s, double out_f[2],
                       int64_t *out_index);                   /* src/mc_code.rs:64-79 */
double oracle_direction(double u);                            /* src/mc_code.rs:34-37 */
void oracle_scat_mat_calc(int32_t energygroups, int32_t matid, int32_t neutron_energy,
                          double inv_sigs, const double *scat_matrix,
                          double *out);                       /* src/mc_code.rs:81-111 */
int32_t oracle_cdf_index(const double *cdf, int32_t n, double u); /* partition_point().min(n-1) */

/* PCG32, written independently of the product (src/rand.rs:36-107). */
void oracle_pcg32_stream(uint64_t seed, uint64_t stream, int32_t n, uint32_t *out);
uint64_t oracle_stream_id(uint64_t problem, uint64_t cycle, uint64_t p);
/* SplitMix64 finalizer and the draws of the generator of history `id`
   (initial state seed ^ mix(id), stream mix(id + golden)). */
uint64_t oracle_mix64(uint64_t z);
void oracle_history_stream(uint64_t seed, uint64_t id, int32_t n, uint32_t *out);

/* Canonical fixed-order sum used for the k estimator: 256-wide pairwise
   trees, level by level, zero padded (also implemented on the device). */
double oracle_canonical_sum(const double *v, int64_t n);

/* The shared logarithm, exported for its accuracy test. */
double oracle_log(double x);
void oracle_log_many(const double *x, double *y, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
