/*
 * oracle.c — CPU FP64 restatement of the NRAPS Monte Carlo history loop.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the HIP product.
 *
 * Reference anchors (paths are inside the jurskidl/NRAPS tree):
 *   PCG32 ............................ src/rand.rs:36-107
 *   energy (chi CDF) ................. src/mc_code.rs:6-32
 *   direction ........................ src/mc_code.rs:34-37
 *   spawn_neutron .................... src/mc_code.rs:39-53, :228-230
 *   hit_boundary ..................... src/mc_code.rs:55-62
 *   cross_mesh ....................... src/mc_code.rs:64-79
 *   scat_mat_calc .................... src/mc_code.rs:81-111
 *   interaction ...................... src/mc_code.rs:113-132
 *   particle_travel .................. src/mc_code.rs:134-213
 *   particle_lifetime ................ src/mc_code.rs:215-257
 *   average_assembly ................. src/mc_code.rs:259-274
 *   monte_carlo (generation loop) .... src/mc_code.rs:276-380
 *   inv_sigtr ........................ src/process_input.rs:152-156
 *   fuel_indices ..................... src/main.rs:110-115
 *
 * Deviations from the reference, shared with the product:
 *   - one uniform per scatter-group inversion (the reference draws a fresh
 *     random() at every binary-search probe, src/mc_code.rs:124-126);
 *   - exactly N histories per generation (the reference's inclusive ranges
 *     run N+T, src/mc_code.rs:226);
 *   - a flight draw of 0 is replaced by 2^-32; a zero-length segment scores
 *     0 (the reference scores |0/0| = NaN when mu == 0);
 *   - 32-bit table indices (the reference computes them in u8);
 *   - the birth position is clamped to the cell's right edge;
 *   - the source cell is drawn by one multiply-shift draw.
 * New behaviour (not in the reference): region-by-region surface tracking
 * (ohistory_region: the reference's tracking with the cell crossings inside a
 * material region merged; it draws and scores like ohistory_surface up to
 * rounding, which tests/test_region.py checks history by history), Woodcock
 * tracking, the fission-site bank, per-history k accumulation with a
 * canonical summation order.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "../nraps_amd/csrc/nraps_core.h"

/* ------------------------------------------------------------------------ */
/* PCG32, independent restatement of src/rand.rs                             */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t state, inc;
  int64_t draws;
} orng;

static uint32_t orng_u32(orng *r) {
  uint64_t s = r->state;
  r->state = s * 6364136223846793005ULL + r->inc;
  uint32_t v = (uint32_t)(((s >> 18) ^ s) >> 27);
  unsigned rot = (unsigned)(s >> 59);
  r->draws++;
  return rot ? (v >> rot) | (v << (32u - rot)) : v;
}

static void orng_init(orng *r, uint64_t seed, uint64_t stream) {
  r->state = 0;
  r->inc = stream * 2u + 1u;
  orng_u32(r);
  r->state += seed;
  orng_u32(r);
  r->draws = 0;
}

static double orng_u(orng *r) { return (double)orng_u32(r) / 4294967296.0; }
static double orng_flight(orng *r) {
  uint32_t u = orng_u32(r);
  return (u == 0 ? 1.0 : (double)u) / 4294967296.0;
}

uint64_t oracle_stream_id(uint64_t problem, uint64_t cycle, uint64_t p) {
  return problem * (1ULL << 44) + cycle * (1ULL << 28) + p;
}

/* SplitMix64 finalizer (Steele, Lea & Flood 2014; Stafford's Mix13 constants). */
uint64_t oracle_mix64(uint64_t z) {
  z ^= z >> 30;
  z *= 0xBF58476D1CE4E5B9ULL;
  z ^= z >> 27;
  z *= 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* Generator of history `id`: state and stream both scrambled from the id. */
static void orng_history(orng *r, uint64_t seed, uint64_t id) {
  orng_init(r, seed ^ oracle_mix64(id), oracle_mix64(id + 0x9E3779B97F4A7C15ULL));
}

void oracle_pcg32_stream(uint64_t seed, uint64_t stream, int32_t n, uint32_t *out) {
  orng r;
  orng_init(&r, seed, stream);
  for (int32_t i = 0; i < n; i++) out[i] = orng_u32(&r);
}

void oracle_history_stream(uint64_t seed, uint64_t id, int32_t n, uint32_t *out) {
  orng r;
  orng_history(&r, seed, id);
  for (int32_t i = 0; i < n; i++) out[i] = orng_u32(&r);
}

double oracle_log(double x) { return nraps_log(x); }
void oracle_log_many(const double *x, double *y, int64_t n) {
  for (int64_t i = 0; i < n; i++) y[i] = nraps_log(x[i]);
}

/* ------------------------------------------------------------------------ */
/* Reference helpers, restated                                               */
/* ------------------------------------------------------------------------ */
void oracle_hit_boundary(double mu, double start_x, double delta_s, double bound,
                         double mesh_end, double out[3]) {
  out[0] = mu * (-bound);
  out[1] = (delta_s + (start_x - mesh_end)) * (-bound);
  out[2] = mesh_end;
}

void oracle_cross_mesh(int64_t mesh_index, double mu, double start_x, double mesh_end,
                       double delta_s, double out_f[2], int64_t *out_index) {
  *out_index = mu >= 0.0 ? mesh_index + 1 : mesh_index - 1;
  out_f[0] = delta_s + (start_x - mesh_end);
  out_f[1] = mesh_end;
}

double oracle_direction(double u) { return 2.0 * u - 1.0; }

void oracle_scat_mat_calc(int32_t G, int32_t matid, int32_t g, double inv_sigs,
                          const double *scat, double *out) {
  int64_t base = (int64_t)G * G * matid + (int64_t)G * g;
  double cum = 0.0;
  for (int32_t j = 0; j < G; j++) {
    cum += scat[base + j];
    out[j] = cum * inv_sigs;
  }
}

/* slice.partition_point(|&x| x < u).min(n - 1) for a CDF: the first index
   whose value is not below u, clamped to the last index. */
int32_t oracle_cdf_index(const double *cdf, int32_t n, double u) {
  int32_t i = 0;
  while (i < n && cdf[i] < u) i++;
  return i < n - 1 ? i : n - 1;
}

/* ------------------------------------------------------------------------ */
/* Prepared problem                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nc, G, M, numass;
  double boundl, boundr, XL, XR;
  const int32_t *mat;
  const double *dx, *left, *right;
  int nf;
  int32_t *fuel;
  double *inv_sigtr, *sigw, *abs_ratio, *inv_sigs, *nusigf, *chicdf;
  const double *scat;
  double maj;
  int R;
  int32_t *reg_of, *reg_mat, *reg_first; /* reg_first[r]: first cell, [R] = nc */
  double *reg_lo, *reg_hi, *kreg; /* kreg[g*(R+1) + r] */
} oprob;

static void oprob_free(oprob *p) {
  free(p->fuel);
  free(p->inv_sigtr);
  free(p->sigw);
  free(p->abs_ratio);
  free(p->inv_sigs);
  free(p->nusigf);
  free(p->chicdf);
  free(p->reg_of);
  free(p->reg_mat);
  free(p->reg_lo);
  free(p->reg_hi);
  free(p->reg_first);
  free(p->kreg);
}

static int oprob_init(oprob *p, const nraps_problem *pb) {
  memset(p, 0, sizeof *p);
  int nc = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  if (nc < 1 || G < 1 || M < 1 || !pb->matid || !pb->delta_x || !pb->mesh_left ||
      !pb->mesh_right || !pb->sigt || !pb->sigs || !pb->mu || !pb->siga || !pb->sigf ||
      !pb->nut || !pb->chit || !pb->scat_matrix)
    return -1;
  p->nc = nc;
  p->G = G;
  p->M = M;
  p->numass = pb->numass;
  p->boundl = pb->boundl;
  p->boundr = pb->boundr;
  p->mat = pb->matid;
  p->dx = pb->delta_x;
  p->left = pb->mesh_left;
  p->right = pb->mesh_right;
  p->scat = pb->scat_matrix;
  p->XL = pb->mesh_left[0];
  p->XR = pb->mesh_right[nc - 1];
  for (int c = 0; c < nc; c++)
    if (pb->matid[c] < 0 || pb->matid[c] >= M) return -1;
  /* fuel cells: given, or matid 0/1 (src/main.rs:110-115) */
  if (pb->fuel_indices) {
    p->nf = pb->n_fuel;
    p->fuel = malloc(sizeof(int32_t) * (p->nf > 0 ? p->nf : 1));
    memcpy(p->fuel, pb->fuel_indices, sizeof(int32_t) * p->nf);
  } else {
    p->fuel = malloc(sizeof(int32_t) * nc);
    for (int c = 0; c < nc; c++)
      if (pb->matid[c] == 0 || pb->matid[c] == 1) p->fuel[p->nf++] = c;
  }
  int GM = G * M;
  p->inv_sigtr = malloc(sizeof(double) * GM);
  p->sigw = malloc(sizeof(double) * GM);
  p->abs_ratio = malloc(sizeof(double) * GM);
  p->inv_sigs = malloc(sizeof(double) * GM);
  p->nusigf = malloc(sizeof(double) * GM);
  p->chicdf = malloc(sizeof(double) * GM);
  for (int i = 0; i < GM; i++) {
    p->inv_sigtr[i] = pb->inv_sigtr ? pb->inv_sigtr[i]
                                    : 1.0 / (pb->sigt[i] - pb->mu[i] * pb->sigs[i]);
    p->sigw[i] = 1.0 / p->inv_sigtr[i];
    p->abs_ratio[i] = pb->siga[i] / pb->sigt[i];
    p->inv_sigs[i] = 1.0 / pb->sigs[i];
    p->nusigf[i] = pb->nut[i] * pb->sigf[i];
  }
  for (int m = 0; m < M; m++) {
    double cum = 0.0;
    for (int g = 0; g < G; g++) {
      cum += pb->chit[m + M * g];
      p->chicdf[m * G + g] = cum;
    }
  }
  /* majorant over the materials that occur in the mesh */
  p->maj = 0.0;
  for (int c = 0; c < nc; c++)
    for (int g = 0; g < G; g++) {
      double s = p->sigw[pb->matid[c] + M * g];
      if (s > p->maj) p->maj = s;
    }
  /* regions: maximal runs of one material */
  p->reg_of = malloc(sizeof(int32_t) * nc);
  int R = 0;
  for (int c = 0; c < nc; c++) {
    if (c > 0 && pb->matid[c] != pb->matid[c - 1]) R++;
    p->reg_of[c] = R;
  }
  R++;
  p->R = R;
  p->reg_mat = malloc(sizeof(int32_t) * R);
  p->reg_lo = malloc(sizeof(double) * R);
  p->reg_hi = malloc(sizeof(double) * R);
  p->reg_first = malloc(sizeof(int32_t) * (R + 1));
  p->reg_first[R] = nc;
  for (int c = 0; c < nc; c++) {
    int r = p->reg_of[c];
    if (c == 0 || p->reg_of[c - 1] != r) {
      p->reg_lo[r] = pb->mesh_left[c];
      p->reg_mat[r] = pb->matid[c];
      p->reg_first[r] = c;
    }
    p->reg_hi[r] = pb->mesh_right[c];
  }
  p->kreg = malloc(sizeof(double) * G * (R + 1));
  for (int g = 0; g < G; g++) {
    double acc = 0.0;
    p->kreg[g * (R + 1)] = 0.0;
    for (int r = 0; r < R; r++) {
      acc = acc + p->nusigf[p->reg_mat[r] + M * g] * (p->reg_hi[r] - p->reg_lo[r]);
      p->kreg[g * (R + 1) + r + 1] = acc;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Per-history context                                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  double x;
  int32_t cell;
} osite;

typedef struct {
  osite *v;
  int64_t n, cap;
} osites;

static void osites_push(osites *s, double x, int32_t cell) {
  if (s->n == s->cap) {
    s->cap = s->cap ? s->cap * 2 : 1024;
    s->v = realloc(s->v, sizeof(osite) * s->cap);
  }
  s->v[s->n].x = x;
  s->v[s->n].cell = cell;
  s->n++;
}

typedef struct {
  int64_t collisions, virt, crossings, absorptions, leaks, capped, steps;
} ostats;

typedef struct {
  const oprob *p;
  const nraps_run_opts *o;
  int cycle;
  int from_bank;   /* birth from src bank */
  int make_sites;  /* bank fission sites */
  double inv_k;
  const double *src_x;
  const int32_t *src_cell;
} ocycle;

typedef struct {
  double *tally; /* [G*nc] */
  osites sites;
  ostats st;
  /* trace for this history */
  nraps_trace_event *tr;
  int32_t tr_cap, tr_len;
} owork;

static void otrace(owork *w, int type, int cell, int g, int aux, double x, double mu) {
  if (w->tr && w->tr_len < w->tr_cap) {
    nraps_trace_event *e = &w->tr[w->tr_len];
    e->type = type;
    e->cell = cell;
    e->group = g;
    e->aux = aux;
    e->x = x;
    e->mu = mu;
  }
  w->tr_len++;
}

/* Birth: src/mc_code.rs:39-53 (uniform fuel source) or a bank site. */
static void obirth(const ocycle *cy, int64_t pidx, orng *r, double *x, double *mu,
                   int32_t *cell, int32_t *g) {
  const oprob *p = cy->p;
  if (!cy->from_bank) {
    uint32_t u = orng_u32(r);
    int32_t c = p->fuel[nraps_uindex(u, (uint32_t)p->nf)];
    double frac = orng_u(r);
    *mu = oracle_direction(orng_u(r));
    double ue = orng_u(r);
    *cell = c;
    double xx = p->left[c] + frac * p->dx[c];
    *x = xx > p->right[c] ? p->right[c] : xx;
    *g = oracle_cdf_index(&p->chicdf[p->mat[c] * p->G], p->G, ue);
  } else {
    *mu = oracle_direction(orng_u(r));
    double ue = orng_u(r);
    *cell = cy->src_cell[pidx];
    *x = cy->src_x[pidx];
    *g = oracle_cdf_index(&p->chicdf[p->mat[*cell] * p->G], p->G, ue);
  }
}

/* Collision outcome, src/mc_code.rs:113-132 with the fission-site bank
   added after the reference's three draws.  Returns 1 if the particle
   survives (scattered). */
static int ocollide(const ocycle *cy, owork *w, orng *r, int64_t pidx, int xs, int cell,
                    double x, int32_t *g, double *mu) {
  const oprob *p = cy->p;
  int G = p->G, M = p->M, m = p->mat[cell];
  double u_int = orng_u(r);
  double mu_new = 2.0 * orng_u(r) - 1.0;
  double u_scat = orng_u(r);
  if (cy->make_sites) {
    double nsf = p->nusigf[m + M * (*g)];
    if (nsf > 0.0) {
      double y = nsf * p->inv_sigtr[xs] * cy->inv_k + orng_u(r);
      int n = y < 1000.0 ? (int)y : 1000;
      for (int i = 0; i < n; i++) osites_push(&w->sites, x, cell);
    }
  }
  (void)pidx;
  w->st.collisions++;
  if (u_int < p->abs_ratio[xs]) {
    w->st.absorptions++;
    otrace(w, NRAPS_EV_ABSORB, cell, *g, *g, x, *mu);
    return 0;
  }
  double cdf[64];
  oracle_scat_mat_calc(G, m, *g, p->inv_sigs[xs], p->scat, cdf);
  int32_t gn = oracle_cdf_index(cdf, G, u_scat);
  otrace(w, NRAPS_EV_SCATTER, cell, *g, gn, x, *mu);
  *g = gn;
  *mu = mu_new;
  return 1;
}

static double osegment(double a, double b, double mu) {
  /* |(a - b)/mu|, zero for a zero-length segment */
  return a == b ? 0.0 : fabs((a - b) / mu);
}

/* One history with cell-by-cell surface tracking: particle_lifetime +
   particle_travel, src/mc_code.rs:134-257. */
static double ohistory_surface(const ocycle *cy, owork *w, int64_t pidx, orng *r) {
  const oprob *p = cy->p;
  const int G = p->G, M = p->M, nc = p->nc;
  int xs_every = cy->o->xs_refresh == NRAPS_XS_EVERY_COLLISION;
  int64_t cap = cy->o->max_events > 0 ? cy->o->max_events : 10000000;
  double x, mu;
  int32_t cell, g;
  double kacc = 0.0;
  obirth(cy, pidx, r, &x, &mu, &cell, &g);
  otrace(w, NRAPS_EV_BIRTH, cell, g, 0, x, mu);
  int64_t events = 0;
  for (;;) { /* particle_travel entries */
    int xs = p->mat[cell] + M * g;
    double delta_s = mu * -nraps_log(orng_flight(r)) * p->inv_sigtr[xs];
    int same = 1;
    while (same) {
      if (++events > cap) {
        w->st.capped++;
        otrace(w, NRAPS_EV_CAP, cell, g, 0, x, mu);
        return kacc;
      }
      double end_x = x + delta_s;
      double mesh_end = mu >= 0.0 ? p->right[cell] : p->left[cell];
      if ((mu < 0.0 && mesh_end > end_x && cell == 0) ||
          (mu >= 0.0 && end_x > mesh_end && cell == nc - 1)) {
        double tl = osegment(x, mesh_end, mu);
        w->tally[g * nc + cell] += tl;
        kacc += p->nusigf[p->mat[cell] + M * g] * tl;
        double bound = mu >= 0.0 ? p->boundr : p->boundl;
        w->st.crossings++;
        if (bound > 0.0) {
          double o3[3];
          oracle_hit_boundary(mu, x, delta_s, bound, mesh_end, o3);
          mu = o3[0];
          delta_s = o3[1];
          x = o3[2];
        } else {
          w->st.leaks++;
          otrace(w, NRAPS_EV_LEAK, cell, g, 0, mesh_end, mu);
          return kacc;
        }
      } else if (fabs(end_x - x) > fabs(mesh_end - x)) {
        double tl = osegment(x, mesh_end, mu);
        w->tally[g * nc + cell] += tl;
        kacc += p->nusigf[p->mat[cell] + M * g] * tl;
        int prev = p->mat[cell];
        double o2[2];
        int64_t ni;
        oracle_cross_mesh(cell, mu, x, mesh_end, delta_s, o2, &ni);
        delta_s = o2[0];
        x = o2[1];
        cell = (int32_t)ni;
        w->st.crossings++;
        if (prev != p->mat[cell]) same = 0;
      } else {
        double tl = osegment(x, end_x, mu);
        w->tally[g * nc + cell] += tl;
        kacc += p->nusigf[p->mat[cell] + M * g] * tl;
        if (!ocollide(cy, w, r, pidx, xs, cell, end_x, &g, &mu)) return kacc;
        x = end_x;
        if (xs_every) xs = p->mat[cell] + M * g;
        delta_s = mu * -nraps_log(orng_flight(r)) * p->inv_sigtr[xs];
      }
    }
  }
  (void)G;
}

/* Cell holding the end of a move: moving right (mu >= 0) a point on an edge
   belongs to the cell on its left, moving left to the cell on its right, as
   in the surface tracker's strict crossing test. */
static int32_t olocate(const oprob *p, double xe, double mu) {
  int32_t lo = 0, hi = p->nc - 1; /* largest c with left[c] < xe (<= moving left) */
  while (lo < hi) {
    int32_t mid = lo + (hi - lo + 1) / 2;
    int below = mu >= 0.0 ? p->left[mid] < xe : p->left[mid] <= xe;
    if (below)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* Track-length tally of a straight piece from (xa, ca) to (xb, cb): the
   overlap with every cell, divided by |mu| — computed cell by cell here,
   independently of the product's difference-array scheme. */
static void otally_piece(const oprob *p, owork *w, int g, double xa, int32_t ca, double xb,
                         int32_t cb, double mu) {
  if (xa == xb) return;
  double lo = xa < xb ? xa : xb, hi = xa < xb ? xb : xa;
  int32_t c0 = ca < cb ? ca : cb, c1 = ca < cb ? cb : ca;
  double am = fabs(mu);
  for (int32_t c = c0; c <= c1; c++) {
    double a = p->left[c] > lo ? p->left[c] : lo;
    double b = p->right[c] < hi ? p->right[c] : hi;
    if (b > a) w->tally[g * p->nc + c] += (b - a) / am;
  }
}

/* k-contribution integral at x in cell c for group g (region prefix). */
static double okint(const oprob *p, int g, double x, int32_t c) {
  int r = p->reg_of[c];
  return p->kreg[g * (p->R + 1) + r] + p->nusigf[p->reg_mat[r] + p->M * g] * (x - p->reg_lo[r]);
}

/* The straight path between real events (real collisions, boundary faces,
   the event cap) is one segment: virtual collisions change neither the
   direction nor the group.  A segment is scored where it ends: tally over
   its cells with weight 1/|mu|, and |K(end) - K(start)|/|mu| of k, where K
   is the nu*Sigma_f integral from the left face. */
typedef struct {
  double sx;
  int32_t scell;
} oseg;

static void oscore(const oprob *p, owork *w, oseg *s, int32_t g, double mu, double xe, int32_t ce,
                   double *kacc) {
  otally_piece(p, w, g, s->sx, s->scell, xe, ce, mu);
  if (s->sx != xe) *kacc += fabs(okint(p, g, xe, ce) - okint(p, g, s->sx, s->scell)) * (1.0 / fabs(mu));
  s->sx = xe;
  s->scell = ce;
}

/* One history with Woodcock delta tracking against the global majorant of
   the transport cross section.  Cross-section staleness follows the
   reference (src/mc_code.rs:147): the XS group g_xs is refreshed to the
   current group whenever the path has left the material region in which
   it was last refreshed. */
static double ohistory_woodcock(const ocycle *cy, owork *w, int64_t pidx, orng *r) {
  const oprob *p = cy->p;
  const int M = p->M, nc = p->nc;
  int xs_every = cy->o->xs_refresh == NRAPS_XS_EVERY_COLLISION;
  int64_t cap = cy->o->max_events > 0 ? cy->o->max_events : 10000000;
  double inv_maj = 1.0 / p->maj;
  double x, mu;
  int32_t cell, g;
  double kacc = 0.0;
  obirth(cy, pidx, r, &x, &mu, &cell, &g);
  otrace(w, NRAPS_EV_BIRTH, cell, g, 0, x, mu);
  int32_t gxs = g, rxs = p->reg_of[cell];
  oseg seg = {x, cell};
  int64_t events = 0;
  for (;;) {
    if (++events > cap) {
      oscore(p, w, &seg, g, mu, x, cell, &kacc);
      w->st.capped++;
      otrace(w, NRAPS_EV_CAP, cell, g, 0, x, mu);
      return kacc;
    }
    w->st.steps++;
    double disp = mu * -nraps_log(orng_flight(r)) * inv_maj;
    int left_region = 0;
    for (;;) {
      double end = x + disp;
      if (mu >= 0.0 && end > p->XR) {
        oscore(p, w, &seg, g, mu, p->XR, nc - 1, &kacc);
        if (p->reg_of[nc - 1] != rxs) left_region = 1;
        w->st.crossings++;
        if (p->boundr > 0.0) {
          double o3[3];
          oracle_hit_boundary(mu, x, disp, p->boundr, p->XR, o3);
          mu = o3[0];
          disp = o3[1];
          x = o3[2];
          cell = nc - 1;
          continue;
        }
        w->st.leaks++;
        otrace(w, NRAPS_EV_LEAK, nc - 1, g, 0, p->XR, mu);
        return kacc;
      } else if (mu < 0.0 && p->XL > end) {
        oscore(p, w, &seg, g, mu, p->XL, 0, &kacc);
        if (p->reg_of[0] != rxs) left_region = 1;
        w->st.crossings++;
        if (p->boundl > 0.0) {
          double o3[3];
          oracle_hit_boundary(mu, x, disp, p->boundl, p->XL, o3);
          mu = o3[0];
          disp = o3[1];
          x = o3[2];
          cell = 0;
          continue;
        }
        w->st.leaks++;
        otrace(w, NRAPS_EV_LEAK, 0, g, 0, p->XL, mu);
        return kacc;
      } else {
        if (end != x) {
          cell = olocate(p, end, mu);
          x = end;
        }
        break;
      }
    }
    if (!xs_every && (left_region || p->reg_of[cell] != rxs)) {
      gxs = g;
      rxs = p->reg_of[cell];
    }
    int xs = p->mat[cell] + M * gxs;
    double u_acc = orng_u(r);
    if (u_acc * p->maj < p->sigw[xs]) {
      oscore(p, w, &seg, g, mu, x, cell, &kacc);
      if (!ocollide(cy, w, r, pidx, xs, cell, x, &g, &mu)) return kacc;
      if (xs_every) {
        gxs = g;
        rxs = p->reg_of[cell];
      }
    } else {
      w->st.virt++;
    }
  }
}

/* One history with the reference's surface tracking taken region by region
   (particle_travel, src/mc_code.rs:134-213): the cell crossings inside a
   material region are merged, since they change neither the flight nor the
   XS index.  A flight is sampled after the birth, after a scatter (with
   the XS index the particle holds: stale after a group change) and on entry
   to another region (XS index refreshed to the current group,
   src/mc_code.rs:147).  Faces reflect with the remaining displacement
   (src/mc_code.rs:55-62) or leak.  The draws are those of the cell-by-cell
   tracker, in the same order; only the rounding of positions differs.
   Straight paths are scored where they end, as in ohistory_woodcock. */
static double ohistory_region(const ocycle *cy, owork *w, int64_t pidx, orng *r) {
  const oprob *p = cy->p;
  const int M = p->M, nc = p->nc, R = p->R;
  int xs_every = cy->o->xs_refresh == NRAPS_XS_EVERY_COLLISION;
  int64_t cap = cy->o->max_events > 0 ? cy->o->max_events : 10000000;
  double x, mu;
  int32_t cell, g;
  double kacc = 0.0;
  obirth(cy, pidx, r, &x, &mu, &cell, &g);
  otrace(w, NRAPS_EV_BIRTH, cell, g, 0, x, mu);
  int32_t rr = p->reg_of[cell];
  int xs = p->mat[cell] + M * g;
  oseg seg = {x, cell};
  double ds = mu * -nraps_log(orng_flight(r)) * p->inv_sigtr[xs];
  int64_t events = 0;
  for (;;) {
    if (++events > cap) {
      oscore(p, w, &seg, g, mu, x, cell, &kacc);
      w->st.capped++;
      otrace(w, NRAPS_EV_CAP, cell, g, 0, x, mu);
      return kacc;
    }
    w->st.steps++;
    double end = x + ds;
    int fwd = mu >= 0.0;
    /* region edges: left edge of the region's first cell; the right face */
    int er = fwd ? rr + 1 : rr;
    double edge = er < R ? p->left[p->reg_first[er]] : p->XR;
    if (fwd ? end > edge : edge > end) {
      w->st.crossings++;
      if (fwd ? rr == R - 1 : rr == 0) {
        int32_t fc = fwd ? nc - 1 : 0;
        oscore(p, w, &seg, g, mu, edge, fc, &kacc);
        double bound = fwd ? p->boundr : p->boundl;
        if (!(bound > 0.0)) {
          w->st.leaks++;
          otrace(w, NRAPS_EV_LEAK, fc, g, 0, edge, mu);
          return kacc;
        }
        double o3[3];
        oracle_hit_boundary(mu, x, ds, bound, edge, o3);
        mu = o3[0];
        ds = o3[1];
        x = o3[2];
        cell = fc;
        continue;
      }
      int32_t rn = fwd ? rr + 1 : rr - 1;
      x = edge;
      cell = fwd ? p->reg_first[rn] : p->reg_first[rr] - 1;
      rr = rn;
      xs = p->mat[cell] + M * g;
    } else {
      if (end != x) {
        cell = olocate(p, end, mu);
        x = end;
      }
      oscore(p, w, &seg, g, mu, x, cell, &kacc);
      if (!ocollide(cy, w, r, pidx, xs, cell, x, &g, &mu)) return kacc;
      if (xs_every) xs = p->mat[cell] + M * g;
    }
    ds = mu * -nraps_log(orng_flight(r)) * p->inv_sigtr[xs];
  }
}

/* ------------------------------------------------------------------------ */
/* Canonical k sum                                                           */
/* ------------------------------------------------------------------------ */
double oracle_canonical_sum(const double *v, int64_t n) {
  if (n <= 0) return 0.0;
  int64_t m = n;
  double *buf = malloc(sizeof(double) * (size_t)((n + 255) / 256 * 256));
  memcpy(buf, v, sizeof(double) * n);
  int first = 1;
  while (m > 1 || first) {
    first = 0;
    int64_t groups = (m + 255) / 256;
    for (int64_t gi = 0; gi < groups; gi++) {
      double s[256];
      for (int i = 0; i < 256; i++) {
        int64_t j = gi * 256 + i;
        s[i] = j < m ? buf[j] : 0.0;
      }
      for (int stride = 128; stride >= 1; stride >>= 1)
        for (int i = 0; i < stride; i++) s[i] = s[i] + s[i + stride];
      buf[gi] = s[0];
    }
    m = groups;
  }
  double out = buf[0];
  free(buf);
  return out;
}

/* ------------------------------------------------------------------------ */
/* Generation loop                                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
  const ocycle *cy;
  int64_t N, block, nblocks;
  double *block_tally;  /* [nblocks][G*nc] */
  osites *block_sites;  /* [nblocks] */
  ostats *block_stats;  /* [nblocks] */
  double *kacc;         /* [N] */
  atomic_llong next;
  /* trace */
  nraps_result *res;
  int trace_now;
} ojob;

static void orun_block(ojob *J, int64_t b) {
  const ocycle *cy = J->cy;
  const oprob *p = cy->p;
  owork w;
  memset(&w, 0, sizeof w);
  w.tally = &J->block_tally[(size_t)b * p->G * p->nc];
  memset(w.tally, 0, sizeof(double) * p->G * p->nc);
  int64_t lo = b * J->block, hi = lo + J->block < J->N ? lo + J->block : J->N;
  int method = cy->o->method;
  nraps_result *res = J->res;
  for (int64_t i = lo; i < hi; i++) {
    orng r;
    orng_history(&r, cy->o->seed, oracle_stream_id(cy->o->problem_id, cy->cycle, i));
    int traced = J->trace_now && i < cy->o->trace_histories;
    w.tr = traced && res->trace ? &res->trace[i * cy->o->trace_events] : NULL;
    w.tr_cap = traced ? cy->o->trace_events : 0;
    w.tr_len = 0;
    double k = method == NRAPS_EVENT_WOODCOCK ? ohistory_woodcock(cy, &w, i, &r)
               : method == NRAPS_EVENT_REGION ? ohistory_region(cy, &w, i, &r)
                                              : ohistory_surface(cy, &w, i, &r);
    J->kacc[i] = k;
    if (traced) {
      if (res->trace_len) res->trace_len[i] = w.tr_len;
      if (res->draws) res->draws[i] = r.draws;
      if (res->kacc) res->kacc[i] = k;
    }
  }
  J->block_sites[b] = w.sites;
  J->block_stats[b] = w.st;
}

static void *oworker(void *arg) {
  ojob *J = arg;
  for (;;) {
    int64_t b = atomic_fetch_add(&J->next, 1);
    if (b >= J->nblocks) break;
    orun_block(J, b);
  }
  return NULL;
}

static __thread char g_oerr[256];

static int oracle_solve_impl(const nraps_problem *pb, const nraps_run_opts *o, int threads,
                             nraps_result *res, int32_t bank_cycle, double *bank_x,
                             int32_t *bank_cell, int64_t *bank_n, const double *init_x,
                             const int32_t *init_cell, double init_k) {
  oprob P;
  if (oprob_init(&P, pb)) {
    snprintf(g_oerr, sizeof g_oerr, "bad problem");
    return NRAPS_E_INVALID;
  }
  const int G = P.G, nc = P.nc, M = P.M;
  const int64_t N = o->histories;
  const int gens = o->generations, skip = o->skip;
  int fission = o->source == NRAPS_SOURCE_FISSION_BANK;
  /* diffusion start: generation 0 is born from the given sources */
  int dstart = fission && o->bank_init == NRAPS_BANK_INIT_DIFFUSION;
  if (N < 1 || gens < 1 || skip < 1 || skip >= gens || P.nf < 1 ||
      (fission && !dstart && (skip < 2 || gens < 3)) ||
      (dstart && (!init_x || !init_cell || !(init_k > 0.0))) ||
      (!dstart && o->bank_init != NRAPS_BANK_INIT_UNIFORM_FUEL)) {
    oprob_free(&P);
    return NRAPS_E_INVALID;
  }
  if (threads <= 0) {
    long np = sysconf(_SC_NPROCESSORS_ONLN);
    threads = np > 1 ? (int)np - 1 : 1;
  }
  int64_t block = 4096;
  if (N / block > 4096) block = (N + 4095) / 4096;
  int64_t nblocks = (N + block - 1) / block;
  double *block_tally = malloc(sizeof(double) * (size_t)nblocks * G * nc);
  osites *block_sites = calloc((size_t)nblocks, sizeof(osites));
  ostats *block_stats = calloc((size_t)nblocks, sizeof(ostats));
  double *kacc = malloc(sizeof(double) * N);
  double *tally = malloc(sizeof(double) * G * nc);
  double *flux_acc = calloc((size_t)G * nc, sizeof(double));
  double *fs_acc = calloc((size_t)nc, sizeof(double));
  double *fsum = calloc((size_t)G * nc, sizeof(double));
  double *fsq = calloc((size_t)G * nc, sizeof(double));
  double *kk = calloc((size_t)gens, sizeof(double));
  double *src_x = NULL;
  int32_t *src_cell = NULL;
  if (fission) {
    src_x = malloc(sizeof(double) * N);
    src_cell = malloc(sizeof(int32_t) * N);
    if (dstart) {
      memcpy(src_x, init_x, sizeof(double) * N);
      memcpy(src_cell, init_cell, sizeof(int32_t) * N);
    }
  }
  int status = NRAPS_OK;
  double k_prev = o->k_initial;
  /* nut of material 0, group index 1 in the power conversion of
     src/mc_code.rs:353-357 (group 0 when there is a single group) */
  double nut_conv = pb->nut[0 + M * (G > 1 ? 1 : 0)];
  double fund = 1.0 / (double)(gens - (skip - 1));
  ostats tot;
  memset(&tot, 0, sizeof tot);

  for (int cyc = 0; cyc < gens; cyc++) {
    ocycle cy;
    cy.p = &P;
    cy.o = o;
    cy.cycle = cyc;
    cy.from_bank = fission && (dstart || cyc >= 2);
    cy.make_sites = fission && (dstart || cyc >= 1);
    cy.inv_k = !cy.make_sites ? 0.0 : cyc == 0 ? 1.0 / init_k : 1.0 / kk[cyc - 1];
    cy.src_x = src_x;
    cy.src_cell = src_cell;
    if (bank_x && cyc == bank_cycle && cy.from_bank) {
      memcpy(bank_x, src_x, sizeof(double) * N);
      memcpy(bank_cell, src_cell, sizeof(int32_t) * N);
      *bank_n = N;
    }
    ojob J;
    memset(&J, 0, sizeof J);
    J.cy = &cy;
    J.N = N;
    J.block = block;
    J.nblocks = nblocks;
    J.block_tally = block_tally;
    J.block_sites = block_sites;
    J.block_stats = block_stats;
    J.kacc = kacc;
    J.res = res;
    J.trace_now = cyc == o->trace_cycle && o->trace_histories > 0;
    atomic_init(&J.next, 0);
    int nt = threads < nblocks ? threads : (int)nblocks;
    pthread_t *th = malloc(sizeof(pthread_t) * nt);
    for (int t = 1; t < nt; t++) pthread_create(&th[t], NULL, oworker, &J);
    oworker(&J);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
    free(th);
    /* fixed-order reduction over blocks: independent of the thread count */
    memset(tally, 0, sizeof(double) * G * nc);
    int64_t nsites = 0;
    for (int64_t b = 0; b < nblocks; b++) {
      const double *bt = &block_tally[(size_t)b * G * nc];
      for (int i = 0; i < G * nc; i++) tally[i] += bt[i];
      nsites += block_sites[b].n;
      tot.collisions += block_stats[b].collisions;
      tot.virt += block_stats[b].virt;
      tot.crossings += block_stats[b].crossings;
      tot.absorptions += block_stats[b].absorptions;
      tot.leaks += block_stats[b].leaks;
      tot.capped += block_stats[b].capped;
    }
    double knew = oracle_canonical_sum(kacc, N) / (double)N;
    kk[cyc] = knew;
    if (res->k) res->k[cyc] = knew;
    if (res->sites) res->sites[cyc] = nsites;
    /* reference normalisation, src/mc_code.rs:340-363 */
    double kref = 0.0;
    for (int g = 0; g < G; g++)
      for (int c = 0; c < nc; c++) {
        double dxc = P.dx[c];
        int xi = P.mat[c] + M * g;
        double flux = tally[g * nc + c] / (k_prev * (double)N * dxc);
        double fsrc = pb->nut[xi] * pb->sigf[xi] * flux;
        kref += k_prev * dxc * fsrc;
        if (cyc >= skip) {
          double conversion =
              (3565e6 * k_prev * 36.2) / (200e6 * 1.602176634e-19 * nut_conv * P.right[nc - 1]);
          flux_acc[g * nc + c] += flux * conversion * fund;
          fs_acc[c] += fsrc * fund;
          double fv = tally[g * nc + c] / ((double)N * dxc);
          fsum[g * nc + c] += fv;
          fsq[g * nc + c] += fv * fv;
        }
      }
    if (res->k_ref) res->k_ref[cyc] = kref;
    if (cyc == gens - 1 && res->tally_last) memcpy(res->tally_last, tally, sizeof(double) * G * nc);
    k_prev = knew;
    /* fission bank: sites in (history, sequence) order; systematic
       resampling to N sources with one offset from the service stream */
    if (cy.make_sites) {
      if (nsites == 0) {
        status = NRAPS_E_NO_FISSION;
      } else {
        osite *all = malloc(sizeof(osite) * nsites);
        int64_t at = 0;
        for (int64_t b = 0; b < nblocks; b++) {
          memcpy(all + at, block_sites[b].v, sizeof(osite) * block_sites[b].n);
          at += block_sites[b].n;
        }
        orng rs;
        orng_history(&rs, o->seed, oracle_stream_id(o->problem_id, cyc, NRAPS_SERVICE_PARTICLE));
        double u0 = orng_u(&rs);
        double ratio = (double)nsites / (double)N;
        for (int64_t j = 0; j < N; j++) {
          int64_t idx = (int64_t)(((double)j + u0) * ratio);
          if (idx > nsites - 1) idx = nsites - 1;
          src_x[j] = all[idx].x;
          src_cell[j] = all[idx].cell;
        }
        free(all);
      }
    }
    for (int64_t b = 0; b < nblocks; b++) {
      free(block_sites[b].v);
      memset(&block_sites[b], 0, sizeof(osites));
    }
    if (status) break;
  }
  /* results */
  int nact = gens - skip;
  if (res->flux) memcpy(res->flux, flux_acc, sizeof(double) * G * nc);
  if (res->fission_source) memcpy(res->fission_source, fs_acc, sizeof(double) * nc);
  if (res->assembly_average) {
    memset(res->assembly_average, 0, sizeof(double) * G * nc);
    int na = P.numass > 0 ? P.numass : 1;
    int per = nc / na;
    for (int g = 0; g < G && per > 0; g++)
      for (int a = 0; a < na; a++) {
        double s = 0.0;
        for (int c = a * per; c < (a + 1) * per; c++) s += flux_acc[g * nc + c];
        double avg = s / (double)per;
        for (int c = a * per; c < (a + 1) * per; c++) res->assembly_average[g * nc + c] = avg;
      }
  }
  if (res->k_fund) {
    memset(res->k_fund, 0, sizeof(double) * gens);
    for (int gg = skip; gg < gens; gg++) {
      double s = 0.0;
      for (int x = skip; x <= gg; x++) s += kk[x];
      res->k_fund[gg] = s / (double)(gg - (skip - 1));
    }
  }
  for (int i = 0; i < G * nc; i++) {
    double mean = fsum[i] / nact;
    double var = nact > 1 ? (fsq[i] - fsum[i] * fsum[i] / nact) / (nact - 1) : 0.0;
    if (var < 0) var = 0;
    if (res->flux_mean) res->flux_mean[i] = mean;
    if (res->flux_sigma) res->flux_sigma[i] = sqrt(var / nact);
  }
  {
    double s = 0, s2 = 0;
    for (int x = skip; x < gens; x++) {
      s += kk[x];
      s2 += kk[x] * kk[x];
    }
    res->k_mean = s / nact;
    double var = nact > 1 ? (s2 - s * s / nact) / (nact - 1) : 0.0;
    res->k_sigma = var > 0 ? sqrt(var / nact) : 0.0;
  }
  res->histories_run = N * gens;
  res->collisions = tot.collisions;
  res->virtual_collisions = tot.virt;
  res->crossings = tot.crossings;
  res->absorptions = tot.absorptions;
  res->leaks = tot.leaks;
  res->capped = tot.capped;
  res->mean_flights = (double)(tot.collisions + tot.leaks) / (double)(N * gens);
  if (status == NRAPS_OK && tot.capped) status = NRAPS_E_EVENT_CAP;
  res->status = status;
  free(block_tally);
  free(block_sites);
  free(block_stats);
  free(kacc);
  free(tally);
  free(flux_acc);
  free(fs_acc);
  free(fsum);
  free(fsq);
  free(kk);
  free(src_x);
  free(src_cell);
  oprob_free(&P);
  return status;
}

int oracle_solve(const nraps_problem *pb, const nraps_run_opts *o, int threads,
                 nraps_result *res) {
  return oracle_solve_impl(pb, o, threads, res, -1, NULL, NULL, NULL, NULL, NULL, 0.0);
}

int oracle_solve_init(const nraps_problem *pb, const nraps_run_opts *o, int threads,
                      nraps_result *res, const double *init_x, const int32_t *init_cell,
                      double init_k) {
  return oracle_solve_impl(pb, o, threads, res, -1, NULL, NULL, NULL, init_x, init_cell, init_k);
}

int oracle_solve_banks(const nraps_problem *pb, const nraps_run_opts *o, int threads,
                       nraps_result *res, int32_t bank_cycle, double *bank_x,
                       int32_t *bank_cell, int64_t *bank_n) {
  *bank_n = 0;
  return oracle_solve_impl(pb, o, threads, res, bank_cycle, bank_x, bank_cell, bank_n, NULL,
                           NULL, 0.0);
}
