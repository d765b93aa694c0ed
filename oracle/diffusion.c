/*
 * diffusion.c — CPU FP64 restatement of the reference's finite-difference
 * multigroup diffusion solver (nalgebra_method, src/discrete.rs:169-356).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker for the device
 * diffusion solver and for the diffusion-started fission bank.
 *
 * Reference anchors (paths are inside the jurskidl/NRAPS tree):
 *   matrix_gen ........ src/discrete.rs:4-101
 *   q_gen ............. src/discrete.rs:103-126
 *   scat_calc ......... src/discrete.rs:128-149
 *   average_assembly .. src/discrete.rs:151-167
 *   power iteration ... src/discrete.rs:169-291
 *   normalisation ..... src/discrete.rs:293-343
 *
 * The reference inverts each group matrix densely (try_inverse, :201) and
 * forms A^-1 rhs row by row (:240-284).  A is tridiagonal, so this
 * restatement solves A x = rhs by Gaussian elimination without pivoting
 * (the Thomas algorithm): the same vector, with less rounding.
 *
 * Kept exactly as the reference has them:
 *   - D/dx = (1/3) inv_sigtr / dx per cell and the harmonic-mean coupling
 *     2 D_i D_j / (D_i + D_j) of neighbours (:18-24, :52-63);
 *   - beta = 1 for bound == 1, 0.25 for bound == 0, otherwise the expression
 *     of :26-33, whose right-hand form takes D/dx of cell 1 (:81-88);
 *   - the last diagonal adds the coupling of cells 0 and 1 (:93-99);
 *   - removal dx (sigt - S[mat][g][g]) (:36-41);
 *   - the in-group sweep in group order, each group using the newest flux
 *     of the others (:208-284);
 *   - the flux-change measure: max with the old value at cell 0, overwritten
 *     by every inner cell, max again at the last cell, all but the first
 *     divided by the new flux of cell 0 (:252-283);
 *   - the k update and change (:286-289), and the loop condition of :207;
 *   - the power constant over chunks of G of the group-major nut*sigf*flux
 *     array, weighted by dx of the chunk index (:315-330).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* D/dx of cell c in group g (src/discrete.rs:18-20). */
static double od(const nraps_problem *pb, const double *inv_sigtr, int c, int g) {
  return (1.0 / 3.0) * inv_sigtr[pb->matid[c] + pb->n_mats * g] / pb->delta_x[c];
}

static double ocouple(double a, double b) { return (2.0 * a * b) / (a + b); }

static double obeta(double bound, double d_next, double d_curr) {
  if (bound == 1.0) return 1.0;
  if (bound == 0.0) return 0.25;
  double r = (1.0 - bound) / (1.0 + bound);
  return (1.0 - 0.25 * (r * (1.0 / d_next))) / (1.0 + 0.25 * (r * (1.0 / d_curr)));
}

static double *oinv_sigtr(const nraps_problem *pb) {
  int GM = pb->n_groups * pb->n_mats;
  double *v = malloc(sizeof(double) * GM);
  for (int i = 0; i < GM; i++)
    v[i] = pb->inv_sigtr ? pb->inv_sigtr[i] : 1.0 / (pb->sigt[i] - pb->mu[i] * pb->sigs[i]);
  return v;
}

static void omatrix(const nraps_problem *pb, const double *inv_sigtr, int g, double *sub,
                    double *diag, double *sup) {
  const int n = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  const double d0 = od(pb, inv_sigtr, 0, g), d1 = od(pb, inv_sigtr, 1, g);
  const double h01 = ocouple(d0, d1);
  for (int c = 0; c < n; c++) {
    int m = pb->matid[c];
    double rem = pb->delta_x[c] * (pb->sigt[m + M * g] - pb->scat_matrix[G * G * m + (G + 1) * g]);
    double dc = od(pb, inv_sigtr, c, g);
    double hl = c > 0 ? ocouple(dc, od(pb, inv_sigtr, c - 1, g)) : 0.0;
    double hr = c < n - 1 ? ocouple(dc, od(pb, inv_sigtr, c + 1, g)) : 0.0;
    sub[c] = -hl;
    sup[c] = -hr;
    if (c == 0)
      diag[c] = 2.0 * dc * (1.0 - obeta(pb->boundl, d1, dc)) + rem + h01;
    else if (c == n - 1)
      diag[c] = 2.0 * dc * (1.0 - obeta(pb->boundr, d1, dc)) + rem + h01;
    else
      diag[c] = hl + rem + hr;
  }
}

void oracle_diffusion_matrix(const nraps_problem *pb, int32_t g, double *sub, double *diag,
                             double *sup) {
  double *ist = oinv_sigtr(pb);
  omatrix(pb, ist, g, sub, diag, sup);
  free(ist);
}

/* Thomas solve of (sub, diag, sup) x = rhs; w is n scratch. */
static void othomas(int n, const double *sub, const double *diag, const double *sup,
                    const double *rhs, double *x, double *w) {
  double m = 1.0 / diag[0];
  w[0] = sup[0] * m;
  x[0] = rhs[0] * m;
  for (int i = 1; i < n; i++) {
    m = 1.0 / (diag[i] - sub[i] * w[i - 1]);
    w[i] = sup[i] * m;
    x[i] = (rhs[i] - sub[i] * x[i - 1]) * m;
  }
  for (int i = n - 2; i >= 0; i--) x[i] = x[i] - w[i] * x[i + 1];
}

/* q_gen (src/discrete.rs:103-126): q[g][c] = (sum_x nut sigf flux) dx chi. */
static void oqgen(const nraps_problem *pb, const double *flux, double *q) {
  const int n = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  for (int g = 0; g < G; g++)
    for (int c = 0; c < n; c++) {
      int m = pb->matid[c];
      double f = 0.0;
      for (int x = 0; x < G; x++) f += pb->nut[m + M * x] * pb->sigf[m + M * x] * flux[x * n + c];
      q[g * n + c] = f * pb->delta_x[c] * pb->chit[m + M * g];
    }
}

int oracle_diffusion(const nraps_problem *pb, const nraps_diffusion_opts *o,
                     nraps_diffusion_result *res) {
  const int n = pb->n_cells, G = pb->n_groups, M = pb->n_mats;
  if (n < 2 || G < 1 || M < 1) return NRAPS_E_INVALID;
  const double ftol = o && o->flux_tol > 0 ? o->flux_tol : 1e-5;
  const double ktol = o && o->k_tol > 0 ? o->k_tol : 1e-6;
  const int maxit = o && o->max_iterations > 0 ? o->max_iterations : 1000000;
  double *ist = oinv_sigtr(pb);
  double *sub = malloc(sizeof(double) * G * n), *diag = malloc(sizeof(double) * G * n);
  double *sup = malloc(sizeof(double) * G * n);
  for (int g = 0; g < G; g++) omatrix(pb, ist, g, sub + g * n, diag + g * n, sup + g * n);
  double *flux = malloc(sizeof(double) * G * n), *q = malloc(sizeof(double) * G * n);
  double *qold = malloc(sizeof(double) * G * n);
  double *rhs = malloc(sizeof(double) * n), *x = malloc(sizeof(double) * n);
  double *w = malloc(sizeof(double) * n);
  for (int i = 0; i < G * n; i++) flux[i] = 1.0;
  oqgen(pb, flux, q);
  double k = 1.0, dflux = 1.0, dk = 1.0;
  int it = 0, status = NRAPS_OK;
  while (dflux >= ftol && dk >= ktol) {
    if (it >= maxit) {
      status = NRAPS_E_NOT_CONVERGED;
      break;
    }
    it++;
    memcpy(qold, q, sizeof(double) * G * n);
    for (int g = 0; g < G; g++) {
      /* scat_calc (src/discrete.rs:128-149) and the right-hand side */
      for (int c = 0; c < n; c++) {
        int m = pb->matid[c];
        double s = 0.0;
        for (int e = 0; e < G; e++)
          if (e != g) s += pb->scat_matrix[G * G * m + G * e + g] * flux[e * n + c] * pb->delta_x[c];
        rhs[c] = q[g * n + c] * (1.0 / k) + s;
      }
      othomas(n, sub + g * n, diag + g * n, sup + g * n, rhs, x, w);
      double *f = flux + g * n;
      dflux = fmax(fabs((f[0] - x[0]) / f[0]), dflux);
      f[0] = x[0];
      for (int c = 1; c < n - 1; c++) {
        dflux = fabs((f[c] - x[c]) / f[0]);
        f[c] = x[c];
      }
      dflux = fmax(fabs((f[n - 1] - x[n - 1]) / f[0]), dflux);
      f[n - 1] = x[n - 1];
    }
    oqgen(pb, flux, q);
    double sq = 0.0, sqo = 0.0;
    for (int i = 0; i < G * n; i++) sq += q[i];
    for (int i = 0; i < G * n; i++) sqo += qold[i];
    double kold = k;
    k = kold * (sq / sqo);
    dk = fabs((k - kold) / kold);
  }
  /* power normalisation (src/discrete.rs:293-343) */
  double *t = malloc(sizeof(double) * G * n);
  for (int g = 0; g < G; g++)
    for (int c = 0; c < n; c++) {
      int m = pb->matid[c];
      t[g * n + c] = flux[g * n + c] * pb->nut[m + M * g] * pb->sigf[m + M * g];
    }
  double P = 0.0;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int u = 0; u < G; u++) s += t[j * G + u];
    P += s * pb->delta_x[j];
  }
  double pc = 3565e6 / (1.6022e-13 * 200.0 * P);
  for (int i = 0; i < G * n; i++) flux[i] *= pc;
  if (res->flux) memcpy(res->flux, flux, sizeof(double) * G * n);
  if (res->fission_source)
    for (int c = 0; c < n; c++) {
      int m = pb->matid[c];
      double s = 0.0;
      for (int g = 0; g < G; g++) s += pb->nut[m + M * g] * pb->sigf[m + M * g] * flux[g * n + c];
      res->fission_source[c] = s;
    }
  if (res->assembly_average) {
    memset(res->assembly_average, 0, sizeof(double) * G * n);
    int na = pb->numass > 0 ? pb->numass : 1;
    int per = n / na;
    for (int g = 0; g < G && per > 0; g++)
      for (int a = 0; a < na; a++) {
        double s = 0.0;
        for (int c = a * per; c < (a + 1) * per; c++) s += flux[g * n + c];
        double avg = s / (double)per;
        for (int c = a * per; c < (a + 1) * per; c++) res->assembly_average[g * n + c] = avg;
      }
  }
  res->k = k;
  res->delta_flux = dflux;
  res->delta_k = dk;
  res->iterations = it;
  res->status = status;
  free(t);
  free(ist);
  free(sub);
  free(diag);
  free(sup);
  free(flux);
  free(q);
  free(qold);
  free(rhs);
  free(x);
  free(w);
  return status;
}

/* Generation-0 sources from a per-cell fission source: cell weights
   s_c dx_c summed in cell order, source p at the quantile (p + 1/2)/N of
   the total, placed linearly inside its cell and clamped to its right edge. */
int oracle_diffusion_bank(const nraps_problem *pb, const double *fission_source, int64_t N,
                          double *x, int32_t *cell) {
  const int n = pb->n_cells;
  double *cdf = malloc(sizeof(double) * n);
  double acc = 0.0;
  for (int c = 0; c < n; c++) {
    double s = fission_source[c];
    acc = acc + (s > 0.0 ? s * pb->delta_x[c] : 0.0);
    cdf[c] = acc;
  }
  if (!(acc > 0.0)) {
    free(cdf);
    return NRAPS_E_NO_FISSION;
  }
  for (int64_t p = 0; p < N; p++) {
    double t = (((double)p + 0.5) / (double)N) * acc;
    int lo = 0, hi = n - 1; /* first c with cdf[c] > t */
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (cdf[mid] > t) hi = mid;
      else lo = mid + 1;
    }
    int c = lo;
    double below = c > 0 ? cdf[c - 1] : 0.0;
    double w = cdf[c] - below;
    double frac = w > 0.0 ? (t - below) / w : 0.5;
    double xx = pb->mesh_left[c] + frac * pb->delta_x[c];
    x[p] = xx > pb->mesh_right[c] ? pb->mesh_right[c] : xx;
    cell[p] = c;
  }
  free(cdf);
  return NRAPS_OK;
}
