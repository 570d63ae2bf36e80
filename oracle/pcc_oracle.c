/*
 * pcc_oracle.c -- CPU oracle for the 1-D subsurface heat conduction of the
 * time-dependent thermal model.  TEST INFRASTRUCTURE ONLY: tests/ use it as
 * the checker; the product never calls it.
 *
 * Restates, in plain C (FP64, -ffp-contract=off, the reference's operation
 * order), for nf independent facets:
 *   - oracle_conduction_q   pcc/conductionQ.c:10-186 (radiative surface
 *                           boundary, Crank-Nicolson, predictor-corrector on
 *                           the linearised emission)
 *   - oracle_conduction_t   pcc/conductionT.c:4-80 (prescribed surface
 *                           temperature)
 *   - the tridiagonal solve pcc/tridag.c:5-34 (Thomas algorithm)
 * as thermal.pyx:218-246 calls them, one facet at a time.
 *
 * Layout: T is nf x (nz + 1) row-major (facet f's profile T[f (nz+1) + l],
 * l = 0 the surface), as thermal.pyx keeps it.  z, ti, rhoc have nz + 1
 * entries with z[0] = 0.  Where the reference exits the process (tridag's
 * zero pivots) these functions return a nonzero code instead.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#define O_SIGSB 5.6704e-8

/* Thomas algorithm on rows 1..n (index 0 unused): u = A^-1 r with
 * sub-diagonal a, diagonal b, super-diagonal c.  Returns 1 or 2 on a zero
 * pivot (the reference's two fatal errors), else 0.  gam has n + 1 slots. */
static int o_tridag(const double *a, const double *b, const double *c, const double *r,
                    double *u, int n, double *gam) {
  double bet = b[1];
  if (bet == 0.0) return 1;
  u[1] = r[1] / bet;
  for (int j = 2; j <= n; ++j) {
    gam[j] = c[j - 1] / bet;
    bet = b[j] - a[j] * gam[j];
    if (bet == 0.0) return 2;
    u[j] = (r[j] - a[j] * u[j - 1]) / bet;
  }
  for (int j = n - 1; j >= 1; --j) u[j] -= gam[j + 1] * u[j + 1];
  return 0;
}

/* Interior coefficients shared by both boundary conditions (rows 2..nz-1)
 * and the bottom row nz. */
static void o_interior(int nz, const double *z, double dt, const double *rhoc, const double *k,
                       double *alpha, double *gamma) {
  for (int i = 2; i < nz; ++i) {
    const double buf = dt / (z[i + 1] - z[i - 1]);
    alpha[i] = 2 * k[i + 1] * buf / (rhoc[i] + rhoc[i + 1]) / (z[i + 1] - z[i]);
    gamma[i] = 2 * k[i] * buf / (rhoc[i] + rhoc[i + 1]) / (z[i] - z[i - 1]);
  }
  const double buf = dt / (z[nz] - z[nz - 1]) / (z[nz] - z[nz - 1]);
  gamma[nz] = k[nz] * buf / (2 * rhoc[nz]);
}

/* Returns the first nonzero tridag code (0 if none).  iters[f] (optional)
 * receives the number of predictor-corrector repeats of facet f. */
int oracle_conduction_q(int64_t nf, int nz, const double *z, double dt, const double *Qn,
                        const double *Qnp1, double *T, const double *ti, const double *rhoc,
                        const double *emiss, const double *Fgeotherm, double *Fsurf,
                        int *iters) {
  const int n1 = nz + 1;
  double *w = (double *)malloc(sizeof(double) * 9 * n1);
  double *k = w, *alpha = w + n1, *gamma = w + 2 * n1, *a = w + 3 * n1, *b = w + 4 * n1,
         *c = w + 5 * n1, *r = w + 6 * n1, *Told = w + 7 * n1, *gam = w + 8 * n1;
  int err = 0;
  for (int i = 1; i <= nz; ++i) k[i] = ti[i] * ti[i] / rhoc[i];
  const double dz = 2. * z[1];
  const double beta = dt / rhoc[1] / (2. * dz * dz);
  alpha[1] = beta * k[2];
  gamma[1] = beta * k[1];
  o_interior(nz, z, dt, rhoc, k, alpha, gamma);
  const double k1 = k[1] / dz;
  for (int i = 1; i <= nz; ++i) {
    a[i] = -gamma[i];
    b[i] = 1. + alpha[i] + gamma[i];
    c[i] = -alpha[i];
  }
  b[nz] = 1. + gamma[nz];
  for (int64_t f = 0; f < nf; ++f) {
    double *Tf = T + f * n1;
    const double em = emiss[f];
    double Tr = Tf[0];
    int iter = 0;
    for (int i = 1; i <= nz; ++i) Told[i] = Tf[i];
    for (;;) {
      const double arad = -3 * em * O_SIGSB * Tr * Tr * Tr * Tr;
      const double brad = 2 * em * O_SIGSB * Tr * Tr * Tr;
      const double ann = (Qn[f] - arad) / (k1 + brad);
      const double annp1 = (Qnp1[f] - arad) / (k1 + brad);
      const double bn = (k1 - brad) / (k1 + brad);
      b[1] = 1. + alpha[1] + gamma[1] - gamma[1] * bn;
      r[1] = gamma[1] * (annp1 + ann) + (1. - alpha[1] - gamma[1] + gamma[1] * bn) * Tf[1] +
             alpha[1] * Tf[2];
      for (int i = 2; i < nz; ++i)
        r[i] = gamma[i] * Tf[i - 1] + (1 - alpha[i] - gamma[i]) * Tf[i] + alpha[i] * Tf[i + 1];
      r[nz] = gamma[nz] * Tf[nz - 1] + (1. - gamma[nz]) * Tf[nz] +
              dt / rhoc[nz] * Fgeotherm[f] / (z[nz] - z[nz - 1]);
      const int e = o_tridag(a, b, c, r, Tf, nz, gam);
      if (e && !err) err = e;
      Tf[0] = 0.5 * (annp1 + bn * Tf[1] + Tf[1]);
      if ((Tf[0] > 1.2 * Tr || Tf[0] < 0.8 * Tr) && iter < 10) {
        ++iter;
        Tr = sqrt(Tr * Tf[0]);
        for (int i = 1; i <= nz; ++i) Tf[i] = Told[i];
        continue;
      }
      break;
    }
    if (iters) iters[f] = iter;
    Fsurf[f] = -k[1] * (Tf[1] - Tf[0]) / z[1];
  }
  free(w);
  return err;
}

int oracle_conduction_t(int64_t nf, int nz, const double *z, double dt, double *T,
                        const double *Tsurf, const double *Tsurfp1, const double *ti,
                        const double *rhoc, const double *Fgeotherm, double *Fsurf) {
  const int n1 = nz + 1;
  double *w = (double *)malloc(sizeof(double) * 8 * n1);
  double *k = w, *alpha = w + n1, *gamma = w + 2 * n1, *a = w + 3 * n1, *b = w + 4 * n1,
         *c = w + 5 * n1, *r = w + 6 * n1, *gam = w + 7 * n1;
  int err = 0;
  for (int i = 1; i <= nz; ++i) k[i] = ti[i] * ti[i] / rhoc[i];
  alpha[1] = k[2] * dt / rhoc[1] / (z[2] - z[1]) / z[2];
  gamma[1] = k[1] * dt / rhoc[1] / z[1] / z[2];
  o_interior(nz, z, dt, rhoc, k, alpha, gamma);
  for (int i = 1; i <= nz; ++i) {
    a[i] = -gamma[i];
    b[i] = 1. + alpha[i] + gamma[i];
    c[i] = -alpha[i];
  }
  b[nz] = 1. + gamma[nz];
  for (int64_t f = 0; f < nf; ++f) {
    double *Tf = T + f * n1;
    r[1] = alpha[1] * Tf[2] + (1. - alpha[1] - gamma[1]) * Tf[1] +
           gamma[1] * (Tsurf[f] + Tsurfp1[f]);
    for (int i = 2; i < nz; ++i)
      r[i] = gamma[i] * Tf[i - 1] + (1. - alpha[i] - gamma[i]) * Tf[i] + alpha[i] * Tf[i + 1];
    r[nz] = gamma[nz] * Tf[nz - 1] + (1. - gamma[nz]) * Tf[nz] +
            dt / rhoc[nz] * Fgeotherm[f] / (z[nz] - z[nz - 1]);
    const int e = o_tridag(a, b, c, r, Tf, nz, gam);
    if (e && !err) err = e;
    Fsurf[f] = -k[1] * (Tf[1] - Tsurfp1[f]) / z[1];
  }
  free(w);
  return err;
}
