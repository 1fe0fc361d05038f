/* The CPU oracle (oracle/tprf_oracle.c) under the address and
 * undefined-behaviour sanitizers on small inputs: each restated transform
 * runs, and the round trips the mathematics fixes are checked
 * (omega -> tau -> omega of a pole sum, k -> r -> k, Dyson -> bubble -> RPA).
 * Exit status 0 when every check holds. */
#define _GNU_SOURCE
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "tprf_oracle.h"

static int failures = 0;
static void check(int ok, const char *what, double v) {
  if (!ok) {
    printf("FAIL %s (%.3e)\n", what, v);
    ++failures;
  }
}

int main(void) {
  /* omega -> tau -> omega of a two-pole G */
  const double beta = 10.0;
  const int nmax = 64, ntau = 6 * nmax + 1, nmesh = 2 * nmax;
  double complex *gw = malloc(sizeof(double complex) * nmesh), *gt = malloc(sizeof(double complex) * ntau),
                 *gw2 = malloc(sizeof(double complex) * nmesh);
  for (int i = 0; i < nmesh; ++i) {
    double complex iw = I * M_PI * (2.0 * (i - nmax) + 1) / beta;
    gw[i] = 0.6 / (iw - 0.7) + 0.4 / (iw + 1.3);
  }
  double err = 0;
  check(or_fourier_w_to_t(0, beta, nmax, ntau, 1, gw, gt, NULL, NULL, &err) == 0, "w_to_t", err);
  check(or_fourier_t_to_w(0, beta, ntau, nmax, 1, gt, gw2) == 0, "t_to_w", 0);
  double d = 0;
  for (int i = 0; i < nmesh; ++i) d = fmax(d, cabs(gw2[i] - gw[i]));
  check(d < 1e-6, "omega -> tau -> omega", d);
  /* tail moments of the same G: t1 = 1, t2 = 0.6 * 0.7 - 0.4 * 1.3 */
  double complex t[4];
  check(or_tail_moments(0, beta, nmax, 1, gw, t, &err) == 0, "tail moments", err);
  check(cabs(t[1] - 1.0) < 1e-12 && cabs(t[2] - (0.42 - 0.52)) < 1e-9, "tail moments values", cabs(t[2] + 0.1));
  /* k -> r -> k */
  const int dims[3] = {4, 3, 2}, nk = 24, n = 2;
  double complex *x = malloc(sizeof(double complex) * nk * 4), *y = malloc(sizeof(double complex) * nk * 4),
                 *z = malloc(sizeof(double complex) * nk * 4);
  for (int i = 0; i < nk * 4; ++i) x[i] = sin(1.0 + i) + I * cos(2.0 * i);
  check(or_fourier_lattice(1, dims, 4, x, y, -1) == 0 && or_fourier_lattice(1, dims, 4, y, z, +1) == 0, "lattice", 0);
  d = 0;
  for (int i = 0; i < nk * 4; ++i) d = fmax(d, cabs(z[i] - x[i]));
  check(d < 1e-13, "k -> r -> k", d);
  /* Dyson -> k -> r -> tau -> bubble -> r -> q -> RPA on a two-orbital lattice */
  double complex *ek = malloc(sizeof(double complex) * nk * n * n);
  for (int k = 0; k < nk; ++k) {
    ek[k * 4 + 0] = -2.0 * cos(2 * M_PI * (k % 4) / 4.0);
    ek[k * 4 + 3] = 0.5 - cos(2 * M_PI * (k % 3) / 3.0);
    ek[k * 4 + 1] = 0.1;
    ek[k * 4 + 2] = 0.1;
  }
  double complex *gwk = malloc(sizeof(double complex) * nmesh * nk * n * n),
                 *gwr = malloc(sizeof(double complex) * nmesh * nk * n * n),
                 *gtr = malloc(sizeof(double complex) * ntau * nk * n * n);
  check(or_lattice_dyson_g_wk(beta, nmax, 0.2, nk, n, ek, 0, NULL, gwk) == 0, "dyson", 0);
  check(or_fourier_lattice(nmesh, dims, n * n, gwk, gwr, -1) == 0, "lattice g", 0);
  check(or_fourier_w_to_t(0, beta, nmax, ntau, (long)nk * n * n, gwr, gtr, NULL, NULL, &err) == 0, "w_to_t g", err);
  const int nb = 4, nnu = 2 * nb - 1, n4 = 16;
  double complex *chi0r = malloc(sizeof(double complex) * nnu * nk * n4),
                 *chi0k = malloc(sizeof(double complex) * nnu * nk * n4),
                 *chi = malloc(sizeof(double complex) * nnu * nk * n4), U[16] = {0};
  for (int a = 0; a < 2; ++a) U[(a * 2 + a) * 4 + a * 2 + a] = 1.0;
  check(or_chi0_wr_from_grt_PH(beta, ntau, dims, n, nb, gtr, chi0r) == 0, "bubble", 0);
  check(or_fourier_lattice(nnu, dims, n4, chi0r, chi0k, +1) == 0, "lattice chi", 0);
  check(or_solve_rpa_PH((long)nnu * nk, n, chi0k, U, chi) == 0, "rpa", 0);
  int fin = 1;
  for (int i = 0; i < nnu * nk * n4; ++i) fin = fin && isfinite(creal(chi[i])) && isfinite(cimag(chi[i]));
  check(fin, "chain finite", 0);
  free(gw); free(gt); free(gw2); free(x); free(y); free(z); free(ek); free(gwk); free(gwr); free(gtr);
  free(chi0r); free(chi0k); free(chi);
  printf("%s: %d failure(s)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
