/*
 * tprf_oracle.c -- CPU restatement of the TRIQS/tprf 3.2 lattice bubble + RPA
 * chain (C99 + OpenMP).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP product path is
 * compared with, and the CPU baseline timed beside it.  The product library
 * never links, loads or calls it.
 *
 * Every function names the reference lines it restates.  File paths are
 * relative to the TRIQS/tprf 3.2 source tree.  Where the reference relies on
 * the third-party TRIQS library (fit_tail, make_adjoint_mesh), the behaviour
 * restated here is documented at the function and is not pinned by any
 * reference test.
 */
#include "tprf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef or_c128 cplx;
typedef long double _Complex lcplx;

#define OR_OK 0
#define OR_ERR_ARG (-1)
#define OR_ERR_MESH_TOO_SHORT (-3)
#define OR_ERR_TAIL_FIT (-4)
#define OR_ERR_TAIL_ORDER (-5)
#define OR_ERR_SINGULAR (-6)
#define OR_ERR_NOMEM (-8)

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* The fit-error assertion of the omega->tau transform can be switched off so
 * that the transform's matrix can be read column by column from unit inputs
 * (the map is complex-linear; the assertion is a check on the data only). */
static int g_fit_check = 1;
void or_set_fit_check(int on) { g_fit_check = on != 0; }

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void or_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* fermi(e) -- c++/triqs_tprf/lattice/lattice_utility.cpp:146-153 */
double or_fermi(double e) {
  if (e < 0) return 1. / (exp(e) + 1.);
  double exp_me = exp(-e);
  return exp_me / (1 + exp_me);
}

/* ------------------------------------------------------------------------ */
/* 1-D complex DFT: recursive mixed-radix decimation in time with a twiddle
 * table.  out[k] = sum_j in[j] * exp(sign * 2 pi i j k / n).  This plays the
 * role of fftw_execute_dft in c++/triqs_tprf/fourier/fourier_common.cpp:48-54
 * (FFTW_FORWARD = sign -1, FFTW_BACKWARD = sign +1, both unnormalised).      */

typedef struct {
  int n;
  cplx *w; /* w[k] = exp(sign 2 pi i k / n) */
} fft_plan;

static int smallest_factor(int n) {
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

static int fft_plan_make(fft_plan *p, int n, int sign) {
  p->n = n;
  p->w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
  if (!p->w) return OR_ERR_NOMEM;
  for (int k = 0; k < n; ++k) {
    /* long double angle reduction keeps every table entry correctly rounded */
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p->w[k] = (double)cosl(a) + (double)(sign * sinl(a)) * I;
  }
  return OR_OK;
}

static void fft_plan_free(fft_plan *p) { free(p->w); p->w = NULL; }

/* sub-transform of length n whose root of unity is w[tw] (tw = N / n) */
static void dft_rec(const fft_plan *p, int n, int tw, const cplx *in, long is, cplx *out) {
  if (n == 1) { out[0] = in[0]; return; }
  int r = smallest_factor(n);
  int m = n / r;
  for (int q = 0; q < r; ++q) dft_rec(p, m, tw * r, in + q * is, is * r, out + (long)q * m);
  const int N = p->n;
  cplx t[64];
  cplx *tt = (r <= 64) ? t : (cplx *)malloc(sizeof(cplx) * (size_t)r);
  for (int k1 = 0; k1 < m; ++k1) {
    for (int q = 0; q < r; ++q) {
      long e = ((long)q * k1 * tw) % N;
      tt[q] = out[(long)q * m + k1] * p->w[e];
    }
    for (int k2 = 0; k2 < r; ++k2) {
      cplx s = 0;
      for (int q = 0; q < r; ++q) {
        long e = ((long)((q * k2) % r) * (N / r)) % N;
        s += tt[q] * p->w[e];
      }
      out[(long)k2 * m + k1] = s;
    }
  }
  if (tt != t) free(tt);
}

static void fft_exec(const fft_plan *p, const cplx *in, cplx *out) { dft_rec(p, p->n, 1, in, 1, out); }

int or_fft1d(int n, int sign, const cplx *in, cplx *out) {
  if (n < 1 || (sign != 1 && sign != -1)) return OR_ERR_ARG;
  fft_plan p;
  if (fft_plan_make(&p, n, sign)) return OR_ERR_NOMEM;
  fft_exec(&p, in, out);
  fft_plan_free(&p);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Jordan inverse with partial pivoting (the role of nda `inverse`,
 * LAPACK getrf+getri).  a is destroyed; returns OR_ERR_SINGULAR on a zero
 * pivot.                                                                   */
static int gj_inverse(int n, cplx *a, cplx *inv) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) inv[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = cabs(a[k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      double v = cabs(a[i * n + k]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) return OR_ERR_SINGULAR;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        cplx t = a[k * n + j]; a[k * n + j] = a[p * n + j]; a[p * n + j] = t;
        t = inv[k * n + j]; inv[k * n + j] = inv[p * n + j]; inv[p * n + j] = t;
      }
    }
    cplx piv = 1.0 / a[k * n + k];
    for (int j = 0; j < n; ++j) { a[k * n + j] *= piv; inv[k * n + j] *= piv; }
    for (int i = 0; i < n; ++i) {
      if (i == k) continue;
      cplx f = a[i * n + k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; ++j) {
        a[i * n + j] -= f * a[k * n + j];
        inv[i * n + j] -= f * inv[k * n + j];
      }
    }
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* a1: lattice Dyson equation.
 * c++/triqs_tprf/lattice/gf.cpp:91-123 (lattice_dyson_g_Xk), entry points
 * gf.cpp:126-132; Sigma = 0 twin gf.cpp:41-58.
 *   G[w,k] = inverse((i w_n + mu) 1 - e_k[k] - Sigma[w] or Sigma[w,k])       */
int or_lattice_dyson_g_wk(double beta, int n_iw, double mu, int nk, int n, const cplx *e_k, int sigma_kind,
                          const cplx *sigma, cplx *g_wk) {
  if (beta <= 0 || n_iw < 1 || nk < 1 || n < 1 || n > 16) return OR_ERR_ARG;
  if (sigma_kind != OR_SIGMA_NONE && !sigma) return OR_ERR_ARG;
  const long nw = 2L * n_iw, n2 = (long)n * n, npts = nw * nk;
  int status = OR_OK;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < npts; ++idx) {
    long iw = idx / nk, k = idx % nk;
    double wn = M_PI * (2.0 * (double)(iw - n_iw) + 1.0) / beta;
    cplx a[256], inv[256];
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        cplx v = -e_k[k * n2 + i * n + j];
        if (i == j) v += (cplx)(mu + wn * I);
        if (sigma_kind == OR_SIGMA_W) v -= sigma[iw * n2 + i * n + j];
        else if (sigma_kind == OR_SIGMA_WK) v -= sigma[(iw * nk + k) * n2 + i * n + j];
        a[i * n + j] = v;
      }
    if (gj_inverse(n, a, inv) != OR_OK) {
#pragma omp atomic write
      status = OR_ERR_SINGULAR;
    }
    memcpy(g_wk + idx * n2, inv, sizeof(cplx) * (size_t)n2);
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* a2 / a6: batched 3-D lattice FFT with interleaved targets.
 * c++/triqs_tprf/lattice/fourier.hpp:165-231 (loop over the outer mesh),
 * c++/triqs_tprf/fourier/fourier_lattice.cpp:28-58 (k->r: FFTW_FORWARD then
 * divide by N_k; r->k: FFTW_BACKWARD, no scale),
 * c++/triqs_tprf/fourier/fourier_common.cpp:26-54 (rank-3 plan, howmany =
 * ntarget, stride = ntarget, dist = 1).                                     */
int or_fourier_lattice(long nouter, const int dims[3], long ntarget, const cplx *in, cplx *out, int direction) {
  if (nouter < 1 || ntarget < 1 || (direction != -1 && direction != 1)) return OR_ERR_ARG;
  int maxlen = 1;
  for (int d = 0; d < 3; ++d) {
    if (dims[d] < 1 || dims[d] > 65536) return OR_ERR_ARG;
    if (dims[d] > maxlen) maxlen = dims[d];
  }
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long slab = nk * ntarget;
  fft_plan plans[3];
  for (int d = 0; d < 3; ++d)
    if (fft_plan_make(&plans[d], dims[d], direction)) return OR_ERR_NOMEM;
  const double scale = (direction == -1) ? 1.0 / (double)nk : 1.0;
  if (out != in) memcpy(out, in, sizeof(cplx) * (size_t)(nouter * slab));
  /* axis strides (in elements) for axes 0,1,2 */
  const long stride[3] = {(long)dims[1] * dims[2] * ntarget, (long)dims[2] * ntarget, ntarget};
#pragma omp parallel
  {
    cplx *bi = (cplx *)malloc(sizeof(cplx) * (size_t)maxlen), *bo = (cplx *)malloc(sizeof(cplx) * (size_t)maxlen);
#pragma omp for schedule(static)
    for (long w = 0; w < nouter; ++w) {
      cplx *x = out + w * slab;
      for (int ax = 0; ax < 3; ++ax) {
        const int len = dims[ax];
        if (len == 1) continue;
        const long nlines = nk / len * ntarget;
        for (long line = 0; line < nlines; ++line) {
          /* decompose line into the indices of the other axes and the target */
          long t = line % ntarget, rest = line / ntarget;
          long base = t;
          int oth[2], no = 0;
          for (int d = 0; d < 3; ++d)
            if (d != ax) oth[no++] = d;
          long i_b = rest % dims[oth[1]], i_a = rest / dims[oth[1]];
          base += i_a * stride[oth[0]] + i_b * stride[oth[1]];
          for (int j = 0; j < len; ++j) bi[j] = x[base + j * stride[ax]];
          fft_exec(&plans[ax], bi, bo);
          for (int j = 0; j < len; ++j) x[base + j * stride[ax]] = bo[j];
        }
      }
      if (scale != 1.0)
        for (long i = 0; i < slab; ++i) x[i] *= scale;
    }
    free(bi);
    free(bo);
  }
  for (int d = 0; d < 3; ++d) fft_plan_free(&plans[d]);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Tail-fit window.  Restates the window of the TRIQS 3.2 tail_fitter used by
 * fit_tail (third-party, called at c++/triqs_tprf/fourier/fourier_matsubara.cpp:226):
 * tail fraction 0.2 of each half-mesh, at most 30 points per side taken from
 * the outermost frequency inward with an integer step, expansion order
 * K = min(9, floor(ln 1e16 / ln omega_max), n_fit - 1), Vandermonde in the
 * normalised variable z = omega_max / (i omega_n).  The exact TRIQS window
 * and order rule are recalled, not read; this is the policy both the oracle
 * and the product implement.                                               */
int or_tail_window_make(int statistic, double beta, int n_max, or_tail_window *w) {
  if (beta <= 0 || n_max < 2) return OR_ERR_ARG;
  const int N = (statistic == OR_FERMION) ? 2 * n_max : 2 * n_max - 1;
  w->omega_max = (statistic == OR_FERMION) ? M_PI * (2 * n_max - 1) / beta : 2 * M_PI * (n_max - 1) / beta;
  int W = (int)lround(0.2 * N / 2.0);
  if (W < 1) W = 1;
  int npts = W < 30 ? W : 30;
  int step = W / npts;
  if (step < 1) step = 1;
  /* the two sides must not overlap and must not touch omega = 0 (bosons) */
  if ((npts - 1) * step >= (N - 1) / 2) return OR_ERR_TAIL_ORDER;
  for (int i = 0; i < npts; ++i) {
    w->idx[i] = i * step;
    w->idx[npts + i] = N - 1 - i * step;
  }
  w->n_fit = 2 * npts;
  w->step = step;
  int K = 9;
  if (w->omega_max > 1.0) {
    int k2 = (int)floor(log(1e16) / log(w->omega_max));
    if (k2 < K) K = k2;
  }
  if (K > w->n_fit - 1) K = w->n_fit - 1;
  w->order = K;
  /* the omega->tau transform needs a proper 3rd moment:
     fourier_matsubara.cpp:235-237 asserts more than 4 tail coefficients */
  if (K < 4) return OR_ERR_TAIL_ORDER;
  return OR_OK;
}

static double matsubara(int statistic, long n, double beta) {
  return (statistic == OR_FERMION) ? M_PI * (2.0 * n + 1.0) / beta : 2.0 * M_PI * (double)n / beta;
}

/* Householder QR of the n_fit x (K+1) normalised Vandermonde in long double */
typedef struct {
  int m, k;          /* rows, columns */
  lcplx v[128 * 10];  /* reflectors, column-major v[c*m + r] */
  lcplx R[10 * 10];
  lcplx V[128 * 10];  /* the Vandermonde itself, row-major */
} ls_qr;

static void ls_qr_make(const or_tail_window *tw, int statistic, int first, double beta, ls_qr *q) {
  const int m = tw->n_fit, k = tw->order + 1;
  q->m = m; q->k = k;
  lcplx A[128 * 10];
  for (int r = 0; r < m; ++r) {
    long n = first + tw->idx[r];
    lcplx z = (lcplx)tw->omega_max / ((lcplx)(matsubara(statistic, n, beta) * I));
    lcplx p = 1;
    for (int c = 0; c < k; ++c) { A[r * k + c] = p; q->V[r * k + c] = p; p *= z; }
  }
  for (int c = 0; c < k; ++c) {
    long double nx = 0;
    for (int r = c; r < m; ++r) nx += creall(A[r * k + c] * conjl(A[r * k + c]));
    nx = sqrtl(nx);
    lcplx x0 = A[c * k + c];
    long double ax0 = cabsl(x0);
    lcplx ph = (ax0 > 0) ? x0 / ax0 : 1;
    lcplx alpha = -ph * nx;
    for (int r = 0; r < m; ++r) q->v[c * m + r] = 0;
    for (int r = c; r < m; ++r) q->v[c * m + r] = A[r * k + c];
    q->v[c * m + c] -= alpha;
    long double nv = 0;
    for (int r = c; r < m; ++r) nv += creall(q->v[c * m + r] * conjl(q->v[c * m + r]));
    nv = sqrtl(nv);
    if (nv > 0)
      for (int r = c; r < m; ++r) q->v[c * m + r] /= nv;
    for (int cc = c; cc < k; ++cc) {
      lcplx s = 0;
      for (int r = c; r < m; ++r) s += conjl(q->v[c * m + r]) * A[r * k + cc];
      for (int r = c; r < m; ++r) A[r * k + cc] -= 2 * q->v[c * m + r] * s;
    }
  }
  for (int r = 0; r < k; ++r)
    for (int c = 0; c < k; ++c) q->R[r * k + c] = (c >= r) ? A[r * k + c] : 0;
}

/* least-squares coefficients c (K+1) and RMS residual for one column g */
static double ls_qr_solve(const ls_qr *q, const cplx *g, lcplx *c) {
  const int m = q->m, k = q->k;
  lcplx y[128];
  for (int r = 0; r < m; ++r) y[r] = g[r];
  for (int cc = 0; cc < k; ++cc) {
    lcplx s = 0;
    for (int r = cc; r < m; ++r) s += conjl(q->v[cc * m + r]) * y[r];
    for (int r = cc; r < m; ++r) y[r] -= 2 * q->v[cc * m + r] * s;
  }
  for (int r = k - 1; r >= 0; --r) {
    lcplx s = y[r];
    for (int cc = r + 1; cc < k; ++cc) s -= q->R[r * k + cc] * c[cc];
    c[r] = s / q->R[r * k + r];
  }
  long double res = 0;
  for (int r = 0; r < m; ++r) {
    lcplx s = 0;
    for (int cc = 0; cc < k; ++cc) s += q->V[r * k + cc] * c[cc];
    lcplx d = (lcplx)g[r] - s;
    res += creall(d * conjl(d));
  }
  return (double)sqrtl(res / m);
}

/* oneFermion / oneBoson: c++/triqs_tprf/fourier/fourier_matsubara.cpp:31-37 */
static double one_fermion(double b, double tau, double beta) {
  return -(b >= 0 ? exp(-b * tau) / (1 + exp(-beta * b)) : exp(b * (beta - tau)) / (1 + exp(beta * b)));
}
static double one_boson(double b, double tau, double beta) {
  return (b >= 0 ? exp(-b * tau) / (exp(-beta * b) - 1) : exp(b * (beta - tau)) / (1 - exp(b * beta)));
}

/* pole model: fourier_matsubara.cpp:150-170 and :274-288 */
static void pole_model(int statistic, cplx m1, cplx m2, cplx m3, double b[3], cplx a[3]) {
  if (statistic == OR_FERMION) {
    b[0] = 0; b[1] = 1; b[2] = -1;
    a[0] = m1 - m3; a[1] = (m2 + m3) / 2; a[2] = (m3 - m2) / 2;
  } else {
    b[0] = -0.5; b[1] = -1; b[2] = 1;
    a[0] = 4.0 * (m1 - m3) / 3.0; a[1] = m3 - (m1 + m2) / 2.0; a[2] = m1 / 6.0 + m2 / 2.0 + m3 / 3.0;
  }
}

/* ------------------------------------------------------------------------ */
/* a3: Matsubara frequency -> imaginary time, per column.
 * c++/triqs_tprf/lattice/fourier.hpp:98-130 (loop over r),
 * c++/triqs_tprf/fourier/fourier_matsubara.cpp:194-308 (_fourier_impl for
 * imtime): moments from fit_tail (:225-246) unless given; 3-pole model
 * subtraction (:274-290); FFTW_FORWARD of length L = ntau-1 (:213,:294);
 * phase exp(-i pi tau/beta) for fermions and the pole model added back
 * (:298-302); endpoint G(tau_L) = -/+ (G(tau_0) + m1) (:304-305).           */
int or_fourier_w_to_t(int statistic, double beta, int n_max, int ntau, long ncol, const cplx *gw, cplx *gt,
                      const cplx *moments, cplx *moments_out, double *fit_error) {
  if (beta <= 0 || n_max < 1 || ncol < 1 || ntau < 2) return OR_ERR_ARG;
  const long L = ntau - 1;
  const int nmesh = (statistic == OR_FERMION) ? 2 * n_max : 2 * n_max - 1;
  const int first = (statistic == OR_FERMION) ? -n_max : -(n_max - 1);
  if (L < 2L * n_max) return OR_ERR_MESH_TOO_SHORT; /* fourier_matsubara.cpp:201-204, :251-254 */
  or_tail_window tw;
  ls_qr *qr = NULL;
  if (!moments) {
    int st = or_tail_window_make(statistic, beta, n_max, &tw);
    if (st) return st;
    qr = (ls_qr *)malloc(sizeof(ls_qr));
    if (!qr) return OR_ERR_NOMEM;
    ls_qr_make(&tw, statistic, first, beta, qr);
  }
  fft_plan plan;
  if (fft_plan_make(&plan, (int)L, -1)) { free(qr); return OR_ERR_NOMEM; }
  double max_err = 0;
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)L), *y = (cplx *)malloc(sizeof(cplx) * (size_t)L);
    double my_err = 0;
#pragma omp for schedule(static)
    for (long c = 0; c < ncol; ++c) {
      cplx m1, m2, m3;
      if (moments) {
        m1 = moments[c]; m2 = moments[ncol + c]; m3 = moments[2 * ncol + c];
      } else {
        cplx gfit[128];
        lcplx coef[10];
        for (int r = 0; r < tw.n_fit; ++r) gfit[r] = gw[(long)tw.idx[r] * ncol + c];
        double e = ls_qr_solve(qr, gfit, coef);
        if (e > my_err) my_err = e;
        m1 = (cplx)(coef[1] * (long double)tw.omega_max);
        m2 = (cplx)(coef[2] * (long double)tw.omega_max * (long double)tw.omega_max);
        m3 = (cplx)(coef[3] * (long double)tw.omega_max * (long double)tw.omega_max * (long double)tw.omega_max);
      }
      if (moments_out) { moments_out[c] = m1; moments_out[ncol + c] = m2; moments_out[2 * ncol + c] = m3; }
      double b[3];
      cplx a[3];
      pole_model(statistic, m1, m2, m3, b, a);
      for (long j = 0; j < L; ++j) x[j] = 0;
      for (int i = 0; i < nmesh; ++i) {
        long n = first + i;
        cplx iw = matsubara(statistic, n, beta) * I;
        cplx model = a[0] / (iw - b[0]) + a[1] / (iw - b[1]) + a[2] / (iw - b[2]);
        x[(n + L) % L] = (1.0 / beta) * (gw[(long)i * ncol + c] - model);
      }
      fft_exec(&plan, x, y);
      for (long j = 0; j < L; ++j) {
        double tau = (double)j * (beta / (double)L);
        cplx v;
        if (statistic == OR_FERMION) {
          v = y[j] * cexp(-I * (M_PI / beta) * tau);
          for (int p = 0; p < 3; ++p) v += a[p] * one_fermion(b[p], tau, beta);
        } else {
          v = y[j];
          for (int p = 0; p < 3; ++p) v += a[p] * one_boson(b[p], tau, beta);
        }
        gt[j * ncol + c] = v;
      }
      double pm = (statistic == OR_FERMION) ? -1.0 : 1.0;
      gt[L * ncol + c] = pm * (gt[c] + m1);
    }
#pragma omp critical
    if (my_err > max_err) max_err = my_err;
    free(x);
    free(y);
  }
  fft_plan_free(&plan);
  free(qr);
  if (fit_error) *fit_error = max_err;
  if (!moments && max_err >= 1e-3 && g_fit_check) return OR_ERR_TAIL_FIT; /* fourier_matsubara.cpp:227-230 */
  return OR_OK;
}

/* Tail-fit moments t_k = c_k omega_max^k (k = 0..3) of every column, by the
 * fit of or_fourier_w_to_t (the TRIQS fit_tail window and order, third-party,
 * called at c++/triqs_tprf/lattice/lattice_utility.cpp:71 and by density()
 * for rho_k_from_g_wk, gw.cpp:46).  t: [4][ncol].                          */
int or_tail_moments(int statistic, double beta, int n_max, long ncol, const cplx *gw, cplx *t, double *fit_error) {
  if (beta <= 0 || n_max < 1 || ncol < 1) return OR_ERR_ARG;
  const int first = (statistic == OR_FERMION) ? -n_max : -(n_max - 1);
  or_tail_window tw;
  int st = or_tail_window_make(statistic, beta, n_max, &tw);
  if (st) return st;
  ls_qr *qr = (ls_qr *)malloc(sizeof(ls_qr));
  if (!qr) return OR_ERR_NOMEM;
  ls_qr_make(&tw, statistic, first, beta, qr);
  double max_err = 0;
  for (long c = 0; c < ncol; ++c) {
    cplx gfit[128];
    lcplx coef[10];
    for (int r = 0; r < tw.n_fit; ++r) gfit[r] = gw[(long)tw.idx[r] * ncol + c];
    double e = ls_qr_solve(qr, gfit, coef);
    if (e > max_err) max_err = e;
    long double wp = 1;
    for (int k = 0; k < 4; ++k) {
      t[(long)k * ncol + c] = (cplx)(coef[k] * wp);
      wp *= (long double)tw.omega_max;
    }
  }
  free(qr);
  if (fit_error) *fit_error = max_err;
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* exact inverse of V_{m,j} = m^{j-1} (m, j = 1..8): row j-1 of V^{-1} holds the
 * coefficients of t^{j-1} in the Lagrange basis polynomials l_m(t) =
 * prod_{k != m} (t - k)/(m - k).  Each entry is an integer ratio, so one
 * division gives the correctly rounded double.  (The reference hard-codes a
 * rounded version of this matrix at fourier_matsubara.cpp:54-71.)          */
void or_vandermonde_inverse8(double *vinv) {
  for (int m = 1; m <= 8; ++m) {
    double poly[9] = {1, 0, 0, 0, 0, 0, 0, 0, 0}; /* coefficients, ascending */
    double den = 1;
    int deg = 0;
    for (int k = 1; k <= 8; ++k) {
      if (k == m) continue;
      for (int d = deg + 1; d >= 1; --d) poly[d] = poly[d - 1] - k * poly[d];
      poly[0] = -k * poly[0];
      ++deg;
      den *= (double)(m - k);
    }
    for (int j = 0; j < 8; ++j) vinv[j * 8 + (m - 1)] = poly[j] / den;
  }
}

/* fit_derivatives: c++/triqs_tprf/fourier/fourier_matsubara.cpp:42-97.
 * returns m2, m3 from 8-point one-sided differences at both ends.          */
static void fit_derivatives(int statistic, long L, double dtau, const double *vinv, const cplx *g, long gs, cplx *m2,
                            cplx *m3) {
  cplx dl[8], dr[8];
  for (int m = 1; m <= 8; ++m) {
    double tm = (double)m * dtau;
    dl[m - 1] = (g[m * gs] - g[0]) / tm;
    dr[m - 1] = (g[L * gs] - g[(L - m) * gs]) / tm;
  }
  cplx gl0 = 0, gl1 = 0, gr0 = 0, gr1 = 0;
  for (int m = 0; m < 8; ++m) {
    gl0 += vinv[0 * 8 + m] * dl[m]; gl1 += vinv[1 * 8 + m] * dl[m];
    gr0 += vinv[0 * 8 + m] * dr[m]; gr1 += vinv[1 * 8 + m] * dr[m];
  }
  double sign = (statistic == OR_FERMION) ? -1 : 1;
  *m2 = gl0 - sign * gr0;
  *m3 = -(gl1 + sign * gr1) * 2.0 / dtau;
}

/* one column of the imaginary time -> Matsubara transform.
 * c++/triqs_tprf/fourier/fourier_matsubara.cpp:123-184 (_fourier_impl for
 * imfreq): m1 from the endpoints (:151,:163), pole model (:150-171),
 * x_j = (beta/L) [exp(i pi tau_j / beta)] (g_j - model) for j < L, FFTW_BACKWARD
 * of length L (:118,:176), G(i w_n) = y[(n+L)%L] + sum a/(iw - b) (:181).   */
static void t_to_w_column(int statistic, double beta, long L, int nw, const double *vinv, const fft_plan *plan,
                          const cplx *g, long gs, cplx *out, long os, cplx *x, cplx *y) {
  const double dtau = beta / (double)L;
  cplx m1, m2, m3;
  fit_derivatives(statistic, L, dtau, vinv, g, gs, &m2, &m3);
  m1 = (statistic == OR_FERMION) ? -(g[0] + g[L * gs]) : -(g[0] - g[L * gs]);
  double b[3];
  cplx a[3];
  pole_model(statistic, m1, m2, m3, b, a);
  const double fact = beta / (double)L;
  for (long j = 0; j < L; ++j) {
    double tau = (double)j * dtau;
    cplx model = 0;
    if (statistic == OR_FERMION) {
      for (int p = 0; p < 3; ++p) model += a[p] * one_fermion(b[p], tau, beta);
      x[j] = fact * cexp(I * (M_PI / beta) * tau) * (g[j * gs] - model);
    } else {
      for (int p = 0; p < 3; ++p) model += a[p] * one_boson(b[p], tau, beta);
      x[j] = fact * (g[j * gs] - model);
    }
  }
  fft_exec(plan, x, y);
  const int nout = (statistic == OR_FERMION) ? 2 * nw : 2 * nw - 1;
  const int first = (statistic == OR_FERMION) ? -nw : -(nw - 1);
  for (int i = 0; i < nout; ++i) {
    long n = first + i;
    cplx iw = matsubara(statistic, n, beta) * I;
    out[i * os] = y[(n + L) % L] + a[0] / (iw - b[0]) + a[1] / (iw - b[1]) + a[2] / (iw - b[2]);
  }
}

/* a5: c++/triqs_tprf/lattice/fourier.hpp:132-163 (loop over r) ->
 * fourier_matsubara.cpp:102-184.  Precondition L >= 2 (last_index + 1)
 * (:105-108, :128-131).                                                      */
int or_fourier_t_to_w(int statistic, double beta, int ntau, int nw, long ncol, const cplx *gt, cplx *gw) {
  if (beta <= 0 || nw < 1 || ncol < 1 || ntau < 10) return OR_ERR_ARG;
  const long L = ntau - 1;
  if (L < 2L * nw) return OR_ERR_MESH_TOO_SHORT;
  double vinv[64];
  or_vandermonde_inverse8(vinv);
  fft_plan plan;
  if (fft_plan_make(&plan, (int)L, +1)) return OR_ERR_NOMEM;
#pragma omp parallel
  {
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)L), *y = (cplx *)malloc(sizeof(cplx) * (size_t)L);
#pragma omp for schedule(static)
    for (long c = 0; c < ncol; ++c) t_to_w_column(statistic, beta, L, nw, vinv, &plan, gt + c, ncol, gw + c, ncol, x, y);
    free(x);
    free(y);
  }
  fft_plan_free(&plan);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
static long minus_r(long r, const int dims[3]) {
  long i2 = r % dims[2], i1 = (r / dims[2]) % dims[1], i0 = r / ((long)dims[1] * dims[2]);
  long j0 = (dims[0] - i0) % dims[0], j1 = (dims[1] - i1) % dims[1], j2 = (dims[2] - i2) % dims[2];
  return (j0 * dims[1] + j1) * dims[2] + j2;
}

/* chi0_abcd(tau_j, r) = G_da(tau_j, r) * G_bc(tau_{L-j}, -r)
 * c++/triqs_tprf/lattice/chi_imtime.cpp:127 (g_mr_t(beta - t) evaluated on
 * the mesh point tau_{L-j}).                                                */
static void product_column_block(int ntau, long nr, int n, const cplx *g_tr, long r, long mr, cplx *chi_t) {
  const long n2 = (long)n * n, n4 = n2 * n2, L = ntau - 1;
  for (long j = 0; j < ntau; ++j) {
    const cplx *gp = g_tr + (j * nr + r) * n2;
    const cplx *gm = g_tr + ((L - j) * nr + mr) * n2;
    cplx *o = chi_t + j * n4;
    for (int a = 0; a < n; ++a)
      for (int b = 0; b < n; ++b)
        for (int c = 0; c < n; ++c)
          for (int d = 0; d < n; ++d) o[((a * n + b) * n + c) * n + d] = gp[d * n + a] * gm[b * n + c];
  }
}

/* a4: c++/triqs_tprf/lattice/chi_imtime.cpp:88-136 */
int or_chi0_tr_from_grt_PH(int ntau, const int dims[3], int n, const cplx *g_tr, cplx *chi0_tr) {
  if (ntau < 2 || n < 1) return OR_ERR_ARG;
  const long nr = (long)dims[0] * dims[1] * dims[2];
  const long n2 = (long)n * n, n4 = n2 * n2;
#pragma omp parallel
  {
    cplx *chi_t = (cplx *)malloc(sizeof(cplx) * (size_t)(ntau * n4));
#pragma omp for schedule(static)
    for (long r = 0; r < nr; ++r) {
      product_column_block(ntau, nr, n, g_tr, r, minus_r(r, dims), chi_t);
      for (long j = 0; j < ntau; ++j) memcpy(chi0_tr + (j * nr + r) * n4, chi_t + j * n4, sizeof(cplx) * (size_t)n4);
    }
    free(chi_t);
  }
  return OR_OK;
}

/* a4': fused product and tau->nu per r, c++/triqs_tprf/lattice/chi_imtime.cpp:139-182.
 * The reference's fused variant calls TRIQS make_gf_from_fourier (:175); this
 * restatement applies the a5 transform (fourier_matsubara.cpp:102-184) so the
 * fused result equals the a4 -> a5 chain.                                   */
int or_chi0_wr_from_grt_PH(double beta, int ntau, const int dims[3], int n, int nw, const cplx *g_tr, cplx *chi0_wr) {
  if (beta <= 0 || ntau < 10 || n < 1 || nw < 1) return OR_ERR_ARG;
  const long L = ntau - 1;
  if (L < 2L * nw) return OR_ERR_MESH_TOO_SHORT;
  const long nr = (long)dims[0] * dims[1] * dims[2];
  const long n2 = (long)n * n, n4 = n2 * n2;
  double vinv[64];
  or_vandermonde_inverse8(vinv);
  fft_plan plan;
  if (fft_plan_make(&plan, (int)L, +1)) return OR_ERR_NOMEM;
#pragma omp parallel
  {
    cplx *chi_t = (cplx *)malloc(sizeof(cplx) * (size_t)(ntau * n4));
    cplx *x = (cplx *)malloc(sizeof(cplx) * (size_t)L), *y = (cplx *)malloc(sizeof(cplx) * (size_t)L);
#pragma omp for schedule(static)
    for (long r = 0; r < nr; ++r) {
      product_column_block(ntau, nr, n, g_tr, r, minus_r(r, dims), chi_t);
      for (long c = 0; c < n4; ++c)
        t_to_w_column(OR_BOSON, beta, L, nw, vinv, &plan, chi_t + c, n4, chi0_wr + r * n4 + c, nr * n4, x, y);
    }
    free(chi_t);
    free(x);
    free(y);
  }
  fft_plan_free(&plan);
  return OR_OK;
}

/* chi_trapz_tau: c++/triqs_tprf/lattice/chi_imtime.cpp:230-247 */
static void trapz_column(int ntau, double beta, const cplx *g, long gs, long ncomp, cplx *out) {
  for (long c = 0; c < ncomp; ++c) {
    cplx s = 0;
    for (long j = 0; j < ntau; ++j) s += g[j * gs + c];
    s -= 0.5 * g[c];
    s -= 0.5 * g[(long)(ntau - 1) * gs + c];
    out[c] = s * (beta / (double)(ntau - 1));
  }
}

/* chi0_w0r_from_grt_PH: c++/triqs_tprf/lattice/chi_imtime.cpp:185-228 */
int or_chi0_w0r_from_grt_PH(double beta, int ntau, const int dims[3], int n, const cplx *g_tr, cplx *chi0_wr) {
  if (beta <= 0 || ntau < 2 || n < 1) return OR_ERR_ARG;
  const long nr = (long)dims[0] * dims[1] * dims[2];
  const long n2 = (long)n * n, n4 = n2 * n2;
#pragma omp parallel
  {
    cplx *chi_t = (cplx *)malloc(sizeof(cplx) * (size_t)(ntau * n4));
#pragma omp for schedule(static)
    for (long r = 0; r < nr; ++r) {
      product_column_block(ntau, nr, n, g_tr, r, minus_r(r, dims), chi_t);
      trapz_column(ntau, beta, chi_t, n4, n4, chi0_wr + r * n4);
    }
    free(chi_t);
  }
  return OR_OK;
}

/* chi_w0r_from_chi_tr: c++/triqs_tprf/lattice/chi_imtime.cpp:250-277 */
int or_chi_w0r_from_chi_tr(double beta, int ntau, long nr, int n, const cplx *chi_tr, cplx *chi_wr) {
  if (beta <= 0 || ntau < 2 || n < 1 || nr < 1) return OR_ERR_ARG;
  const long n4 = (long)n * n * n * n;
#pragma omp parallel for schedule(static)
  for (long r = 0; r < nr; ++r) trapz_column(ntau, beta, chi_tr + r * n4, nr * n4, n4, chi_wr + r * n4);
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* a7: c++/triqs_tprf/lattice/rpa.cpp:29-64.  M_chi = inverse(1 - M_chi0 M_U)
 * M_chi0 (:57) on the matrices of group_indices_view(X, idx_group<0,1>,
 * idx_group<3,2>) (:40, :54-55).  A view of a row-major tensor can only merge
 * the pair (c, d) in memory order, so M_X[(a,b),(c,d)] = X_abcd; the
 * reference's own dynamical_screened_interaction_w.py test, whose exchange
 * entries (0,1,0,1) and (1,0,1,0) differ, holds only with this grouping (the
 * same nda call there).  For U_abcd = U_badc (every real Hermitian
 * interaction) it coincides with M[(a,b),(d,c)].                            */
int or_solve_rpa_PH(long npts, int n, const cplx *chi0, const cplx *U, cplx *chi) {
  if (npts < 1 || n < 1 || n > 8) return OR_ERR_ARG;
  const int N = n * n;
  const long n4 = (long)N * N;
  cplx *MU = (cplx *)malloc(sizeof(cplx) * (size_t)n4);
  if (!MU) return OR_ERR_NOMEM;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int c = 0; c < n; ++c)
        for (int d = 0; d < n; ++d) MU[(a * n + b) * N + (c * n + d)] = U[((a * n + b) * n + c) * n + d];
  int status = OR_OK;
#pragma omp parallel
  {
    cplx *X = (cplx *)malloc(sizeof(cplx) * 3 * (size_t)n4);
    cplx *A = X + n4, *Ainv = A + n4;
#pragma omp for schedule(static)
    for (long p = 0; p < npts; ++p) {
      if (!X) {
#pragma omp atomic write
        status = OR_ERR_NOMEM;
        continue;
      }
      const cplx *x = chi0 + p * n4;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
          for (int c = 0; c < n; ++c)
            for (int d = 0; d < n; ++d) X[(a * n + b) * N + (c * n + d)] = x[((a * n + b) * n + c) * n + d];
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          cplx s = (i == j) ? 1.0 : 0.0;
          for (int l = 0; l < N; ++l) s -= X[i * N + l] * MU[l * N + j];
          A[i * N + j] = s;
        }
      if (gj_inverse(N, A, Ainv) != OR_OK) {
#pragma omp atomic write
        status = OR_ERR_SINGULAR;
      }
      cplx *o = chi + p * n4;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b)
          for (int c = 0; c < n; ++c)
            for (int d = 0; d < n; ++d) {
              int i = a * n + b, j = c * n + d;
              cplx s = 0;
              for (int l = 0; l < N; ++l) s += Ainv[i * N + l] * X[l * N + j];
              o[((a * n + b) * n + c) * n + d] = s;
            }
    }
    free(X);
  }
  free(MU);
  return status;
}

/* ------------------------------------------------------------------------ */
/* f4: dynamical_screened_interaction_W, c++/triqs_tprf/lattice/
 * dynamical_screened_interaction.cpp.  Per (frequency, k) point:
 *   dynamic = 0 (static V_k, dynamical_screened_interaction_W_opt_bubble,
 *   :79-116):  denom = 1 - sum_ef PI_abef V_efcd (:100-106), inv_denom its
 *   matrix inverse (:108), W_abcd = sum_ef V_abef inv_denom_efcd (:110-114);
 *   dynamic = 1 (V per point, screened_interaction_from_generic_susceptibility
 *   <bubble>, :33-74): M_W = M_V inverse(1 - M_PI M_V) (:57-62).
 * Matrices M_X[(a,b),(c,d)] = X_abcd: the grouping of group_indices_view(X,
 * idx_group<0,1>, idx_group<3,2>) on a row-major tensor (see a7 above).
 * V is [nk][n^4] (dynamic = 0, point p uses V[p % nk]) or [npts][n^4].     */
int or_screened_interaction_W(long npts, long nk, int n, int dynamic, const cplx *PI, const cplx *V, cplx *W) {
  if (npts < 1 || nk < 1 || n < 1 || n > 8) return OR_ERR_ARG;
  const int N = n * n;
  const long n4 = (long)N * N;
  int status = OR_OK;
#pragma omp parallel
  {
    cplx *D = (cplx *)malloc(sizeof(cplx) * 2 * (size_t)n4);
    cplx *Dinv = D + n4;
#pragma omp for schedule(static)
    for (long p = 0; p < npts; ++p) {
      if (!D) {
#pragma omp atomic write
        status = OR_ERR_NOMEM;
        continue;
      }
      const cplx *pi = PI + p * n4;
      const cplx *v = V + (dynamic ? p : p % nk) * n4;
      cplx *w = W + p * n4;
      /* both paths: denominator 1 - M_PI M_V, then M_W = M_V inverse(denominator)
       * (the static path writes the same sums as explicit index loops) */
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          cplx s = (i == j) ? 1.0 : 0.0;
          for (int l = 0; l < N; ++l) s -= pi[i * N + l] * v[l * N + j];
          D[i * N + j] = s;
        }
      if (gj_inverse(N, D, Dinv) != OR_OK) {
#pragma omp atomic write
        status = OR_ERR_SINGULAR;
      }
      for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
          cplx s = 0;
          for (int l = 0; l < N; ++l) s += v[i * N + l] * Dinv[l * N + j];
          w[i * N + j] = s;
        }
    }
    free(D);
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* Hermitian eigen-decomposition by cyclic complex Jacobi rotations.
 * a = evecs diag(evals) evecs^H, eigenvalues ascending.                     */
int or_eigh(int n, const cplx *a_in, double *evals, cplx *V) {
  if (n < 1 || n > 16) return OR_ERR_ARG;
  cplx A[256];
  memcpy(A, a_in, sizeof(cplx) * (size_t)(n * n));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? 1.0 : 0.0;
  double nrm = 0;
  for (int i = 0; i < n * n; ++i) nrm += creal(A[i] * conj(A[i]));
  nrm = sqrt(nrm);
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += creal(A[p * n + q] * conj(A[p * n + q]));
    if (sqrt(off) <= 1e-17 * (nrm > 0 ? nrm : 1)) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        cplx bpq = A[p * n + q];
        double ab = cabs(bpq);
        if (ab == 0) continue;
        cplx ph = bpq / ab; /* e^{i phi} */
        double alpha = creal(A[p * n + p]), gamma = creal(A[q * n + q]);
        double theta = 0.5 * atan2(2 * ab, alpha - gamma);
        double c = cos(theta), s = sin(theta);
        /* G = diag(1, e^{-i phi}) [[c, -s], [s, c]] acting on (p, q) */
        cplx Gpp = c, Gpq = -s, Gqp = s * conj(ph), Gqq = c * conj(ph);
        for (int k = 0; k < n; ++k) { /* A <- A G */
          cplx akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = akp * Gpp + akq * Gqp;
          A[k * n + q] = akp * Gpq + akq * Gqq;
          cplx vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = vkp * Gpp + vkq * Gqp;
          V[k * n + q] = vkp * Gpq + vkq * Gqq;
        }
        for (int k = 0; k < n; ++k) { /* A <- G^H A */
          cplx apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = conj(Gpp) * apk + conj(Gqp) * aqk;
          A[q * n + k] = conj(Gpq) * apk + conj(Gqq) * aqk;
        }
        A[p * n + q] = 0;
        A[q * n + p] = 0;
      }
  }
  for (int i = 0; i < n; ++i) evals[i] = creal(A[i * n + i]);
  /* sort ascending (selection sort on columns) */
  for (int i = 0; i < n; ++i) {
    int m = i;
    for (int j = i + 1; j < n; ++j)
      if (evals[j] < evals[m]) m = j;
    if (m != i) {
      double t = evals[i]; evals[i] = evals[m]; evals[m] = t;
      for (int k = 0; k < n; ++k) { cplx tv = V[k * n + i]; V[k * n + i] = V[k * n + m]; V[k * n + m] = tv; }
    }
  }
  return OR_OK;
}

/* Lindhard bubble: c++/triqs_tprf/lattice/lindhard_chi00.cpp:34-94.
 * chi00_abcd(nu, q) = 1/N_k sum_{k,i,j} U_k(a,i) U_k^*(d,i) U_kq(c,j) U_kq^*(b,j) F_ij,
 * F_ij = [f(beta e_ki) - f(beta e_{k+q,j})] / (i nu + e_{k+q,j} - e_ki), with the
 * nu = 0, |de| < 1e-10 limit beta / (4 cosh^2(beta e_ki / 2)) (:64-74).      */
int or_lindhard_chi00(double beta, double mu, int nw, const int dims[3], int n, const cplx *e_k, cplx *chi_wk) {
  if (beta <= 0 || nw < 1 || n < 1 || n > 8) return OR_ERR_ARG;
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long n2 = (long)n * n, n4 = n2 * n2;
  const int nnu = 2 * nw - 1;
  double *ev = (double *)malloc(sizeof(double) * (size_t)(nk * n));
  cplx *U = (cplx *)malloc(sizeof(cplx) * (size_t)(nk * n2));
  if (!ev || !U) { free(ev); free(U); return OR_ERR_NOMEM; }
  for (long k = 0; k < nk; ++k) {
    cplx m[64];
    for (int i = 0; i < n2; ++i) m[i] = e_k[k * n2 + i];
    for (int i = 0; i < n; ++i) m[i * n + i] -= mu;
    or_eigh(n, m, ev + k * n, U + k * n2);
  }
  memset(chi_wk, 0, sizeof(cplx) * (size_t)(nnu * nk * n4));
#pragma omp parallel for schedule(dynamic)
  for (long q = 0; q < nk; ++q) {
    long q2 = q % dims[2], q1 = (q / dims[2]) % dims[1], q0 = q / ((long)dims[1] * dims[2]);
    for (long k = 0; k < nk; ++k) {
      long k2 = k % dims[2], k1 = (k / dims[2]) % dims[1], k0 = k / ((long)dims[1] * dims[2]);
      long kq = (((k0 + q0) % dims[0]) * dims[1] + (k1 + q1) % dims[1]) * dims[2] + (k2 + q2) % dims[2];
      const double *ek = ev + k * n, *ekq = ev + kq * n;
      const cplx *Uk = U + k * n2, *Ukq = U + kq * n2;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          double de = ekq[j] - ek[i];
          double dn = or_fermi(ek[i] * beta) - or_fermi(ekq[j] * beta);
          for (int mi = 0; mi < nnu; ++mi) {
            double nu = 2 * M_PI * (double)(mi - (nw - 1)) / beta;
            cplx F;
            if (fabs(nu) < 1e-10 && fabs(de) < 1e-10) {
              double ch = cosh(0.5 * beta * ek[i]);
              F = beta / (4. * ch * ch);
            } else {
              F = dn / (nu * I + de);
            }
            cplx *o = chi_wk + ((long)mi * nk + q) * n4;
            for (int a = 0; a < n; ++a)
              for (int b = 0; b < n; ++b)
                for (int c = 0; c < n; ++c)
                  for (int d = 0; d < n; ++d)
                    o[((a * n + b) * n + c) * n + d] +=
                        Uk[a * n + i] * conj(Uk[d * n + i]) * Ukq[c * n + j] * conj(Ukq[b * n + j]) * F;
          }
        }
    }
  }
  const long tot = (long)nnu * nk * n4;
  for (long i = 0; i < tot; ++i) chi_wk[i] /= (double)nk;
  free(ev);
  free(U);
  return OR_OK;
}
