/*
 * tprf_oracle.h -- CPU restatement of the TRIQS/tprf 3.2 lattice bubble + RPA
 * chain.  TEST INFRASTRUCTURE ONLY: this library is the oracle that the HIP
 * product path is compared against (and the timed CPU baseline).  It is never
 * linked into, loaded by, or called from the product library.
 *
 * Conventions (all arrays row-major, in the reference's own order):
 *   complex numbers are C99 `double complex` (= {re, im} pairs);
 *   lattice meshes are (d0, d1, d2) with linear index (i0*d1 + i1)*d2 + i2;
 *   fermionic Matsubara mesh with n_max = n_iw has 2*n_iw points n = -n_iw..n_iw-1;
 *   bosonic mesh with n_max = nw has 2*nw-1 points m = -(nw-1)..nw-1;
 *   imaginary-time mesh has ntau points tau_j = j*beta/(ntau-1), j = 0..ntau-1.
 *
 * Return codes follow include/tprf_hip.h (0 = success, negative = error).
 */
#ifndef TPRF_ORACLE_H
#define TPRF_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex or_c128;

enum { OR_FERMION = 0, OR_BOSON = 1 };
enum { OR_SIGMA_NONE = 0, OR_SIGMA_W = 1, OR_SIGMA_WK = 2 };

/* Tail-fit window (restatement of the TRIQS tail_fitter defaults). */
typedef struct {
  int n_fit;       /* total fitted frequencies (both sides) */
  int order;       /* expansion order K (coefficients 0..K) */
  int step;        /* spacing of the fitted points (in mesh indices) */
  double omega_max;/* |largest frequency| used to normalise the Vandermonde */
  int idx[128];    /* data indices of the fitted frequencies */
} or_tail_window;

int or_tail_window_make(int statistic, double beta, int n_max, or_tail_window *w);

/* a1 */
int or_lattice_dyson_g_wk(double beta, int n_iw, double mu, int nk, int n, const or_c128 *e_k,
                          int sigma_kind, const or_c128 *sigma, or_c128 *g_wk);
/* a2 / a6: direction -1 = k->r (FFTW_FORWARD, then /N_k), +1 = r->k (FFTW_BACKWARD, no scale) */
int or_fourier_lattice(long nouter, const int dims[3], long ntarget, const or_c128 *in, or_c128 *out,
                       int direction);
/* a3: omega -> tau, tail moments fitted (moments == NULL) or given ([3][ncol], m1,m2,m3) */
int or_fourier_w_to_t(int statistic, double beta, int n_max, int ntau, long ncol, const or_c128 *gw,
                      or_c128 *gt, const or_c128 *moments, or_c128 *moments_out, double *fit_error);
/* a5: tau -> omega (fit_derivatives boundary moments), nw = n_max of the output mesh */
/* tail-fit moments t_0..t_3 ([4][ncol]) of every column (split, density) */
int or_tail_moments(int statistic, double beta, int n_max, long ncol, const or_c128 *gw, or_c128 *t,
                    double *fit_error);
int or_fourier_t_to_w(int statistic, double beta, int ntau, int nw, long ncol, const or_c128 *gt,
                      or_c128 *gw);
/* a4 */
int or_chi0_tr_from_grt_PH(int ntau, const int dims[3], int n, const or_c128 *g_tr, or_c128 *chi0_tr);
/* a4' : fused product + the a5 transform */
int or_chi0_wr_from_grt_PH(double beta, int ntau, const int dims[3], int n, int nw, const or_c128 *g_tr,
                           or_c128 *chi0_wr);
/* static nu=0 trapezoid variants */
int or_chi0_w0r_from_grt_PH(double beta, int ntau, const int dims[3], int n, const or_c128 *g_tr,
                            or_c128 *chi0_wr);
int or_chi_w0r_from_chi_tr(double beta, int ntau, long nr, int n, const or_c128 *chi_tr, or_c128 *chi_wr);
/* a7 */
int or_solve_rpa_PH(long npts, int n, const or_c128 *chi0, const or_c128 *U, or_c128 *chi);
/* f4: dynamical_screened_interaction_W (static V: nk entries, dynamic: per point) */
int or_screened_interaction_W(long npts, long nk, int n, int dynamic, const or_c128 *PI, const or_c128 *V, or_c128 *W);
/* closed-form checker: Lindhard bubble (bosonic mesh n_max = nw) */
int or_lindhard_chi00(double beta, double mu, int nw, const int dims[3], int n, const or_c128 *e_k,
                      or_c128 *chi_wk);
/* Hermitian eigen-decomposition (Jacobi), exposed for its own test */
int or_eigh(int n, const or_c128 *a, double *evals, or_c128 *evecs);
/* plain 1-D DFT, exposed so the FFT can be pinned against a brute-force sum */
int or_fft1d(int n, int sign, const or_c128 *in, or_c128 *out);
/* exact inverse Vandermonde V_{m,j} = m^{j-1}, m,j = 1..8 (row-major 8x8) */
void or_vandermonde_inverse8(double *vinv);

double or_fermi(double e);
int or_num_threads(void);
void or_set_num_threads(int n);
void or_set_fit_check(int on); /* 0: no TAIL_FIT error on bad fits (for reading off the linear map) */

#ifdef __cplusplus
}
#endif
#endif
