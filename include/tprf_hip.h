/*
 * tprf_hip.h -- C ABI of the MI355X-native lattice bubble + RPA library
 * (libtprf_hip.so).
 *
 * One entry point per function of the TRIQS/tprf 3.2 hot path
 * (c++/triqs_tprf/lattice).  The reference takes TRIQS gf objects; here each
 * gf is exploded into its mesh parameters and a data pointer laid out exactly
 * as the reference stores gf data: row-major [mesh0][mesh1][target...].
 *
 *   - every tprf_c128 pointer is a DEVICE pointer (hipMalloc'd or a torch CUDA
 *     tensor's data_ptr) on the context's device; the library never frees
 *     caller memory;
 *   - calls are asynchronous and ordered on the context's stream
 *     (tprf_ctx_set_stream); none synchronises unless stated;
 *   - a context is not thread-safe; use one per host thread;
 *   - fermionic Matsubara mesh (n_max = N): 2N points, n = -N..N-1;
 *     bosonic Matsubara mesh (n_max = N): 2N-1 points, m = -(N-1)..N-1;
 *     imaginary-time mesh: ntau points tau_j = j beta/(ntau-1);
 *     lattice dims (d0,d1,d2), linear index (i0*d1 + i1)*d2 + i2, d2 = 1 in 2-D.
 *
 * Return value: TPRF_OK or a negative TPRF_ERR_* code; tprf_ctx_last_error()
 * gives the message.  The reference raises C++ exceptions at the matching
 * places (cited per code).
 */
#ifndef TPRF_HIP_H
#define TPRF_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } tprf_c128; /* layout of std::complex<double> */
typedef struct tprf_ctx tprf_ctx;

#define TPRF_FERMION 0
#define TPRF_BOSON 1

#define TPRF_SIGMA_NONE 0
#define TPRF_SIGMA_W 1  /* sigma[w][a][b]     (g_w_cvt overload, gf.hpp:123)  */
#define TPRF_SIGMA_WK 2 /* sigma[w][k][a][b]  (g_wk_cvt overload, gf.hpp:83)  */

#define TPRF_OK 0
#define TPRF_ERR_ARG (-1)            /* bad shape or argument                                  */
#define TPRF_ERR_POSITIVE_ONLY (-2)  /* fourier_matsubara.cpp:196,221 (not representable here)  */
#define TPRF_ERR_MESH_TOO_SHORT (-3) /* L < 2 (last_index+1): fourier_matsubara.cpp:105,128,201,251 */
#define TPRF_ERR_TAIL_FIT (-4)       /* tail fit error >= 1e-3: fourier_matsubara.cpp:227      */
#define TPRF_ERR_TAIL_ORDER (-5)     /* fewer than 5 tail coefficients: fourier_matsubara.cpp:235 */
#define TPRF_ERR_SINGULAR (-6)       /* zero pivot in a Dyson or RPA inversion                  */
#define TPRF_ERR_HIP (-7)            /* HIP runtime error                                      */
#define TPRF_ERR_NOMEM (-8)
#define TPRF_ERR_UNSUPPORTED (-9)    /* size outside the kernels' limits (see each function)   */

/* ---- context ---------------------------------------------------------- */
/* The reference has no context object: plans (fourier.hpp:177, :111) are made
 * per call.  Here plans, twiddle and tail tables live in the context.       */
int tprf_ctx_create(tprf_ctx **ctx, int device);
int tprf_ctx_destroy(tprf_ctx *ctx);
int tprf_ctx_set_stream(tprf_ctx *ctx, void *hip_stream); /* hipStream_t; NULL = default stream */
/* Options (no counterpart in the reference).  TPRF_OPT_GENERIC_KERNELS = 1
 * routes every call through the generic full-length kernels instead of the
 * pruned specialisations (same results to rounding; used to cross-check). */
#define TPRF_OPT_GENERIC_KERNELS 1
/* TPRF_OPT_BUBBLE_VARIANT = v >= 0 fixes the workgroup shape of the fused
 * bubble kernel (same results); -1 (default) autotunes it: the first call per
 * shape times every shape and keeps the fastest (not while the stream is being
 * captured into a graph).                                                   */
#define TPRF_OPT_BUBBLE_VARIANT 2
/* TPRF_OPT_W2T_VARIANT: the same for the workgroup shape of the pruned
 * omega -> tau kernel (0, 1 = 256, 512 threads per workgroup; 2 = for
 * tprf_fourier_wr_to_tr_herm, one persistent workgroup per CU whose transform
 * waves never load from global memory while four loader waves stage the next
 * tile in LDS; elsewhere as 1).                                             */
#define TPRF_OPT_W2T_VARIANT 3
/* TPRF_OPT_LATTICE_VARIANT: the same for the number of target columns per
 * workgroup of the single-pass 16x16x16 lattice FFT (0, 1, 2 = 1, 2, 4
 * columns; the fused Dyson + k -> r kernel of tprf_lattice_dyson_g_wr takes
 * the value mod 2: one orbital row or one column per workgroup).  In-place
 * calls (in == out) are never timed; they use the cached or default choice. */
#define TPRF_OPT_LATTICE_VARIANT 4
/* TPRF_OPT_TRACE = 1: every entry point opens a roctx range named after it
 * (and tprf_forward_batched one per instance) over its host-side span, for
 * rocprofv3 --marker-trace next to the kernel trace.  roctx is loaded from
 * the rocprofiler SDK on first use; TPRF_ERR_UNSUPPORTED if it is absent.   */
#define TPRF_OPT_TRACE 5
int tprf_ctx_set_option(tprf_ctx *ctx, int option, int value);
/* Device scratch (no counterpart in the reference).  The static bubble, the
 * Dyson adjoint and lindhard_chi00 need scratch beyond their outputs.  By
 * default the context allocates it; that buffer only grows, and an outgrown
 * one is kept, not freed, until tprf_ctx_destroy (a graph captured earlier may
 * still use it).  A caller that captures such calls into a HIP graph binds a
 * workspace of its own for the capture, so that each graph owns its buffer and
 * two graphs never share one: tprf_ctx_set_workspace(ctx, ptr, bytes) (device
 * memory; NULL unbinds); while bound, an entry that needs more scratch than
 * `bytes` returns TPRF_ERR_NOMEM.  tprf_ctx_workspace_size reports the largest
 * scratch any call asked for since the context was made or the last reset
 * (reset != 0 clears it after reading), e.g. after an eager warm-up of the
 * calls to be captured.                                                       */
int tprf_ctx_set_workspace(tprf_ctx *ctx, void *work, size_t bytes);
int tprf_ctx_workspace_size(tprf_ctx *ctx, size_t *bytes, int reset);
const char *tprf_ctx_last_error(const tprf_ctx *ctx);
/* Largest RMS tail-fit residual of the fitted omega->tau calls since the last
 * reset (blocking: synchronises the stream).  The reference asserts < 1e-3
 * and warns above 1e-6 (fourier_matsubara.cpp:227-234); the fitted zeroth
 * moment is reported too (warning above 1e-6, :238-239).                   */
int tprf_ctx_tail_status(tprf_ctx *ctx, double *max_fit_error, double *max_abs_m0);
/* Clears the status words with a small kernel on the context's stream (no
 * memset node, so the call can be captured into a HIP graph).              */
int tprf_ctx_tail_status_reset(tprf_ctx *ctx);

/* ---- a1: Dyson --------------------------------------------------------- */
/* lattice_dyson_g_wk(mu, e_k, sigma_w | sigma_wk), c++/triqs_tprf/lattice/gf.hpp:83,123,
 * gf.cpp:91-132; Sigma = 0 (sigma_kind NONE) is lattice_dyson_g0_wk, gf.cpp:41-64.
 * e_k[nk][n][n], sigma per sigma_kind, out g_wk[2 n_iw][nk][n][n].  n <= 8.  */
int tprf_lattice_dyson_g_wk(tprf_ctx *ctx, double beta, int n_iw, double mu, long nk, int norb,
                            const tprf_c128 *e_k, int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_wk);
/* The same on a mesh given by its frequencies z[nz] (device): G[z][k] =
 * inverse((z + mu) 1 - e_k - Sigma).  With z = omega + i delta this is
 * lattice_dyson_g0_fk / lattice_dyson_g_fk (gf.hpp:103,204,225; gf.cpp:71-86,
 * 92-123, 142-148, the real-frequency mesh with its broadening); sigma per sigma_kind
 * ([nz][n][n] or [nz][nk][n][n]).  out g_zk[nz][nk][n][n].  n <= 8.         */
int tprf_lattice_dyson_g_zk(tprf_ctx *ctx, long nz, const tprf_c128 *z, double mu, long nk, int norb,
                            const tprf_c128 *e_k, int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_zk);
/* Its k average, the local Green's function g_z[nz][n][n] = (1/nk) sum_k G[z][k]
 * (lattice_dyson_g_w, lattice_dyson_g_f: gf.hpp:245,267, gf.cpp:153-204; the
 * k sum in index order).  Uses nz nk n^2 complex entries of scratch.        */
int tprf_lattice_dyson_g_z(tprf_ctx *ctx, long nz, const tprf_c128 *z, double mu, long nk, int norb,
                           const tprf_c128 *e_k, int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_z);

/* a1 followed by a2 in one call, Sigma = 0: lattice_dyson_g0_wk (gf.cpp:41-58)
 * then fourier_wk_to_wr (gf.cpp:209, fourier_lattice.cpp:39,49), as the
 * reference's driver runs them one after the other; no single counterpart.
 * e_k[nk][n][n] -> g_wr[2 n_iw][nk][n][n] (nk = dims[0] dims[1] dims[2]).
 * For a 16 x 16 x 16 lattice and n <= 3 the inversion is done inside the
 * transform's load phase and G(iw, k) is never stored; otherwise g_wr first
 * holds G(iw, k) and is transformed in place.  n <= 8.                      */
int tprf_lattice_dyson_g_wr(tprf_ctx *ctx, double beta, int n_iw, double mu, const int dims[3], int norb,
                            const tprf_c128 *e_k, tprf_c128 *g_wr);
/* The same for a Hermitian e_k (the caller's promise): on the fused 16^3 path
 * only the frequencies w_n >= 0 are inverted and transformed, and
 * G_ba(i w_{-n-1}, -r) is written as conj(G_ab(i w_n, r)) (from
 * G(i w_{-n-1}, k) = G(i w_n, k)^H); other shapes take the general path.
 * full = 0: on the fused path only the columns of r <= -r are written (all
 * that tprf_fourier_wr_to_tr_herm reads); full = 1: every column.           */
int tprf_lattice_dyson_g_wr_herm(tprf_ctx *ctx, double beta, int n_iw, double mu, const int dims[3], int norb,
                                 const tprf_c128 *e_k, tprf_c128 *g_wr, int full);

/* ---- a2 / a6: lattice FFT ---------------------------------------------- */
/* fourier_wk_to_wr, gf.hpp:276 / gf.cpp:209 -> fourier.hpp:165-197:
 *   out[w][r][t] = (1/N_k) sum_k exp(-2 pi i k.r / dims) in[w][k][t].
 * Also chi_wr_from_chi_wk (chi_imtime.cpp:294).  ntarget = n^2 (G) or n^4 (chi).
 * Every dimension must factor into primes <= 13 and be <= 4096.            */
int tprf_fourier_wk_to_wr(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *in,
                          tprf_c128 *out);
/* fourier_wr_to_wk, gf.cpp:214 -> fourier.hpp:199-231 (no scaling,
 * exp(+2 pi i k.r)); also chi_wk_from_chi_wr, chi_imtime.cpp:289-292.      */
int tprf_fourier_wr_to_wk(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *in,
                          tprf_c128 *out);

/* ---- a3: Matsubara -> imaginary time ----------------------------------- */
/* fourier_wr_to_tr(g_wr, nt), gf.hpp:296 / gf.cpp:262 -> fourier.hpp:98-130 ->
 * fourier_matsubara.cpp:194-308.  Also chi_tr_from_chi_wr (statistic BOSON,
 * chi_imtime.cpp:284).  in[nmesh][ncol] -> out[ntau][ncol] where ncol is the
 * product of the remaining dims (r and target).  ntau <= 0 selects the default
 * 6 n_max + 1.  moments: NULL = fit the tail per column (the reference's
 * fit_tail call, :225-246), else device array [3][ncol] of m1, m2, m3 (the
 * reference's explicit-moment overload, fourier.hpp:46).  L = ntau-1 must factor
 * into primes <= 13 and be <= 6144.                                          */
int tprf_fourier_wr_to_tr(tprf_ctx *ctx, int statistic, double beta, int n_max, int ntau, long ncol,
                          const tprf_c128 *in, tprf_c128 *out, const tprf_c128 *moments);
/* fourier_wr_to_tr (fermions, fitted tail) of a lattice G(i w, r)[a][b] with
 * G_ab(i w_n, r)^* = G_ba(i w_{-n-1}, -r) (the caller's promise; it holds for a
 * Hermitian e_k with no self-energy): in[2 n_max][nk][n][n] ->
 * out[ntau][nk][n][n].  No counterpart in the reference: only the columns
 * (r, a, b) with r < -r, or r = -r and a <= b, are transformed.  full = 1:
 * G_ba(tau, -r) is written as the conjugate of G_ab(tau, r), which is what the
 * full transform gives for such a G, to rounding; full = 0: the columns of
 * r > -r are not written (only r <= -r hold data; the Hermitian bubble
 * tprf_chi0_wr_from_grt_PH_herm reads no others).  TPRF_ERR_UNSUPPORTED where
 * fourier_wr_to_tr would not use its input-pruned kernel (ntau - 1 a multiple
 * of the smallest power of two M >= 2 n_max, 64 <= M <= 1024).            */
int tprf_fourier_wr_to_tr_herm(tprf_ctx *ctx, double beta, int n_max, int ntau, const int dims[3], int norb,
                               const tprf_c128 *in, tprf_c128 *out, int full);

/* ---- a5: imaginary time -> Matsubara ----------------------------------- */
/* chi_wr_from_chi_tr(chi_tr, nw), chi_imtime.hpp:87 / chi_imtime.cpp:279-282 ->
 * fourier.hpp:132-163 -> fourier_matsubara.cpp:42-184.  Also fourier_tr_to_wr
 * (statistic FERMION, gf.cpp:257).  in[ntau][ncol] -> out[nmesh(nw)][ncol].  */
int tprf_fourier_tr_to_wr(tprf_ctx *ctx, int statistic, double beta, int ntau, int nw, long ncol,
                          const tprf_c128 *in, tprf_c128 *out);

/* ---- a4: bubble product ------------------------------------------------ */
/* chi0_tr_from_grt_PH(g_tr), chi_imtime.hpp:38 / chi_imtime.cpp:88-136:
 * chi0[j][r][a][b][c][d] = G[j][r][d][a] * G[ntau-1-j][-r][b][c].            */
int tprf_chi0_tr_from_grt_PH(tprf_ctx *ctx, int ntau, const int dims[3], int norb, const tprf_c128 *g_tr,
                             tprf_c128 *chi0_tr);
/* chi0_wr_from_grt_PH(g_tr, nw), chi_imtime.hpp:40 / chi_imtime.cpp:139-182:
 * the product fused with the bosonic tau->nu transform of a5 (chi0_tr is never
 * stored).  out[2nw-1][nr][n^4].                                            */
int tprf_chi0_wr_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, int nw,
                             const tprf_c128 *g_tr, tprf_c128 *chi0_wr);
/* The same for a G with G_ab(tau, r)^* = G_ba(tau, -r) at every mesh point (the
 * caller's promise): G from a Hermitian e_k with no self-energy, or with one
 * obeying Sigma(i w)^H = Sigma(-i w).  No counterpart in the reference: it
 * reads only the columns of r <= -r (G(tau, -r) = G(tau, r)^H), transforms
 * only the orbital pairs a <= b and forms the columns of b > a from the
 * symmetry chi0_badc(nu, r) = conj of the tau-reversed product transform of
 * abcd; for such a G the result equals tprf_chi0_wr_from_grt_PH's to
 * rounding.  For a G without the symmetry the result is wrong.  full = 1:
 * every frequency is written; full = 0: the rows nu_m, m >= 0 (nw - 1 ..
 * 2 nw - 2) are, the others may be left unwritten (for a caller that uses
 * chi0(-nu) = chi0(nu)^H in the PH grouping, as tprf_solve_rpa_PH_herm does).  */
int tprf_chi0_wr_from_grt_PH_herm(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, int nw,
                                  const tprf_c128 *g_tr, tprf_c128 *chi0_wr, int full);
/* chi0_w0r_from_grt_PH(g_tr), chi_imtime.cpp:185-228 (trapezoid, nu = 0). out[1][nr][n^4] */
int tprf_chi0_w0r_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb,
                              const tprf_c128 *g_tr, tprf_c128 *chi0_wr);
/* chi_w0r_from_chi_tr(chi_tr), chi_imtime.cpp:250-277 (trapezoid, nu = 0). */
int tprf_chi_w0r_from_chi_tr(tprf_ctx *ctx, double beta, int ntau, long nr, int norb, const tprf_c128 *chi_tr,
                             tprf_c128 *chi_wr);

/* ---- a7: RPA ----------------------------------------------------------- */
/* solve_rpa_PH(chi0, U), rpa.hpp:43 / rpa.cpp:29-68: for every mesh point,
 * M_chi = inverse(1 - M_chi0 M_U) M_chi0 with M_X[(a,b),(c,d)] = X_abcd, the
 * matrix that group_indices_view(X, idx_group<0,1>, idx_group<3,2>) makes of
 * a row-major tensor (rpa.cpp:40,54-55; for U_abcd = U_badc it equals the
 * (a,b),(d,c) grouping of rpa.hpp's formula).  chi0[npts][n^4], U[n^4]
 * (device), out chi[npts][n^4].  n <= 8.                                   */
int tprf_solve_rpa_PH(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi0, const tprf_c128 *U,
                      tprf_c128 *chi);
/* solve_rpa_PH on a bosonic Matsubara mesh of 2 nb - 1 frequencies
 * (chi0, chi [2 nb - 1][nk][n^4]) when chi0_abcd(-nu, q) = conj(chi0_dcba(nu, q))
 * and U_abcd = conj(U_dcba) (the caller's promise; it holds for chi0 from a
 * Hermitian e_k and any Hermitian interaction, such as the Kanamori tensors:
 * in the grouping M[(a,b),(c,d)] = X_abcd both read M(-nu) = S M(nu)^H S with
 * S the swap (a,b) -> (b,a), and the push-through identity carries it to
 * chi = (1 - chi0 U)^-1 chi0).  No counterpart in the reference: only the nb
 * frequencies nu >= 0 of chi0 are read and solved; chi_dcba(-nu, q) is
 * written as conj(chi_abcd(nu, q)), which is what the full solve gives for
 * such inputs, to rounding.                                                 */
int tprf_solve_rpa_PH_herm(tprf_ctx *ctx, int nb, long nk, int norb, const tprf_c128 *chi0, const tprf_c128 *U,
                           tprf_c128 *chi);

/* ---- f4: screened interaction ------------------------------------------ */
/* dynamical_screened_interaction_W(PI_wk, V_k) and (PI_wk, V_wk),
 * dynamical_screened_interaction.hpp:44,82 / dynamical_screened_interaction.cpp
 * :79-128 (static V: the _opt_bubble loops) and :33-74 (dynamic V): for every
 * (frequency, k) point M_W = M_V inverse(1 - M_PI M_V), M_X[(a,b),(c,d)] =
 * X_abcd (the grouping of group_indices_view on the row-major tensors; the
 * two overloads give the same matrices).  PI[nw][nk][n^4]; V[nk][n^4]
 * (dynamic = 0) or [nw][nk][n^4] (dynamic = 1); out W[nw][nk][n^4].  The
 * frequency mesh may be Matsubara or real (the real-frequency overloads,
 * dynamical_screened_interaction.hpp:63,100, do the same per point).  n <= 8;
 * a singular denominator sets the status word as for the RPA.               */
int tprf_dynamical_screened_interaction_W(tprf_ctx *ctx, long nw, long nk, int norb, const tprf_c128 *PI,
                                          const tprf_c128 *V, int dynamic, tprf_c128 *W);

/* ---- f4: GW self-energy pieces (c++/triqs_tprf/lattice/gw.cpp) ---------- */
/* split_into_dynamic_wk_and_constant_k (lattice_utility.cpp:57-84): per column
 * of chi_wk [mesh][ncol] the zeroth coefficient of the tail fit (the window
 * and order of fourier_wr_to_tr's fit) is chi_const_k [ncol], and
 * chi_dyn_wk = chi_wk - chi_const_k.  statistic: TPRF_FERMION / TPRF_BOSON.  */
int tprf_split_into_dynamic_wk_and_constant_k(tprf_ctx *ctx, int statistic, double beta, int n_max, long ncol,
                                              const tprf_c128 *chi_wk, tprf_c128 *chi_dyn_wk,
                                              tprf_c128 *chi_const_k);
/* The density of each column of a fermionic g_w [2 n_max][ncol] (the TRIQS
 * density() that rho_k_from_g_wk, gw.cpp:36-52, applies per k): the fitted
 * moments t1..t3 regularise the Matsubara sum,
 *   rho = (1/beta) sum_n [g_n - t1/iw - t2/(iw)^2 - t3/(iw)^3] + t1/2 - t2 beta/4. */
int tprf_density(tprf_ctx *ctx, double beta, int n_max, long ncol, const tprf_c128 *g_w, tprf_c128 *rho);
/* gw_dynamic_sigma (gw.cpp:66-102): sigma_ab(tau, r) = -sum_cd W_acdb(tau, r)
 * G_cd(tau, r) at npts = ntau nr points; W [npts][n^4], g/sigma [npts][n^2].  */
int tprf_gw_dynamic_sigma(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *W_tr, const tprf_c128 *g_tr,
                          tprf_c128 *sigma_tr);
/* fock_sigma(v_r, rho_r) (gw.cpp:149-167): sigma_ab(r) = -sum_cd V_acdb(r) rho_dc(r). */
int tprf_fock_sigma_r(tprf_ctx *ctx, long nr, int norb, const tprf_c128 *v_r, const tprf_c128 *rho_r,
                      tprf_c128 *sigma_r);
/* hartree_sigma(v_k, rho_r) (gw.cpp:113-122): sigma_ab(r = 0) = sum_cd
 * V_abcd(k = 0) rho_cd(r = 0), zero elsewhere; v_k [nk][n^4] (row 0 read).     */
int tprf_hartree_sigma_r(tprf_ctx *ctx, long nr, int norb, const tprf_c128 *v_k, const tprf_c128 *rho_r,
                         tprf_c128 *sigma_r);
/* hartree_sigma(v_k, g_wk) / fock_sigma(v_k, g_wk) on the k mesh by the direct
 * sum over q (gw.cpp:124-147 / :169-195), with rho_k the density per k:
 *   fock = 0: sigma_ab(k) =  (1/N) sum_q sum_cd V_abcd(q) rho_cd(k + q)
 *   fock = 1: sigma_ab(k) = -(1/N) sum_q sum_cd V_acdb(q) rho_dc(k + q).  O(N_k^2). */
int tprf_static_sigma_k(tprf_ctx *ctx, int fock, const int dims[3], int norb, const tprf_c128 *v_k,
                        const tprf_c128 *rho_k, tprf_c128 *sigma_k);
/* add_dynamic_and_static (lattice_utility.cpp:86-101), in place:
 * x_w[w][col] += s[col].                                                       */
int tprf_add_dynamic_and_static(tprf_ctx *ctx, long nw, long ncol, tprf_c128 *x_w, const tprf_c128 *s);

/* ---- f4: linearised Eliashberg (c++/triqs_tprf/lattice/eliashberg.cpp) -- */
/* eliashberg_g_delta_g_product (eliashberg.cpp:31-58): on Delta's fermionic
 * mesh (n_max_d <= n_max_g; G read at the same frequency)
 *   F_dc(w, k) = sum_ef G_cf(w, k) conj(G_ed(w, -k)) Delta_ef(w, k).
 * g_wk [2 n_max_g][nk][n^2], delta_wk and F_wk [2 n_max_d][nk][n^2].          */
int tprf_eliashberg_g_delta_g_product(tprf_ctx *ctx, int n_max_g, int n_max_d, const int dims[3], int norb,
                                      const tprf_c128 *g_wk, const tprf_c128 *delta_wk, tprf_c128 *F_wk);
/* eliashberg_product by summation (eliashberg.cpp:92-120):
 *   Delta_ab(w, k) = -1/(2 N beta) sum_{n, q} sum_cd Gamma_cadb(w - n, k - q) F_dc(n, q)
 * with F from tprf_eliashberg_g_delta_g_product; gamma_wk on a bosonic mesh
 * of 2 n_max_gamma - 1 >= 2 (2 n_max_d) points.  O(N_w^2 N_k^2).              */
int tprf_eliashberg_product(tprf_ctx *ctx, double beta, int n_max_gamma, int n_max_g, int n_max_d, const int dims[3],
                            int norb, const tprf_c128 *gamma_wk, const tprf_c128 *g_wk, const tprf_c128 *delta_wk,
                            tprf_c128 *delta_out);
/* eliashberg_dynamic_gamma_f_product / eliashberg_constant_gamma_f_product
 * (eliashberg.cpp:175-186, 214-240): delta_ab = -1/2 sum_cd Gamma_cadb F_dc
 * per point; npts = ntau nr (dynamic: Gamma_dyn_tr and F_tr) or nr (constant:
 * Gamma_const_r and the tau = 0 row of F_tr).                                  */
int tprf_eliashberg_gamma_f_product(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *gamma, const tprf_c128 *F,
                                    tprf_c128 *delta);
/* construct_phi_wk (eliashberg.cpp:336-360): Phi = U chi U per point on the
 * matrices M[(a,b),(d,c)] = X_abcd; chi/phi [npts][n^4], U [n^4].             */
int tprf_construct_phi_wk(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi, const tprf_c128 *U,
                          tprf_c128 *phi);

/* ---- analytic bubble ----------------------------------------------------- */
/* lindhard_chi00(e_k, mesh, mu) and lindhard_chi00(e_k, mesh, beta, mu, delta),
 * lattice_desc.py:411,495 / lindhard_chi00.cpp:34-108:
 *   chi[m][q][a][b][c][d] = 1/N_k sum_{k,i,j} U_k(a,i) U_k^*(d,i) U_{k+q}(c,j) U_{k+q}^*(b,j)
 *                           [f(beta e_ki) - f(beta e_{k+q,j})] / (z_m + e_{k+q,j} - e_ki)
 * with (e_ki, U_k) the eigenpairs of e_k - mu; z[nz] (device) holds the
 * complex frequencies: i nu_m for a bosonic Matsubara mesh, omega + i delta
 * on a real-frequency mesh (the reference's w + idelta); where |z_m| and
 * |e_{k+q,j} - e_ki| are both below 1e-10 the factor is the limit
 * beta / (4 cosh^2(beta e_ki / 2)).  e_k[nk][n][n] Hermitian, chi[nz][nk][n^4].
 * O(N_k^2) work.  n <= 8.                                                    */
int tprf_lindhard_chi00(tprf_ctx *ctx, double beta, double mu, long nz, const tprf_c128 *z, const int dims[3],
                        int norb, const tprf_c128 *e_k, tprf_c128 *chi);
/* No counterpart in the reference.  tprf_lindhard_chi00 for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): the same kernels and the same result, with
 * the workspace (eigenvalues and band projectors, nk (8 n + 16 n^3) bytes) in
 * a fixed 32 MiB file-scope device array of the library instead of the
 * context's allocated scratch.  The call allocates nothing and frees nothing;
 * larger sizes return TPRF_ERR_UNSUPPORTED.  Each GPU holds one such array
 * until the process ends, shared by every call on that GPU: calls that
 * overlap on one GPU (other streams, other contexts or threads) race.        */
int tprf_lindhard_chi00_noalloc(tprf_ctx *ctx, double beta, double mu, long nz, const tprf_c128 *z,
                                const int dims[3], int norb, const tprf_c128 *e_k, tprf_c128 *chi);

/* ---- batched forward chain ---------------------------------------------- */
/* Many lattices of uneven (dims, norb, n_iw, beta, nb) through the whole
 * forward chain in one call (no counterpart in
 * the reference, whose driver handles one lattice per call,
 * lattice_utils.py:191-269, followed by solve_rpa_PH, rpa.cpp:29-68).  Per
 * instance: Dyson + k -> r (tprf_lattice_dyson_g_wr), omega -> tau
 * (tprf_fourier_wr_to_tr), the fused bubble (tprf_chi0_wr_from_grt_PH), r -> q
 * and the RPA, each in the symmetric form its flags allow:
 *   TPRF_CHAIN_HERMITIAN_EK  e_k Hermitian (G(tau, r) is then stored for
 *                            r <= -r only where the pruned omega -> tau fits);
 *   TPRF_CHAIN_HERMITIAN_U   with the above, U_abcd = conj(U_dcba) (chi0_wr and
 *                            chi0_wk then hold the rows nu >= 0 only; chi_wk
 *                            holds every frequency).
 * Every pointer is device memory the caller owns, in the reference layout;
 * tprf_chain_buffer_sizes gives the element counts of g_wr, g_tr, chi0_wr,
 * chi0_wk and chi_wk.  With nstreams > 1 the instances (largest first, each on
 * the least-loaded lane) run on nstreams - 1 library streams next to the
 * context's, forked from and joined to it by events, so the call as a whole
 * is ordered on the context's stream and can be captured into one HIP graph
 * from it.  nstreams <= 16.  The kernel autotuner settles its choices on the
 * first call that is not being captured.                                    */
#define TPRF_CHAIN_HERMITIAN_EK 1
#define TPRF_CHAIN_HERMITIAN_U 2
typedef struct {
  double beta;
  int n_iw;    /* fermionic n_max */
  int ntau;    /* <= 0: 6 n_iw + 1 */
  int nb;      /* bosonic n_max: 2 nb - 1 frequencies */
  int dims[3];
  int norb;
  double mu;
  int flags;
} tprf_chain_desc;
typedef struct {
  const tprf_c128 *e_k; /* [nk][n][n] */
  const tprf_c128 *U;   /* [n][n][n][n] */
  tprf_c128 *g_wr;      /* [2 n_iw][nk][n][n]      workspace */
  tprf_c128 *g_tr;      /* [ntau][nk][n][n]        workspace */
  tprf_c128 *chi0_wr;   /* [2 nb - 1][nk][n^4]     workspace */
  tprf_c128 *chi0_wk;   /* [2 nb - 1][nk][n^4]     output    */
  tprf_c128 *chi_wk;    /* [2 nb - 1][nk][n^4]     output    */
} tprf_chain_ptrs;
int tprf_chain_buffer_sizes(const tprf_chain_desc *desc, long counts[5]);
int tprf_forward_batched(tprf_ctx *ctx, int n_inst, const tprf_chain_desc *desc, const tprf_chain_ptrs *ptrs,
                         int nstreams);

/* ---- differentiation: static chi_RPA(Q, i nu = 0) and its adjoints ------
 * No counterpart in the reference (TPRF has no derivatives).  Each adjoint is
 * the conjugate transpose of the complex-linear map named, or the reverse-mode
 * derivative of a nonlinear one, in the convention dL = Re sum conj(xbar) dx
 * (xbar = dL/dRe x + i dL/dIm x, torch's complex gradient).  All arrays are
 * device memory; scalars returned through device pointers.                  */

/* Batches.  nbatch lattices with the same mesh (beta, n_iw, dims, norb) are
 * folded into the k (or r) index, k = inst * nk + k_local, with the mesh
 * index outermost: G[w][inst][k][n][n], G(tau)[j][inst][r][n][n],
 * chi0_r[inst][r][n^4], chi_q[inst][n^4].  Every lattice transform and the
 * omega <-> tau transforms then treat the batch as nbatch times more columns
 * (tprf_fourier_wk_to_wr with nw * nbatch; tprf_fourier_wr_to_tr with
 * ncol = nbatch * nk * n^2).                                                 */

/* lattice_dyson_g_wk (gf.cpp:91-132) for a batch, one mu per lattice:
 * mu[nbatch] (device), e_k[nbatch][nk][n][n] -> g_wk[2 n_iw][nbatch][nk][n][n]. */
int tprf_lattice_dyson_g_wk_batched(tprf_ctx *ctx, double beta, int n_iw, const double *mu, long nbatch, long nk,
                                    int norb, const tprf_c128 *e_k, tprf_c128 *g_wk);
/* Adjoint of lattice_dyson_g_wk(mu, e_k) (gf.cpp:91-132) given G and Gbar:
 * ebar_k[k] = sum_w G^H Gbar G^H, mubar[inst] = -Re sum_{w,k} tr(G^H Gbar G^H).
 * g_wk, gbar_wk [2 n_iw][nbatch][nk][n][n]; ebar_k [nbatch][nk][n][n];
 * mubar [nbatch] doubles.  norb <= 3.                                       */
int tprf_lattice_dyson_g_wk_adjoint(tprf_ctx *ctx, int n_iw, long nk, int norb, long nbatch, const tprf_c128 *g_wk,
                                    const tprf_c128 *gbar_wk, tprf_c128 *ebar_k, double *mubar);
/* Adjoint of fourier_wk_to_wr (fourier_lattice.cpp:49): gbar_wk =
 * FFT_{+}(gbar_wr) / N_k.                                                    */
int tprf_fourier_wk_to_wr_adjoint(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *gbar_wr,
                                  tprf_c128 *gbar_wk);
/* Adjoint of fourier_wr_to_tr (fourier_matsubara.cpp:218-308) with fitted
 * moments (fitted = 1: the least-squares tail fit is differentiated too) or
 * given moments (fitted = 0: the moments are constants).  gbar_tr[ntau][ncol]
 * -> gbar_wr[nmesh][ncol].  TPRF_ERR_UNSUPPORTED unless ntau - 1 is a multiple
 * of the smallest power of two M >= 2 n_max, 64 <= M <= 1024.                */
int tprf_fourier_wr_to_tr_adjoint(tprf_ctx *ctx, int statistic, double beta, int n_max, int ntau, long ncol,
                                  const tprf_c128 *gbar_tr, tprf_c128 *gbar_wr, int fitted);
/* nu = 0 row of chi0_wr_from_grt_PH (chi_imtime.cpp:139-182): the bosonic
 * tau -> nu transform at nu = 0 is a real linear functional
 * chi0_abcd(r) = sum_j c_j G_da(tau_j, r) G_bc(tau_{L-j}, -r);
 * g_tr[ntau][nbatch][nr][n][n] -> chi0_r[nbatch][nr][n^4].  norb <= 4.       */
int tprf_chi0_static_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, long nbatch,
                                 const tprf_c128 *g_tr, tprf_c128 *chi0_r);
/* Adjoint of tprf_chi0_static_from_grt_PH with respect to g_tr:
 * chibar_r[nbatch][nr][n^4] -> gbar_tr[ntau][nbatch][nr][n^2] (g_tr is the
 * forward input).                                                            */
int tprf_chi0_static_from_grt_PH_adjoint(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb,
                                         long nbatch, const tprf_c128 *g_tr, const tprf_c128 *chibar_r,
                                         tprf_c128 *gbar_tr);
/* The same nu = 0 bubble of the fitted omega -> tau transform (fourier_wr_to_tr,
 * fourier_matsubara.cpp:218-308) followed by chi0_wr_from_grt_PH at nu = 0,
 * from G(i w, r) directly: G(tau, r) is never formed.  The composite map is
 * a bilinear form in Matsubara space, chi0_abcd(r) = -(1/beta) sum_n
 * G_da(i w_n, r) G_bc(i w_n, -r) plus a correction of rank about 30 from the
 * tail model and the boundary-derivative fit (exact algebra, equal to the
 * tau route to rounding).  g_wr[2 n_iw][nbatch][nr][n][n] -> chi0_r as above;
 * the tail fit's residual and zeroth moment go to the status words as for
 * fourier_wr_to_tr.  ntau <= 0: 6 n_iw + 1; ntau >= 18; norb <= 3.  The
 * adjoint takes chibar_r and writes gbar_wr [2 n_iw][nbatch][nr][n][n].      */
int tprf_chi0_static_from_gwr_PH(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], int norb,
                                 long nbatch, const tprf_c128 *g_wr, tprf_c128 *chi0_r);
int tprf_chi0_static_from_gwr_PH_adjoint(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], int norb,
                                         long nbatch, const tprf_c128 *g_wr, const tprf_c128 *chibar_r,
                                         tprf_c128 *gbar_wr);
/* The static chain at one momentum in one call, straight from e_k:
 * lattice_dyson_g_wk (gf.cpp:91-132) -> fourier_wk_to_wr (fourier_lattice.cpp:49)
 * -> fourier_wr_to_tr (fourier_matsubara.cpp:218-308) -> the nu = 0 row of
 * chi0_wr_from_grt_PH (chi_imtime.cpp:139-182) -> one momentum of
 * chi_wk_from_chi_wr (fourier_lattice.cpp:55-58), as
 *   chi0_abcd(Q, 0) = (1/nk) sum_k B(G_da(iw, k), G_bc(iw, k - Q))
 * with B the nu = 0 bilinear form of tprf_chi0_static_from_gwr_PH; G(iw, k)
 * is evaluated in registers and never stored.  mu[nbatch] and e_k
 * [nbatch][nk][n][n] are device arrays, Q = 2 pi q / dims (q integers, any
 * sign), chi0_q [nbatch][n^4].  proj (NULL or [29][nbatch][nk][n][n] complex)
 * receives the per-k tail-fit projections, which the adjoint then reuses
 * instead of recomputing them.  The adjoint takes chibar_q [nbatch][n^4] and
 * writes ebar_k [nbatch][nk][n][n] and mubar[nbatch] (dL = Re sum conj(xbar) dx).
 * The tail-fit status words are those of fourier_wr_to_tr.  n = 1..3.        */
int tprf_chi0_static_q_from_ek(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], const int q[3],
                               int norb, long nbatch, const double *mu, const tprf_c128 *e_k, tprf_c128 *chi0_q,
                               tprf_c128 *proj);
int tprf_chi0_static_q_from_ek_adjoint(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3],
                                       const int q[3], int norb, long nbatch, const double *mu, const tprf_c128 *e_k,
                                       const tprf_c128 *chibar_q, const tprf_c128 *proj, tprf_c128 *ebar_k,
                                       double *mubar);
/* One momentum of fourier_wr_to_wk (fourier_lattice.cpp:55-58):
 * out[inst][t] = sum_r exp(+2 pi i Q.r / dims) in[inst][r][t], Q = q[0..2]
 * (integers), and its adjoint gbar_r[inst][r][t] = exp(-2 pi i Q.r / dims)
 * gbar_q[inst][t].  ntarget <= 65535.                                        */
int tprf_fourier_r_to_q(tprf_ctx *ctx, const int dims[3], const int q[3], long ntarget, long nbatch,
                        const tprf_c128 *in_r, tprf_c128 *out_q);
int tprf_fourier_r_to_q_adjoint(tprf_ctx *ctx, const int dims[3], const int q[3], long ntarget, long nbatch,
                                const tprf_c128 *gbar_q, tprf_c128 *gbar_r);
/* solve_rpa_PH (rpa.cpp:29-64) at npts points with a U per point, and its
 * reverse-mode derivative: with chibar given (may be NULL for forward only),
 * chi0bar = A^-H chibar (1 + U chi)^H and Ubar = chi0^H A^-H chibar chi^H in
 * the PH matrix grouping, A = 1 - chi0 U.  chi0, U, chibar, chi, chi0bar,
 * Ubar [npts][n^4].  norb <= 4.                                              */
int tprf_solve_rpa_PH_adjoint(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi0, const tprf_c128 *U,
                              const tprf_c128 *chibar, tprf_c128 *chi, tprf_c128 *chi0bar, tprf_c128 *Ubar);

#ifdef __cplusplus
}
#endif
#endif
