// common.hpp -- shared device helpers and launcher declarations for libtprf_hip.
//
// Complex numbers are double2 {x = re, y = im}, the layout of std::complex<double>
// and of tprf_c128.  All arithmetic is FP64.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tprf {

using c2 = double2;

__host__ __device__ __forceinline__ c2 mk(double re, double im) { return c2{re, im}; }
__host__ __device__ __forceinline__ c2 cadd(c2 a, c2 b) { return c2{a.x + b.x, a.y + b.y}; }
__host__ __device__ __forceinline__ c2 csub(c2 a, c2 b) { return c2{a.x - b.x, a.y - b.y}; }
__host__ __device__ __forceinline__ c2 cmul(c2 a, c2 b) { return c2{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ __forceinline__ c2 cscale(c2 a, double s) { return c2{a.x * s, a.y * s}; }
__host__ __device__ __forceinline__ c2 cconj(c2 a) { return c2{a.x, -a.y}; }
__host__ __device__ __forceinline__ double cabs2(c2 a) { return a.x * a.x + a.y * a.y; }
// a / b, computed as a * conj(b) / |b|^2
__host__ __device__ __forceinline__ c2 cdiv(c2 a, c2 b) {
  double d = 1.0 / (b.x * b.x + b.y * b.y);
  return c2{(a.x * b.x + a.y * b.y) * d, (a.y * b.x - a.x * b.y) * d};
}
// a * conj(b)
__host__ __device__ __forceinline__ c2 cmulc(c2 a, c2 b) { return c2{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y}; }
// a * (i s): multiply by a purely imaginary number
__host__ __device__ __forceinline__ c2 cmul_i(c2 a, double s) { return c2{-a.y * s, a.x * s}; }
// acc + a b and acc - a b as two chained fused multiply-adds per component
// (four FP64 instructions instead of a multiply, a fused multiply-add and an
// add per component)
__host__ __device__ __forceinline__ c2 cfma(c2 acc, c2 a, c2 b) {
  return c2{fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y))};
}
__host__ __device__ __forceinline__ c2 cfms(c2 acc, c2 a, c2 b) {
  return c2{fma(-a.x, b.x, fma(a.y, b.y, acc.x)), fma(-a.x, b.y, fma(-a.y, b.x, acc.y))};
}

// Native two-double vector for values held in registers across a loop
// iteration (prefetches): copies of double2 are struct copies that can be left
// in private memory when they are assigned under a condition inside a loop,
// while this type always stays in VGPRs.
typedef double v2d __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2d ld_v2d(const c2 *p) { return *reinterpret_cast<const v2d *>(p); }
__device__ __forceinline__ void st_v2d(c2 *p, v2d v) { *reinterpret_cast<v2d *>(p) = v; }

constexpr int kMaxPass = 24;

// One FFT length: Stockham passes with the listed radices, twiddle table
// tw[k] = exp(-2 pi i k / n) in device memory.
struct FftDev {
  int n;
  int npass;
  int radix[kMaxPass];
  const c2 *tw;
};

// Least-squares tail fit of one Matsubara mesh (see tail_fit_make in capi.cpp).
constexpr int kMaxFit = 64;
constexpr int kMaxOrder = 10;  // coefficients 0..9
struct TailFitDev {
  int n_fit;             // fitted frequencies
  int ncoef;             // K + 1
  const int *idx;        // [n_fit] data indices into the Matsubara mesh
  const c2 *pinv;        // [ncoef][n_fit] least-squares operator
  const c2 *vander;      // [n_fit][ncoef] normalised Vandermonde
  double wpow[4];        // omega_max^p, p = 0..3
};

// Per (statistic, beta, L) tables of the Matsubara transforms.
struct MatsDev {
  int statistic;         // 0 fermion, 1 boson
  int L;                 // ntau - 1
  double beta;
  const double *psi;     // [3][L] pole-model functions of tau_j (a = 1)
  const c2 *phase;       // [L] exp(-i pi j / L) (fermions), unused for bosons
};

// The nu = 0 bubble in Matsubara space (static_mats.hip): per n >= 0 a table
// row of kStaticTabWidth reals (cos(w_n tau_j) / beta, j = 0..8;
// sin(w_n tau_j) / beta, j = 1..8; Re and Im of Y_q and Y'_q, q = 0..3), and
// the coefficients of the small bilinear part, passed by value.
constexpr int kStaticTabWidth = 33;
struct StaticMatsDev {
  double inv_beta;
  double e[9];   // -(dc_j + dc_{L-j})
  double f[9];   // dc_j - dc_{L-j} (f[0] unused)
  double Q[16];  // Q_pq = sum_j c_j Z_jp Z_{L-j,q}
};

// Inverse 8x8 Vandermonde of the boundary-derivative fit, passed by value.
struct Vinv8 {
  double v[64];
};

// Tables of the pruned bubble transform (bubble.hip): tau_j with j = N1*j2 + j1,
// inner DFT of length P over j2, outer twiddles over j1.
struct BubbleDev {
  int P;                 // inner DFT length (power of two, >= 2 nb)
  int ncanon;            // number of canonical r (r <= -r)
  const c2 *tw1;         // [N1][P] exp(+2 pi i j1 m(k) / L), m(k) = k or k - P
  const c2 *twp;         // [P] exp(+2 pi i p / P)
  const c2 *hmod;        // [3][2nb-1] 1 / (i nu_m - b_q) - (beta/L) sum_{j<L} psi_q(tau_j) exp(+2 pi i j m / L)
  const int *rcanon;     // [ncanon] canonical r in increasing order
};

// Tables of the input-pruned Matsubara -> tau transform (w2t.hip): L = Q M,
// the 2 n_max nonzero inputs fit one length-M DFT per output residue t < Q.
struct W2tDev {
  int M, Q;
  const c2 *twL;         // [L] exp(-2 pi i k / L)
  const c2 *twM;         // [M] exp(-2 pi i k / M)
  const double *dtail;   // [3][L] psi_q(tau_j) - (1/beta) sum_{n in mesh} exp(-i w_n tau_j) / (i w_n - b_q), real:
                         // the part of the pole model beyond the mesh, added at the outputs
  const c2 *otab;        // [L][3] per-tau output records [phase, (D_0, D_1), (D_2, 0)] (forward kernel)
};

// Canonical columns of a symmetric transform: col[i] (increasing) are
// transformed, and mcol[i] (-1: none) receives the conjugate of col[i]'s result.
struct SymCols {
  const int *col;
  const int *mcol;
  long n;
};

// Status words written by kernels (device memory, 4 doubles):
//   [0] max RMS tail-fit residual, [1] max |m0|, [2] singular-matrix flag.
struct Status {
  double *d;
};

// ---- launchers (kernels.hip) -------------------------------------------
// mu_arr (optional, device): one mu per lattice of a batch folded into k, nk_inst
// k-points each (k = inst * nk_inst + k_local); mu is then ignored
// zarr != nullptr: the nz frequencies z[i] in place of the Matsubara mesh
hipError_t launch_dyson(hipStream_t s, double beta, int n_iw, double mu, long nk, int n, const c2 *ek, int sigma_kind,
                        const c2 *sigma, c2 *g, Status st, const double *mu_arr = nullptr, long nk_inst = 1,
                        const c2 *zarr = nullptr, long nz = 0);
// out[z][ab] = (1/nk) sum_k g[z][k][ab]
hipError_t launch_k_average(hipStream_t s, long nz, long nk, int n, const c2 *g, c2 *out);
// Dyson (Sigma = 0) fused into the 16^3 k -> r transform, n <= 3:
// e_k[4096][n][n] -> g_wr[2 n_iw][4096][n][n]
// variant: 0 = one orbital row per workgroup, 1 = one column (same results)
// sym (e_k Hermitian): 1 = the frequencies w_n >= 0 only, G_ba(i w_{-n-1}, -r)
// written as conj(G_ab(i w_n, r)); 2 = the same, writing only the columns of
// r <= -r
hipError_t launch_dyson_lattice16r(hipStream_t s, double beta, int n_iw, double mu, int n, const c2 *ek, c2 *out,
                                   Status st, int variant, int sym = 0);
int dyson_lattice16r_variants(int n);
// lindhard.hip: eigenpairs of e_k - mu into ev[nk][n] and proj[nk][n][n][n]
// (workspace), then chi[nz][nk][n^4] at the nz complex frequencies z (device)
hipError_t launch_lindhard(hipStream_t s, double beta, double mu, long nz, const c2 *z, const int dims[3], int n,
                           const c2 *ek, double *ev, c2 *proj, c2 *chi);
// Size of the file-scope device array behind tprf_lindhard_chi00_noalloc, and
// its device address on `device` (which must be the current device; the
// address is looked up once per device and cached).
constexpr size_t kNoallocScratchBytes = size_t(32) << 20;
hipError_t noalloc_scratch_address(int device, void **ptr);
hipError_t launch_lattice_axis(hipStream_t s, int sign, const c2 *in, c2 *out, long outer, int len, long inner,
                               const FftDev &p, double scale);
hipError_t launch_w2t(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                      const c2 *moments, const TailFitDev &fit, const MatsDev &mt, const FftDev &p, Status st);
hipError_t launch_t2w(hipStream_t s, int statistic, double beta, int nw, long ncol, const c2 *in, c2 *out,
                      const MatsDev &mt, const FftDev &p, const Vinv8 &vinv);
// herm: G_ab(tau, r)^* = G_ba(tau, -r), only the columns of r <= -r read
hipError_t launch_chi0_wr(hipStream_t s, double beta, int nw, const int dims[3], int n, const c2 *gtr, c2 *out,
                          const MatsDev &mt, const FftDev &p, const Vinv8 &vinv, bool herm = false);
hipError_t launch_chi0_tr(hipStream_t s, int ntau, const int dims[3], int n, const c2 *gtr, c2 *out);
hipError_t launch_chi0_w0r(hipStream_t s, double beta, int ntau, const int dims[3], int n, const c2 *gtr, c2 *out);
hipError_t launch_chi_w0r(hipStream_t s, double beta, int ntau, long nr, int n, const c2 *chi_tr, c2 *out);
hipError_t launch_rpa(hipStream_t s, long npts, int n, const c2 *chi0, const c2 *U, c2 *chi, Status st);
// RPA, its mirrored form and W for 5 <= n <= kWideMaxOrb: one workgroup per
// point, [1 - Lm Rm | Lm] in LDS (32 n^4 bytes)
constexpr int kWideMaxOrb = 8;
hipError_t launch_solve_wide(hipStream_t s, int n, long npts, const c2 *Lm, int lmode, const c2 *Rm, int rmode,
                             c2 *out, Status st, long base, long nk, int mirror);
// RPA at the nonnegative bosonic frequencies of chi0[2 nb - 1][nk][n^4], each
// result also written at -nu as its conjugate transpose in the PH grouping
// screened interaction W = V inverse(1 - PI V) per point (V per k when
// dynamic = 0, per point otherwise)
hipError_t launch_screened_W(hipStream_t s, long npts, long nk, int n, int dynamic, const c2 *PI, const c2 *V, c2 *W,
                             Status st);
hipError_t launch_rpa_mirror(hipStream_t s, long nb, long nk, int n, const c2 *chi0, const c2 *U, c2 *chi,
                             Status st);
// bubble.hip: returns hipErrorNotSupported when (n, P) has no specialisation
// variant < bubble_pair_variants() selects the workgroup shape (same results);
// herm: the form for G_ab(tau, r)^* = G_ba(tau, -r), which transforms only the
// orbital pairs a <= b (hipErrorNotSupported for shapes without orbital-pair
// workgroups); nonneg: only the frequencies nu >= 0 are written
hipError_t launch_bubble_pair(hipStream_t s, double beta, int L, int nb, const int dims[3], int n, const c2 *gtr,
                              c2 *out, const BubbleDev &bd, const Vinv8 &vinv, int variant, bool herm = false,
                              bool nonneg = false);
bool bubble_pair_supported(int n, int P);
int bubble_pair_variants();
// bubble_wg.hip: the Hermitian form with one workgroup per canonical (r, -r)
// owning every canonical column (n = 2, 3; P = 128; hipErrorNotSupported
// otherwise)
hipError_t launch_bubble_wg(hipStream_t s, double beta, int L, int nb, const int dims[3], int n, const c2 *gtr,
                            c2 *out, const BubbleDev &bd, const Vinv8 &vinv, bool nonneg);
// variants usable with herm, written to list[0..count), count returned (<= 8)
int bubble_pair_herm_variants(int n, int P, int *list);
// w2t.hip: returns hipErrorNotSupported when M has no specialisation
// variant < w2t_pruned_variants() selects the workgroup width (same results)
hipError_t launch_w2t_pruned(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                             const c2 *moments, const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd,
                             Status st, int variant);
bool w2t_pruned_supported(int M);
int w2t_pruned_variants();
// the same on the canonical columns sym.col only (fermions, ncol the row
// length), writing the conjugate of each result to sym.mcol as well
hipError_t launch_w2t_pruned_sym(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                                 const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st, int variant,
                                 const SymCols &sym);
// w2t.hip: transpose of the pruned w -> tau map (same M specialisations)
hipError_t launch_w2t_adjoint(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *gbar_t,
                              c2 *gbar_w, int fitted, const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd);
// adjoint.hip: reverse-mode kernels of the static bubble + RPA at one momentum.
// nbatch lattices of equal dims are folded into the k / r index
// (k = inst * nk + k_local); per-lattice outputs (chi_q, mubar) are [nbatch][...].
hipError_t launch_q_project(hipStream_t s, const c2 *chi_r, c2 *chi_q, const int dims[3], const int q[3], int ntarget,
                            long nbatch);
hipError_t launch_q_project_adjoint(hipStream_t s, const c2 *chibar_q, c2 *chibar_r, const int dims[3], const int q[3],
                                    int ntarget, long nbatch);
hipError_t launch_bubble_static_adjoint(hipStream_t s, int ntau, const int dims[3], int n, long nbatch, const c2 *gtr,
                                        const c2 *chibar_r, const double *cw, c2 *gbar);
// part_e: nchunk * nbatch * nk * n^2 values; mu_part: nbatch * nchunk * ceil(nk / 256) doubles
hipError_t launch_dyson_adjoint(hipStream_t s, int nw, long nk, int n, long nbatch, const c2 *g, const c2 *gbar,
                                c2 *ebar, c2 *part_e, double *mu_part, double *mubar, int nchunk);
// part: nchunk * nbatch * nr * n^4 values
hipError_t launch_bubble_static(hipStream_t s, int ntau, const int dims[3], int n, long nbatch, const c2 *gtr,
                                const double *cw, c2 *part, int nchunk, c2 *chi_r);
hipError_t launch_rpa_point(hipStream_t s, long npts, int n, const c2 *chi0, const c2 *U, const c2 *chibar, c2 *chi,
                            c2 *xbar, c2 *ubar, Status st);
// kernels.hip: single-pass 3-D lattice FFT; hipErrorNotSupported for other dims
// device workspace (bytes) of the launchers below
size_t static_mats_work_bytes(long nbatch, const int dims[3], int n, int n_iw, bool adjoint);
hipError_t launch_static_mats(hipStream_t s, long nbatch, const int dims[3], int n, int n_iw, const c2 *g,
                              const double *tab, const TailFitDev &fit, const StaticMatsDev &sm, c2 *work, c2 *chi,
                              double *status);
hipError_t launch_static_mats_adjoint(hipStream_t s, long nbatch, const int dims[3], int n, int n_iw, const c2 *g,
                                      const double *tab, const TailFitDev &fit, const StaticMatsDev &sm, c2 *work,
                                      const c2 *chibar, c2 *gbar, double *status);
// static_q.hip: chi0(Q, 0) straight from e_k (G formed in registers), and its
// adjoint to (e_k, mu); proj[29][nbatch nk n^2] optional (nullptr: in work)
size_t static_q_work_bytes(long nbatch, const int dims[3], int n, int n_iw, bool adjoint);
hipError_t launch_static_q(hipStream_t s, long nbatch, const int dims[3], const int q[3], int n, int n_iw, double beta,
                           const double *mu, const c2 *ek, const double *tab, const int *fitpos, const TailFitDev &fit,
                           const StaticMatsDev &sm, c2 *work, c2 *proj, bool reuse, const c2 *chibar, c2 *chi0_q,
                           c2 *ebar, double *mubar, double *status);
// consumers.hip: GW and Eliashberg pieces (mode of launch_vertex_contract:
// 0 gw_dynamic_sigma, 1 fock (r), 2 hartree (r), 3 eliashberg gamma-F)
hipError_t launch_vertex_contract(hipStream_t s, int mode, int n, long npts, const c2 *V, long vstride, const c2 *y,
                                  c2 *out);
hipError_t launch_static_sigma_k(hipStream_t s, bool fock, int n, const int dims[3], const c2 *V, const c2 *rho,
                                 c2 *out);
hipError_t launch_tail_split(hipStream_t s, long ncol, int nmesh, const TailFitDev &fit, const c2 *x, c2 *dyn,
                             c2 *cst);
hipError_t launch_density(hipStream_t s, long ncol, int n_max, double beta, const TailFitDev &fit, const c2 *x,
                          c2 *rho, double *status);
hipError_t launch_g_delta_g(hipStream_t s, int n, long nwd, int off, const int dims[3], const c2 *g, const c2 *delta,
                            c2 *F);
hipError_t launch_eliashberg_sum(hipStream_t s, int n, int nd, int ng, const int dims[3], double scale,
                                 const c2 *gamma, const c2 *F, c2 *out);
hipError_t launch_construct_phi(hipStream_t s, int n, long npts, const c2 *chi, const c2 *U, c2 *phi);
hipError_t launch_add_static(hipStream_t s, long nw, long ncol, c2 *x, const c2 *st);
hipError_t launch_zero(hipStream_t s, long n, c2 *x);
// adjoint.hip: out[b] = sum of part[b n .. b n + n), b < nbatch
hipError_t launch_sum_partials(hipStream_t s, const double *part, long nbatch, int n, double *out);
hipError_t launch_lattice3d(hipStream_t s, int sign, const c2 *in, c2 *out, long nw, const int dims[3], long ntarget,
                            double scale, int variant = 0);
int lattice3d_variants(const int dims[3]);
int lattice3d_default_variant(const int dims[3]);

// largest LDS footprint (bytes) a Matsubara-axis kernel may request
constexpr int kMatsLdsBytes = 128 * 1024;
int mats_cols_per_block(int L);
// Lengths whose two LDS buffers exceed kMatsLdsBytes run one column per
// workgroup of kBigThreads, transformed in place in one buffer of at most
// kMatsBigLdsBytes (L + 1 <= kBigMaxN).
constexpr int kMatsBigLdsBytes = 152 * 1024;
constexpr int kBigMaxN = kMatsBigLdsBytes / 16;
constexpr int kBigThreads = 512;
bool mats_single_column(int L);
// whether the Matsubara-axis kernels take a mesh of length L on this device
bool mats_supported(int L);

}  // namespace tprf
