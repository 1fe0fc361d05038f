// static_mats.hip -- the nu = 0 bubble of the imaginary-time route evaluated in
// Matsubara space, and its adjoint.
//
// The static bubble of the chain fourier_wr_to_tr (fitted tail,
// fourier_matsubara.cpp:218-308) then chi0_wr_from_grt_PH at nu = 0
// (chi_imtime.cpp:139-182 with the bosonic transform of
// fourier_matsubara.cpp:42-184) is
//   chi0_abcd(r) = sum_{j=0..L} c_j (T g1)_j (T g2)_{L-j},
//   g1 = G_da(i w, r), g2 = G_bc(i w, -r),
// with T the fermionic omega -> tau map and c_j the real weights of the tau ->
// nu = 0 transform.  T is the scaled DFT F_jn = exp(-i w_n tau_j) / beta (row L
// = -row 0, the endpoint rule) plus the pole model, rank four through the
// fitted moments mu = (a_1, a_2, a_3, m_1): T = F + Z mu with the real table
// Z_jq = D_q(tau_j) (j < L), Z_Lq = -D_q(0), Z_L3 = -1.  Since sum_{j<L}
// F_jn F_{L-j,m} = -delta_nm / beta (discrete orthogonality; 2 n_iw <= L),
//   chi0 = -(1/beta) sum_n g1_n g2_n
//        + sum_{j in S} dc_j (F_j g1) (F_{L-j} g2)      dc_j = c_j - beta/L [j < L]
//        + mu1 . (Y g2) + (Y' g1) . mu2 + mu1^T Q mu2
// with S = {0..8, L-8..L} (the boundary-derivative corrections of c) and the
// small tables Y_q = sum_j c_j Z_jq F_{L-j}, Y'_q = sum_j c_j Z_{L-j,q} F_j,
// Q_pq = sum_j c_j Z_jp Z_{L-j,q}.  Every row of F, Y, Y' has
// R_{-n-1} = conj(R_n), so each projection is two real dot products over the
// n >= 0 half with sigma_n = g_n + g_{-n-1} and delta_n = g_n - g_{-n-1}.
// G(tau, r) is never formed: one pass over G(i w, r) gives per column the 29
// projections q (C_j = sum cos_j sigma, S_j = sum sin_j delta, y = Y g,
// y' = Y' g, mu), and a second pass forms the Matsubara sums and the small
// bilinear correction of every (r, -r) pair.  Both passes split the frequency
// sum into chunks (a grid dimension) whose partial sums a finalising kernel
// adds in a fixed order, so that one lattice fills the GPU as a batch does and
// results do not depend on scheduling.
#include "common.hpp"
#include "mats_device.hpp"
#include "static_device.hpp"

namespace tprf {

namespace {
constexpr int kTab = kStaticTabWidth;  // per n >= 0: cos 9, sin 8, ReY 4, ImY 4, ReY' 4, ImY' 4
}  // namespace

// 1a. Partial projections of every column over the frequency chunk
//     blockIdx.y (n >= 0 in [k chunk, (k + 1) chunk)); g[2 n_iw][ncol],
//     part[sk][33][ncol]: C_0..8, S_1..8, YA, YB, PA, PB (y = YA + i YB,
//     y' = PA + i PB).
__global__ void __launch_bounds__(256) k_static_proj_part(long ncol, int n_iw, int chunk, const c2 *__restrict__ g,
                                                          const double *__restrict__ tab, c2 *__restrict__ part) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int n0 = blockIdx.y * chunk, n1 = min(n_iw, n0 + chunk);
  c2 C[9], S[8], YA[4], YB[4], PA[4], PB[4];
#pragma unroll
  for (int i = 0; i < 9; ++i) C[i] = c2{0, 0};
#pragma unroll
  for (int i = 0; i < 8; ++i) S[i] = c2{0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) YA[i] = YB[i] = PA[i] = PB[i] = c2{0, 0};
  const c2 *gp = g + (long)n_iw * ncol + col, *gm = g + (long)(n_iw - 1) * ncol + col;
#pragma unroll 2
  for (int n = n0; n < n1; ++n) {
    const v2d a = ld_v2d(gp + (long)n * ncol), b = ld_v2d(gm - (long)n * ncol);
    const c2 sg = c2{a.x + b.x, a.y + b.y}, dl = c2{a.x - b.x, a.y - b.y};
    const double *t = tab + (long)n * kTab;
#pragma unroll
    for (int j = 0; j < 9; ++j) C[j] = cfma(C[j], c2{t[j], 0.0}, sg);
#pragma unroll
    for (int j = 0; j < 8; ++j) S[j] = cfma(S[j], c2{t[9 + j], 0.0}, dl);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      YA[i] = cfma(YA[i], c2{t[17 + i], 0.0}, sg);
      YB[i] = cfma(YB[i], c2{t[21 + i], 0.0}, dl);
      PA[i] = cfma(PA[i], c2{t[25 + i], 0.0}, sg);
      PB[i] = cfma(PB[i], c2{t[29 + i], 0.0}, dl);
    }
  }
  c2 *o = part + (long)blockIdx.y * 33 * ncol + col;
#pragma unroll
  for (int j = 0; j < 9; ++j) o[(long)j * ncol] = C[j];
#pragma unroll
  for (int j = 0; j < 8; ++j) o[(long)(9 + j) * ncol] = S[j];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    o[(long)(17 + i) * ncol] = YA[i];
    o[(long)(21 + i) * ncol] = YB[i];
    o[(long)(25 + i) * ncol] = PA[i];
    o[(long)(29 + i) * ncol] = PB[i];
  }
}

// 1b. The projections q[29][ncol] of every column: the chunks' partial sums
//     added in order, and the fitted tail (fourier_matsubara.cpp:225-246):
//     coef = pinv g[idx], m_k = omega_max^k coef_k, mu = (a_1, a_2, a_3, m_1),
//     with the fit's residual and zeroth moment reported as omega -> tau does.
__global__ void __launch_bounds__(256) k_static_proj_final(long ncol, int sk, const c2 *__restrict__ part,
                                                           const c2 *__restrict__ g, TailFitDev fit,
                                                           c2 *__restrict__ q, double *status) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  auto sum = [&](int slot) {
    c2 v = c2{0, 0};
    for (int k = 0; k < sk; ++k) v = cadd(v, part[((long)k * 33 + slot) * ncol + col]);
    return v;
  };
  c2 *o = q + col;
  for (int j = 0; j < 17; ++j) o[(long)j * ncol] = sum(j);  // C_0..8 and S_1..8 keep their slots
  for (int i = 0; i < 4; ++i) {
    o[(long)(kQY + i) * ncol] = cadd(sum(17 + i), cmul_i(sum(21 + i)));
    o[(long)(kQYp + i) * ncol] = cadd(sum(25 + i), cmul_i(sum(29 + i)));
  }
  c2 coef[kMaxOrder];
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k) coef[k] = c2{0, 0};
  for (int f = 0; f < fit.n_fit; ++f) {
    const c2 x = g[(long)fit.idx[f] * ncol + col];
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k)
      if (k < fit.ncoef) coef[k] = cfma(coef[k], fit.pinv[(long)k * fit.n_fit + f], x);
  }
  double r2 = 0.0;
  for (int f = 0; f < fit.n_fit; ++f) {
    c2 v = g[(long)fit.idx[f] * ncol + col];
#pragma unroll
    for (int k = 0; k < kMaxOrder; ++k)
      if (k < fit.ncoef) v = cfms(v, fit.vander[(long)f * fit.ncoef + k], coef[k]);
    r2 += cabs2(v);
  }
  atomic_max_nonneg(&status[0], sqrt(r2 / fit.n_fit));
  atomic_max_nonneg(&status[1], sqrt(cabs2(coef[0])));
  const c2 m1 = cscale(coef[1], fit.wpow[1]);
  c2 a[3];
  Poles<0>::coef(m1, cscale(coef[2], fit.wpow[2]), cscale(coef[3], fit.wpow[3]), a);
#pragma unroll
  for (int i = 0; i < 3; ++i) o[(long)(kQM + i) * ncol] = a[i];
  o[(long)(kQM + 3) * ncol] = m1;
}

// 2a. Partial Matsubara sums of every (instance, r, d) over the mesh chunk
//     blockIdx.y: part2[k][inst][r][abcd] = sum_m G_da(i w_m, r) G_bc(i w_m, -r).
template <int N>
__global__ void __launch_bounds__(256) k_static_pair_part(long nbatch, long nk, int d0, int d1, int d2, int nw,
                                                          int chunk, const c2 *__restrict__ g, c2 *__restrict__ part2) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nbatch * nk * N) return;
  const int d = (int)(t % N);
  const long ir = t / N, inst = ir / nk, r = ir - inst * nk;
  const long mr = minus_r(r, d0, d1, d2);
  const long ncol = nbatch * nk * n2;
  const long base_r = (inst * nk + r) * n2, base_m = (inst * nk + mr) * n2;
  const int m0 = blockIdx.y * chunk, m1 = min(nw, m0 + chunk);
  c2 acc[N][n2];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int s = 0; s < n2; ++s) acc[a][s] = c2{0, 0};
  for (int m = m0; m < m1; ++m) {
    const c2 *row = g + (long)m * ncol;
    v2d x[N], y[n2];
#pragma unroll
    for (int a = 0; a < N; ++a) x[a] = ld_v2d(row + base_r + d * N + a);
#pragma unroll
    for (int s = 0; s < n2; ++s) y[s] = ld_v2d(row + base_m + s);
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int s = 0; s < n2; ++s) acc[a][s] = cfma(acc[a][s], c2{x[a].x, x[a].y}, c2{y[s].x, y[s].y});
  }
  c2 *out = part2 + ((long)blockIdx.y * nbatch * nk + inst * nk + r) * n4;
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int s = 0; s < n2; ++s) out[((a * N + s / N) * N + s % N) * N + d] = acc[a][s];
}

// 2b. chi0_abcd(r) per output: -(1/beta) x the chunks' sums + the correction.
template <int N>
__global__ void __launch_bounds__(256) k_static_pair_final(long nbatch, long nk, int d0, int d1, int d2, int sk,
                                                           const c2 *__restrict__ part2, const c2 *__restrict__ q,
                                                           StaticMatsDev sm, c2 *__restrict__ chi) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nout = nbatch * nk * n4;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout) return;
  const int abcd = (int)(t % n4);
  const long ir = t / n4, inst = ir / nk, r = ir - inst * nk;
  const int d = abcd % N, c = (abcd / N) % N, b = (abcd / n2) % N, a = abcd / (n2 * N);
  const long mr = minus_r(r, d0, d1, d2);
  const long ncol = nbatch * nk * n2;
  c2 v = c2{0, 0};
  for (int k = 0; k < sk; ++k) v = cadd(v, part2[(long)k * nout + t]);
  chi[t] = cadd(cscale(v, -sm.inv_beta),
                corr_value(q + (inst * nk + r) * n2 + d * N + a, q + (inst * nk + mr) * n2 + b * N + c, ncol, sm));
}

// 3a. Adjoint, per column (x, y) of r: kappa_s = chibar(r)[y, s1, s2, x] +
//     chibar(-r)[s2, x, y, s1] (the Matsubara sums' weights of the partner
//     columns s of -r) and the correction's gradient qbar (29 slots),
//     qbar += chibar conj(d corr / d q) over both roles
//     (dL = Re sum conj(xbar) dx).
template <int N>
__global__ void __launch_bounds__(256) k_static_qbar(long nbatch, long nk, int d0, int d1, int d2,
                                                     const c2 *__restrict__ q, const c2 *__restrict__ chibar,
                                                     StaticMatsDev sm, c2 *__restrict__ qbar, c2 *__restrict__ kap) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int xy = (int)(col % n2), x = xy / N, y = xy - x * N;
  const long ir = col / n2, inst = ir / nk, r = ir - inst * nk;
  const long mr = minus_r(r, d0, d1, d2);
  const long base_m = (inst * nk + mr) * n2;
  const c2 *cb_r = chibar + (inst * nk + r) * n4, *cb_m = chibar + (inst * nk + mr) * n4;
  auto w_role1 = [&](int s) { return cb_r[((y * N + s / N) * N + s % N) * N + x]; };
  auto w_role2 = [&](int s) { return cb_m[((s % N * N + x) * N + y) * N + s / N]; };
  for (int s = 0; s < n2; ++s) kap[(long)s * ncol + col] = cadd(w_role1(s), w_role2(s));
  c2 qb[kQSlots];
#pragma unroll
  for (int k = 0; k < kQSlots; ++k) qb[k] = c2{0, 0};
#pragma unroll 1
  for (int s = 0; s < n2; ++s) {
    qbar_role1(qb, w_role1(s), q + base_m + s, ncol, sm);
    qbar_role2(qb, w_role2(s), q + base_m + s, ncol, sm);
  }
#pragma unroll
  for (int k = 0; k < 29; ++k) qbar[(long)k * ncol + col] = qb[k];
}

// 3b. gbar of every column over the frequency chunk blockIdx.y (n >= 0 and
//     its partner -n-1): -(1/beta) sum_s kappa_s conj(g_s(-r)) and qbar spread
//     back over the sigma / delta dot products.
template <int N>
__global__ void __launch_bounds__(256) k_static_expand(long nbatch, long nk, int d0, int d1, int d2, int n_iw,
                                                       int chunk, const c2 *__restrict__ g,
                                                       const double *__restrict__ tab, const c2 *__restrict__ qbar,
                                                       const c2 *__restrict__ kap, double inv_beta,
                                                       c2 *__restrict__ gbar) {
  constexpr int n2 = N * N;
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const long ir = col / n2, inst = ir / nk, r = ir - inst * nk;
  const long mr = minus_r(r, d0, d1, d2);
  const c2 *gm_base = g + (inst * nk + mr) * n2;
  c2 qb[29], kp[n2];
#pragma unroll
  for (int k = 0; k < 29; ++k) qb[k] = qbar[(long)k * ncol + col];
#pragma unroll
  for (int s = 0; s < n2; ++s) kp[s] = kap[(long)s * ncol + col];
  const int n0 = blockIdx.y * chunk, n1 = min(n_iw, n0 + chunk);
  c2 *o = gbar + col;
  for (int n = n0; n < n1; ++n) {
    const long ip = (long)(n_iw + n) * ncol, im = (long)(n_iw - 1 - n) * ncol;
    v2d yp[n2], ym[n2];
#pragma unroll
    for (int s = 0; s < n2; ++s) {
      yp[s] = ld_v2d(gm_base + ip + s);
      ym[s] = ld_v2d(gm_base + im + s);
    }
    c2 vp = c2{0, 0}, vm = c2{0, 0};
#pragma unroll
    for (int s = 0; s < n2; ++s) {
      vp = cadd(vp, cmulc(kp[s], c2{yp[s].x, yp[s].y}));
      vm = cadd(vm, cmulc(kp[s], c2{ym[s].x, ym[s].y}));
    }
    c2 sb, db;
    qbar_spread(qb, tab + (long)n * kTab, sb, db);
    o[ip] = cadd(cscale(vp, -inv_beta), cadd(sb, db));
    o[im] = cadd(cscale(vm, -inv_beta), csub(sb, db));
  }
}

// 3c. The fit's part of gbar, after 3b: mu = (a_1, a_2, a_3, m_1),
//     a = (m1 - m3, (m2 + m3)/2, (m3 - m2)/2), m_k = omega_max^k coef_k,
//     coef = pinv g[idx]: gbar[idx_f] += sum_k conj(pinv_kf) omega_max^k mbar_k.
__global__ void __launch_bounds__(256) k_static_fit_adjoint(long ncol, TailFitDev fit, const c2 *__restrict__ qbar,
                                                            c2 *__restrict__ gbar) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  c2 qm[kQSlots], mb[4];
#pragma unroll
  for (int k = kQM; k < kQSlots; ++k) qm[k] = qbar[(long)k * ncol + col];
  mu_to_moment_bar(qm, mb);
  for (int f = 0; f < fit.n_fit; ++f) {
    c2 v = c2{0, 0};
#pragma unroll
    for (int k = 1; k <= 3; ++k) v = cadd(v, cmulc(cscale(mb[k], fit.wpow[k]), fit.pinv[(long)k * fit.n_fit + f]));
    c2 *p = gbar + (long)fit.idx[f] * ncol + col;
    *p = cadd(*p, v);
  }
}

namespace {
constexpr long kStaticTargetThreads = 1L << 18;

// chunks of a frequency sum of length len for `threads` threads: enough to put
// about 2^18 threads in flight, but at least min_len frequencies per chunk
int chunks_for(long threads, int len, int min_len, int *chunk) {
  long sk = (kStaticTargetThreads + threads - 1) / threads;
  const long cap = len / min_len > 1 ? len / min_len : 1;
  if (sk > cap) sk = cap;
  if (sk < 1) sk = 1;
  *chunk = (int)((len + sk - 1) / sk);
  return (len + *chunk - 1) / *chunk;
}
}  // namespace

size_t static_mats_work_bytes(long nbatch, const int dims[3], int n, int n_iw, bool adjoint) {
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long ncol = nbatch * nk * n * n, nout = nbatch * nk * n * n * n * n;
  int c1, c2_;
  const int sk = chunks_for(ncol, n_iw, 32, &c1);
  const int sk2 = chunks_for(nbatch * nk * n, 2 * n_iw, 32, &c2_);
  size_t elems = (size_t)(33 * sk + 29) * ncol;
  elems += adjoint ? (size_t)(29 + n * n) * ncol : (size_t)sk2 * nout;
  return elems * sizeof(c2);
}

hipError_t launch_static_mats(hipStream_t s, long nbatch, const int dims[3], int n, int n_iw, const c2 *g,
                              const double *tab, const TailFitDev &fit, const StaticMatsDev &sm, c2 *work, c2 *chi,
                              double *status) {
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long ncol = nbatch * nk * n * n, nout = nbatch * nk * n * n * n * n;
  int chunk, chunk2;
  const int sk = chunks_for(ncol, n_iw, 32, &chunk);
  const int sk2 = chunks_for(nbatch * nk * n, 2 * n_iw, 32, &chunk2);
  c2 *part = work, *q = work + (long)33 * sk * ncol, *part2 = q + 29 * ncol;
  const unsigned bc = (unsigned)((ncol + 255) / 256);
  hipLaunchKernelGGL(k_static_proj_part, dim3(bc, sk), dim3(256), 0, s, ncol, n_iw, chunk, g, tab, part);
  hipLaunchKernelGGL(k_static_proj_final, dim3(bc), dim3(256), 0, s, ncol, sk, part, g, fit, q, status);
  const dim3 gp((unsigned)((nbatch * nk * n + 255) / 256), sk2), gf((unsigned)((nout + 255) / 256));
  switch (n) {
#define TPRF_SM_PAIR(N_)                                                                                               case N_:                                                                                                               hipLaunchKernelGGL(k_static_pair_part<N_>, gp, dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2], 2 * n_iw,                        chunk2, g, part2);                                                                                hipLaunchKernelGGL(k_static_pair_final<N_>, gf, dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2], sk2,                            part2, q, sm, chi);                                                                               break;
    TPRF_SM_PAIR(1)
    TPRF_SM_PAIR(2)
    TPRF_SM_PAIR(3)
#undef TPRF_SM_PAIR
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_static_mats_adjoint(hipStream_t s, long nbatch, const int dims[3], int n, int n_iw, const c2 *g,
                                      const double *tab, const TailFitDev &fit, const StaticMatsDev &sm, c2 *work,
                                      const c2 *chibar, c2 *gbar, double *status) {
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long ncol = nbatch * nk * n * n;
  int chunk, chunk3;
  const int sk = chunks_for(ncol, n_iw, 32, &chunk);
  const int sk3 = chunks_for(ncol, n_iw, 16, &chunk3);
  c2 *part = work, *q = work + (long)33 * sk * ncol, *qbar = q + 29 * ncol, *kap = qbar + 29 * ncol;
  const unsigned bc = (unsigned)((ncol + 255) / 256);
  hipLaunchKernelGGL(k_static_proj_part, dim3(bc, sk), dim3(256), 0, s, ncol, n_iw, chunk, g, tab, part);
  hipLaunchKernelGGL(k_static_proj_final, dim3(bc), dim3(256), 0, s, ncol, sk, part, g, fit, q, status);
  switch (n) {
#define TPRF_SM_ADJ(N_)                                                                                                case N_:                                                                                                               hipLaunchKernelGGL(k_static_qbar<N_>, dim3(bc), dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2], q, chibar,                        sm, qbar, kap);                                                                                   hipLaunchKernelGGL(k_static_expand<N_>, dim3(bc, sk3), dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2],                          n_iw, chunk3, g, tab, qbar, kap, sm.inv_beta, gbar);                                              break;
    TPRF_SM_ADJ(1)
    TPRF_SM_ADJ(2)
    TPRF_SM_ADJ(3)
#undef TPRF_SM_ADJ
    default: return hipErrorInvalidValue;
  }
  hipLaunchKernelGGL(k_static_fit_adjoint, dim3(bc), dim3(256), 0, s, ncol, fit, qbar, gbar);
  return hipGetLastError();
}

}  // namespace tprf
