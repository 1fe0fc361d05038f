// kernels.hip -- CDNA4 (gfx950) kernels of the lattice bubble + RPA chain.
//
// Every kernel computes in complex FP64.  Data are in the reference's
// row-major gf layout [mesh0][mesh1][target], so each kernel's global
// accesses are contiguous along the innermost (orbital / r) index.
#include "common.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

// ------------------------------------------------------------------------
// a1 Dyson: G[w,k] = inverse((i w_n + mu) 1 - e_k - Sigma), one thread per
// (w, k); Gauss-Jordan with partial pivoting on [M | 1] in registers.
// zarr: the frequencies z[iw] of any mesh (real frequencies with their i delta,
// or given Matsubara points) in place of i w_n: M = (z + mu) 1 - e_k - Sigma.
template <int N>
__global__ void __launch_bounds__(256) k_dyson(double beta, int n_iw, double mu, const double *__restrict__ mu_arr,
                                               long nk_inst, long nk, const c2 *__restrict__ ek, int sigma_kind,
                                               const c2 *__restrict__ sigma, c2 *__restrict__ g, long npts,
                                               double *status, const c2 *__restrict__ zarr) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= npts) return;
  const long iw = idx / nk, k = idx - iw * nk;
  if (mu_arr) mu = mu_arr[k / nk_inst];  // a batch of lattices folded into k
  double zr = 0.0, wn;
  if (zarr) {
    zr = zarr[iw].x;
    wn = zarr[iw].y;
  } else {
    wn = kPi * (double)(2 * (iw - n_iw) + 1) / beta;
  }
  c2 A[N][N], B[N][N];
  const c2 *e = ek + k * N * N;
  // all n^2 entries of e_k (and of Sigma) are loaded before the first is used
  v2d ev[N * N], sv[N * N];
#pragma unroll
  for (int t = 0; t < N * N; ++t) ev[t] = ld_v2d(e + t);
  if (sigma_kind != 0) {
    const c2 *sg = sigma + (sigma_kind == 1 ? iw : idx) * N * N;
#pragma unroll
    for (int t = 0; t < N * N; ++t) sv[t] = ld_v2d(sg + t);
  }
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      c2 v = c2{-ev[i * N + j].x, -ev[i * N + j].y};
      if (i == j) { v.x += mu + zr; v.y += wn; }
      if (sigma_kind != 0) v = csub(v, c2{sv[i * N + j].x, sv[i * N + j].y});
      A[i][j] = v;
      B[i][j] = (i == j) ? c2{1.0, 0.0} : c2{0.0, 0.0};
    }
  bool singular = false;
#pragma unroll
  for (int kk = 0; kk < N; ++kk) {
    int p = kk;
    double best = cabs2(A[kk][kk]);
#pragma unroll
    for (int i = kk + 1; i < N; ++i) {
      double v = cabs2(A[i][kk]);
      if (v > best) { best = v; p = i; }
    }
    if (best == 0.0) singular = true;
#pragma unroll
    for (int i = kk + 1; i < N; ++i) {
      if (p == i) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          c2 t = A[kk][j]; A[kk][j] = A[i][j]; A[i][j] = t;
          t = B[kk][j]; B[kk][j] = B[i][j]; B[i][j] = t;
        }
      }
    }
    const c2 inv = cdiv(c2{1.0, 0.0}, A[kk][kk]);
#pragma unroll
    for (int j = 0; j < N; ++j) {
      A[kk][j] = cmul(A[kk][j], inv);
      B[kk][j] = cmul(B[kk][j], inv);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
      if (i == kk) continue;
      const c2 f = A[i][kk];
#pragma unroll
      for (int j = 0; j < N; ++j) {
        A[i][j] = cfms(A[i][j], f, A[kk][j]);
        B[i][j] = cfms(B[i][j], f, B[kk][j]);
      }
    }
  }
  if (singular) status[2] = 1.0;
  c2 *o = g + idx * N * N;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) o[i * N + j] = B[i][j];
}

// n = 7, 8: N lanes own a point (P = 64 / N points per wave), lane i holding
// row i of [M | 1]; Gauss-Jordan with partial pivoting, the pivot row read from
// its lane by shuffles, rows left in place (the row that pivoted column k
// holds, after the sweep, pivot times row k of the inverse).
template <int N>
__global__ void __launch_bounds__(256) k_dyson_rows(double beta, int n_iw, double mu, const double *__restrict__ mu_arr,
                                                    long nk_inst, long nk, const c2 *__restrict__ ek, int sigma_kind,
                                                    const c2 *__restrict__ sigma, c2 *__restrict__ g, long npts,
                                                    double *status, const c2 *__restrict__ zarr) {
  constexpr int P = 64 / N;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gidx = lane / N, i = lane - gidx * N;
  const long idx = ((long)blockIdx.x * 4 + wave) * P + gidx;
  const bool valid = gidx < P && idx < npts;
  c2 A[N], B[N];
  if (valid) {
    const long iw = idx / nk, k = idx - iw * nk;
    const double m = (mu_arr ? mu_arr[k / nk_inst] : mu) + (zarr ? zarr[iw].x : 0.0);
    const double wn = zarr ? zarr[iw].y : kPi * (double)(2 * (iw - n_iw) + 1) / beta;
    const c2 *e = ek + k * N * N + i * N;
    const c2 *sg = sigma_kind == 0 ? nullptr : sigma + (sigma_kind == 1 ? iw : idx) * N * N + i * N;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      c2 v = c2{-e[j].x, -e[j].y};
      if (j == i) { v.x += m; v.y += wn; }
      if (sg) v = csub(v, sg[j]);
      A[j] = v;
      B[j] = (j == i) ? c2{1.0, 0.0} : c2{0.0, 0.0};
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) { A[j] = c2{0.0, 0.0}; B[j] = c2{0.0, 0.0}; }
  }
  const int lane0 = lane - i;  // first lane of this point's group
  bool used = false, singular = false;
  int myk = 0;
  c2 pv = c2{1.0, 0.0};
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double key = (!used && valid) ? cabs2(A[k]) : -1.0;
    double best = -2.0;
    int best_i = 0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
      const double o = __shfl(key, lane0 + r, 64);
      if (o > best) { best = o; best_i = r; }
    }
    if (valid && !(best > 0.0)) singular = true;
    const int src = lane0 + best_i;
    c2 pr[N], pb[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      pr[j] = c2{__shfl(A[j].x, src, 64), __shfl(A[j].y, src, 64)};
      pb[j] = c2{__shfl(B[j].x, src, 64), __shfl(B[j].y, src, 64)};
    }
    if (i == best_i) {
      used = true;
      myk = k;
      pv = A[k];
    } else {
      const c2 f = cdiv(A[k], pr[k]);
#pragma unroll
      for (int j = 0; j < N; ++j) {
        if (j > k) A[j] = cfms(A[j], f, pr[j]);
        B[j] = cfms(B[j], f, pb[j]);
      }
      A[k] = c2{0.0, 0.0};
    }
  }
  if (singular) status[2] = 1.0;
  if (valid) {
    const c2 inv = cdiv(c2{1.0, 0.0}, pv);
    c2 *o = g + idx * N * N + myk * N;
#pragma unroll
    for (int j = 0; j < N; ++j) o[j] = cmul(B[j], inv);
  }
}

// out[z][ab] = (1/nk) sum_k g[z][k][ab], the k sum in index order (the order of
// lattice_dyson_g_X's accumulation, gf.cpp:153-190), thread per (z, ab).
__global__ void __launch_bounds__(256) k_k_average(long nz, long nk, int n2, const c2 *__restrict__ g,
                                                   c2 *__restrict__ out) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nz * n2) return;
  const long z = t / n2;
  const int ab = (int)(t - z * n2);
  const c2 *p = g + z * nk * n2 + ab;
  c2 acc = c2{0.0, 0.0};
  for (long k = 0; k < nk; ++k) acc = cadd(acc, p[k * n2]);
  out[t] = cscale(acc, 1.0 / (double)nk);
}

hipError_t launch_k_average(hipStream_t s, long nz, long nk, int n, const c2 *g, c2 *out) {
  const long nt = nz * n * n;
  hipLaunchKernelGGL(k_k_average, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, nz, nk, n * n, g, out);
  return hipGetLastError();
}

hipError_t launch_dyson(hipStream_t s, double beta, int n_iw, double mu, long nk, int n, const c2 *ek, int sigma_kind,
                        const c2 *sigma, c2 *g, Status st, const double *mu_arr, long nk_inst, const c2 *zarr,
                        long nz) {
  const long npts = (zarr ? nz : 2L * n_iw) * nk;
  const int bs = 256;
  const long nb = (npts + bs - 1) / bs;
  switch (n) {
#define TPRF_DYSON_CASE(N)                                                                               \
  case N:                                                                                                \
    hipLaunchKernelGGL(k_dyson<N>, dim3(nb), dim3(bs), 0, s, beta, n_iw, mu, mu_arr, nk_inst, nk, ek, sigma_kind, \
                       sigma, g, npts, st.d, zarr);                                                      \
    break;
    TPRF_DYSON_CASE(1)
    TPRF_DYSON_CASE(2)
    TPRF_DYSON_CASE(3)
    TPRF_DYSON_CASE(4)
    TPRF_DYSON_CASE(5)
    TPRF_DYSON_CASE(6)
#undef TPRF_DYSON_CASE
#define TPRF_DYSON_ROWS(N)                                                                                   \
  case N: {                                                                                                  \
    const long nbr = (npts + 4 * (64 / N) - 1) / (4 * (64 / N));                                             \
    hipLaunchKernelGGL(k_dyson_rows<N>, dim3(nbr), dim3(256), 0, s, beta, n_iw, mu, mu_arr, nk_inst, nk, ek,  \
                       sigma_kind, sigma, g, npts, st.d, zarr);                                              \
    break;                                                                                                   \
  }
    TPRF_DYSON_ROWS(7)
    TPRF_DYSON_ROWS(8)
#undef TPRF_DYSON_ROWS
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// a2 / a6: one axis of the batched lattice FFT.  The tensor is viewed as
// [outer][len][inner]; a workgroup loads C consecutive inner columns of one
// outer index (contiguous C*16 bytes per row), transforms them in LDS and
// writes them back scaled.
template <int S>
__global__ void __launch_bounds__(256) k_lattice_axis(const c2 *__restrict__ in, c2 *__restrict__ out, long outer,
                                                      int len, long inner, int C, FftDev p, double scale) {
  extern __shared__ c2 sm[];
  const int ld = len | 1;
  c2 *a = sm, *b = sm + C * ld;
  const long ntile_in = (inner + C - 1) / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long o = tile / ntile_in;
  const long c0 = (tile - o * ntile_in) * C;
  const int cn = (int)((inner - c0) < C ? (inner - c0) : C);
  const c2 *src = in + o * len * inner + c0;
  for (int t = threadIdx.x; t < len * cn; t += blockDim.x) {
    const int j = t / cn, c = t - j * cn;
    a[c * ld + j] = src[(long)j * inner + c];
  }
  __syncthreads();
  c2 *r = fft_lds<S>(a, b, ld, cn, p, threadIdx.x, blockDim.x);
  c2 *dst = out + o * len * inner + c0;
  for (int t = threadIdx.x; t < len * cn; t += blockDim.x) {
    const int j = t / cn, c = t - j * cn;
    dst[(long)j * inner + c] = cscale(r[c * ld + j], scale);
  }
}

hipError_t launch_lattice_axis(hipStream_t s, int sign, const c2 *in, c2 *out, long outer, int len, long inner,
                               const FftDev &p, double scale) {
  const int ld = len | 1;
  long C = 2048 / ld;
  if (C < 1) C = 1;
  if (C > inner) C = inner;
  const long ntile = outer * ((inner + C - 1) / C);
  const size_t lds = 2 * C * ld * sizeof(c2);
  if (sign < 0)
    hipLaunchKernelGGL(k_lattice_axis<-1>, dim3(ntile), dim3(256), lds, s, in, out, outer, len, inner, (int)C, p,
                       scale);
  else
    hipLaunchKernelGGL(k_lattice_axis<1>, dim3(ntile), dim3(256), lds, s, in, out, outer, len, inner, (int)C, p,
                       scale);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// Matsubara transforms (pole model and helpers in mats_device.hpp).

int mats_cols_per_block(int L) {
  // two LDS buffers of C columns of (L+1) rounded to odd
  const int ld = (L + 1) | 1;
  int C = (int)(kMatsLdsBytes / (2 * ld * (int)sizeof(c2)));
  if (C > 16) C = 16;
  return C < 1 ? 1 : C;
}

// set once the device has granted kMatsBigLdsBytes to the single-column kernels
static bool g_mats_big_ok = false;

bool mats_single_column(int L) { return 2L * ((L + 1) | 1) * (long)sizeof(c2) > kMatsLdsBytes; }

bool mats_supported(int L) {
  return !mats_single_column(L) || (g_mats_big_ok && ((L + 1) | 1) * (long)sizeof(c2) <= kMatsBigLdsBytes);
}

// a3: Matsubara -> imaginary time for C columns (fourier_matsubara.cpp:218-308).
// BIG: one column per workgroup of kBigThreads, transformed in place.
template <int STAT, bool BIG>
__global__ void __launch_bounds__(BIG ? kBigThreads : 256)
    k_w2t(double beta, int n_max, long ncol, int C, const c2 *__restrict__ in, c2 *__restrict__ out,
          const c2 *__restrict__ moments, TailFitDev fit, MatsDev mt, FftDev p, double *status) {
  extern __shared__ c2 sm[];
  const int L = mt.L;
  const int ld = (L + 1) | 1;
  c2 *a = sm, *b = BIG ? nullptr : sm + C * ld;
  __shared__ c2 s_coef[16][kMaxOrder];
  __shared__ c2 s_a[16][3];
  __shared__ c2 s_m1[16];
  __shared__ double s_res[16];
  const long ntile = (ncol + C - 1) / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long c0 = tile * C;
  const int cn = (int)((ncol - c0) < C ? (ncol - c0) : C);
  (void)ntile;
  const int nmesh = (STAT == 0) ? 2 * n_max : 2 * n_max - 1;
  const int first = (STAT == 0) ? -n_max : -(n_max - 1);

  // 1. tail moments
  if (moments == nullptr) {
    for (int t = threadIdx.x; t < cn * fit.ncoef; t += blockDim.x) {
      const int c = t / fit.ncoef, q = t - c * fit.ncoef;
      c2 s = c2{0.0, 0.0};
      for (int f = 0; f < fit.n_fit; ++f) s = cfma(s, fit.pinv[q * fit.n_fit + f], in[(long)fit.idx[f] * ncol + c0 + c]);
      s_coef[c][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < cn) {
      const int c = threadIdx.x;
      double r2 = 0.0;
      for (int f = 0; f < fit.n_fit; ++f) {
        c2 v = in[(long)fit.idx[f] * ncol + c0 + c];
        for (int q = 0; q < fit.ncoef; ++q) v = cfms(v, fit.vander[f * fit.ncoef + q], s_coef[c][q]);
        r2 += cabs2(v);
      }
      s_res[c] = sqrt(r2 / fit.n_fit);
      c2 m1 = cscale(s_coef[c][1], fit.wpow[1]);
      c2 m2 = cscale(s_coef[c][2], fit.wpow[2]);
      c2 m3 = cscale(s_coef[c][3], fit.wpow[3]);
      Poles<STAT>::coef(m1, m2, m3, s_a[c]);
      s_m1[c] = m1;
      atomic_max_nonneg(&status[0], s_res[c]);
      atomic_max_nonneg(&status[1], sqrt(cabs2(s_coef[c][0])));
    }
  } else if (threadIdx.x < cn) {
    const int c = threadIdx.x;
    c2 m1 = moments[c0 + c], m2 = moments[ncol + c0 + c], m3 = moments[2 * ncol + c0 + c];
    Poles<STAT>::coef(m1, m2, m3, s_a[c]);
    s_m1[c] = m1;
  }
  // 2. zero the FFT input, then scatter (1/beta)(G - model) to (n + L) % L
  for (int t = threadIdx.x; t < cn * ld; t += blockDim.x) a[t] = c2{0.0, 0.0};
  __syncthreads();
  const double ib = 1.0 / beta;
  for (int t = threadIdx.x; t < nmesh * cn; t += blockDim.x) {
    const int i = t / cn, c = t - i * cn;
    const long n = first + i;
    const double w = matsubara_w<STAT>(n, beta);
    c2 v = csub(in[(long)i * ncol + c0 + c], pole_sum<STAT>(s_a[c], w));
    a[c * ld + (int)((n + L) % L)] = cscale(v, ib);
  }
  __syncthreads();
  // 3. forward FFT of length L
  c2 *y;
  if constexpr (BIG)
    y = fft_lds_inplace<-1, kBigThreads, kBigMaxN>(a, p, threadIdx.x);
  else
    y = fft_lds<-1>(a, b, ld, cn, p, threadIdx.x, blockDim.x);
  // 4. G(tau_j) = y_j [exp(-i pi j/L)] + sum a_i psi_i(tau_j); G(tau_L) = -/+ (G(tau_0) + m1)
  const double pm = (STAT == 0) ? -1.0 : 1.0;
  for (int t = threadIdx.x; t < (L + 1) * cn; t += blockDim.x) {
    const int j = t / cn, c = t - j * cn;
    const int jj = (j == L) ? 0 : j;
    c2 v = y[c * ld + jj];
    if (STAT == 0) v = cmul(v, mt.phase[jj]);
    const c2 *ac = s_a[c];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double ps = mt.psi[q * L + jj];
      v.x += ac[q].x * ps;
      v.y += ac[q].y * ps;
    }
    if (j == L) v = cscale(cadd(v, s_m1[c]), pm);
    out[(long)j * ncol + c0 + c] = v;
  }
}

hipError_t launch_w2t(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                      const c2 *moments, const TailFitDev &fit, const MatsDev &mt, const FftDev &p, Status st) {
  const int ld = (mt.L + 1) | 1;
  if (mats_single_column(mt.L)) {
    const size_t lds = (size_t)ld * sizeof(c2);
    if (statistic == 0)
      hipLaunchKernelGGL((k_w2t<0, true>), dim3(ncol), dim3(kBigThreads), lds, s, beta, n_max, ncol, 1, in, out,
                         moments, fit, mt, p, st.d);
    else
      hipLaunchKernelGGL((k_w2t<1, true>), dim3(ncol), dim3(kBigThreads), lds, s, beta, n_max, ncol, 1, in, out,
                         moments, fit, mt, p, st.d);
    return hipGetLastError();
  }
  const int C = mats_cols_per_block(mt.L);
  const long nb = (ncol + C - 1) / C;
  const size_t lds = 2 * (size_t)C * ld * sizeof(c2);
  if (statistic == 0)
    hipLaunchKernelGGL((k_w2t<0, false>), dim3(nb), dim3(256), lds, s, beta, n_max, ncol, C, in, out, moments, fit,
                       mt, p, st.d);
  else
    hipLaunchKernelGGL((k_w2t<1, false>), dim3(nb), dim3(256), lds, s, beta, n_max, ncol, C, in, out, moments, fit,
                       mt, p, st.d);
  return hipGetLastError();
}

// a5 core, shared by k_t2w and the fused product kernel: given the L+1 tau
// samples of each column in a[c*ld + j], compute the boundary moments
// (fit_derivatives, fourier_matsubara.cpp:42-97), subtract the pole model,
// transform and write the 2nw-1 (bosons) or 2nw (fermions) frequencies.
template <int STAT, bool BIG>
__device__ void t2w_core(c2 *a, c2 *b, int ld, int cn, double beta, int nw, const MatsDev &mt, const FftDev &p,
                         const Vinv8 &vinv, c2 (*s_a)[3], c2 *dst, long dst_row, long dst_col_stride) {
  const int L = mt.L;
  const double dtau = beta / (double)L;
  if (threadIdx.x < cn) {
    const int c = threadIdx.x;
    const c2 *g = a + c * ld;
    boundary_poles<STAT>([&](int m) { return g[m]; }, [&](int m) { return g[L - m]; }, dtau, vinv, s_a[c],
                         (c2 *)nullptr);
  }
  __syncthreads();
  const double fact = beta / (double)L;
  for (int t = threadIdx.x; t < cn * L; t += blockDim.x) {
    const int c = t / L, j = t - c * L;
    const c2 *ac = s_a[c];
    c2 model = c2{0, 0};
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double ps = mt.psi[q * L + j];
      model.x += ac[q].x * ps;
      model.y += ac[q].y * ps;
    }
    c2 v = csub(a[c * ld + j], model);
    if (STAT == 0) v = cmul(cconj(mt.phase[j]), v);  // exp(+i pi tau_j / beta)
    a[c * ld + j] = cscale(v, fact);
  }
  __syncthreads();
  c2 *y;
  if constexpr (BIG)
    y = fft_lds_inplace<1, kBigThreads, kBigMaxN>(a, p, threadIdx.x);
  else
    y = fft_lds<1>(a, b, ld, cn, p, threadIdx.x, blockDim.x);
  const int nout = (STAT == 0) ? 2 * nw : 2 * nw - 1;
  const int first = (STAT == 0) ? -nw : -(nw - 1);
  for (int t = threadIdx.x; t < nout * cn; t += blockDim.x) {
    const int i = t / cn, c = t - i * cn;
    const long n = first + i;
    const double w = matsubara_w<STAT>(n, beta);
    c2 v = cadd(y[c * ld + (int)((n + L) % L)], pole_sum<STAT>(s_a[c], w));
    dst[(long)i * dst_row + c * dst_col_stride] = v;
  }
}

template <int STAT, bool BIG>
__global__ void __launch_bounds__(BIG ? kBigThreads : 256)
    k_t2w(double beta, int nw, long ncol, int C, const c2 *__restrict__ in, c2 *__restrict__ out, MatsDev mt,
          FftDev p, Vinv8 vinv) {
  extern __shared__ c2 sm[];
  const int L = mt.L;
  const int ld = (L + 1) | 1;
  c2 *a = sm, *b = BIG ? nullptr : sm + C * ld;
  __shared__ c2 s_a[16][3];
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long c0 = tile * C;
  const int cn = (int)((ncol - c0) < C ? (ncol - c0) : C);
  for (int t = threadIdx.x; t < (L + 1) * cn; t += blockDim.x) {
    const int j = t / cn, c = t - j * cn;
    a[c * ld + j] = in[(long)j * ncol + c0 + c];
  }
  __syncthreads();
  t2w_core<STAT, BIG>(a, b, ld, cn, beta, nw, mt, p, vinv, s_a, out + c0, ncol, 1);
}

hipError_t launch_t2w(hipStream_t s, int statistic, double beta, int nw, long ncol, const c2 *in, c2 *out,
                      const MatsDev &mt, const FftDev &p, const Vinv8 &vinv) {
  const int ld = (mt.L + 1) | 1;
  if (mats_single_column(mt.L)) {
    const size_t lds = (size_t)ld * sizeof(c2);
    if (statistic == 0)
      hipLaunchKernelGGL((k_t2w<0, true>), dim3(ncol), dim3(kBigThreads), lds, s, beta, nw, ncol, 1, in, out, mt, p,
                         vinv);
    else
      hipLaunchKernelGGL((k_t2w<1, true>), dim3(ncol), dim3(kBigThreads), lds, s, beta, nw, ncol, 1, in, out, mt, p,
                         vinv);
    return hipGetLastError();
  }
  const int C = mats_cols_per_block(mt.L);
  const long nb = (ncol + C - 1) / C;
  const size_t lds = 2 * (size_t)C * ld * sizeof(c2);
  if (statistic == 0)
    hipLaunchKernelGGL((k_t2w<0, false>), dim3(nb), dim3(256), lds, s, beta, nw, ncol, C, in, out, mt, p, vinv);
  else
    hipLaunchKernelGGL((k_t2w<1, false>), dim3(nb), dim3(256), lds, s, beta, nw, ncol, C, in, out, mt, p, vinv);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// a4 / a4': bubble product chi0_abcd(tau_j, r) = G_da(tau_j, r) G_bc(tau_{L-j}, -r)
// (chi_imtime.cpp:127).

// fused product + bosonic tau->nu: one workgroup per (r, C consecutive abcd)
// herm: G_ab(tau, s)^* = G_ba(tau, -s) and only the columns of s <= -s are
// read (the others as conj(G_ba(tau, -s))).
template <bool BIG>
__global__ void __launch_bounds__(BIG ? kBigThreads : 256)
    k_chi0_wr(double beta, int nw, int d0, int d1, int d2, int n, int C, const c2 *__restrict__ gtr,
              c2 *__restrict__ out, MatsDev mt, FftDev p, Vinv8 vinv, int herm) {
  extern __shared__ c2 sm[];
  const int L = mt.L;
  const int ld = (L + 1) | 1;
  c2 *a = sm, *b = BIG ? nullptr : sm + C * ld;
  __shared__ c2 s_a[16][3];
  const long nr = (long)d0 * d1 * d2;
  const int n2 = n * n, n4 = n2 * n2;
  const int tiles_per_r = (n4 + C - 1) / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long r = tile / tiles_per_r;
  const int q0 = (int)(tile - r * tiles_per_r) * C;
  const int cn = (n4 - q0) < C ? (n4 - q0) : C;
  const long mr = minus_r(r, d0, d1, d2);
  for (int t = threadIdx.x; t < (L + 1) * cn; t += blockDim.x) {
    const int j = t / cn, c = t - j * cn;
    const int q = q0 + c;
    const int dd = q % n, cc = (q / n) % n, bb = (q / n2) % n, aa = q / (n2 * n);
    // G_xy(tau_i, s), read from -s (transposed, conjugated) when herm and s > -s
    auto g = [&](long i, long sv, long msv, int x, int y) {
      if (herm && sv > msv) return cconj(gtr[(i * nr + msv) * n2 + y * n + x]);
      return gtr[(i * nr + sv) * n2 + x * n + y];
    };
    const c2 gp = g(j, r, mr, dd, aa);
    const c2 gm = g(L - j, mr, r, bb, cc);
    a[c * ld + j] = cmul(gp, gm);
  }
  __syncthreads();
  t2w_core<1, BIG>(a, b, ld, cn, beta, nw, mt, p, vinv, s_a, out + r * n4 + q0, nr * n4, 1);
}

hipError_t launch_chi0_wr(hipStream_t s, double beta, int nw, const int dims[3], int n, const c2 *gtr, c2 *out,
                          const MatsDev &mt, const FftDev &p, const Vinv8 &vinv, bool herm) {
  const int n4 = n * n * n * n;
  const int ld = (mt.L + 1) | 1;
  const long nr = (long)dims[0] * dims[1] * dims[2];
  if (mats_single_column(mt.L)) {
    hipLaunchKernelGGL(k_chi0_wr<true>, dim3(nr * n4), dim3(kBigThreads), (size_t)ld * sizeof(c2), s, beta, nw,
                       dims[0], dims[1], dims[2], n, 1, gtr, out, mt, p, vinv, herm ? 1 : 0);
    return hipGetLastError();
  }
  int C = mats_cols_per_block(mt.L);
  if (C > n4) C = n4;
  const long nb = nr * ((n4 + C - 1) / C);
  const size_t lds = 2 * (size_t)C * ld * sizeof(c2);
  hipLaunchKernelGGL(k_chi0_wr<false>, dim3(nb), dim3(256), lds, s, beta, nw, dims[0], dims[1], dims[2], n, C, gtr,
                     out, mt, p, vinv, herm ? 1 : 0);
  return hipGetLastError();
}

// materialised product, one thread per output element
__global__ void __launch_bounds__(256) k_chi0_tr(int ntau, int d0, int d1, int d2, int n, const c2 *__restrict__ gtr,
                                                 c2 *__restrict__ out, long total) {
  const long nr = (long)d0 * d1 * d2;
  const int n2 = n * n, n4 = n2 * n2;
  const int L = ntau - 1;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int q = (int)(idx % n4);
    const long jr = idx / n4;
    const long r = jr % nr, j = jr / nr;
    const long mr = minus_r(r, d0, d1, d2);
    const int dd = q % n, cc = (q / n) % n, bb = (q / n2) % n, aa = q / (n2 * n);
    const c2 gp = gtr[(j * nr + r) * n2 + dd * n + aa];
    const c2 gm = gtr[((L - j) * nr + mr) * n2 + bb * n + cc];
    out[idx] = cmul(gp, gm);
  }
}

hipError_t launch_chi0_tr(hipStream_t s, int ntau, const int dims[3], int n, const c2 *gtr, c2 *out) {
  const long nr = (long)dims[0] * dims[1] * dims[2];
  const long total = (long)ntau * nr * n * n * n * n;
  long nb = (total + 255) / 256;
  if (nb > 65536L * 8) nb = 65536L * 8;
  hipLaunchKernelGGL(k_chi0_tr, dim3(nb), dim3(256), 0, s, ntau, dims[0], dims[1], dims[2], n, gtr, out, total);
  return hipGetLastError();
}

// chi_trapz_tau (chi_imtime.cpp:230-247) of the product, one thread per (r, abcd)
__global__ void __launch_bounds__(256) k_chi0_w0r(double beta, int ntau, int d0, int d1, int d2, int n,
                                                  const c2 *__restrict__ gtr, c2 *__restrict__ out) {
  const long nr = (long)d0 * d1 * d2;
  const int n2 = n * n, n4 = n2 * n2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nr * n4) return;
  const int q = (int)(idx % n4);
  const long r = idx / n4;
  const long mr = minus_r(r, d0, d1, d2);
  const int dd = q % n, cc = (q / n) % n, bb = (q / n2) % n, aa = q / (n2 * n);
  const int L = ntau - 1;
  c2 s = c2{0, 0}, first = c2{0, 0}, last = c2{0, 0};
  for (int j = 0; j < ntau; ++j) {
    const c2 v = cmul(gtr[((long)j * nr + r) * n2 + dd * n + aa], gtr[((long)(L - j) * nr + mr) * n2 + bb * n + cc]);
    s = cadd(s, v);
    if (j == 0) first = v;
    if (j == L) last = v;
  }
  s = csub(s, cscale(first, 0.5));
  s = csub(s, cscale(last, 0.5));
  out[idx] = cscale(s, beta / (double)L);
}

hipError_t launch_chi0_w0r(hipStream_t s, double beta, int ntau, const int dims[3], int n, const c2 *gtr, c2 *out) {
  const long nr = (long)dims[0] * dims[1] * dims[2];
  const long total = nr * n * n * n * n;
  hipLaunchKernelGGL(k_chi0_w0r, dim3((total + 255) / 256), dim3(256), 0, s, beta, ntau, dims[0], dims[1], dims[2], n,
                     gtr, out);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_chi_w0r(double beta, int ntau, long ncol, const c2 *__restrict__ chi_tr,
                                                 c2 *__restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncol) return;
  c2 s = c2{0, 0};
  for (int j = 0; j < ntau; ++j) s = cadd(s, chi_tr[(long)j * ncol + idx]);
  s = csub(s, cscale(chi_tr[idx], 0.5));
  s = csub(s, cscale(chi_tr[(long)(ntau - 1) * ncol + idx], 0.5));
  out[idx] = cscale(s, beta / (double)(ntau - 1));
}

hipError_t launch_chi_w0r(hipStream_t s, double beta, int ntau, long nr, int n, const c2 *chi_tr, c2 *out) {
  const long ncol = nr * n * n * n * n;
  hipLaunchKernelGGL(k_chi_w0r, dim3((ncol + 255) / 256), dim3(256), 0, s, beta, ntau, ncol, chi_tr, out);
  return hipGetLastError();
}

// Order the LDS traffic of one wave: writes issued before are complete and
// visible to every lane of the wave after it.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------------
// a7 RPA: for each point, M_chi = inverse(1 - M_chi0 M_U) M_chi0 (rpa.cpp:57)
// with M_X[(a,b),(c,d)] = X_abcd (the memory-order grouping that
// group_indices_view makes of the row-major tensors, rpa.cpp:40,54-55),
// solved as (1 - X U) Y = X by Gauss-Jordan with partial pivoting.  A group
// of N2 = n^2 lanes owns one point, lane i holding row i of [A | X]; the pivot
// row is exchanged through LDS.
// n = 3: registers held to four waves per SIMD.
// MIRROR: the points are the nb * nk points of the nonnegative bosonic
// frequencies m = pt / nk >= 0 of arrays [2 nb - 1][nk][n^4] (read and written
// at base + pt, base = (nb - 1) nk), and each result is also written at -m
// by chi_dcba(-nu, q) = conj(chi_abcd(nu, q)), the identity that holds when
// chi0 obeys it and U_abcd = conj(U_dcba).
template <int N, bool MIRROR = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N == 3 ? 4 : 1))) k_rpa(long npts, const c2 *__restrict__ chi0, const c2 *__restrict__ U,
                                             c2 *__restrict__ chi, double *status, long base = 0, long nk = 1) {
  constexpr int N2 = N * N;
  constexpr int P = 64 / N2;  // points per wave
  constexpr int NMW = (N2 * N2 + 63) / 64;  // words of the nonzero mask of M_U
  __shared__ c2 slot[4 * P][2 * N2];
  __shared__ unsigned long long sMask[NMW];
  // M_U[l][j] with l = (a, b), j = (c, d) is U[((a N + b) N + c) N + d]
  auto u_idx = [](int l, int j) { return l * N2 + j; };
  // Bit l N2 + j of the mask: M_U[l][j] != 0.  Interactions of physical models
  // are sparse in the particle-hole grouping (Kanamori: 21 of 81 entries for
  // three orbitals); zero entries add exact zeros to 1 - X U and are skipped
  // with uniform branches.
  if (threadIdx.x < 64 * NMW) {
    const int t = threadIdx.x;
    bool nz = false;
    if (t < N2 * N2) {
      const c2 u = U[u_idx(t / N2, t % N2)];
      nz = u.x != 0.0 || u.y != 0.0;
    }
    const unsigned long long m = __ballot(nz);
    if ((t & 63) == 0) sMask[t >> 6] = m;
  }
  __syncthreads();
  unsigned long long umask[NMW];
#pragma unroll
  for (int w = 0; w < NMW; ++w) {
    const unsigned long long m = sMask[w];
    umask[w] = ((unsigned long long)__builtin_amdgcn_readfirstlane((unsigned)(m >> 32)) << 32) |
               (unsigned)__builtin_amdgcn_readfirstlane((unsigned)m);
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gidx = lane / N2, i = lane - gidx * N2;
  const int slot_id = wave * P + gidx;
  const long pt = ((long)blockIdx.x * 4 + wave) * P + gidx;
  const bool valid = (gidx < P) && (pt < npts);
  c2 X[N2], A[N2];
  if (valid) {
    const c2 *row = chi0 + (base + pt) * N2 * N2 + i * N2;
#pragma unroll
    for (int c = 0; c < N; ++c)
#pragma unroll
      for (int d = 0; d < N; ++d) X[c * N + d] = row[c * N + d];
  } else {
#pragma unroll
    for (int j = 0; j < N2; ++j) X[j] = c2{0, 0};
  }
  // A = 1 - X M_U, column by column: the n^2 entries of column j of M_U sit
  // at wave-uniform addresses (scalar loads, all issued before the products)
#pragma unroll
  for (int j = 0; j < N2; ++j) {
    c2 uj[N2];
#pragma unroll
    for (int l = 0; l < N2; ++l) uj[l] = U[u_idx(l, j)];
    c2 s = (i == j) ? c2{1.0, 0.0} : c2{0.0, 0.0};
#pragma unroll
    for (int l = 0; l < N2; ++l) {
      const int bit = l * N2 + j;
      if ((umask[bit >> 6] >> (bit & 63)) & 1ull) s = cfms(s, X[l], uj[l]);
    }
    A[j] = s;
  }
  bool used = false;
  int myk = 0;
  c2 pv = c2{1.0, 0.0};
  bool singular = false;
#pragma unroll
  for (int k = 0; k < N2; ++k) {
    // Pivot key: the upper 32 bits of |A[k]|^2 (sign, exponent and 20 mantissa
    // bits of a non-negative double order like the value), one 32-bit shuffle
    // per candidate; -1 for rows that cannot pivot.  Magnitudes within ~1e-6
    // of each other may pick either row, which only changes rounding.
    const int key = (!used && valid) ? (int)(__double_as_longlong(cabs2(A[k])) >> 32) : -1;
    int best_i = 0, best = -2;
#pragma unroll
    for (int s = 0; s < N2; ++s) {
      const int m = __shfl(key, (lane - i) + s, 64);
      if (m > best) { best = m; best_i = s; }
    }
    const bool is_piv = (i == best_i);
    if (valid && best <= 0) singular = true;  // |A[k]|^2 below ~1e-314 in every candidate row
    if (is_piv && gidx < P) {
      used = true;
      myk = k;
      pv = A[k];
#pragma unroll
      for (int j = k; j < N2; ++j) slot[slot_id][j] = A[j];  // columns < k are zero in every unused row
#pragma unroll
      for (int j = 0; j < N2; ++j) slot[slot_id][N2 + j] = X[j];
    }
    wave_lds_sync();  // slots are private to a wave
    if (!is_piv && gidx < P) {
      const c2 f = cdiv(A[k], slot[slot_id][k]);
#pragma unroll
      for (int j = k + 1; j < N2; ++j) A[j] = cfms(A[j], f, slot[slot_id][j]);
      A[k] = c2{0.0, 0.0};
#pragma unroll
      for (int j = 0; j < N2; ++j) X[j] = cfms(X[j], f, slot[slot_id][N2 + j]);
    }
    wave_lds_sync();
  }
  if (singular) status[2] = 1.0;
  if (valid) {
    const c2 inv = cdiv(c2{1.0, 0.0}, pv);
    c2 *row = chi + (base + pt) * N2 * N2 + myk * N2;
#pragma unroll
    for (int c = 0; c < N; ++c)
#pragma unroll
      for (int d = 0; d < N; ++d) row[c * N + d] = cmul(X[c * N + d], inv);
    if constexpr (MIRROR) {
      // this lane's row (a, b) of M_chi(nu) is column (a, b) of M_chi(-nu):
      // chi_dcba(-nu) = conj(chi_abcd(nu)) for every (c, d)
      const long m = pt / nk;
      if (m > 0) {
        const int a = myk / N, b = myk - (myk / N) * N;
        c2 *mp = chi + (base + pt - 2 * m * nk) * N2 * N2;
#pragma unroll
        for (int c = 0; c < N; ++c)
#pragma unroll
          for (int d = 0; d < N; ++d) mp[((d * N + c) * N + b) * N + a] = cconj(cmul(X[c * N + d], inv));
      }
    }
  }
}

// ------------------------------------------------------------------------
// f4 screened interaction (dynamical_screened_interaction_W,
// dynamical_screened_interaction.cpp:33-128): per (frequency, k) point
//   M_W = M_V inverse(1 - M_PI M_V) = inverse(1 - M_V M_PI) M_V
// (push-through), M_X[(a,b),(c,d)] = X_abcd.  The second form is the RPA
// solve above with V as the right-hand side and PI, which changes from point
// to point, in place of U: N2 lanes own a point, lane i holding row i of
// [A | V]; the point's PI is staged in a wave-private LDS slot to form
// A = 1 - V PI, and the pivot row goes through LDS.  V is [nk][n^4] and read
// at k = pt % nk (static), or [npts][n^4] (dynamic).
template <int N>
__global__ void __launch_bounds__(256) k_screened_W(long npts, long nk, int dynamic, const c2 *__restrict__ PI,
                                                   const c2 *__restrict__ V, c2 *__restrict__ W, double *status) {
  constexpr int N2 = N * N;
  constexpr int P = 64 / N2;  // points per wave
  __shared__ c2 sPi[4 * P][N2 * N2];
  __shared__ c2 slot[4 * P][2 * N2];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gidx = lane / N2, i = lane - gidx * N2;
  const int slot_id = wave * P + (gidx < P ? gidx : 0);
  const long pt = ((long)blockIdx.x * 4 + wave) * P + gidx;
  const bool valid = (gidx < P) && (pt < npts);
  c2 X[N2], A[N2];
  if (valid) {
    const c2 *vrow = V + (dynamic ? pt : pt % nk) * N2 * N2 + i * N2;
    const c2 *prow = PI + pt * N2 * N2 + i * N2;
#pragma unroll
    for (int j = 0; j < N2; ++j) {
      X[j] = vrow[j];
      sPi[slot_id][i * N2 + j] = prow[j];
    }
  } else {
#pragma unroll
    for (int j = 0; j < N2; ++j) X[j] = c2{0.0, 0.0};
  }
  wave_lds_sync();
  // A = 1 - V PI, row i
#pragma unroll
  for (int j = 0; j < N2; ++j) {
    c2 s = (i == j) ? c2{1.0, 0.0} : c2{0.0, 0.0};
    if (gidx < P) {
#pragma unroll
      for (int l = 0; l < N2; ++l) s = cfms(s, X[l], sPi[slot_id][l * N2 + j]);
    }
    A[j] = s;
  }
  bool used = false;
  int myk = 0;
  c2 pv = c2{1.0, 0.0};
  bool singular = false;
#pragma unroll
  for (int k = 0; k < N2; ++k) {
    // pivot: the upper 32 bits of |A[k]|^2 as a key, compared by shuffles
    const int key = (!used && valid) ? (int)(__double_as_longlong(cabs2(A[k])) >> 32) : -1;
    int best_i = 0, best = -2;
#pragma unroll
    for (int s = 0; s < N2; ++s) {
      const int m = __shfl(key, (lane - i) + s, 64);
      if (m > best) { best = m; best_i = s; }
    }
    const bool is_piv = (i == best_i);
    if (valid && best <= 0) singular = true;
    if (is_piv && gidx < P) {
      used = true;
      myk = k;
      pv = A[k];
#pragma unroll
      for (int j = k; j < N2; ++j) slot[slot_id][j] = A[j];
#pragma unroll
      for (int j = 0; j < N2; ++j) slot[slot_id][N2 + j] = X[j];
    }
    wave_lds_sync();
    if (!is_piv && gidx < P) {
      const c2 f = cdiv(A[k], slot[slot_id][k]);
#pragma unroll
      for (int j = k + 1; j < N2; ++j) A[j] = cfms(A[j], f, slot[slot_id][j]);
      A[k] = c2{0.0, 0.0};
#pragma unroll
      for (int j = 0; j < N2; ++j) X[j] = cfms(X[j], f, slot[slot_id][N2 + j]);
    }
    wave_lds_sync();
  }
  if (singular) status[2] = 1.0;
  if (valid) {
    const c2 inv = cdiv(c2{1.0, 0.0}, pv);
    c2 *row = W + pt * N2 * N2 + myk * N2;
#pragma unroll
    for (int j = 0; j < N2; ++j) row[j] = cmul(X[j], inv);
  }
}

// ------------------------------------------------------------------------
// RPA and W for 5 <= n <= 8 (M = n^2 = 25..64 rows): one workgroup per point,
// the augmented matrix [A | B] (M x 2M, 32 M^2 bytes <= 128 KiB) in LDS,
// Gauss-Jordan with partial pivoting and row swaps.  A = 1 - Lm Rm and B = Lm,
// so that B ends as inverse(1 - Lm Rm) Lm: (Lm, Rm) = (chi0, U) gives M_chi
// (rpa.cpp:57), (V, PI) gives M_W (the push-through form above).
//   lmode 0: Lm per point at base + pt; 1: at pt % nk (static V)
//   rmode 0: Rm per point at base + pt; 1: one matrix for every point (U)
//   mirror: also write chi_dcba(-nu) = conj(chi_abcd(nu)) as k_rpa<N, true>
__global__ void __launch_bounds__(256) k_solve_wide(int n, long npts, const c2 *__restrict__ Lm, int lmode,
                                                    const c2 *__restrict__ Rm, int rmode, c2 *__restrict__ out,
                                                    double *status, long base, long nk, int mirror) {
  extern __shared__ c2 sm[];
  __shared__ int s_p;
  __shared__ c2 s_piv;
  const int M = n * n, W2 = 2 * M;
  c2 *T = sm;  // T[i W2 + j]: columns j < M of A, j >= M of B
  const long pt = xcd_remap(blockIdx.x, gridDim.x);
  if (pt >= npts) return;
  const long MM = (long)M * M;
  const c2 *lp = Lm + (lmode == 0 ? base + pt : pt % nk) * MM;
  const c2 *rp = Rm + (rmode == 0 ? (base + pt) * MM : 0);
  for (int t = threadIdx.x; t < M * M; t += blockDim.x) {
    const int i = t / M, j = t - i * M;
    T[i * W2 + M + j] = lp[t];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < M * M; t += blockDim.x) {
    const int i = t / M, j = t - i * M;
    c2 acc = (i == j) ? c2{1.0, 0.0} : c2{0.0, 0.0};
    const c2 *li = T + i * W2 + M;
    for (int l = 0; l < M; ++l) acc = cfms(acc, li[l], rp[(long)l * M + j]);
    T[i * W2 + j] = acc;
  }
  __syncthreads();
  bool singular = false;
  for (int k = 0; k < M; ++k) {
    if (threadIdx.x < 64) {
      // pivot: largest |A_ik|^2 over rows i >= k, lowest row on ties
      const int i = k + threadIdx.x;
      double v = (i < M) ? cabs2(T[i * W2 + k]) : -1.0;
      int idx = (i < M) ? i : M;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(idx, off, 64);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
      }
      if (threadIdx.x == 0) {
        s_p = idx;
        s_piv = T[idx * W2 + k];
        if (!(v > 0.0)) singular = true;
      }
    }
    __syncthreads();
    const int p = s_p;
    const c2 inv = cdiv(c2{1.0, 0.0}, s_piv);
    // swap rows k and p over the columns >= k, the new row k scaled by 1 / pivot
    for (int j = k + threadIdx.x; j < W2; j += blockDim.x) {
      const c2 a = T[p * W2 + j], b = T[k * W2 + j];
      T[k * W2 + j] = cmul(a, inv);
      if (p != k) T[p * W2 + j] = b;
    }
    __syncthreads();
    // eliminate column k from every other row (columns > k; column k is not
    // read again)
    const int w = W2 - k - 1;
    for (int t = threadIdx.x; t < M * w; t += blockDim.x) {
      const int i = t / w;
      if (i == k) continue;
      const int j = k + 1 + (t - i * w);
      T[i * W2 + j] = cfms(T[i * W2 + j], T[i * W2 + k], T[k * W2 + j]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && singular) status[2] = 1.0;
  c2 *o = out + (base + pt) * MM;
  for (int t = threadIdx.x; t < M * M; t += blockDim.x) {
    const int i = t / M, j = t - i * M;
    o[t] = T[i * W2 + M + j];
  }
  if (mirror) {
    const long m = pt / nk;
    if (m > 0) {
      c2 *mp = out + (base + pt - 2 * m * nk) * MM;
      for (int t = threadIdx.x; t < M * M; t += blockDim.x) {
        const int i = t / M, j = t - i * M;
        const int a = i / n, b = i - a * n, c = j / n, d = j - c * n;
        mp[((d * n + c) * n + b) * n + a] = cconj(T[i * W2 + M + j]);
      }
    }
  }
}

hipError_t launch_solve_wide(hipStream_t s, int n, long npts, const c2 *Lm, int lmode, const c2 *Rm, int rmode,
                             c2 *out, Status st, long base, long nk, int mirror) {
  if (n < 5 || n > kWideMaxOrb) return hipErrorInvalidValue;
  const size_t lds = (size_t)32 * n * n * n * n;
  hipLaunchKernelGGL(k_solve_wide, dim3(npts), dim3(256), lds, s, n, npts, Lm, lmode, Rm, rmode, out, st.d, base, nk,
                     mirror);
  return hipGetLastError();
}

hipError_t launch_screened_W(hipStream_t s, long npts, long nk, int n, int dynamic, const c2 *PI, const c2 *V, c2 *W,
                             Status st) {
  switch (n) {
#define TPRF_SW_CASE(N)                                                                                  \
  case N: {                                                                                              \
    constexpr int P = 64 / (N * N);                                                                      \
    const long nblk = (npts + 4 * P - 1) / (4 * P);                                                      \
    hipLaunchKernelGGL(k_screened_W<N>, dim3(nblk), dim3(256), 0, s, npts, nk, dynamic, PI, V, W, st.d); \
    break;                                                                                               \
  }
    TPRF_SW_CASE(1)
    TPRF_SW_CASE(2)
    TPRF_SW_CASE(3)
    TPRF_SW_CASE(4)
#undef TPRF_SW_CASE
    default: return launch_solve_wide(s, n, npts, V, dynamic ? 0 : 1, PI, 0, W, st, 0, nk, 0);
  }
  return hipGetLastError();
}

hipError_t launch_rpa_mirror(hipStream_t s, long nb, long nk, int n, const c2 *chi0, const c2 *U, c2 *chi,
                             Status st) {
  const long npts = nb * nk, base = (nb - 1) * nk;
  switch (n) {
#define TPRF_RPA_MCASE(N)                                                                            \
  case N: {                                                                                         \
    constexpr int P = 64 / (N * N);                                                                 \
    const long nblk = (npts + 4 * P - 1) / (4 * P);                                                 \
    hipLaunchKernelGGL((k_rpa<N, true>), dim3(nblk), dim3(256), 0, s, npts, chi0, U, chi, st.d, base, nk); \
    break;                                                                                          \
  }
    TPRF_RPA_MCASE(1)
    TPRF_RPA_MCASE(2)
    TPRF_RPA_MCASE(3)
    TPRF_RPA_MCASE(4)
#undef TPRF_RPA_MCASE
    default: return launch_solve_wide(s, n, npts, chi0, 0, U, 1, chi, st, base, nk, 1);
  }
  return hipGetLastError();
}

hipError_t launch_rpa(hipStream_t s, long npts, int n, const c2 *chi0, const c2 *U, c2 *chi, Status st) {
  switch (n) {
#define TPRF_RPA_CASE(N)                                                                           \
  case N: {                                                                                        \
    constexpr int P = 64 / (N * N);                                                                \
    const long nb = (npts + 4 * P - 1) / (4 * P);                                                  \
    hipLaunchKernelGGL(k_rpa<N>, dim3(nb), dim3(256), 0, s, npts, chi0, U, chi, st.d);              \
    break;                                                                                         \
  }
    TPRF_RPA_CASE(1)
    TPRF_RPA_CASE(2)
    TPRF_RPA_CASE(3)
    TPRF_RPA_CASE(4)
#undef TPRF_RPA_CASE
    default: return launch_solve_wide(s, n, npts, chi0, 0, U, 1, chi, st, 0, 1, 0);
  }
  return hipGetLastError();
}

}  // namespace tprf

namespace tprf {

// Kernels that request more than 64 KiB of dynamic LDS opt in once per process.
// The single-column kernels ask for up to kMatsBigLdsBytes; they are used only
// where the device grants that much to one workgroup (with room for their
// static arrays).
hipError_t set_kernel_attributes() {
  const int bytes = kMatsLdsBytes;
  const void *fns[] = {(const void *)k_w2t<0, false>, (const void *)k_w2t<1, false>, (const void *)k_t2w<0, false>,
                       (const void *)k_t2w<1, false>, (const void *)k_chi0_wr<false>};
  for (const void *f : fns) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  {
    hipError_t e = hipFuncSetAttribute((const void *)k_solve_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       32 * kWideMaxOrb * kWideMaxOrb * kWideMaxOrb * kWideMaxOrb);
    if (e != hipSuccess) return e;
  }
  int dev = 0, lds_max = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  if (e != hipSuccess) return e;
  if (lds_max < kMatsBigLdsBytes + 8 * 1024) return hipSuccess;
  const void *big[] = {(const void *)k_w2t<0, true>, (const void *)k_w2t<1, true>, (const void *)k_t2w<0, true>,
                       (const void *)k_t2w<1, true>, (const void *)k_chi0_wr<true>};
  for (const void *f : big) {
    e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, kMatsBigLdsBytes);
    if (e != hipSuccess) return e;
  }
  g_mats_big_ok = true;
  return hipSuccess;
}

}  // namespace tprf

namespace tprf {

// ------------------------------------------------------------------------
// a2 / a6 single pass: the whole 3-D transform of C target components of one
// outer index in LDS (layout [k][c], C = 4096 / N_k), each axis as register
// DFTs of length D <= 16 along the lines of that axis.
// LDS index of element (k, c): one pad slot per 16 k keeps the lanes of a
// contiguous line (stride 16 elements) on distinct banks
__device__ __forceinline__ int l3_idx(int k, int c, int C) { return (k + (k >> 4)) * C + c; }

// Radix plan of one lattice axis of length D: in-place decimation-in-frequency
// stages of radix R0 then R1 (R1 = 1: one stage).  Lengths whose factors are
// not all in {2, 3, 4, 5, 8, 16} have no plan (R0 = 0).
template <int D>
struct AxisPlan {
  static constexpr int R0 = (D == 1) ? 1
                            : (D == 2 || D == 3 || D == 4 || D == 5 || D == 8 || D == 16) ? D
                            : (D == 6) ? 3 : (D == 10) ? 5 : (D == 12) ? 4 : (D == 20) ? 5 : (D == 24) ? 8
                            : (D == 32) ? 8 : (D == 40) ? 8 : (D == 48) ? 16 : (D == 64) ? 8 : (D == 80) ? 16 : 0;
  static constexpr int R1 = (R0 == 0 || R0 == D) ? 1 : D / R0;
  static constexpr bool ok = R0 != 0 && (R1 == 1 || R1 == 2 || R1 == 3 || R1 == 4 || R1 == 5 || R1 == 8 ||
                                         R1 == 16);
};

// One DIF stage over every line of an axis held in LDS: span S (= D for the
// first stage), radix R, sub-length m = S / R; the R points b*S + j + r*m of a
// butterfly are transformed in place and twiddled by W_S^{j r'} = W_D^{(D/S) j r'}.
template <int R, int D, int SIGN>
__device__ __forceinline__ void lds_axis_stage(c2 *sm, const c2 *__restrict__ tw, int S, int stride, int nlines,
                                               int line_mul0, int line_div, int line_mul1, int C) {
  const int m = S / R;
  const int nbf = D / R;
  const int items = C * nlines * nbf;
  for (int it = threadIdx.x; it < items; it += blockDim.x) {
    const int c = it % C;
    const int rest = it / C;
    const int line = rest / nbf, b = rest - (rest / nbf) * nbf;
    // base offset of the line: (line / line_div) * line_mul0 + (line % line_div) * line_mul1
    const int base = (line / line_div) * line_mul0 + (line % line_div) * line_mul1;
    const int blk = b / m, j = b - blk * m;
    const int a0 = blk * S + j;
    c2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = sm[l3_idx(base + (a0 + r * m) * stride, c, C)];
    SmallDft<R, SIGN>::run(v, nullptr, 0);
    if (m > 1) {
      const int step = (D / S) * j;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const c2 w = tw[(step * r) % D];  // exp(+2 pi i e / D)
        v[r] = cmul(v[r], SIGN > 0 ? w : cconj(w));
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) sm[l3_idx(base + (a0 + r * m) * stride, c, C)] = v[r];
  }
}

// frequency k of an axis -> its LDS position after the DIF stages (digit reversal)
template <int D>
__device__ __forceinline__ int axis_pos(int k) {
  using Pl = AxisPlan<D>;
  if (Pl::R1 == 1) return k;
  const int r0 = k % Pl::R0, r1 = k / Pl::R0;  // k = r0 + R0 r1
  return r0 * (D / Pl::R0) + r1;                // position r0 m0 + r1, m0 = D / R0
}

template <int D, int SIGN>
__device__ __forceinline__ void lds_axis(c2 *sm, const c2 *tw, int stride, int nlines, int line_mul0, int line_div,
                                         int line_mul1, int C) {
  using Pl = AxisPlan<D>;
  lds_axis_stage<Pl::R0, D, SIGN>(sm, tw, D, stride, nlines, line_mul0, line_div, line_mul1, C);
  if constexpr (Pl::R1 > 1) {
    __syncthreads();
    lds_axis_stage<Pl::R1, D, SIGN>(sm, tw, D / Pl::R0, stride, nlines, line_mul0, line_div, line_mul1, C);
  }
}

// Single-pass 3-D lattice FFT: a workgroup holds the whole D0 x D1 x D2 grid of C
// target columns of one outer index in LDS (padded index k + k / 16), transforms
// every axis in place with its radix plan and writes the result back in natural
// order, scaled.
template <int D0, int D1, int D2, int SIGN>
__global__ void __launch_bounds__(256) k_lattice3d(const c2 *__restrict__ in, c2 *__restrict__ out, long ntarget,
                                                   double scale) {
  constexpr int NK = D0 * D1 * D2;
  constexpr int C = NK >= 4096 ? 1 : 4096 / NK;
  __shared__ c2 sm[(NK + NK / 16 + 1) * C];
  __shared__ c2 tw0[D0], tw1[D1], tw2[D2];
  const long ntc = (ntarget + C - 1) / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long w = tile / ntc;
  const long c0 = (tile - w * ntc) * C;
  const int cn = (ntarget - c0) < C ? (int)(ntarget - c0) : C;
  for (int t = threadIdx.x; t < D0 + D1 + D2; t += blockDim.x) {
    // exp(+2 pi i e / D): sincospi of an exactly reduced fraction
    double sn, cs;
    if (t < D0) { sincospi(2.0 * t / D0, &sn, &cs); tw0[t] = c2{cs, sn}; }
    else if (t < D0 + D1) { const int e = t - D0; sincospi(2.0 * e / D1, &sn, &cs); tw1[e] = c2{cs, sn}; }
    else { const int e = t - D0 - D1; sincospi(2.0 * e / D2, &sn, &cs); tw2[e] = c2{cs, sn}; }
  }
  const c2 *src = in + w * NK * ntarget + c0;
  for (int t = threadIdx.x; t < NK * C; t += blockDim.x) {
    const int k = t / C, c = t - k * C;
    sm[l3_idx(k, c, C)] = c < cn ? src[(long)k * ntarget + c] : c2{0.0, 0.0};
  }
  __syncthreads();
  // axis 2 (stride 1): line = i0 * D1 + i1, base = line * D2
  if constexpr (D2 > 1) {
    lds_axis<D2, SIGN>(sm, tw2, 1, D0 * D1, D2, 1, 0, C);
    __syncthreads();
  }
  // axis 1 (stride D2): line = i0 * D2 + i2, base = i0 * D1 * D2 + i2
  if constexpr (D1 > 1) {
    lds_axis<D1, SIGN>(sm, tw1, D2, D0 * D2, D1 * D2, D2, 1, C);
    __syncthreads();
  }
  // axis 0 (stride D1 D2): line = i1 * D2 + i2, base = line
  if constexpr (D0 > 1) {
    lds_axis<D0, SIGN>(sm, tw0, D1 * D2, D1 * D2, 1, 1, 0, C);
    __syncthreads();
  }
  c2 *dst = out + w * NK * ntarget + c0;
  for (int t = threadIdx.x; t < NK * cn; t += blockDim.x) {
    const int k = t / cn, c = t - k * cn;
    const int k2 = k % D2, k1 = (k / D2) % D1, k0 = k / (D1 * D2);
    const int p = (axis_pos<D0>(k0) * D1 + axis_pos<D1>(k1)) * D2 + axis_pos<D2>(k2);
    dst[(long)k * ntarget + c] = cscale(sm[l3_idx(p, c, C)], scale);
  }
}

// Register-resident 16^3 lattice FFT of C adjacent target columns of one
// outer index.  Thread t = C p + cc (p = 16 hi + lo) holds a 16-point line of
// column cc in registers and transforms it with one radix-16 register DFT per
// axis; lanes of adjacent columns are adjacent, so each row of the grid is
// read and written as C * 16 contiguous bytes.  LDS (columns interleaved)
// only carries the two transposes between axes, each in two rounds over half
// of the grid; in each round the same threads write all of their 16 values
// and then read their next line, so no second register array is live.
//   axis 0: line p = (hi, lo) = (i1, i2) holds i0 = 0..15 (coalesced loads)
//   axis 1: line p = (k0, i2) holds i1; rounds split i2 = lo in halves
//   axis 2: line p = (k0, k1) holds i2; rounds split k0 = hi in halves
// The three axis transforms and two transposes of the register-resident 16^3
// FFT: thread t = C p + cc (p = 16 hi + lo) enters with the 16-point line
// i0 = 0..15 at (i1, i2) = (hi, lo) of column cc and leaves with the line
// k2 = 0..15 at (k0, k1) = (hi, lo).
template <int C, int SIGN>
__device__ __forceinline__ void l16r_transform(c2 *v, c2 *sm, int cc, int hi, int lo) {
  SmallDft<16, SIGN>::run(v, nullptr, 0);
  // transpose 1: element (k0, i1, i2) of half h (i2 in [8h, 8h + 8)) at
  // e = (k0 * 16 + i1) * 8 + (i2 - 8h), one pad slot per 128 entries
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if ((lo >> 3) == h) {
      const int b = lo - 8 * h;
#pragma unroll
      for (int k0 = 0; k0 < 16; ++k0) {
        const int e = (k0 * 16 + hi) * 8 + b;
        sm[(e + (e >> 7)) * C + cc] = v[k0];
      }
    }
    __syncthreads();
    if ((lo >> 3) == h) {
      const int b = lo - 8 * h;
#pragma unroll
      for (int i1 = 0; i1 < 16; ++i1) {
        const int e = (hi * 16 + i1) * 8 + b;
        v[i1] = sm[(e + (e >> 7)) * C + cc];
      }
    }
    __syncthreads();
  }
  SmallDft<16, SIGN>::run(v, nullptr, 0);
  // transpose 2: element (k0, k1, i2) of half h (k0 in [8h, 8h + 8)) at
  // e = ((k0 - 8h) * 16 + k1) * 16 + i2, one pad slot per 16 entries
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if ((hi >> 3) == h) {
      const int a = hi - 8 * h;
#pragma unroll
      for (int k1 = 0; k1 < 16; ++k1) {
        const int e = (a * 16 + k1) * 16 + lo;
        sm[(e + (e >> 4)) * C + cc] = v[k1];
      }
    }
    __syncthreads();
    if ((hi >> 3) == h) {
      const int a = hi - 8 * h;
#pragma unroll
      for (int i2 = 0; i2 < 16; ++i2) {
        const int e = (a * 16 + lo) * 16 + i2;
        v[i2] = sm[(e + (e >> 4)) * C + cc];
      }
    }
    __syncthreads();
  }
  SmallDft<16, SIGN>::run(v, nullptr, 0);
}

// Register-resident 16^3 lattice FFT of C adjacent target columns of one
// outer index.  Thread t = C p + cc (p = 16 hi + lo) holds a 16-point line of
// column cc in registers and transforms it with one radix-16 register DFT per
// axis; lanes of adjacent columns are adjacent, so each row of the grid is
// read and written as C * 16 contiguous bytes.  LDS (columns interleaved)
// only carries the two transposes between axes, each in two rounds over half
// of the grid; in each round the same threads write all of their 16 values
// and then read their next line, so no second register array is live.
//   axis 0: line p = (hi, lo) = (i1, i2) holds i0 = 0..15 (coalesced loads)
//   axis 1: line p = (k0, i2) holds i1; rounds split i2 = lo in halves
//   axis 2: line p = (k0, k1) holds i2; rounds split k0 = hi in halves
template <int C, int SIGN>
__global__ void __launch_bounds__(256 * C) k_lattice16r(const c2 *__restrict__ in, c2 *__restrict__ out,
                                                        long ntarget, double scale) {
  constexpr int NSM = 2048 + 2048 / 16;  // half grid, one pad slot per 16 entries (the larger of the two paddings)
  __shared__ c2 sm[NSM * C];
  const long ntc = (ntarget + C - 1) / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long w = tile / ntc;
  const long c0 = (tile - w * ntc) * C;
  const int t = threadIdx.x, cc = t % C, p = t / C, hi = p >> 4, lo = p & 15;
  const bool live = c0 + cc < ntarget;  // columns past the end are neither read nor written
  const c2 *src = in + w * 4096 * ntarget + c0 + cc;
  c2 *dst = out + w * 4096 * ntarget + c0 + cc;
  c2 v[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = live ? src[(long)(i * 256 + p) * ntarget] : c2{0.0, 0.0};
  l16r_transform<C, SIGN>(v, sm, cc, hi, lo);
  if (live) {
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) dst[(long)(p * 16 + k2) * ntarget] = cscale(v[k2], scale);
  }
}

// Entry [0][0] of Y^-1 for an n x n complex matrix Y, n <= 3, by cofactors:
// (Y^-1)[0][0] = C[0][0] / det Y, det Y = sum_i Y[i][0] C[i][0], with the cyclic
// cofactor formula C[i][j] = Y[i+1][j+1] Y[i+2][j+2] - Y[i+1][j+2] Y[i+2][j+1]
// (indices mod 3).  A thread that needs G[a][b] = (M^-1)[a][b] passes the
// cyclically relabelled Y[i][j] = M[(b + i) mod n][(a + j) mod n]: then
// Y^-1[0][0] = G[a][b], and every register index stays a compile-time constant.
template <int N>
__device__ __forceinline__ c2 inv00(const v2d (&y)[N][N], bool &singular) {
  auto Y = [&](int i, int j) { return c2{y[i % N][j % N].x, y[i % N][j % N].y}; };
  c2 num, det;
  if constexpr (N == 1) {
    num = c2{1.0, 0.0};
    det = Y(0, 0);
  } else if constexpr (N == 2) {
    num = Y(1, 1);
    det = cfms(cmul(Y(0, 0), Y(1, 1)), Y(0, 1), Y(1, 0));
  } else {
    c2 cof[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) cof[i] = cfms(cmul(Y(i + 1, 1), Y(i + 2, 2)), Y(i + 1, 2), Y(i + 2, 1));
    num = cof[0];
    det = c2{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; ++i) det = cfma(det, Y(i, 0), cof[i]);
  }
  if (det.x == 0.0 && det.y == 0.0) singular = true;
  return cdiv(num, det);
}

// Dyson fused into the 16^3 k -> r transform (lattice_dyson_g0_wk, gf.cpp:41-58,
// then fourier_wk_to_wr, fourier_lattice.cpp:39,49), Sigma = 0, n <= 3: the
// workgroup of frequency w and orbital row a forms G[a][b](w, k) =
// inverse((i w_n + mu) 1 - e_k)[a][b] of its k-points from e_k (held in L2)
// instead of reading a stored G(w, k), then transforms the columns (a, b)
// exactly as k_lattice16r<C, -1>.  G(w, k) is never written.  C = n: one
// workgroup per (w, a), its n columns (a, b) adjacent (n * 16-byte row
// pieces); C = 1: one workgroup per (w, a, b) (smaller workgroups, more of them
// resident per CU).
// SYM (e_k Hermitian): only the frequencies w_n >= 0 (index w >= n_iw) are
// formed and transformed; G(i w_{-n-1}, k) = G(i w_n, k)^H gives
// G_ba(i w_{-n-1}, -r) = conj(G_ab(i w_n, r)), written alongside.  half: only
// the columns of r <= -r are written (at every frequency).
template <int N, int C, bool SYM = false>
__global__ void __launch_bounds__(256 * C) k_dyson_lattice16r(double beta, int n_iw, double mu,
                                                              const c2 *__restrict__ ek, c2 *__restrict__ out,
                                                              double scale, double *status, int half = 0) {
  static_assert(C == 1 || C == N, "one column or one orbital row per workgroup");
  constexpr int NT = N * N, NG = NT / C;  // workgroups per frequency
  constexpr int NSM = 2048 + 2048 / 16;
  __shared__ c2 sm[NSM * C];
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long w = tile / NG + (SYM ? n_iw : 0);
  const int g = (int)(tile - (tile / NG) * NG);
  const int a = (C == N) ? g : g / N;
  const int t = threadIdx.x, cc = t % C, p = t / C, hi = p >> 4, lo = p & 15;
  const int b = (C == N) ? cc : g % N;  // this thread's column (a, b)
  const double wn = kPi * (double)(2 * (w - n_iw) + 1) / beta;
  // entry (i, j) of the relabelled matrix: M[(b + i) mod n][(a + j) mod n]
  int off[N][N];
  bool diag[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int r = (b + i) % N, c = (a + j) % N;
      off[i][j] = r * N + c;
      diag[i][j] = r == c;
    }
  bool singular = false;
  c2 v[16];
  // KB k-points at a time: their KB n^2 entries of e_k are loaded before any is used
  constexpr int KB = N == 3 ? 2 : 4;
#pragma unroll
  for (int i4 = 0; i4 < 16; i4 += KB) {
    v2d x[KB][N][N];
#pragma unroll
    for (int u = 0; u < KB; ++u) {
      const c2 *e = ek + (long)((i4 + u) * 256 + p) * NT;
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) x[u][i][j] = ld_v2d(e + off[i][j]);
    }
#pragma unroll
    for (int u = 0; u < KB; ++u) {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          v2d m = -x[u][i][j];
          if (diag[i][j]) { m.x += mu; m.y += wn; }
          x[u][i][j] = m;
        }
      v[i4 + u] = inv00<N>(x[u], singular);
    }
  }
  if (singular) status[2] = 1.0;
  l16r_transform<C, -1>(v, sm, cc, hi, lo);
  c2 *dst = out + w * 4096 * NT + a * N + b;
  if constexpr (SYM) {
    // r = (hi, lo, k2) -> -r = (-hi, -lo, -k2) mod 16, frequency index 2 n_iw - 1 - w
    c2 *mdst = out + (2L * n_iw - 1 - w) * 4096 * NT + b * N + a;
    const int mh = (16 - hi) & 15, ml = (16 - lo) & 15;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) {
      const int r = p * 16 + k2, mr = (mh * 16 + ml) * 16 + ((16 - k2) & 15);
      if (!half || r <= mr) dst[(long)r * NT] = cscale(v[k2], scale);
      if (!half || mr <= r) mdst[(long)mr * NT] = cconj(cscale(v[k2], scale));
    }
  } else {
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) dst[(long)(p * 16 + k2) * NT] = cscale(v[k2], scale);
  }
}

template <int N, int C>
static hipError_t launch_dl16(hipStream_t s, double beta, int n_iw, double mu, const c2 *ek, c2 *out, Status st,
                              int sym) {
  const long nb = (sym ? 1L : 2L) * n_iw * (N * N / C);
  if (sym)
    hipLaunchKernelGGL((k_dyson_lattice16r<N, C, true>), dim3(nb), dim3(256 * C), 0, s, beta, n_iw, mu, ek, out,
                       1.0 / 4096.0, st.d, sym == 2 ? 1 : 0);
  else
    hipLaunchKernelGGL((k_dyson_lattice16r<N, C, false>), dim3(nb), dim3(256 * C), 0, s, beta, n_iw, mu, ek, out,
                       1.0 / 4096.0, st.d);
  return hipGetLastError();
}

// variants (same results): 0 = one orbital row per workgroup, 1 = one column
int dyson_lattice16r_variants(int n) { return n == 1 ? 1 : 2; }

hipError_t launch_dyson_lattice16r(hipStream_t s, double beta, int n_iw, double mu, int n, const c2 *ek, c2 *out,
                                   Status st, int variant, int sym) {
  switch (n) {
    case 1: return launch_dl16<1, 1>(s, beta, n_iw, mu, ek, out, st, sym);
    case 2: return variant == 1 ? launch_dl16<2, 1>(s, beta, n_iw, mu, ek, out, st, sym)
                                : launch_dl16<2, 2>(s, beta, n_iw, mu, ek, out, st, sym);
    case 3: return variant == 1 ? launch_dl16<3, 1>(s, beta, n_iw, mu, ek, out, st, sym)
                                : launch_dl16<3, 3>(s, beta, n_iw, mu, ek, out, st, sym);
    default: return hipErrorInvalidValue;
  }
}

template <int C>
static hipError_t launch_l16r(hipStream_t s, const c2 *in, c2 *out, long nw, long ntarget, int sign, double scale) {
  const long nb = nw * ((ntarget + C - 1) / C);
  if (sign < 0)
    hipLaunchKernelGGL((k_lattice16r<C, -1>), dim3(nb), dim3(256 * C), 0, s, in, out, ntarget, scale);
  else
    hipLaunchKernelGGL((k_lattice16r<C, 1>), dim3(nb), dim3(256 * C), 0, s, in, out, ntarget, scale);
  return hipGetLastError();
}

template <int D0, int D1, int D2>
static hipError_t launch_l3(hipStream_t s, const c2 *in, c2 *out, long nw, long ntarget, int sign, double scale,
                            int variant) {
  if constexpr (D0 == 16 && D1 == 16 && D2 == 16) {
    // columns per workgroup 1, 2 or 4 (same results)
    if (variant == 1) return launch_l16r<2>(s, in, out, nw, ntarget, sign, scale);
    if (variant == 2) return launch_l16r<4>(s, in, out, nw, ntarget, sign, scale);
    return launch_l16r<1>(s, in, out, nw, ntarget, sign, scale);
  } else {
    (void)variant;
    static_assert(AxisPlan<D0>::ok && AxisPlan<D1>::ok && AxisPlan<D2>::ok, "no radix plan");
    constexpr int NK = D0 * D1 * D2;
    constexpr int C = NK >= 4096 ? 1 : 4096 / NK;
    const long nb = nw * ((ntarget + C - 1) / C);
    if (sign < 0)
      hipLaunchKernelGGL((k_lattice3d<D0, D1, D2, -1>), dim3(nb), dim3(256), 0, s, in, out, ntarget, scale);
    else
      hipLaunchKernelGGL((k_lattice3d<D0, D1, D2, 1>), dim3(nb), dim3(256), 0, s, in, out, ntarget, scale);
    return hipGetLastError();
  }
}

// lattice shapes with a single-pass kernel
static const int kLattice3dShapes[][3] = {{16, 16, 16}, {12, 12, 12}, {8, 8, 8}, {6, 6, 6},  {4, 4, 4},
                                          {2, 2, 2},    {64, 64, 1},  {48, 48, 1}, {32, 32, 1}, {24, 24, 1},
                                          {16, 16, 1},  {12, 12, 1},  {8, 8, 1},  {6, 6, 1},   {4, 4, 1},
                                          {2, 2, 1},    {16, 1, 1},   {8, 1, 1},  {4, 4, 2},   {8, 8, 4}};

// number of single-pass kernel variants (same results) for a lattice shape, 0 without one
int lattice3d_variants(const int dims[3]) {
  if (dims[0] == 16 && dims[1] == 16 && dims[2] == 16) return 3;
  for (const auto &d : kLattice3dShapes)
    if (d[0] == dims[0] && d[1] == dims[1] && d[2] == dims[2]) return 1;
  return 0;
}
// variant used when it cannot be timed (in-place calls before the first tuned call)
int lattice3d_default_variant(const int dims[3]) {
  return (dims[0] == 16 && dims[1] == 16 && dims[2] == 16) ? 1 : 0;
}

// the lattice shapes with a single-pass kernel; hipErrorNotSupported otherwise
hipError_t launch_lattice3d(hipStream_t s, int sign, const c2 *in, c2 *out, long nw, const int dims[3], long ntarget,
                            double scale, int variant) {
  const int a = dims[0], b = dims[1], c = dims[2];
#define TPRF_L3(A, B, C_)                                                              \
  if (a == A && b == B && c == C_) return launch_l3<A, B, C_>(s, in, out, nw, ntarget, sign, scale, variant);
  TPRF_L3(16, 16, 16)
  TPRF_L3(12, 12, 12)
  TPRF_L3(8, 8, 8)
  TPRF_L3(6, 6, 6)
  TPRF_L3(4, 4, 4)
  TPRF_L3(2, 2, 2)
  TPRF_L3(64, 64, 1)
  TPRF_L3(48, 48, 1)
  TPRF_L3(32, 32, 1)
  TPRF_L3(24, 24, 1)
  TPRF_L3(16, 16, 1)
  TPRF_L3(12, 12, 1)
  TPRF_L3(8, 8, 1)
  TPRF_L3(6, 6, 1)
  TPRF_L3(4, 4, 1)
  TPRF_L3(2, 2, 1)
  TPRF_L3(16, 1, 1)
  TPRF_L3(8, 1, 1)
  TPRF_L3(4, 4, 2)
  TPRF_L3(8, 8, 4)
#undef TPRF_L3
  return hipErrorNotSupported;
}

}  // namespace tprf
