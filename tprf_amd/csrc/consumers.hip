// consumers.hip -- the consumers of chi in the lattice module: GW self-energy
// pieces (gw.cpp) and the linearised Eliashberg products (eliashberg.cpp).
// Each kernel is one pointwise contraction, reduction or sum; the Fourier
// transforms around them are the lattice and Matsubara kernels of the chain.
#include "common.hpp"
#include "mats_device.hpp"

namespace tprf {

namespace {

// k + s q on the lattice (i0 d1 + i1) d2 + i2, s = +1 or -1
__device__ __forceinline__ long add_k(long k, long q, int d0, int d1, int d2, int s) {
  const long k2 = k % d2, k1 = (k / d2) % d1, k0 = k / ((long)d1 * d2);
  const long q2 = q % d2, q1 = (q / d2) % d1, q0 = q / ((long)d1 * d2);
  const long j0 = ((k0 + s * q0) % d0 + d0) % d0, j1 = ((k1 + s * q1) % d1 + d1) % d1, j2 = ((k2 + s * q2) % d2 + d2) % d2;
  return (j0 * d1 + j1) * d2 + j2;
}

}  // namespace

// Pointwise vertex contractions, thread per (point, a, b), V[p] of n^4 entries
// (vstride 0: one V for every point), y[p] and out[p] of n^2:
//   MODE 0  gw_dynamic_sigma (gw.cpp:88-102)       out_ab = -sum_cd V_acdb y_cd
//   MODE 1  fock_sigma, r space (gw.cpp:149-167)     out_ab = -sum_cd V_acdb y_dc
//   MODE 2  hartree_sigma (gw.cpp:113-122)           out_ab =  sum_cd V_abcd y_cd
//   MODE 3  eliashberg gamma-F products (eliashberg.cpp:175-186, 214-240)
//                                                    out_ab = -1/2 sum_cd V_cadb y_dc
template <int N, int MODE>
__global__ void __launch_bounds__(256) k_vertex_contract(long npts, const c2 *__restrict__ V, long vstride,
                                                         const c2 *__restrict__ y, c2 *__restrict__ out) {
  constexpr int n2 = N * N;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts * n2) return;
  const long p = t / n2;
  const int ab = (int)(t - p * n2), a = ab / N, b = ab - a * N;
  const c2 *v = V + p * vstride;
  const c2 *yp = y + p * n2;
  c2 s = c2{0, 0};
#pragma unroll
  for (int c = 0; c < N; ++c)
#pragma unroll
    for (int d = 0; d < N; ++d) {
      if (MODE == 0) s = cfma(s, v[((a * N + c) * N + d) * N + b], yp[c * N + d]);
      if (MODE == 1) s = cfma(s, v[((a * N + c) * N + d) * N + b], yp[d * N + c]);
      if (MODE == 2) s = cfma(s, v[((a * N + b) * N + c) * N + d], yp[c * N + d]);
      if (MODE == 3) s = cfma(s, v[((c * N + a) * N + d) * N + b], yp[d * N + c]);
    }
  const double scale = MODE == 0 || MODE == 1 ? -1.0 : MODE == 2 ? 1.0 : -0.5;
  out[t] = cscale(s, scale);
}

// Static self-energies on the k mesh by the direct sum over q
// (hartree_sigma gw.cpp:124-147, fock_sigma gw.cpp:169-195):
//   HARTREE  sigma_ab(k) =  (1/N) sum_q sum_cd V_abcd(q) rho_cd(k + q)
//   FOCK     sigma_ab(k) = -(1/N) sum_q sum_cd V_acdb(q) rho_dc(k + q)
template <int N, bool FOCK>
__global__ void __launch_bounds__(256) k_static_sigma_k(int d0, int d1, int d2, const c2 *__restrict__ V,
                                                        const c2 *__restrict__ rho, c2 *__restrict__ out) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nk = (long)d0 * d1 * d2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk * n2) return;
  const long k = t / n2;
  const int ab = (int)(t - k * n2), a = ab / N, b = ab - a * N;
  c2 s = c2{0, 0};
  for (long q = 0; q < nk; ++q) {
    const c2 *v = V + q * n4;
    const c2 *r = rho + add_k(k, q, d0, d1, d2, +1) * n2;
#pragma unroll
    for (int c = 0; c < N; ++c)
#pragma unroll
      for (int d = 0; d < N; ++d) {
        if (FOCK) s = cfma(s, v[((a * N + c) * N + d) * N + b], r[d * N + c]);
        else s = cfma(s, v[((a * N + b) * N + c) * N + d], r[c * N + d]);
      }
  }
  out[t] = cscale(s, (FOCK ? -1.0 : 1.0) / (double)nk);
}

// split_into_dynamic_wk_and_constant_k (lattice_utility.cpp:57-84): per column
// the zeroth coefficient of the tail fit is the constant part, the rest the
// dynamic part.  x, dyn: [nmesh][ncol]; cst: [ncol].
__global__ void __launch_bounds__(256) k_tail_split(long ncol, int nmesh, TailFitDev fit, const c2 *__restrict__ x,
                                                    c2 *__restrict__ dyn, c2 *__restrict__ cst) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  c2 c0 = c2{0, 0};
  for (int f = 0; f < fit.n_fit; ++f) c0 = cfma(c0, fit.pinv[f], x[(long)fit.idx[f] * ncol + col]);
  cst[col] = c0;
  for (int m = 0; m < nmesh; ++m) dyn[(long)m * ncol + col] = csub(x[(long)m * ncol + col], c0);
}

// density of a fermionic Matsubara function per column (the TRIQS density of
// an imfreq Green's function used by rho_k_from_g_wk, gw.cpp:36-52): with the
// fitted moments t1..t3,
//   rho = (1/beta) sum_n [g_n - t1/(iw) - t2/(iw)^2 - t3/(iw)^3] + t1/2 - t2 beta/4,
// the closed-form sums of the three tail terms being 1/2, -beta/4 and 0.
__global__ void __launch_bounds__(256) k_density(long ncol, int n_max, double beta, TailFitDev fit,
                                                 const c2 *__restrict__ x, c2 *__restrict__ rho,
                                                 double *status) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  c2 coef[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) coef[k] = c2{0, 0};
  for (int f = 0; f < fit.n_fit; ++f) {
    const c2 v = x[(long)fit.idx[f] * ncol + col];
#pragma unroll
    for (int k = 0; k < 4; ++k) coef[k] = cfma(coef[k], fit.pinv[(long)k * fit.n_fit + f], v);
  }
  const c2 t1 = cscale(coef[1], fit.wpow[1]), t2 = cscale(coef[2], fit.wpow[2]), t3 = cscale(coef[3], fit.wpow[3]);
  c2 s = c2{0, 0};
  for (int m = 0; m < 2 * n_max; ++m) {
    const double w = kPi * (double)(2 * (m - n_max) + 1) / beta;
    const double iw = 1.0 / w;  // 1/(i w) = -i / w
    // t1/(iw) = -i t1 / w ; t2/(iw)^2 = -t2 / w^2 ; t3/(iw)^3 = i t3 / w^3
    c2 tail = cmul_i(t1, -iw);
    tail = cadd(tail, cscale(t2, -iw * iw));
    tail = cadd(tail, cmul_i(t3, iw * iw * iw));
    s = cadd(s, csub(x[(long)m * ncol + col], tail));
  }
  rho[col] = cadd(cadd(cscale(s, 1.0 / beta), cscale(t1, 0.5)), cscale(t2, -0.25 * beta));
  atomic_max_nonneg(&status[1], sqrt(cabs2(coef[0])));
}

// F_dc(w, k) = sum_ef G_cf(w, k) conj(G_ed(w, -k)) Delta_ef(w, k)
// (eliashberg.cpp:31-58); w runs over Delta's mesh, G is read at the same
// Matsubara frequency (mesh offset off = n_max_g - n_max_d).
template <int N>
__global__ void __launch_bounds__(256) k_g_delta_g(long nwd, int off, int d0, int d1, int d2,
                                                   const c2 *__restrict__ g, const c2 *__restrict__ delta,
                                                   c2 *__restrict__ F) {
  constexpr int n2 = N * N;
  const long nk = (long)d0 * d1 * d2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nwd * nk * n2) return;
  const long wk = t / n2;
  const int dc = (int)(t - wk * n2), d = dc / N, c = dc - d * N;
  const long w = wk / nk, k = wk - w * nk;
  const long mk = minus_r(k, d0, d1, d2);
  const c2 *gp = g + ((w + off) * nk + k) * n2, *gm = g + ((w + off) * nk + mk) * n2;
  const c2 *dl = delta + wk * n2;
  c2 s = c2{0, 0};
#pragma unroll
  for (int e = 0; e < N; ++e)
#pragma unroll
    for (int f = 0; f < N; ++f) s = cfma(s, cmulc(gp[c * N + f], gm[e * N + d]), dl[e * N + f]);
  F[t] = s;
}

// eliashberg_product by summation (eliashberg.cpp:92-120):
//   Delta_ab(w, k) = -1/(2 N beta) sum_{n, q} sum_cd Gamma_cadb(w - n, k - q) F_dc(n, q)
// on Delta's fermionic mesh (2 nd points); Gamma's bosonic mesh holds
// m = -(ng - 1)..ng - 1, ng >= 2 nd + 1.
template <int N>
__global__ void __launch_bounds__(256) k_eliashberg_sum(int nd, int ng, int d0, int d1, int d2, double scale,
                                                        const c2 *__restrict__ gamma, const c2 *__restrict__ F,
                                                        c2 *__restrict__ out) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nk = (long)d0 * d1 * d2;
  const long nw = 2L * nd;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nw * nk * n2) return;
  const long wk = t / n2;
  const int ab = (int)(t - wk * n2), a = ab / N, b = ab - a * N;
  const long w = wk / nk, k = wk - w * nk;
  c2 s = c2{0, 0};
  for (long n = 0; n < nw; ++n) {
    const long gm = (w - n) + (ng - 1);
    for (long q = 0; q < nk; ++q) {
      const c2 *G = gamma + (gm * nk + add_k(k, q, d0, d1, d2, -1)) * n4;
      const c2 *f = F + (n * nk + q) * n2;
#pragma unroll
      for (int c = 0; c < N; ++c)
#pragma unroll
        for (int d = 0; d < N; ++d) s = cfma(s, G[((c * N + a) * N + d) * N + b], f[d * N + c]);
    }
  }
  out[t] = cscale(s, scale);
}

// construct_phi_wk (eliashberg.cpp:336-360): Phi = U chi U on the matrices
// M[(a,b),(d,c)] = X_abcd, thread per (point, row (a, b)).
template <int N>
__global__ void __launch_bounds__(256) k_construct_phi(long npts, const c2 *__restrict__ chi, const c2 *__restrict__ U,
                                                       c2 *__restrict__ phi) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts * n2) return;
  const long p = t / n2;
  const int i = (int)(t - p * n2);
  // flat index of M[r][s]: r n^2 + (s % n) n + s / n
  auto at = [](int r, int s) { return r * n2 + (s % N) * N + s / N; };
  const c2 *x = chi + p * n4;
  c2 row[n2];
#pragma unroll
  for (int l = 0; l < n2; ++l) {
    c2 v = c2{0, 0};
#pragma unroll
    for (int k = 0; k < n2; ++k) v = cfma(v, U[at(i, k)], x[at(k, l)]);
    row[l] = v;
  }
  c2 *o = phi + p * n4;
#pragma unroll
  for (int j = 0; j < n2; ++j) {
    c2 v = c2{0, 0};
#pragma unroll
    for (int l = 0; l < n2; ++l) v = cfma(v, row[l], U[at(l, j)]);
    o[at(i, j)] = v;
  }
}

// The same for 5 <= n <= 8: one workgroup per point, T = M_U M_chi in LDS
// (16 n^4 bytes), then Phi = T M_U.
__global__ void __launch_bounds__(256) k_construct_phi_wide(int n, long npts, const c2 *__restrict__ chi,
                                                            const c2 *__restrict__ U, c2 *__restrict__ phi) {
  extern __shared__ c2 T[];
  const int n2 = n * n;
  const long n4 = (long)n2 * n2;
  const long p = blockIdx.x;
  if (p >= npts) return;
  auto at = [n, n2](int r, int s) { return r * n2 + (s % n) * n + s / n; };
  const c2 *x = chi + p * n4;
  for (int t = threadIdx.x; t < n2 * n2; t += blockDim.x) {
    const int i = t / n2, l = t - i * n2;
    c2 v = c2{0, 0};
    for (int k = 0; k < n2; ++k) v = cfma(v, U[at(i, k)], x[at(k, l)]);
    T[t] = v;
  }
  __syncthreads();
  c2 *o = phi + p * n4;
  for (int t = threadIdx.x; t < n2 * n2; t += blockDim.x) {
    const int i = t / n2, j = t - i * n2;
    c2 v = c2{0, 0};
    for (int l = 0; l < n2; ++l) v = cfma(v, T[i * n2 + l], U[at(l, j)]);
    o[at(i, j)] = v;
  }
}

// add_dynamic_and_static (lattice_utility.cpp:86-101): x[w][col] += s[col]
__global__ void __launch_bounds__(256) k_add_static(long nw, long ncol, c2 *__restrict__ x, const c2 *__restrict__ s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nw * ncol) return;
  x[t] = cadd(x[t], s[t % ncol]);
}

// x[i] = 0 (a kernel rather than a memset node, so the call may be captured)
__global__ void __launch_bounds__(256) k_zero(long n, c2 *__restrict__ x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = c2{0, 0};
}

// ---------------------------------------------------------------- launchers
namespace {
inline unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }
}  // namespace

hipError_t launch_vertex_contract(hipStream_t s, int mode, int n, long npts, const c2 *V, long vstride, const c2 *y,
                                  c2 *out) {
  const long nt = npts * n * n;
#define TPRF_VC(NN, MM) hipLaunchKernelGGL((k_vertex_contract<NN, MM>), dim3(nblocks(nt)), dim3(256), 0, s, npts, V, \
                                           vstride, y, out)
#define TPRF_VCN(NN)        \
  switch (mode) {           \
    case 0: TPRF_VC(NN, 0); break; \
    case 1: TPRF_VC(NN, 1); break; \
    case 2: TPRF_VC(NN, 2); break; \
    case 3: TPRF_VC(NN, 3); break; \
    default: return hipErrorInvalidValue; \
  }
  switch (n) {
    case 1: TPRF_VCN(1); break;
    case 2: TPRF_VCN(2); break;
    case 3: TPRF_VCN(3); break;
    case 4: TPRF_VCN(4); break;
    case 5: TPRF_VCN(5); break;
    case 6: TPRF_VCN(6); break;
    case 7: TPRF_VCN(7); break;
    case 8: TPRF_VCN(8); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_VCN
#undef TPRF_VC
  return hipGetLastError();
}

hipError_t launch_static_sigma_k(hipStream_t s, bool fock, int n, const int dims[3], const c2 *V, const c2 *rho,
                                 c2 *out) {
  const long nt = (long)dims[0] * dims[1] * dims[2] * n * n;
#define TPRF_SK(NN)                                                                                              \
  if (fock) hipLaunchKernelGGL((k_static_sigma_k<NN, true>), dim3(nblocks(nt)), dim3(256), 0, s, dims[0], dims[1], \
                               dims[2], V, rho, out);                                                           \
  else hipLaunchKernelGGL((k_static_sigma_k<NN, false>), dim3(nblocks(nt)), dim3(256), 0, s, dims[0], dims[1],    \
                          dims[2], V, rho, out)
  switch (n) {
    case 1: TPRF_SK(1); break;
    case 2: TPRF_SK(2); break;
    case 3: TPRF_SK(3); break;
    case 4: TPRF_SK(4); break;
    case 5: TPRF_SK(5); break;
    case 6: TPRF_SK(6); break;
    case 7: TPRF_SK(7); break;
    case 8: TPRF_SK(8); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_SK
  return hipGetLastError();
}

hipError_t launch_tail_split(hipStream_t s, long ncol, int nmesh, const TailFitDev &fit, const c2 *x, c2 *dyn,
                             c2 *cst) {
  hipLaunchKernelGGL(k_tail_split, dim3(nblocks(ncol)), dim3(256), 0, s, ncol, nmesh, fit, x, dyn, cst);
  return hipGetLastError();
}

hipError_t launch_density(hipStream_t s, long ncol, int n_max, double beta, const TailFitDev &fit, const c2 *x,
                          c2 *rho, double *status) {
  hipLaunchKernelGGL(k_density, dim3(nblocks(ncol)), dim3(256), 0, s, ncol, n_max, beta, fit, x, rho, status);
  return hipGetLastError();
}

hipError_t launch_g_delta_g(hipStream_t s, int n, long nwd, int off, const int dims[3], const c2 *g, const c2 *delta,
                            c2 *F) {
  const long nt = nwd * dims[0] * dims[1] * dims[2] * n * n;
#define TPRF_GDG(NN) hipLaunchKernelGGL(k_g_delta_g<NN>, dim3(nblocks(nt)), dim3(256), 0, s, nwd, off, dims[0], \
                                        dims[1], dims[2], g, delta, F)
  switch (n) {
    case 1: TPRF_GDG(1); break;
    case 2: TPRF_GDG(2); break;
    case 3: TPRF_GDG(3); break;
    case 4: TPRF_GDG(4); break;
    case 5: TPRF_GDG(5); break;
    case 6: TPRF_GDG(6); break;
    case 7: TPRF_GDG(7); break;
    case 8: TPRF_GDG(8); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_GDG
  return hipGetLastError();
}

hipError_t launch_eliashberg_sum(hipStream_t s, int n, int nd, int ng, const int dims[3], double scale,
                                 const c2 *gamma, const c2 *F, c2 *out) {
  const long nt = 2L * nd * dims[0] * dims[1] * dims[2] * n * n;
#define TPRF_ES(NN) hipLaunchKernelGGL(k_eliashberg_sum<NN>, dim3(nblocks(nt)), dim3(256), 0, s, nd, ng, dims[0], \
                                       dims[1], dims[2], scale, gamma, F, out)
  switch (n) {
    case 1: TPRF_ES(1); break;
    case 2: TPRF_ES(2); break;
    case 3: TPRF_ES(3); break;
    case 4: TPRF_ES(4); break;
    case 5: TPRF_ES(5); break;
    case 6: TPRF_ES(6); break;
    case 7: TPRF_ES(7); break;
    case 8: TPRF_ES(8); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_ES
  return hipGetLastError();
}

hipError_t launch_construct_phi(hipStream_t s, int n, long npts, const c2 *chi, const c2 *U, c2 *phi) {
  const long nt = npts * n * n;
#define TPRF_CP(NN) hipLaunchKernelGGL(k_construct_phi<NN>, dim3(nblocks(nt)), dim3(256), 0, s, npts, chi, U, phi)
  switch (n) {
    case 1: TPRF_CP(1); break;
    case 2: TPRF_CP(2); break;
    case 3: TPRF_CP(3); break;
    case 4: TPRF_CP(4); break;
    case 5:
    case 6:
    case 7:
    case 8: {
      hipLaunchKernelGGL(k_construct_phi_wide, dim3(npts), dim3(256), (size_t)16 * n * n * n * n, s, n, npts, chi, U,
                         phi);
      break;
    }
    default: return hipErrorInvalidValue;
  }
#undef TPRF_CP
  return hipGetLastError();
}

// 16 n^4 bytes of LDS for k_construct_phi_wide reach 64 KiB at n = 8
hipError_t consumers_set_attributes() {
  return hipFuncSetAttribute((const void *)k_construct_phi_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                             16 * 4096);
}

hipError_t launch_zero(hipStream_t s, long n, c2 *x) {
  hipLaunchKernelGGL(k_zero, dim3(nblocks(n)), dim3(256), 0, s, n, x);
  return hipGetLastError();
}

hipError_t launch_add_static(hipStream_t s, long nw, long ncol, c2 *x, const c2 *st) {
  hipLaunchKernelGGL(k_add_static, dim3(nblocks(nw * ncol)), dim3(256), 0, s, nw, ncol, x, st);
  return hipGetLastError();
}

}  // namespace tprf
