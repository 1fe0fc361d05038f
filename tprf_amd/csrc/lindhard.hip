// lindhard.hip -- the analytic (Lindhard) particle-hole bubble
// (lindhard_chi00, c++/triqs_tprf/lattice/lindhard_chi00.cpp:34-94):
//
//   chi00_abcd(z, q) = 1/N_k sum_{k, i, j} U_k(a,i) U_k^*(d,i) U_{k+q}(c,j) U_{k+q}^*(b,j) F_ij,
//   F_ij = [f(beta e_ki) - f(beta e_{k+q,j})] / (z + e_{k+q,j} - e_ki),
//
// with e_ki, U_k the eigenpairs of e_k - mu, z = i nu_m (Matsubara) or
// omega + i delta (real frequency), and F_ij = beta / (4 cosh^2(beta e_ki / 2))
// when |z| and |e_{k+q,j} - e_ki| are both below 1e-10 (:64-74).
//
// Two kernels: per k the eigenpairs of the Hermitian n x n matrix by cyclic
// complex Jacobi rotations in registers, stored as eigenvalues and band
// projectors P_i(k)[a][d] = U_k(a,i) U_k^*(d,i) (independent of the phase and,
// summed over a degenerate subspace, of the basis of the eigenvectors); then
// one workgroup per (q, block of frequencies) sums over k in LDS tiles.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "fft_device.hpp"

namespace tprf {

// fermi(e) of lattice_utility.cpp:146-153 (the form that does not overflow)
__device__ __forceinline__ double fermi_fn(double e) {
  if (e < 0) return 1.0 / (exp(e) + 1.0);
  const double em = exp(-e);
  return em / (1.0 + em);
}

// Eigenpairs of H = e_k - mu (Hermitian, n <= 4) by cyclic Jacobi sweeps.  For
// the pair (p, q) with H_pq = |b| e^{i phi}, D = diag(1, e^{-i phi}) makes the
// 2 x 2 block real symmetric [[a, |b|], [|b|, d]]; the rotation R with
// t = tan(theta) = sgn(tau) / (|tau| + sqrt(1 + tau^2)), tau = (d - a) / (2 |b|)
// diagonalises it, and G = D R = [[c, s], [-s e^{-i phi}, c e^{-i phi}]] acts as
// H <- G^H H G, V <- V G.
template <int N>
__global__ void __launch_bounds__(256) k_eigh_proj(long nk, double mu, const c2 *__restrict__ ek,
                                                   double *__restrict__ ev, c2 *__restrict__ proj) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  c2 H[N][N], V[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) {
      H[i][j] = ek[(k * N + i) * N + j];
      V[i][j] = c2{i == j ? 1.0 : 0.0, 0.0};
    }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    H[i][i].x -= mu;
    H[i][i].y = 0.0;
  }
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = 0.0, diag = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
      diag += H[p][p].x * H[p][p].x;
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += cabs2(H[p][q]);
    }
    if (off <= 1e-34 * diag || off == 0.0) break;
#pragma unroll
    for (int p = 0; p < N; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double bb = sqrt(cabs2(H[p][q]));
        if (bb == 0.0) continue;
        const c2 ph = c2{H[p][q].x / bb, -H[p][q].y / bb};  // e^{-i phi}
        const double tau = (H[q][q].x - H[p][p].x) / (2.0 * bb);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        // G = [[c, s], [-s ph, c ph]]: columns p, q of H and V
        const c2 g10 = cscale(ph, -s), g11 = cscale(ph, c);
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const c2 hp = H[r][p], hq = H[r][q];
          H[r][p] = cadd(cscale(hp, c), cmul(hq, g10));
          H[r][q] = cadd(cscale(hp, s), cmul(hq, g11));
          const c2 vp = V[r][p], vq = V[r][q];
          V[r][p] = cadd(cscale(vp, c), cmul(vq, g10));
          V[r][q] = cadd(cscale(vp, s), cmul(vq, g11));
        }
        // rows p, q of H: multiply by G^H
#pragma unroll
        for (int r = 0; r < N; ++r) {
          const c2 hp = H[p][r], hq = H[q][r];
          H[p][r] = cadd(cscale(hp, c), cmul(hq, cconj(g10)));
          H[q][r] = cadd(cscale(hp, s), cmul(hq, cconj(g11)));
        }
        H[p][q] = c2{0.0, 0.0};
        H[q][p] = c2{0.0, 0.0};
        H[p][p].y = 0.0;
        H[q][q].y = 0.0;
      }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    ev[k * N + i] = H[i][i].x;
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int d = 0; d < N; ++d) proj[((k * N + i) * N + a) * N + d] = cmul(V[a][i], cconj(V[d][i]));
  }
}

// One workgroup per (q, block of NV frequencies); thread t owns output
// (frequency z_{v0 + t / n^4}, abcd = t % n^4).  The k-sum runs over tiles of
// KT k-points: eigenvalues and projectors of k and k + q, and the factors
// F[v][kk][i][j] of the block's frequencies, staged in LDS.
//   chi_abcd += sum_i P_i(k)[a][d] sum_j F_ij P_j(k + q)[c][b]
template <int N>
__global__ void __launch_bounds__(256) k_lindhard(double beta, long nz, const c2 *__restrict__ z, int d0, int d1,
                                                  int d2, const double *__restrict__ ev, const c2 *__restrict__ proj,
                                                  c2 *__restrict__ chi) {
  constexpr int N2 = N * N, N4 = N2 * N2;
  constexpr int NV = 256 / N4 > 64 ? 64 : 256 / N4;  // frequencies per workgroup
  constexpr int KT = 16;                           // k-points per tile
  __shared__ double sE[2][KT][N];
  __shared__ c2 sP[2][KT][N][N2];
  __shared__ c2 sF[NV][KT][N2];
  __shared__ c2 sZ[NV];
  const long nk = (long)d0 * d1 * d2;
  const long nvb = (nz + NV - 1) / NV;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long q = tile / nvb;
  const long v0 = (tile - q * nvb) * NV;
  const int t = threadIdx.x;
  const int v = t / N4, abcd = t - v * N4;
  const int a = abcd / (N * N2), b = (abcd / N2) % N, c = (abcd / N) % N, d = abcd % N;
  const bool live = v < NV && v0 + v < nz;
  if (t < NV) sZ[t] = (v0 + t < nz) ? z[v0 + t] : c2{0.0, 0.0};
  const long q0 = q / ((long)d1 * d2), q1 = (q / d2) % d1, q2 = q % d2;
  c2 acc = c2{0.0, 0.0};
  for (long kt = 0; kt < nk; kt += KT) {
    const int kn = (nk - kt) < KT ? (int)(nk - kt) : KT;
    __syncthreads();
    // stage eigenvalues and projectors of k and k + q
    for (int e = t; e < 2 * KT * N * N2; e += blockDim.x) {
      const int side = e / (KT * N * N2), r = e - side * (KT * N * N2);
      const int kk = r / (N * N2), w = r - kk * (N * N2);
      if (kk < kn) {
        long k = kt + kk;
        if (side) {
          const long k0 = k / ((long)d1 * d2), k1 = (k / d2) % d1, k2 = k % d2;
          k = (((k0 + q0) % d0) * d1 + (k1 + q1) % d1) * d2 + (k2 + q2) % d2;
        }
        sP[side][kk][w / N2][w % N2] = proj[k * N * N2 + w];
        if (w < N) sE[side][kk][w] = ev[k * N + w];
      }
    }
    __syncthreads();
    // F[v][kk][i][j] for the block's frequencies
    for (int e = t; e < NV * KT * N2; e += blockDim.x) {
      const int vv = e / (KT * N2), r = e - vv * (KT * N2);
      const int kk = r / N2, ij = r - kk * N2, i = ij / N, j = ij % N;
      if (kk < kn) {
        const double eki = sE[0][kk][i], ekqj = sE[1][kk][j];
        const double de = ekqj - eki;
        const double dn = fermi_fn(eki * beta) - fermi_fn(ekqj * beta);
        const c2 zz = sZ[vv];
        c2 F;
        if (sqrt(cabs2(zz)) < 1e-10 && fabs(de) < 1e-10) {
          const double ch = cosh(0.5 * beta * eki);
          F = c2{beta / (4.0 * ch * ch), 0.0};
        } else {
          F = cdiv(c2{dn, 0.0}, c2{zz.x + de, zz.y});
        }
        sF[vv][kk][ij] = F;
      }
    }
    __syncthreads();
    if (live) {
      for (int kk = 0; kk < kn; ++kk) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          c2 s = c2{0.0, 0.0};
#pragma unroll
          for (int j = 0; j < N; ++j) s = cfma(s, sF[v][kk][i * N + j], sP[1][kk][j][c * N + b]);
          acc = cfma(acc, sP[0][kk][i][a * N + d], s);
        }
      }
    }
  }
  if (live) chi[((v0 + v) * nk + q) * N4 + abcd] = cscale(acc, 1.0 / (double)nk);
}

// ------------------------------------------------------------------------
// 5 <= n <= 8.  Eigenpairs: the same Jacobi sweeps as k_eigh_proj, thread per
// k, with H and V in a private LDS region of 2 n^2 entries instead of
// registers.
constexpr int kEighWideThreads = 32;
__global__ void __launch_bounds__(kEighWideThreads) k_eigh_proj_wide(int n, long nk, double mu,
                                                                     const c2 *__restrict__ ek,
                                                                     double *__restrict__ ev, c2 *__restrict__ proj) {
  extern __shared__ c2 smw[];
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nk) return;
  c2 *H = smw + threadIdx.x * 2 * n * n, *V = H + n * n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      H[i * n + j] = ek[(k * n + i) * n + j];
      V[i * n + j] = c2{i == j ? 1.0 : 0.0, 0.0};
    }
  for (int i = 0; i < n; ++i) {
    H[i * n + i].x -= mu;
    H[i * n + i].y = 0.0;
  }
  for (int sweep = 0; sweep < 16; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < n; ++p) {
      diag += H[p * n + p].x * H[p * n + p].x;
      for (int q = p + 1; q < n; ++q) off += cabs2(H[p * n + q]);
    }
    if (off <= 1e-34 * diag || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double bb = sqrt(cabs2(H[p * n + q]));
        if (bb == 0.0) continue;
        const c2 ph = c2{H[p * n + q].x / bb, -H[p * n + q].y / bb};  // e^{-i phi}
        const double tau = (H[q * n + q].x - H[p * n + p].x) / (2.0 * bb);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), sn = t * c;
        const c2 g10 = cscale(ph, -sn), g11 = cscale(ph, c);
        for (int r = 0; r < n; ++r) {
          const c2 hp = H[r * n + p], hq = H[r * n + q];
          H[r * n + p] = cadd(cscale(hp, c), cmul(hq, g10));
          H[r * n + q] = cadd(cscale(hp, sn), cmul(hq, g11));
          const c2 vp = V[r * n + p], vq = V[r * n + q];
          V[r * n + p] = cadd(cscale(vp, c), cmul(vq, g10));
          V[r * n + q] = cadd(cscale(vp, sn), cmul(vq, g11));
        }
        for (int r = 0; r < n; ++r) {
          const c2 hp = H[p * n + r], hq = H[q * n + r];
          H[p * n + r] = cadd(cscale(hp, c), cmul(hq, cconj(g10)));
          H[q * n + r] = cadd(cscale(hp, sn), cmul(hq, cconj(g11)));
        }
        H[p * n + q] = c2{0.0, 0.0};
        H[q * n + p] = c2{0.0, 0.0};
        H[p * n + p].y = 0.0;
        H[q * n + q].y = 0.0;
      }
  }
  for (int i = 0; i < n; ++i) {
    ev[k * n + i] = H[i * n + i].x;
    for (int a = 0; a < n; ++a)
      for (int d = 0; d < n; ++d) proj[((k * n + i) * n + a) * n + d] = cmul(V[a * n + i], cconj(V[d * n + i]));
  }
}

// The k sum for 5 <= n <= 8: one workgroup per (q, frequency), thread t owning
// the outputs abcd = t + 256 s (s < 16); tiles of KT = 2 k-points of
// eigenvalues, projectors of k and k + q and the factors F_ij in LDS.
constexpr int kLhWideKT = 2;
__global__ void __launch_bounds__(256) k_lindhard_wide(int n, double beta, long nz, const c2 *__restrict__ z, int d0,
                                                       int d1, int d2, const double *__restrict__ ev,
                                                       const c2 *__restrict__ proj, c2 *__restrict__ chi) {
  extern __shared__ c2 smw[];
  constexpr int KT = kLhWideKT, NS = 16;
  const int n2 = n * n, n3 = n2 * n, n4 = n2 * n2;
  c2 *sP = smw;                        // [2][KT][n][n2]
  c2 *sF = sP + 2 * KT * n3;           // [KT][n2]
  double *sE = (double *)(sF + KT * n2);  // [2][KT][n]
  const long nk = (long)d0 * d1 * d2;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long q = tile / nz, v = tile - q * nz;
  const c2 zz = z[v];
  const long q0 = q / ((long)d1 * d2), q1 = (q / d2) % d1, q2 = q % d2;
  const int t = threadIdx.x;
  c2 acc[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) acc[s] = c2{0.0, 0.0};
  for (long kt = 0; kt < nk; kt += KT) {
    const int kn = (nk - kt) < KT ? (int)(nk - kt) : KT;
    __syncthreads();
    for (int e = t; e < 2 * KT * n3; e += blockDim.x) {
      const int side = e / (KT * n3), r = e - side * (KT * n3);
      const int kk = r / n3, w = r - kk * n3;
      if (kk < kn) {
        long k = kt + kk;
        if (side) {
          const long k0 = k / ((long)d1 * d2), k1 = (k / d2) % d1, k2 = k % d2;
          k = (((k0 + q0) % d0) * d1 + (k1 + q1) % d1) * d2 + (k2 + q2) % d2;
        }
        sP[e] = proj[k * n3 + w];
        if (w < n) sE[(side * KT + kk) * n + w] = ev[k * n + w];
      }
    }
    __syncthreads();
    for (int e = t; e < KT * n2; e += blockDim.x) {
      const int kk = e / n2, ij = e - kk * n2, i = ij / n, j = ij - i * n;
      if (kk < kn) {
        const double eki = sE[kk * n + i], ekqj = sE[(KT + kk) * n + j];
        const double de = ekqj - eki;
        const double dn = fermi_fn(eki * beta) - fermi_fn(ekqj * beta);
        c2 F;
        if (sqrt(cabs2(zz)) < 1e-10 && fabs(de) < 1e-10) {
          const double ch = cosh(0.5 * beta * eki);
          F = c2{beta / (4.0 * ch * ch), 0.0};
        } else {
          F = cdiv(c2{dn, 0.0}, c2{zz.x + de, zz.y});
        }
        sF[e] = F;
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const int abcd = t + 256 * s;
      if (abcd < n4) {
        const int a = abcd / n3, b = (abcd / n2) % n, c = (abcd / n) % n, d = abcd % n;
        for (int kk = 0; kk < kn; ++kk) {
          const c2 *P0 = sP + kk * n3, *P1 = sP + (KT + kk) * n3, *F = sF + kk * n2;
          for (int i = 0; i < n; ++i) {
            c2 sum = c2{0.0, 0.0};
            for (int j = 0; j < n; ++j) sum = cfma(sum, F[i * n + j], P1[j * n2 + c * n + b]);
            acc[s] = cfma(acc[s], P0[i * n2 + a * n + d], sum);
          }
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int abcd = t + 256 * s;
    if (abcd < n4) chi[(v * nk + q) * n4 + abcd] = cscale(acc[s], 1.0 / (double)nk);
  }
}

size_t lindhard_wide_lds(int n) {
  const int KT = kLhWideKT;
  return sizeof(c2) * (size_t)(2 * KT * n * n * n + KT * n * n) + sizeof(double) * 2 * KT * n;
}

// dynamic LDS of the eigen kernel: 2 n^2 entries per thread, 64 KiB at n = 8
hipError_t lindhard_set_attributes() {
  return hipFuncSetAttribute((const void *)k_eigh_proj_wide, hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(sizeof(c2) * 2 * 64 * kEighWideThreads));
}

hipError_t launch_lindhard(hipStream_t s, double beta, double mu, long nz, const c2 *z, const int dims[3], int n,
                           const c2 *ek, double *ev, c2 *proj, c2 *chi) {
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const long nbk = (nk + 255) / 256;
  switch (n) {
#define TPRF_LH_CASE(N)                                                                                  \
  case N: {                                                                                              \
    constexpr int N4 = N * N * N * N;                                                                    \
    constexpr int NV = 256 / N4 > 64 ? 64 : 256 / N4;                                                    \
    hipLaunchKernelGGL(k_eigh_proj<N>, dim3(nbk), dim3(256), 0, s, nk, mu, ek, ev, proj);              \
    const long nb = nk * ((nz + NV - 1) / NV);                                                           \
    hipLaunchKernelGGL(k_lindhard<N>, dim3(nb), dim3(256), 0, s, beta, nz, z, dims[0], dims[1], dims[2], \
                       ev, proj, chi);                                                                   \
    break;                                                                                               \
  }
    TPRF_LH_CASE(1)
    TPRF_LH_CASE(2)
    TPRF_LH_CASE(3)
    TPRF_LH_CASE(4)
#undef TPRF_LH_CASE
    case 5:
    case 6:
    case 7:
    case 8: {
      const long nbe = (nk + kEighWideThreads - 1) / kEighWideThreads;
      hipLaunchKernelGGL(k_eigh_proj_wide, dim3(nbe), dim3(kEighWideThreads),
                         sizeof(c2) * 2 * n * n * kEighWideThreads, s, n, nk, mu, ek, ev, proj);
      hipLaunchKernelGGL(k_lindhard_wide, dim3(nk * nz), dim3(256), lindhard_wide_lds(n), s, n, beta, nz, z, dims[0],
                         dims[1], dims[2], ev, proj, chi);
      break;
    }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// Device workspace of the no-allocation entry point tprf_lindhard_chi00_noalloc:
// one file-scope array per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy (on the host it is only a
// stand-in for the symbol); it uses the device address that
// hipGetSymbolAddress returns for the device that is current, cached here per
// device and kept by each context.  Every call on one GPU shares the array,
// so calls that overlap on one GPU race.
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

hipError_t noalloc_scratch_address(int device, void **ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *addr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices || !ptr) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess) return e;
    addr[device] = p;
  }
  *ptr = addr[device];
  return hipSuccess;
}

}  // namespace tprf
