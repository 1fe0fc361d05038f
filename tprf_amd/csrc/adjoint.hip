// adjoint.hip -- reverse-mode kernels of the static (nu = 0) bubble + RPA chain
// at one lattice momentum Q:
//   L = Re tr chi_RPA(Q, 0),  chi_RPA = (1 - chi0 U)^-1 chi0 (PH grouping),
//   chi0(Q, 0) = sum_r exp(i Q.r) chi0(r, 0),  chi0(r, 0) = sum_j c_j x(tau_j, r),
//   x_abcd(tau_j, r) = G_da(tau_j, r) G_bc(tau_{L-j}, -r).
// The weights c_j are the nu = 0 row of the tau -> nu map (real).  All
// adjoints follow dL = Re sum conj(xbar) dx, i.e. xbar = dL/dRe x + i dL/dIm x.
#include "common.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

// chi0(Q) = sum_r exp(2 pi i Q.r / dims) chi0(r): one workgroup per (target
// component, lattice of the batch), deterministic tree reduction over r.
__global__ void __launch_bounds__(256) k_q_project(const c2 *__restrict__ chi_r, c2 *__restrict__ chi_q, int d0, int d1,
                                                   int d2, int q0, int q1, int q2, int ntarget) {
  __shared__ c2 red[256];
  const int comp = blockIdx.x;
  const long nr = (long)d0 * d1 * d2;
  const c2 *src = chi_r + (long)blockIdx.y * nr * ntarget;
  c2 s = c2{0.0, 0.0};
  for (long r = threadIdx.x; r < nr; r += blockDim.x) {
    const long i2 = r % d2, i1 = (r / d2) % d1, i0 = r / ((long)d1 * d2);
    // phase fraction in [0, 1) computed exactly in integers per dimension
    const double f = (double)((q0 * i0) % d0) / d0 + (double)((q1 * i1) % d1) / d1 + (double)((q2 * i2) % d2) / d2;
    double sn, cs;
    sincospi(2.0 * f, &sn, &cs);
    s = cadd(s, cmul(c2{cs, sn}, src[r * ntarget + comp]));
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) chi_q[(long)blockIdx.y * ntarget + comp] = red[0];
}

// chibar(r) = exp(-2 pi i Q.r / dims) chibar(Q), per lattice of the batch
__global__ void __launch_bounds__(256) k_q_project_adjoint(const c2 *__restrict__ chibar_q, c2 *__restrict__ chibar_r,
                                                           int d0, int d1, int d2, int q0, int q1, int q2,
                                                           int ntarget, long total) {
  const long nr = (long)d0 * d1 * d2;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long rt = idx / ntarget;
  const int comp = (int)(idx - rt * ntarget);
  const long inst = rt / nr, r = rt - inst * nr;
  const long i2 = r % d2, i1 = (r / d2) % d1, i0 = r / ((long)d1 * d2);
  const double f = (double)((q0 * i0) % d0) / d0 + (double)((q1 * i1) % d1) / d1 + (double)((q2 * i2) % d2) / d2;
  double sn, cs;
  sincospi(2.0 * f, &sn, &cs);
  chibar_r[idx] = cmul(c2{cs, -sn}, chibar_q[inst * ntarget + comp]);
}

// Static bubble adjoint: gbar_xy(tau_j, r) =
//   c_j     sum_{b,c} chibar_{y b c x}(r)  conj(G_bc(tau_{L-j}, -r))
// + c_{L-j} sum_{a,d} chibar_{a x y d}(-r) conj(G_da(tau_{L-j}, -r)).
// Thread (r, x, y) of tau chunk blockIdx.y keeps its 2 n^2 chibar weights in
// registers and walks its chunk of tau; consecutive threads write consecutive
// gbar entries.  A batch of lattices (same dims) is folded into r:
// r = inst * nr + r_local.
template <int N>
__global__ void __launch_bounds__(256) k_bubble_static_adjoint(int ntau, int d0, int d1, int d2, long nbatch,
                                                               int jchunk, const c2 *__restrict__ gtr,
                                                               const c2 *__restrict__ chibar_r,
                                                               const double *__restrict__ cw, c2 *__restrict__ gbar) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nri = (long)d0 * d1 * d2;
  const long nr = nri * nbatch;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nr * n2) return;
  const long r = idx / n2;
  const int xy = (int)(idx - r * n2), x = xy / N, y = xy % N;
  const int L = ntau - 1;
  const long inst = r / nri;
  const long mr = inst * nri + minus_r(r - inst * nri, d0, d1, d2);
  const c2 *cb = chibar_r + r * n4, *cbm = chibar_r + mr * n4;
  c2 A[n2], B[n2];
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      A[b * N + c] = cb[((y * N + b) * N + c) * N + x];   // pairs with G_bc
      B[c * N + b] = cbm[((b * N + x) * N + y) * N + c];  // (a, d) = (b, c): pairs with G_cb
    }
  const int j0 = blockIdx.y * jchunk;
  const int j1 = (j0 + jchunk) < ntau ? (j0 + jchunk) : ntau;
  for (int j = j0; j < j1; ++j) {
    const c2 *gm = gtr + ((long)(L - j) * nr + mr) * n2;  // G(tau_{L-j}, -r)
    c2 s1 = c2{0.0, 0.0}, s2 = c2{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < n2; ++i) {
      const c2 g = cconj(gm[i]);
      s1 = cfma(s1, A[i], g);
      s2 = cfma(s2, B[i], g);
    }
    gbar[((long)j * nr + r) * n2 + xy] = cadd(cscale(s1, cw[j]), cscale(s2, cw[L - j]));
  }
}

// Dyson adjoint: Mbar = -G^H Gbar G^H for every (w, k); ebar_k = -sum_w Mbar,
// mubar = Re sum tr Mbar.  A batch of lattices is folded into k (k = inst * nk
// + k_local); blockIdx.x = inst * nbk + block of k_local, so no block straddles
// two lattices.  Thread (k, chunk) sums its chunk of frequencies into
// part_e[chunk][k][n^2]; per-block sums of Re tr Mbar go to
// mu_part[(inst * nchunk + chunk) * nbk + block].
template <int N>
__global__ void __launch_bounds__(256) k_dyson_adjoint(int nw, long nk, long nbatch, int nbk, int wchunk,
                                                       const c2 *__restrict__ g, const c2 *__restrict__ gbar,
                                                       c2 *__restrict__ part_e, double *__restrict__ mu_part) {
  __shared__ double red[256];
  const long inst = blockIdx.x / nbk;
  const int kb = (int)(blockIdx.x - inst * nbk);
  const long kl = (long)kb * blockDim.x + threadIdx.x;
  const long ktot = nk * nbatch;
  const long k = inst * nk + kl;
  const int w0 = blockIdx.y * wchunk;
  const int w1 = (w0 + wchunk) < nw ? (w0 + wchunk) : nw;
  double mu = 0.0;
  if (kl < nk) {
    c2 e[N][N];
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) e[a][b] = c2{0.0, 0.0};
    for (int w = w0; w < w1; ++w) {
      const c2 *G = g + ((long)w * ktot + k) * N * N;
      const c2 *B = gbar + ((long)w * ktot + k) * N * N;
      c2 Gv[N][N], Bv[N][N];
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = 0; b < N; ++b) { Gv[a][b] = G[a * N + b]; Bv[a][b] = B[a * N + b]; }
      // T = G^H Gbar ; S = T G^H = -Mbar
      c2 T[N][N];
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = 0; b < N; ++b) {
          c2 s = c2{0.0, 0.0};
#pragma unroll
          for (int l = 0; l < N; ++l) s = cfma(s, cconj(Gv[l][a]), Bv[l][b]);
          T[a][b] = s;
        }
#pragma unroll
      for (int a = 0; a < N; ++a)
#pragma unroll
        for (int b = 0; b < N; ++b) {
          c2 s = c2{0.0, 0.0};
#pragma unroll
          for (int l = 0; l < N; ++l) s = cfma(s, T[a][l], cconj(Gv[b][l]));
          e[a][b] = cadd(e[a][b], s);
          if (a == b) mu -= s.x;
        }
    }
    c2 *o = part_e + ((long)blockIdx.y * ktot + k) * N * N;
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int b = 0; b < N; ++b) o[a * N + b] = e[a][b];
  }
  red[threadIdx.x] = mu;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) mu_part[(inst * gridDim.y + blockIdx.y) * nbk + kb] = red[0];
}

// out[i] = sum_c part[c * stride + i], i < stride (fixed order: deterministic)
__global__ void __launch_bounds__(256) k_sum_chunks(const c2 *__restrict__ part, int nchunk, long stride,
                                                    c2 *__restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= stride) return;
  c2 s = c2{0.0, 0.0};
  for (int c = 0; c < nchunk; ++c) s = cadd(s, part[(long)c * stride + i]);
  out[i] = s;
}

// Static bubble: chi0_abcd(r) = sum_j c_j G_da(tau_j, r) G_bc(tau_{L-j}, -r).
// Thread (r, d, a) of tau chunk y accumulates the n^2 entries (b, c) into
// part[y][r][abcd].
template <int N>
__global__ void __launch_bounds__(256) k_bubble_static(int ntau, int d0, int d1, int d2, long nbatch, int jchunk,
                                                       const c2 *__restrict__ gtr, const double *__restrict__ cw,
                                                       c2 *__restrict__ part) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nri = (long)d0 * d1 * d2;
  const long nr = nri * nbatch;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nr * n2) return;
  const long r = idx / n2;
  const int da = (int)(idx - r * n2), d = da / N, a = da % N;
  const int L = ntau - 1;
  const long inst = r / nri;
  const long mr = inst * nri + minus_r(r - inst * nri, d0, d1, d2);
  const int j0 = blockIdx.y * jchunk;
  const int j1 = (j0 + jchunk) < ntau ? (j0 + jchunk) : ntau;
  c2 acc[n2];
#pragma unroll
  for (int i = 0; i < n2; ++i) acc[i] = c2{0.0, 0.0};
  for (int j = j0; j < j1; ++j) {
    const c2 x = cscale(gtr[((long)j * nr + r) * n2 + da], cw[j]);
    const c2 *gm = gtr + ((long)(L - j) * nr + mr) * n2;
#pragma unroll
    for (int bc = 0; bc < n2; ++bc) acc[bc] = cfma(acc[bc], x, gm[bc]);
  }
  c2 *o = part + ((long)blockIdx.y * nr + r) * n4;
#pragma unroll
  for (int b = 0; b < N; ++b)
#pragma unroll
    for (int c = 0; c < N; ++c) o[((a * N + b) * N + c) * N + d] = acc[b * N + c];
}

// out[b] = sum of part[b * n .. b * n + n) (one workgroup per segment, deterministic)
__global__ void __launch_bounds__(256) k_sum_partials(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double red[256];
  const double *p = part + (long)blockIdx.x * n;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// RPA at single points, forward and reverse, one thread per point (n <= 2):
// X = M(chi0), A = 1 - X U, chi = A^-1 X; with chibar = M(dL/dchi):
// Xbar = A^-H chibar (1 + U chi)^H, Ubar = X^H A^-H chibar chi^H.
template <int N>
__global__ void __launch_bounds__(64) k_rpa_point(long npts, const c2 *__restrict__ chi0, const c2 *__restrict__ U,
                                                  const c2 *__restrict__ chibar_t, c2 *__restrict__ chi,
                                                  c2 *__restrict__ xbar_t, c2 *__restrict__ ubar_t, double *status) {
  constexpr int M = N * N, n4 = M * M;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  auto tidx = [](int row, int col) { return row * M + col; };  // M[(a,b),(c,d)] <-> tensor abcd
  c2 X[M][M], MU[M][M], A[M][M], Ai[M][M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      X[i][jj] = chi0[p * n4 + tidx(i, jj)];
      MU[i][jj] = U[p * n4 + tidx(i, jj)];
    }
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = (i == jj) ? c2{1.0, 0.0} : c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfms(s, X[i][l], MU[l][jj]);
      A[i][jj] = s;
      Ai[i][jj] = (i == jj) ? c2{1.0, 0.0} : c2{0.0, 0.0};
    }
  // Gauss-Jordan inverse with partial pivoting
  bool singular = false;
#pragma unroll
  for (int k = 0; k < M; ++k) {
    int pv = k;
    double best = cabs2(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < M; ++i)
      if (cabs2(A[i][k]) > best) { best = cabs2(A[i][k]); pv = i; }
    if (best == 0.0) singular = true;
#pragma unroll
    for (int i = k + 1; i < M; ++i)
      if (pv == i) {
#pragma unroll
        for (int jj = 0; jj < M; ++jj) {
          c2 t = A[k][jj]; A[k][jj] = A[i][jj]; A[i][jj] = t;
          t = Ai[k][jj]; Ai[k][jj] = Ai[i][jj]; Ai[i][jj] = t;
        }
      }
    const c2 inv = cdiv(c2{1.0, 0.0}, A[k][k]);
#pragma unroll
    for (int jj = 0; jj < M; ++jj) { A[k][jj] = cmul(A[k][jj], inv); Ai[k][jj] = cmul(Ai[k][jj], inv); }
#pragma unroll
    for (int i = 0; i < M; ++i) {
      if (i == k) continue;
      const c2 f = A[i][k];
#pragma unroll
      for (int jj = 0; jj < M; ++jj) { A[i][jj] = cfms(A[i][jj], f, A[k][jj]); Ai[i][jj] = cfms(Ai[i][jj], f, Ai[k][jj]); }
    }
  }
  if (singular && status) status[2] = 1.0;
  c2 Y[M][M];  // chi
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, Ai[i][l], X[l][jj]);
      Y[i][jj] = s;
      chi[p * n4 + tidx(i, jj)] = s;
    }
  if (!chibar_t) return;
  c2 Cb[M][M], Z[M][M], W[M][M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) Cb[i][jj] = chibar_t[p * n4 + tidx(i, jj)];
  // Z = A^-H Cb
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, cconj(Ai[l][i]), Cb[l][jj]);
      Z[i][jj] = s;
    }
  // W = 1 + U chi
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = (i == jj) ? c2{1.0, 0.0} : c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, MU[i][l], Y[l][jj]);
      W[i][jj] = s;
    }
  // Xbar = Z W^H ; Ubar = X^H Z chi^H
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, Z[i][l], cconj(W[jj][l]));
      xbar_t[p * n4 + tidx(i, jj)] = s;
    }
  c2 T[M][M];
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, cconj(X[l][i]), Z[l][jj]);
      T[i][jj] = s;
    }
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int jj = 0; jj < M; ++jj) {
      c2 s = c2{0.0, 0.0};
#pragma unroll
      for (int l = 0; l < M; ++l) s = cfma(s, T[i][l], cconj(Y[jj][l]));
      ubar_t[p * n4 + tidx(i, jj)] = s;
    }
}

// RPA at single points for n = 3, 4 (M = n^2 = 9, 16 rows): one workgroup per
// point, the matrices in LDS, Gauss-Jordan with partial pivoting on
// [A | 1] (A = 1 - X U), then chi = A^-1 X and, with chibar given,
// Xbar = A^-H chibar (1 + U chi)^H, Ubar = X^H A^-H chibar chi^H, as k_rpa_point.
template <int M>
__global__ void __launch_bounds__(256) k_rpa_point_wg(const c2 *__restrict__ chi0, const c2 *__restrict__ U,
                                                      const c2 *__restrict__ chibar_t, c2 *__restrict__ chi,
                                                      c2 *__restrict__ xbar_t, c2 *__restrict__ ubar_t,
                                                      double *status) {
  constexpr int MM = M * M;
  __shared__ c2 X[M][M], MU[M][M], A[M][M], Ai[M][M], Y[M][M], T1[M][M], T2[M][M];
  __shared__ c2 fcol[M], pinv_s;
  __shared__ int piv;
  const int tid = threadIdx.x;
  const long p = blockIdx.x;
  const c2 *x = chi0 + p * MM, *u = U + p * MM;
  for (int t = tid; t < MM; t += blockDim.x) {
    X[t / M][t % M] = x[t];
    MU[t / M][t % M] = u[t];
  }
  __syncthreads();
  for (int t = tid; t < MM; t += blockDim.x) {
    const int i = t / M, j = t % M;
    c2 s = c2{i == j ? 1.0 : 0.0, 0.0};
    for (int l = 0; l < M; ++l) s = cfms(s, X[i][l], MU[l][j]);
    A[i][j] = s;
    Ai[i][j] = c2{i == j ? 1.0 : 0.0, 0.0};
  }
  __syncthreads();
  bool singular = false;
  for (int k = 0; k < M; ++k) {
    if (tid == 0) {
      int pv = k;
      double best = cabs2(A[k][k]);
      for (int i = k + 1; i < M; ++i)
        if (cabs2(A[i][k]) > best) { best = cabs2(A[i][k]); pv = i; }
      piv = pv;
      singular = singular || best == 0.0;
    }
    __syncthreads();
    const int pv = piv;
    if (pv != k)
      for (int t = tid; t < 2 * M; t += blockDim.x) {
        c2 *ra = t < M ? &A[k][t] : &Ai[k][t - M], *rb = t < M ? &A[pv][t] : &Ai[pv][t - M];
        const c2 tmp = *ra;
        *ra = *rb;
        *rb = tmp;
      }
    __syncthreads();
    if (tid == 0) pinv_s = cdiv(c2{1.0, 0.0}, A[k][k]);
    __syncthreads();
    for (int t = tid; t < 2 * M; t += blockDim.x) {
      c2 *r = t < M ? &A[k][t] : &Ai[k][t - M];
      *r = cmul(*r, pinv_s);
    }
    __syncthreads();
    for (int i = tid; i < M; i += blockDim.x) fcol[i] = A[i][k];
    __syncthreads();
    for (int t = tid; t < M * 2 * M; t += blockDim.x) {
      const int i = t / (2 * M), j = t % (2 * M);
      if (i == k) continue;
      if (j < M) A[i][j] = cfms(A[i][j], fcol[i], A[k][j]);
      else Ai[i][j - M] = cfms(Ai[i][j - M], fcol[i], Ai[k][j - M]);
    }
    __syncthreads();
  }
  if (tid == 0 && singular && status) status[2] = 1.0;
  // Y = chi = Ai X
  for (int t = tid; t < MM; t += blockDim.x) {
    const int i = t / M, j = t % M;
    c2 s = c2{0, 0};
    for (int l = 0; l < M; ++l) s = cfma(s, Ai[i][l], X[l][j]);
    Y[i][j] = s;
    chi[p * MM + t] = s;
  }
  if (!chibar_t) return;  // uniform across the workgroup
  __syncthreads();
  const c2 *cb = chibar_t + p * MM;
  // T1 = Z = Ai^H chibar ; T2 = W = 1 + U chi
  for (int t = tid; t < MM; t += blockDim.x) {
    const int i = t / M, j = t % M;
    c2 z = c2{0, 0}, w = c2{i == j ? 1.0 : 0.0, 0.0};
    for (int l = 0; l < M; ++l) {
      z = cfma(z, cconj(Ai[l][i]), cb[l * M + j]);
      w = cfma(w, MU[i][l], Y[l][j]);
    }
    T1[i][j] = z;
    T2[i][j] = w;
  }
  __syncthreads();
  // Xbar = Z W^H ; A <- X^H Z (A is free now)
  for (int t = tid; t < MM; t += blockDim.x) {
    const int i = t / M, j = t % M;
    c2 xb = c2{0, 0}, xz = c2{0, 0};
    for (int l = 0; l < M; ++l) {
      xb = cfma(xb, T1[i][l], cconj(T2[j][l]));
      xz = cfma(xz, cconj(X[l][i]), T1[l][j]);
    }
    xbar_t[p * MM + t] = xb;
    A[i][j] = xz;
  }
  __syncthreads();
  // Ubar = (X^H Z) chi^H
  for (int t = tid; t < MM; t += blockDim.x) {
    const int i = t / M, j = t % M;
    c2 s = c2{0, 0};
    for (int l = 0; l < M; ++l) s = cfma(s, A[i][l], cconj(Y[j][l]));
    ubar_t[p * MM + t] = s;
  }
}

// ---- launchers ---------------------------------------------------------
hipError_t launch_q_project(hipStream_t s, const c2 *chi_r, c2 *chi_q, const int dims[3], const int q[3], int ntarget,
                            long nbatch) {
  hipLaunchKernelGGL(k_q_project, dim3(ntarget, nbatch), dim3(256), 0, s, chi_r, chi_q, dims[0], dims[1], dims[2],
                     q[0], q[1], q[2], ntarget);
  return hipGetLastError();
}

hipError_t launch_q_project_adjoint(hipStream_t s, const c2 *chibar_q, c2 *chibar_r, const int dims[3], const int q[3],
                                    int ntarget, long nbatch) {
  const long total = (long)dims[0] * dims[1] * dims[2] * ntarget * nbatch;
  hipLaunchKernelGGL(k_q_project_adjoint, dim3((total + 255) / 256), dim3(256), 0, s, chibar_q, chibar_r, dims[0],
                     dims[1], dims[2], q[0], q[1], q[2], ntarget, total);
  return hipGetLastError();
}

hipError_t launch_bubble_static_adjoint(hipStream_t s, int ntau, const int dims[3], int n, long nbatch, const c2 *gtr,
                                        const c2 *chibar_r, const double *cw, c2 *gbar) {
  const long nthr = (long)dims[0] * dims[1] * dims[2] * nbatch * n * n;
  const long nb = (nthr + 255) / 256;
  // tau chunks: about 2^20 threads in flight, each chunk at least 16 tau points
  long nchunk = (1L << 20) / nthr;
  if (nchunk < 1) nchunk = 1;
  if (nchunk > ntau / 16) nchunk = ntau / 16 > 0 ? ntau / 16 : 1;
  const int jchunk = (int)((ntau + nchunk - 1) / nchunk);
  nchunk = (ntau + jchunk - 1) / jchunk;
  const dim3 grid(nb, nchunk);
#define TPRF_SA(N) hipLaunchKernelGGL(k_bubble_static_adjoint<N>, grid, dim3(256), 0, s, ntau, dims[0], dims[1], \
                                      dims[2], nbatch, jchunk, gtr, chibar_r, cw, gbar)
  switch (n) {
    case 1: TPRF_SA(1); break;
    case 2: TPRF_SA(2); break;
    case 3: TPRF_SA(3); break;
    case 4: TPRF_SA(4); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_SA
  return hipGetLastError();
}

hipError_t launch_dyson_adjoint(hipStream_t s, int nw, long nk, int n, long nbatch, const c2 *g, const c2 *gbar,
                                c2 *ebar, c2 *part_e, double *mu_part, double *mubar, int nchunk) {
  const int nbk = (int)((nk + 255) / 256);
  const int wchunk = (nw + nchunk - 1) / nchunk;
  const dim3 grid(nbk * nbatch, nchunk);
#define TPRF_DA(N) hipLaunchKernelGGL(k_dyson_adjoint<N>, grid, dim3(256), 0, s, nw, nk, nbatch, nbk, wchunk, g, gbar, \
                                      part_e, mu_part)
  switch (n) {
    case 1: TPRF_DA(1); break;
    case 2: TPRF_DA(2); break;
    case 3: TPRF_DA(3); break;
    case 4: TPRF_DA(4); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_DA
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long stride = nk * nbatch * n * n;
  hipLaunchKernelGGL(k_sum_chunks, dim3((stride + 255) / 256), dim3(256), 0, s, part_e, nchunk, stride, ebar);
  hipLaunchKernelGGL(k_sum_partials, dim3(nbatch), dim3(256), 0, s, mu_part, nbk * nchunk, mubar);
  return hipGetLastError();
}

hipError_t launch_sum_partials(hipStream_t s, const double *part, long nbatch, int n, double *out) {
  hipLaunchKernelGGL(k_sum_partials, dim3((unsigned)nbatch), dim3(256), 0, s, part, n, out);
  return hipGetLastError();
}

hipError_t launch_bubble_static(hipStream_t s, int ntau, const int dims[3], int n, long nbatch, const c2 *gtr,
                                const double *cw, c2 *part, int nchunk, c2 *chi_r) {
  const long nr = (long)dims[0] * dims[1] * dims[2] * nbatch;
  const long nb = (nr * n * n + 255) / 256;
  const int jchunk = (ntau + nchunk - 1) / nchunk;
  const dim3 grid(nb, nchunk);
#define TPRF_SF(N) hipLaunchKernelGGL(k_bubble_static<N>, grid, dim3(256), 0, s, ntau, dims[0], dims[1], dims[2], \
                                      nbatch, jchunk, gtr, cw, part)
  switch (n) {
    case 1: TPRF_SF(1); break;
    case 2: TPRF_SF(2); break;
    case 3: TPRF_SF(3); break;
    case 4: TPRF_SF(4); break;
    default: return hipErrorInvalidValue;
  }
#undef TPRF_SF
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long stride = nr * n * n * n * n;
  hipLaunchKernelGGL(k_sum_chunks, dim3((stride + 255) / 256), dim3(256), 0, s, part, nchunk, stride, chi_r);
  return hipGetLastError();
}

hipError_t launch_rpa_point(hipStream_t s, long npts, int n, const c2 *chi0, const c2 *U, const c2 *chibar, c2 *chi,
                            c2 *xbar, c2 *ubar, Status st) {
  const long nb = (npts + 63) / 64;
  switch (n) {
    case 1: hipLaunchKernelGGL(k_rpa_point<1>, dim3(nb), dim3(64), 0, s, npts, chi0, U, chibar, chi, xbar, ubar, st.d); break;
    case 2: hipLaunchKernelGGL(k_rpa_point<2>, dim3(nb), dim3(64), 0, s, npts, chi0, U, chibar, chi, xbar, ubar, st.d); break;
    case 3: hipLaunchKernelGGL(k_rpa_point_wg<9>, dim3(npts), dim3(128), 0, s, chi0, U, chibar, chi, xbar, ubar, st.d); break;
    case 4: hipLaunchKernelGGL(k_rpa_point_wg<16>, dim3(npts), dim3(256), 0, s, chi0, U, chibar, chi, xbar, ubar, st.d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace tprf
