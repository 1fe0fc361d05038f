// static_q.hip -- chi0(Q, nu = 0) of the whole static chain straight from e_k,
// and its adjoint with respect to e_k and mu.
//
// The chain lattice_dyson_g_wk (gf.cpp:91-132) -> fourier_wk_to_wr
// (fourier_lattice.cpp:39,49) -> fourier_wr_to_tr with the fitted tail
// (fourier_matsubara.cpp:218-308) -> chi0_wr_from_grt_PH at nu = 0
// (chi_imtime.cpp:139-182) -> one momentum of chi_wk_from_chi_wr
// (fourier_lattice.cpp:55-58) is, per r, the Matsubara bilinear form
// B(g1, g2) = -(1/beta) g1 . g2 + corr(q(g1), q(g2)) of G_da(r) and G_bc(-r)
// (static_mats.hip).  Every projection q is linear per column, so it commutes
// with the lattice transform, and sum_r e^{iQr} x(r) y(-r) with
// x(r) = (1/N) sum_k e^{-ikr} x(k) is (1/N) sum_k x(k) y(k - Q):
//   chi0_abcd(Q) = (1/N_k) sum_k B(G_da(i w, k), G_bc(i w, k - Q)).
// The tail fit runs on the k columns (its residual and zeroth moment are
// reported as for the r columns of fourier_wr_to_tr).
//
// G(i w, k) is never stored.  Per k the kernels take the characteristic
// polynomial p(z) = det(z - A) of A = e_k - s 1 (s = Re tr e_k / N, which
// centres the spectrum) and the adjugate coefficients adj(z - A) =
// z^{N-1} 1 + sum_{j < N-1} z^j B_j (Faddeev-LeVerrier), so that
//   G_xy(i w, k) = num_xy(z) / p(z),  z = i w + mu - s,
// costs two Horner evaluations of degree N per element.
#include "common.hpp"
#include "mats_device.hpp"
#include "static_device.hpp"

namespace tprf {

namespace {

constexpr int kTabW = kStaticTabWidth;

template <int N>
struct Resolvent {
  double s;                          // shift Re tr e / N
  c2 c[N];                           // p(z) = z^N + sum_j c_j z^j
  c2 B[N > 1 ? N - 1 : 1][N][N];     // adjugate coefficients of z^j, j < N - 1
};

// Faddeev-LeVerrier on A = e - s 1: M_1 = 1, c_{N-k} = -tr(A M_k) / k,
// M_{k+1} = A M_k + c_{N-k} 1; adj(z - A) = sum_k z^{N-k} M_k, so B_j = M_{N-j}.
template <int N>
__device__ __forceinline__ void resolvent_make(const c2 *__restrict__ ek, Resolvent<N> &R) {
  c2 A[N][N];
  double tr = 0.0;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) A[i][j] = ek[i * N + j];
#pragma unroll
  for (int i = 0; i < N; ++i) tr += A[i][i].x;
  R.s = tr / N;
#pragma unroll
  for (int i = 0; i < N; ++i) A[i][i].x -= R.s;
  c2 M[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) M[i][j] = c2{i == j ? 1.0 : 0.0, 0.0};
#pragma unroll
  for (int k = 1; k <= N; ++k) {
    c2 AM[N][N];
    c2 t = c2{0, 0};
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int j = 0; j < N; ++j) {
        c2 v = c2{0, 0};
#pragma unroll
        for (int l = 0; l < N; ++l) v = cfma(v, A[i][l], M[l][j]);
        AM[i][j] = v;
      }
#pragma unroll
    for (int i = 0; i < N; ++i) t = cadd(t, AM[i][i]);
    const c2 ck = cscale(t, -1.0 / k);
    R.c[N - k] = ck;
    if (k < N) {
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) {
          M[i][j] = AM[i][j];
          if (i == j) M[i][j] = cadd(M[i][j], ck);
        }
      // M is now M_{k+1}: the coefficient of z^{N-k-1}
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) R.B[N - k - 1][i][j] = M[i][j];
    }
  }
}

// the N coefficients (low to high) of num_xy: B_j[x][y], then delta_xy
template <int N>
__device__ __forceinline__ void resolvent_elem(const Resolvent<N> &R, int x, int y, c2 (&a)[N]) {
#pragma unroll
  for (int j = 0; j + 1 < N; ++j) {
    c2 v = c2{0, 0};
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int k = 0; k < N; ++k)
        if (i == x && k == y) v = R.B[j][i][k];
    a[j] = v;
  }
  a[N - 1] = c2{x == y ? 1.0 : 0.0, 0.0};
}

// sum_j a_j z^j (N coefficients)
template <int N>
__device__ __forceinline__ c2 horner(const c2 (&a)[N], c2 z) {
  c2 v = a[N - 1];
#pragma unroll
  for (int j = N - 2; j >= 0; --j) v = cfma(a[j], v, z);
  return v;
}

// p(z) (monic)
template <int N>
__device__ __forceinline__ c2 charpoly(const c2 (&c)[N], c2 z) {
  c2 v = c2{1.0, 0.0};
#pragma unroll
  for (int j = N - 1; j >= 0; --j) v = cfma(c[j], v, z);
  return v;
}

// 1 / b: the reciprocal of |b|^2 by v_rcp_f64 refined with two Newton steps
// (each squares the relative error, leaving rounding), instead of an IEEE
// division; |b|^2 = |p(i w + mu - s)|^2 stays in the normal range here
__device__ __forceinline__ c2 crecip(c2 b) {
  const double d = fma(b.x, b.x, b.y * b.y);
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return c2{b.x * r, -b.y * r};
}

// k - Q (sgn = -1) and k + Q (sgn = +1) on the lattice (i0 d1 + i1) d2 + i2
__device__ __forceinline__ long shift_k(long k, int d0, int d1, int d2, int q0, int q1, int q2, int sgn) {
  const long i2 = k % d2, i1 = (k / d2) % d1, i0 = k / ((long)d1 * d2);
  const long j0 = ((i0 + sgn * q0) % d0 + d0) % d0, j1 = ((i1 + sgn * q1) % d1 + d1) % d1,
             j2 = ((i2 + sgn * q2) % d2 + d2) % d2;
  return (j0 * d1 + j1) * d2 + j2;
}

__device__ __forceinline__ double wn(int n, double beta) { return kPi * (double)(2 * n + 1) / beta; }

// one element G_xy(z) of a k point: coefficients and the characteristic polynomial
template <int N>
struct Elem {
  c2 a[N], c[N];
  double zr;  // Re z = mu - s
  __device__ __forceinline__ void make(const c2 *ek, double mu, int x, int y) {
    Resolvent<N> R;
    resolvent_make<N>(ek, R);
#pragma unroll
    for (int j = 0; j < N; ++j) c[j] = R.c[j];
    resolvent_elem<N>(R, x, y, a);
    zr = mu - R.s;
  }
  __device__ __forceinline__ c2 at(double w) const {
    const c2 z = c2{zr, w};
    return cmul(horner<N>(a, z), crecip(charpoly<N>(c, z)));
  }
};

}  // namespace

// F1. Partial projections of every k column (inst, k, x, y) of G over the
//     frequency chunk blockIdx.y (n >= 0): part[sk][33][ncol] as in
//     static_mats.hip, or with one chunk the slots 0..24 of q themselves.
//     mu[nbatch], e_k[nbatch][nk][N][N].
template <int N>
__global__ void __launch_bounds__(256) k_sq_proj_part(long nbatch, long nk, int n_iw, int chunk, double beta,
                                                      const double *__restrict__ mu, const c2 *__restrict__ ek,
                                                      const double *__restrict__ tab, c2 *__restrict__ part,
                                                      c2 *__restrict__ q) {
  constexpr int n2 = N * N;
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int xy = (int)(col % n2), x = xy / N, y = xy - x * N;
  const long ik = col / n2, inst = ik / nk;
  Elem<N> E;
  E.make(ek + ik * n2, mu[inst], x, y);
  const int n0 = blockIdx.y * chunk, n1 = min(n_iw, n0 + chunk);
  c2 C[9], S[8], YA[4], YB[4], PA[4], PB[4];
#pragma unroll
  for (int i = 0; i < 9; ++i) C[i] = c2{0, 0};
#pragma unroll
  for (int i = 0; i < 8; ++i) S[i] = c2{0, 0};
#pragma unroll
  for (int i = 0; i < 4; ++i) YA[i] = YB[i] = PA[i] = PB[i] = c2{0, 0};
  for (int n = n0; n < n1; ++n) {
    const double w = wn(n, beta);
    const c2 a = E.at(w), b = E.at(-w);
    const c2 sg = cadd(a, b), dl = csub(a, b);
    const double *t = tab + (long)n * kTabW;
#pragma unroll
    for (int j = 0; j < 9; ++j) C[j] = cfma_r(C[j], t[j], sg);
#pragma unroll
    for (int j = 0; j < 8; ++j) S[j] = cfma_r(S[j], t[9 + j], dl);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      YA[i] = cfma_r(YA[i], t[17 + i], sg);
      YB[i] = cfma_r(YB[i], t[21 + i], dl);
      PA[i] = cfma_r(PA[i], t[25 + i], sg);
      PB[i] = cfma_r(PB[i], t[29 + i], dl);
    }
  }
  if (gridDim.y == 1) {  // one chunk: the projections themselves
    c2 *o = q + col;
#pragma unroll
    for (int j = 0; j < 9; ++j) o[(long)j * ncol] = C[j];
#pragma unroll
    for (int j = 0; j < 8; ++j) o[(long)(9 + j) * ncol] = S[j];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      o[(long)(kQY + i) * ncol] = cadd(YA[i], cmul_i(YB[i]));
      o[(long)(kQYp + i) * ncol] = cadd(PA[i], cmul_i(PB[i]));
    }
  } else {
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
}

// F1b. The tail fit of every k column.  A block holds 64 columns (lanes) and
//      kFitG waves; wave g evaluates G at the fitted frequencies f = g, g +
//      kFitG, ... (the same f across the wave, so the fit operator is read by
//      scalar loads) and applies its share of the least-squares operator.  The
//      partial coefficients meet in LDS (added in wave order), and each wave
//      takes the residual on the values it kept.  Writes the moment slots
//      q[kQM..kQM+3] and the status words.
constexpr int kFitG = 4;
template <int N>
__global__ void __launch_bounds__(64 * kFitG) k_sq_fit(long nbatch, long nk, int n_iw, double beta,
                                                       const double *__restrict__ mu, const c2 *__restrict__ ek,
                                                       TailFitDev fit, c2 *__restrict__ q, double *status) {
  constexpr int n2 = N * N, FP = (kMaxFit + kFitG - 1) / kFitG;
  __shared__ c2 sc[kFitG][kMaxOrder][64];
  __shared__ double sr[kFitG][64];
  // g is the wave's index: uniform across it (readfirstlane makes that known)
  const int lane = threadIdx.x & 63, g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * 64 + lane;
  const bool live = col < ncol;
  c2 v[FP], pc[kMaxOrder];
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k) pc[k] = c2{0, 0};
#pragma unroll
  for (int i = 0; i < FP; ++i) v[i] = c2{0, 0};
  Elem<N> E;
  if (live) {
    const int xy = (int)(col % n2), x = xy / N, y = xy - x * N;
    const long ik = col / n2;
    E.make(ek + ik * n2, mu[ik / nk], x, y);
  }
#pragma unroll
  for (int i = 0; i < FP; ++i) {
    const int f = g + i * kFitG;
    if (f < fit.n_fit) {
      if (live) v[i] = E.at(wn(fit.idx[f] - n_iw, beta));
#pragma unroll
      for (int k = 0; k < kMaxOrder; ++k)
        if (k < fit.ncoef) pc[k] = cfma(pc[k], fit.pinv[(long)k * fit.n_fit + f], v[i]);
    }
  }
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k) sc[g][k][lane] = pc[k];
  __syncthreads();
  c2 coef[kMaxOrder];
#pragma unroll
  for (int k = 0; k < kMaxOrder; ++k) {
    c2 t = c2{0, 0};
#pragma unroll
    for (int h = 0; h < kFitG; ++h) t = cadd(t, sc[h][k][lane]);
    coef[k] = t;
  }
  double r2 = 0.0;
#pragma unroll
  for (int i = 0; i < FP; ++i) {
    const int f = g + i * kFitG;
    if (f < fit.n_fit) {
      c2 d = v[i];
#pragma unroll
      for (int k = 0; k < kMaxOrder; ++k)
        if (k < fit.ncoef) d = cfms(d, fit.vander[(long)f * fit.ncoef + k], coef[k]);
      r2 += cabs2(d);
    }
  }
  sr[g][lane] = r2;
  __syncthreads();
  if (!live || g != 0) return;
  double rt = 0.0;
#pragma unroll
  for (int h = 0; h < kFitG; ++h) rt += sr[h][lane];
  atomic_max_nonneg(&status[0], sqrt(rt / fit.n_fit));
  atomic_max_nonneg(&status[1], sqrt(cabs2(coef[0])));
  const c2 m1 = cscale(coef[1], fit.wpow[1]);
  c2 a[3];
  Poles<0>::coef(m1, cscale(coef[2], fit.wpow[2]), cscale(coef[3], fit.wpow[3]), a);
  c2 *o = q + col;
#pragma unroll
  for (int i = 0; i < 3; ++i) o[(long)(kQM + i) * ncol] = a[i];
  o[(long)(kQM + 3) * ncol] = m1;
}

// F2. With several chunks: q slot blockIdx.y (< 25) of every column from the
//     chunk partials, added in chunk order (y and y' combine two partial
//     slots: sum over cos-rows + i times sum over sin-rows).
__global__ void __launch_bounds__(256) k_sq_proj_sum(long ncol, int sk, const c2 *__restrict__ part,
                                                     c2 *__restrict__ q) {
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int s = blockIdx.y;
  const int s1 = s < 17 ? s : s < 21 ? s : 25 + (s - 21);
  const int s2 = s < 17 ? -1 : s < 21 ? 21 + (s - 17) : 29 + (s - 21);
  c2 v1 = c2{0, 0}, v2 = c2{0, 0};
  for (int k = 0; k < sk; ++k) {
    v1 = cadd(v1, part[((long)k * 33 + s1) * ncol + col]);
    if (s2 >= 0) v2 = cadd(v2, part[((long)k * 33 + s2) * ncol + col]);
  }
  q[(long)s * ncol + col] = s2 < 0 ? v1 : cadd(v1, cmul_i(v2));
}

// F3. Thread (k, d) of one instance, mesh chunk blockIdx.y:
//     acc[a][bc] = -(1/beta) sum_m G_da(i w_m, k) G_bc(i w_m, k - Q), plus (chunk
//     0) corr(q(G_da(k)), q(G_bc(k - Q))); the block's sum of every output
//     abcd goes to part[inst][chunk][block][n^4].
template <int N, int BT>
__global__ void __launch_bounds__(BT) k_sq_pair(long nbatch, long nk, int d0, int d1, int d2, int q0, int q1, int q2,
                                                int n_iw, int chunk, int nblk, double beta,
                                                const double *__restrict__ mu, const c2 *__restrict__ ek,
                                                const c2 *__restrict__ q, StaticMatsDev sm, c2 *__restrict__ part) {
  constexpr int n2 = N * N, n4 = n2 * n2, N3 = N * n2;
  __shared__ c2 red[BT][N3];
  const int tid = threadIdx.x;
  const long inst = blockIdx.x / nblk, blk = blockIdx.x - inst * nblk;
  const long tk = blk * BT + tid;  // (k, d) within the instance
  const long k = tk / N;
  const int d = (int)(tk - k * N);
  const bool live = k < nk;
  c2 acc[N][n2];
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int s = 0; s < n2; ++s) acc[a][s] = c2{0, 0};
  long kq = 0;
  if (live) {
    kq = shift_k(k, d0, d1, d2, q0, q1, q2, -1);
    const double m = mu[inst];
    c2 r1[N][N], c1[N];  // row d of G(k): coefficients of num_da
    double z1;
    {
      Resolvent<N> R;
      resolvent_make<N>(ek + (inst * nk + k) * n2, R);
#pragma unroll
      for (int a = 0; a < N; ++a) resolvent_elem<N>(R, d, a, r1[a]);
#pragma unroll
      for (int j = 0; j < N; ++j) c1[j] = R.c[j];
      z1 = m - R.s;
    }
    Resolvent<N> R2;
    resolvent_make<N>(ek + (inst * nk + kq) * n2, R2);
    const double z2 = m - R2.s;
    const int nw = 2 * n_iw;
    const int m0 = blockIdx.y * chunk, m1 = min(nw, m0 + chunk);
    for (int mm = m0; mm < m1; ++mm) {
      const double w = wn(mm - n_iw, beta);
      const c2 za = c2{z1, w}, zb = c2{z2, w};
      // 1 / (p1 p2) once, folded into the row of G(k)
      const c2 inv = crecip(cmul(charpoly<N>(c1, za), charpoly<N>(R2.c, zb)));
      c2 g1[N];
#pragma unroll
      for (int a = 0; a < N; ++a) g1[a] = cmul(horner<N>(r1[a], za), inv);
      // numerators of G(k - Q): z^{N-1} delta + sum_j z^j B_j
      c2 zp[N];
      zp[0] = c2{1.0, 0.0};
#pragma unroll
      for (int j = 1; j < N; ++j) zp[j] = cmul(zp[j - 1], zb);
#pragma unroll
      for (int s = 0; s < n2; ++s) {
        const int b = s / N, c = s % N;
        c2 num = b == c ? zp[N - 1] : c2{0, 0};
#pragma unroll
        for (int j = 0; j + 1 < N; ++j) num = cfma(num, R2.B[j][b][c], zp[j]);
#pragma unroll
        for (int a = 0; a < N; ++a) acc[a][s] = cfma(acc[a][s], g1[a], num);
      }
    }
#pragma unroll
    for (int a = 0; a < N; ++a)
#pragma unroll
      for (int s = 0; s < n2; ++s) acc[a][s] = cscale(acc[a][s], -sm.inv_beta);
  }
#pragma unroll
  for (int a = 0; a < N; ++a)
#pragma unroll
    for (int s = 0; s < n2; ++s) red[tid][a * n2 + s] = acc[a][s];
  if (live && blockIdx.y == 0) {
    const long ncol = nbatch * nk * n2;
    for (int u = 0; u < N3; ++u) {
      const int a = u / n2, s = u - a * n2;
      red[tid][u] = cadd(red[tid][u],
                         corr_value(q + (inst * nk + k) * n2 + d * N + a, q + (inst * nk + kq) * n2 + s, ncol, sm));
    }
  }
  __syncthreads();
  // output abcd = ((a N + b) N + c) N + d: the threads of the block whose d matches
  for (int o = tid; o < n4; o += BT) {
    const int dd = o % N, a = o / (n2 * N), s = (o / N) % n2;
    const int t0 = (int)(((dd - (blk * BT) % N) % N + N) % N);  // first tid with d == dd
    c2 v = c2{0, 0};
    for (int t = t0; t < BT; t += N) v = cadd(v, red[t][a * n2 + s]);
    part[((inst * gridDim.y + blockIdx.y) * nblk + blk) * n4 + o] = v;
  }
}

// F4. chi0_q[inst][abcd] = (1/N_k) sum over chunks and blocks: one wave per
//     output, a fixed-order tree over the wave's partial sums.
__global__ void __launch_bounds__(64) k_sq_pair_final(int n4, int nparts, double inv_nk, const c2 *__restrict__ part,
                                                      c2 *__restrict__ chi0_q) {
  __shared__ c2 red[64];
  const long inst = blockIdx.x / n4;
  const int o = (int)(blockIdx.x - inst * n4), tid = threadIdx.x;
  c2 v = c2{0, 0};
  for (int p = tid; p < nparts; p += 64) v = cadd(v, part[(inst * nparts + p) * n4 + o]);
  red[tid] = v;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (tid < w) red[tid] = cadd(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) chi0_q[inst * n4 + o] = cscale(red[0], inv_nk);
}

// A1. qbar[29][ncol] of every k column (x, y): role 1 in the outputs
//     (a, b, c, d) = (y, s1, s2, x) with the partner column s at k - Q, role 2
//     in (s2, x, y, s1) with the partner s at k + Q; weights chibar / N_k.
template <int N>
__global__ void __launch_bounds__(256) k_sq_qbar(long nbatch, long nk, int d0, int d1, int d2, int q0, int q1, int q2,
                                                 double inv_nk, const c2 *__restrict__ q,
                                                 const c2 *__restrict__ chibar, StaticMatsDev sm,
                                                 c2 *__restrict__ qbar) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int xy = (int)(col % n2), x = xy / N, y = xy - x * N;
  const long ik = col / n2, inst = ik / nk, k = ik - inst * nk;
  const long km = shift_k(k, d0, d1, d2, q0, q1, q2, -1), kp = shift_k(k, d0, d1, d2, q0, q1, q2, +1);
  const c2 *cb = chibar + inst * n4;
  c2 qb[kQSlots];
#pragma unroll
  for (int i = 0; i < kQSlots; ++i) qb[i] = c2{0, 0};
#pragma unroll 1
  for (int s = 0; s < n2; ++s) {
    const int s1 = s / N, s2 = s - s1 * N;
    qbar_role1(qb, cscale(cb[((y * N + s1) * N + s2) * N + x], inv_nk), q + (inst * nk + km) * n2 + s, ncol, sm);
    qbar_role2(qb, cscale(cb[((s2 * N + x) * N + y) * N + s1], inv_nk), q + (inst * nk + kp) * n2 + s, ncol, sm);
  }
#pragma unroll
  for (int i = 0; i < kQSlots; ++i) qbar[(long)i * ncol + col] = qb[i];
}

// A2. For every k column (x, y) and frequency chunk (n >= 0 with its partner
//     -n-1): gbar_xy(i w, k) = -(1/(beta N_k)) sum_s (chibar(y,s1,s2,x)
//     conj G_s(k - Q) + chibar(s2,x,y,s1) conj G_s(k + Q)) + the spread qbar
//     + the fit's part at the fitted frequencies, and its contributions to the
//     Dyson adjoint ebar_k = sum_w G^H Gbar G^H, mubar = -Re sum tr(G^H Gbar G^H):
//     epart[chunk][col][i][j], mpart[inst][chunk][k][xy].
//     sum_s w_s conj(G_s(z)) = conj(W(z) / p(z)) with W(z) = sum_s conj(w_s) num_s(z),
//     one polynomial per neighbour.
template <int N>
__global__ void __launch_bounds__(256) k_sq_ebar(long nbatch, long nk, int d0, int d1, int d2, int q0, int q1, int q2,
                                                 int n_iw, int chunk, double beta, double inv_nk,
                                                 const double *__restrict__ mu, const c2 *__restrict__ ek,
                                                 const c2 *__restrict__ chibar, const double *__restrict__ tab,
                                                 const int *__restrict__ fitpos, TailFitDev fit,
                                                 const c2 *__restrict__ qbar, c2 *__restrict__ epart,
                                                 double *__restrict__ mpart) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long ncol = nbatch * nk * n2;
  const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const int xy = (int)(col % n2), x = xy / N, y = xy - x * N;
  const long ik = col / n2, inst = ik / nk, k = ik - inst * nk;
  const long km = shift_k(k, d0, d1, d2, q0, q1, q2, -1), kp = shift_k(k, d0, d1, d2, q0, q1, q2, +1);
  const double m = mu[inst];
  const c2 *cb = chibar + inst * n4;
  const double sc = -inv_nk / beta;
  // row x and column y of G(k); p of k, k - Q, k + Q; the neighbour polynomials
  c2 rx[N][N], cy[N][N], c0[N], cm[N], cp[N], W1[N], W2[N];
  double z0, zm = 0.0, zp = 0.0;
  {
    Resolvent<N> R;
    resolvent_make<N>(ek + (inst * nk + k) * n2, R);
#pragma unroll
    for (int i = 0; i < N; ++i) {
      resolvent_elem<N>(R, x, i, rx[i]);
      resolvent_elem<N>(R, i, y, cy[i]);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) c0[j] = R.c[j];
    z0 = m - R.s;
  }
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    Resolvent<N> R;
    resolvent_make<N>(ek + (inst * nk + (side ? kp : km)) * n2, R);
    c2 W[N];
#pragma unroll
    for (int j = 0; j < N; ++j) W[j] = c2{0, 0};
#pragma unroll
    for (int s = 0; s < n2; ++s) {
      const int s1 = s / N, s2 = s - s1 * N;
      const c2 w = side ? cb[((s2 * N + x) * N + y) * N + s1] : cb[((y * N + s1) * N + s2) * N + x];
      const c2 wc = c2{w.x * sc, -w.y * sc};  // conj(w_s)
#pragma unroll
      for (int j = 0; j + 1 < N; ++j) W[j] = cfma(W[j], wc, R.B[j][s1][s2]);
      if (s1 == s2) W[N - 1] = cadd(W[N - 1], wc);
    }
#pragma unroll
    for (int j = 0; j < N; ++j) {
      if (side) {
        W2[j] = W[j];
        cp[j] = R.c[j];
      } else {
        W1[j] = W[j];
        cm[j] = R.c[j];
      }
    }
    if (side) zp = m - R.s;
    else zm = m - R.s;
  }
  c2 qb[kQSlots], mb[4];
#pragma unroll
  for (int i = 0; i < kQSlots; ++i) qb[i] = qbar[(long)i * ncol + col];
  mu_to_moment_bar(qb, mb);
#pragma unroll
  for (int j = 1; j <= 3; ++j) mb[j] = cscale(mb[j], fit.wpow[j]);
  c2 eb[N][N];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) eb[i][j] = c2{0, 0};
  double mub = 0.0;
  // gbar at mesh index mm with its spread part sd; accumulate the Dyson adjoint
  auto visit = [&](int mm, c2 sd) {
    const double w = wn(mm - n_iw, beta);
    c2 gb = sd;
    {
      const c2 za = c2{zm, w}, zb = c2{zp, w};
      gb = cadd(gb, cconj(cmul(horner<N>(W1, za), crecip(charpoly<N>(cm, za)))));
      gb = cadd(gb, cconj(cmul(horner<N>(W2, zb), crecip(charpoly<N>(cp, zb)))));
    }
    const int f = fitpos[mm];
    if (f >= 0) {
#pragma unroll
      for (int kk = 1; kk <= 3; ++kk) gb = cadd(gb, cmulc(mb[kk], fit.pinv[(long)kk * fit.n_fit + f]));
    }
    const c2 z = c2{z0, w};
    const c2 inv = crecip(charpoly<N>(c0, z));
    c2 gx[N], gy[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      gx[i] = cmul(horner<N>(rx[i], z), inv);  // G[x][i]
      gy[i] = cmul(horner<N>(cy[i], z), inv);  // G[i][y]
    }
    // G^H (gb e_x e_y^T) G^H: [i][j] += gb conj(G[x][i] G[j][y])
    c2 tr = c2{0, 0};
#pragma unroll
    for (int i = 0; i < N; ++i) {
      tr = cfma(tr, gx[i], gy[i]);
      const c2 gh = cmulc(gb, gx[i]);
#pragma unroll
      for (int j = 0; j < N; ++j) eb[i][j] = cadd(eb[i][j], cmulc(gh, gy[j]));
    }
    mub -= cmulc(gb, tr).x;
  };
  const int n0 = blockIdx.y * chunk, n1 = min(n_iw, n0 + chunk);
  for (int n = n0; n < n1; ++n) {
    c2 sb, db;
    qbar_spread(qb, tab + (long)n * kTabW, sb, db);
    visit(n_iw + n, cadd(sb, db));
    visit(n_iw - 1 - n, csub(sb, db));
  }
  c2 *o = epart + ((long)blockIdx.y * ncol + col) * n2;
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int j = 0; j < N; ++j) o[i * N + j] = eb[i][j];
  mpart[(inst * gridDim.y + blockIdx.y) * nk * n2 + (k * n2 + xy)] = mub;
}

// A3. ebar_k[inst][k][i][j] = sum over chunks and the n^2 columns of k; the
//     thread of (k, i j = 0) also sums k's mubar partials into mk[inst][k].
__global__ void __launch_bounds__(256) k_sq_ebar_final(long nk, long nk_total, int n2, int sk, long ncol,
                                                       const c2 *__restrict__ epart, const double *__restrict__ mpart,
                                                       c2 *__restrict__ ebar, double *__restrict__ mk) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nk_total * n2) return;
  const long ik = t / n2;
  const int ij = (int)(t - ik * n2);
  c2 v = c2{0, 0};
  for (int yk = 0; yk < sk; ++yk)
    for (int xy = 0; xy < n2; ++xy) v = cadd(v, epart[((long)yk * ncol + ik * n2 + xy) * n2 + ij]);
  ebar[t] = v;
  if (ij == 0) {
    const long inst = ik / nk, k = ik - inst * nk;
    double m = 0.0;
    for (int yk = 0; yk < sk; ++yk)
      for (int xy = 0; xy < n2; ++xy) m += mpart[((inst * sk + yk) * nk + k) * n2 + xy];
    mk[ik] = m;
  }
}

namespace {
constexpr long kSqTargetThreads = 1L << 18;
// split len into chunks of at least min_len so that threads * chunks reaches the target
int sq_chunks(long threads, int len, int min_len, int *chunk) {
  long sk = (kSqTargetThreads + threads - 1) / threads;
  const long cap = len / min_len > 1 ? len / min_len : 1;
  if (sk > cap) sk = cap;
  if (sk < 1) sk = 1;
  *chunk = (int)((len + sk - 1) / sk);
  return (len + *chunk - 1) / *chunk;
}
constexpr int pair_bt(int n) { return n == 3 ? 64 : n == 2 ? 128 : 256; }
struct SqPlan {
  long ncol;
  int sk, chunk;    // projection chunks over n >= 0
  long nblk;        // pair blocks per instance
  int sk2, chunk2;  // pair chunks over the whole mesh
  int sk3, chunk3;  // adjoint chunks over n >= 0
  SqPlan(long nbatch, const int dims[3], int n, int n_iw) {
    const long nk = (long)dims[0] * dims[1] * dims[2];
    ncol = nbatch * nk * n * n;
    sk = sq_chunks(ncol, n_iw, 32, &chunk);
    const int bt = pair_bt(n);
    nblk = (nk * n + bt - 1) / bt;
    sk2 = sq_chunks(nbatch * nk * n, 2 * n_iw, 32, &chunk2);
    sk3 = sq_chunks(ncol, n_iw, 16, &chunk3);
  }
};
}  // namespace

// workspace (bytes): partial projections, the projections (unused when the
// caller keeps them), then the pair partials, or the adjoint's qbar, ebar
// partials and mubar partials
size_t static_q_work_bytes(long nbatch, const int dims[3], int n, int n_iw, bool adjoint) {
  const SqPlan P(nbatch, dims, n, n_iw);
  size_t elems = (size_t)(33 * P.sk + kQSlots) * P.ncol;
  if (!adjoint)
    elems += (size_t)nbatch * P.sk2 * P.nblk * n * n * n * n;
  else
    elems += (size_t)kQSlots * P.ncol + (size_t)P.sk3 * P.ncol * n * n + ((size_t)P.sk3 * P.ncol + 1) / 2 +
             ((size_t)nbatch * dims[0] * dims[1] * dims[2] + 1) / 2;
  return elems * sizeof(c2);
}

template <int N>
static hipError_t static_q_run(hipStream_t s, long nbatch, const int dims[3], const int qv[3], int n_iw, double beta,
                               const double *mu, const c2 *ek, const double *tab, const int *fitpos,
                               const TailFitDev &fit, const StaticMatsDev &sm, c2 *work, c2 *proj, bool reuse,
                               const c2 *chibar, c2 *chi0_q, c2 *ebar, double *mubar, double *status) {
  constexpr int n2 = N * N, n4 = n2 * n2;
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const SqPlan P(nbatch, dims, N, n_iw);
  const long ncol = P.ncol;
  c2 *part = work, *q = proj ? proj : work + (long)33 * P.sk * ncol;
  c2 *rest = work + (long)(33 * P.sk + kQSlots) * ncol;
  const unsigned bc = (unsigned)((ncol + 255) / 256);
  if (!reuse) {
    hipLaunchKernelGGL(k_sq_proj_part<N>, dim3(bc, P.sk), dim3(256), 0, s, nbatch, nk, n_iw, P.chunk, beta, mu, ek,
                       tab, part, q);
    hipLaunchKernelGGL(k_sq_fit<N>, dim3((unsigned)((ncol + 63) / 64)), dim3(64 * kFitG), 0, s, nbatch, nk, n_iw,
                       beta, mu, ek, fit, q, status);
    if (P.sk > 1)
      hipLaunchKernelGGL(k_sq_proj_sum, dim3(bc, kQM), dim3(256), 0, s, ncol, P.sk, part, q);
  }
  if (!chibar) {  // forward
    constexpr int BT = pair_bt(N);
    hipLaunchKernelGGL((k_sq_pair<N, BT>), dim3((unsigned)(nbatch * P.nblk), P.sk2), dim3(BT), 0, s, nbatch, nk,
                       dims[0], dims[1], dims[2], qv[0], qv[1], qv[2], n_iw, P.chunk2, (int)P.nblk, beta, mu, ek, q,
                       sm, rest);
    hipLaunchKernelGGL(k_sq_pair_final, dim3((unsigned)(nbatch * n4)), dim3(64), 0, s, n4, (int)(P.sk2 * P.nblk),
                       1.0 / (double)nk, rest, chi0_q);
    return hipGetLastError();
  }
  c2 *qbar = rest, *epart = rest + (long)kQSlots * ncol;
  double *mpart = (double *)(epart + (long)P.sk3 * ncol * n2);
  hipLaunchKernelGGL(k_sq_qbar<N>, dim3(bc), dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2], qv[0], qv[1],
                     qv[2], 1.0 / (double)nk, q, chibar, sm, qbar);
  hipLaunchKernelGGL(k_sq_ebar<N>, dim3(bc, P.sk3), dim3(256), 0, s, nbatch, nk, dims[0], dims[1], dims[2], qv[0],
                     qv[1], qv[2], n_iw, P.chunk3, beta, 1.0 / (double)nk, mu, ek, chibar, tab, fitpos, fit, qbar,
                     epart, mpart);
  double *mk = mpart + ((P.sk3 * ncol + 1) / 2) * 2;  // after mpart, c2-aligned
  hipLaunchKernelGGL(k_sq_ebar_final, dim3((unsigned)((nbatch * nk * n2 + 255) / 256)), dim3(256), 0, s, nk,
                     nbatch * nk, n2, P.sk3, ncol, epart, mpart, ebar, mk);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_sum_partials(s, mk, nbatch, (int)nk, mubar);
}

hipError_t launch_static_q(hipStream_t s, long nbatch, const int dims[3], const int q[3], int n, int n_iw, double beta,
                           const double *mu, const c2 *ek, const double *tab, const int *fitpos, const TailFitDev &fit,
                           const StaticMatsDev &sm, c2 *work, c2 *proj, bool reuse, const c2 *chibar, c2 *chi0_q,
                           c2 *ebar, double *mubar, double *status) {
  switch (n) {
    case 1: return static_q_run<1>(s, nbatch, dims, q, n_iw, beta, mu, ek, tab, fitpos, fit, sm, work, proj, reuse, chibar, chi0_q, ebar, mubar, status);
    case 2: return static_q_run<2>(s, nbatch, dims, q, n_iw, beta, mu, ek, tab, fitpos, fit, sm, work, proj, reuse, chibar, chi0_q, ebar, mubar, status);
    case 3: return static_q_run<3>(s, nbatch, dims, q, n_iw, beta, mu, ek, tab, fitpos, fit, sm, work, proj, reuse, chibar, chi0_q, ebar, mubar, status);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace tprf
