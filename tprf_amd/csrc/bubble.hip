// bubble.hip -- fused bubble product + bosonic tau -> nu transform for the
// |m| < nb output frequencies (chi0_wr_from_grt_PH, chi_imtime.cpp:139-182,
// with the tau -> nu math of fourier_matsubara.cpp:42-184).
//
// For one canonical r (r <= -r) a workgroup forms the products
//   x_abcd(tau_j) = G_da(tau_j, r) G_bc(tau_{L-j}, -r)
// and their raw sums S(m) = sum_{j<L} x_j exp(2 pi i j m / L) by splitting
// j = N1 j2 + j1 (L = N1 P): for every j1 a length-P DFT over j2 (P >= 2 nb, a
// power of two, done as R1 x R2 in registers with one LDS transpose) is
// twiddled by exp(2 pi i j1 m / L) and accumulated.  Only the 2nb-1 needed
// frequencies are ever formed; G(tau, r) and G(tau, -r) are read once per
// workgroup, P tau-points at a time.
//
// The columns of -r are the columns of r reversed in tau:
//   x^{-r}_cdab(tau_j) = x^{r}_abcd(tau_{L-j}),
// whose sums are S^{-r}(m) = S^{r}(-m) - x_0 + x_L exactly, so each pair
// (r, -r) is transformed once.  The pole model is applied after the transform
// (it is linear): y(m) = (beta/L) (S(m) - sum_i a_i Phi_i(m)) + sum_i a_i/(i nu_m - b_i)
// with Phi_i(m) the transform of the model's tau samples, and a_i from each
// column's own boundary derivatives.
#include "common.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

// AB = true: a workgroup owns the CG = n^2 columns of one orbital pair (a, b)
// (abcd with c, d free) and stages only the n entries G_da(r), d < n, and
// G_bc(-r), c < n, per tau point; AB = false: CG consecutive columns, all n^2
// entries of G(r) and G(-r) staged.
template <int N, int R1, int R2, int CG, int NT, bool AB>
__global__ void __launch_bounds__(NT) k_bubble_pair(double beta, int L, int nb, int d0, int d1, int d2,
                                                    const c2 *__restrict__ gtr, c2 *__restrict__ out, BubbleDev bd,
                                                    Vinv8 vinv) {
  constexpr int P = R1 * R2;
  constexpr int n2 = N * N, n4 = n2 * n2;
  static_assert(!AB || CG == n2, "one orbital pair per workgroup");
  static_assert(NT >= P, "the outer twiddles of a step are staged one per thread");
  constexpr int NG = (n4 + CG - 1) / CG;
  constexpr int U2 = (CG * R1 + NT - 1) / NT;
  constexpr int TLD = R2 + 1;  // padded LDS row of the transpose
  constexpr int NE = AB ? N : n2;  // staged entries of G per tau point and side
  constexpr int NSTAGE = P * NE;
  constexpr int PF = (NSTAGE + NT - 1) / NT;
  static_assert(CG * R1 * TLD >= CG * P, "transpose buffer reused for the final exchange");

  __shared__ c2 sA[NSTAGE], sB[NSTAGE];
  __shared__ c2 sT[CG * R1 * TLD];
  __shared__ c2 sE[CG][18];
  __shared__ c2 sPole[CG][2][3];
  __shared__ c2 sTwp[P];   // exp(2 pi i p / P)
  __shared__ c2 sTw1[P];   // exp(2 pi i j1 m(k) / L) of the current j1

  const int tid = threadIdx.x;
  const long nr = (long)d0 * d1 * d2;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long pair = tile / NG;
  const int grp = (int)(tile - pair * NG);
  const long r = bd.rcanon[pair];
  const long mr = minus_r(r, d0, d1, d2);
  const int q0 = grp * CG;
  const int cn = (n4 - q0) < CG ? (n4 - q0) : CG;
  const int a0 = q0 / (n2 * N), b0 = (q0 / n2) % N;  // the orbital pair of an AB workgroup
  const int N1 = L / P;
  const double dtau = beta / (double)L;

  // products at tau_0..tau_8 and tau_{L-8}..tau_L: sE[c][e], e < 9 -> j = e, e >= 9 -> j = L - 17 + e
  for (int t = tid; t < cn * 18; t += NT) {
    const int c = t / 18, e = t - c * 18;
    const int j = e < 9 ? e : L - 17 + e;
    const int q = q0 + c;
    const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
    sE[c][e] = cmul(gtr[((long)j * nr + r) * n2 + dd * N + aa], gtr[((long)(L - j) * nr + mr) * n2 + bb * N + cc]);
  }
  __syncthreads();
  if (tid < cn) {
    const c2 *x = sE[tid];
    boundary_poles<1>([&](int m) { return x[m]; }, [&](int m) { return x[17 - m]; }, dtau, vinv, sPole[tid][0],
                      (c2 *)nullptr);
    boundary_poles<1>([&](int m) { return x[17 - m]; }, [&](int m) { return x[m]; }, dtau, vinv, sPole[tid][1],
                      (c2 *)nullptr);
  }

  for (int t = tid; t < P; t += NT) sTwp[t] = bd.twp[t];
  c2 pa[PF], pb[PF], pw = c2{0.0, 0.0};
  // loads of step j1 + 1 are issued right after step j1's tile is staged
  auto fetch = [&](int j1) {
    if (tid < P) pw = bd.tw1[(long)j1 * P + tid];
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = tid + i * NT;
      if (t < NSTAGE) {
        const int j2 = t / NE, e = t - j2 * NE;
        const long tau = (long)N1 * j2 + j1;
        // AB: e = d of G_da(r) (a = a0) and e = c of G_bc(-r) (b = b0)
        const int ea = AB ? e * N + a0 : e, eb = AB ? b0 * N + e : e;
        pa[i] = gtr[(tau * nr + r) * n2 + ea];
        pb[i] = gtr[((L - tau) * nr + mr) * n2 + eb];
      }
    }
  };
  fetch(0);

  c2 acc[U2][R2];
#pragma unroll
  for (int i = 0; i < U2; ++i)
#pragma unroll
    for (int k = 0; k < R2; ++k) acc[i][k] = c2{0.0, 0.0};

  for (int j1 = 0; j1 < N1; ++j1) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = tid + i * NT;
      if (t < NSTAGE) {
        sA[t] = pa[i];
        sB[t] = pb[i];
      }
    }
    if (tid < P) sTw1[tid] = pw;
    __syncthreads();
    if (j1 + 1 < N1) fetch(j1 + 1);
    // stage 1: unit (c, s): DFT-R1 over k of x at j2 = R2 k + s, twiddled by exp(2 pi i s k1 / P)
    for (int u = tid; u < cn * R2; u += NT) {
      const int c = u / R2, s = u - c * R2;
      const int q = q0 + c;
      const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
      c2 v[R1];
#pragma unroll
      for (int k = 0; k < R1; ++k) {
        const int j2 = R2 * k + s;
        v[k] = AB ? cmul(sA[j2 * NE + dd], sB[j2 * NE + cc])
                  : cmul(sA[j2 * n2 + dd * N + aa], sB[j2 * n2 + bb * N + cc]);
      }
      SmallDft<R1, 1>::run(v, nullptr, 0);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], sTwp[s * k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) sT[(c * R1 + k1) * TLD + s] = v[k1];
    }
    __syncthreads();
    // stage 2: unit (c, k1): DFT-R2 over s gives bins k = k1 + R1 k2; accumulate with exp(2 pi i j1 m / L)
    const c2 *w = sTw1;
#pragma unroll
    for (int i = 0; i < U2; ++i) {
      const int u = tid + i * NT;
      if (u < cn * R1) {
        const int c = u / R1, k1 = u - c * R1;
        c2 v[R2];
#pragma unroll
        for (int s = 0; s < R2; ++s) v[s] = sT[(c * R1 + k1) * TLD + s];
        SmallDft<R2, 1>::run(v, nullptr, 0);
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) acc[i][k2] = cadd(acc[i][k2], cmul(w[k1 + R1 * k2], v[k2]));
      }
    }
  }
  __syncthreads();
  c2 *sS = sT;  // S(m) of every (column, bin) for the mirrored columns
#pragma unroll
  for (int i = 0; i < U2; ++i) {
    const int u = tid + i * NT;
    if (u < cn * R1) {
      const int c = u / R1, k1 = u - c * R1;
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) sS[c * P + k1 + R1 * k2] = acc[i][k2];
    }
  }
  __syncthreads();
  const double fact = beta / (double)L;
  const bool mirror = (mr != r);
  const int nnu = 2 * nb - 1;
#pragma unroll
  for (int i = 0; i < U2; ++i) {
    const int u = tid + i * NT;
    if (u < cn * R1) {
      const int c = u / R1, k1 = u - c * R1;
      const int q = q0 + c;
      const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
      const int qm = ((cc * N + dd) * N + aa) * N + bb;
      const c2 *a0 = sPole[c][0], *a1 = sPole[c][1];
      const c2 xl_minus_x0 = csub(sE[c][17], sE[c][0]);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const int k = k1 + R1 * k2;
        const int m = k < P / 2 ? k : k - P;
        if (m <= -nb || m >= nb) continue;
        const int im = m + nb - 1;
        const c2 f0 = bd.phi[im], f1 = bd.phi[nnu + im], f2 = bd.phi[2 * nnu + im];
        c2 model = cadd(cadd(cmul(a0[0], f0), cmul(a0[1], f1)), cmul(a0[2], f2));
        const c2 g0 = bd.pinv_nu[im], g1 = bd.pinv_nu[nnu + im], g2 = bd.pinv_nu[2 * nnu + im];
        c2 v = cadd(cscale(csub(acc[i][k2], model), fact),
                    cadd(cadd(cmul(a0[0], g0), cmul(a0[1], g1)), cmul(a0[2], g2)));
        out[((long)im * nr + r) * n4 + q] = v;
        if (mirror) {
          const c2 sm = cadd(sS[c * P + ((P - k) & (P - 1))], xl_minus_x0);
          model = cadd(cadd(cmul(a1[0], f0), cmul(a1[1], f1)), cmul(a1[2], f2));
          v = cadd(cscale(csub(sm, model), fact), cadd(cadd(cmul(a1[0], g0), cmul(a1[1], g1)), cmul(a1[2], g2)));
          out[((long)im * nr + mr) * n4 + qm] = v;
        }
      }
    }
  }
}

template <int N, int R1, int R2, int CG, int NT, bool AB>
static hipError_t launch_one(hipStream_t s, double beta, int L, int nb, const int dims[3], const c2 *gtr, c2 *out,
                             const BubbleDev &bd, const Vinv8 &vinv) {
  constexpr int n4 = N * N * N * N;
  constexpr int NG = (n4 + CG - 1) / CG;
  const long nb_blocks = (long)bd.ncanon * NG;
  hipLaunchKernelGGL((k_bubble_pair<N, R1, R2, CG, NT, AB>), dim3(nb_blocks), dim3(NT), 0, s, beta, L, nb, dims[0],
                     dims[1], dims[2], gtr, out, bd, vinv);
  return hipGetLastError();
}

bool bubble_pair_supported(int n, int P) { return (n == 2 || n == 3) && (P == 64 || P == 128); }

int bubble_pair_variants() { return 3; }

// Workgroup shapes (same results, picked by the autotuner): variant 0 owns one
// orbital pair; 1 is a narrower or differently grouped shape; 2 stages every G
// entry for CG = n^4 / 3 (n = 3) or n^4 (n = 2) columns.
hipError_t launch_bubble_pair(hipStream_t s, double beta, int L, int nb, const int dims[3], int n, const c2 *gtr,
                              c2 *out, const BubbleDev &bd, const Vinv8 &vinv, int variant) {
#define TPRF_BP(N_, R1_, R2_, CG_, NT_, AB_) \
  return launch_one<N_, R1_, R2_, CG_, NT_, AB_>(s, beta, L, nb, dims, gtr, out, bd, vinv)
  if (n == 3 && bd.P == 128) {
    if (variant == 0) TPRF_BP(3, 16, 8, 9, 256, true);
    if (variant == 1) TPRF_BP(3, 16, 8, 9, 128, true);
    TPRF_BP(3, 16, 8, 27, 256, false);
  }
  if (n == 3 && bd.P == 64) {
    if (variant == 0) TPRF_BP(3, 8, 8, 9, 256, true);
    if (variant == 1) TPRF_BP(3, 8, 8, 9, 128, true);
    TPRF_BP(3, 8, 8, 27, 256, false);
  }
  if (n == 2 && bd.P == 128) {
    if (variant == 0) TPRF_BP(2, 16, 8, 4, 128, true);
    if (variant == 1) TPRF_BP(2, 16, 8, 8, 128, false);
    TPRF_BP(2, 16, 8, 16, 256, false);
  }
  if (n == 2 && bd.P == 64) {
    if (variant == 0) TPRF_BP(2, 8, 8, 4, 64, true);
    if (variant == 1) TPRF_BP(2, 8, 8, 4, 128, true);
    TPRF_BP(2, 8, 8, 16, 256, false);
  }
#undef TPRF_BP
  return hipErrorNotSupported;
}

}  // namespace tprf
