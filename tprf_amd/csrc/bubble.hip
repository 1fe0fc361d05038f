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
#include <type_traits>

#include "common.hpp"
#include "lds_dma.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

// AB = true: a workgroup owns the CG = n^2 columns of one orbital pair (a, b)
// (abcd with c, d free) and stages only the n entries G_da(r), d < n, and
// G_bc(-r), c < n, per tau point; AB = false: CG consecutive columns, all n^2
// entries of G(r) and G(-r) staged.
// DMA = false: the tile of step j1 + 1 is loaded into registers during step j1
// and written to LDS at the top of step j1 + 1.  DMA = true: tiles go straight
// from global memory to LDS through two buffers, two steps ahead; each wave
// issues 64-entry chunks of the tile and waits for its own older chunks with
// vmcnt before the barrier that publishes them.
// WPE: waves per SIMD the register allocation must allow (1: no limit).
// HERM (with AB): G satisfies G_ab(tau, r)^* = G_ba(tau, -r) (a Hermitian e_k
// and no self-energy, or one with Sigma(i w)^H = Sigma(-i w)).  Only the
// columns of r <= -r are read: the -r side is conj(G_cb(tau, r)).  The
// products obey x^r_badc(tau_j) = x^r_abcd(tau_{L-j})^* and
// x^{-r}_dcba(tau_j) = x^r_abcd(tau_j)^*, so only the orbital pairs a <= b are
// transformed: a workgroup with a < b also writes the columns badc of r and
// dcba of -r, with S_badc(m) = (S(m) - x_0 + x_L)^*, S^{-r}_dcba(m) = S(-m)^*
// and the conjugated pole coefficients of the mirrored and the direct column.
// For a == b the same identities pair (a, a, c, d) with (a, a, d, c), and only
// the columns c <= d are transformed.
// Outputs are written for the frequencies m_lo <= m < nb (m_lo = 1 - nb: all;
// 0: nu >= 0 only).
// MG (HERM, n = 3): a workgroup owns the off-diagonal pair (a1, b1) = (0,1),
// (0,2) or (1,2) and the diagonal pair (a2, a2) = (0,0), (1,1) or (2,2): 9 + 6
// columns, three workgroups per r pair, so that the DFT units fill the lanes
// of 256 threads; the tile holds G_da for a in {a1, a2} and G_eb for b in
// {b1, a2}.
template <int N, int R1, int R2, int CG, int NT, bool AB, bool DMA, int WPE, bool HERM = false, bool MG = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) k_bubble_pair(double beta, int L, int nb, int d0, int d1, int d2,
                                                    const c2 *__restrict__ gtr, c2 *__restrict__ out, BubbleDev bd,
                                                    Vinv8 vinv, int m_lo) {
  constexpr int P = R1 * R2;
  constexpr int n2 = N * N, n4 = n2 * n2;
  static_assert(!AB || CG == n2 || MG, "one orbital pair per workgroup");
  static_assert(!MG || (HERM && N == 3 && CG == 15), "merged pairs: Hermitian form, three orbitals");
  static_assert(!HERM || AB, "the Hermitian form works on orbital-pair workgroups");
  static_assert(NT >= P, "the outer twiddles of a step are staged one per thread");
  constexpr int NG = MG ? N : (HERM ? N * (N + 1) / 2 : (n4 + CG - 1) / CG);
  constexpr int U2 = (CG * R1 + NT - 1) / NT;
  constexpr int TLD = R2 + 1;  // padded LDS row of the transpose
  constexpr int NE = MG ? 2 * N : (AB ? N : n2);  // staged entries of G per tau point and side
  constexpr int NSTAGE = P * NE;
  constexpr int PF = (NSTAGE + NT - 1) / NT;
  static_assert(CG * R1 * TLD >= CG * P, "transpose buffer reused for the final exchange");
  constexpr int NW = NT / 64;                  // waves per workgroup
  constexpr int NCH = NSTAGE / 64;             // 64-entry chunks of one side of a tile
  constexpr int KD = (2 * NCH) / NW;           // chunks each wave issues per tile
  static_assert(!DMA || (NSTAGE % 64 == 0 && NT % 64 == 0 && (2 * NCH) % NW == 0),
                "whole-wave chunks, the same number for every wave");

  __shared__ c2 sA[NSTAGE], sB[NSTAGE];
  __shared__ c2 sA2[DMA ? NSTAGE : 1], sB2[DMA ? NSTAGE : 1];  // second buffers of the DMA pipeline
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
  // the orbital pair of an AB workgroup: (grp / n, grp % n), or with HERM the
  // grp-th pair a <= b in row-major order
  int a0 = grp / N, b0 = grp % N;
  if constexpr (HERM) {
    a0 = 0;
    int rest = grp;
    while (rest >= N - a0) { rest -= N - a0; ++a0; }
    b0 = a0 + rest;
  }
  int a2 = 0;  // MG: the diagonal pair's orbital
  if constexpr (MG) {
    a0 = grp < 2 ? 0 : 1;
    b0 = grp == 0 ? 1 : 2;
    a2 = grp;
  }
  const int q0 = AB ? (a0 * N + b0) * n2 : grp * CG;
  // HERM, a == b: (a, a, c, d) and (a, a, d, c) are conjugate partners, so only
  // the n (n + 1) / 2 columns c <= d are transformed; column slot c' maps to
  // the c'-th pair c <= d in row-major order
  const bool diag_pair = HERM && !MG && a0 == b0;
  const int cn = MG ? CG : (diag_pair ? N * (N + 1) / 2 : ((n4 - q0) < CG ? (n4 - q0) : CG));
  auto col_q = [&](int c) {
    if (MG && c >= n2) {  // the diagonal pair's columns c <= d
      int cc = 0, rest = c - n2;
      while (rest >= N - cc) { rest -= N - cc; ++cc; }
      return (a2 * N + a2) * n2 + cc * N + cc + rest;
    }
    if (!diag_pair) return q0 + c;
    int cc = 0, rest = c;
    while (rest >= N - cc) { rest -= N - cc; ++cc; }
    return q0 + cc * N + cc + rest;
  };
  // MG: the tile slot (0: a1 / b1, 1: a2) of column slot c
  auto tslot = [&](int c) { return (MG && c >= n2) ? N : 0; };
  const int N1 = L / P;
  const double dtau = beta / (double)L;

  // products at tau_0..tau_8 and tau_{L-8}..tau_L: sE[c][e], e < 9 -> j = e, e >= 9 -> j = L - 17 + e,
  // and the inner twiddles sTwp; every load is issued before the first LDS store
  // (clamped indices: entries past cn * 18 load a valid element and are dropped)
  {
    constexpr int NEB = (CG * 18 + NT - 1) / NT, NTP = (P + NT - 1) / NT;
    v2d ga[NEB], gb[NEB], tp[NTP];
#pragma unroll
    for (int w = 0; w < NEB; ++w) {
      const int t = min(tid + w * NT, cn * 18 - 1);
      const int c = t / 18, e = t - c * 18;
      const int j = e < 9 ? e : L - 17 + e;
      const int q = col_q(c);
      const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
      ga[w] = ld_v2d(gtr + ((long)j * nr + r) * n2 + dd * N + aa);
      // HERM: G_bc(tau, -r) = conj(G_cb(tau, r)), read from r (conjugated below)
      gb[w] = HERM ? ld_v2d(gtr + ((long)(L - j) * nr + r) * n2 + cc * N + bb)
                   : ld_v2d(gtr + ((long)(L - j) * nr + mr) * n2 + bb * N + cc);
    }
#pragma unroll
    for (int w = 0; w < NTP; ++w) tp[w] = ld_v2d(bd.twp + min(tid + w * NT, P - 1));
#pragma unroll
    for (int w = 0; w < NEB; ++w) {
      const int t = tid + w * NT;
      if (t < cn * 18)
        sE[t / 18][t % 18] = HERM ? cmulc(c2{ga[w].x, ga[w].y}, c2{gb[w].x, gb[w].y})
                                  : cmul(c2{ga[w].x, ga[w].y}, c2{gb[w].x, gb[w].y});
    }
#pragma unroll
    for (int w = 0; w < NTP; ++w)
      if (tid + w * NT < P) st_v2d(sTwp + tid + w * NT, tp[w]);
  }
  __syncthreads();
  if (tid < cn) {
    const c2 *x = sE[tid];
    boundary_poles<1>([&](int m) { return x[m]; }, [&](int m) { return x[17 - m]; }, dtau, vinv, sPole[tid][0],
                      (c2 *)nullptr);
    boundary_poles<1>([&](int m) { return x[17 - m]; }, [&](int m) { return x[m]; }, dtau, vinv, sPole[tid][1],
                      (c2 *)nullptr);
  }

  // Element offset of staged entry t of side A (G_da(tau, r)) or side B
  // (G_bc(L - tau, -r)) at tau = j2 N1 (+ j1 for step j1); AB: e = d with
  // a = a0, and e = c with b = b0.
  const long stepa = nr * n2;  // one tau point
  auto off_a = [&](int t) {
    const int j2 = t / NE, e = t - j2 * NE;
    if constexpr (MG) {
      const int x = e % N, a = e < N ? a0 : a2;
      return ((long)N1 * j2 * nr + r) * n2 + x * N + a;
    }
    return ((long)N1 * j2 * nr + r) * n2 + (AB ? e * N + a0 : e);
  };
  auto off_b = [&](int t) {
    const int j2 = t / NE, e = t - j2 * NE;
    if constexpr (MG) {
      const int x = e % N, b = e < N ? b0 : a2;
      return (((long)L - (long)N1 * j2) * nr + r) * n2 + x * N + b;  // conj(G_xb(r))
    }
    if constexpr (HERM) return (((long)L - (long)N1 * j2) * nr + r) * n2 + e * N + b0;  // conj(G_eb(r))
    return (((long)L - (long)N1 * j2) * nr + mr) * n2 + (AB ? b0 * N + e : e);
  };

  c2 acc[U2][R2];
#pragma unroll
  for (int i = 0; i < U2; ++i)
#pragma unroll
    for (int k = 0; k < R2; ++k) acc[i][k] = c2{0.0, 0.0};

  // stage 1: unit (c, s): DFT-R1 over k of x at j2 = R2 k + s, twiddled by exp(2 pi i s k1 / P)
  auto stage1 = [&](const c2 *tA, const c2 *tB) {
    for (int u = tid; u < cn * R2; u += NT) {
      const int c = u / R2, s = u - c * R2;
      const int q = col_q(c);
      const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
      c2 v[R1];
#pragma unroll
      for (int k = 0; k < R1; ++k) {
        const int j2 = R2 * k + s;
        v[k] = HERM ? cmulc(tA[j2 * NE + tslot(c) + dd], tB[j2 * NE + tslot(c) + cc])
             : AB ? cmul(tA[j2 * NE + dd], tB[j2 * NE + cc])
                  : cmul(tA[j2 * n2 + dd * N + aa], tB[j2 * n2 + bb * N + cc]);
      }
      SmallDft<R1, 1>::run(v, nullptr, 0);
#pragma unroll
      for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], sTwp[s * k1]);
#pragma unroll
      for (int k1 = 0; k1 < R1; ++k1) sT[(c * R1 + k1) * TLD + s] = v[k1];
    }
  };
  // stage 2: unit (c, k1): DFT-R2 over s gives bins k = k1 + R1 k2; accumulate with exp(2 pi i j1 m / L)
  auto stage2 = [&]() {
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
        for (int k2 = 0; k2 < R2; ++k2) acc[i][k2] = cfma(acc[i][k2], w[k1 + R1 * k2], v[k2]);
      }
    }
  };

  // The twiddles of step j1 + 1 travel in a register (v2d: a double2 copy
  // assigned under a branch in a loop can be left in private memory).
  v2d pw = v2d{0.0, 0.0};
  if constexpr (DMA) {
    // chunk c = wave + NW i (i < KD) of a tile: side A for c < NCH, side B
    // after.  Selects rather than branches, so that every wave issues exactly
    // KD loads per tile on every path.
    const int wave = tid >> 6, lane = tid & 63;
    long offd[KD];
#pragma unroll
    for (int i = 0; i < KD; ++i) {
      const int c = wave + NW * i;
      offd[i] = c < NCH ? off_a(c * 64 + lane) : off_b((c - NCH) * 64 + lane);
    }
    auto issue = [&](int j1, c2 *tA, c2 *tB) {
      const long da = (long)j1 * stepa;
#pragma unroll
      for (int i = 0; i < KD; ++i) {
        const int c = wave + NW * i;
        const bool sb = c >= NCH;
        dma16(gtr + offd[i] + (sb ? -da : da), sb ? tB + (c - NCH) * 64 : tA + c * 64);
      }
    };
    // Step j1 reads tile j1 from (tA, tB).  ISSUE: tile j1 + 2 goes into the
    // same buffers once every wave is past stage 1, behind the twiddles of
    // step j1 + 1, so each wave's newest KD or more loads are always tile
    // j1 + 2's and vmcnt(KD) at the next step waits for nothing younger than
    // that step's own tile and twiddles.  The last two steps issue nothing
    // and are peeled off the loop.
    auto step = [&](auto issue_tag, int j1, c2 *tA, c2 *tB) {
      constexpr bool ISSUE = decltype(issue_tag)::value;
      if (ISSUE || j1 + 1 < N1)
        wait_vmcnt<KD>();
      else
        wait_vmcnt<0>();
      barrier_keep_vm();
      if (tid < P) st_v2d(sTw1 + tid, pw);
      stage1(tA, tB);
      barrier_keep_vm();
      if ((ISSUE || j1 + 1 < N1) && tid < P) pw = ld_v2d(bd.tw1 + (long)(j1 + 1) * P + tid);
      if constexpr (ISSUE) issue(j1 + 2, tA, tB);
      stage2();
    };
    using Issue = std::integral_constant<bool, true>;
    using NoIssue = std::integral_constant<bool, false>;
    // N1 >= 2 (the launcher sends N1 = 1 to the register variant)
    issue(0, sA, sB);
    if (tid < P) pw = ld_v2d(bd.tw1 + tid);
    issue(1, sA2, sB2);
    int j1 = 0;
    for (; j1 + 3 < N1; j1 += 2) {
      step(Issue{}, j1, sA, sB);
      step(Issue{}, j1 + 1, sA2, sB2);
    }
    if (N1 - j1 == 3) {
      step(Issue{}, j1, sA, sB);
      step(NoIssue{}, j1 + 1, sA2, sB2);
      step(NoIssue{}, j1 + 2, sA, sB);
    } else {
      step(NoIssue{}, j1, sA, sB);
      step(NoIssue{}, j1 + 1, sA2, sB2);
    }
  } else {
    // Loads of step j1 + 1 are issued right after step j1's tile is staged and
    // stay in flight in registers through step j1's DFTs.
    long offa[PF], offb[PF];
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int t = tid + i * NT;
      offa[i] = off_a(t < NSTAGE ? t : 0);
      offb[i] = off_b(t < NSTAGE ? t : 0);
    }
    v2d pa[PF], pb[PF];
    if (tid < P) pw = ld_v2d(bd.tw1 + tid);
#pragma unroll
    for (int i = 0; i < PF; ++i)
      if (tid + i * NT < NSTAGE) {
        pa[i] = ld_v2d(gtr + offa[i]);
        pb[i] = ld_v2d(gtr + offb[i]);
      }
    for (int j1 = 0; j1 < N1; ++j1) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < PF; ++i)
        if (tid + i * NT < NSTAGE) {
          st_v2d(sA + tid + i * NT, pa[i]);
          st_v2d(sB + tid + i * NT, pb[i]);
        }
      if (tid < P) st_v2d(sTw1 + tid, pw);
      __syncthreads();
      if (j1 + 1 < N1) {
        if (tid < P) pw = ld_v2d(bd.tw1 + (long)(j1 + 1) * P + tid);
        const long da = (long)(j1 + 1) * stepa;
#pragma unroll
        for (int i = 0; i < PF; ++i)
          if (tid + i * NT < NSTAGE) {
            pa[i] = ld_v2d(gtr + offa[i] + da);
            pb[i] = ld_v2d(gtr + offb[i] - da);
          }
      }
      stage1(sA, sB);
      __syncthreads();
      stage2();
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
  // The pole-model table H_q(m) (3 x nnu entries) goes into the tile buffers,
  // free now (sA, then sB): all of a thread's loads are issued before its LDS
  // stores, and the outputs below read LDS instead of waiting on a global load
  // before each store.
  const int nnu = 2 * nb - 1;
  static_assert(3 * P <= 2 * NSTAGE, "the tile buffers hold the pole-model table");
  {
    constexpr int NH = (3 * P + NT - 1) / NT;
    v2d hv[NH];
#pragma unroll
    for (int w = 0; w < NH; ++w) {
      const int f = tid + w * NT;
      hv[w] = ld_v2d(bd.hmod + (f < 3 * nnu ? f : 0));
    }
#pragma unroll
    for (int w = 0; w < NH; ++w) {
      const int f = tid + w * NT;
      if (f < 3 * nnu) st_v2d(f < NSTAGE ? sA + f : sB + (f - NSTAGE), hv[w]);
    }
  }
  __syncthreads();
  auto hmod = [&](int f) { return f < NSTAGE ? sA[f] : sB[f - NSTAGE]; };
  const double fact = beta / (double)L;
  const bool mirror = (mr != r);
#pragma unroll
  for (int i = 0; i < U2; ++i) {
    const int u = tid + i * NT;
    if (u < cn * R1) {
      const int c = u / R1, k1 = u - c * R1;
      const int q = col_q(c);
      const int dd = q % N, cc = (q / N) % N, bb = (q / n2) % N, aa = q / (n2 * N);
      const int qm = ((cc * N + dd) * N + aa) * N + bb;
      const c2 *pa0 = sPole[c][0], *pa1 = sPole[c][1];
      const c2 xl_minus_x0 = csub(sE[c][17], sE[c][0]);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) {
        const int k = k1 + R1 * k2;
        const int m = k < P / 2 ? k : k - P;
        if (m < m_lo || m >= nb) continue;
        const int im = m + nb - 1;
        // y(m) = fact S(m) + sum_q a_q H_q(m), H_q = 1 / (i nu_m - b_q) - fact Phi_q(m)
        const c2 h0 = hmod(im), h1 = hmod(nnu + im), h2 = hmod(2 * nnu + im);
        c2 v = cadd(cscale(acc[i][k2], fact), cfma(cfma(cmul(pa0[0], h0), pa0[1], h1), pa0[2], h2));
        out[((long)im * nr + r) * n4 + q] = v;
        if (mirror) {
          const c2 sm = cadd(sS[c * P + ((P - k) & (P - 1))], xl_minus_x0);
          v = cadd(cscale(sm, fact), cfma(cfma(cmul(pa1[0], h0), pa1[1], h1), pa1[2], h2));
          out[((long)im * nr + mr) * n4 + qm] = v;
        }
        if constexpr (HERM) {
          if (aa < bb || cc != dd) {
            // column badc of r: S = (S(m) - x_0 + x_L)^*, pole coefficients conj(a1)
            const int qb = ((bb * N + aa) * N + dd) * N + cc;
            const c2 sb = cconj(cadd(acc[i][k2], xl_minus_x0));
            v = cadd(cscale(sb, fact),
                     cfma(cfma(cmul(cconj(pa1[0]), h0), cconj(pa1[1]), h1), cconj(pa1[2]), h2));
            out[((long)im * nr + r) * n4 + qb] = v;
            if (mirror) {
              // column dcba of -r: S = S(-m)^*, pole coefficients conj(a0)
              const int qd = ((dd * N + cc) * N + bb) * N + aa;
              const c2 sd = cconj(sS[c * P + ((P - k) & (P - 1))]);
              v = cadd(cscale(sd, fact),
                       cfma(cfma(cmul(cconj(pa0[0]), h0), cconj(pa0[1]), h1), cconj(pa0[2]), h2));
              out[((long)im * nr + mr) * n4 + qd] = v;
            }
          }
        }
      }
    }
  }
}

template <int N, int R1, int R2, int CG, int NT, bool AB, bool DMA, int WPE, bool HERM, bool MG = false>
static hipError_t launch_one_h(hipStream_t s, double beta, int L, int nb, const int dims[3], const c2 *gtr, c2 *out,
                               const BubbleDev &bd, const Vinv8 &vinv, int m_lo) {
  constexpr int n4 = N * N * N * N;
  constexpr int NG = MG ? N : (HERM ? N * (N + 1) / 2 : (n4 + CG - 1) / CG);
  const long nb_blocks = (long)bd.ncanon * NG;
  if (DMA && L / (R1 * R2) < 2) {  // the global -> LDS pipeline needs two tiles
    hipLaunchKernelGGL((k_bubble_pair<N, R1, R2, CG, NT, AB, false, 1, HERM, MG>), dim3(nb_blocks), dim3(NT), 0, s, beta,
                       L, nb, dims[0], dims[1], dims[2], gtr, out, bd, vinv, m_lo);
    return hipGetLastError();
  }
  hipLaunchKernelGGL((k_bubble_pair<N, R1, R2, CG, NT, AB, DMA, WPE, HERM, MG>), dim3(nb_blocks), dim3(NT), 0, s, beta,
                     L, nb, dims[0], dims[1], dims[2], gtr, out, bd, vinv, m_lo);
  return hipGetLastError();
}

// launches the plain form, or with herm the Hermitian form where the shape
// has orbital-pair workgroups (hipErrorNotSupported otherwise)
template <int N, int R1, int R2, int CG, int NT, bool AB, bool DMA, int WPE>
static hipError_t launch_one(hipStream_t s, double beta, int L, int nb, const int dims[3], const c2 *gtr, c2 *out,
                             const BubbleDev &bd, const Vinv8 &vinv, bool herm, int m_lo) {
  if (herm) {
    if constexpr (AB)
      return launch_one_h<N, R1, R2, CG, NT, AB, DMA, WPE, true>(s, beta, L, nb, dims, gtr, out, bd, vinv, m_lo);
    else
      return hipErrorNotSupported;
  }
  return launch_one_h<N, R1, R2, CG, NT, AB, DMA, WPE, false>(s, beta, L, nb, dims, gtr, out, bd, vinv, m_lo);
}

bool bubble_pair_supported(int n, int P) { return (n == 2 || n == 3) && (P == 64 || P == 128); }

int bubble_pair_variants() { return 8; }

// the variants with orbital-pair workgroups (those the Hermitian form has)
int bubble_pair_herm_variants(int n, int P, int *list) {
  static const int v3a[] = {0, 1, 3, 4, 5, 6, 7}, v3b[] = {0, 1, 3, 4, 5}, v2a[] = {0, 3, 7}, v2b[] = {0, 1, 3, 4};
  const int *v = (n == 3) ? (P == 128 ? v3a : v3b) : (P == 128 ? v2a : v2b);
  const int cnt = (n == 3) ? (P == 128 ? 7 : 5) : (P == 128 ? 3 : 4);
  for (int i = 0; i < cnt; ++i) list[i] = v[i];
  return cnt;
}

// Workgroup shapes (same results, picked by the autotuner; variants 5 and 6,
// the merged orbital pairs of the Hermitian form at n = 3 with the inner DFT
// split 16 x 8 or 8 x 16, and variant 7, the Hermitian form with one
// workgroup per (r, -r) pair (bubble_wg.hip, n = 2, 3, P = 128), run as 0
// elsewhere):
// variant 0 owns one
// orbital pair; 1 is a narrower or differently grouped shape; 2 stages every G
// entry for CG = n^4 / 3 (n = 3) or n^4 (n = 2) columns; 3 and 4 load tiles
// straight into LDS (for n = 3, P = 128 the two differ in the occupancy the
// register allocation is held to).
hipError_t launch_bubble_pair(hipStream_t s, double beta, int L, int nb, const int dims[3], int n, const c2 *gtr,
                              c2 *out, const BubbleDev &bd, const Vinv8 &vinv, int variant, bool herm, bool nonneg) {
  const int m_lo = nonneg ? 0 : 1 - nb;
  if (variant == 7) {
    if (herm && (n == 2 || n == 3) && bd.P == 128) return launch_bubble_wg(s, beta, L, nb, dims, n, gtr, out, bd, vinv, nonneg);
    variant = 0;
  }
  if (variant == 5 || variant == 6) {  // merged orbital pairs: the Hermitian form at n = 3 only
    if (herm && n == 3 && bd.P == 128) {
      if (variant == 6)  // inner DFT split 8 x 16 instead of 16 x 8
        return launch_one_h<3, 8, 16, 15, 256, true, false, 1, true, true>(s, beta, L, nb, dims, gtr, out, bd, vinv,
                                                                            m_lo);
      return launch_one_h<3, 16, 8, 15, 256, true, false, 1, true, true>(s, beta, L, nb, dims, gtr, out, bd, vinv, m_lo);
    }
    if (herm && n == 3 && bd.P == 64)
      return launch_one_h<3, 8, 8, 15, 256, true, false, 1, true, true>(s, beta, L, nb, dims, gtr, out, bd, vinv, m_lo);
    variant = 0;
  }
#define TPRF_BP(N_, R1_, R2_, CG_, NT_, AB_, DMA_, WPE_) \
  return launch_one<N_, R1_, R2_, CG_, NT_, AB_, DMA_, WPE_>(s, beta, L, nb, dims, gtr, out, bd, vinv, herm, m_lo)
  if (n == 3 && bd.P == 128) {
    if (variant == 0) TPRF_BP(3, 16, 8, 9, 256, true, false, 1);
    if (variant == 1) TPRF_BP(3, 16, 8, 9, 128, true, false, 1);
    if (variant == 3) TPRF_BP(3, 16, 8, 9, 256, true, true, 3);
    if (variant == 4) TPRF_BP(3, 16, 8, 9, 256, true, false, 3);
    TPRF_BP(3, 16, 8, 27, 256, false, false, 1);
  }
  if (n == 3 && bd.P == 64) {
    if (variant == 0) TPRF_BP(3, 8, 8, 9, 256, true, false, 1);
    if (variant == 1) TPRF_BP(3, 8, 8, 9, 128, true, false, 1);
    if (variant == 3) TPRF_BP(3, 8, 8, 9, 128, true, true, 1);
    if (variant == 4) TPRF_BP(3, 8, 8, 9, 192, true, true, 1);
    TPRF_BP(3, 8, 8, 27, 256, false, false, 1);
  }
  if (n == 2 && bd.P == 128) {
    if (variant == 0) TPRF_BP(2, 16, 8, 4, 128, true, false, 1);
    if (variant == 1) TPRF_BP(2, 16, 8, 8, 128, false, false, 1);
    if (variant == 3) TPRF_BP(2, 16, 8, 4, 128, true, true, 1);
    if (variant == 4) TPRF_BP(2, 16, 8, 8, 128, false, true, 1);
    TPRF_BP(2, 16, 8, 16, 256, false, false, 1);
  }
  if (n == 2 && bd.P == 64) {
    if (variant == 0) TPRF_BP(2, 8, 8, 4, 64, true, false, 1);
    if (variant == 1) TPRF_BP(2, 8, 8, 4, 128, true, false, 1);
    if (variant == 3) TPRF_BP(2, 8, 8, 4, 64, true, true, 1);
    if (variant == 4) TPRF_BP(2, 8, 8, 4, 128, true, true, 1);
    TPRF_BP(2, 8, 8, 16, 256, false, false, 1);
  }
#undef TPRF_BP
  return hipErrorNotSupported;
}

}  // namespace tprf
