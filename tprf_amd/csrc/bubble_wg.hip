// bubble_wg.hip -- fused bubble product + bosonic tau -> nu transform
// (chi0_wr_from_grt_PH, chi_imtime.cpp:139-182, with the tau -> nu math of
// fourier_matsubara.cpp:42-184) for a Hermitian G(tau, r), one workgroup per
// canonical pair (r, -r) owning every canonical column.
//
// Same mathematics as k_bubble_pair<..., HERM> (bubble.hip): for the canonical
// columns abcd (a < b, or a == b and c <= d) the products
//   x_abcd(tau_j) = G_da(tau_j, r) conj(G_cb(tau_{L-j}, r))
// are summed as S(m) = sum_{j<L} x_j exp(2 pi i j m / L) with j = N1 j2 + j1,
// a length-P = 128 DFT over j2 (8 x 16: a DFT-8 per lane, then the DFT-16 over
// lane pairs) twiddled by exp(2 pi i j1 m / L) and accumulated over j1; the
// other columns of r and of -r follow from the conjugation and time-reversal
// identities, and the pole model is applied at the outputs.
//
// Shape: wave w owns the canonical columns 4w .. 4w + 3 (lane = 16 * column
// slot + s in the first stage, 16 * slot + 2 k1 + h in the second), so every
// lane is busy in both DFT stages and the stage-1 -> stage-2 transpose is
// private to the wave (no workgroup barrier between the stages).  The tau rows
// of step j1 (A side, j = N1 j2 + j1) are the B-side rows of step N1 - j1 and
// vice versa, so the steps run in pairs on one staged tile: every row of
// G(tau, r) is read from memory once per workgroup, and there is one
// workgroup barrier per step.  Results go through LDS and are written as whole
// n^4-column rows of chi0(nu, r) and chi0(nu, -r).
#include "common.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

namespace {

constexpr int kWgP = 128;   // inner DFT length
constexpr int kTwRow = 128; // twiddle row: bin k at k
constexpr int kSRow = 129;  // padded row of the per-column sums in the epilogue

template <int N>
struct Canon {
  static constexpr int n2 = N * N;
  static constexpr int noff = N * (N - 1) / 2 * n2;  // columns of the pairs a < b
  static constexpr int per = N * (N + 1) / 2;        // columns c <= d of a pair a == b
  static constexpr int NC = noff + N * per;
  static constexpr int NW = (NC + 3) / 4;            // waves of four columns
};

// canonical column index -> (a, b, c, d): first the pairs a < b in row-major
// order with every (c, d), then the pairs a == b with c <= d in row-major order
template <int N>
__device__ __forceinline__ void canon_col(int col, int &a, int &b, int &c, int &d) {
  constexpr int n2 = N * N, noff = Canon<N>::noff, per = Canon<N>::per;
  if (col < noff) {
    int p = col / n2;
    const int rest = col - p * n2;
    a = 0;
    while (p >= N - 1 - a) { p -= N - 1 - a; ++a; }
    b = a + 1 + p;
    c = rest / N;
    d = rest - c * N;
  } else {
    int rest = col - noff;
    a = rest / per;
    b = a;
    rest -= a * per;
    c = 0;
    while (rest >= N - c) { rest -= N - c; ++c; }
    d = c + rest;
  }
}

// inverse of canon_col for a canonical (a, b, c, d)
template <int N>
__device__ __forceinline__ int canon_index(int a, int b, int c, int d) {
  constexpr int n2 = N * N, noff = Canon<N>::noff, per = Canon<N>::per;
  if (a < b) return (a * (N - 1) - a * (a - 1) / 2 + (b - a - 1)) * n2 + c * N + d;
  return noff + a * per + c * N - c * (c - 1) / 2 + (d - c);
}

// x, opaque to the optimiser: addresses computed from it are formed where they
// are used instead of being hoisted out of the loops and held in registers
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// LDS of one workgroup, in c2 units
template <int N>
struct WgLds {
  static constexpr int n2 = N * N, n4 = n2 * n2;
  static constexpr int NC = Canon<N>::NC, NW = Canon<N>::NW;
  static constexpr int tile = 0;                         // [2][P][n2]: slots T1, T2
  static constexpr int tw = tile + 2 * kWgP * n2;        // [2 buffers][2 steps][kTwRow]
  static constexpr int st = tw + 4 * kTwRow;             // [NW][32][16] wave-private transposes
  static constexpr int pole = st + NW * 32 * 16;         // [NC][7]: a(x), a(reversed x), x_L - x_0
  static constexpr int twp = pole + NC * 7;              // [8][16] exp(2 pi i s k1 / P) at [k1][s]
  static constexpr int hmod = twp + kWgP;                // [3][nnu], nnu = 2 nb - 1 <= P - 1
  static constexpr int cmap = hmod + 3 * (kWgP - 1);     // [2][n^4] ints: 4 canonical column + variant
  static constexpr int cpole = cmap + (2 * n4 + 3) / 4;  // [3][18] doubles: pole coefficients from x_j
  static constexpr int braw = cpole + 28;                // [18][n2]: G(tau_j, r), j = 0..8, L-8..L
  static constexpr int total = braw + 18 * n2;
  // between two pairs: the sums [NC][kSRow] in the transpose region
  static constexpr int sums = st;
  static_assert(NC * kSRow <= NW * 32 * 16, "sums fit the transpose region");
};

}  // namespace

// Persistent: gridDim.x is a multiple of 8 and the workgroups of XCD x (blockIdx
// mod 8) take the x-th eighth of the pairs in rounds of gridDim.x / 8
// consecutive pairs, so that pairs whose rows share cache lines run at the same
// time on one L2.  While a pair's results are written, the first tile of the
// workgroup's next pair is in flight.
template <int N>
__global__ void __launch_bounds__(64 * Canon<N>::NW)
    k_bubble_wg(double beta, int L, int nb, int d0, int d1, int d2, const c2 *__restrict__ gtr, c2 *__restrict__ out,
                BubbleDev bd, Vinv8 vinv, int m_lo) {
  constexpr int P = kWgP;
  constexpr int n2 = N * N, n4 = n2 * n2;
  constexpr int NC = Canon<N>::NC, NW = Canon<N>::NW, NT = 64 * NW;
  using Lay = WgLds<N>;
  constexpr int TILE = 2 * P * n2;
  constexpr int PFT = (TILE + NT - 1) / NT;      // tile entries per thread
  constexpr int PFW = (2 * P + NT - 1) / NT;     // twiddle entries per thread
  constexpr int NBR = 18 * n2;                   // boundary rows' entries
  static_assert(NBR <= NT, "one boundary entry per thread");
  constexpr int GO = NT / n4;                    // output row groups
  static_assert(GO >= 1, "one output row per thread group");
  __shared__ c2 lds[Lay::total];

  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const long nr = (long)d0 * d1 * d2;
  // this workgroup's units: the pairs xstart + wslot + i W, i < R, of its
  // XCD's full rounds; the lo pairs left over are split by waves into KP
  // parts each (part k: the columns of waves [k NW / KP, (k + 1) NW / KP)),
  // taken by the first lo KP workgroups of the XCD, so that the last round
  // costs a fraction of a pair instead of a whole one
  const long npair = bd.ncanon, q8 = npair >> 3, rem8 = npair & 7;
  const int xcd = blockIdx.x & 7, wslot = blockIdx.x >> 3, W = gridDim.x >> 3;
  const long xstart = xcd * q8 + (xcd < rem8 ? xcd : rem8), xcount = q8 + (xcd < rem8 ? 1 : 0);
  // (uniform: held in scalar registers)
  const int R = __builtin_amdgcn_readfirstlane((int)(xcount / W));
  const int lo = __builtin_amdgcn_readfirstlane((int)(xcount - (long)R * W));
  const int KP = __builtin_amdgcn_readfirstlane(lo ? min(NW, W / lo) : 1);
  const int nunit = __builtin_amdgcn_readfirstlane(R + ((lo && wslot < lo * KP) ? 1 : 0));
  if (nunit == 0) return;
  auto unit_pair = [&](int i) { return i < R ? xstart + wslot + (long)i * W : xstart + (long)R * W + wslot / KP; };

  const int N1 = L / P;
  const int ntile = N1 / 2 + 1;                  // tiles j1 = 0 .. N1 / 2
  const double dtau = beta / (double)L;
  const int nnu = 2 * nb - 1;

  // this lane's column (both stages) and its entries: A = G_da, B = G_cb
  const int cw = lane >> 4;
  const int col_raw = wave * 4 + cw;
  const bool col_ok = col_raw < NC;
  const int col = col_ok ? col_raw : NC - 1;
  int ca, cb, cc, cd;
  canon_col<N>(col, ca, cb, cc, cd);
  const int ea = cd * N + ca, eb = cc * N + cb;
  const int s = lane & 15;                       // stage 1: j2 = 16 k + s
  const int k1 = (lane >> 1) & 7, h = lane & 1;  // stage 2: bins k1 + 8 (2 k' + h)

  // ---- tables shared by every pair
  // stage-1 twiddles exp(2 pi i s k1 / P) at [k1][s]: the 16 lanes of one
  // read touch 16 consecutive entries (ds_read_b128 serves 16 lanes at a time,
  // conflict-free when their 16-byte pieces differ modulo 256 bytes)
  for (int p = tid; p < P; p += NT) st_v2d(lds + Lay::twp + p, ld_v2d(bd.twp + (((p >> 4) * (p & 15)) & (P - 1))));
  for (int f = tid; f < 3 * nnu; f += NT) st_v2d(lds + Lay::hmod + f, ld_v2d(bd.hmod + f));
  // The bosonic pole coefficients of a column (fourier_matsubara.cpp:42-97,
  // :274-288) are linear in its boundary differences dl_m = x_m - x_0,
  // dr_m = x_L - x_{L-m} (m = 1..8) and x_L - x_0:
  //   a_i = w1_i (x_L - x_0) + sum_m (wl_im dl_m + wr_im dr_m),
  // with m2 = sum_m V0_m (dl_m - dr_m) / tau_m, m3 = -(2 / dtau) sum_m V1_m
  // (dl_m + dr_m) / tau_m and a = (4/3 (m1 - m3), m3 - (m1 + m2) / 2,
  // m1 / 6 + m2 / 2 + m3 / 3).  Row i of cpl: w1_i, wl_i1..wl_i8, wr_i1..wr_i8.
  double *cpl = reinterpret_cast<double *>(lds + Lay::cpole);
  if (tid < 3) {
    const int i = tid;
    const double d1 = i == 0 ? 4.0 / 3.0 : (i == 1 ? -0.5 : 1.0 / 6.0);
    const double d2 = i == 0 ? 0.0 : (i == 1 ? -0.5 : 0.5);
    const double d3 = i == 0 ? -4.0 / 3.0 : (i == 1 ? 1.0 : 1.0 / 3.0);
    cpl[i * 17] = d1;
    for (int m = 1; m <= 8; ++m) {
      const double itm = 1.0 / ((double)m * dtau);
      const double g2 = d2 * vinv.v[m - 1] * itm, g3 = d3 * (-2.0 / dtau) * vinv.v[8 + m - 1] * itm;
      cpl[i * 17 + m] = g2 + g3;
      cpl[i * 17 + 8 + m] = g3 - g2;
    }
  }
  // column map: side 0 (r) and 1 (-r), q = wxyz -> 4 canonical index + variant
  // (0: abcd of r; 1: cdab of -r; 2: badc of r; 3: dcba of -r)
  {
    int *cmap = reinterpret_cast<int *>(lds + Lay::cmap);
    for (int f = tid; f < 2 * n4; f += NT) {
      const int side = f / n4, q = f - side * n4;
      const int w = q / (n2 * N), x = (q / n2) % N, y = (q / N) % N, z = q % N;
      int code;
      if (side == 0)
        code = (w < x || (w == x && y <= z)) ? 4 * canon_index<N>(w, x, y, z) : 4 * canon_index<N>(x, w, z, y) + 2;
      else
        code = (y < z || (y == z && w <= x)) ? 4 * canon_index<N>(y, z, w, x) + 1 : 4 * canon_index<N>(z, y, x, w) + 3;
      cmap[f] = code;
    }
  }

  // The boundary rows G(tau_j, r), j = 0..8 and L-8..L (row e: j = e or
  // L - 17 + e), pass through the staged tiles: tau = N1 j2 + t is row j2 of
  // slot 0 of tile t when t = tau mod N1 <= N1 / 2, else tau = N1 (j2 + 1) - t
  // is row j2 of slot 1 of tile t = -tau mod N1.  Thread e n2 + ent copies its
  // entry out of tile bsrc >> 16, LDS element bsrc & 0xffff.
  int bsrc = -1;
  if (tid < NBR) {
    const int e = tid / n2, ent = tid - e * n2;
    const int tau = e < 9 ? e : L - 17 + e;
    const int r0 = tau % N1;
    int tt, el;
    if (r0 <= N1 / 2 && tau / N1 < P) {
      tt = r0;
      el = (tau / N1) * n2 + ent;
    } else {
      tt = (N1 - r0) % N1;
      el = (P + (tau + tt) / N1 - 1) * n2 + ent;
    }
    bsrc = (tt << 16) | el;
  }

  // tile t: slot 0 holds G(N1 j2 + t), slot 1 G(N1 (j2 + 1) - t), j2 < P; the
  // twiddle rows of its steps t and N1 - t
  v2d pf[PFT], pt[PFW];
  auto load_tile = [&](long rr, int t) {
    const int tt = opaque(tid);
#pragma unroll
    for (int i = 0; i < PFT; ++i) {
      const int e = min(tt + i * NT, TILE - 1);
      const int slot = e / (P * n2), rem = e - slot * (P * n2);
      const int j2 = rem / n2, ent = rem - j2 * n2;
      const long row = slot == 0 ? (long)N1 * j2 + t : (long)N1 * (j2 + 1) - t;
      pf[i] = ld_v2d(gtr + (row * nr + rr) * n2 + ent);
    }
#pragma unroll
    for (int i = 0; i < PFW; ++i) {
      const int e = min(tt + i * NT, 2 * P - 1);
      const int which = e / P, k = e - which * P;
      const int j1 = which == 0 ? t : (N1 - t) % N1;
      pt[i] = ld_v2d(bd.tw1 + (long)j1 * P + k);
    }
  };
  auto store_tile = [&](int t) {
    const int tt = opaque(tid);
#pragma unroll
    for (int i = 0; i < PFT; ++i) {
      const int e = tt + i * NT;
      if (e < TILE) st_v2d(lds + Lay::tile + e, pf[i]);
    }
#pragma unroll
    for (int i = 0; i < PFW; ++i) {
      const int e = tt + i * NT;
      if (e < 2 * P) {
        const int which = e / P, k = e - which * P;
        st_v2d(lds + Lay::tw + ((t & 1) * 2 + which) * kTwRow + k, pt[i]);
      }
    }
  };

  c2 acc[8];
  c2 *const sT = lds + Lay::st + wave * (32 * 16);

  // stage 1: lane (cw, s): DFT-8 over k of x at j2 = 16 k + s, twiddled by
  // exp(2 pi i s k1 / P), into the wave's transpose row (cw, k1); rows are
  // stored with their 16 entries permuted by s ^ (2 k1), so that the stage-2
  // reads (lanes (k1, h) of one row group each take s = 2 q + h) are
  // conflict-free too
  auto stage1 = [&](int slot_a) {
    __builtin_amdgcn_sched_barrier(0);
    // A at row j2 = 16 k + s, B at row P - 1 - j2 = (15 - s) + 16 (7 - k):
    // one base address per side and constant offsets
    const c2 *TA = lds + Lay::tile + slot_a * (P * n2) + s * n2 + ea;
    const c2 *TB = lds + Lay::tile + (1 - slot_a) * (P * n2) + (15 - s) * n2 + eb;
    // the 16 reads are issued back to back and waited for once (the empty asm
    // takes them all as operands)
    v2d ra[8], rb[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      ra[k] = ld_v2d(TA + 16 * k * n2);
      rb[k] = ld_v2d(TB + 16 * (7 - k) * n2);
    }
    asm volatile("" : "+v"(ra[0]), "+v"(ra[1]), "+v"(ra[2]), "+v"(ra[3]), "+v"(ra[4]), "+v"(ra[5]), "+v"(ra[6]),
                 "+v"(ra[7]), "+v"(rb[0]), "+v"(rb[1]), "+v"(rb[2]), "+v"(rb[3]), "+v"(rb[4]), "+v"(rb[5]),
                 "+v"(rb[6]), "+v"(rb[7]));
    c2 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = cmulc(c2{ra[k].x, ra[k].y}, c2{rb[k].x, rb[k].y});
    const c2 *tw = lds + Lay::twp + s;
    v2d rt[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) rt[q] = ld_v2d(tw + 16 * (q + 1));
    SmallDft<8, 1>::run(v, nullptr, 0);
    asm volatile("" : "+v"(rt[0]), "+v"(rt[1]), "+v"(rt[2]), "+v"(rt[3]), "+v"(rt[4]), "+v"(rt[5]), "+v"(rt[6]));
#pragma unroll
    for (int q = 1; q < 8; ++q) v[q] = cmul(v[q], c2{rt[q - 1].x, rt[q - 1].y});
#pragma unroll
    for (int q = 0; q < 8; ++q) sT[(cw * 8 + q) * 16 + (s ^ (2 * q))] = v[q];
    __builtin_amdgcn_sched_barrier(0);
  };
  // stage 2: lane pair (cw, k1, h = 0/1): the DFT-16 over s of row (cw, k1)
  // split by decimation in frequency: lane h forms u(s') = y(s') +- y(s' + 8),
  // s' < 8, twiddled by exp(2 pi i h s' / 16), whose DFT-8 gives the bins
  // kk = 2 k' + h; it keeps k1 + 8 kk and accumulates them with
  // exp(2 pi i j1 m / L)
  auto stage2 = [&](const c2 *twrow) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_sched_barrier(0);
    const c2 *row = sT + (cw * 8 + k1) * 16;
    v2d ry[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) ry[q] = ld_v2d(row + (q ^ (2 * k1)));
    asm volatile("" : "+v"(ry[0]), "+v"(ry[1]), "+v"(ry[2]), "+v"(ry[3]), "+v"(ry[4]), "+v"(ry[5]), "+v"(ry[6]),
                 "+v"(ry[7]), "+v"(ry[8]), "+v"(ry[9]), "+v"(ry[10]), "+v"(ry[11]), "+v"(ry[12]), "+v"(ry[13]),
                 "+v"(ry[14]), "+v"(ry[15]));
    const double sg = h ? -1.0 : 1.0;
    c2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = c2{fma(sg, ry[q + 8].x, ry[q].x), fma(sg, ry[q + 8].y, ry[q].y)};
    // the accumulation twiddles are read while the DFT-8 runs
    v2d rw[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) rw[q] = ld_v2d(twrow + k1 + 8 * h + 16 * q);
    if (h) {
      // exp(2 pi i s' / 16), s' = 1..7
      const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, hh = 0.70710678118654752440;
      const c2 w16[7] = {c2{c1, s1}, c2{hh, hh}, c2{s1, c1}, c2{0.0, 1.0}, c2{-s1, c1}, c2{-hh, hh}, c2{-c1, s1}};
#pragma unroll
      for (int q = 1; q < 8; ++q) v[q] = cmul(v[q], w16[q - 1]);
    }
    SmallDft<8, 1>::run(v, nullptr, 0);
#pragma unroll
    for (int q = 0; q < 8; ++q) acc[q] = cfma(acc[q], c2{rw[q].x, rw[q].y}, v[q]);
    __builtin_amdgcn_sched_barrier(0);
    // the next stage 1 of this wave overwrites the transpose rows read here
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };

  // first pair: boundary samples and tile 0 in flight together
  long r = bd.rcanon[unit_pair(0)];
  load_tile(r, 0);
  store_tile(0);
  if (ntile > 1) load_tile(r, 1);

  const double fact = beta / (double)L;
  const int im_lo = m_lo + nb - 1, nim = nnu - im_lo;
  for (int it = 0;; ++it) {
    const long mr = minus_r(r, d0, d1, d2);
    // the waves whose columns this unit transforms (all but in a part)
    const int part = wslot % KP;
    const int wlo = it < R ? 0 : part * NW / KP, whi = it < R ? NW : (part + 1) * NW / KP;
    const bool wact = wave >= wlo && wave < whi;
    // the next unit's r, loaded now: its load_tile in the epilogue does not
    // wait for a load just issued
    const bool more = it + 1 < nunit;
    const long rn = more ? bd.rcanon[unit_pair(it + 1)] : r;
#pragma unroll
    for (int i = 0; i < 8; ++i) acc[i] = c2{0.0, 0.0};
    // steps in tile order: j1 = 0, then (1, N1 - 1), (2, N1 - 2), ...; u = 0 is
    // the first step of tile t (A rows in slot 0), u = 1 the second (slot 1)
    int t = 0, u = 0;
    for (int step = 0; step < N1; ++step) {
      const bool two = t > 0 && 2 * t != N1;
      if (u == 0) {
        __syncthreads();  // tile t staged
        if ((bsrc >> 16) == t) lds[Lay::braw + tid] = lds[Lay::tile + (bsrc & 0xffff)];
      }
      if (wact) stage1(u);
      const bool last = !two || u == 1;
      if (last) {
        __syncthreads();  // every wave is done with tile t
        if (t + 1 < ntile) store_tile(t + 1);
        if (t + 2 < ntile) load_tile(r, t + 2);
      }
      if (wact) stage2(lds + Lay::tw + ((t & 1) * 2 + u) * kTwRow);
      if (last) {
        ++t;
        u = 0;
      } else {
        u = 1;
      }
    }
    __syncthreads();  // tiles and transposes free

    // ---- this pair's sums and boundary products to LDS; the next pair's
    // boundary samples and tile 0 in flight during the outputs
    const int te = opaque(tid);  // epilogue indices are formed here, not held through the steps
    if (col_ok && wact) {
#pragma unroll
      for (int q = 0; q < 8; ++q) lds[Lay::sums + col * kSRow + k1 + 8 * h + 16 * q] = acc[q];
    }
    if (more) load_tile(rn, 0);
    __syncthreads();
    // pole coefficients: thread 6 c + 3 d + i forms a_i of column c from its
    // boundary products x_e = G_da(tau_e, r) conj(G_cb(tau_{L-e}, r)) (e = j
    // for j <= 8, e = 17 - (L - j) above), d = 0; for the reversed x (d = 1)
    // dl and dr trade places and every difference changes sign
    if (te < 6 * NC && te / 24 >= wlo && te / 24 < whi) {  // column te / 6 of an active wave
      const int c = te / 6, d = (te / 3) & 1, i = te % 3;
      int a, b, cc2, dd2;
      canon_col<N>(c, a, b, cc2, dd2);
      const c2 *br = lds + Lay::braw;
      c2 x[18];
#pragma unroll
      for (int e = 0; e < 18; ++e) x[e] = cmulc(br[e * n2 + dd2 * N + a], br[(17 - e) * n2 + cc2 * N + b]);
      const double *ci = cpl + i * 17;
      const c2 m1 = csub(x[17], x[0]);
      c2 sum = c2{ci[0] * m1.x, ci[0] * m1.y};
#pragma unroll
      for (int m = 1; m <= 8; ++m) {
        const c2 dl = csub(x[m], x[0]), dr = csub(x[17], x[17 - m]);
        const double wa = ci[d ? 8 + m : m], wb = ci[d ? m : 8 + m];
        sum = c2{fma(wa, dl.x, fma(wb, dr.x, sum.x)), fma(wa, dl.y, fma(wb, dr.y, sum.y))};
      }
      c2 *pc = lds + Lay::pole + c * 7;
      pc[3 * d + i] = d ? c2{-sum.x, -sum.y} : sum;
      if (d == 0 && i == 0) pc[6] = m1;
    }
    // next pair: tile 0 staged (the tile region is unused since the steps
    // ended) and tile 1 in flight before the output stores: vmcnt counts loads
    // and stores in issue order, so a wait for a load issued after the output
    // stores would also wait for them
    if (more) {
      store_tile(0);
      if (ntile > 1) load_tile(rn, 1);
    }
    __syncthreads();

    // outputs: thread (g, q) writes column q of the rows im_lo + g, + GO, ...
    // of r (and of -r): whole rows of n^4 columns per group of threads
    if (te < GO * n4) {
      const int q = te % n4, g = te / n4;
      const int *cmap = reinterpret_cast<const int *>(lds + Lay::cmap);
      const c2 *sH = lds + Lay::hmod;
      for (int side = 0; side < (mr != r ? 2 : 1); ++side) {
        const int code = cmap[side * n4 + q];
        const int c = code >> 2, var = code & 3;
        if (c / 4 < wlo || c / 4 >= whi) continue;  // a column of another part
        const bool addx = var == 1 || var == 2, cj = var >= 2;
        const c2 *pc = lds + Lay::pole + c * 7;
        const c2 xa = addx ? pc[6] : c2{0.0, 0.0};
        const c2 *pa = pc + (addx ? 3 : 0);
        const double sgn = cj ? -1.0 : 1.0;
        const c2 p0 = c2{pa[0].x, sgn * pa[0].y}, p1 = c2{pa[1].x, sgn * pa[1].y}, p2 = c2{pa[2].x, sgn * pa[2].y};
        const c2 *S = lds + Lay::sums + c * kSRow;
        const long ostep = (long)GO * nr * n4;
        c2 *o = out + ((long)(im_lo + g) * nr + (side ? mr : r)) * n4 + q;
        const int msg = side ? -1 : 1;
#pragma unroll 4
        for (int row = g; row < nim; row += GO, o += ostep) {
          const int im = im_lo + row;
          const c2 sv = S[(msg * (im - (nb - 1))) & (P - 1)];
          const c2 sx = c2{(sv.x + xa.x) * fact, sgn * (sv.y + xa.y) * fact};
          *o = cfma(cfma(cfma(sx, p0, sH[im]), p1, sH[nnu + im]), p2, sH[2 * nnu + im]);
        }
      }
    }
    if (!more) break;
    // the first step's barrier orders the output reads of the sums before the
    // transposes are overwritten
    r = rn;
  }
}

// Hermitian form of the fused bubble with one workgroup per canonical pair
// (r, -r) at a time; n = 2 or 3 and an inner DFT length of 128
// (hipErrorNotSupported otherwise).  The grid is one workgroup per CU (a
// multiple of 8, at most one per pair and XCD slot).
hipError_t launch_bubble_wg(hipStream_t s, double beta, int L, int nb, const int dims[3], int n, const c2 *gtr,
                            c2 *out, const BubbleDev &bd, const Vinv8 &vinv, bool nonneg) {
  if (bd.P != kWgP || L % kWgP != 0 || 2 * nb > kWgP) return hipErrorNotSupported;
  const int m_lo = nonneg ? 0 : 1 - nb;
  static int cus[64] = {0};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!cus[dev]) {
    int cu = 0;
    e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    cus[dev] = cu > 8 ? cu : 8;
  }
  const long per_xcd = ((long)bd.ncanon + 7) / 8;
  const long w = per_xcd < cus[dev] / 8 ? per_xcd : cus[dev] / 8;
  const dim3 grid((unsigned)(8 * (w > 0 ? w : 1)));
  if (n == 3) {
    hipLaunchKernelGGL(k_bubble_wg<3>, grid, dim3(64 * Canon<3>::NW), 0, s, beta, L, nb, dims[0], dims[1], dims[2], gtr,
                       out, bd, vinv, m_lo);
    return hipGetLastError();
  }
  if (n == 2) {
    hipLaunchKernelGGL(k_bubble_wg<2>, grid, dim3(64 * Canon<2>::NW), 0, s, beta, L, nb, dims[0], dims[1], dims[2], gtr,
                       out, bd, vinv, m_lo);
    return hipGetLastError();
  }
  return hipErrorNotSupported;
}

}  // namespace tprf
