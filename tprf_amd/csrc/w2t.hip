// w2t.hip -- Matsubara -> imaginary-time transform with a pruned input
// (fourier_wr_to_tr, fourier_matsubara.cpp:218-308).
//
// The length-L input of the reference's FFT is nonzero only at the 2 n_max
// positions (n + L) % L of the mesh.  With L = Q M and M >= 2 n_max a power
// of two, the outputs split by residue t = j mod Q:
//   y_{Q j' + t} = sum_n [x_n exp(-2 pi i t n / L)] exp(-2 pi i j' n / M),
// one length-M DFT per t over the mesh (n taken mod M).  A workgroup holds C
// columns (C * 16 consecutive bytes per row); each column's DFT runs on M/8
// threads, eight values per thread, radix-8 Stockham passes exchanged through
// LDS.
#include "common.hpp"
#include "lds_dma.hpp"
#include "fft_device.hpp"
#include "mats_device.hpp"

namespace tprf {

template <int M>
struct W2tPlan {
  static constexpr int n8 = (M >= 512) ? 3 : (M >= 64 ? 2 : 1);  // radix-8 passes
  static constexpr int last = M / (n8 == 3 ? 512 : (n8 == 2 ? 64 : 8));  // 1, 2, 4 or 8: final pass radix
  static constexpr int npass = n8 + (last > 1 ? 1 : 0);
};

// Twiddle by W_M^{k r (M / (Ns R))}, then the radix-R DFT, for B = 8/R butterflies held in v.
template <int M, int R, int S>
__device__ __forceinline__ void w2t_butterflies(c2 *v, int j, int Ns, const c2 *__restrict__ twM) {
  constexpr int B = 8 / R;
#pragma unroll
  for (int u = 0; u < B; ++u) {
    const int b = j * B + u;
    const int k = b % Ns;
    if (k) {
      const int step = k * (M / (Ns * R));
#pragma unroll
      for (int r = 1; r < R; ++r) v[u * R + r] = cmul(v[u * R + r], twiddle<S>(twM, step * r));
    }
    SmallDft<R, S>::run(v + u * R, nullptr, 0);
  }
}

constexpr int kW2tThreads = 256;

// x, opaque to the optimiser: what is computed from it is computed where it is
// used instead of being hoisted out of loops and held in registers
__device__ __forceinline__ int opaque32(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ long opaque64(long x) {
  asm volatile("" : "+v"(x));
  return x;
}

// NT threads per workgroup: C = NT / (M / 8) columns, i.e. C * 16 contiguous
// bytes of every row written per workgroup.
// Fermionic M = 256, 512 with 256 threads: registers held to the three
// workgroups per CU that the LDS allows (they fit without spilling); two
// waves per SIMD otherwise.
// SYM: only the columns sym.col[] are transformed (tiles over that list), and
// each result is also written, conjugated, to column sym.mcol[] (when >= 0):
// for G_ab(i w_n, r)^* = G_ba(i w_{-n-1}, -r) the transform of column
// (-r, b, a) is the conjugate of that of (r, a, b).
template <int STAT, int M, int NT, bool SYM = false>
__global__ void __launch_bounds__(NT)
    __attribute__((amdgpu_waves_per_eu((STAT == 0 && (M == 256 || M == 512) && NT == 256) ? 3 : 2)))
    k_w2t_pruned(double beta, int n_max, long ncol, const c2 *__restrict__ in, c2 *__restrict__ out,
                 const c2 *__restrict__ moments, TailFitDev fit, MatsDev mt, W2tDev wd, double *status,
                 SymCols sym = SymCols{nullptr, nullptr, 0}, int mom_in_out = 0) {
  constexpr int TPC = M / 8;  // threads per column
  constexpr int C = NT / TPC;  // columns per workgroup
  // C >= 3: the output records of a residue (3 M entries) fit in buf
  constexpr bool OTAB = C >= 3;
  using Plan = W2tPlan<M>;
  __shared__ c2 buf[M * C];
  __shared__ c2 s_a[C][3];
  __shared__ c2 s_m1[C];
  __shared__ c2 s_tw[M];

  const int tid = threadIdx.x;
  const int c = tid % C, j = tid / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long c0 = tile * C;
  const long ntot = SYM ? sym.n : ncol;
  const int cn = (ntot - c0) < C ? (int)(ntot - c0) : C;
  const long cidx = c0 + (c < cn ? c : 0);
  const long col = SYM ? (long)sym.col[cidx] : cidx;
  const long mcol = SYM ? (long)sym.mcol[cidx] : -1;
  const int L = mt.L, Q = wd.Q;
  const int nmesh = (STAT == 0) ? 2 * n_max : 2 * n_max - 1;
  const int first = (STAT == 0) ? -n_max : -(n_max - 1);

  // 1. this thread's eight inputs (the same FFT positions e = j + TPC r for
  //    every t), held in registers.  All loads (inputs and the twiddle table)
  //    are issued before the first LDS store (outside the mesh a valid entry
  //    is loaded and discarded).
  constexpr int NTW = (M + NT - 1) / NT;
  v2d twv[NTW], rv[8];
#pragma unroll
  for (int w = 0; w < NTW; ++w) twv[w] = ld_v2d(wd.twM + min(tid + w * NT, M - 1));
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int e = j + TPC * r;
    const int n = e < M / 2 ? e : e - M;
    const bool in_mesh = n >= first && n < first + nmesh;
    rv[r] = ld_v2d(in + (in_mesh ? (long)(n - first) * ncol + col : col));
  }
#pragma unroll
  for (int w = 0; w < NTW; ++w)
    if (tid + w * NT < M) st_v2d(s_tw + tid + w * NT, twv[w]);
  c2 raw[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int e = j + TPC * r;
    const int n = e < M / 2 ? e : e - M;
    raw[r] = (n >= first && n < first + nmesh) ? c2{rv[r].x, rv[r].y} : c2{0.0, 0.0};
  }
  // 2. tail moments: given, or fitted by k_tail_moments and left in the
  //    column's rows L - 2 .. L
  if (tid < cn) {
    c2 m1, m2, m3;
    if (mom_in_out) {
      // k_tail_moments left this column's moments in its rows L - 2 .. L,
      // which only this workgroup writes, and only after this read
      const long cc = SYM ? (long)sym.col[c0 + tid] : c0 + tid;
      const c2 *mo = out + (long)(L - 2) * ncol + cc;
      m1 = mo[0];
      m2 = mo[ncol];
      m3 = mo[2 * ncol];
    } else {
      m1 = moments[c0 + tid];
      m2 = moments[ncol + c0 + tid];
      m3 = moments[2 * ncol + c0 + tid];
    }
    Poles<STAT>::coef(m1, m2, m3, s_a[tid]);
    s_m1[tid] = m1;
  }
  __syncthreads();
  const c2 a0 = s_a[c][0], a1 = s_a[c][1], a2 = s_a[c][2];
  const double ib = 1.0 / beta;
  const double pm = (STAT == 0) ? -1.0 : 1.0;

  // 3. The pole model is not subtracted from the inputs: by linearity
  //      phase DFT[(G - sum_q a_q pole_q) / beta] + sum_q a_q psi_q
  //    = phase DFT[G / beta] + sum_q a_q D_q(tau),
  //    D_q = psi_q - (1/beta) sum_{n in mesh} exp(-i w_n tau) pole_q(n) (the
  //    real table wd.dtail, long double on the host), so only the scale is
  //    applied here and the model enters at the outputs.
#pragma unroll
  for (int r = 0; r < 8; ++r) raw[r] = cscale(raw[r], ib);

  for (int t = 0; t < Q; ++t) {
    // pass 1 inputs: FFT index e = j + TPC r  <->  n = e or e - M, twiddled by
    // exp(-2 pi i t n / L).  The eight table loads are unconditional (every
    // index t n mod L is in the table) so they are all in flight together.
    c2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) v[r] = raw[r];
    if (t) {
      v2d tw[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = j + TPC * r;
        const int n = e < M / 2 ? e : e - M;
        int tn = t * n;  // |t n| < Q M / 2 <= L
        if (tn < 0) tn += L;
        tw[r] = ld_v2d(wd.twL + tn);
      }
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const int e = j + TPC * r;
        const int n = e < M / 2 ? e : e - M;
        if (n >= first && n < first + nmesh) v[r] = cmul(raw[r], c2{tw[r].x, tw[r].y});
      }
    }
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < Plan::npass; ++p) {
      const int R = (p < Plan::n8) ? 8 : Plan::last;
      if (OTAB && p + 1 == Plan::npass) {
        // buf is free (every thread has read the last exchange): copy this
        // residue's output records otab[Q jp + t], jp < M, straight into it
        // while the last pass computes (64-record chunks per wave)
        const int wave = tid >> 6, lane = tid & 63;
        for (int ch = wave; ch < 3 * M / 64; ch += NT / 64) {
          const int e = ch * 64 + lane, jp = e / 3, part = e - 3 * jp;
          dma16(wd.otab + ((long)(Q * jp + t) * 3 + part), buf + ch * 64);
        }
      }
      if (R == 8) w2t_butterflies<M, 8, -1>(v, j, Ns, s_tw);
      else if (R == 4) w2t_butterflies<M, 4, -1>(v, j, Ns, s_tw);
      else w2t_butterflies<M, 2, -1>(v, j, Ns, s_tw);
      const int B = 8 / R;
      if (p + 1 < Plan::npass) {
        // write this pass's outputs, read the next pass's inputs
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / R, r = u % R;
          const int b = j * B + uu;
          const int k = b % Ns;
          buf[((b - k) * R + k + r * Ns) * C + c] = v[u];
        }
        __syncthreads();
        Ns *= R;
        const int Rn = (p + 1 < Plan::n8) ? 8 : Plan::last;
        const int Bn = 8 / Rn;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / Rn, r = u % Rn;
          v[u] = buf[(j * Bn + uu + r * (M / Rn)) * C + c];
        }
        __syncthreads();
      } else if (OTAB) {
        // outputs: bin j' = o + r Ns  ->  tau index Q j' + t, with the records
        // [phase, (D_0, D_1), (D_2, 0)] of tau_{Q j' + t} read from LDS
        wait_vmcnt<0>();
        __syncthreads();
        if (c < cn) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int uu = u / R, r = u % R;
            const int b = j * B + uu;
            const int k = b % Ns;
            const int jp = (b - k) * R + k + r * Ns;
            const int jt = Q * jp + t;
            const c2 ph = buf[3 * jp], d01 = buf[3 * jp + 1], d2 = buf[3 * jp + 2];
            c2 y = v[u];
            if (STAT == 0) y = cmul(y, ph);
            y.x += a0.x * d01.x + a1.x * d01.y + a2.x * d2.x;
            y.y += a0.y * d01.x + a1.y * d01.y + a2.y * d2.x;
            out[(long)jt * ncol + col] = y;
            if (jt == 0) out[(long)L * ncol + col] = cscale(cadd(y, s_m1[c]), pm);
            if (SYM && mcol >= 0) {
              out[(long)jt * ncol + mcol] = cconj(y);
              if (jt == 0) out[(long)L * ncol + mcol] = cconj(cscale(cadd(y, s_m1[c]), pm));
            }
          }
        }
        __syncthreads();  // the next residue's first exchange overwrites buf
      } else if (c < cn) {
        // outputs: bin j' = o + r Ns  ->  tau index Q j' + t
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / R, r = u % R;
          const int b = j * B + uu;
          const int k = b % Ns;
          const int jp = (b - k) * R + k + r * Ns;
          const int jt = Q * jp + t;
          c2 y = v[u];
          if (STAT == 0) y = cmul(y, mt.phase[jt]);
          const double p0 = wd.dtail[jt], p1 = wd.dtail[L + jt], p2 = wd.dtail[2 * L + jt];
          y.x += a0.x * p0 + a1.x * p1 + a2.x * p2;
          y.y += a0.y * p0 + a1.y * p1 + a2.y * p2;
          out[(long)jt * ncol + col] = y;
          if (jt == 0) out[(long)L * ncol + col] = cscale(cadd(y, s_m1[c]), pm);
          if (SYM && mcol >= 0) {
            out[(long)jt * ncol + mcol] = cconj(y);
            if (jt == 0) out[(long)L * ncol + mcol] = cconj(cscale(cadd(y, s_m1[c]), pm));
          }
        }
      }
    }
  }
}

// Wave-specialised persistent form of the symmetric transform above (fermions,
// the moments left in rows L - 2 .. L of each column by k_tail_moments, Q >= 2,
// M = 256 or 512).  gridDim.x workgroups (at most one per CU) take the tiles
// s, s + gridDim.x, ... of their slot s (XCD-aware order).  Waves 0 .. 7 (the
// transform) read only LDS inside the loop and write the outputs; waves 8 .. 11
// (the loader) copy each residue's output records into buf and stage the next
// tile's inputs (scaled by 1 / beta), columns and pole coefficients in LDS,
// spread over the current tile's residues.  vmcnt counts a wave's loads and
// stores together, in issue order: a load issued after stores can only be
// waited for together with them, and an output store takes microseconds to
// complete under this write load, so no wave here both stores and loads.
// Barriers leave vector memory in flight (barrier_keep_vm); both roles pass
// the same barriers: per residue two per exchange, one before the outputs
// (records in buf) and one after them (buf free).
constexpr int kW2tSpecLoad = 256;  // loader threads

template <int M>
__global__ void __launch_bounds__(512 + kW2tSpecLoad)
    k_w2t_spec(double beta, int n_max, long ncol, const c2 *__restrict__ in, c2 *__restrict__ out, MatsDev mt,
               W2tDev wd, SymCols sym) {
  constexpr int NT = 512, NL = kW2tSpecLoad;
  constexpr int TPC = M / 8;         // transform threads per column
  constexpr int C = NT / TPC;        // columns per tile
  constexpr int NE = M * C / NL;     // staged inputs per loader thread and tile
  constexpr int NR = 3 * M / NL;     // records per loader thread and residue
  static_assert(M == 256 || M == 512, "shapes of the wave-specialised form");
  static_assert(NL % C == 0 && NE % 8 == 0, "loader layout");
  using Plan = W2tPlan<M>;
  __shared__ c2 buf[M * C];
  __shared__ c2 stg[M * C];  // next tile's inputs / beta: [e][c], FFT position e
  __shared__ c2 s_tw[M];
  __shared__ c2 s_b[TPC];    // b_j = exp(-2 pi i j / L)
  __shared__ c2 s_S[8];      // S_r = exp(-2 pi i (TPC r - [r >= 4] M) / L)
  __shared__ c2 s_a[2][C][3];
  __shared__ c2 s_m1[2][C];
  __shared__ int s_col[2][C], s_mcol[2][C], s_cn[2];

  const int tid = threadIdx.x;
  const bool loader = tid >= NT;
  const int L = mt.L, Q = wd.Q;
  const long ntile = (sym.n + C - 1) / C;
  const long G = gridDim.x;
  const int first = -n_max, nmesh = 2 * n_max;
  const double ib = 1.0 / beta;
  auto live = [&](long tl) { return (sym.n - tl * C) < C ? (int)(sym.n - tl * C) : C; };
  // FFT position e <-> mesh index n = e or e - M; rows outside the mesh are zero
  auto in_row = [&](int e) {
    const int n = e < M / 2 ? e : e - M;
    return (n >= first && n < first + nmesh) ? n - first : -1;
  };

  // loader: thread lt = tid - NT owns column lc = lt % C of every tile and the
  // FFT positions e = lt / C + (NL / C) k, k < NE
  const int lt = tid - NT, lc = lt % C, le0 = lt / C;
  v2d pin[NE / 2];  // a residue's share of the next tile's inputs (at most half)
  auto stage_meta = [&](long tl, int pr) {
    const int cn_ = live(tl);
    const long c0 = tl * C;
    if (lt < C) {
      const int ci = lt < cn_ ? lt : 0;
      s_col[pr][lt] = sym.col[c0 + ci];
      s_mcol[pr][lt] = sym.mcol[c0 + ci];
      if (lt < cn_) {
        const c2 *mo = out + (long)(L - 2) * ncol + sym.col[c0 + lt];
        const c2 m1 = mo[0];
        Poles<0>::coef(m1, mo[ncol], mo[2 * ncol], s_a[pr][lt]);
        s_m1[pr][lt] = m1;
      }
      if (lt == 0) s_cn[pr] = cn_;
    }
  };
  // inputs k in [k0, k1) of tile tl: loads, then the LDS stores
  // this loader thread's column of tile tl (entry 0 of the tile past its end)
  auto lcol = [&](long tl, const int *list) {
    const int cn_ = live(tl);
    return (long)list[tl * C + (lc < cn_ ? lc : 0)];
  };
  auto load_in = [&](long col, int k0, int k1) {
    const int ler = opaque32(le0);
#pragma unroll
    for (int k = 0; k < NE / 2; ++k)
      if (k0 + k < k1) {
        const int row = in_row(ler + (NL / C) * (k0 + k));
        pin[k] = ld_v2d(in + (row >= 0 ? (long)row * ncol + col : col));
      }
  };
  auto store_in = [&](int k0, int k1) {
    const int lcr = opaque32(lc), ler = opaque32(le0);
#pragma unroll
    for (int k = 0; k < NE / 2; ++k)
      if (k0 + k < k1) {
        const int e = ler + (NL / C) * (k0 + k);
        st_v2d(stg + e * C + lcr, in_row(e) >= 0 ? v2d{pin[k].x * ib, pin[k].y * ib} : v2d{0.0, 0.0});
      }
  };

  long tile = xcd_remap(blockIdx.x, G);
  const int chunk = (NE + Q - 1) / Q;  // staged inputs per residue (<= NE / 2 for Q >= 2)
  v2d rec[NR];
  // the records of residue t (the same for every tile), into rec
  auto load_rec = [&](int t) {
    const int ltr = opaque32(lt);
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const int e = ltr + NL * i, jp = e / 3, part = e - 3 * jp;
      rec[i] = ld_v2d(wd.otab + (long)(Q * jp + t) * 3 + part);
    }
  };
  long colA = 0, mcolA = -1;  // loader: its column of the next tile and the mirror
  if (loader) {
    stage_meta(tile, 0);
    const long col0 = lcol(tile, sym.col);
    for (int k0 = 0; k0 < NE; k0 += NE / 2) {
      load_in(col0, k0, k0 + NE / 2);
      store_in(k0, k0 + NE / 2);
    }
    if (tile + G < ntile) {
      colA = lcol(tile + G, sym.col);
      mcolA = lcol(tile + G, sym.mcol);
    }
  } else {
    for (int k = tid; k < M; k += NT) st_v2d(s_tw + k, ld_v2d(wd.twM + k));
    if (tid < TPC) st_v2d(s_b + tid, ld_v2d(wd.twL + tid));
    if (tid < 8) {
      int k = TPC * tid - (tid >= 4 ? M : 0);
      if (k < 0) k += L;
      st_v2d(s_S + tid, ld_v2d(wd.twL + k));
    }
  }
  barrier_keep_vm();

  if (loader) {
    // The same barriers as the transform threads below.  At the start of each
    // residue the loader issues that residue's records and a chunk of the next
    // tile's inputs (chunk t in residue t), and stores them to LDS just before
    // the records barrier; in residue 0 it also loads the next tile's moments
    // (stored as pole coefficients with its columns) and the columns of the
    // tile after it.  The columns are loaded a tile ahead, so that no load
    // at a residue's start waits for another.
    long colB = 0, mcolB = -1;
    v2d mom[3];
    for (int par = 0;; par ^= 1) {
      const long next = tile + G, nn = next + G;
      const bool more = next < ntile;
      const int cnx = more ? live(next) : 0;
      for (int t = 0; t < Q; ++t) {
        load_rec(t);
        if (more) load_in(colA, t * chunk, min(NE, (t + 1) * chunk));
        if (t == 0) {
          if (more && lt < C) {
            const c2 *mo = out + (long)(L - 2) * ncol + colA;
#pragma unroll
            for (int q = 0; q < 3; ++q) mom[q] = ld_v2d(mo + q * ncol);
          }
          if (nn < ntile) {
            colB = lcol(nn, sym.col);
            mcolB = lcol(nn, sym.mcol);
          }
        }
#pragma unroll
        for (int p = 0; p + 1 < Plan::npass; ++p) {
          barrier_keep_vm();
          barrier_keep_vm();
        }
        // buf is free: this residue's records; stg was read before this
        // tile's first exchange: chunk t of the next tile's inputs
        const int ltr = opaque32(lt);
#pragma unroll
        for (int i = 0; i < NR; ++i) st_v2d(buf + ltr + NL * i, rec[i]);
        if (more) {
          store_in(t * chunk, min(NE, (t + 1) * chunk));
          if (t == 0 && lt < C) {
            s_col[par ^ 1][lt] = (int)colA;
            s_mcol[par ^ 1][lt] = (int)mcolA;
            if (lt < cnx) {
              const c2 m1{mom[0].x, mom[0].y};
              Poles<0>::coef(m1, c2{mom[1].x, mom[1].y}, c2{mom[2].x, mom[2].y}, s_a[par ^ 1][lt]);
              s_m1[par ^ 1][lt] = m1;
            }
            if (lt == 0) s_cn[par ^ 1] = cnx;
          }
        }
        barrier_keep_vm();
        barrier_keep_vm();
      }
      if (!more) break;
      tile = next;
      colA = colB;
      mcolA = mcolB;
    }
    return;
  }

  const int c = tid % C, j = tid / C;
  for (int par = 0;; par ^= 1) {
    const long next = tile + G;
    const bool more = next < ntile;
    c2 raw[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const v2d x = ld_v2d(stg + (j + TPC * r) * C + c);
      raw[r] = c2{x.x, x.y};
    }
    const int cn = s_cn[par];
    const long col = s_col[par][c], mcol = s_mcol[par][c];
    for (int t = 0; t < Q; ++t) {
      // the thread's indices laundered per residue: the addresses formed from
      // them are formed here instead of being held in registers across residues
      const int jr = opaque32(j), cr = opaque32(c);
      const long colr = opaque64(col), mcolr = opaque64(mcol);
      c2 v[8];
      if (t == 0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = raw[r];
      } else {
        const c2 bj = s_b[jr];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
          const c2 w1 = cmul(bj, s_S[r]);
          c2 w = w1;
          for (int k = 1; k < t; ++k) w = cmul(w, w1);
          v[r] = cmul(raw[r], w);  // raw is zero outside the mesh
        }
      }
      int Ns = 1;
#pragma unroll
      for (int p = 0; p < Plan::npass; ++p) {
        const int R = (p < Plan::n8) ? 8 : Plan::last;
        const int B = 8 / R;
        if (R == 8) w2t_butterflies<M, 8, -1>(v, jr, Ns, s_tw);
        else if (R == 4) w2t_butterflies<M, 4, -1>(v, jr, Ns, s_tw);
        else w2t_butterflies<M, 2, -1>(v, jr, Ns, s_tw);
        if (p + 1 < Plan::npass) {
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int uu = u / R, r = u % R;
            const int b = jr * B + uu;
            const int k = b % Ns;
            buf[((b - k) * R + k + r * Ns) * C + cr] = v[u];
          }
          barrier_keep_vm();
          Ns *= R;
          const int Rn = (p + 1 < Plan::n8) ? 8 : Plan::last;
          const int Bn = 8 / Rn;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int uu = u / Rn, r = u % Rn;
            v[u] = buf[(jr * Bn + uu + r * (M / Rn)) * C + cr];
          }
          barrier_keep_vm();
        } else {
          barrier_keep_vm();  // the records are in buf
          if (cr < cn) {
            // outputs: bin jr' -> tau index Q jr' + t, records [phase, (D_0, D_1), (D_2, 0)]
            const c2 a0 = s_a[par][cr][0], a1 = s_a[par][cr][1], a2 = s_a[par][cr][2];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const int uu = u / R, r = u % R;
              const int b = jr * B + uu;
              const int k = b % Ns;
              const int jp = (b - k) * R + k + r * Ns;
              const int jt = Q * jp + t;
              const c2 ph = buf[3 * jp], d01 = buf[3 * jp + 1], d2 = buf[3 * jp + 2];
              c2 y = cmul(v[u], ph);
              y.x += a0.x * d01.x + a1.x * d01.y + a2.x * d2.x;
              y.y += a0.y * d01.x + a1.y * d01.y + a2.y * d2.x;
              out[(long)jt * ncol + colr] = y;
              if (jt == 0) out[(long)L * ncol + colr] = cscale(cadd(y, s_m1[par][cr]), -1.0);
              if (mcolr >= 0) {
                out[(long)jt * ncol + mcolr] = cconj(y);
                if (jt == 0) out[(long)L * ncol + mcolr] = cconj(cscale(cadd(y, s_m1[par][cr]), -1.0));
              }
            }
          }
          barrier_keep_vm();  // buf is free for the next residue's exchanges
        }
      }
    }
    if (!more) break;
    tile = next;
  }
}

// The least-squares tail fit of every column as a kernel of its own: a
// workgroup takes 64 columns (column colmap[i], or i), lane = column, and its
// four waves 16 of the (at most 64) fitted frequencies each (the table reads
// are then wave-uniform); partial coefficients and residuals are summed
// through LDS.  The moments m_q = w_max^q s_q, q = 1..3, go to mom[(q - 1) ncol + col],
// the fit residual and |s_0| to the status words (one atomic per wave).  The
// pruned transform then runs without its in-workgroup fit, whose chain of
// dependent steps otherwise holds every wave of a workgroup.
constexpr int kTmParts = 4;
constexpr int kTmPer = kMaxFit / kTmParts;  // fitted frequencies per wave (at most)
__global__ void __launch_bounds__(64 * kTmParts) k_tail_moments(long ntot, const int *__restrict__ colmap, long ncol,
                                                                TailFitDev fit, const c2 *__restrict__ in,
                                                                c2 *__restrict__ mom, double *status) {
  __shared__ c2 s_part[kTmParts][kMaxOrder][64];
  __shared__ double s_r2[kTmParts][64];
  // the operator and the Vandermonde rows, zero past n_fit rows and ncoef
  // coefficients: every loop below has compile-time bounds and no branch, so
  // its LDS reads are issued together
  __shared__ c2 s_pinv[kMaxOrder][kMaxFit], s_vander[kMaxFit][kMaxOrder];
  const int nc = fit.ncoef, nf = fit.n_fit;
  for (int t = threadIdx.x; t < kMaxOrder * kMaxFit; t += blockDim.x) {
    const int q = t / kMaxFit, f = t - q * kMaxFit;
    s_pinv[q][f] = (q < nc && f < nf) ? fit.pinv[q * nf + f] : c2{0.0, 0.0};
    const int fv = t / kMaxOrder, qv = t - fv * kMaxOrder;
    s_vander[fv][qv] = (qv < nc && fv < nf) ? fit.vander[fv * nc + qv] : c2{0.0, 0.0};
  }
  const int lane = threadIdx.x & 63;
  const int part = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long i = (long)blockIdx.x * 64 + lane;
  const bool valid = i < ntot;
  const long col = valid ? (colmap ? (long)colmap[i] : i) : (colmap ? (long)colmap[0] : 0);
  // wave `part` takes the fitted rows f = kTmPer part + u (rows past n_fit
  // load row n_fit - 1 and have zero weight)
  const int f0 = part * kTmPer;
  c2 x[kTmPer];
#pragma unroll
  for (int u = 0; u < kTmPer; ++u) x[u] = in[(long)fit.idx[min(f0 + u, nf - 1)] * ncol + col];
  __syncthreads();  // the tables are in LDS (read as broadcasts below)
  c2 s[kMaxOrder];
#pragma unroll
  for (int q = 0; q < kMaxOrder; ++q) s[q] = c2{0.0, 0.0};
#pragma unroll
  for (int u = 0; u < kTmPer; ++u)
#pragma unroll
    for (int q = 0; q < kMaxOrder; ++q) s[q] = cfma(s[q], s_pinv[q][f0 + u], x[u]);
#pragma unroll
  for (int q = 0; q < kMaxOrder; ++q) s_part[part][q][lane] = s[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kMaxOrder; ++q) {
    c2 t = s_part[0][q][lane];
#pragma unroll
    for (int w = 1; w < kTmParts; ++w) t = cadd(t, s_part[w][q][lane]);
    s[q] = t;
  }
  double r2 = 0.0;
#pragma unroll
  for (int u = 0; u < kTmPer; ++u) {
    c2 v = x[u];
#pragma unroll
    for (int q = 0; q < kMaxOrder; ++q) v = cfms(v, s_vander[f0 + u][q], s[q]);
    r2 += f0 + u < nf ? cabs2(v) : 0.0;
  }
  s_r2[part][lane] = r2;
  __syncthreads();
  if (part != 0) return;
  r2 = s_r2[0][lane];
#pragma unroll
  for (int w = 1; w < kTmParts; ++w) r2 += s_r2[w][lane];
  if (valid) {
    mom[col] = cscale(s[1], fit.wpow[1]);
    mom[ncol + col] = cscale(s[2], fit.wpow[2]);
    mom[2 * ncol + col] = cscale(s[3], fit.wpow[3]);
  }
  double res = valid ? sqrt(r2 / nf) : 0.0, m0 = valid ? sqrt(cabs2(s[0])) : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    res = fmax(res, __shfl_xor(res, off, 64));
    m0 = fmax(m0, __shfl_xor(m0, off, 64));
  }
  if (lane == 0) {
    atomic_max_nonneg(&status[0], res);
    atomic_max_nonneg(&status[1], m0);
  }
}

hipError_t launch_tail_moments(hipStream_t s, long ntot, const int *colmap, long ncol, const TailFitDev &fit,
                               const c2 *in, c2 *mom, Status st) {
  hipLaunchKernelGGL(k_tail_moments, dim3((unsigned)((ntot + 63) / 64)), dim3(64 * kTmParts), 0, s, ntot, colmap,
                     ncol, fit, in, mom, st.d);
  return hipGetLastError();
}

bool w2t_pruned_supported(int M) { return M == 64 || M == 128 || M == 256 || M == 512 || M == 1024; }

template <int STAT, int M, int NT>
static hipError_t launch_m(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out, const c2 *moments,
                           const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st) {
  constexpr int C = NT / (M / 8);
  const long nb = (ncol + C - 1) / C;
  // without given moments: the fit kernel first, its moments parked in out
  const int in_out = moments == nullptr;
  if (in_out) {
    hipError_t e = launch_tail_moments(s, ncol, nullptr, ncol, fit, in, out + (long)(mt.L - 2) * ncol, st);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_w2t_pruned<STAT, M, NT>), dim3(nb), dim3(NT), 0, s, beta, n_max, ncol, in, out, moments, fit,
                     mt, wd, st.d, SymCols{nullptr, nullptr, 0}, in_out);
  return hipGetLastError();
}

template <int STAT, int NT>
static hipError_t launch_stat(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                              const c2 *moments, const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd,
                              Status st) {
  switch (wd.M) {
    case 64: return launch_m<STAT, 64, NT>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    case 128: return launch_m<STAT, 128, NT>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    case 256: return launch_m<STAT, 256, NT>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    case 512: return launch_m<STAT, 512, NT>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    case 1024: return launch_m<STAT, 1024, NT>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    default: return hipErrorNotSupported;
  }
}

int w2t_pruned_variants() { return 3; }

template <int M, int NT>
static hipError_t launch_sym_m(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                               const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st,
                               const SymCols &sym) {
  constexpr int C = NT / (M / 8);
  const long nb = (sym.n + C - 1) / C;
  hipError_t e = launch_tail_moments(s, sym.n, sym.col, ncol, fit, in, out + (long)(mt.L - 2) * ncol, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_w2t_pruned<0, M, NT, true>), dim3(nb), dim3(NT), 0, s, beta, n_max, ncol, in, out,
                     (const c2 *)nullptr, fit, mt, wd, st.d, sym, 1);
  return hipGetLastError();
}

template <int NT>
static hipError_t launch_sym_nt(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                                const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st,
                                const SymCols &sym) {
  switch (wd.M) {
    case 64: return launch_sym_m<64, NT>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
    case 128: return launch_sym_m<128, NT>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
    case 256: return launch_sym_m<256, NT>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
    case 512: return launch_sym_m<512, NT>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
    case 1024: return launch_sym_m<1024, NT>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
    default: return hipErrorNotSupported;
  }
}

// The number of CUs of the current device (cached per device): the grid of
// the persistent form.
static hipError_t device_cus(int &cus) {
  static int cache[64] = {0};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  if (!cache[dev]) {
    int cu = 0;
    e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    cache[dev] = cu > 1 ? cu : 1;
  }
  cus = cache[dev];
  return hipSuccess;
}

template <int M>
static hipError_t launch_sym_spec(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                                  const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st,
                                  const SymCols &sym) {
  constexpr int C = 512 / (M / 8);
  int cus = 0;
  hipError_t e = device_cus(cus);
  if (e != hipSuccess) return e;
  const long ntile = (sym.n + C - 1) / C;
  const long nb = ntile < cus ? ntile : cus;
  e = launch_tail_moments(s, sym.n, sym.col, ncol, fit, in, out + (long)(mt.L - 2) * ncol, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_w2t_spec<M>), dim3(nb), dim3(512 + kW2tSpecLoad), 0, s, beta, n_max, ncol, in, out, mt, wd,
                     sym);
  return hipGetLastError();
}

// variant 0: 256 threads per workgroup, 1: 512, 2: the wave-specialised
// persistent form where it applies (M = 256 or 512, Q >= 2), else as 1
hipError_t launch_w2t_pruned_sym(hipStream_t s, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                                 const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd, Status st, int variant,
                                 const SymCols &sym) {
  if (variant == 2 && wd.Q >= 2 && wd.M == 512)
    return launch_sym_spec<512>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
  if (variant == 2 && wd.Q >= 2 && wd.M == 256)
    return launch_sym_spec<256>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
  if (variant >= 1) return launch_sym_nt<512>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
  return launch_sym_nt<256>(s, beta, n_max, ncol, in, out, fit, mt, wd, st, sym);
}

// variant 0: 256 threads per workgroup; 1: 512 threads (twice the columns per
// workgroup, twice the contiguous bytes per written row)
hipError_t launch_w2t_pruned(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *in, c2 *out,
                             const c2 *moments, const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd,
                             Status st, int variant) {
  if (variant >= 1) {  // 2: the wave-specialised form is for the symmetric call; here as 1
    if (statistic == 0) return launch_stat<0, 512>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
    return launch_stat<1, 512>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
  }
  if (statistic == 0) return launch_stat<0, 256>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
  return launch_stat<1, 256>(s, beta, n_max, ncol, in, out, moments, fit, mt, wd, st);
}

}  // namespace tprf

namespace tprf {

// ------------------------------------------------------------------------
// Adjoint of the Matsubara -> tau transform above: given gbar(tau_j), j = 0..L,
// computes gbar(i w_n) = T^H gbar(tau) for the complex-linear map T of
// k_w2t_pruned (moments fitted, or given when `fitted` is 0).  Transposed
// steps in reverse order: the endpoint G(tau_L) = pm (G(tau_0) + m1), the
// phase, the length-M DFTs per residue t (conjugate sign), the pole model
// (added in tau, subtracted in frequency: together sum_q a_q D_q(tau_j)), the
// moment -> pole map and the least-squares fit.
template <int STAT, int M>
__global__ void __launch_bounds__(kW2tThreads, 2) k_w2t_adjoint(double beta, int n_max, long ncol,
                                                          const c2 *__restrict__ gbar_t, c2 *__restrict__ gbar_w,
                                                          int fitted, TailFitDev fit, MatsDev mt, W2tDev wd) {
  constexpr int NT = kW2tThreads;
  constexpr int TPC = M / 8;
  constexpr int C = NT / TPC;
  using Plan = W2tPlan<M>;
  __shared__ c2 buf[M * C];
  __shared__ c2 s_tw[M];
  __shared__ c2 s_part[NT][3];
  __shared__ c2 s_part2[C][8][3];
  __shared__ c2 s_mbar[C][3];

  const int tid = threadIdx.x;
  const int c = tid % C, j = tid / C;
  const long tile = xcd_remap(blockIdx.x, gridDim.x);
  const long c0 = tile * C;
  const int cn = (ncol - c0) < C ? (int)(ncol - c0) : C;
  const long col = c0 + (c < cn ? c : 0);
  const int L = mt.L, Q = wd.Q;
  const int nmesh = (STAT == 0) ? 2 * n_max : 2 * n_max - 1;
  const int first = (STAT == 0) ? -n_max : -(n_max - 1);
  const double pm = (STAT == 0) ? -1.0 : 1.0;
  for (int t = tid; t < M; t += NT) s_tw[t] = wd.twM[t];
  __syncthreads();

  const c2 gL = gbar_t[(long)L * ncol + col];
  c2 xbar[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) xbar[u] = c2{0.0, 0.0};
  c2 ab0 = c2{0.0, 0.0}, ab1 = c2{0.0, 0.0}, ab2 = c2{0.0, 0.0};

  for (int t = 0; t < Q; ++t) {
    c2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int jt = Q * (j + TPC * r) + t;
      c2 g = gbar_t[(long)jt * ncol + col];
      if (jt == 0) g = cadd(g, cscale(gL, pm));
      // pole model net of its in-mesh part: abar_q += D_q(tau_j) gbar_j (D_q real)
      ab0 = cadd(ab0, cscale(g, wd.dtail[jt]));
      ab1 = cadd(ab1, cscale(g, wd.dtail[L + jt]));
      ab2 = cadd(ab2, cscale(g, wd.dtail[2 * L + jt]));
      v[r] = (STAT == 0) ? cmul(cconj(mt.phase[jt]), g) : g;
    }
    int Ns = 1;
#pragma unroll
    for (int p = 0; p < Plan::npass; ++p) {
      const int R = (p < Plan::n8) ? 8 : Plan::last;
      if (R == 8) w2t_butterflies<M, 8, 1>(v, j, Ns, s_tw);
      else if (R == 4) w2t_butterflies<M, 4, 1>(v, j, Ns, s_tw);
      else w2t_butterflies<M, 2, 1>(v, j, Ns, s_tw);
      const int B = 8 / R;
      if (p + 1 < Plan::npass) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / R, r = u % R;
          const int b = j * B + uu;
          const int k = b % Ns;
          buf[((b - k) * R + k + r * Ns) * C + c] = v[u];
        }
        __syncthreads();
        Ns *= R;
        const int Rn = (p + 1 < Plan::n8) ? 8 : Plan::last;
        const int Bn = 8 / Rn;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / Rn, r = u % Rn;
          v[u] = buf[(j * Bn + uu + r * (M / Rn)) * C + c];
        }
        __syncthreads();
      } else {
        // bins e -> mesh n = e or e - M, twiddled by exp(+2 pi i t n / L) and summed over t
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int uu = u / R, r = u % R;
          const int b = j * B + uu;
          const int k = b % Ns;
          const int e = (b - k) * R + k + r * Ns;
          const int n = e < M / 2 ? e : e - M;
          c2 z = v[u];
          if (t) {
            int tn = t * n;
            if (tn < 0) tn += L;
            z = cmul(z, cconj(wd.twL[tn]));
          }
          xbar[u] = cadd(xbar[u], z);
        }
      }
    }
  }
  // gbar(i w_n) = xbar_n / beta (the in-mesh part of the pole model is already in D_q)
  const double ib = 1.0 / beta;
  {
    const int R = Plan::npass > Plan::n8 ? Plan::last : 8;
    const int B = 8 / R;
    int Ns = M / R;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int uu = u / R, r = u % R;
      const int b = j * B + uu;
      const int k = b % Ns;
      const int e = (b - k) * R + k + r * Ns;
      const int n = e < M / 2 ? e : e - M;
      c2 g = c2{0.0, 0.0};
      if (n >= first && n < first + nmesh) g = cscale(xbar[u], ib);
      buf[e * C + c] = g;
    }
  }
  s_part[tid][0] = ab0;
  s_part[tid][1] = ab1;
  s_part[tid][2] = ab2;
  __syncthreads();
  // the TPC partial sums of each column in two levels (8 groups, then their
  // total) rather than one chain of TPC dependent LDS reads
  for (int t = tid; t < cn * 8; t += NT) {
    const int cc = t >> 3, part = t & 7;
    c2 b0 = c2{0, 0}, b1 = c2{0, 0}, b2 = c2{0, 0};
#pragma unroll
    for (int w = 0; w < TPC / 8; ++w) {
      const int src = (part + 8 * w) * C + cc;
      b0 = cadd(b0, s_part[src][0]);
      b1 = cadd(b1, s_part[src][1]);
      b2 = cadd(b2, s_part[src][2]);
    }
    s_part2[cc][part][0] = b0;
    s_part2[cc][part][1] = b1;
    s_part2[cc][part][2] = b2;
  }
  __syncthreads();
  if (tid < cn) {
    c2 a0 = c2{0, 0}, a1 = c2{0, 0}, a2 = c2{0, 0};
#pragma unroll
    for (int part = 0; part < 8; ++part) {
      a0 = cadd(a0, s_part2[tid][part][0]);
      a1 = cadd(a1, s_part2[tid][part][1]);
      a2 = cadd(a2, s_part2[tid][part][2]);
    }
    c2 m1, m2, m3;
    if (STAT == 0) {  // a = (m1 - m3, (m2 + m3)/2, (m3 - m2)/2)
      m1 = a0;
      m2 = cscale(csub(a1, a2), 0.5);
      m3 = cadd(cscale(a0, -1.0), cscale(cadd(a1, a2), 0.5));
    } else {  // a = (4/3 (m1 - m3), m3 - (m1 + m2)/2, m1/6 + m2/2 + m3/3)
      m1 = cadd(cadd(cscale(a0, 4.0 / 3.0), cscale(a1, -0.5)), cscale(a2, 1.0 / 6.0));
      m2 = cadd(cscale(a1, -0.5), cscale(a2, 0.5));
      m3 = cadd(cadd(cscale(a0, -4.0 / 3.0), a1), cscale(a2, 1.0 / 3.0));
    }
    const c2 gLc = gbar_t[(long)L * ncol + c0 + tid];
    m1 = cadd(m1, cscale(gLc, pm));
    s_mbar[tid][0] = m1;
    s_mbar[tid][1] = m2;
    s_mbar[tid][2] = m3;
  }
  __syncthreads();
  if (fitted) {
    // moments m_p = w^p sum_f P[p][f] g_f  ->  gbar_f += w^p conj(P[p][f]) mbar_p
    for (int t = tid; t < cn * fit.n_fit; t += NT) {
      const int cc = t / fit.n_fit, f = t - cc * fit.n_fit;
      int e = first + fit.idx[f];
      if (e < 0) e += M;
      c2 add = c2{0.0, 0.0};
#pragma unroll
      for (int p = 1; p <= 3; ++p)
        add = cadd(add, cscale(cmul(cconj(fit.pinv[p * fit.n_fit + f]), s_mbar[cc][p - 1]), fit.wpow[p]));
      buf[e * C + cc] = cadd(buf[e * C + cc], add);
    }
    __syncthreads();
  }
  for (int t = tid; t < nmesh * C; t += NT) {
    const int i = t / C, cc = t - i * C;
    if (cc >= cn) continue;
    int e = first + i;
    if (e < 0) e += M;
    gbar_w[(long)i * ncol + c0 + cc] = buf[e * C + cc];
  }
}

template <int STAT, int M>
static hipError_t launch_adj_m(hipStream_t s, double beta, int n_max, long ncol, const c2 *gt, c2 *gw, int fitted,
                               const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd) {
  constexpr int C = kW2tThreads / (M / 8);
  const long nb = (ncol + C - 1) / C;
  hipLaunchKernelGGL((k_w2t_adjoint<STAT, M>), dim3(nb), dim3(kW2tThreads), 0, s, beta, n_max, ncol, gt, gw, fitted,
                     fit, mt, wd);
  return hipGetLastError();
}

template <int STAT>
static hipError_t launch_adj_stat(hipStream_t s, double beta, int n_max, long ncol, const c2 *gt, c2 *gw, int fitted,
                                  const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd) {
  switch (wd.M) {
    case 64: return launch_adj_m<STAT, 64>(s, beta, n_max, ncol, gt, gw, fitted, fit, mt, wd);
    case 128: return launch_adj_m<STAT, 128>(s, beta, n_max, ncol, gt, gw, fitted, fit, mt, wd);
    case 256: return launch_adj_m<STAT, 256>(s, beta, n_max, ncol, gt, gw, fitted, fit, mt, wd);
    case 512: return launch_adj_m<STAT, 512>(s, beta, n_max, ncol, gt, gw, fitted, fit, mt, wd);
    case 1024: return launch_adj_m<STAT, 1024>(s, beta, n_max, ncol, gt, gw, fitted, fit, mt, wd);
    default: return hipErrorNotSupported;
  }
}

hipError_t launch_w2t_adjoint(hipStream_t s, int statistic, double beta, int n_max, long ncol, const c2 *gbar_t,
                              c2 *gbar_w, int fitted, const TailFitDev &fit, const MatsDev &mt, const W2tDev &wd) {
  if (statistic == 0) return launch_adj_stat<0>(s, beta, n_max, ncol, gbar_t, gbar_w, fitted, fit, mt, wd);
  return launch_adj_stat<1>(s, beta, n_max, ncol, gbar_t, gbar_w, fitted, fit, mt, wd);
}

}  // namespace tprf
