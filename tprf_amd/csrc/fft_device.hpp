// fft_device.hpp -- batched mixed-radix Stockham FFT on LDS (FP64 complex).
//
// A workgroup transforms `ncol` columns of length n held in LDS at
// a[c * ld + j].  Radices 2, 3, 4, 5, 7, 8, 11 and 13 have unrolled
// butterflies; any other prime factor runs as a run-time radix of O(R^2)
// work.  Each pass reads one buffer and writes the other (autosort, no bit
// reversal), so two buffers of ncol * ld elements are needed; the function
// returns the buffer holding the result.  S = -1 computes
// y_k = sum_j x_j exp(-2 pi i jk/n) (FFTW_FORWARD), S = +1 the conjugate
// kernel (FFTW_BACKWARD); neither scales.
#pragma once

#include "common.hpp"

namespace tprf {

template <int S>
__device__ __forceinline__ c2 twiddle(const c2 *__restrict__ tw, int e) {
  c2 w = tw[e];
  if (S > 0) w.y = -w.y;
  return w;
}

// in-place small DFT: v_q <- sum_r v_r exp(S 2 pi i r q / R)
template <int R, int S>
struct SmallDft;

template <int S>
struct SmallDft<2, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    c2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S>
struct SmallDft<3, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    const double s3 = 0.86602540378443864676 * S;  // S sin(2 pi / 3)
    c2 t = cadd(v[1], v[2]);
    c2 d = csub(v[1], v[2]);
    c2 m = c2{v[0].x - 0.5 * t.x, v[0].y - 0.5 * t.y};
    c2 id = cmul_i(d, s3);
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, id);
    v[2] = csub(m, id);
  }
};

template <int S>
struct SmallDft<4, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    c2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    c2 c = cadd(v[1], v[3]);
    c2 d = cmul_i(csub(v[1], v[3]), (double)S);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, d);
    v[3] = csub(b, d);
  }
};

template <int S>
struct SmallDft<5, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    const double c1 = 0.30901699437494742410, c2_ = -0.80901699437494742410;
    const double s1 = 0.95105651629515357212 * S, s2 = 0.58778525229247312917 * S;
    c2 p14 = cadd(v[1], v[4]), m14 = csub(v[1], v[4]);
    c2 p23 = cadd(v[2], v[3]), m23 = csub(v[2], v[3]);
    c2 x0 = v[0];
    c2 r1 = c2{x0.x + c1 * p14.x + c2_ * p23.x, x0.y + c1 * p14.y + c2_ * p23.y};
    c2 r2 = c2{x0.x + c2_ * p14.x + c1 * p23.x, x0.y + c2_ * p14.y + c1 * p23.y};
    c2 i1 = c2{s1 * m14.x + s2 * m23.x, s1 * m14.y + s2 * m23.y};
    c2 i2 = c2{s2 * m14.x - s1 * m23.x, s2 * m14.y - s1 * m23.y};
    c2 j1 = cmul_i(i1, 1.0), j2 = cmul_i(i2, 1.0);
    v[0] = cadd(x0, cadd(p14, p23));
    v[1] = cadd(r1, j1);
    v[4] = csub(r1, j1);
    v[2] = cadd(r2, j2);
    v[3] = csub(r2, j2);
  }
};

template <int S>
struct SmallDft<8, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    const double h = 0.70710678118654752440;
    c2 e[4] = {v[0], v[2], v[4], v[6]};
    c2 o[4] = {v[1], v[3], v[5], v[7]};
    SmallDft<4, S>::run(e, nullptr, 0);
    SmallDft<4, S>::run(o, nullptr, 0);
    // o_q *= W8^q, W8 = exp(S i pi / 4)
    c2 o1 = c2{h * (o[1].x - S * o[1].y), h * (o[1].y + S * o[1].x)};
    c2 o2 = cmul_i(o[2], (double)S);
    c2 o3 = c2{h * (-o[3].x - S * o[3].y), h * (-o[3].y + S * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);
    v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);
    v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);
    v[7] = csub(e[3], o3);
  }
};

// 16 = 4 x 4: DFT-4 over n1 for each n2, twiddle W16^{n2 k1}, DFT-4 over n2
template <int S>
struct SmallDft<16, S> {
  __device__ __forceinline__ static void run(c2 *v, const c2 *, int) {
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, h = 0.70710678118654752440;
    c2 y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      c2 t[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
      SmallDft<4, S>::run(t, nullptr, 0);
#pragma unroll
      for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = t[k1];
    }
    // W16^p = exp(S 2 pi i p / 16) for p = n2 * k1
    const c2 w1 = c2{c1, S * s1}, w2 = c2{h, S * h}, w3 = c2{s1, S * c1};
    const c2 w6 = c2{-h, S * h}, w9 = c2{-c1, -S * s1};
    y[1][1] = cmul(y[1][1], w1);
    y[1][2] = cmul(y[1][2], w2);
    y[1][3] = cmul(y[1][3], w3);
    y[2][1] = cmul(y[2][1], w2);
    y[2][2] = cmul_i(y[2][2], (double)S);
    y[2][3] = cmul(y[2][3], w6);
    y[3][1] = cmul(y[3][1], w3);
    y[3][2] = cmul(y[3][2], w6);
    y[3][3] = cmul(y[3][3], w9);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      c2 t[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
      SmallDft<4, S>::run(t, nullptr, 0);
#pragma unroll
      for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
    }
  }
};

// generic odd prime radix via the length-n twiddle table (R divides n)
template <int R, int S>
struct SmallDftGeneric {
  __device__ __forceinline__ static void run(c2 *v, const c2 *tw, int n) {
    c2 out[R];
    const int step = n / R;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      c2 s = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) s = cadd(s, cmul(v[r], twiddle<S>(tw, ((r * q) % R) * step)));
      out[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = out[q];
  }
};
template <int S>
struct SmallDft<7, S> : SmallDftGeneric<7, S> {};
template <int S>
struct SmallDft<11, S> : SmallDftGeneric<11, S> {};
template <int S>
struct SmallDft<13, S> : SmallDftGeneric<13, S> {};

// one radix-R Stockham butterfly of the pass with stride Ns
template <int R, int S>
__device__ __forceinline__ void stockham_bfly(const c2 *__restrict__ src, c2 *__restrict__ dst, int j, int m, int Ns,
                                              int twstep, const c2 *__restrict__ tw, int n) {
  c2 v[R];
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = src[j + r * m];
  const int k = j % Ns;
  if (k) {
    const int e = k * twstep;
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twiddle<S>(tw, e * r));
  }
  SmallDft<R, S>::run(v, tw, n);
  const int o = (j - k) * R + k;
#pragma unroll
  for (int r = 0; r < R; ++r) dst[o + r * Ns] = v[r];
}

// Output q of the radix-R butterfly j of the pass with stride Ns for a radix
// known only at run time (a prime above 13): the twiddle w^(k twstep r) and the
// DFT root w^((r q mod R) n / R) taken together from the length-n table, R
// terms read from src.  O(R^2) per butterfly.
template <int S>
__device__ __forceinline__ c2 runtime_radix_out(const c2 *src, int j, int q, int R, int m, int Ns, int twstep,
                                                const c2 *__restrict__ tw, int n) {
  const int e = (j % Ns) * twstep;
  c2 acc = src[j];
  int rq = 0, er = 0;
  for (int r = 1; r < R; ++r) {
    rq += q;
    if (rq >= R) rq -= R;
    er += e;
    if (er >= n) er -= n;
    int idx = er + rq * m;
    if (idx >= n) idx -= n;
    acc = cadd(acc, cmul(src[j + r * m], twiddle<S>(tw, idx)));
  }
  return acc;
}

template <int S>
__device__ c2 *fft_lds(c2 *a, c2 *b, int ld, int ncol, const FftDev &p, int tid, int nthr) {
  int Ns = 1;
  for (int ps = 0; ps < p.npass; ++ps) {
    const int R = p.radix[ps];
    const int m = p.n / R;
    const int twstep = p.n / (Ns * R);
    const int total = m * ncol;
    for (int t = tid; t < total; t += nthr) {
      const int c = t / m;
      const int j = t - c * m;
      const c2 *src = a + c * ld;
      c2 *dst = b + c * ld;
      switch (R) {
        case 2: stockham_bfly<2, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 3: stockham_bfly<3, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 4: stockham_bfly<4, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 5: stockham_bfly<5, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 7: stockham_bfly<7, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 8: stockham_bfly<8, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 11: stockham_bfly<11, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        case 13: stockham_bfly<13, S>(src, dst, j, m, Ns, twstep, p.tw, p.n); break;
        default: {
          const int o = (j - j % Ns) * R + j % Ns;
          for (int q = 0; q < R; ++q) dst[o + q * Ns] = runtime_radix_out<S>(src, j, q, R, m, Ns, twstep, p.tw, p.n);
        }
      }
    }
    __syncthreads();
    c2 *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// One column of length n <= NMAX transformed in place in a single LDS buffer
// by NT threads: each Stockham pass first takes every butterfly of the thread
// into registers (at most ceil(NMAX / (R NT)) of them, indices fixed at compile
// time), then after a barrier writes them to their autosorted positions.  For
// lengths whose two buffers do not fit in LDS.
template <int R, int S, int NT, int NMAX>
__device__ __forceinline__ void inplace_pass(c2 *a, int n, int Ns, const c2 *__restrict__ tw, int tid) {
  constexpr int KB = (NMAX + R * NT - 1) / (R * NT);
  const int m = n / R;
  const int twstep = n / (Ns * R);
  c2 v[KB][R];
#pragma unroll
  for (int s = 0; s < KB; ++s) {
    const int j = tid + s * NT;
    if (j < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[s][r] = a[j + r * m];
      const int k = j % Ns;
      if (k) {
        const int e = k * twstep;
#pragma unroll
        for (int r = 1; r < R; ++r) v[s][r] = cmul(v[s][r], twiddle<S>(tw, e * r));
      }
      SmallDft<R, S>::run(v[s], tw, n);
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < KB; ++s) {
    const int j = tid + s * NT;
    if (j < m) {
      const int k = j % Ns;
      const int o = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) a[o + r * Ns] = v[s][r];
    }
  }
  __syncthreads();
}

// in-place pass of a run-time radix: each thread keeps its outputs
// t = q m + j (at most ceil(NMAX / NT)) in registers until every read is done
template <int S, int NT, int NMAX>
__device__ __forceinline__ void inplace_pass_runtime(c2 *a, int n, int R, int Ns, const c2 *__restrict__ tw,
                                                     int tid) {
  constexpr int KB = (NMAX + NT - 1) / NT;
  const int m = n / R;
  const int twstep = n / (Ns * R);
  c2 v[KB];
#pragma unroll
  for (int s = 0; s < KB; ++s) {
    const int t = tid + s * NT;
    if (t < n) v[s] = runtime_radix_out<S>(a, t % m, t / m, R, m, Ns, twstep, tw, n);
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < KB; ++s) {
    const int t = tid + s * NT;
    if (t < n) {
      const int j = t % m, q = t / m;
      a[(j - j % Ns) * R + j % Ns + q * Ns] = v[s];
    }
  }
  __syncthreads();
}

template <int S, int NT, int NMAX>
__device__ c2 *fft_lds_inplace(c2 *a, const FftDev &p, int tid) {
  int Ns = 1;
  for (int ps = 0; ps < p.npass; ++ps) {
    const int R = p.radix[ps];
    switch (R) {
      case 2: inplace_pass<2, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 3: inplace_pass<3, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 4: inplace_pass<4, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 5: inplace_pass<5, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 7: inplace_pass<7, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 8: inplace_pass<8, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 11: inplace_pass<11, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      case 13: inplace_pass<13, S, NT, NMAX>(a, p.n, Ns, p.tw, tid); break;
      default: inplace_pass_runtime<S, NT, NMAX>(a, p.n, R, Ns, p.tw, tid); break;
    }
    Ns *= R;
  }
  return a;
}

// Map a hardware workgroup id to a logical tile so that consecutive logical
// tiles run on the same XCD (workgroups are dealt round-robin over 8 XCDs,
// each with its own L2): neighbours that share cache lines then share an L2.
__device__ __forceinline__ long xcd_remap(long b, long nb) {
  const long q = nb / 8, rem = nb % 8;
  const long x = b % 8, i = b / 8;
  return x * q + (x < rem ? x : rem) + i;
}

}  // namespace tprf
