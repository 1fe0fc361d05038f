// mats_device.hpp -- device helpers shared by the Matsubara-transform kernels.
#pragma once

#include "common.hpp"

namespace tprf {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax(reinterpret_cast<unsigned long long *>(addr), (unsigned long long)__double_as_longlong(v));
}

// Pole model of fourier_matsubara.cpp:150-171,274-288:
// fermions b = (0, 1, -1), bosons b = (-1/2, -1, 1).
template <int STAT>
struct Poles;
template <>
struct Poles<0> {
  static constexpr double b0 = 0.0, b1 = 1.0, b2 = -1.0;
  __device__ static void coef(c2 m1, c2 m2, c2 m3, c2 *a) {
    a[0] = csub(m1, m3);
    a[1] = cscale(cadd(m2, m3), 0.5);
    a[2] = cscale(csub(m3, m2), 0.5);
  }
};
template <>
struct Poles<1> {
  static constexpr double b0 = -0.5, b1 = -1.0, b2 = 1.0;
  __device__ static void coef(c2 m1, c2 m2, c2 m3, c2 *a) {
    a[0] = cscale(csub(m1, m3), 4.0 / 3.0);
    a[1] = csub(m3, cscale(cadd(m1, m2), 0.5));
    a[2] = cadd(cadd(cscale(m1, 1.0 / 6.0), cscale(m2, 0.5)), cscale(m3, 1.0 / 3.0));
  }
};

// sum_i a_i / (i w - b_i)
template <int STAT>
__device__ __forceinline__ c2 pole_sum(const c2 *a, double w) {
  c2 s = cdiv(a[0], c2{-Poles<STAT>::b0, w});
  s = cadd(s, cdiv(a[1], c2{-Poles<STAT>::b1, w}));
  s = cadd(s, cdiv(a[2], c2{-Poles<STAT>::b2, w}));
  return s;
}

template <int STAT>
__device__ __forceinline__ double matsubara_w(long n, double beta) {
  return STAT == 0 ? kPi * (double)(2 * n + 1) / beta : 2.0 * kPi * (double)n / beta;
}

// -r on the cyclic lattice (d0, d1, d2): component-wise (-i mod d)
__device__ __forceinline__ long minus_r(long r, int d0, int d1, int d2) {
  const long i2 = r % d2, i1 = (r / d2) % d1, i0 = r / ((long)d1 * d2);
  const long j0 = (d0 - i0) % d0, j1 = (d1 - i1) % d1, j2 = (d2 - i2) % d2;
  return (j0 * d1 + j1) * d2 + j2;
}

// Boundary moments of a tau function from its first and last nine samples
// (fit_derivatives, fourier_matsubara.cpp:42-97, with m1 from :151,:163):
// lo(m) = g(tau_m), hi(m) = g(tau_{L-m}), m = 0..8.  Returns the pole
// coefficients a[3] of the statistic's model.
// RCP: the differences are multiplied by 1 / tau_m (one division per m
// instead of four; results within an ulp of the quotients).
template <int STAT, bool RCP = false, typename Lo, typename Hi>
__device__ __forceinline__ void boundary_poles(Lo lo, Hi hi, double dtau, const Vinv8 &vinv, c2 *a, c2 *m1out) {
  c2 gl0 = c2{0, 0}, gl1 = c2{0, 0}, gr0 = c2{0, 0}, gr1 = c2{0, 0};
  const c2 g0 = lo(0), gL = hi(0);
#pragma unroll
  for (int m = 1; m <= 8; ++m) {
    const double tm = (double)m * dtau;
    const c2 gm = lo(m), gLm = hi(m);
    c2 dl, dr;
    if constexpr (RCP) {
      const double itm = 1.0 / tm;
      dl = c2{(gm.x - g0.x) * itm, (gm.y - g0.y) * itm};
      dr = c2{(gL.x - gLm.x) * itm, (gL.y - gLm.y) * itm};
    } else {
      dl = c2{(gm.x - g0.x) / tm, (gm.y - g0.y) / tm};
      dr = c2{(gL.x - gLm.x) / tm, (gL.y - gLm.y) / tm};
    }
    gl0 = cadd(gl0, cscale(dl, vinv.v[0 * 8 + m - 1]));
    gl1 = cadd(gl1, cscale(dl, vinv.v[1 * 8 + m - 1]));
    gr0 = cadd(gr0, cscale(dr, vinv.v[0 * 8 + m - 1]));
    gr1 = cadd(gr1, cscale(dr, vinv.v[1 * 8 + m - 1]));
  }
  const double sign = (STAT == 0) ? -1.0 : 1.0;
  const c2 m2 = csub(gl0, cscale(gr0, sign));
  const c2 m3 = cscale(cadd(gl1, cscale(gr1, sign)), -2.0 / dtau);
  const c2 m1 = (STAT == 0) ? cscale(cadd(g0, gL), -1.0) : cscale(csub(g0, gL), -1.0);
  Poles<STAT>::coef(m1, m2, m3, a);
  if (m1out) *m1out = m1;
}

}  // namespace tprf
