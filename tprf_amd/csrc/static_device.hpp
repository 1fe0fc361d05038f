// static_device.hpp -- device helpers of the nu = 0 bubble in Matsubara space
// (static_mats.hip, static_q.hip): the 29 projection slots of a column, the
// small bilinear correction corr(q1, q2) and its two partial gradients, and
// the spreading of a slot gradient back over the frequencies.
//
// Slots: C_0..C_8 (sum cos_j sigma), S_1..S_8 (sum sin_j delta), y_0..y_3
// (Y g), y'_0..y'_3 (Y' g), mu_0..mu_3 (a_1, a_2, a_3, m_1 of the fitted tail);
// a column's slot k lives at q[k * ncol].  With e_j, f_j, Q of StaticMatsDev,
//   corr = sum_j e_j (C1_j C2_j + S1_j S2_j) + i f_j (S1_j C2_j - C1_j S2_j)
//        + mu1 . (y2 + Q mu2) + y'1 . mu2          (S_0 = 0, f_0 unused).
#pragma once

#include "common.hpp"

namespace tprf {

constexpr int kQC = 0, kQS = 9, kQY = 17, kQYp = 21, kQM = 25, kQSlots = 29;  // S_j at kQS + j - 1

__device__ __forceinline__ c2 cmul_i(c2 a) { return c2{-a.y, a.x}; }  // i a
// acc + t a for real t (two FMAs)
__device__ __forceinline__ c2 cfma_r(c2 acc, double t, c2 a) { return c2{fma(t, a.x, acc.x), fma(t, a.y, acc.y)}; }

__device__ __forceinline__ c2 corr_value(const c2 *q1, const c2 *q2, long ncol, const StaticMatsDev &sm) {
  auto a = [&](int k) { return q1[(long)k * ncol]; };
  auto b = [&](int k) { return q2[(long)k * ncol]; };
  c2 s = cscale(cmul(a(kQC), b(kQC)), sm.e[0]);
#pragma unroll
  for (int j = 1; j <= 8; ++j) {
    const c2 c1 = a(kQC + j), c2_ = b(kQC + j), s1 = a(kQS + j - 1), s2 = b(kQS + j - 1);
    s = cadd(s, cscale(cadd(cmul(c1, c2_), cmul(s1, s2)), sm.e[j]));
    s = cadd(s, cscale(cmul_i(csub(cmul(s1, c2_), cmul(c1, s2))), sm.f[j]));
  }
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    c2 t = b(kQY + qq);  // y2_q + (Q mu2)_q
#pragma unroll
    for (int p = 0; p < 4; ++p) t = cfma(t, c2{sm.Q[qq * 4 + p], 0.0}, b(kQM + p));
    s = cfma(s, a(kQM + qq), t);
    s = cfma(s, a(kQYp + qq), b(kQM + qq));
  }
  return s;
}

// qb += w conj(d corr / d q1) with the partner q2 = qs (this column in role 1)
__device__ __forceinline__ void qbar_role1(c2 *qb, c2 w, const c2 *qs, long ncol, const StaticMatsDev &sm) {
  auto b = [&](int k) { return qs[(long)k * ncol]; };
  qb[kQC] = cadd(qb[kQC], cmulc(w, cscale(b(kQC), sm.e[0])));
#pragma unroll
  for (int j = 1; j <= 8; ++j) {
    const c2 cs = b(kQC + j), ss = b(kQS + j - 1);
    qb[kQC + j] = cadd(qb[kQC + j], cmulc(w, csub(cscale(cs, sm.e[j]), cmul_i(cscale(ss, sm.f[j])))));
    qb[kQS + j - 1] = cadd(qb[kQS + j - 1], cmulc(w, cadd(cscale(ss, sm.e[j]), cmul_i(cscale(cs, sm.f[j])))));
  }
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    c2 t = b(kQY + qq);  // y2 + Q mu2
#pragma unroll
    for (int p = 0; p < 4; ++p) t = cfma(t, c2{sm.Q[qq * 4 + p], 0.0}, b(kQM + p));
    qb[kQM + qq] = cadd(qb[kQM + qq], cmulc(w, t));
    qb[kQYp + qq] = cadd(qb[kQYp + qq], cmulc(w, b(kQM + qq)));  // d / d y'1 = mu2
  }
}

// qb += w conj(d corr / d q2) with the partner q1 = qs (this column in role 2)
__device__ __forceinline__ void qbar_role2(c2 *qb, c2 w, const c2 *qs, long ncol, const StaticMatsDev &sm) {
  auto b = [&](int k) { return qs[(long)k * ncol]; };
  qb[kQC] = cadd(qb[kQC], cmulc(w, cscale(b(kQC), sm.e[0])));
#pragma unroll
  for (int j = 1; j <= 8; ++j) {
    const c2 cs = b(kQC + j), ss = b(kQS + j - 1);
    qb[kQC + j] = cadd(qb[kQC + j], cmulc(w, cadd(cscale(cs, sm.e[j]), cmul_i(cscale(ss, sm.f[j])))));
    qb[kQS + j - 1] = cadd(qb[kQS + j - 1], cmulc(w, csub(cscale(ss, sm.e[j]), cmul_i(cscale(cs, sm.f[j])))));
  }
#pragma unroll
  for (int qq = 0; qq < 4; ++qq) {
    c2 t = b(kQYp + qq);  // y'1 + Q^T mu1
#pragma unroll
    for (int p = 0; p < 4; ++p) t = cfma(t, c2{sm.Q[p * 4 + qq], 0.0}, b(kQM + p));
    qb[kQM + qq] = cadd(qb[kQM + qq], cmulc(w, t));
    qb[kQY + qq] = cadd(qb[kQY + qq], cmulc(w, b(kQM + qq)));  // d / d y2 = mu1
  }
}

// The slot gradient spread over the n-th sigma / delta (n >= 0) of the table
// row t: sigma-bar and delta-bar, so that gbar(n) += sb + db and
// gbar(-n-1) += sb - db.  (y = sum Re Y sigma + i Im Y delta, hence
// delta-bar += conj(i Im Y) ybar.)
__device__ __forceinline__ void qbar_spread(const c2 *qb, const double *t, c2 &sb, c2 &db) {
  sb = c2{0, 0};
  db = c2{0, 0};
#pragma unroll
  for (int j = 0; j < 9; ++j) sb = cfma_r(sb, t[j], qb[kQC + j]);
#pragma unroll
  for (int j = 0; j < 8; ++j) db = cfma_r(db, t[9 + j], qb[kQS + j]);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    sb = cfma_r(sb, t[17 + i], qb[kQY + i]);
    sb = cfma_r(sb, t[25 + i], qb[kQYp + i]);
    // db -= i (Im Y) qbar: (x, y) -> (x + t y, y - t x)
    db = c2{fma(t[21 + i], qb[kQY + i].y, db.x), fma(-t[21 + i], qb[kQY + i].x, db.y)};
    db = c2{fma(t[29 + i], qb[kQYp + i].y, db.x), fma(-t[29 + i], qb[kQYp + i].x, db.y)};
  }
}

// d mu -> d m: mu = (m1 - m3, (m2 + m3)/2, (m3 - m2)/2, m1); mb[1..3] = mbar
__device__ __forceinline__ void mu_to_moment_bar(const c2 *qb, c2 *mb) {
  const c2 a0 = qb[kQM], a1 = qb[kQM + 1], a2 = qb[kQM + 2];
  mb[0] = c2{0, 0};
  mb[1] = cadd(a0, qb[kQM + 3]);
  mb[2] = cscale(csub(a1, a2), 0.5);
  mb[3] = cadd(cscale(a0, -1.0), cscale(cadd(a1, a2), 0.5));
}

}  // namespace tprf
