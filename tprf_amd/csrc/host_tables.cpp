// host_tables.cpp -- see host_tables.hpp.  Host code only (g++ or hipcc).
#include "host_tables.hpp"

#include <cmath>
#include <complex>
#include <string>
#include <vector>

namespace tprf_host {

namespace {
double matsubara_l(int statistic, long n, double beta) {
  return statistic == kFermion ? M_PI * (double)(2 * n + 1) / beta : 2.0 * M_PI * (double)n / beta;
}
}  // namespace

// Tail-fit operator of one Matsubara mesh.  Window and order restate the
// TRIQS 3.2 tail_fitter defaults that fit_tail uses (called at
// c++/triqs_tprf/fourier/fourier_matsubara.cpp:226): 20% of each half mesh,
// at most 30 points per side from the outermost frequency inward with an
// integer step, expansion order min(9, floor(ln 1e16 / ln omega_max)), fit in
// z = omega_max / (i omega_n).  The least-squares operator pinv = R^-1 Q^H is
// built once by modified Gram-Schmidt (two passes) in long double.
int tail_fit_build(int statistic, double beta, int n_max, TailTables &t, std::string &msg) {
  const int N = statistic == kFermion ? 2 * n_max : 2 * n_max - 1;
  const int first = statistic == kFermion ? -n_max : -(n_max - 1);
  const double wmax = statistic == kFermion ? M_PI * (2 * n_max - 1) / beta : 2 * M_PI * (n_max - 1) / beta;
  int W = (int)std::lround(0.2 * N / 2.0);
  if (W < 1) W = 1;
  const int npts = W < 30 ? W : 30;
  int step = W / npts;
  if (step < 1) step = 1;
  if ((npts - 1) * step >= (N - 1) / 2)
    return msg = "Matsubara mesh too small for the high-frequency tail fit", 1;
  const int nfit = 2 * npts;
  int K = 9;
  if (wmax > 1.0) {
    int k2 = (int)std::floor(std::log(1e16) / std::log(wmax));
    if (k2 < K) K = k2;
  }
  if (K > nfit - 1) K = nfit - 1;
  if (K < 4)
    return msg = "ERROR: Inverse Fourier implementation requires at least a proper 3rd high-frequency moment", 1;
  const int nc = K + 1;
  std::vector<int> idx(nfit);
  for (int i = 0; i < npts; ++i) { idx[i] = i * step; idx[npts + i] = N - 1 - i * step; }
  typedef std::complex<long double> lc;
  std::vector<lc> V(nfit * nc), Q(nfit * nc), R(nc * nc, lc(0));
  for (int r = 0; r < nfit; ++r) {
    lc z = lc(wmax) / lc(0.0L, (long double)matsubara_l(statistic, first + idx[r], beta));
    lc pw = 1;
    for (int c = 0; c < nc; ++c) { V[r * nc + c] = pw; pw *= z; }
  }
  for (int c = 0; c < nc; ++c) {
    std::vector<lc> v(nfit);
    for (int r = 0; r < nfit; ++r) v[r] = V[r * nc + c];
    for (int pass = 0; pass < 2; ++pass)
      for (int p = 0; p < c; ++p) {
        lc s = 0;
        for (int r = 0; r < nfit; ++r) s += std::conj(Q[r * nc + p]) * v[r];
        for (int r = 0; r < nfit; ++r) v[r] -= s * Q[r * nc + p];
        R[p * nc + c] += s;
      }
    long double nv = 0;
    for (int r = 0; r < nfit; ++r) nv += std::norm(v[r]);
    nv = std::sqrt(nv);
    R[c * nc + c] = nv;
    for (int r = 0; r < nfit; ++r) Q[r * nc + c] = v[r] / nv;
  }
  // pinv = R^-1 Q^H  (nc x nfit)
  struct cd { double x, y; };
  std::vector<cd> pinv(nc * nfit), vd(nfit * nc);
  for (int f = 0; f < nfit; ++f) {
    std::vector<lc> x(nc);
    for (int r = nc - 1; r >= 0; --r) {
      lc s = std::conj(Q[f * nc + r]);
      for (int c = r + 1; c < nc; ++c) s -= R[r * nc + c] * x[c];
      x[r] = s / R[r * nc + r];
    }
    for (int r = 0; r < nc; ++r) pinv[r * nfit + f] = cd{(double)x[r].real(), (double)x[r].imag()};
  }
  for (int i = 0; i < nfit * nc; ++i) vd[i] = cd{(double)V[i].real(), (double)V[i].imag()};
  t.nfit = nfit;
  t.ncoef = nc;
  t.wmax = wmax;
  t.idx = idx;
  t.fitpos.assign(N, -1);
  for (int f = 0; f < nfit; ++f) t.fitpos[idx[f]] = f;
  t.pinv.resize(pinv.size());
  for (size_t i = 0; i < pinv.size(); ++i) t.pinv[i] = std::complex<double>(pinv[i].x, pinv[i].y);
  t.vander.resize(vd.size());
  for (size_t i = 0; i < vd.size(); ++i) t.vander[i] = std::complex<double>(vd[i].x, vd[i].y);
  long double wp = 1;
  for (int p = 0; p < 4; ++p) { t.wpow[p] = (double)wp; wp *= (long double)wmax; }
  return 0;
}

// oneFermion / oneBoson with a = 1, c++/triqs_tprf/fourier/fourier_matsubara.cpp:31-37
double one_fermion(double b, double tau, double beta) {
  return -(b >= 0 ? std::exp(-b * tau) / (1 + std::exp(-beta * b))
                  : std::exp(b * (beta - tau)) / (1 + std::exp(beta * b)));
}
double one_boson(double b, double tau, double beta) {
  return (b >= 0 ? std::exp(-b * tau) / (std::exp(-beta * b) - 1) : std::exp(b * (beta - tau)) / (1 - std::exp(b * beta)));
}

// exact inverse of V_{m,j} = m^{j-1}, m, j = 1..8: row j-1 holds the t^{j-1}
// coefficients of the Lagrange polynomials on the nodes 1..8 (integer ratios).
void vandermonde_inverse8(double *vinv) {
  for (int m = 1; m <= 8; ++m) {
    double poly[9] = {1, 0, 0, 0, 0, 0, 0, 0, 0};
    double den = 1;
    int deg = 0;
    for (int k = 1; k <= 8; ++k) {
      if (k == m) continue;
      for (int d = deg + 1; d >= 1; --d) poly[d] = poly[d - 1] - k * poly[d];
      poly[0] = -k * poly[0];
      ++deg;
      den *= (double)(m - k);
    }
    for (int j = 0; j < 8; ++j) vinv[j * 8 + (m - 1)] = poly[j] / den;
  }
}

// D_q(tau_j) = psi_q(tau_j) - (1/beta) sum_{n in mesh} exp(-i w_n tau_j) / (i w_n - b_q),
// j = 0..L-1: what the pole model adds in tau beyond its own in-mesh transform.
// The omega -> tau map is its FFT of G / beta plus sum_q a_q D_q(tau_j), so the
// forward kernel adds this table at its outputs and the adjoint needs
// a_q-bar = sum_j D_q(tau_j) gbar_j.  D is small (the pole tail beyond the mesh)
// while both of its terms are O(1); it is formed here in long double so that
// neither kernel takes the difference of two large sums.
std::vector<long double> dtail_values_ld(int statistic, double beta, int n_max, int L) {
  typedef std::complex<long double> lc;
  const long double bf[3] = {0.0L, 1.0L, -1.0L}, bb[3] = {-0.5L, -1.0L, 1.0L};
  const long double *b = statistic == kFermion ? bf : bb;
  const int first = statistic == kFermion ? -n_max : -(n_max - 1);
  const int nmesh = statistic == kFermion ? 2 * n_max : 2 * n_max - 1;
  const long double lb = beta;
  std::vector<long double> d(3 * (size_t)L);
  std::vector<lc> pole(3 * (size_t)nmesh);
  for (int q = 0; q < 3; ++q)
    for (int i = 0; i < nmesh; ++i) {
      const long n = first + i;
      const long double w = statistic == kFermion ? kPiL * (2 * n + 1) / lb : 2 * kPiL * n / lb;
      pole[(size_t)q * nmesh + i] = 1.0L / lc(-b[q], w);
    }
  std::vector<lc> ph(2 * (size_t)L);  // exp(-i pi k / L), k < 2L
  for (long k = 0; k < 2L * L; ++k) {
    const long double ang = kPiL * (long double)k / L;
    ph[k] = lc(cosl(ang), -sinl(ang));
  }
  for (int j = 0; j < L; ++j) {
    const long double tau = lb * j / L;
    lc s[3] = {0, 0, 0};
    for (int i = 0; i < nmesh; ++i) {
      const long n = first + i;
      // w_n tau_j = pi (2n+1) j / L (fermions) or 2 pi n j / L (bosons), reduced exactly mod 2 L
      long num = statistic == kFermion ? (2 * n + 1) * (long)j : 2 * n * (long)j;
      num %= 2L * L;
      if (num < 0) num += 2L * L;
      const lc e = ph[num];
      for (int q = 0; q < 3; ++q) s[q] += e * pole[(size_t)q * nmesh + i];
    }
    for (int q = 0; q < 3; ++q) {
      long double psi;
      const long double bq = b[q];
      if (statistic == kFermion)
        psi = bq >= 0 ? -expl(-bq * tau) / (1 + expl(-lb * bq)) : -expl(bq * (lb - tau)) / (1 + expl(lb * bq));
      else
        psi = bq >= 0 ? expl(-bq * tau) / (expl(-lb * bq) - 1) : expl(bq * (lb - tau)) / (1 - expl(bq * lb));
      // real: the mesh is symmetric under w -> -w and the terms of w and -w are
      // complex conjugates (b_q real), so only rounding is left in v.imag()
      const lc v = lc(psi, 0) - s[q] / lb;
      d[(size_t)q * L + j] = v.real();
    }
  }
  return d;
}


// Weights c_j, j = 0..L, of the bosonic tau -> nu transform at nu = 0
// (fourier_matsubara.cpp:42-184 restated by t2w_core in kernels.hip):
//   chi(0) = (beta/L) sum_{j<L} (x_j - sum_q a_q psi_q(tau_j)) + sum_q a_q / (-b_q)
// with a = A m the pole coefficients of the boundary moments
//   m1 = x_L - x_0,  m2 = gl0 - gr0,  m3 = -(2/dtau)(gl1 + gr1),
//   gl_i = sum_m V_{i,m} (x_m - x_0)/(m dtau),  gr_i = sum_m V_{i,m} (x_L - x_{L-m})/(m dtau),
// all linear in x with real coefficients, so chi(0) = sum_j c_j x_j.
std::vector<long double> static_weights_ld(const double *vinv8, double beta, int L) {
  const double bb[3] = {-0.5, -1.0, 1.0};
  // a_q = sum_p A[q][p] m_p (bosonic three-pole model)
  const long double A[3][3] = {{4.0L / 3, 0, -4.0L / 3}, {-0.5L, -0.5L, 1}, {1.0L / 6, 0.5L, 1.0L / 3}};
  const long double fact = (long double)beta / L;
  const double dtau = beta / (double)L;
  long double phi[3];
  for (int q = 0; q < 3; ++q) {
    long double sp = 0;
    for (int j = 0; j < L; ++j) sp += one_boson(bb[q], (double)j * dtau, beta);
    phi[q] = -fact * sp + 1.0L / (-(long double)bb[q]);
  }
  long double kap[3];
  for (int p = 0; p < 3; ++p) kap[p] = phi[0] * A[0][p] + phi[1] * A[1][p] + phi[2] * A[2][p];
  std::vector<long double> c(L + 1, 0.0L);
  for (int j = 0; j < L; ++j) c[j] = fact;
  // m1
  c[0] -= kap[0];
  c[L] += kap[0];
  // m2, m3 through the boundary derivatives
  for (int m = 1; m <= 8; ++m) {
    const long double tm = (long double)m * dtau;
    const long double v0 = vinv8[0 * 8 + m - 1] / tm, v1 = vinv8[1 * 8 + m - 1] / tm;
    const long double k2 = kap[1], k3 = kap[2] * (-2.0L / dtau);
    // gl: +x_m - x_0 ; gr: +x_L - x_{L-m} ; m2 = gl0 - gr0 ; m3 = -(2/dtau)(gl1 + gr1)
    c[m] += k2 * v0 + k3 * v1;
    c[0] -= k2 * v0 + k3 * v1;
    c[L] += -k2 * v0 + k3 * v1;
    c[L - m] -= -k2 * v0 + k3 * v1;
  }
  return c;
}

// Tables of the nu = 0 bubble evaluated in Matsubara space (static_mats.hip)
// for a fermionic mesh of n_max = n_iw and L = ntau - 1, in long double:
// per n >= 0 the row [cos(w_n tau_j) / beta (j = 0..8), sin(w_n tau_j) / beta
// (j = 1..8), Re Y_q, Im Y_q, Re Y'_q, Im Y'_q (q = 0..3)] with
//   Y_qn  = -(1/beta) sum_{j=0..L} c_j Z_jq exp(+i w_n tau_j)
//   Y'_qn =  (1/beta) sum_{j=0..L} c_j Z_{L-j,q} exp(-i w_n tau_j),
// Z_jq = D_q(tau_j) (j < L, q < 3), Z_Lq = -D_q(0), Z_L3 = -1, Z_j3 = 0 (j < L);
// and e_j = -(dc_j + dc_{L-j}), f_j = dc_j - dc_{L-j}, dc_j = c_j - beta/L [j < L],
// Q_pq = sum_j c_j Z_jp Z_{L-j,q}.  Needs L >= 17 (disjoint boundary sets).
void static_mats_build(const double *vinv8, double beta, int n_iw, int L, std::vector<double> &t,
                       StaticMatsHost &d) {
  {
    typedef long double ld;
    const std::vector<ld> c = static_weights_ld(vinv8, beta, L);
    const std::vector<ld> D = dtail_values_ld(kFermion, beta, n_iw, L);
    const ld lb = beta;
    std::vector<ld> Z((size_t)(L + 1) * 4, 0.0L);
    for (int j = 0; j < L; ++j)
      for (int q = 0; q < 3; ++q) Z[(size_t)j * 4 + q] = D[(size_t)q * L + j];
    for (int q = 0; q < 3; ++q) Z[(size_t)L * 4 + q] = -D[(size_t)q * L];
    Z[(size_t)L * 4 + 3] = -1.0L;
    std::vector<ld> cs(2 * (size_t)L), sn(2 * (size_t)L);  // angle pi k / L, k < 2L
    for (long k = 0; k < 2L * L; ++k) {
      cs[k] = cosl(kPiL * (ld)k / (ld)L);
      sn[k] = sinl(kPiL * (ld)k / (ld)L);
    }
    t.assign((size_t)n_iw * 33, 0.0);
    for (int n = 0; n < n_iw; ++n) {
      double *row = &t[(size_t)n * 33];
      auto ang = [&](long j) { return ((2L * n + 1) * j) % (2L * L); };  // w_n tau_j = pi ang / L
      for (int j = 0; j <= 8; ++j) row[j] = (double)(cs[ang(j)] / lb);
      for (int j = 1; j <= 8; ++j) row[8 + j] = (double)(sn[ang(j)] / lb);
      for (int q = 0; q < 4; ++q) {
        ld yr = 0, yi = 0, pr = 0, pi = 0;
        for (long j = 0; j <= L; ++j) {
          const long a = ang(j);
          const ld wz = c[j] * Z[(size_t)j * 4 + q], wp = c[j] * Z[(size_t)(L - j) * 4 + q];
          yr += wz * cs[a];
          yi += wz * sn[a];
          pr += wp * cs[a];
          pi += wp * sn[a];
        }
        row[17 + q] = (double)(-yr / lb);
        row[21 + q] = (double)(-yi / lb);
        row[25 + q] = (double)(pr / lb);
        row[29 + q] = (double)(-pi / lb);
      }
    }
    d = StaticMatsHost{};
    d.inv_beta = 1.0 / beta;
    const ld fact = lb / (ld)L;
    auto dc = [&](long j) { return c[j] - (j < L ? fact : 0.0L); };
    d.e[0] = (double)(-(dc(0) + dc(L)));
    d.f[0] = 0.0;
    for (int j = 1; j <= 8; ++j) {
      d.e[j] = (double)(-(dc(j) + dc(L - j)));
      d.f[j] = (double)(dc(j) - dc(L - j));
    }
    for (int p = 0; p < 4; ++p)
      for (int q = 0; q < 4; ++q) {
        ld acc = 0;
        for (long j = 0; j <= L; ++j) acc += c[j] * Z[(size_t)j * 4 + p] * Z[(size_t)(L - j) * 4 + q];
        d.Q[p * 4 + q] = (double)acc;
      }
  }
}

}  // namespace tprf_host
