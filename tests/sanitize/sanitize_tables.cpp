// Checks of the host-side tables (tprf_amd/csrc/host_tables.cpp), built with
// the address and undefined-behaviour sanitizers: every builder runs on a set
// of meshes, and its output is held to what the mathematics fixes.
//   tail fit: pinv V = 1, and a polynomial in z = w_max/(i w) is recovered;
//   nu = 0 weights: sum_j c_j = beta (a constant), and the transform of each
//     model pole 1/(i nu - b), b in {-1/2, -1, 1}, is -1/b (fine tau meshes);
//   D_q tails and the Matsubara bubble tables: finite.
// Exit status 0 when every check holds.
#include <cmath>
#include <complex>
#include <cstdio>
#include <string>
#include <vector>

#include "host_tables.hpp"

using namespace tprf_host;

static int failures = 0;
static void check(bool ok, const char *what, double v) {
  if (!ok) {
    std::printf("FAIL %s (%.3e)\n", what, v);
    ++failures;
  }
}

int main() {
  double vinv[64];
  vandermonde_inverse8(vinv);
  const struct { int stat; double beta; int n_max; } fits[] = {
      {kFermion, 10.0, 64}, {kFermion, 40.0, 256}, {kFermion, 20.0, 512}, {kBoson, 10.0, 64}, {kBoson, 1.0, 30},
      {kFermion, 300.0, 500}};
  for (const auto &f : fits) {
    TailTables t;
    std::string msg;
    if (tail_fit_build(f.stat, f.beta, f.n_max, t, msg)) {
      check(false, msg.c_str(), 0);
      continue;
    }
    double e = 0;
    for (int a = 0; a < t.ncoef; ++a)
      for (int b = 0; b < t.ncoef; ++b) {
        std::complex<double> s = 0;
        for (int r = 0; r < t.nfit; ++r) s += t.pinv[(size_t)a * t.nfit + r] * t.vander[(size_t)r * t.ncoef + b];
        e = std::fmax(e, std::abs(s - (a == b ? 1.0 : 0.0)));
      }
    check(e < 1e-9, "tail fit: pinv V = 1", e);
    int bad = 0;
    for (int i = 0; i < (int)t.fitpos.size(); ++i)
      if (t.fitpos[i] >= 0 && t.idx[t.fitpos[i]] != i) ++bad;
    check(bad == 0, "tail fit: fitpos inverts idx", bad);
    check(std::fabs(t.wpow[1] - t.wmax) <= 1e-12 * t.wmax, "tail fit: wpow", t.wpow[1]);
  }
  // tail fit of a mesh too short for it reports an error instead of failing
  {
    TailTables t;
    std::string msg;
    check(tail_fit_build(kFermion, 10.0, 2, t, msg) != 0, "tail fit: mesh of 4 points refused", 0);
  }
  for (int L : {18, 64, 384, 1536, 3072}) {
    const double beta = 20.0;
    std::vector<long double> c = static_weights_ld(vinv, beta, L);
    long double s = 0;
    for (long double x : c) s += x;
    check(std::fabs((double)(s - beta)) < 1e-12 * beta, "static weights: sum = beta", (double)(s - beta));
    // the model poles pass through the 8-point boundary-derivative fit of
    // m2, m3, exact only to O(dtau^7): checked where dtau is small
    if (L >= 384)
      for (double b : {-0.5, -1.0, 1.0}) {
        long double v = 0;
        for (int j = 0; j <= L; ++j) v += c[j] * one_boson(b, j * beta / L, beta);
        check(std::fabs((double)(v + 1.0 / b)) < 1e-10, "static weights: model pole", (double)(v + 1.0 / b));
      }
  }
  for (int stat : {kFermion, kBoson})
    for (int L : {64, 1536}) {
      std::vector<long double> d = dtail_values_ld(stat, 10.0, 16, L);
      bool fin = d.size() == 3 * (size_t)L;
      for (long double x : d) fin = fin && std::isfinite((double)x) && std::fabs((double)x) < 10;
      check(fin, "D_q tails finite", 0);
    }
  {
    std::vector<double> tab;
    StaticMatsHost sm;
    static_mats_build(vinv, 20.0, 256, 1536, tab, sm);
    bool fin = tab.size() == (size_t)256 * 33 && std::isfinite(sm.inv_beta);
    for (double x : tab) fin = fin && std::isfinite(x);
    for (int j = 0; j < 9; ++j) fin = fin && std::isfinite(sm.e[j]) && std::isfinite(sm.f[j]);
    for (int j = 0; j < 16; ++j) fin = fin && std::isfinite(sm.Q[j]);
    check(fin, "static Matsubara tables finite", 0);
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
