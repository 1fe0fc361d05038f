// host_tables.hpp -- host-side tables of the Matsubara transforms, built in
// long double from the mathematics alone (no HIP): the tail-fit operator, the
// out-of-mesh pole tails D_q(tau), the nu = 0 weights of the bosonic
// tau -> nu map, and the tables of the nu = 0 bubble in Matsubara space.
// capi.cpp uploads them; a CPU build of this file under the address and
// undefined-behaviour sanitizers checks them (tests/cpp/sanitize).
#pragma once

#include <complex>
#include <string>
#include <vector>

namespace tprf_host {

constexpr long double kPiL = 3.141592653589793238462643383279502884L;
constexpr int kFermion = 0, kBoson = 1;  // TPRF_FERMION / TPRF_BOSON

// oneFermion / oneBoson with a = 1 (fourier_matsubara.cpp:31-37)
double one_fermion(double b, double tau, double beta);
double one_boson(double b, double tau, double beta);

// exact inverse of V_{m,j} = m^{j-1}, m, j = 1..8, row-major [j][m]
void vandermonde_inverse8(double *vinv);

// Least-squares tail fit of one Matsubara mesh: nfit data indices into the
// mesh, pinv [ncoef][nfit] = R^-1 Q^H, vander [nfit][ncoef] in
// z = omega_max / (i w), wpow = omega_max^p (p < 4), fitpos [mesh] = position
// of each mesh index in idx or -1.
struct TailTables {
  int nfit = 0, ncoef = 0;
  double wmax = 0;
  std::vector<int> idx, fitpos;
  std::vector<std::complex<double>> pinv, vander;
  double wpow[4] = {1, 0, 0, 0};
};
// 0 on success, else 1 and msg (mesh too small / fewer than 4 coefficients)
int tail_fit_build(int statistic, double beta, int n_max, TailTables &t, std::string &msg);

// D_q(tau_j), j < L, q < 3 ([q][j]): what the pole model adds in tau beyond
// its own in-mesh transform
std::vector<long double> dtail_values_ld(int statistic, double beta, int n_max, int L);

// weights c_j (j = 0..L) of the bosonic tau -> nu map at nu = 0
std::vector<long double> static_weights_ld(const double *vinv8, double beta, int L);

// the Matsubara-space nu = 0 bubble: table rows of width 33 per n < n_iw and
// the small bilinear part (e, f, Q)
struct StaticMatsHost {
  double inv_beta = 0;
  double e[9] = {}, f[9] = {}, Q[16] = {};
};
void static_mats_build(const double *vinv8, double beta, int n_iw, int L, std::vector<double> &tab,
                       StaticMatsHost &sm);

}  // namespace tprf_host
