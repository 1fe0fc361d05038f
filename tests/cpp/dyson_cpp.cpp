// A caller of the real-frequency and local Dyson signatures of
// include/tprf_hip.hpp (gf.hpp:103,204,225,245,267,316,325): reads one problem
// from a binary file, runs
//   lattice_dyson_g0_fk(mu, e_k, mesh, delta), lattice_dyson_g_fk(mu, e_k,
//   sigma_fk, delta), lattice_dyson_g_fk(mu, e_k, sigma_f, delta),
//   lattice_dyson_g_f(mu, e_k, sigma_f, delta), lattice_dyson_g_w(mu, e_k,
//   sigma_w) and fourier_fr_to_fk(fourier_fk_to_fr(g0_fk)),
// and writes the results to a binary file.
//
// Input: int32 dims[3], norb, n_w, n_max, 0, 0; float64 omega_min, omega_max,
//        delta, beta, mu; complex128 e_k[nk][n^2], sigma_fk[n_w][nk][n^2],
//        sigma_f[n_w][n^2], sigma_w[2 n_max][n^2].
// Output: complex128 g0_fk, g_fk (sigma_fk), g_fk (sigma_f), g_f, g_w, round trip.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "tprf_hip.hpp"

using namespace tprf_hip;

static void write(FILE *f, const std::vector<dcomplex> &v) {
  if (fwrite(v.data(), sizeof(dcomplex), v.size(), f) != v.size()) throw std::runtime_error("short write");
}

template <class G>
static G load(FILE *in, G g) {
  std::vector<dcomplex> h((size_t)g.size());
  if (fread(h.data(), sizeof(dcomplex), h.size(), in) != h.size()) throw std::runtime_error("short data");
  g.from_host(h);
  return g;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  try {
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) throw std::runtime_error("cannot open input");
    int32_t ip[8];
    double dp[5];
    if (fread(ip, sizeof(int32_t), 8, in) != 8 || fread(dp, sizeof(double), 5, in) != 5)
      throw std::runtime_error("short header");
    const int n = ip[3];
    mesh::brzone kmesh{{{ip[0], ip[1], ip[2]}}};
    const mesh::refreq fmesh{dp[0], dp[1], ip[4]};
    const mesh::imfreq wmesh{dp[3], mesh::Fermion, ip[5]};
    const double delta = dp[2], mu = dp[4];
    e_k_t e_k = load(in, e_k_t(kmesh, mesh::none{}, n));
    g_fk_t sigma_fk = load(in, g_fk_t(fmesh, kmesh, n));
    g_f_t sigma_f = load(in, g_f_t(fmesh, mesh::none{}, n));
    g_w_t sigma_w = load(in, g_w_t(wmesh, mesh::none{}, n));
    std::fclose(in);

    g_fk_t g0_fk = lattice_dyson_g0_fk(mu, e_k, fmesh, delta);
    g_fk_t g_fk1 = lattice_dyson_g_fk(mu, e_k, sigma_fk, delta);
    g_fk_t g_fk2 = lattice_dyson_g_fk(mu, e_k, sigma_f, delta);
    g_f_t g_f = lattice_dyson_g_f(mu, e_k, sigma_f, delta);
    g_w_t g_w = lattice_dyson_g_w(mu, e_k, sigma_w);
    g_fk_t back = fourier_fr_to_fk(fourier_fk_to_fr(g0_fk));

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    for (const auto &v : {g0_fk.to_host(), g_fk1.to_host(), g_fk2.to_host(), g_f.to_host(), g_w.to_host(),
                          back.to_host()})
      write(out, v);
    std::fclose(out);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "dyson_cpp: %s\n", e.what());
    return 1;
  }
  return 0;
}
