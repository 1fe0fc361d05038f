// A caller of the GW and Eliashberg signatures of include/tprf_hip.hpp
// (gw.hpp, eliashberg.hpp): reads one problem from a binary file, runs
//   gw_sigma(W_wk, g_wk), gw_sigma(v_k, g_wk), construct_phi_wk(chi, U),
//   eliashberg_product(Gamma_big, g, delta) and, after the split of Gamma and
//   dynamic_and_constant_to_tr, eliashberg_product_fft(Gamma_dyn_tr,
//   Gamma_const_r, g, delta),
// and writes the results to a binary file.
//
// Input: int32 dims[3], norb, n_max (fermionic mesh of g and delta, bosonic of
//        W and Gamma), n_big (bosonic mesh of Gamma_big), 0, 0; float64 beta;
//        complex128 g_wk[2 n_max][nk][n^2], W_wk[2 n_max - 1][nk][n^4],
//        v_k[nk][n^4], delta_wk[2 n_max][nk][n^2], Gamma_wk[2 n_max - 1][nk][n^4],
//        Gamma_big[2 n_big - 1][nk][n^4], U[n^4].
// Output: complex128 sigma_wk, sigma_k (static), phi_wk (of Gamma_wk),
//         delta (summation), delta (FFT).
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
    double beta;
    if (fread(ip, sizeof(int32_t), 8, in) != 8 || fread(&beta, sizeof(double), 1, in) != 1)
      throw std::runtime_error("short header");
    const int n = ip[3], n_max = ip[4], n_big = ip[5];
    mesh::brzone kmesh{{{ip[0], ip[1], ip[2]}}};
    const mesh::imfreq wf{beta, mesh::Fermion, n_max}, wb{beta, mesh::Boson, n_max}, wbig{beta, mesh::Boson, n_big};
    g_wk_t g_wk = load(in, g_wk_t(wf, kmesh, n));
    chi_wk_t W_wk = load(in, chi_wk_t(wb, kmesh, n));
    chi_k_t v_k = load(in, chi_k_t(kmesh, mesh::none{}, n));
    g_wk_t delta_wk = load(in, g_wk_t(wf, kmesh, n));
    chi_wk_t Gamma_wk = load(in, chi_wk_t(wb, kmesh, n));
    chi_wk_t Gamma_big = load(in, chi_wk_t(wbig, kmesh, n));
    std::vector<dcomplex> U((size_t)n * n * n * n);
    if (fread(U.data(), sizeof(dcomplex), U.size(), in) != U.size()) throw std::runtime_error("short data");
    std::fclose(in);

    g_wk_t sigma_wk = gw_sigma(W_wk, g_wk);
    e_k_t sigma_k = gw_sigma(v_k, g_wk);
    chi_wk_t phi_wk = construct_phi_wk(Gamma_wk, array4_view{U.data(), n});
    g_wk_t d_sum = eliashberg_product(Gamma_big, g_wk, delta_wk);
    auto [dyn_wk, const_k] = split_into_dynamic_wk_and_constant_k(Gamma_wk);
    auto [dyn_tr, const_r] = dynamic_and_constant_to_tr(dyn_wk, const_k);
    g_wk_t d_fft = eliashberg_product_fft(dyn_tr, const_r, g_wk, delta_wk);

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    write(out, sigma_wk.to_host());
    write(out, sigma_k.to_host());
    write(out, phi_wk.to_host());
    write(out, d_sum.to_host());
    write(out, d_fft.to_host());
    std::fclose(out);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "consumers_cpp: %s\n", e.what());
    return 1;
  }
  return 0;
}
