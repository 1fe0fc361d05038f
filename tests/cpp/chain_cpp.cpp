// A caller of the reference's C++ signatures (include/tprf_hip.hpp): reads one
// lattice (e_k, U and the mesh parameters) from a binary file, runs the bubble
// + RPA chain through lattice_dyson_g0_wk, fourier_wk_to_wr, fourier_wr_to_tr,
// chi0_wr_from_grt_PH, chi_wk_from_chi_wr and solve_rpa_PH, the driver
// imtime_bubble_chi0_wk without save_memory, and the same lattice as a batch of
// two through forward_batched, and writes the results to a binary file.
//
// Input: int32 dims[3], norb, n_iw, nb, flags; float64 beta, mu;
//        complex128 e_k[nk][n][n], U[n][n][n][n].
// Output: complex128 chi0_wk (fused), chi0_wk (stored product), chi_wk,
//         chi_wk of batch instance 0 (flags), chi_wk of instance 1 (no flags),
//         each [2 nb - 1][nk][n^4].
#include <cstdio>
#include <cstdint>
#include <vector>

#include "tprf_hip.hpp"

using namespace tprf_hip;

static void write(FILE *f, const std::vector<dcomplex> &v) {
  if (fwrite(v.data(), sizeof(dcomplex), v.size(), f) != v.size()) throw std::runtime_error("short write");
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
    double dp[2];
    if (fread(ip, sizeof(int32_t), 8, in) != 8 || fread(dp, sizeof(double), 2, in) != 2)
      throw std::runtime_error("short header");
    const int n = ip[3], n_iw = ip[4], nb = ip[5], flags = ip[6];
    const double beta = dp[0], mu = dp[1];
    mesh::brzone kmesh{{{ip[0], ip[1], ip[2]}}};
    const long nk = kmesh.size();
    std::vector<dcomplex> ek((size_t)(nk * n * n)), U((size_t)n * n * n * n);
    if (fread(ek.data(), sizeof(dcomplex), ek.size(), in) != ek.size() ||
        fread(U.data(), sizeof(dcomplex), U.size(), in) != U.size())
      throw std::runtime_error("short data");
    std::fclose(in);

    e_k_t e_k(kmesh, mesh::none{}, n);
    e_k.from_host(ek);
    g_wk_t g_wk = lattice_dyson_g0_wk(mu, e_k, mesh::imfreq{beta, mesh::Fermion, n_iw});
    g_wr_t g_wr = fourier_wk_to_wr(g_wk);
    g_tr_t g_tr = fourier_wr_to_tr(g_wr);
    chi_wk_t chi0_wk = chi_wk_from_chi_wr(chi0_wr_from_grt_PH(g_tr, nb));
    chi_wk_t chi0_wk_stored = imtime_bubble_chi0_wk(g_wk, nb, false);
    chi_wk_t chi_wk = solve_rpa_PH(chi0_wk, array4_view{U.data(), n});

    tprf_chain_desc d{};
    d.beta = beta;
    d.n_iw = n_iw;
    d.ntau = -1;
    d.nb = nb;
    for (int i = 0; i < 3; ++i) d.dims[i] = ip[i];
    d.norb = n;
    d.mu = mu;
    d.flags = flags;
    std::vector<chain_instance> batch;
    batch.emplace_back(d, ek, U);
    d.flags = 0;
    batch.emplace_back(d, ek, U);
    forward_batched(batch, 2);

    FILE *out = std::fopen(argv[2], "wb");
    if (!out) throw std::runtime_error("cannot open output");
    write(out, chi0_wk.to_host());
    write(out, chi0_wk_stored.to_host());
    write(out, chi_wk.to_host());
    write(out, batch[0].download(batch[0].chi_wk, batch[0].counts[4]));
    write(out, batch[1].download(batch[1].chi_wk, batch[1].counts[4]));
    std::fclose(out);
  } catch (const error &e) {
    std::fprintf(stderr, "tprf_hip::error %d: %s\n", e.code, e.what());
    return 1;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
