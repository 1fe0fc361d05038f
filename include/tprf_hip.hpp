/*
 * tprf_hip.hpp -- C++ signatures of the TRIQS/tprf 3.2 lattice bubble + RPA hot
 * path over the C ABI of libtprf_hip.so (tprf_hip.h).  Header only.
 *
 * The reference's functions (c++/triqs_tprf/lattice/gf.hpp, chi_imtime.hpp,
 * rpa.hpp, lindhard_chi00.hpp) take TRIQS gf const views and return owning gf
 * objects.  TRIQS is not a dependency here, so the same names and argument
 * lists are declared over a light gf type: its meshes (imfreq, imtime,
 * brzone, cyclat) and a complex128 array in DEVICE memory laid out as the
 * reference stores gf data, [mesh0][mesh1][target...] row-major
 * (types.hpp:163-214).  A gf either owns its memory (hipMalloc'd, freed with
 * the last copy) or views a caller's device pointer.
 *
 * Every call goes to a per-thread context of the current device, ordered on
 * the stream set by tprf_hip::set_stream (default: the null stream).  Calls
 * are asynchronous, except that the functions whose reference raises or warns
 * on a condition found in the data (the tail fit, fourier_matsubara.cpp:227-239;
 * a singular inversion) synchronise the stream to read the status.  Errors
 * throw tprf_hip::error (a std::runtime_error) with the library's message,
 * where the reference throws TRIQS_RUNTIME_ERROR; warnings go to std::cerr.
 *
 * Link with -ltprf_hip -lamdhip64 (the host compiler needs
 * -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include).
 */
#ifndef TPRF_HIP_HPP
#define TPRF_HIP_HPP

#include <hip/hip_runtime_api.h>

#include <array>
#include <cmath>
#include <complex>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "tprf_hip.h"

namespace tprf_hip {

using dcomplex = std::complex<double>;

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

namespace detail {

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw error(TPRF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

struct stream_state {
  hipStream_t stream = nullptr;
};
inline stream_state &state() {
  thread_local stream_state s;
  return s;
}

// one context per (thread, device), destroyed with the thread
struct ctx_table {
  std::vector<tprf_ctx *> ctx;
  ~ctx_table() {
    for (tprf_ctx *c : ctx)
      if (c) tprf_ctx_destroy(c);
  }
};

inline tprf_ctx *context() {
  thread_local ctx_table table;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if ((int)table.ctx.size() <= dev) table.ctx.resize(dev + 1, nullptr);
  if (!table.ctx[dev]) {
    tprf_ctx *c = nullptr;
    const int st = tprf_ctx_create(&c, dev);
    if (st != TPRF_OK) throw error(st, "tprf_ctx_create failed");
    table.ctx[dev] = c;
  }
  tprf_ctx_set_stream(table.ctx[dev], (void *)state().stream);
  return table.ctx[dev];
}

inline void check(int st, tprf_ctx *c, const char *what) {
  if (st != TPRF_OK) throw error(st, std::string(what) + ": " + tprf_ctx_last_error(c));
}

// the reference's fit error / tail warnings (fourier_matsubara.cpp:227-239), and
// the singular-inversion status
inline void check_tail(tprf_ctx *c, bool fitted) {
  double err = 0, m0 = 0;
  check(tprf_ctx_tail_status(c, &err, &m0), c, "status");
  if (!fitted) return;
  if (err >= 1e-3)
    throw error(TPRF_ERR_TAIL_FIT,
                "ERROR: High frequency moments have an error greater than 1e-3.\n  Error = " + std::to_string(err));
  if (err > 1e-6)
    std::cerr << "WARNING: High frequency moments have an error greater than 1e-6.\n Error = " << err << "\n";
  if (m0 > 1e-6) std::cerr << "WARNING: High frequency tail is not zero: " << m0 << "\n";
}

}  // namespace detail

/* Stream of the calls made by this thread (hipStream_t; nullptr = null stream). */
inline void set_stream(hipStream_t s) { detail::state().stream = s; }

// ---- meshes (types.hpp; TRIQS mesh conventions) --------------------------
namespace mesh {
enum statistic_enum { Fermion = TPRF_FERMION, Boson = TPRF_BOSON };

/* Matsubara mesh, n_max = N: 2N fermionic points (n = -N..N-1) or 2N-1 bosonic
 * (m = -(N-1)..N-1). */
struct imfreq {
  double beta = 1;
  statistic_enum statistic = Fermion;
  int n_max = 1;
  long size() const { return statistic == Fermion ? 2L * n_max : 2L * n_max - 1; }
  bool operator==(const imfreq &o) const { return beta == o.beta && statistic == o.statistic && n_max == o.n_max; }
};
/* Imaginary-time mesh: n_tau points tau_j = j beta / (n_tau - 1). */
struct imtime {
  double beta = 1;
  statistic_enum statistic = Fermion;
  int n_tau = 2;
  long size() const { return n_tau; }
};
struct lattice_dims {
  std::array<int, 3> dims{1, 1, 1};
  long size() const { return (long)dims[0] * dims[1] * dims[2]; }
};
/* Brillouin-zone and cyclic-lattice meshes: d0 x d1 x d2, linear index
 * (i0 d1 + i1) d2 + i2 (d2 = 1 in 2-D). */
struct brzone : lattice_dims {};
struct cyclat : lattice_dims {};
/* Real-frequency mesh: n_w points omega_min .. omega_max, both included. */
struct refreq {
  double omega_min = -1, omega_max = 1;
  int n_w = 2;
  long size() const { return n_w; }
  double value(long i) const {
    return n_w == 1 ? omega_min : omega_min + (omega_max - omega_min) * (double)i / (double)(n_w - 1);
  }
};
/* The single point of a mesh-less factor (e_k, Sigma(w) have one mesh). */
struct none {
  long size() const { return 1; }
};
}  // namespace mesh

/* gf on the product mesh M0 x M1 (M1 = mesh::none for a single mesh) with a
 * target of rank R (2: matrix_valued n x n; 4: tensor_valued<4> n^4). */
template <class M0, class M1, int R>
class gf {
 public:
  gf() = default;
  /* Owning: device memory for mesh0 x mesh1 x n^R entries (uninitialised). */
  gf(M0 m0, M1 m1, int norb) : m0_(m0), m1_(m1), n_(norb) {
    dcomplex *p = nullptr;
    detail::hip_check(hipMalloc((void **)&p, sizeof(dcomplex) * (size_t)size()), "hipMalloc");
    data_ = std::shared_ptr<dcomplex>(p, [](dcomplex *q) { (void)hipFree(q); });
  }
  /* View of a caller's device array (not freed here). */
  gf(M0 m0, M1 m1, int norb, dcomplex *device_data) : m0_(m0), m1_(m1), n_(norb) {
    data_ = std::shared_ptr<dcomplex>(device_data, [](dcomplex *) {});
  }
  const M0 &mesh0() const { return m0_; }
  const M1 &mesh1() const { return m1_; }
  int norb() const { return n_; }
  long target_size() const {
    long t = 1;
    for (int i = 0; i < R; ++i) t *= n_;
    return t;
  }
  long size() const { return m0_.size() * m1_.size() * target_size(); }
  dcomplex *data() const { return data_.get(); }

  /* Synchronous copies between host and device (on the current thread's stream). */
  std::vector<dcomplex> to_host() const {
    std::vector<dcomplex> h((size_t)size());
    detail::hip_check(hipMemcpyAsync(h.data(), data(), sizeof(dcomplex) * h.size(), hipMemcpyDeviceToHost,
                                     detail::state().stream),
                      "hipMemcpyAsync");
    detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");
    return h;
  }
  void from_host(const std::vector<dcomplex> &h) {
    if ((long)h.size() != size()) throw error(TPRF_ERR_ARG, "gf::from_host: size mismatch");
    detail::hip_check(hipMemcpyAsync(data(), h.data(), sizeof(dcomplex) * h.size(), hipMemcpyHostToDevice,
                                     detail::state().stream),
                      "hipMemcpyAsync");
    detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");
  }

 private:
  M0 m0_{};
  M1 m1_{};
  int n_ = 1;
  std::shared_ptr<dcomplex> data_;
};

// the reference's type names (types.hpp:163-214); const views are const references
using e_k_t = gf<mesh::brzone, mesh::none, 2>;
using g_w_t = gf<mesh::imfreq, mesh::none, 2>;
using g_wk_t = gf<mesh::imfreq, mesh::brzone, 2>;
using g_wr_t = gf<mesh::imfreq, mesh::cyclat, 2>;
using g_tr_t = gf<mesh::imtime, mesh::cyclat, 2>;
using chi_tr_t = gf<mesh::imtime, mesh::cyclat, 4>;
using chi_wr_t = gf<mesh::imfreq, mesh::cyclat, 4>;
using chi_wk_t = gf<mesh::imfreq, mesh::brzone, 4>;
using g_f_t = gf<mesh::refreq, mesh::none, 2>;
using g_fk_t = gf<mesh::refreq, mesh::brzone, 2>;
using g_fr_t = gf<mesh::refreq, mesh::cyclat, 2>;
using g_f_cvt = const g_f_t &;
using g_fk_cvt = const g_fk_t &;
using g_fr_cvt = const g_fr_t &;
using e_k_cvt = const e_k_t &;
using g_w_cvt = const g_w_t &;
using g_wk_cvt = const g_wk_t &;
using g_wr_cvt = const g_wr_t &;
using g_tr_cvt = const g_tr_t &;
using chi_tr_cvt = const chi_tr_t &;
using chi_wr_cvt = const chi_wr_t &;
using chi_wk_cvt = const chi_wk_t &;
using chi_wk_vt = const chi_wk_t &;

/* The rank-4 interaction of solve_rpa_PH (array_contiguous_view<dcomplex, 4>,
 * rpa.hpp:43): n^4 entries U[a][b][c][d] in HOST memory. */
struct array4_view {
  const dcomplex *data;
  int n;
};

namespace detail {
inline mesh::cyclat to_r(const mesh::brzone &k) { return mesh::cyclat{{k.dims}}; }
inline mesh::brzone to_k(const mesh::cyclat &r) { return mesh::brzone{{r.dims}}; }
inline const tprf_c128 *c(const dcomplex *p) { return reinterpret_cast<const tprf_c128 *>(p); }
inline tprf_c128 *c(dcomplex *p) { return reinterpret_cast<tprf_c128 *>(p); }
}  // namespace detail

// ---- a1: Dyson (gf.hpp:44,83,123; gf.cpp:41-132) ---------------------------
inline g_wk_t lattice_dyson_g0_wk(double mu, e_k_cvt e_k, mesh::imfreq mesh) {
  if (mesh.statistic != mesh::Fermion) throw error(TPRF_ERR_ARG, "lattice_dyson_g0_wk: fermionic mesh required");
  tprf_ctx *cx = detail::context();
  g_wk_t g(mesh, e_k.mesh0(), e_k.norb());
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_lattice_dyson_g_wk(cx, mesh.beta, mesh.n_max, mu, e_k.mesh0().size(), e_k.norb(),
                                        detail::c(e_k.data()), TPRF_SIGMA_NONE, nullptr, detail::c(g.data())),
                cx, "lattice_dyson_g0_wk");
  detail::check_tail(cx, false);
  return g;
}

inline g_wk_t lattice_dyson_g_wk(double mu, e_k_cvt e_k, g_w_cvt sigma_w) {
  tprf_ctx *cx = detail::context();
  const mesh::imfreq &wm = sigma_w.mesh0();
  g_wk_t g(wm, e_k.mesh0(), e_k.norb());
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_lattice_dyson_g_wk(cx, wm.beta, wm.n_max, mu, e_k.mesh0().size(), e_k.norb(),
                                        detail::c(e_k.data()), TPRF_SIGMA_W, detail::c(sigma_w.data()),
                                        detail::c(g.data())),
                cx, "lattice_dyson_g_wk");
  detail::check_tail(cx, false);
  return g;
}

inline g_wk_t lattice_dyson_g_wk(double mu, e_k_cvt e_k, g_wk_cvt sigma_wk) {
  tprf_ctx *cx = detail::context();
  const mesh::imfreq &wm = sigma_wk.mesh0();
  if (sigma_wk.mesh1().size() != e_k.mesh0().size()) throw error(TPRF_ERR_ARG, "lattice_dyson_g_wk: k-mesh mismatch");
  g_wk_t g(wm, e_k.mesh0(), e_k.norb());
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_lattice_dyson_g_wk(cx, wm.beta, wm.n_max, mu, e_k.mesh0().size(), e_k.norb(),
                                        detail::c(e_k.data()), TPRF_SIGMA_WK, detail::c(sigma_wk.data()),
                                        detail::c(g.data())),
                cx, "lattice_dyson_g_wk");
  detail::check_tail(cx, false);
  return g;
}

// ---- a1 on given frequencies: real frequencies and the local G (gf.hpp:103,204,225,245,267)
namespace detail {
/* The frequencies z_i (+ i delta) of a mesh as a device array. */
inline std::shared_ptr<dcomplex> device_freqs(const std::vector<dcomplex> &z) {
  dcomplex *p = nullptr;
  hip_check(hipMalloc((void **)&p, sizeof(dcomplex) * z.size()), "hipMalloc");
  std::shared_ptr<dcomplex> d(p, [](dcomplex *q) { (void)hipFree(q); });
  hip_check(hipMemcpyAsync(p, z.data(), sizeof(dcomplex) * z.size(), hipMemcpyHostToDevice, state().stream),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(state().stream), "hipStreamSynchronize");
  return d;
}
inline std::vector<dcomplex> freqs(const mesh::refreq &m, double delta) {
  std::vector<dcomplex> z((size_t)m.size());
  for (long i = 0; i < m.size(); ++i) z[(size_t)i] = dcomplex(m.value(i), delta);
  return z;
}
inline std::vector<dcomplex> freqs(const mesh::imfreq &m) {
  std::vector<dcomplex> z((size_t)m.size());
  const long first = m.statistic == mesh::Fermion ? -m.n_max : -(m.n_max - 1);
  const int odd = m.statistic == mesh::Fermion ? 1 : 0;
  for (long i = 0; i < m.size(); ++i) z[(size_t)i] = dcomplex(0.0, M_PI * (double)(2 * (first + i) + odd) / m.beta);
  return z;
}
template <class G>
inline void dyson_z(const std::vector<dcomplex> &z, double mu, e_k_cvt e_k, int kind, const dcomplex *sigma, G &out,
                    bool local, const char *what) {
  tprf_ctx *cx = context();
  auto zd = device_freqs(z);
  check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  auto f = local ? tprf_lattice_dyson_g_z : tprf_lattice_dyson_g_zk;
  check(f(cx, (long)z.size(), c(zd.get()), mu, e_k.mesh0().size(), e_k.norb(), c(e_k.data()), kind, c(sigma),
          c(out.data())),
        cx, what);
  check_tail(cx, false);
  hip_check(hipStreamSynchronize(state().stream), "hipStreamSynchronize");  // z is freed on return
}
}  // namespace detail

inline g_fk_t lattice_dyson_g0_fk(double mu, e_k_cvt e_k, mesh::refreq mesh, double delta) {
  g_fk_t g(mesh, e_k.mesh0(), e_k.norb());
  detail::dyson_z(detail::freqs(mesh, delta), mu, e_k, TPRF_SIGMA_NONE, nullptr, g, false, "lattice_dyson_g0_fk");
  return g;
}
inline g_fk_t lattice_dyson_g_fk(double mu, e_k_cvt e_k, g_fk_cvt sigma_fk, double delta) {
  if (sigma_fk.mesh1().size() != e_k.mesh0().size()) throw error(TPRF_ERR_ARG, "lattice_dyson_g_fk: k-mesh mismatch");
  g_fk_t g(sigma_fk.mesh0(), e_k.mesh0(), e_k.norb());
  detail::dyson_z(detail::freqs(sigma_fk.mesh0(), delta), mu, e_k, TPRF_SIGMA_WK, sigma_fk.data(), g, false,
                  "lattice_dyson_g_fk");
  return g;
}
inline g_fk_t lattice_dyson_g_fk(double mu, e_k_cvt e_k, g_f_cvt sigma_f, double delta) {
  g_fk_t g(sigma_f.mesh0(), e_k.mesh0(), e_k.norb());
  detail::dyson_z(detail::freqs(sigma_f.mesh0(), delta), mu, e_k, TPRF_SIGMA_W, sigma_f.data(), g, false,
                  "lattice_dyson_g_fk");
  return g;
}
inline g_w_t lattice_dyson_g_w(double mu, e_k_cvt e_k, g_w_cvt sigma_w) {
  if (sigma_w.mesh0().statistic != mesh::Fermion)
    throw error(TPRF_ERR_ARG, "lattice_dyson_g_w: fermionic mesh required");
  g_w_t g(sigma_w.mesh0(), mesh::none{}, e_k.norb());
  detail::dyson_z(detail::freqs(sigma_w.mesh0()), mu, e_k, TPRF_SIGMA_W, sigma_w.data(), g, true,
                  "lattice_dyson_g_w");
  return g;
}
inline g_f_t lattice_dyson_g_f(double mu, e_k_cvt e_k, g_f_cvt sigma_f, double delta) {
  g_f_t g(sigma_f.mesh0(), mesh::none{}, e_k.norb());
  detail::dyson_z(detail::freqs(sigma_f.mesh0(), delta), mu, e_k, TPRF_SIGMA_W, sigma_f.data(), g, true,
                  "lattice_dyson_g_f");
  return g;
}

// ---- a2 / a6 and inverses: lattice FFTs (gf.hpp:276,286; chi_imtime.hpp:123,141)
inline g_wr_t fourier_wk_to_wr(g_wk_cvt g_wk) {
  tprf_ctx *cx = detail::context();
  g_wr_t out(g_wk.mesh0(), detail::to_r(g_wk.mesh1()), g_wk.norb());
  detail::check(tprf_fourier_wk_to_wr(cx, g_wk.mesh0().size(), g_wk.mesh1().dims.data(), g_wk.target_size(),
                                      detail::c(g_wk.data()), detail::c(out.data())),
                cx, "fourier_wk_to_wr");
  return out;
}
inline g_wk_t fourier_wr_to_wk(g_wr_cvt g_wr) {
  tprf_ctx *cx = detail::context();
  g_wk_t out(g_wr.mesh0(), detail::to_k(g_wr.mesh1()), g_wr.norb());
  detail::check(tprf_fourier_wr_to_wk(cx, g_wr.mesh0().size(), g_wr.mesh1().dims.data(), g_wr.target_size(),
                                      detail::c(g_wr.data()), detail::c(out.data())),
                cx, "fourier_wr_to_wk");
  return out;
}
/* fourier_fk_to_fr / fourier_fr_to_fk (gf.hpp:316,325): the same transforms on a
 * real-frequency mesh. */
inline g_fr_t fourier_fk_to_fr(g_fk_cvt g_fk) {
  tprf_ctx *cx = detail::context();
  g_fr_t out(g_fk.mesh0(), detail::to_r(g_fk.mesh1()), g_fk.norb());
  detail::check(tprf_fourier_wk_to_wr(cx, g_fk.mesh0().size(), g_fk.mesh1().dims.data(), g_fk.target_size(),
                                      detail::c(g_fk.data()), detail::c(out.data())),
                cx, "fourier_fk_to_fr");
  return out;
}
inline g_fk_t fourier_fr_to_fk(g_fr_cvt g_fr) {
  tprf_ctx *cx = detail::context();
  g_fk_t out(g_fr.mesh0(), detail::to_k(g_fr.mesh1()), g_fr.norb());
  detail::check(tprf_fourier_wr_to_wk(cx, g_fr.mesh0().size(), g_fr.mesh1().dims.data(), g_fr.target_size(),
                                      detail::c(g_fr.data()), detail::c(out.data())),
                cx, "fourier_fr_to_fk");
  return out;
}
inline chi_wk_t chi_wk_from_chi_wr(chi_wr_cvt chi_wr) {
  tprf_ctx *cx = detail::context();
  chi_wk_t out(chi_wr.mesh0(), detail::to_k(chi_wr.mesh1()), chi_wr.norb());
  detail::check(tprf_fourier_wr_to_wk(cx, chi_wr.mesh0().size(), chi_wr.mesh1().dims.data(), chi_wr.target_size(),
                                      detail::c(chi_wr.data()), detail::c(out.data())),
                cx, "chi_wk_from_chi_wr");
  return out;
}
inline chi_wr_t chi_wr_from_chi_wk(chi_wk_cvt chi_wk) {
  tprf_ctx *cx = detail::context();
  chi_wr_t out(chi_wk.mesh0(), detail::to_r(chi_wk.mesh1()), chi_wk.norb());
  detail::check(tprf_fourier_wk_to_wr(cx, chi_wk.mesh0().size(), chi_wk.mesh1().dims.data(), chi_wk.target_size(),
                                      detail::c(chi_wk.data()), detail::c(out.data())),
                cx, "chi_wr_from_chi_wk");
  return out;
}

// ---- a3: Matsubara -> imaginary time (gf.hpp:296; chi_imtime.hpp:105) ------
namespace detail {
template <int R>
gf<mesh::imtime, mesh::cyclat, R> w_to_t(const gf<mesh::imfreq, mesh::cyclat, R> &g, int nt) {
  tprf_ctx *cx = context();
  const mesh::imfreq &wm = g.mesh0();
  const int ntau = nt > 0 ? nt : 6 * wm.n_max + 1;  // TRIQS make_adjoint_mesh default
  gf<mesh::imtime, mesh::cyclat, R> out(mesh::imtime{wm.beta, wm.statistic, ntau}, g.mesh1(), g.norb());
  check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  check(tprf_fourier_wr_to_tr(cx, wm.statistic, wm.beta, wm.n_max, ntau, g.mesh1().size() * g.target_size(),
                              c(g.data()), c(out.data()), nullptr),
        cx, "fourier_wr_to_tr");
  check_tail(cx, true);
  return out;
}
template <int R>
gf<mesh::imfreq, mesh::cyclat, R> t_to_w(const gf<mesh::imtime, mesh::cyclat, R> &g, int nw) {
  tprf_ctx *cx = context();
  const mesh::imtime &tm = g.mesh0();
  if (nw <= 0) nw = (tm.n_tau - 1) / 6;
  gf<mesh::imfreq, mesh::cyclat, R> out(mesh::imfreq{tm.beta, tm.statistic, nw}, g.mesh1(), g.norb());
  check(tprf_fourier_tr_to_wr(cx, tm.statistic, tm.beta, tm.n_tau, nw, g.mesh1().size() * g.target_size(),
                              c(g.data()), c(out.data())),
        cx, "fourier_tr_to_wr");
  return out;
}
}  // namespace detail

inline g_tr_t fourier_wr_to_tr(g_wr_cvt g_wr, int nt = -1) { return detail::w_to_t(g_wr, nt); }
inline chi_tr_t chi_tr_from_chi_wr(chi_wr_cvt chi_wr, int ntau = -1) { return detail::w_to_t(chi_wr, ntau); }

// ---- a5: imaginary time -> Matsubara (gf.hpp:306; chi_imtime.hpp:87) -------
inline g_wr_t fourier_tr_to_wr(g_tr_cvt g_tr, int nw = -1) { return detail::t_to_w(g_tr, nw); }
inline chi_wr_t chi_wr_from_chi_tr(chi_tr_cvt chi_tr, int nw) { return detail::t_to_w(chi_tr, nw); }

// ---- a4 / a4' / f1: bubble (chi_imtime.hpp:38,40,54,70) -------------------
inline chi_tr_t chi0_tr_from_grt_PH(g_tr_cvt g_tr) {
  tprf_ctx *cx = detail::context();
  const mesh::imtime &tm = g_tr.mesh0();
  chi_tr_t out(mesh::imtime{tm.beta, mesh::Boson, tm.n_tau}, g_tr.mesh1(), g_tr.norb());
  detail::check(tprf_chi0_tr_from_grt_PH(cx, tm.n_tau, g_tr.mesh1().dims.data(), g_tr.norb(),
                                         detail::c(g_tr.data()), detail::c(out.data())),
                cx, "chi0_tr_from_grt_PH");
  return out;
}
inline chi_wr_t chi0_wr_from_grt_PH(g_tr_cvt g_tr, int nw) {
  tprf_ctx *cx = detail::context();
  const mesh::imtime &tm = g_tr.mesh0();
  chi_wr_t out(mesh::imfreq{tm.beta, mesh::Boson, nw}, g_tr.mesh1(), g_tr.norb());
  detail::check(tprf_chi0_wr_from_grt_PH(cx, tm.beta, tm.n_tau, g_tr.mesh1().dims.data(), g_tr.norb(), nw,
                                         detail::c(g_tr.data()), detail::c(out.data())),
                cx, "chi0_wr_from_grt_PH");
  return out;
}
inline chi_wr_t chi0_w0r_from_grt_PH(g_tr_cvt g_tr) {
  tprf_ctx *cx = detail::context();
  const mesh::imtime &tm = g_tr.mesh0();
  chi_wr_t out(mesh::imfreq{tm.beta, mesh::Boson, 1}, g_tr.mesh1(), g_tr.norb());
  detail::check(tprf_chi0_w0r_from_grt_PH(cx, tm.beta, tm.n_tau, g_tr.mesh1().dims.data(), g_tr.norb(),
                                          detail::c(g_tr.data()), detail::c(out.data())),
                cx, "chi0_w0r_from_grt_PH");
  return out;
}
inline chi_wr_t chi_w0r_from_chi_tr(chi_tr_cvt chi_tr) {
  tprf_ctx *cx = detail::context();
  const mesh::imtime &tm = chi_tr.mesh0();
  chi_wr_t out(mesh::imfreq{tm.beta, mesh::Boson, 1}, chi_tr.mesh1(), chi_tr.norb());
  detail::check(tprf_chi_w0r_from_chi_tr(cx, tm.beta, tm.n_tau, chi_tr.mesh1().size(), chi_tr.norb(),
                                         detail::c(chi_tr.data()), detail::c(out.data())),
                cx, "chi_w0r_from_chi_tr");
  return out;
}

// ---- a7: RPA (rpa.hpp:43) ------------------------------------------------
inline chi_wk_t solve_rpa_PH(chi_wk_vt chi0, array4_view U) {
  tprf_ctx *cx = detail::context();
  const int n = chi0.norb();
  if (U.n != n) throw error(TPRF_ERR_ARG, "solve_rpa_PH: U does not match chi0's target");
  const size_t un = (size_t)n * n * n * n;
  dcomplex *ud = nullptr;
  detail::hip_check(hipMalloc((void **)&ud, sizeof(dcomplex) * un), "hipMalloc");
  std::shared_ptr<dcomplex> hold(ud, [](dcomplex *q) { (void)hipFree(q); });
  detail::hip_check(hipMemcpyAsync(ud, U.data, sizeof(dcomplex) * un, hipMemcpyHostToDevice, detail::state().stream),
                    "hipMemcpyAsync");
  chi_wk_t out(chi0.mesh0(), chi0.mesh1(), n);
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_solve_rpa_PH(cx, chi0.mesh0().size() * chi0.mesh1().size(), n, detail::c(chi0.data()),
                                  detail::c(ud), detail::c(out.data())),
                cx, "solve_rpa_PH");
  detail::check_tail(cx, false);  // synchronises: U's device copy is no longer used
  return out;
}

// ---- f3: Lindhard (lindhard_chi00.hpp:97) ----------------------------------
inline chi_wk_t lindhard_chi00(e_k_cvt e_k, mesh::imfreq mesh, double mu) {
  if (mesh.statistic != mesh::Boson) throw error(TPRF_ERR_ARG, "lindhard_chi00: statistic is incorrect.");
  tprf_ctx *cx = detail::context();
  const long nz = mesh.size();
  std::vector<dcomplex> z((size_t)nz);
  for (long i = 0; i < nz; ++i) z[i] = dcomplex(0.0, 2.0 * M_PI * (double)(i - (mesh.n_max - 1)) / mesh.beta);
  dcomplex *zd = nullptr;
  detail::hip_check(hipMalloc((void **)&zd, sizeof(dcomplex) * z.size()), "hipMalloc");
  std::shared_ptr<dcomplex> hold(zd, [](dcomplex *q) { (void)hipFree(q); });
  detail::hip_check(hipMemcpyAsync(zd, z.data(), sizeof(dcomplex) * z.size(), hipMemcpyHostToDevice,
                                   detail::state().stream),
                    "hipMemcpyAsync");
  chi_wk_t out(mesh, e_k.mesh0(), e_k.norb());
  detail::check(tprf_lindhard_chi00(cx, mesh.beta, mu, nz, detail::c(zd), e_k.mesh0().dims.data(), e_k.norb(),
                                    detail::c(e_k.data()), detail::c(out.data())),
                cx, "lindhard_chi00");
  detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");
  return out;
}

// ---- driver (python/triqs_tprf/lattice_utils.py:191-269) ------------------
/* imtime_bubble_chi0_wk(g_wk, nw, save_memory): the reference driver's stage
 * order and branches: the static trapezoid (nw == 1), the stored product +
 * tau -> nu, or the fused product (save_memory). */
inline chi_wk_t imtime_bubble_chi0_wk(g_wk_cvt g_wk, int nw = 1, bool save_memory = false) {
  g_tr_t g_tr = fourier_wr_to_tr(fourier_wk_to_wr(g_wk));
  if (nw == 1) return chi_wk_from_chi_wr(chi0_w0r_from_grt_PH(g_tr));
  if (!save_memory) return chi_wk_from_chi_wr(chi_wr_from_chi_tr(chi0_tr_from_grt_PH(g_tr), nw));
  return chi_wk_from_chi_wr(chi0_wr_from_grt_PH(g_tr, nw));
}

// ---- f4: consumers of chi -- GW (gw.hpp) and Eliashberg (eliashberg.hpp) ----
using e_r_t = gf<mesh::cyclat, mesh::none, 2>;
using chi_k_t = gf<mesh::brzone, mesh::none, 4>;
using chi_r_t = gf<mesh::cyclat, mesh::none, 4>;
using e_r_cvt = const e_r_t &;
using chi_k_cvt = const chi_k_t &;
using chi_r_cvt = const chi_r_t &;
using chi_k_vt = const chi_k_t &;
using chi_r_vt = const chi_r_t &;
using chi_tr_vt = const chi_tr_t &;
using g_wk_vt = const g_wk_t &;
using g_tr_vt = const g_tr_t &;

namespace detail {
// one-frequency lattice transform of a k (to_r) or r function (make_gf_from_fourier)
template <class MI, class MO, int R>
gf<MO, mesh::none, R> lattice1(const gf<MI, mesh::none, R> &x, MO mo, bool to_r) {
  tprf_ctx *cx = context();
  gf<MO, mesh::none, R> out(mo, mesh::none{}, x.norb());
  check(to_r ? tprf_fourier_wk_to_wr(cx, 1, x.mesh0().dims.data(), x.target_size(), c(x.data()), c(out.data()))
             : tprf_fourier_wr_to_wk(cx, 1, x.mesh0().dims.data(), x.target_size(), c(x.data()), c(out.data())),
        cx, "lattice fourier");
  return out;
}
template <class M0, class M1, int R>
gf<M0, M1, R> plus_static(const gf<M0, M1, R> &dyn, const gf<M1, mesh::none, R> &st) {
  tprf_ctx *cx = context();
  gf<M0, M1, R> out(dyn.mesh0(), dyn.mesh1(), dyn.norb());
  hip_check(hipMemcpyAsync(out.data(), dyn.data(), sizeof(dcomplex) * (size_t)dyn.size(), hipMemcpyDeviceToDevice,
                           state().stream),
            "hipMemcpyAsync");
  check(tprf_add_dynamic_and_static(cx, dyn.mesh0().size(), dyn.mesh1().size() * dyn.target_size(), c(out.data()),
                                    c(st.data())),
        cx, "add_dynamic_and_static");
  return out;
}
}  // namespace detail

/* split_into_dynamic_wk_and_constant_k (lattice_utility.hpp; lattice_utility.cpp:57-84) */
inline std::tuple<chi_wk_t, chi_k_t> split_into_dynamic_wk_and_constant_k(chi_wk_cvt chi_wk) {
  tprf_ctx *cx = detail::context();
  const mesh::imfreq &wm = chi_wk.mesh0();
  chi_wk_t dyn(wm, chi_wk.mesh1(), chi_wk.norb());
  chi_k_t cst(chi_wk.mesh1(), mesh::none{}, chi_wk.norb());
  detail::check(tprf_split_into_dynamic_wk_and_constant_k(cx, wm.statistic, wm.beta, wm.n_max,
                                                          chi_wk.mesh1().size() * chi_wk.target_size(),
                                                          detail::c(chi_wk.data()), detail::c(dyn.data()),
                                                          detail::c(cst.data())),
                cx, "split_into_dynamic_wk_and_constant_k");
  return {dyn, cst};
}

/* rho_k_from_g_wk (gw.hpp:36; gw.cpp:36-52) */
inline e_k_t rho_k_from_g_wk(g_wk_cvt g_wk) {
  tprf_ctx *cx = detail::context();
  const mesh::imfreq &wm = g_wk.mesh0();
  e_k_t rho(g_wk.mesh1(), mesh::none{}, g_wk.norb());
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_density(cx, wm.beta, wm.n_max, g_wk.mesh1().size() * g_wk.target_size(), detail::c(g_wk.data()),
                             detail::c(rho.data())),
                cx, "rho_k_from_g_wk");
  detail::check_tail(cx, false);
  return rho;
}

/* gw_dynamic_sigma (gw.hpp:158; gw.cpp:66-102) */
inline g_tr_t gw_dynamic_sigma(chi_tr_cvt W_tr, g_tr_cvt g_tr) {
  if (W_tr.mesh0().n_tau != g_tr.mesh0().n_tau || W_tr.mesh0().beta != g_tr.mesh0().beta)
    throw error(TPRF_ERR_ARG, "gw_sigma_tr: tau meshes are not the same.");
  if (W_tr.mesh0().statistic != mesh::Boson || g_tr.mesh0().statistic != mesh::Fermion)
    throw error(TPRF_ERR_ARG, "gw_sigma_tr: statistics are incorrect.");
  if (W_tr.mesh1().dims != g_tr.mesh1().dims) throw error(TPRF_ERR_ARG, "gw_sigma_tr: real-space meshes are not the same.");
  tprf_ctx *cx = detail::context();
  g_tr_t out(g_tr.mesh0(), g_tr.mesh1(), g_tr.norb());
  detail::check(tprf_gw_dynamic_sigma(cx, g_tr.mesh0().size() * g_tr.mesh1().size(), g_tr.norb(),
                                      detail::c(W_tr.data()), detail::c(g_tr.data()), detail::c(out.data())),
                cx, "gw_dynamic_sigma");
  return out;
}

/* hartree_sigma / fock_sigma on the k mesh (gw.hpp:121,148; gw.cpp:124-147, 169-195) */
inline e_k_t hartree_sigma(chi_k_cvt v_k, g_wk_cvt g_wk) {
  tprf_ctx *cx = detail::context();
  e_k_t rho = rho_k_from_g_wk(g_wk), out(g_wk.mesh1(), mesh::none{}, g_wk.norb());
  detail::check(tprf_static_sigma_k(cx, 0, g_wk.mesh1().dims.data(), g_wk.norb(), detail::c(v_k.data()),
                                    detail::c(rho.data()), detail::c(out.data())),
                cx, "hartree_sigma");
  return out;
}
inline e_k_t fock_sigma(chi_k_cvt v_k, g_wk_cvt g_wk) {
  tprf_ctx *cx = detail::context();
  e_k_t rho = rho_k_from_g_wk(g_wk), out(g_wk.mesh1(), mesh::none{}, g_wk.norb());
  detail::check(tprf_static_sigma_k(cx, 1, g_wk.mesh1().dims.data(), g_wk.norb(), detail::c(v_k.data()),
                                    detail::c(rho.data()), detail::c(out.data())),
                cx, "fock_sigma");
  return out;
}
/* hartree_sigma / fock_sigma in r space (gw.hpp:122,149; gw.cpp:113-122, 149-167) */
inline e_r_t hartree_sigma(chi_k_cvt v_k, e_r_cvt rho_r) {
  tprf_ctx *cx = detail::context();
  e_r_t out(rho_r.mesh0(), mesh::none{}, rho_r.norb());
  detail::check(tprf_hartree_sigma_r(cx, rho_r.mesh0().size(), rho_r.norb(), detail::c(v_k.data()),
                                     detail::c(rho_r.data()), detail::c(out.data())),
                cx, "hartree_sigma");
  return out;
}
inline e_r_t fock_sigma(chi_r_cvt v_r, e_r_cvt rho_r) {
  tprf_ctx *cx = detail::context();
  e_r_t out(rho_r.mesh0(), mesh::none{}, rho_r.norb());
  detail::check(tprf_fock_sigma_r(cx, rho_r.mesh0().size(), rho_r.norb(), detail::c(v_r.data()),
                                  detail::c(rho_r.data()), detail::c(out.data())),
                cx, "fock_sigma");
  return out;
}

/* gw_sigma (gw.hpp:92,160; gw.cpp:198-244): static (the Fock term) or dynamic */
inline e_k_t gw_sigma(chi_k_cvt v_k, g_wk_cvt g_wk) { return fock_sigma(v_k, g_wk); }
inline g_wk_t gw_sigma(chi_wk_cvt W_wk, g_wk_cvt g_wk) {
  if (W_wk.mesh0().beta != g_wk.mesh0().beta) throw error(TPRF_ERR_ARG, "gw_sigma: inverse temperatures are not the same.");
  if (W_wk.mesh0().statistic != mesh::Boson || g_wk.mesh0().statistic != mesh::Fermion)
    throw error(TPRF_ERR_ARG, "gw_sigma: statistics are incorrect.");
  if (W_wk.mesh1().dims != g_wk.mesh1().dims) throw error(TPRF_ERR_ARG, "gw_sigma: k-space meshes are not the same.");
  auto [W_dyn_wk, W_const_k] = split_into_dynamic_wk_and_constant_k(W_wk);
  g_tr_t g_tr = fourier_wr_to_tr(fourier_wk_to_wr(g_wk));
  chi_tr_t W_dyn_tr = chi_tr_from_chi_wr(chi_wr_from_chi_wk(W_dyn_wk));
  g_tr_t sigma_dyn_tr = gw_dynamic_sigma(W_dyn_tr, g_tr);
  const mesh::cyclat rm = detail::to_r(g_wk.mesh1());
  chi_r_t W_const_r = detail::lattice1(W_const_k, rm, true);
  e_r_t rho_r = detail::lattice1(rho_k_from_g_wk(g_wk), rm, true);
  e_k_t sigma_fock_k = detail::lattice1(fock_sigma(W_const_r, rho_r), g_wk.mesh1(), false);
  g_wk_t sigma_wk = fourier_wr_to_wk(fourier_tr_to_wr(sigma_dyn_tr, g_wk.mesh0().n_max));
  return detail::plus_static(sigma_wk, sigma_fock_k);
}

/* eliashberg_g_delta_g_product (eliashberg.hpp:150; eliashberg.cpp:31-58) */
inline g_wk_t eliashberg_g_delta_g_product(g_wk_vt g_wk, g_wk_vt delta_wk) {
  tprf_ctx *cx = detail::context();
  g_wk_t F(delta_wk.mesh0(), delta_wk.mesh1(), delta_wk.norb());
  detail::check(tprf_eliashberg_g_delta_g_product(cx, g_wk.mesh0().n_max, delta_wk.mesh0().n_max,
                                                  delta_wk.mesh1().dims.data(), delta_wk.norb(), detail::c(g_wk.data()),
                                                  detail::c(delta_wk.data()), detail::c(F.data())),
                cx, "eliashberg_g_delta_g_product");
  return F;
}
/* eliashberg_product by summation (eliashberg.hpp:52; eliashberg.cpp:92-120) */
inline g_wk_t eliashberg_product(chi_wk_vt Gamma_pp, g_wk_vt g_wk, g_wk_vt delta_wk) {
  tprf_ctx *cx = detail::context();
  g_wk_t out(delta_wk.mesh0(), delta_wk.mesh1(), delta_wk.norb());
  detail::check(tprf_eliashberg_product(cx, delta_wk.mesh0().beta, Gamma_pp.mesh0().n_max, g_wk.mesh0().n_max,
                                        delta_wk.mesh0().n_max, delta_wk.mesh1().dims.data(), delta_wk.norb(),
                                        detail::c(Gamma_pp.data()), detail::c(g_wk.data()),
                                        detail::c(delta_wk.data()), detail::c(out.data())),
                cx, "eliashberg_product");
  return out;
}
/* dynamic_and_constant_to_tr (eliashberg.hpp:159; eliashberg.cpp:122-134) */
inline std::tuple<chi_tr_t, chi_r_t> dynamic_and_constant_to_tr(chi_wk_vt Gamma_pp_dyn_wk, chi_k_vt Gamma_pp_const_k) {
  chi_tr_t dyn_tr = chi_tr_from_chi_wr(chi_wr_from_chi_wk(Gamma_pp_dyn_wk));
  chi_r_t const_r = detail::lattice1(Gamma_pp_const_k, detail::to_r(Gamma_pp_const_k.mesh0()), true);
  return {dyn_tr, const_r};
}
/* eliashberg_dynamic_gamma_f_product / eliashberg_constant_gamma_f_product
 * (eliashberg.hpp:169,177; eliashberg.cpp:175-186, 214-240) */
inline g_tr_t eliashberg_dynamic_gamma_f_product(chi_tr_vt Gamma_pp_dyn_tr, g_tr_vt F_tr) {
  if (Gamma_pp_dyn_tr.mesh0().n_tau != F_tr.mesh0().n_tau)
    throw error(TPRF_ERR_ARG, "The size of the imaginary time mesh of Gamma must be the size of the mesh of Delta.");
  tprf_ctx *cx = detail::context();
  g_tr_t out(F_tr.mesh0(), F_tr.mesh1(), F_tr.norb());
  detail::check(tprf_eliashberg_gamma_f_product(cx, F_tr.mesh0().size() * F_tr.mesh1().size(), F_tr.norb(),
                                                detail::c(Gamma_pp_dyn_tr.data()), detail::c(F_tr.data()),
                                                detail::c(out.data())),
                cx, "eliashberg_dynamic_gamma_f_product");
  return out;
}
inline e_r_t eliashberg_constant_gamma_f_product(chi_r_vt Gamma_pp_const_r, g_tr_vt F_tr) {
  tprf_ctx *cx = detail::context();
  e_r_t out(F_tr.mesh1(), mesh::none{}, F_tr.norb());
  detail::check(tprf_eliashberg_gamma_f_product(cx, F_tr.mesh1().size(), F_tr.norb(), detail::c(Gamma_pp_const_r.data()),
                                                detail::c(F_tr.data()), detail::c(out.data())),
                cx, "eliashberg_constant_gamma_f_product");
  return out;
}
/* eliashberg_product_fft (eliashberg.hpp:146; eliashberg.cpp:242-268) */
inline g_wk_t eliashberg_product_fft(chi_tr_vt Gamma_pp_dyn_tr, chi_r_vt Gamma_pp_const_r, g_wk_vt g_wk,
                                     g_wk_vt delta_wk) {
  g_tr_t F_tr = fourier_wr_to_tr(fourier_wk_to_wr(eliashberg_g_delta_g_product(g_wk, delta_wk)));
  g_tr_t delta_tr = eliashberg_dynamic_gamma_f_product(Gamma_pp_dyn_tr, F_tr);
  e_r_t delta_r = eliashberg_constant_gamma_f_product(Gamma_pp_const_r, F_tr);
  return fourier_wr_to_wk(detail::plus_static(fourier_tr_to_wr(delta_tr), delta_r));
}
/* eliashberg_product_fft_constant (eliashberg.hpp:148; eliashberg.cpp:279-302) */
inline g_wk_t eliashberg_product_fft_constant(chi_r_vt Gamma_pp_const_r, g_wk_vt g_wk, g_wk_vt delta_wk) {
  g_wk_t F_wk = eliashberg_g_delta_g_product(g_wk, delta_wk);
  g_tr_t F_tr = fourier_wr_to_tr(fourier_wk_to_wr(F_wk));
  e_k_t delta_k = detail::lattice1(eliashberg_constant_gamma_f_product(Gamma_pp_const_r, F_tr), delta_wk.mesh1(), false);
  g_wk_t zero(F_wk.mesh0(), F_wk.mesh1(), F_wk.norb());
  detail::hip_check(hipMemsetAsync(zero.data(), 0, sizeof(dcomplex) * (size_t)zero.size(), detail::state().stream),
                    "hipMemsetAsync");
  return detail::plus_static(zero, delta_k);
}
/* construct_phi_wk (eliashberg.hpp:199; eliashberg.cpp:336-360); U in host memory */
inline chi_wk_t construct_phi_wk(chi_wk_vt chi, array4_view U) {
  tprf_ctx *cx = detail::context();
  const int n = chi.norb();
  if (U.n != n) throw error(TPRF_ERR_ARG, "construct_phi_wk: U does not match chi's target");
  const size_t un = (size_t)n * n * n * n;
  dcomplex *ud = nullptr;
  detail::hip_check(hipMalloc((void **)&ud, sizeof(dcomplex) * un), "hipMalloc");
  std::shared_ptr<dcomplex> hold(ud, [](dcomplex *q) { (void)hipFree(q); });
  detail::hip_check(hipMemcpyAsync(ud, U.data, sizeof(dcomplex) * un, hipMemcpyHostToDevice, detail::state().stream),
                    "hipMemcpyAsync");
  chi_wk_t out(chi.mesh0(), chi.mesh1(), n);
  detail::check(tprf_construct_phi_wk(cx, chi.mesh0().size() * chi.mesh1().size(), n, detail::c(chi.data()),
                                      detail::c(ud), detail::c(out.data())),
                cx, "construct_phi_wk");
  detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");  // U's copy in use
  return out;
}

// ---- batched forward chain (tprf_forward_batched) --------------------------
/* One lattice of a batch: e_k and U in, chi0_wk and chi_wk out (over 2 nb - 1
 * bosonic frequencies; with the symmetric flags chi0_wk holds nu >= 0 only),
 * plus the workspace the chain needs, all owned by the instance. */
struct chain_instance {
  tprf_chain_desc desc{};
  std::shared_ptr<dcomplex> e_k, U, g_wr, g_tr, chi0_wr, chi0_wk, chi_wk;
  long counts[5] = {0, 0, 0, 0, 0};
  long nk() const { return (long)desc.dims[0] * desc.dims[1] * desc.dims[2]; }

  chain_instance(const tprf_chain_desc &d, const std::vector<dcomplex> &e_k_host, const std::vector<dcomplex> &U_host)
      : desc(d) {
    if (tprf_chain_buffer_sizes(&desc, counts) != TPRF_OK) throw error(TPRF_ERR_ARG, "chain_instance: bad descriptor");
    const long n = d.norb;
    if ((long)e_k_host.size() != nk() * n * n || (long)U_host.size() != n * n * n * n)
      throw error(TPRF_ERR_ARG, "chain_instance: input sizes");
    e_k = alloc(e_k_host.size());
    U = alloc(U_host.size());
    g_wr = alloc(counts[0]);
    g_tr = alloc(counts[1]);
    chi0_wr = alloc(counts[2]);
    chi0_wk = alloc(counts[3]);
    chi_wk = alloc(counts[4]);
    upload(e_k.get(), e_k_host);
    upload(U.get(), U_host);
  }
  tprf_chain_ptrs ptrs() const {
    return tprf_chain_ptrs{detail::c(e_k.get()), detail::c(U.get()),       detail::c(g_wr.get()),
                           detail::c(g_tr.get()), detail::c(chi0_wr.get()), detail::c(chi0_wk.get()),
                           detail::c(chi_wk.get())};
  }
  std::vector<dcomplex> download(const std::shared_ptr<dcomplex> &p, long count) const {
    std::vector<dcomplex> h((size_t)count);
    detail::hip_check(hipMemcpyAsync(h.data(), p.get(), sizeof(dcomplex) * h.size(), hipMemcpyDeviceToHost,
                                     detail::state().stream),
                      "hipMemcpyAsync");
    detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");
    return h;
  }

 private:
  static std::shared_ptr<dcomplex> alloc(size_t count) {
    dcomplex *p = nullptr;
    detail::hip_check(hipMalloc((void **)&p, sizeof(dcomplex) * count), "hipMalloc");
    return std::shared_ptr<dcomplex>(p, [](dcomplex *q) { (void)hipFree(q); });
  }
  static void upload(dcomplex *d, const std::vector<dcomplex> &h) {
    detail::hip_check(
        hipMemcpyAsync(d, h.data(), sizeof(dcomplex) * h.size(), hipMemcpyHostToDevice, detail::state().stream),
        "hipMemcpyAsync");
    detail::hip_check(hipStreamSynchronize(detail::state().stream), "hipStreamSynchronize");
  }
};

/* The whole batch in one library call (nstreams lanes), then the tail-fit and
 * singularity status of the batch as for fourier_wr_to_tr. */
inline void forward_batched(const std::vector<chain_instance> &batch, int nstreams = 4) {
  tprf_ctx *cx = detail::context();
  std::vector<tprf_chain_desc> d;
  std::vector<tprf_chain_ptrs> p;
  for (const chain_instance &b : batch) {
    d.push_back(b.desc);
    p.push_back(b.ptrs());
  }
  detail::check(tprf_ctx_tail_status_reset(cx), cx, "status reset");
  detail::check(tprf_forward_batched(cx, (int)batch.size(), d.data(), p.data(), nstreams), cx, "forward_batched");
  detail::check_tail(cx, true);
}

}  // namespace tprf_hip

#endif
