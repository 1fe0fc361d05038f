// capi.cpp -- host side of libtprf_hip: context, plan and table caches, and the
// extern "C" entry points declared in include/tprf_hip.h.
//
// Entry points validate shapes the way the reference's preconditions do,
// build (once per shape) the FFT plans and Matsubara tables in device memory,
// and launch the kernels of kernels.hip on the context's stream.
#include "../../include/tprf_hip.h"

#include <dlfcn.h>

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "common.hpp"
#include "host_tables.hpp"

using tprf::c2;

namespace tprf {
hipError_t set_kernel_attributes();
hipError_t consumers_set_attributes();
hipError_t lindhard_set_attributes();
}

struct FftEntry {
  tprf::FftDev dev;
  c2 *tw = nullptr;
};

struct TailEntry {
  tprf::TailFitDev dev;
  int *idx = nullptr;
  c2 *pinv = nullptr;
  c2 *vander = nullptr;
  int *fitpos = nullptr;  // [mesh size]: position f of the mesh index in idx, or -1
};

struct MatsEntry {
  tprf::MatsDev dev;
  double *psi = nullptr;
  c2 *phase = nullptr;
};

struct BubbleEntry {
  tprf::BubbleDev dev;
  c2 *tw1 = nullptr, *twp = nullptr, *hmod = nullptr;
  int *rcanon = nullptr;
};

struct SymEntry {
  int *col = nullptr, *mcol = nullptr;
  long n = 0;
};

struct tprf_ctx {
  int device = 0;
  bool generic_kernels = false;
  hipStream_t stream = nullptr;
  std::string err;
  double *d_status = nullptr;
  std::map<int, FftEntry> fft;
  std::map<std::tuple<int, double, int>, TailEntry> fits;
  std::map<std::tuple<int, double, int>, MatsEntry> mats;
  tprf::Vinv8 vinv;
  std::map<std::tuple<double, int, int, int, int, int>, BubbleEntry> bubbles;
  std::map<std::tuple<double, int>, double *> static_w;  // nu = 0 weights c_j
  std::map<std::tuple<double, int, int>, std::pair<double *, tprf::StaticMatsDev>> static_mats;
  std::map<std::tuple<int, double, int, int>, double *> dtail;  // out-of-mesh pole model in tau (real)
  std::map<std::tuple<int, double, int, int>, c2 *> otab;       // per-tau output records of omega -> tau
  void *scratch = nullptr;  // this device's copy of the no-allocation entry point's file-scope array
  void *work = nullptr;                                   // grow-only scratch
  size_t work_bytes = 0;
  std::vector<void *> work_retired;  // outgrown scratch, kept until destroy: a captured graph may still use it
  void *ext_work = nullptr;          // caller-owned workspace (tprf_ctx_set_workspace)
  size_t ext_work_bytes = 0;
  size_t work_peak = 0;              // largest scratch asked for (tprf_ctx_workspace_size)
  int bubble_variant = -1;                                // -1: autotuned per shape
  std::map<std::tuple<int, int, int, int, long>, int> bubble_tuned;
  std::map<std::tuple<int, int, int, int, long>, int> bubble_herm_tuned;
  int w2t_variant = -1;
  std::map<std::tuple<int, int, int, int, long>, int> w2t_tuned;
  int lattice_variant = -1;
  std::map<std::tuple<long, long, int, int, int, int>, int> lattice_tuned;
  std::map<std::tuple<int, int>, int> dyson_lattice_tuned;
  std::map<std::tuple<int, int, int, int>, SymEntry> symcols;  // canonical columns per (dims, n)
  std::vector<hipStream_t> lanes;      // streams of tprf_forward_batched (default priority)
  std::vector<hipEvent_t> lane_events; // fork event + one join event per lane
  bool trace = false;                  // roctx range per entry point (TPRF_OPT_TRACE)
};

namespace {

// roctx from the rocprofiler SDK, loaded on first use (the library then has no
// link-time dependency on the profiler).
struct RoctxApi {
  int (*push)(const char *) = nullptr;
  int (*pop)() = nullptr;
};

const RoctxApi *roctx_api() {
  static RoctxApi api;
  static std::once_flag once;
  std::call_once(once, [] {
    void *h = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("librocprofiler-sdk-roctx.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    api.push = (int (*)(const char *))dlsym(h, "roctxRangePushA");
    api.pop = (int (*)())dlsym(h, "roctxRangePop");
    if (!api.push || !api.pop) api.push = nullptr;
  });
  return api.push ? &api : nullptr;
}

// A roctx range over the host-side span of one entry point (argument checks,
// table set-up and the enqueueing of its kernels) when the context traces.
struct TraceScope {
  const RoctxApi *api;
  TraceScope(const tprf_ctx *ctx, const char *name) : api(ctx && ctx->trace ? roctx_api() : nullptr) {
    if (api) api->push(name);
  }
  ~TraceScope() {
    if (api) api->pop();
  }
  TraceScope(const TraceScope &) = delete;
  TraceScope &operator=(const TraceScope &) = delete;
};

}  // namespace

#define TPRF_TRACE(ctx, name) TraceScope tprf_trace_scope_((ctx), (name))

namespace {

using tprf_host::kPiL;

int fail(tprf_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(tprf_ctx *ctx, hipError_t e, const char *where) {
  return fail(ctx, TPRF_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define TPRF_HIP_TRY(ctx, expr, where)          \
  do {                                          \
    hipError_t _e = (expr);                     \
    if (_e != hipSuccess) return hip_fail(ctx, _e, where); \
  } while (0)

bool factorize(int n, std::vector<int> &f) {
  f.clear();
  while (n % 8 == 0) { f.push_back(8); n /= 8; }
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0) { f.push_back(p); n /= p; }
  // larger primes: run-time radix passes
  for (int p = 17; (long)p * p <= n; p += 2)
    while (n % p == 0) { f.push_back(p); n /= p; }
  if (n > 1) f.push_back(n);
  return (int)f.size() <= tprf::kMaxPass;
}

template <typename T>
int upload(tprf_ctx *ctx, T **dst, const std::vector<T> &src) {
  TPRF_HIP_TRY(ctx, hipMalloc((void **)dst, sizeof(T) * src.size()), "hipMalloc");
  TPRF_HIP_TRY(ctx, hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ctx->stream),
               "hipMemcpyAsync");
  TPRF_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return TPRF_OK;
}

// FFT plan of length n (twiddles exp(-2 pi i k/n), long double angles)
int get_fft(tprf_ctx *ctx, int n, const tprf::FftDev **out) {
  auto it = ctx->fft.find(n);
  if (it != ctx->fft.end()) { *out = &it->second.dev; return TPRF_OK; }
  std::vector<int> f;
  if (n < 1 || !factorize(n, f))
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "FFT length " + std::to_string(n) + " has too many factors");
  std::vector<c2> tw(n);
  for (int k = 0; k < n; ++k) {
    long double a = 2.0L * kPiL * (long double)k / (long double)n;
    tw[k] = c2{(double)cosl(a), -(double)sinl(a)};
  }
  FftEntry e;
  int st = upload(ctx, &e.tw, tw);
  if (st) return st;
  e.dev.n = n;
  e.dev.npass = (int)f.size();
  for (int i = 0; i < tprf::kMaxPass; ++i) e.dev.radix[i] = i < (int)f.size() ? f[i] : 1;
  e.dev.tw = e.tw;
  auto res = ctx->fft.emplace(n, e);
  *out = &res.first->second.dev;
  return TPRF_OK;
}

// Tail-fit operator of one Matsubara mesh (host_tables.cpp: the TRIQS
// fit_tail window and order, least squares in long double), uploaded once.
int get_tail(tprf_ctx *ctx, int statistic, double beta, int n_max, const tprf::TailFitDev **out) {
  auto key = std::make_tuple(statistic, beta, n_max);
  auto it = ctx->fits.find(key);
  if (it != ctx->fits.end()) { *out = &it->second.dev; return TPRF_OK; }
  tprf_host::TailTables t;
  std::string msg;
  if (tprf_host::tail_fit_build(statistic, beta, n_max, t, msg)) return fail(ctx, TPRF_ERR_TAIL_ORDER, msg);
  std::vector<c2> pinv(t.pinv.size()), vd(t.vander.size());
  for (size_t i = 0; i < pinv.size(); ++i) pinv[i] = c2{t.pinv[i].real(), t.pinv[i].imag()};
  for (size_t i = 0; i < vd.size(); ++i) vd[i] = c2{t.vander[i].real(), t.vander[i].imag()};
  TailEntry e;
  int st;
  if ((st = upload(ctx, &e.idx, t.idx))) return st;
  if ((st = upload(ctx, &e.pinv, pinv))) return st;
  if ((st = upload(ctx, &e.vander, vd))) return st;
  if ((st = upload(ctx, &e.fitpos, t.fitpos))) return st;
  e.dev.n_fit = t.nfit;
  e.dev.ncoef = t.ncoef;
  e.dev.idx = e.idx;
  e.dev.pinv = e.pinv;
  e.dev.vander = e.vander;
  for (int p = 0; p < 4; ++p) e.dev.wpow[p] = t.wpow[p];
  auto res = ctx->fits.emplace(key, e);
  *out = &res.first->second.dev;
  return TPRF_OK;
}

using tprf_host::one_boson;
using tprf_host::one_fermion;

int get_mats(tprf_ctx *ctx, int statistic, double beta, int L, const tprf::MatsDev **out) {
  auto key = std::make_tuple(statistic, beta, L);
  auto it = ctx->mats.find(key);
  if (it != ctx->mats.end()) { *out = &it->second.dev; return TPRF_OK; }
  const double bf[3] = {0.0, 1.0, -1.0}, bb[3] = {-0.5, -1.0, 1.0};
  std::vector<double> psi(3 * (size_t)L);
  std::vector<c2> phase(L);
  for (int j = 0; j < L; ++j) {
    const double tau = (double)j * (beta / (double)L);
    for (int q = 0; q < 3; ++q)
      psi[q * L + j] = statistic == TPRF_FERMION ? one_fermion(bf[q], tau, beta) : one_boson(bb[q], tau, beta);
    long double a = kPiL * (long double)j / (long double)L;
    phase[j] = c2{(double)cosl(a), -(double)sinl(a)};
  }
  MatsEntry e;
  int st;
  if ((st = upload(ctx, &e.psi, psi))) return st;
  if ((st = upload(ctx, &e.phase, phase))) return st;
  e.dev.statistic = statistic;
  e.dev.L = L;
  e.dev.beta = beta;
  e.dev.psi = e.psi;
  e.dev.phase = e.phase;
  auto res = ctx->mats.emplace(key, e);
  *out = &res.first->second.dev;
  return TPRF_OK;
}

using tprf_host::vandermonde_inverse8;

using tprf_host::dtail_values_ld;

static std::vector<double> dtail_values(int statistic, double beta, int n_max, int L) {
  const std::vector<long double> d = dtail_values_ld(statistic, beta, n_max, L);
  return std::vector<double>(d.begin(), d.end());
}

int get_dtail(tprf_ctx *ctx, int statistic, double beta, int n_max, int L, const double **out) {
  auto key = std::make_tuple(statistic, beta, n_max, L);
  auto it = ctx->dtail.find(key);
  if (it != ctx->dtail.end()) { *out = it->second; return TPRF_OK; }
  const std::vector<double> d = dtail_values(statistic, beta, n_max, L);
  double *dd;
  int st = upload(ctx, &dd, d);
  if (st) return st;
  ctx->dtail.emplace(key, dd);
  *out = dd;
  return TPRF_OK;
}

// Per-tau output table of the pruned omega -> tau kernel, one 48-byte record
// per tau_j: [phase_j, (D_0, D_1), (D_2, 0)] with phase_j = exp(-i pi j / L)
// for fermions (1 for bosons) and D_q the table above; the kernel copies the
// records of one output residue into LDS in one pass.
int get_otab(tprf_ctx *ctx, int statistic, double beta, int n_max, int L, const c2 **out) {
  auto key = std::make_tuple(statistic, beta, n_max, L);
  auto it = ctx->otab.find(key);
  if (it != ctx->otab.end()) { *out = it->second; return TPRF_OK; }
  const std::vector<double> d = dtail_values(statistic, beta, n_max, L);
  std::vector<c2> o(3 * (size_t)L);
  for (int j = 0; j < L; ++j) {
    c2 ph = c2{1.0, 0.0};
    if (statistic == TPRF_FERMION) {
      const long double a = kPiL * (long double)j / (long double)L;
      ph = c2{(double)cosl(a), -(double)sinl(a)};
    }
    o[3 * (size_t)j] = ph;
    o[3 * (size_t)j + 1] = c2{d[j], d[(size_t)L + j]};
    o[3 * (size_t)j + 2] = c2{d[2 * (size_t)L + j], 0.0};
  }
  c2 *dd;
  int st = upload(ctx, &dd, o);
  if (st) return st;
  ctx->otab.emplace(key, dd);
  *out = dd;
  return TPRF_OK;
}

// Tables of the pruned bubble transform for (beta, L, nb, dims) and inner DFT
// length P (see bubble.hip).  Twiddles and the model transforms Phi_q(m) are
// evaluated in long double from exactly reduced angles.
int get_bubble(tprf_ctx *ctx, double beta, int L, int nb, const int dims[3], int P, const tprf::BubbleDev **out) {
  auto key = std::make_tuple(beta, L, nb, dims[0], dims[1], dims[2]);
  auto it = ctx->bubbles.find(key);
  if (it != ctx->bubbles.end() && it->second.dev.P == P) { *out = &it->second.dev; return TPRF_OK; }
  const int N1 = L / P, nnu = 2 * nb - 1;
  std::vector<long double> cs(L), sn(L);
  for (int j = 0; j < L; ++j) {
    long double a = 2.0L * kPiL * (long double)j / (long double)L;
    cs[j] = cosl(a);
    sn[j] = sinl(a);
  }
  auto e = [&](long jm) {  // exp(+2 pi i jm / L)
    long r = ((jm % L) + L) % L;
    return std::complex<long double>(cs[r], sn[r]);
  };
  std::vector<c2> tw1((size_t)N1 * P), twp(P), hmod(3 * (size_t)nnu);
  for (int j1 = 0; j1 < N1; ++j1)
    for (int k = 0; k < P; ++k) {
      const long m = k < P / 2 ? k : k - P;
      auto w = e((long)j1 * m);
      tw1[(size_t)j1 * P + k] = c2{(double)w.real(), (double)w.imag()};
    }
  for (int p = 0; p < P; ++p) {
    long double a = 2.0L * kPiL * (long double)p / (long double)P;
    twp[p] = c2{(double)cosl(a), (double)sinl(a)};
  }
  // H_q(m) = 1 / (i nu_m - b_q) - (beta / L) Phi_q(m), Phi_q(m) the transform of
  // the model's tau samples: the kernel's output is (beta/L) S(m) + sum_q a_q H_q(m)
  const double bb[3] = {-0.5, -1.0, 1.0};
  const long double fact = (long double)beta / (long double)L;
  for (int q = 0; q < 3; ++q) {
    std::vector<double> ps(L);
    for (int j = 0; j < L; ++j) ps[j] = one_boson(bb[q], (double)j * (beta / (double)L), beta);
    for (int im = 0; im < nnu; ++im) {
      const long m = im - (nb - 1);
      std::complex<long double> s = 0;
      for (int j = 0; j < L; ++j) s += (long double)ps[j] * e((long)j * m);
      const std::complex<long double> z =
          1.0L / std::complex<long double>(-(long double)bb[q], 2.0L * kPiL * (long double)m / (long double)beta);
      const std::complex<long double> h = z - fact * s;
      hmod[(size_t)q * nnu + im] = c2{(double)h.real(), (double)h.imag()};
    }
  }
  const long nr = (long)dims[0] * dims[1] * dims[2];
  std::vector<int> canon;
  for (long r = 0; r < nr; ++r) {
    long i2 = r % dims[2], i1 = (r / dims[2]) % dims[1], i0 = r / ((long)dims[1] * dims[2]);
    long mr = (((dims[0] - i0) % dims[0]) * dims[1] + (dims[1] - i1) % dims[1]) * dims[2] + (dims[2] - i2) % dims[2];
    if (r <= mr) canon.push_back((int)r);
  }
  if (it != ctx->bubbles.end()) {
    (void)hipFree(it->second.tw1); (void)hipFree(it->second.twp); (void)hipFree(it->second.hmod);
    (void)hipFree(it->second.rcanon);
    ctx->bubbles.erase(it);
  }
  BubbleEntry b;
  int st;
  if ((st = upload(ctx, &b.tw1, tw1))) return st;
  if ((st = upload(ctx, &b.twp, twp))) return st;
  if ((st = upload(ctx, &b.hmod, hmod))) return st;
  if ((st = upload(ctx, &b.rcanon, canon))) return st;
  b.dev.P = P;
  b.dev.ncanon = (int)canon.size();
  b.dev.tw1 = b.tw1;
  b.dev.twp = b.twp;
  b.dev.hmod = b.hmod;
  b.dev.rcanon = b.rcanon;
  auto res = ctx->bubbles.emplace(key, b);
  *out = &res.first->second.dev;
  return TPRF_OK;
}

// Scratch of at least `bytes`: the caller's bound workspace if there is one,
// else the context's own buffer.  That one only grows; an outgrown buffer is
// kept (not freed) until tprf_ctx_destroy, because kernels queued on it, or a
// graph captured with its address, may still use it.
int get_work(tprf_ctx *ctx, size_t bytes, void **out) {
  if (bytes > ctx->work_peak) ctx->work_peak = bytes;
  if (ctx->ext_work) {
    if (ctx->ext_work_bytes < bytes)
      return fail(ctx, TPRF_ERR_NOMEM,
                  "bound workspace too small: " + std::to_string(bytes) + " bytes needed, " +
                      std::to_string(ctx->ext_work_bytes) + " bound");
    *out = ctx->ext_work;
    return TPRF_OK;
  }
  if (ctx->work_bytes < bytes) {
    const size_t grow = bytes > 2 * ctx->work_bytes ? bytes : 2 * ctx->work_bytes;
    void *p = nullptr;
    if (hipMalloc(&p, grow) != hipSuccess) return fail(ctx, TPRF_ERR_NOMEM, "scratch allocation failed");
    if (ctx->work) ctx->work_retired.push_back(ctx->work);
    ctx->work = p;
    ctx->work_bytes = grow;
  }
  *out = ctx->work;
  return TPRF_OK;
}

std::vector<long double> static_weights_ld(const tprf_ctx *ctx, double beta, int L) {
  return tprf_host::static_weights_ld(ctx->vinv.v, beta, L);
}

int get_static_weights(tprf_ctx *ctx, double beta, int L, const double **out) {
  auto key = std::make_tuple(beta, L);
  auto it = ctx->static_w.find(key);
  if (it != ctx->static_w.end()) { *out = it->second; return TPRF_OK; }
  const std::vector<long double> c = static_weights_ld(ctx, beta, L);
  std::vector<double> cd(L + 1);
  for (int j = 0; j <= L; ++j) cd[j] = (double)c[j];
  double *d;
  int st = upload(ctx, &d, cd);
  if (st) return st;
  ctx->static_w.emplace(key, d);
  *out = d;
  return TPRF_OK;
}

// Tables of the nu = 0 bubble evaluated in Matsubara space (static_mats.hip;
// built in host_tables.cpp), uploaded once per (beta, n_iw, L).
int get_static_mats(tprf_ctx *ctx, double beta, int n_iw, int L, const double **tab, tprf::StaticMatsDev *sm) {
  auto key = std::make_tuple(beta, n_iw, L);
  auto it = ctx->static_mats.find(key);
  if (it == ctx->static_mats.end()) {
    std::vector<double> t;
    tprf_host::StaticMatsHost h;
    tprf_host::static_mats_build(ctx->vinv.v, beta, n_iw, L, t, h);
    tprf::StaticMatsDev d{};
    d.inv_beta = h.inv_beta;
    for (int j = 0; j < 9; ++j) { d.e[j] = h.e[j]; d.f[j] = h.f[j]; }
    for (int j = 0; j < 16; ++j) d.Q[j] = h.Q[j];
    double *dt;
    int st = upload(ctx, &dt, t);
    if (st) return st;
    it = ctx->static_mats.emplace(key, std::make_pair(dt, d)).first;
  }
  *tab = it->second.first;
  *sm = it->second.second;
  return TPRF_OK;
}

bool dims_ok(const int dims[3]) {
  if (!dims) return false;
  for (int d = 0; d < 3; ++d)
    if (dims[d] < 1 || dims[d] > 4096) return false;
  return true;
}

int check_launch(tprf_ctx *ctx, hipError_t e, const char *what) {
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return TPRF_OK;
}

// Kernel-variant autotuner.  Every variant of a launch writes the same output;
// on the first call per shape key all variants are timed with HIP events on
// the context's stream (two launches each, the second timed) and the fastest
// is cached.  While the stream is being captured into a graph nothing can be
// timed, so variant 0 is used for that call and nothing is cached.
template <class Key, class F>
int tuned_launch(tprf_ctx *ctx, std::map<Key, int> &cache, const Key &key, int nvar, int forced, F launch,
                 const char *what) {
  int variant = forced;
  if (variant < 0) {
    auto it = cache.find(key);
    if (it != cache.end()) {
      variant = it->second;
    } else {
      hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
      TPRF_HIP_TRY(ctx, hipStreamIsCapturing(ctx->stream, &cs), "hipStreamIsCapturing");
      if (cs != hipStreamCaptureStatusNone || nvar == 1) {
        variant = 0;
      } else {
        hipEvent_t e0, e1;
        TPRF_HIP_TRY(ctx, hipEventCreate(&e0), "hipEventCreate");
        TPRF_HIP_TRY(ctx, hipEventCreate(&e1), "hipEventCreate");
        float best = 1e30f;
        int st = TPRF_OK;
        for (int v = 0; v < nvar && st == TPRF_OK; ++v) {
          if ((st = check_launch(ctx, launch(v), what))) break;  // warm-up
          hipError_t e = hipEventRecord(e0, ctx->stream);
          if (e == hipSuccess) e = launch(v);
          if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
          if (e == hipSuccess) e = hipEventSynchronize(e1);
          float ms = 0;
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) { st = hip_fail(ctx, e, what); break; }
          if (ms < best) { best = ms; variant = v; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (st) return st;
        cache.emplace(key, variant);
      }
    }
  }
  return check_launch(ctx, launch(variant), what);
}

int lattice_fft(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const c2 *in, c2 *out, int sign,
                bool scaled) {
  if (!ctx) return TPRF_ERR_ARG;
  if (nw < 1 || ntarget < 1 || !dims_ok(dims) || !in || !out) return fail(ctx, TPRF_ERR_ARG, "lattice fft: bad shape");
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const double scale = scaled ? 1.0 / (double)nk : 1.0;
  int last = -1;
  for (int d = 0; d < 3; ++d)
    if (dims[d] > 1) last = d;
  if (last < 0) {  // a single k-point: copy (N_k = 1, so the scale is 1)
    TPRF_HIP_TRY(ctx, hipMemcpyAsync(out, in, sizeof(c2) * nw * ntarget, hipMemcpyDeviceToDevice, ctx->stream),
                 "hipMemcpyAsync");
    return TPRF_OK;
  }
  const int nvar = ctx->generic_kernels ? 0 : tprf::lattice3d_variants(dims);
  if (nvar > 0) {
    // single-pass kernel; with several variants the first call per shape times
    // them all, except in place (in == out), where a timing launch would
    // transform the data again: the cached or default variant runs then
    auto launch = [&](int v) { return tprf::launch_lattice3d(ctx->stream, sign, in, out, nw, dims, ntarget, scale, v); };
    const auto key = std::make_tuple(nw, ntarget, dims[0], dims[1], dims[2], sign);
    int forced = (nvar == 1 || ctx->lattice_variant >= nvar) ? 0 : ctx->lattice_variant;
    if (forced < 0 && in == out) {
      auto it = ctx->lattice_tuned.find(key);
      forced = it != ctx->lattice_tuned.end() ? it->second : tprf::lattice3d_default_variant(dims);
    }
    return tuned_launch(ctx, ctx->lattice_tuned, key, nvar, forced, launch, "lattice 3-d kernel");
  }
  const c2 *src = in;
  for (int d = 2; d >= 0; --d) {
    if (dims[d] == 1) continue;
    const tprf::FftDev *p;
    int st = get_fft(ctx, dims[d], &p);
    if (st) return st;
    long outer = nw, inner = ntarget;
    for (int e = 0; e < d; ++e) outer *= dims[e];
    for (int e = d + 1; e < 3; ++e) inner *= dims[e];
    // the last axis transformed is axis 0's lowest non-trivial index: apply the scale there
    int lowest = 0;
    while (dims[lowest] == 1) ++lowest;
    st = check_launch(ctx, tprf::launch_lattice_axis(ctx->stream, sign, src, out, outer, dims[d], inner, *p,
                                                     d == lowest ? scale : 1.0),
                      "lattice axis kernel");
    if (st) return st;
    src = out;
  }
  return TPRF_OK;
}

}  // namespace

extern "C" {

int tprf_ctx_create(tprf_ctx **out, int device) {
  if (!out) return TPRF_ERR_ARG;
  *out = nullptr;
  tprf_ctx *ctx = new tprf_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipMalloc((void **)&ctx->d_status, 4 * sizeof(double));
  if (e == hipSuccess) e = hipMemset(ctx->d_status, 0, 4 * sizeof(double));
  if (e == hipSuccess) e = tprf::set_kernel_attributes();
  if (e == hipSuccess) e = tprf::consumers_set_attributes();
  if (e == hipSuccess) e = tprf::lindhard_set_attributes();
  if (e == hipSuccess) e = tprf::noalloc_scratch_address(device, &ctx->scratch);
  if (e != hipSuccess) {
    delete ctx;
    return TPRF_ERR_HIP;
  }
  vandermonde_inverse8(ctx->vinv.v);
  *out = ctx;
  return TPRF_OK;
}

int tprf_ctx_destroy(tprf_ctx *ctx) {
  if (!ctx) return TPRF_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  for (auto &kv : ctx->fft) (void)hipFree(kv.second.tw);
  for (auto &kv : ctx->fits) { (void)hipFree(kv.second.idx); (void)hipFree(kv.second.pinv); (void)hipFree(kv.second.vander); (void)hipFree(kv.second.fitpos); }
  for (auto &kv : ctx->mats) { (void)hipFree(kv.second.psi); (void)hipFree(kv.second.phase); }
  for (auto &kv : ctx->static_w) (void)hipFree(kv.second);
  for (auto &kv : ctx->static_mats) (void)hipFree(kv.second.first);
  for (auto &kv : ctx->dtail) (void)hipFree(kv.second);
  for (auto &kv : ctx->otab) (void)hipFree(kv.second);
  for (auto &kv : ctx->symcols) { (void)hipFree(kv.second.col); (void)hipFree(kv.second.mcol); }
  if (ctx->work) (void)hipFree(ctx->work);
  for (void *p : ctx->work_retired) (void)hipFree(p);
  for (auto &kv : ctx->bubbles) {
    (void)hipFree(kv.second.tw1); (void)hipFree(kv.second.twp); (void)hipFree(kv.second.hmod);
    (void)hipFree(kv.second.rcanon);
  }
  for (hipEvent_t e : ctx->lane_events) (void)hipEventDestroy(e);
  for (hipStream_t l : ctx->lanes) (void)hipStreamDestroy(l);
  (void)hipFree(ctx->d_status);
  delete ctx;
  return TPRF_OK;
}

int tprf_ctx_set_option(tprf_ctx *ctx, int option, int value) {
  if (!ctx) return TPRF_ERR_ARG;
  if (option == TPRF_OPT_GENERIC_KERNELS) {
    ctx->generic_kernels = value != 0;
    return TPRF_OK;
  }
  if (option == TPRF_OPT_LATTICE_VARIANT) {
    if (value < -1 || value > 2) return fail(ctx, TPRF_ERR_ARG, "bad lattice variant");
    ctx->lattice_variant = value;
    return TPRF_OK;
  }
  if (option == TPRF_OPT_W2T_VARIANT) {
    if (value < -1 || value >= tprf::w2t_pruned_variants()) return fail(ctx, TPRF_ERR_ARG, "bad w2t variant");
    ctx->w2t_variant = value;
    return TPRF_OK;
  }
  if (option == TPRF_OPT_TRACE) {
    if (value && !roctx_api()) return fail(ctx, TPRF_ERR_UNSUPPORTED, "roctx library not found");
    ctx->trace = value != 0;
    return TPRF_OK;
  }
  if (option == TPRF_OPT_BUBBLE_VARIANT) {
    if (value < -1 || value >= tprf::bubble_pair_variants()) return fail(ctx, TPRF_ERR_ARG, "bad bubble variant");
    ctx->bubble_variant = value;
    return TPRF_OK;
  }
  return fail(ctx, TPRF_ERR_ARG, "unknown option");
}

int tprf_ctx_set_stream(tprf_ctx *ctx, void *stream) {
  if (!ctx) return TPRF_ERR_ARG;
  ctx->stream = (hipStream_t)stream;
  return TPRF_OK;
}

int tprf_ctx_set_workspace(tprf_ctx *ctx, void *work, size_t bytes) {
  if (!ctx) return TPRF_ERR_ARG;
  if (work && bytes == 0) return fail(ctx, TPRF_ERR_ARG, "set_workspace: zero bytes");
  ctx->ext_work = work;
  ctx->ext_work_bytes = work ? bytes : 0;
  return TPRF_OK;
}

int tprf_ctx_workspace_size(tprf_ctx *ctx, size_t *bytes, int reset) {
  if (!ctx || !bytes) return TPRF_ERR_ARG;
  *bytes = ctx->work_peak;
  if (reset) ctx->work_peak = 0;
  return TPRF_OK;
}

const char *tprf_ctx_last_error(const tprf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int tprf_ctx_tail_status(tprf_ctx *ctx, double *max_fit_error, double *max_abs_m0) {
  if (!ctx) return TPRF_ERR_ARG;
  double h[4];
  TPRF_HIP_TRY(ctx, hipMemcpyAsync(h, ctx->d_status, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), "hipMemcpyAsync");
  TPRF_HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  if (max_fit_error) *max_fit_error = h[0];
  if (max_abs_m0) *max_abs_m0 = h[1];
  if (h[2] != 0.0) return fail(ctx, TPRF_ERR_SINGULAR, "singular matrix in a Dyson or RPA inversion");
  return TPRF_OK;
}

int tprf_ctx_tail_status_reset(tprf_ctx *ctx) {
  if (!ctx) return TPRF_ERR_ARG;
  // a fill kernel, not hipMemsetAsync: callers capture this into HIP graphs
  return check_launch(ctx, tprf::launch_zero(ctx->stream, 2, reinterpret_cast<c2 *>(ctx->d_status)),
                      "status reset kernel");
}

int tprf_lattice_dyson_g_wk(tprf_ctx *ctx, double beta, int n_iw, double mu, long nk, int norb, const tprf_c128 *e_k,
                            int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_wk) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_wk");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || n_iw < 1 || nk < 1 || !e_k || !g_wk) return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_wk: bad argument");
  if (sigma_kind < 0 || sigma_kind > 2 || (sigma_kind && !sigma))
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_wk: bad sigma");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_wk: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_dyson(ctx->stream, beta, n_iw, mu, nk, norb, (const c2 *)e_k, sigma_kind,
                                         (const c2 *)sigma, (c2 *)g_wk, tprf::Status{ctx->d_status}),
                      "dyson kernel");
}

int tprf_lattice_dyson_g_zk(tprf_ctx *ctx, long nz, const tprf_c128 *z, double mu, long nk, int norb,
                            const tprf_c128 *e_k, int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_zk) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_zk");
  if (!ctx) return TPRF_ERR_ARG;
  if (nz < 1 || !z || nk < 1 || !e_k || !g_zk) return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_zk: bad argument");
  if (sigma_kind < 0 || sigma_kind > 2 || (sigma_kind && !sigma))
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_zk: bad sigma");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_zk: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_dyson(ctx->stream, 1.0, 1, mu, nk, norb, (const c2 *)e_k, sigma_kind,
                                         (const c2 *)sigma, (c2 *)g_zk, tprf::Status{ctx->d_status}, nullptr, 1,
                                         (const c2 *)z, nz),
                      "dyson kernel");
}

int tprf_lattice_dyson_g_z(tprf_ctx *ctx, long nz, const tprf_c128 *z, double mu, long nk, int norb,
                           const tprf_c128 *e_k, int sigma_kind, const tprf_c128 *sigma, tprf_c128 *g_z) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_z");
  if (!ctx) return TPRF_ERR_ARG;
  if (nz < 1 || !z || nk < 1 || !e_k || !g_z) return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_z: bad argument");
  if (sigma_kind < 0 || sigma_kind > 2 || (sigma_kind && !sigma))
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_z: bad sigma");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_z: norb must be 1..8");
  void *w;
  int st = get_work(ctx, sizeof(c2) * (size_t)nz * nk * norb * norb, &w);
  if (st) return st;
  st = check_launch(ctx,
                    tprf::launch_dyson(ctx->stream, 1.0, 1, mu, nk, norb, (const c2 *)e_k, sigma_kind,
                                       (const c2 *)sigma, (c2 *)w, tprf::Status{ctx->d_status}, nullptr, 1,
                                       (const c2 *)z, nz),
                    "dyson kernel");
  if (st) return st;
  return check_launch(ctx, tprf::launch_k_average(ctx->stream, nz, nk, norb, (const c2 *)w, (c2 *)g_z),
                      "k average kernel");
}

namespace {

// sym: 0 general, 1 Hermitian e_k (full output), 2 Hermitian e_k, columns of r <= -r only
int dyson_g_wr(tprf_ctx *ctx, double beta, int n_iw, double mu, const int dims[3], int norb, const tprf_c128 *e_k,
               tprf_c128 *g_wr, int sym) {
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || n_iw < 1 || !dims_ok(dims) || !e_k || !g_wr)
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_wr: bad argument");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_wr: norb must be 1..8");
  const long nk = (long)dims[0] * dims[1] * dims[2];
  if (!ctx->generic_kernels && dims[0] == 16 && dims[1] == 16 && dims[2] == 16 && norb <= 3) {
    auto launch = [&](int v) {
      return tprf::launch_dyson_lattice16r(ctx->stream, beta, n_iw, mu, norb, (const c2 *)e_k, (c2 *)g_wr,
                                           tprf::Status{ctx->d_status}, v, sym);
    };
    const int nvar = tprf::dyson_lattice16r_variants(norb);
    const int forced = ctx->lattice_variant >= 0 ? ctx->lattice_variant % nvar : -1;
    return tuned_launch(ctx, ctx->dyson_lattice_tuned, std::make_tuple(n_iw, norb + 8 * sym), nvar, forced, launch,
                        "fused dyson lattice kernel");
  }
  // other shapes: Dyson into g_wr, then the lattice transform in place
  int st = check_launch(ctx,
                        tprf::launch_dyson(ctx->stream, beta, n_iw, mu, nk, norb, (const c2 *)e_k, 0, nullptr,
                                           (c2 *)g_wr, tprf::Status{ctx->d_status}),
                        "dyson kernel");
  if (st) return st;
  return lattice_fft(ctx, 2L * n_iw, dims, (long)norb * norb, (const c2 *)g_wr, (c2 *)g_wr, -1, true);
}

}  // namespace

int tprf_lattice_dyson_g_wr(tprf_ctx *ctx, double beta, int n_iw, double mu, const int dims[3], int norb,
                            const tprf_c128 *e_k, tprf_c128 *g_wr) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_wr");
  return dyson_g_wr(ctx, beta, n_iw, mu, dims, norb, e_k, g_wr, 0);
}

int tprf_lattice_dyson_g_wr_herm(tprf_ctx *ctx, double beta, int n_iw, double mu, const int dims[3], int norb,
                                 const tprf_c128 *e_k, tprf_c128 *g_wr, int full) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_wr_herm");
  return dyson_g_wr(ctx, beta, n_iw, mu, dims, norb, e_k, g_wr, full ? 1 : 2);
}

int tprf_fourier_wk_to_wr(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *in,
                          tprf_c128 *out) {
  TPRF_TRACE(ctx, "tprf_fourier_wk_to_wr");
  return lattice_fft(ctx, nw, dims, ntarget, (const c2 *)in, (c2 *)out, -1, true);
}

int tprf_fourier_wr_to_wk(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *in,
                          tprf_c128 *out) {
  TPRF_TRACE(ctx, "tprf_fourier_wr_to_wk");
  return lattice_fft(ctx, nw, dims, ntarget, (const c2 *)in, (c2 *)out, +1, false);
}

namespace {

// Canonical columns (r, a, b) of G(., r)[a][b] under (r, a, b) <-> (-r, b, a):
// r < -r with every (a, b), and r = -r with a <= b; the partner of each
// (-1 when it is the column itself, and with full = false also for r < -r,
// whose partners are then not written).
int get_symcols(tprf_ctx *ctx, const int dims[3], int n, bool full, tprf::SymCols *out) {
  const auto key = std::make_tuple(dims[0], dims[1], dims[2], full ? n : -n);
  auto it = ctx->symcols.find(key);
  if (it == ctx->symcols.end()) {
    const long nk = (long)dims[0] * dims[1] * dims[2], n2 = (long)n * n;
    if (nk * n2 > 0x7fffffffL) return fail(ctx, TPRF_ERR_UNSUPPORTED, "symmetric transform: too many columns");
    std::vector<int> col, mcol;
    for (long r = 0; r < nk; ++r) {
      const long i2 = r % dims[2], i1 = (r / dims[2]) % dims[1], i0 = r / ((long)dims[1] * dims[2]);
      const long mr = (((dims[0] - i0) % dims[0]) * dims[1] + (dims[1] - i1) % dims[1]) * dims[2] +
                      (dims[2] - i2) % dims[2];
      if (mr < r) continue;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
          if (mr == r && b < a) continue;
          col.push_back((int)(r * n2 + a * n + b));
          mcol.push_back(((mr == r && a == b) || (mr != r && !full)) ? -1 : (int)(mr * n2 + b * n + a));
        }
    }
    SymEntry e;
    int st = upload(ctx, &e.col, col);
    if (st) return st;
    if ((st = upload(ctx, &e.mcol, mcol))) return st;
    e.n = (long)col.size();
    it = ctx->symcols.emplace(key, e).first;
  }
  *out = tprf::SymCols{it->second.col, it->second.mcol, it->second.n};
  return TPRF_OK;
}

// The input-pruned omega -> tau kernel takes L = Q M, M the smallest power of
// two >= 2 n_max, 64 <= M <= 1024, Q <= 64.
int w2t_pruned_M(int n_max) {
  int M = 64;
  while (M < 2 * n_max) M *= 2;
  return M;
}
bool w2t_pruned_fits(const tprf_ctx *ctx, int n_max, int L) {
  const int M = w2t_pruned_M(n_max);
  return !ctx->generic_kernels && L % M == 0 && tprf::w2t_pruned_supported(M) && L / M <= 64;
}

int wr_to_tr(tprf_ctx *ctx, int statistic, double beta, int n_max, int ntau, long ncol, const tprf_c128 *in,
             tprf_c128 *out, const tprf_c128 *moments, const tprf::SymCols *sym) {
  if (!ctx) return TPRF_ERR_ARG;
  if ((statistic != TPRF_FERMION && statistic != TPRF_BOSON) || !(beta > 0) || n_max < 1 || ncol < 1 || !in || !out)
    return fail(ctx, TPRF_ERR_ARG, "fourier_wr_to_tr: bad argument");
  if (ntau <= 0) ntau = 6 * n_max + 1;  // TRIQS make_adjoint_mesh default
  const int L = ntau - 1;
  if (L < 2 * n_max)
    return fail(ctx, TPRF_ERR_MESH_TOO_SHORT,
                "Inverse Fourier: The time mesh mush be at least twice as long as the freq mesh");
  const tprf::FftDev *p;
  const tprf::MatsDev *mt;
  tprf::TailFitDev fit{};
  int st;
  if (!moments) {
    const tprf::TailFitDev *f;
    if ((st = get_tail(ctx, statistic, beta, n_max, &f))) return st;
    fit = *f;
  }
  // input-pruned kernel: L = Q M with M a power of two holding the whole mesh
  const int M = w2t_pruned_M(n_max);
  if (w2t_pruned_fits(ctx, n_max, L)) {
    const tprf::FftDev *pL, *pM;
    if ((st = get_fft(ctx, L, &pL))) return st;
    if ((st = get_fft(ctx, M, &pM))) return st;
    if ((st = get_mats(ctx, statistic, beta, L, &mt))) return st;
    const double *dtail;
    if ((st = get_dtail(ctx, statistic, beta, n_max, L, &dtail))) return st;
    const c2 *otab;
    if ((st = get_otab(ctx, statistic, beta, n_max, L, &otab))) return st;
    tprf::W2tDev wd{M, L / M, pL->tw, pM->tw, dtail, otab};
    if (sym) {
      auto launch = [&](int v) {
        return tprf::launch_w2t_pruned_sym(ctx->stream, beta, n_max, ncol, (const c2 *)in, (c2 *)out, fit, *mt, wd,
                                           tprf::Status{ctx->d_status}, v, *sym);
      };
      const auto key = std::make_tuple(statistic, M, L, 2, ncol);
      return tuned_launch(ctx, ctx->w2t_tuned, key, tprf::w2t_pruned_variants(), ctx->w2t_variant, launch,
                          "w2t pruned kernel (symmetric)");
    }
    auto launch = [&](int v) {
      return tprf::launch_w2t_pruned(ctx->stream, statistic, beta, n_max, ncol, (const c2 *)in, (c2 *)out,
                                     (const c2 *)moments, fit, *mt, wd, tprf::Status{ctx->d_status}, v);
    };
    const auto key = std::make_tuple(statistic, M, L, moments ? 1 : 0, ncol);
    return tuned_launch(ctx, ctx->w2t_tuned, key, tprf::w2t_pruned_variants(), ctx->w2t_variant, launch,
                        "w2t pruned kernel");
  }
  if (sym) return fail(ctx, TPRF_ERR_UNSUPPORTED, "fourier_wr_to_tr_herm: no pruned kernel for this mesh");
  if (!tprf::mats_supported(L))
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "fourier_wr_to_tr: ntau too large for one workgroup");
  if ((st = get_fft(ctx, L, &p))) return st;
  if ((st = get_mats(ctx, statistic, beta, L, &mt))) return st;
  return check_launch(ctx,
                      tprf::launch_w2t(ctx->stream, statistic, beta, n_max, ncol, (const c2 *)in, (c2 *)out,
                                       (const c2 *)moments, fit, *mt, *p, tprf::Status{ctx->d_status}),
                      "w2t kernel");
}

}  // namespace

int tprf_fourier_wr_to_tr(tprf_ctx *ctx, int statistic, double beta, int n_max, int ntau, long ncol,
                          const tprf_c128 *in, tprf_c128 *out, const tprf_c128 *moments) {
  TPRF_TRACE(ctx, "tprf_fourier_wr_to_tr");
  return wr_to_tr(ctx, statistic, beta, n_max, ntau, ncol, in, out, moments, nullptr);
}

int tprf_fourier_wr_to_tr_herm(tprf_ctx *ctx, double beta, int n_max, int ntau, const int dims[3], int norb,
                               const tprf_c128 *in, tprf_c128 *out, int full) {
  TPRF_TRACE(ctx, "tprf_fourier_wr_to_tr_herm");
  if (!ctx) return TPRF_ERR_ARG;
  if (!dims_ok(dims) || norb < 1) return fail(ctx, TPRF_ERR_ARG, "fourier_wr_to_tr_herm: bad argument");
  tprf::SymCols sym;
  int st = get_symcols(ctx, dims, norb, full != 0, &sym);
  if (st) return st;
  const long ncol = (long)dims[0] * dims[1] * dims[2] * norb * norb;
  return wr_to_tr(ctx, TPRF_FERMION, beta, n_max, ntau, ncol, in, out, nullptr, &sym);
}

int tprf_fourier_tr_to_wr(tprf_ctx *ctx, int statistic, double beta, int ntau, int nw, long ncol,
                          const tprf_c128 *in, tprf_c128 *out) {
  TPRF_TRACE(ctx, "tprf_fourier_tr_to_wr");
  if (!ctx) return TPRF_ERR_ARG;
  if ((statistic != TPRF_FERMION && statistic != TPRF_BOSON) || !(beta > 0) || ntau < 10 || ncol < 1 || !in || !out)
    return fail(ctx, TPRF_ERR_ARG, "fourier_tr_to_wr: bad argument");
  const int L = ntau - 1;
  if (nw <= 0) nw = L / 6;  // TRIQS make_adjoint_mesh default
  if (nw < 1 || L < 2 * nw)
    return fail(ctx, TPRF_ERR_MESH_TOO_SHORT,
                "Fourier: The time mesh mush be at least twice as long as the number of positive frequencies");
  if (!tprf::mats_supported(L))
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "fourier_tr_to_wr: ntau too large for one workgroup");
  const tprf::FftDev *p;
  const tprf::MatsDev *mt;
  int st;
  if ((st = get_fft(ctx, L, &p))) return st;
  if ((st = get_mats(ctx, statistic, beta, L, &mt))) return st;
  return check_launch(ctx,
                      tprf::launch_t2w(ctx->stream, statistic, beta, nw, ncol, (const c2 *)in, (c2 *)out, *mt, *p,
                                       ctx->vinv),
                      "t2w kernel");
}

int tprf_chi0_tr_from_grt_PH(tprf_ctx *ctx, int ntau, const int dims[3], int norb, const tprf_c128 *g_tr,
                             tprf_c128 *chi0_tr) {
  TPRF_TRACE(ctx, "tprf_chi0_tr_from_grt_PH");
  if (!ctx) return TPRF_ERR_ARG;
  if (ntau < 2 || !dims_ok(dims) || norb < 1 || norb > 8 || !g_tr || !chi0_tr)
    return fail(ctx, TPRF_ERR_ARG, "chi0_tr_from_grt_PH: bad argument");
  return check_launch(ctx, tprf::launch_chi0_tr(ctx->stream, ntau, dims, norb, (const c2 *)g_tr, (c2 *)chi0_tr),
                      "chi0_tr kernel");
}

namespace {

int chi0_wr_from_grt(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, int nw, const tprf_c128 *g_tr,
                     tprf_c128 *chi0_wr, bool herm, bool nonneg = false) {
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || ntau < 10 || !dims_ok(dims) || norb < 1 || nw < 1 || !g_tr || !chi0_wr)
    return fail(ctx, TPRF_ERR_ARG, "chi0_wr_from_grt_PH: bad argument");
  if (norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "chi0_wr_from_grt_PH: norb must be 1..8");
  const int L = ntau - 1;
  if (L < 2 * nw)
    return fail(ctx, TPRF_ERR_MESH_TOO_SHORT,
                "Fourier: The time mesh mush be at least twice as long as the number of positive frequencies");
  int st;
  // pruned pair kernel: inner DFT length P >= 2 nw dividing L
  int P = 0;
  for (int cand : {64, 128})
    if (cand >= 2 * nw && L % cand == 0) { P = cand; break; }
  if (!ctx->generic_kernels && P && tprf::bubble_pair_supported(norb, P)) {
    const tprf::BubbleDev *bd;
    if ((st = get_bubble(ctx, beta, L, nw, dims, P, &bd))) return st;
    auto launch = [&](int v) {
      return tprf::launch_bubble_pair(ctx->stream, beta, L, nw, dims, norb, (const c2 *)g_tr, (c2 *)chi0_wr, *bd,
                                      ctx->vinv, v);
    };
    const auto key = std::make_tuple(norb, P, L, nw, (long)dims[0] * dims[1] * dims[2]);
    if (!herm)
      return tuned_launch(ctx, ctx->bubble_tuned, key, tprf::bubble_pair_variants(), ctx->bubble_variant, launch,
                          "bubble pair kernel");
    int list[8];
    const int cnt = tprf::bubble_pair_herm_variants(norb, P, list);
    auto launch_h = [&](int i) {
      return tprf::launch_bubble_pair(ctx->stream, beta, L, nw, dims, norb, (const c2 *)g_tr, (c2 *)chi0_wr, *bd,
                                      ctx->vinv, list[i], true, nonneg);
    };
    int forced = -1;  // a forced plain variant without a Hermitian form uses the first one
    if (ctx->bubble_variant >= 0) {
      forced = 0;
      for (int i = 0; i < cnt; ++i)
        if (list[i] == ctx->bubble_variant) forced = i;
    }
    return tuned_launch(ctx, ctx->bubble_herm_tuned, key, cnt, forced, launch_h, "bubble pair kernel (Hermitian)");
  }
  if (!tprf::mats_supported(L))
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "chi0_wr_from_grt_PH: ntau too large for one workgroup");
  const tprf::FftDev *p;
  const tprf::MatsDev *mt;
  if ((st = get_fft(ctx, L, &p))) return st;
  if ((st = get_mats(ctx, TPRF_BOSON, beta, L, &mt))) return st;
  return check_launch(ctx,
                      tprf::launch_chi0_wr(ctx->stream, beta, nw, dims, norb, (const c2 *)g_tr, (c2 *)chi0_wr, *mt,
                                           *p, ctx->vinv, herm),
                      "chi0_wr kernel");
}

}  // namespace

int tprf_chi0_wr_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, int nw,
                             const tprf_c128 *g_tr, tprf_c128 *chi0_wr) {
  TPRF_TRACE(ctx, "tprf_chi0_wr_from_grt_PH");
  return chi0_wr_from_grt(ctx, beta, ntau, dims, norb, nw, g_tr, chi0_wr, false);
}

int tprf_chi0_wr_from_grt_PH_herm(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, int nw,
                                  const tprf_c128 *g_tr, tprf_c128 *chi0_wr, int full) {
  TPRF_TRACE(ctx, "tprf_chi0_wr_from_grt_PH_herm");
  return chi0_wr_from_grt(ctx, beta, ntau, dims, norb, nw, g_tr, chi0_wr, true, full == 0);
}

int tprf_chi0_w0r_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb,
                              const tprf_c128 *g_tr, tprf_c128 *chi0_wr) {
  TPRF_TRACE(ctx, "tprf_chi0_w0r_from_grt_PH");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || ntau < 2 || !dims_ok(dims) || norb < 1 || norb > 8 || !g_tr || !chi0_wr)
    return fail(ctx, TPRF_ERR_ARG, "chi0_w0r_from_grt_PH: bad argument");
  return check_launch(ctx, tprf::launch_chi0_w0r(ctx->stream, beta, ntau, dims, norb, (const c2 *)g_tr, (c2 *)chi0_wr),
                      "chi0_w0r kernel");
}

int tprf_chi_w0r_from_chi_tr(tprf_ctx *ctx, double beta, int ntau, long nr, int norb, const tprf_c128 *chi_tr,
                             tprf_c128 *chi_wr) {
  TPRF_TRACE(ctx, "tprf_chi_w0r_from_chi_tr");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || ntau < 2 || nr < 1 || norb < 1 || norb > 8 || !chi_tr || !chi_wr)
    return fail(ctx, TPRF_ERR_ARG, "chi_w0r_from_chi_tr: bad argument");
  return check_launch(ctx, tprf::launch_chi_w0r(ctx->stream, beta, ntau, nr, norb, (const c2 *)chi_tr, (c2 *)chi_wr),
                      "chi_w0r kernel");
}

int tprf_solve_rpa_PH(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi0, const tprf_c128 *U, tprf_c128 *chi) {
  TPRF_TRACE(ctx, "tprf_solve_rpa_PH");
  if (!ctx) return TPRF_ERR_ARG;
  if (npts < 1 || !chi0 || !U || !chi) return fail(ctx, TPRF_ERR_ARG, "solve_rpa_PH: bad argument");
  if (norb < 1 || norb > tprf::kWideMaxOrb) return fail(ctx, TPRF_ERR_UNSUPPORTED, "solve_rpa_PH: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_rpa(ctx->stream, npts, norb, (const c2 *)chi0, (const c2 *)U, (c2 *)chi,
                                       tprf::Status{ctx->d_status}),
                      "rpa kernel");
}

int tprf_lindhard_chi00(tprf_ctx *ctx, double beta, double mu, long nz, const tprf_c128 *z, const int dims[3],
                        int norb, const tprf_c128 *e_k, tprf_c128 *chi) {
  TPRF_TRACE(ctx, "tprf_lindhard_chi00");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || nz < 1 || !z || !dims_ok(dims) || !e_k || !chi)
    return fail(ctx, TPRF_ERR_ARG, "lindhard_chi00: bad argument");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lindhard_chi00: norb must be 1..8");
  const long nk = (long)dims[0] * dims[1] * dims[2];
  const size_t ev_bytes = ((sizeof(double) * nk * norb + 255) / 256) * 256;
  void *w;
  int st = get_work(ctx, ev_bytes + sizeof(c2) * (size_t)nk * norb * norb * norb, &w);
  if (st) return st;
  return check_launch(ctx,
                      tprf::launch_lindhard(ctx->stream, beta, mu, nz, (const c2 *)z, dims, norb, (const c2 *)e_k,
                                            (double *)w, (c2 *)((char *)w + ev_bytes), (c2 *)chi),
                      "lindhard kernels");
}

int tprf_lindhard_chi00_noalloc(tprf_ctx *ctx, double beta, double mu, long nz, const tprf_c128 *z, const int dims[3],
                                int norb, const tprf_c128 *e_k, tprf_c128 *chi) {
  TPRF_TRACE(ctx, "tprf_lindhard_chi00_noalloc");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || nz < 1 || !z || !dims_ok(dims) || !e_k || !chi)
    return fail(ctx, TPRF_ERR_ARG, "lindhard_chi00_noalloc: bad argument");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lindhard_chi00_noalloc: norb must be 1..8");
  const long nk = (long)dims[0] * dims[1] * dims[2];
  // the same workspace layout as tprf_lindhard_chi00, in the file-scope array
  const size_t ev_bytes = ((sizeof(double) * nk * norb + 255) / 256) * 256;
  const size_t need = ev_bytes + sizeof(c2) * (size_t)nk * norb * norb * norb;
  if (need > tprf::kNoallocScratchBytes)
    return fail(ctx, TPRF_ERR_UNSUPPORTED,
                "lindhard_chi00_noalloc: " + std::to_string(need) + " bytes of workspace needed, the fixed maximum is " +
                    std::to_string(tprf::kNoallocScratchBytes) + " (nk * (8 n + 16 n^3) bytes, rounded)");
  if (!ctx->scratch) return fail(ctx, TPRF_ERR_HIP, "lindhard_chi00_noalloc: no device workspace address");
  char *w = (char *)ctx->scratch;
  return check_launch(ctx,
                      tprf::launch_lindhard(ctx->stream, beta, mu, nz, (const c2 *)z, dims, norb, (const c2 *)e_k,
                                            (double *)w, (c2 *)(w + ev_bytes), (c2 *)chi),
                      "lindhard kernels");
}

int tprf_solve_rpa_PH_herm(tprf_ctx *ctx, int nb, long nk, int norb, const tprf_c128 *chi0, const tprf_c128 *U,
                           tprf_c128 *chi) {
  TPRF_TRACE(ctx, "tprf_solve_rpa_PH_herm");
  if (!ctx) return TPRF_ERR_ARG;
  if (nb < 1 || nk < 1 || !chi0 || !U || !chi) return fail(ctx, TPRF_ERR_ARG, "solve_rpa_PH_herm: bad argument");
  if (norb < 1 || norb > tprf::kWideMaxOrb) return fail(ctx, TPRF_ERR_UNSUPPORTED, "solve_rpa_PH_herm: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_rpa_mirror(ctx->stream, nb, nk, norb, (const c2 *)chi0, (const c2 *)U, (c2 *)chi,
                                              tprf::Status{ctx->d_status}),
                      "rpa kernel (mirrored)");
}

int tprf_dynamical_screened_interaction_W(tprf_ctx *ctx, long nw, long nk, int norb, const tprf_c128 *PI,
                                          const tprf_c128 *V, int dynamic, tprf_c128 *W) {
  TPRF_TRACE(ctx, "tprf_dynamical_screened_interaction_W");
  if (!ctx) return TPRF_ERR_ARG;
  if (nw < 1 || nk < 1 || !PI || !V || !W || (dynamic != 0 && dynamic != 1))
    return fail(ctx, TPRF_ERR_ARG, "dynamical_screened_interaction_W: bad argument");
  if (norb < 1 || norb > tprf::kWideMaxOrb) return fail(ctx, TPRF_ERR_UNSUPPORTED, "dynamical_screened_interaction_W: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_screened_W(ctx->stream, nw * nk, nk, norb, dynamic, (const c2 *)PI, (const c2 *)V,
                                              (c2 *)W, tprf::Status{ctx->d_status}),
                      "screened interaction kernel");
}

// ---- consumers of chi: GW and Eliashberg ----------------------------------

namespace {
int consumer_n(tprf_ctx *ctx, int norb, const char *what) {
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, std::string(what) + ": norb must be 1..8");
  return TPRF_OK;
}
}  // namespace

int tprf_split_into_dynamic_wk_and_constant_k(tprf_ctx *ctx, int statistic, double beta, int n_max, long ncol,
                                              const tprf_c128 *chi_wk, tprf_c128 *chi_dyn_wk,
                                              tprf_c128 *chi_const_k) {
  TPRF_TRACE(ctx, "tprf_split_into_dynamic_wk_and_constant_k");
  if (!ctx) return TPRF_ERR_ARG;
  if ((statistic != TPRF_FERMION && statistic != TPRF_BOSON) || !(beta > 0) || n_max < 1 || ncol < 1 || !chi_wk ||
      !chi_dyn_wk || !chi_const_k)
    return fail(ctx, TPRF_ERR_ARG, "split_into_dynamic_wk_and_constant_k: bad argument");
  const tprf::TailFitDev *fit;
  int st;
  if ((st = get_tail(ctx, statistic, beta, n_max, &fit))) return st;
  const int nmesh = statistic == TPRF_FERMION ? 2 * n_max : 2 * n_max - 1;
  return check_launch(ctx,
                      tprf::launch_tail_split(ctx->stream, ncol, nmesh, *fit, (const c2 *)chi_wk, (c2 *)chi_dyn_wk,
                                              (c2 *)chi_const_k),
                      "tail split kernel");
}

int tprf_density(tprf_ctx *ctx, double beta, int n_max, long ncol, const tprf_c128 *g_w, tprf_c128 *rho) {
  TPRF_TRACE(ctx, "tprf_density");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || n_max < 1 || ncol < 1 || !g_w || !rho) return fail(ctx, TPRF_ERR_ARG, "density: bad argument");
  const tprf::TailFitDev *fit;
  int st;
  if ((st = get_tail(ctx, TPRF_FERMION, beta, n_max, &fit))) return st;
  return check_launch(ctx,
                      tprf::launch_density(ctx->stream, ncol, n_max, beta, *fit, (const c2 *)g_w, (c2 *)rho,
                                           ctx->d_status),
                      "density kernel");
}

int tprf_gw_dynamic_sigma(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *W_tr, const tprf_c128 *g_tr,
                          tprf_c128 *sigma_tr) {
  TPRF_TRACE(ctx, "tprf_gw_dynamic_sigma");
  if (!ctx) return TPRF_ERR_ARG;
  if (npts < 1 || !W_tr || !g_tr || !sigma_tr) return fail(ctx, TPRF_ERR_ARG, "gw_dynamic_sigma: bad argument");
  int st = consumer_n(ctx, norb, "gw_dynamic_sigma");
  if (st) return st;
  const long n4 = (long)norb * norb * norb * norb;
  return check_launch(ctx,
                      tprf::launch_vertex_contract(ctx->stream, 0, norb, npts, (const c2 *)W_tr, n4, (const c2 *)g_tr,
                                                   (c2 *)sigma_tr),
                      "gw dynamic sigma kernel");
}

int tprf_fock_sigma_r(tprf_ctx *ctx, long nr, int norb, const tprf_c128 *v_r, const tprf_c128 *rho_r,
                      tprf_c128 *sigma_r) {
  TPRF_TRACE(ctx, "tprf_fock_sigma_r");
  if (!ctx) return TPRF_ERR_ARG;
  if (nr < 1 || !v_r || !rho_r || !sigma_r) return fail(ctx, TPRF_ERR_ARG, "fock_sigma: bad argument");
  int st = consumer_n(ctx, norb, "fock_sigma");
  if (st) return st;
  const long n4 = (long)norb * norb * norb * norb;
  return check_launch(ctx,
                      tprf::launch_vertex_contract(ctx->stream, 1, norb, nr, (const c2 *)v_r, n4, (const c2 *)rho_r,
                                                   (c2 *)sigma_r),
                      "fock sigma kernel");
}

int tprf_hartree_sigma_r(tprf_ctx *ctx, long nr, int norb, const tprf_c128 *v_k, const tprf_c128 *rho_r,
                         tprf_c128 *sigma_r) {
  TPRF_TRACE(ctx, "tprf_hartree_sigma_r");
  if (!ctx) return TPRF_ERR_ARG;
  if (nr < 1 || !v_k || !rho_r || !sigma_r) return fail(ctx, TPRF_ERR_ARG, "hartree_sigma: bad argument");
  int st = consumer_n(ctx, norb, "hartree_sigma");
  if (st) return st;
  if ((st = check_launch(ctx, tprf::launch_zero(ctx->stream, nr * norb * norb, (c2 *)sigma_r), "zero kernel")))
    return st;
  // only r = 0 carries the Hartree term: V(q = 0) rho(r = 0)
  return check_launch(ctx,
                      tprf::launch_vertex_contract(ctx->stream, 2, norb, 1, (const c2 *)v_k, 0, (const c2 *)rho_r,
                                                   (c2 *)sigma_r),
                      "hartree sigma kernel");
}

int tprf_static_sigma_k(tprf_ctx *ctx, int fock, const int dims[3], int norb, const tprf_c128 *v_k,
                        const tprf_c128 *rho_k, tprf_c128 *sigma_k) {
  TPRF_TRACE(ctx, "tprf_static_sigma_k");
  if (!ctx) return TPRF_ERR_ARG;
  if (!dims_ok(dims) || (fock != 0 && fock != 1) || !v_k || !rho_k || !sigma_k)
    return fail(ctx, TPRF_ERR_ARG, "hartree/fock sigma (k): bad argument");
  int st = consumer_n(ctx, norb, "hartree/fock sigma (k)");
  if (st) return st;
  return check_launch(ctx,
                      tprf::launch_static_sigma_k(ctx->stream, fock != 0, norb, dims, (const c2 *)v_k,
                                                  (const c2 *)rho_k, (c2 *)sigma_k),
                      "static sigma kernel");
}

int tprf_eliashberg_gamma_f_product(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *gamma, const tprf_c128 *F,
                                    tprf_c128 *delta) {
  TPRF_TRACE(ctx, "tprf_eliashberg_gamma_f_product");
  if (!ctx) return TPRF_ERR_ARG;
  if (npts < 1 || !gamma || !F || !delta) return fail(ctx, TPRF_ERR_ARG, "eliashberg gamma-F product: bad argument");
  int st = consumer_n(ctx, norb, "eliashberg gamma-F product");
  if (st) return st;
  const long n4 = (long)norb * norb * norb * norb;
  return check_launch(ctx,
                      tprf::launch_vertex_contract(ctx->stream, 3, norb, npts, (const c2 *)gamma, n4, (const c2 *)F,
                                                   (c2 *)delta),
                      "eliashberg gamma-F kernel");
}

int tprf_eliashberg_g_delta_g_product(tprf_ctx *ctx, int n_max_g, int n_max_d, const int dims[3], int norb,
                                      const tprf_c128 *g_wk, const tprf_c128 *delta_wk, tprf_c128 *F_wk) {
  TPRF_TRACE(ctx, "tprf_eliashberg_g_delta_g_product");
  if (!ctx) return TPRF_ERR_ARG;
  if (n_max_g < 1 || n_max_d < 1 || !dims_ok(dims) || !g_wk || !delta_wk || !F_wk)
    return fail(ctx, TPRF_ERR_ARG, "eliashberg_g_delta_g_product: bad argument");
  if (n_max_d > n_max_g)
    return fail(ctx, TPRF_ERR_ARG,
                "The size of the Matsubara frequency mesh of the Green's function (" + std::to_string(2 * n_max_g) +
                    ") must be atleast the size of the mesh of Delta (" + std::to_string(2 * n_max_d) + ").");
  int st = consumer_n(ctx, norb, "eliashberg_g_delta_g_product");
  if (st) return st;
  return check_launch(ctx,
                      tprf::launch_g_delta_g(ctx->stream, norb, 2L * n_max_d, n_max_g - n_max_d, dims,
                                             (const c2 *)g_wk, (const c2 *)delta_wk, (c2 *)F_wk),
                      "eliashberg G delta G kernel");
}

int tprf_eliashberg_product(tprf_ctx *ctx, double beta, int n_max_gamma, int n_max_g, int n_max_d, const int dims[3],
                            int norb, const tprf_c128 *gamma_wk, const tprf_c128 *g_wk, const tprf_c128 *delta_wk,
                            tprf_c128 *delta_out) {
  TPRF_TRACE(ctx, "tprf_eliashberg_product");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || n_max_gamma < 1 || n_max_g < 1 || n_max_d < 1 || !dims_ok(dims) || !gamma_wk || !g_wk ||
      !delta_wk || !delta_out)
    return fail(ctx, TPRF_ERR_ARG, "eliashberg_product: bad argument");
  if (2 * (2 * n_max_d) > 2 * n_max_gamma - 1)
    return fail(ctx, TPRF_ERR_ARG,
                "The size of the Matsubara frequency mesh of Gamma (" + std::to_string(2 * n_max_gamma - 1) +
                    ") must be atleast TWICE the size of the mesh of Delta (" + std::to_string(2 * n_max_d) + ").");
  int st = consumer_n(ctx, norb, "eliashberg_product");
  if (st) return st;
  const long nk = (long)dims[0] * dims[1] * dims[2];
  void *w;
  if ((st = get_work(ctx, sizeof(c2) * (size_t)(2L * n_max_d * nk * norb * norb), &w))) return st;
  if ((st = tprf_eliashberg_g_delta_g_product(ctx, n_max_g, n_max_d, dims, norb, g_wk, delta_wk, (tprf_c128 *)w)))
    return st;
  return check_launch(ctx,
                      tprf::launch_eliashberg_sum(ctx->stream, norb, n_max_d, n_max_gamma, dims,
                                                  -1.0 / (2.0 * beta * (double)nk), (const c2 *)gamma_wk,
                                                  (const c2 *)w, (c2 *)delta_out),
                      "eliashberg summation kernel");
}

int tprf_construct_phi_wk(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi, const tprf_c128 *U,
                          tprf_c128 *phi) {
  TPRF_TRACE(ctx, "tprf_construct_phi_wk");
  if (!ctx) return TPRF_ERR_ARG;
  if (npts < 1 || !chi || !U || !phi) return fail(ctx, TPRF_ERR_ARG, "construct_phi_wk: bad argument");
  int st = consumer_n(ctx, norb, "construct_phi_wk");
  if (st) return st;
  return check_launch(ctx,
                      tprf::launch_construct_phi(ctx->stream, norb, npts, (const c2 *)chi, (const c2 *)U, (c2 *)phi),
                      "construct phi kernel");
}

int tprf_add_dynamic_and_static(tprf_ctx *ctx, long nw, long ncol, tprf_c128 *x_w, const tprf_c128 *s) {
  TPRF_TRACE(ctx, "tprf_add_dynamic_and_static");
  if (!ctx) return TPRF_ERR_ARG;
  if (nw < 1 || ncol < 1 || !x_w || !s) return fail(ctx, TPRF_ERR_ARG, "add_dynamic_and_static: bad argument");
  return check_launch(ctx, tprf::launch_add_static(ctx->stream, nw, ncol, (c2 *)x_w, (const c2 *)s),
                      "add static kernel");
}

// ---- differentiation -----------------------------------------------------

int tprf_lattice_dyson_g_wk_batched(tprf_ctx *ctx, double beta, int n_iw, const double *mu, long nbatch, long nk,
                                    int norb, const tprf_c128 *e_k, tprf_c128 *g_wk) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_wk_batched");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || n_iw < 1 || nk < 1 || nbatch < 1 || !mu || !e_k || !g_wk)
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_wk_batched: bad argument");
  if (norb < 1 || norb > 8) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_wk_batched: norb must be 1..8");
  return check_launch(ctx,
                      tprf::launch_dyson(ctx->stream, beta, n_iw, 0.0, nk * nbatch, norb, (const c2 *)e_k, 0, nullptr,
                                         (c2 *)g_wk, tprf::Status{ctx->d_status}, mu, nk),
                      "dyson kernel");
}

int tprf_lattice_dyson_g_wk_adjoint(tprf_ctx *ctx, int n_iw, long nk, int norb, long nbatch, const tprf_c128 *g_wk,
                                    const tprf_c128 *gbar_wk, tprf_c128 *ebar_k, double *mubar) {
  TPRF_TRACE(ctx, "tprf_lattice_dyson_g_wk_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  if (n_iw < 1 || nk < 1 || nbatch < 1 || !g_wk || !gbar_wk || !ebar_k || !mubar)
    return fail(ctx, TPRF_ERR_ARG, "lattice_dyson_g_wk_adjoint: bad argument");
  if (norb < 1 || norb > 4) return fail(ctx, TPRF_ERR_UNSUPPORTED, "lattice_dyson_g_wk_adjoint: norb must be 1..4");
  const int nw = 2 * n_iw;
  // frequency chunks: enough (k, chunk) threads to fill the chip
  const long ktot = nk * nbatch;
  int nchunk = (int)((1L << 20) / ktot);
  if (nchunk < 1) nchunk = 1;
  if (nchunk > nw) nchunk = nw;
  if (nchunk > 256) nchunk = 256;
  const long nbk = (nk + 255) / 256;
  const size_t part_bytes = sizeof(c2) * (size_t)nchunk * ktot * norb * norb;
  const size_t mu_bytes = sizeof(double) * (size_t)nchunk * nbk * nbatch;
  void *w;
  int st = get_work(ctx, part_bytes + mu_bytes, &w);
  if (st) return st;
  c2 *part_e = (c2 *)w;
  double *mu_part = (double *)((char *)w + part_bytes);
  return check_launch(ctx,
                      tprf::launch_dyson_adjoint(ctx->stream, nw, nk, norb, nbatch, (const c2 *)g_wk,
                                                 (const c2 *)gbar_wk, (c2 *)ebar_k, part_e, mu_part, mubar, nchunk),
                      "dyson adjoint kernel");
}

int tprf_fourier_wk_to_wr_adjoint(tprf_ctx *ctx, long nw, const int dims[3], long ntarget, const tprf_c128 *gbar_wr,
                                  tprf_c128 *gbar_wk) {
  TPRF_TRACE(ctx, "tprf_fourier_wk_to_wr_adjoint");
  return lattice_fft(ctx, nw, dims, ntarget, (const c2 *)gbar_wr, (c2 *)gbar_wk, +1, true);
}

int tprf_fourier_wr_to_tr_adjoint(tprf_ctx *ctx, int statistic, double beta, int n_max, int ntau, long ncol,
                                  const tprf_c128 *gbar_tr, tprf_c128 *gbar_wr, int fitted) {
  TPRF_TRACE(ctx, "tprf_fourier_wr_to_tr_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  if ((statistic != TPRF_FERMION && statistic != TPRF_BOSON) || !(beta > 0) || n_max < 1 || ncol < 1 || !gbar_tr ||
      !gbar_wr)
    return fail(ctx, TPRF_ERR_ARG, "fourier_wr_to_tr_adjoint: bad argument");
  if (ntau <= 0) ntau = 6 * n_max + 1;
  const int L = ntau - 1;
  if (L < 2 * n_max)
    return fail(ctx, TPRF_ERR_MESH_TOO_SHORT,
                "Inverse Fourier: The time mesh mush be at least twice as long as the freq mesh");
  int M = 64;
  while (M < 2 * n_max) M *= 2;
  if (L % M != 0 || !tprf::w2t_pruned_supported(M) || L / M > 64)
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "fourier_wr_to_tr_adjoint: ntau-1 must be a multiple of M in 64..1024");
  int st;
  tprf::TailFitDev fit{};
  if (fitted) {
    const tprf::TailFitDev *f;
    if ((st = get_tail(ctx, statistic, beta, n_max, &f))) return st;
    fit = *f;
  }
  const tprf::FftDev *pL, *pM;
  const tprf::MatsDev *mt;
  if ((st = get_fft(ctx, L, &pL))) return st;
  if ((st = get_fft(ctx, M, &pM))) return st;
  if ((st = get_mats(ctx, statistic, beta, L, &mt))) return st;
  const double *dtail;
  if ((st = get_dtail(ctx, statistic, beta, n_max, L, &dtail))) return st;
  tprf::W2tDev wd{M, L / M, pL->tw, pM->tw, dtail, nullptr};
  return check_launch(ctx,
                      tprf::launch_w2t_adjoint(ctx->stream, statistic, beta, n_max, ncol, (const c2 *)gbar_tr,
                                               (c2 *)gbar_wr, fitted ? 1 : 0, fit, *mt, wd),
                      "w2t adjoint kernel");
}

int tprf_chi0_static_from_grt_PH(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb, long nbatch,
                                 const tprf_c128 *g_tr, tprf_c128 *chi0_r) {
  TPRF_TRACE(ctx, "tprf_chi0_static_from_grt_PH");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || ntau < 18 || !dims_ok(dims) || nbatch < 1 || !g_tr || !chi0_r)
    return fail(ctx, TPRF_ERR_ARG, "chi0_static_from_grt_PH: bad argument");
  if (norb < 1 || norb > 4) return fail(ctx, TPRF_ERR_UNSUPPORTED, "chi0_static_from_grt_PH: norb must be 1..4");
  const int L = ntau - 1;
  const double *cw;
  int st;
  if ((st = get_static_weights(ctx, beta, L, &cw))) return st;
  const long nr = (long)dims[0] * dims[1] * dims[2] * nbatch;
  const long n4 = (long)norb * norb * norb * norb;
  // tau chunks: enough (r, chunk) threads to fill the chip
  int nchunk = (int)((1L << 20) / (nr * norb * norb));
  if (nchunk < 1) nchunk = 1;
  if (nchunk > 64) nchunk = 64;
  void *w;
  if ((st = get_work(ctx, sizeof(c2) * (size_t)nchunk * nr * n4, &w))) return st;
  return check_launch(ctx,
                      tprf::launch_bubble_static(ctx->stream, ntau, dims, norb, nbatch, (const c2 *)g_tr, cw, (c2 *)w,
                                                 nchunk, (c2 *)chi0_r),
                      "static bubble kernel");
}

int tprf_chi0_static_from_grt_PH_adjoint(tprf_ctx *ctx, double beta, int ntau, const int dims[3], int norb,
                                         long nbatch, const tprf_c128 *g_tr, const tprf_c128 *chibar_r,
                                         tprf_c128 *gbar_tr) {
  TPRF_TRACE(ctx, "tprf_chi0_static_from_grt_PH_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  if (!(beta > 0) || ntau < 18 || !dims_ok(dims) || nbatch < 1 || !g_tr || !chibar_r || !gbar_tr)
    return fail(ctx, TPRF_ERR_ARG, "chi0_static_from_grt_PH_adjoint: bad argument");
  if (norb < 1 || norb > 4)
    return fail(ctx, TPRF_ERR_UNSUPPORTED, "chi0_static_from_grt_PH_adjoint: norb must be 1..4");
  const double *cw;
  int st;
  if ((st = get_static_weights(ctx, beta, ntau - 1, &cw))) return st;
  return check_launch(ctx,
                      tprf::launch_bubble_static_adjoint(ctx->stream, ntau, dims, norb, nbatch, (const c2 *)g_tr,
                                                         (const c2 *)chibar_r, cw, (c2 *)gbar_tr),
                      "static bubble adjoint kernel");
}

namespace {
int static_gwr_check(tprf_ctx *ctx, double beta, int n_iw, int &ntau, const int dims[3], int norb, long nbatch,
                     const char *what) {
  if (!(beta > 0) || n_iw < 1 || !dims_ok(dims) || nbatch < 1) return fail(ctx, TPRF_ERR_ARG, std::string(what) + ": bad argument");
  if (norb < 1 || norb > 3) return fail(ctx, TPRF_ERR_UNSUPPORTED, std::string(what) + ": norb must be 1..3");
  if (ntau <= 0) ntau = 6 * n_iw + 1;
  const int L = ntau - 1;
  if (L < 2 * n_iw)
    return fail(ctx, TPRF_ERR_MESH_TOO_SHORT,
                "Inverse Fourier: The time mesh mush be at least twice as long as the freq mesh");
  if (L < 17) return fail(ctx, TPRF_ERR_UNSUPPORTED, std::string(what) + ": ntau must be at least 18");
  return TPRF_OK;
}
}  // namespace

int tprf_chi0_static_from_gwr_PH(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], int norb,
                                 long nbatch, const tprf_c128 *g_wr, tprf_c128 *chi0_r) {
  TPRF_TRACE(ctx, "tprf_chi0_static_from_gwr_PH");
  if (!ctx) return TPRF_ERR_ARG;
  int st = static_gwr_check(ctx, beta, n_iw, ntau, dims, norb, nbatch, "chi0_static_from_gwr_PH");
  if (st) return st;
  if (!g_wr || !chi0_r) return fail(ctx, TPRF_ERR_ARG, "chi0_static_from_gwr_PH: null pointer");
  const tprf::TailFitDev *fit;
  if ((st = get_tail(ctx, TPRF_FERMION, beta, n_iw, &fit))) return st;
  const double *tab;
  tprf::StaticMatsDev sm;
  if ((st = get_static_mats(ctx, beta, n_iw, ntau - 1, &tab, &sm))) return st;
  void *w;
  if ((st = get_work(ctx, tprf::static_mats_work_bytes(nbatch, dims, norb, n_iw, false), &w))) return st;
  return check_launch(ctx,
                      tprf::launch_static_mats(ctx->stream, nbatch, dims, norb, n_iw, (const c2 *)g_wr, tab, *fit, sm,
                                               (c2 *)w, (c2 *)chi0_r, ctx->d_status),
                      "static bubble (Matsubara) kernels");
}

int tprf_chi0_static_from_gwr_PH_adjoint(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], int norb,
                                         long nbatch, const tprf_c128 *g_wr, const tprf_c128 *chibar_r,
                                         tprf_c128 *gbar_wr) {
  TPRF_TRACE(ctx, "tprf_chi0_static_from_gwr_PH_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  int st = static_gwr_check(ctx, beta, n_iw, ntau, dims, norb, nbatch, "chi0_static_from_gwr_PH_adjoint");
  if (st) return st;
  if (!g_wr || !chibar_r || !gbar_wr) return fail(ctx, TPRF_ERR_ARG, "chi0_static_from_gwr_PH_adjoint: null pointer");
  const tprf::TailFitDev *fit;
  if ((st = get_tail(ctx, TPRF_FERMION, beta, n_iw, &fit))) return st;
  const double *tab;
  tprf::StaticMatsDev sm;
  if ((st = get_static_mats(ctx, beta, n_iw, ntau - 1, &tab, &sm))) return st;
  void *w;
  if ((st = get_work(ctx, tprf::static_mats_work_bytes(nbatch, dims, norb, n_iw, true), &w))) return st;
  return check_launch(ctx,
                      tprf::launch_static_mats_adjoint(ctx->stream, nbatch, dims, norb, n_iw, (const c2 *)g_wr, tab,
                                                       *fit, sm, (c2 *)w, (const c2 *)chibar_r, (c2 *)gbar_wr,
                                                       ctx->d_status),
                      "static bubble (Matsubara) adjoint kernels");
}

namespace {
// common checks and tables of the two fused static entry points; qr = q reduced into [0, dims)
int static_q_setup(tprf_ctx *ctx, double beta, int n_iw, int &ntau, const int dims[3], const int q[3], int norb,
                   long nbatch, const char *what, int qr[3], const tprf::TailFitDev **fit, const int **fitpos,
                   const double **tab, tprf::StaticMatsDev *sm) {
  int st = static_gwr_check(ctx, beta, n_iw, ntau, dims, norb, nbatch, what);
  if (st) return st;
  if (!q) return fail(ctx, TPRF_ERR_ARG, std::string(what) + ": null momentum");
  for (int d = 0; d < 3; ++d) qr[d] = (int)(((long)q[d] % dims[d] + dims[d]) % dims[d]);
  if ((st = get_tail(ctx, TPRF_FERMION, beta, n_iw, fit))) return st;
  *fitpos = ctx->fits.find(std::make_tuple((int)TPRF_FERMION, beta, n_iw))->second.fitpos;
  return get_static_mats(ctx, beta, n_iw, ntau - 1, tab, sm);
}
}  // namespace

int tprf_chi0_static_q_from_ek(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3], const int q[3],
                               int norb, long nbatch, const double *mu, const tprf_c128 *e_k, tprf_c128 *chi0_q,
                               tprf_c128 *proj) {
  TPRF_TRACE(ctx, "tprf_chi0_static_q_from_ek");
  if (!ctx) return TPRF_ERR_ARG;
  const char *what = "chi0_static_q_from_ek";
  int qr[3];
  const tprf::TailFitDev *fit;
  const int *fitpos;
  const double *tab;
  tprf::StaticMatsDev sm;
  int st = static_q_setup(ctx, beta, n_iw, ntau, dims, q, norb, nbatch, what, qr, &fit, &fitpos, &tab, &sm);
  if (st) return st;
  if (!mu || !e_k || !chi0_q) return fail(ctx, TPRF_ERR_ARG, std::string(what) + ": null pointer");
  void *w;
  if ((st = get_work(ctx, tprf::static_q_work_bytes(nbatch, dims, norb, n_iw, false), &w))) return st;
  return check_launch(ctx,
                      tprf::launch_static_q(ctx->stream, nbatch, dims, qr, norb, n_iw, beta, mu, (const c2 *)e_k, tab,
                                            fitpos, *fit, sm, (c2 *)w, (c2 *)proj, false, nullptr, (c2 *)chi0_q,
                                            nullptr, nullptr, ctx->d_status),
                      "static bubble (k space) kernels");
}

int tprf_chi0_static_q_from_ek_adjoint(tprf_ctx *ctx, double beta, int n_iw, int ntau, const int dims[3],
                                       const int q[3], int norb, long nbatch, const double *mu, const tprf_c128 *e_k,
                                       const tprf_c128 *chibar_q, const tprf_c128 *proj, tprf_c128 *ebar_k,
                                       double *mubar) {
  TPRF_TRACE(ctx, "tprf_chi0_static_q_from_ek_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  const char *what = "chi0_static_q_from_ek_adjoint";
  int qr[3];
  const tprf::TailFitDev *fit;
  const int *fitpos;
  const double *tab;
  tprf::StaticMatsDev sm;
  int st = static_q_setup(ctx, beta, n_iw, ntau, dims, q, norb, nbatch, what, qr, &fit, &fitpos, &tab, &sm);
  if (st) return st;
  if (!mu || !e_k || !chibar_q || !ebar_k || !mubar) return fail(ctx, TPRF_ERR_ARG, std::string(what) + ": null pointer");
  void *w;
  if ((st = get_work(ctx, tprf::static_q_work_bytes(nbatch, dims, norb, n_iw, true), &w))) return st;
  return check_launch(ctx,
                      tprf::launch_static_q(ctx->stream, nbatch, dims, qr, norb, n_iw, beta, mu, (const c2 *)e_k, tab,
                                            fitpos, *fit, sm, (c2 *)w, (c2 *)proj, proj != nullptr,
                                            (const c2 *)chibar_q, nullptr, (c2 *)ebar_k, mubar, ctx->d_status),
                      "static bubble (k space) adjoint kernels");
}

int tprf_fourier_r_to_q(tprf_ctx *ctx, const int dims[3], const int q[3], long ntarget, long nbatch,
                        const tprf_c128 *in_r, tprf_c128 *out_q) {
  TPRF_TRACE(ctx, "tprf_fourier_r_to_q");
  if (!ctx) return TPRF_ERR_ARG;
  if (!dims_ok(dims) || !q || ntarget < 1 || ntarget > 65535 || nbatch < 1 || nbatch > 65535 || !in_r || !out_q)
    return fail(ctx, TPRF_ERR_ARG, "fourier_r_to_q: bad argument");
  int qq[3];
  for (int d = 0; d < 3; ++d) qq[d] = ((q[d] % dims[d]) + dims[d]) % dims[d];
  return check_launch(ctx,
                      tprf::launch_q_project(ctx->stream, (const c2 *)in_r, (c2 *)out_q, dims, qq, (int)ntarget,
                                             nbatch),
                      "r->q kernel");
}

int tprf_fourier_r_to_q_adjoint(tprf_ctx *ctx, const int dims[3], const int q[3], long ntarget, long nbatch,
                                const tprf_c128 *gbar_q, tprf_c128 *gbar_r) {
  TPRF_TRACE(ctx, "tprf_fourier_r_to_q_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  if (!dims_ok(dims) || !q || ntarget < 1 || ntarget > 65535 || nbatch < 1 || !gbar_q || !gbar_r)
    return fail(ctx, TPRF_ERR_ARG, "fourier_r_to_q_adjoint: bad argument");
  int qq[3];
  for (int d = 0; d < 3; ++d) qq[d] = ((q[d] % dims[d]) + dims[d]) % dims[d];
  return check_launch(ctx,
                      tprf::launch_q_project_adjoint(ctx->stream, (const c2 *)gbar_q, (c2 *)gbar_r, dims, qq,
                                                     (int)ntarget, nbatch),
                      "r->q adjoint kernel");
}

int tprf_solve_rpa_PH_adjoint(tprf_ctx *ctx, long npts, int norb, const tprf_c128 *chi0, const tprf_c128 *U,
                              const tprf_c128 *chibar, tprf_c128 *chi, tprf_c128 *chi0bar, tprf_c128 *Ubar) {
  TPRF_TRACE(ctx, "tprf_solve_rpa_PH_adjoint");
  if (!ctx) return TPRF_ERR_ARG;
  if (npts < 1 || !chi0 || !U || !chi || (chibar && (!chi0bar || !Ubar)))
    return fail(ctx, TPRF_ERR_ARG, "solve_rpa_PH_adjoint: bad argument");
  if (norb < 1 || norb > 4) return fail(ctx, TPRF_ERR_UNSUPPORTED, "solve_rpa_PH_adjoint: norb must be 1..4");
  if (npts > 2147483647L) return fail(ctx, TPRF_ERR_UNSUPPORTED, "solve_rpa_PH_adjoint: too many points");
  return check_launch(ctx,
                      tprf::launch_rpa_point(ctx->stream, npts, norb, (const c2 *)chi0, (const c2 *)U,
                                             (const c2 *)chibar, (c2 *)chi, (c2 *)chi0bar, (c2 *)Ubar,
                                             tprf::Status{ctx->d_status}),
                      "rpa point kernel");
}


// ---- batched forward chain ------------------------------------------------

namespace {

long chain_nk(const tprf_chain_desc &d) { return (long)d.dims[0] * d.dims[1] * d.dims[2]; }
int chain_ntau(const tprf_chain_desc &d) { return d.ntau > 0 ? d.ntau : 6 * d.n_iw + 1; }

// One instance of the forward chain on ctx->stream: Dyson + k -> r, omega -> tau,
// fused bubble, r -> q, RPA (the stages of the reference driver,
// lattice_utils.py:241-266, then solve_rpa_PH), with the symmetric forms the
// descriptor's flags allow.
int chain_forward(tprf_ctx *ctx, const tprf_chain_desc &d, const tprf_chain_ptrs &p) {
  TPRF_TRACE(ctx, "tprf chain instance");
  const int n = d.norb, ntau = chain_ntau(d), L = ntau - 1;
  const long nk = chain_nk(d), n4 = (long)n * n * n * n;
  const bool herm = (d.flags & TPRF_CHAIN_HERMITIAN_EK) != 0;
  const bool herm_nu = herm && (d.flags & TPRF_CHAIN_HERMITIAN_U) != 0;
  const bool herm_w2t = herm && w2t_pruned_fits(ctx, d.n_iw, L);
  int st = dyson_g_wr(ctx, d.beta, d.n_iw, d.mu, d.dims, n, p.e_k, p.g_wr, herm ? (herm_w2t ? 2 : 1) : 0);
  if (st) return st;
  if (herm_w2t)
    st = tprf_fourier_wr_to_tr_herm(ctx, d.beta, d.n_iw, ntau, d.dims, n, p.g_wr, p.g_tr, 0);
  else
    st = wr_to_tr(ctx, TPRF_FERMION, d.beta, d.n_iw, ntau, nk * n * n, p.g_wr, p.g_tr, nullptr, nullptr);
  if (st) return st;
  st = chi0_wr_from_grt(ctx, d.beta, ntau, d.dims, n, d.nb, p.g_tr, p.chi0_wr, herm, herm_nu);
  if (st) return st;
  const long off = herm_nu ? (long)(d.nb - 1) * nk * n4 : 0;
  const long rows = herm_nu ? d.nb : 2L * d.nb - 1;
  st = lattice_fft(ctx, rows, d.dims, n4, (const c2 *)(p.chi0_wr + off), (c2 *)(p.chi0_wk + off), +1, false);
  if (st) return st;
  if (herm_nu) return tprf_solve_rpa_PH_herm(ctx, d.nb, nk, n, p.chi0_wk, p.U, p.chi_wk);
  return tprf_solve_rpa_PH(ctx, (2L * d.nb - 1) * nk, n, p.chi0_wk, p.U, p.chi_wk);
}

// Bytes the instance's chain moves (materialised tensors written and read
// once): the weight that balances the lanes.
double chain_cost(const tprf_chain_desc &d) {
  const double n = d.norb, nk = (double)chain_nk(d);
  return 16.0 * (5.0 * 2 * d.n_iw * nk * n * n + chain_ntau(d) * nk * n * n + 5.0 * (2 * d.nb - 1) * nk * n * n * n * n);
}

}  // namespace

int tprf_chain_buffer_sizes(const tprf_chain_desc *d, long counts[5]) {
  if (!d || !counts || d->norb < 1 || d->n_iw < 1 || d->nb < 1) return TPRF_ERR_ARG;
  const long nk = chain_nk(*d), n2 = (long)d->norb * d->norb, n4 = n2 * n2;
  counts[0] = 2L * d->n_iw * nk * n2;
  counts[1] = (long)chain_ntau(*d) * nk * n2;
  counts[2] = counts[3] = counts[4] = (2L * d->nb - 1) * nk * n4;
  return TPRF_OK;
}

int tprf_forward_batched(tprf_ctx *ctx, int n_inst, const tprf_chain_desc *desc, const tprf_chain_ptrs *ptrs,
                         int nstreams) {
  TPRF_TRACE(ctx, "tprf_forward_batched");
  if (!ctx) return TPRF_ERR_ARG;
  if (n_inst < 1 || !desc || !ptrs || nstreams < 1 || nstreams > 16)
    return fail(ctx, TPRF_ERR_ARG, "forward_batched: bad argument");
  for (int i = 0; i < n_inst; ++i) {
    const tprf_chain_desc &d = desc[i];
    const tprf_chain_ptrs &p = ptrs[i];
    if (!(d.beta > 0) || d.n_iw < 1 || d.nb < 1 || d.norb < 1 || !dims_ok(d.dims) || !p.e_k || !p.U || !p.g_wr ||
        !p.g_tr || !p.chi0_wr || !p.chi0_wk || !p.chi_wk)
      return fail(ctx, TPRF_ERR_ARG, "forward_batched: bad descriptor or pointer for instance " + std::to_string(i));
  }
  if (nstreams > n_inst) nstreams = n_inst;
  if (nstreams == 1) {
    for (int i = 0; i < n_inst; ++i) {
      const int st = chain_forward(ctx, desc[i], ptrs[i]);
      if (st) return st;
    }
    return TPRF_OK;
  }
  // lanes: nstreams - 1 library streams next to the caller's, forked from and
  // joined to it by events (so that the whole batch can be captured into one
  // graph from the caller's stream); largest instances first, each on the
  // least-loaded lane
  while ((int)ctx->lanes.size() < nstreams - 1) {
    hipStream_t l;
    TPRF_HIP_TRY(ctx, hipStreamCreateWithFlags(&l, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    ctx->lanes.push_back(l);
  }
  while ((int)ctx->lane_events.size() < nstreams) {
    hipEvent_t e;
    TPRF_HIP_TRY(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
    ctx->lane_events.push_back(e);
  }
  std::vector<int> order(n_inst);
  for (int i = 0; i < n_inst; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) { return chain_cost(desc[a]) > chain_cost(desc[b]); });
  std::vector<double> load(nstreams, 0.0);
  std::vector<std::vector<int>> lane_of(nstreams);
  for (int i : order) {
    const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    lane_of[j].push_back(i);
    load[j] += chain_cost(desc[i]);
  }
  hipStream_t main = ctx->stream;
  TPRF_HIP_TRY(ctx, hipEventRecord(ctx->lane_events[0], main), "hipEventRecord");
  int st = TPRF_OK;
  for (int j = 0; j < nstreams && st == TPRF_OK; ++j) {
    hipStream_t s = j == 0 ? main : ctx->lanes[j - 1];
    if (j > 0) {
      hipError_t e = hipStreamWaitEvent(s, ctx->lane_events[0], 0);
      if (e != hipSuccess) { st = hip_fail(ctx, e, "hipStreamWaitEvent"); break; }
    }
    ctx->stream = s;
    for (int i : lane_of[j]) {
      st = chain_forward(ctx, desc[i], ptrs[i]);
      if (st) break;
    }
    ctx->stream = main;
    if (st == TPRF_OK && j > 0) {
      hipError_t e = hipEventRecord(ctx->lane_events[j], s);
      if (e == hipSuccess) e = hipStreamWaitEvent(main, ctx->lane_events[j], 0);
      if (e != hipSuccess) st = hip_fail(ctx, e, "lane join");
    }
  }
  ctx->stream = main;
  return st;
}

}  // extern "C"
