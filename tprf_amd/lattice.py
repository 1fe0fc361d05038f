"""Drop-in for the hot-path functions of ``triqs_tprf.lattice``.

Same names, argument meanings and defaults as the reference's cpp2py module
(python/triqs_tprf/lattice_desc.py); each call runs on the GPU through the C
ABI of libtprf_hip.so and returns a new ``Gf`` whose data stays on the GPU.
Inputs may hold numpy or torch data; they are moved to the current device.

Errors the reference raises as C++ exceptions raise ``TprfError`` (a
RuntimeError) here; its stderr warnings are Python warnings.
"""
import warnings

import numpy as np
import torch

from . import _lib
from .gf import BOSON, FERMION, Gf, MeshBrZone, MeshCycLat, MeshImFreq, MeshImTime, MeshProduct, MeshReFreq

__all__ = [
    "lattice_dyson_g0_wk", "lattice_dyson_g_wk", "fourier_wk_to_wr", "fourier_wr_to_wk", "fourier_wr_to_tr",
    "fourier_tr_to_wr", "chi0_tr_from_grt_PH", "chi0_wr_from_grt_PH", "chi0_w0r_from_grt_PH",
    "chi_w0r_from_chi_tr", "chi_wr_from_chi_tr", "chi_tr_from_chi_wr", "chi_wk_from_chi_wr",
    "chi_wr_from_chi_wk", "solve_rpa_PH", "lindhard_chi00", "dynamical_screened_interaction_W",
    "split_into_dynamic_wk_and_constant_k", "add_dynamic_and_static", "rho_k_from_g_wk", "gw_dynamic_sigma",
    "hartree_sigma", "fock_sigma", "gw_sigma", "eliashberg_g_delta_g_product", "eliashberg_product",
    "eliashberg_product_fft", "eliashberg_product_fft_constant", "dynamic_and_constant_to_tr",
    "eliashberg_dynamic_gamma_f_product", "eliashberg_constant_gamma_f_product", "construct_phi_wk",
]


def _dev(t):
    if isinstance(t, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(t))
    if not t.is_cuda:
        t = t.to("cuda")
    if t.dtype != torch.complex128:
        t = t.to(torch.complex128)
    return t.contiguous()


def _stat_code(mesh):
    return _lib.FERMION if mesh.statistic == FERMION else _lib.BOSON


# ---------------------------------------------------------------------------
# a1 Dyson -- lattice_desc.py:27,57,119 -> c++/triqs_tprf/lattice/gf.cpp:41-132

def lattice_dyson_g0_wk(mu, e_k, mesh):
    """G0[w,k] = [(i w_n + mu) 1 - e_k]^-1 (gf.cpp:41-64)."""
    return _dyson(mu, e_k, mesh, None)


def lattice_dyson_g_wk(mu, e_k, sigma_w=None, sigma_wk=None):
    """G[w,k] = [(i w_n + mu) 1 - e_k - Sigma]^-1 with Sigma(w) or Sigma(w,k) (gf.cpp:91-132)."""
    if (sigma_w is None) == (sigma_wk is None):
        raise TypeError("give exactly one of sigma_w, sigma_wk")
    sigma = sigma_w if sigma_w is not None else sigma_wk
    wmesh = sigma.mesh[0] if isinstance(sigma.mesh, MeshProduct) else sigma.mesh
    return _dyson(mu, e_k, wmesh, sigma)


def _dyson(mu, e_k, wmesh, sigma):
    if not isinstance(wmesh, MeshImFreq) or wmesh.statistic != FERMION:
        raise ValueError("Dyson equation needs a fermionic Matsubara mesh")
    kmesh = e_k.mesh
    n = e_k.target_shape[0]
    ek = _dev(e_k.data)
    c = _lib.ctx(ek.device.index)
    out = torch.empty((wmesh.size, kmesh.size, n, n), dtype=torch.complex128, device=ek.device)
    kind, sp = _lib.SIGMA_NONE, None
    if sigma is not None:
        sd = _dev(sigma.data)
        kind = _lib.SIGMA_WK if isinstance(sigma.mesh, MeshProduct) else _lib.SIGMA_W
        if kind == _lib.SIGMA_WK and sigma.mesh[1].size != kmesh.size:
            raise ValueError("sigma_wk k-mesh does not match e_k")
        sp = _lib.ptr(sd)
    _lib.tail_status_reset(c)
    _lib.call("tprf_lattice_dyson_g_wk", c, wmesh.beta, wmesh.n_max, float(mu), kmesh.size, n, _lib.ptr(ek), kind,
              sp, _lib.ptr(out))
    _lib.tail_status(c)  # raises on a singular matrix
    return Gf(MeshProduct(wmesh, kmesh), data=out)


def _dyson_z(mu, e_k, fmesh, z, sigma, local):
    """G[z,k] = [(z + mu) 1 - e_k - Sigma]^-1 on the frequencies z of fmesh, or
    its k average (local)."""
    kmesh = e_k.mesh
    n = e_k.target_shape[0]
    ek = _dev(e_k.data)
    c = _lib.ctx(ek.device.index)
    zt = torch.from_numpy(np.ascontiguousarray(z, dtype=np.complex128)).to(ek.device)
    kind, sp = _lib.SIGMA_NONE, None
    if sigma is not None:
        sd = _dev(sigma.data)
        kind = _lib.SIGMA_WK if isinstance(sigma.mesh, MeshProduct) else _lib.SIGMA_W
        smesh = sigma.mesh[0] if kind == _lib.SIGMA_WK else sigma.mesh
        if smesh.size != fmesh.size or (kind == _lib.SIGMA_WK and sigma.mesh[1].size != kmesh.size):
            raise ValueError("sigma's mesh does not match the frequency mesh and e_k")
        sp = _lib.ptr(sd)
    shape = (fmesh.size, n, n) if local else (fmesh.size, kmesh.size, n, n)
    out = torch.empty(shape, dtype=torch.complex128, device=ek.device)
    _lib.tail_status_reset(c)
    _lib.call("tprf_lattice_dyson_g_z" if local else "tprf_lattice_dyson_g_zk", c, fmesh.size, _lib.ptr(zt),
              float(mu), kmesh.size, n, _lib.ptr(ek), kind, sp, _lib.ptr(out))
    _lib.tail_status(c)  # raises on a singular matrix
    return Gf(fmesh if local else MeshProduct(fmesh, kmesh), data=out)


def _real_freq_sigma(sigma, sigma_fk, sigma_f):
    given = [x for x in (sigma, sigma_fk, sigma_f) if x is not None]
    if len(given) != 1:
        raise TypeError("give exactly one self-energy")
    return given[0]


def lattice_dyson_g0_fk(mu, e_k, mesh, delta):
    """G0[f,k] = [(omega + i delta + mu) 1 - e_k]^-1 on a real-frequency mesh
    (gf.cpp:71-86)."""
    if not isinstance(mesh, MeshReFreq):
        raise ValueError("lattice_dyson_g0_fk needs a real-frequency mesh")
    return _dyson_z(mu, e_k, mesh, mesh.values() + 1j * float(delta), None, False)


def lattice_dyson_g_fk(mu, e_k, sigma=None, delta=0.0, sigma_fk=None, sigma_f=None):
    """G[f,k] = [(omega + i delta + mu) 1 - e_k - Sigma]^-1 with Sigma(f, k) or
    Sigma(f) (gf.cpp:92-123, 142-148)."""
    s = _real_freq_sigma(sigma, sigma_fk, sigma_f)
    fmesh = s.mesh[0] if isinstance(s.mesh, MeshProduct) else s.mesh
    if not isinstance(fmesh, MeshReFreq):
        raise ValueError("lattice_dyson_g_fk needs a real-frequency self-energy")
    return _dyson_z(mu, e_k, fmesh, fmesh.values() + 1j * float(delta), s, False)


def lattice_dyson_g_w(mu, e_k, sigma_w):
    """Local G(i w_n) = (1/N_k) sum_k [(i w_n + mu) 1 - e_k - Sigma(i w_n)]^-1
    (gf.cpp:153-196)."""
    wmesh = sigma_w.mesh
    if not isinstance(wmesh, MeshImFreq) or wmesh.statistic != FERMION:
        raise ValueError("lattice_dyson_g_w needs a fermionic Matsubara self-energy")
    return _dyson_z(mu, e_k, wmesh, wmesh.values(), sigma_w, True)


def lattice_dyson_g_f(mu, e_k, sigma_f, delta):
    """Local G(omega) = (1/N_k) sum_k [(omega + i delta + mu) 1 - e_k - Sigma(omega)]^-1
    (gf.cpp:153-204)."""
    fmesh = sigma_f.mesh
    if not isinstance(fmesh, MeshReFreq):
        raise ValueError("lattice_dyson_g_f needs a real-frequency self-energy")
    return _dyson_z(mu, e_k, fmesh, fmesh.values() + 1j * float(delta), sigma_f, True)


# ---------------------------------------------------------------------------
# a2 / a6 lattice FFTs -- gf.cpp:209-215, chi_imtime.cpp:289-297 -> fourier.hpp:165-231

def _lattice(g, to_r):
    wmesh, lmesh = g.mesh[0], g.mesh[1]
    if to_r and not isinstance(lmesh, MeshBrZone):
        raise ValueError("expected a (freq, k) Green's function")
    if not to_r and not isinstance(lmesh, MeshCycLat):
        raise ValueError("expected a (freq, r) Green's function")
    x = _dev(g.data)
    c = _lib.ctx(x.device.index)
    out = torch.empty_like(x)
    ntarget = int(np.prod(g.target_shape)) if g.target_shape else 1
    name = "tprf_fourier_wk_to_wr" if to_r else "tprf_fourier_wr_to_wk"
    _lib.call(name, c, wmesh.size, _lib.dims3(lmesh.dims), ntarget, _lib.ptr(x), _lib.ptr(out))
    new = MeshCycLat(lmesh.dims) if to_r else MeshBrZone(lmesh.dims)
    return Gf(MeshProduct(wmesh, new), data=out)


def fourier_wk_to_wr(g_wk):
    """G(w, r) = (1/N_k) sum_k e^{-i k.r} G(w, k) (FFTW_FORWARD then /N_k)."""
    return _lattice(g_wk, True)


def fourier_wr_to_wk(g_wr):
    """G(w, k) = sum_r e^{+i k.r} G(w, r) (FFTW_BACKWARD, no scale)."""
    return _lattice(g_wr, False)


def fourier_fk_to_fr(g_fk):
    """The k -> r transform on a real-frequency mesh (gf.cpp:229-232)."""
    return _lattice(g_fk, True)


def fourier_fr_to_fk(g_fr):
    """The r -> k transform on a real-frequency mesh (gf.cpp:234-237)."""
    return _lattice(g_fr, False)


def chi_wk_from_chi_wr(chi_wr):
    return _lattice(chi_wr, False)


def chi_wr_from_chi_wk(chi_wk):
    return _lattice(chi_wk, True)


# ---------------------------------------------------------------------------
# a3 Matsubara -> imaginary time -- gf.cpp:262-265, chi_imtime.cpp:284-287 ->
# fourier.hpp:98-130 -> fourier_matsubara.cpp:194-308

def _w_to_t(g, nt, moments=None):
    wmesh, rmesh = g.mesh[0], g.mesh[1]
    if not isinstance(wmesh, MeshImFreq):
        raise ValueError("expected a Matsubara mesh first")
    x = _dev(g.data)
    c = _lib.ctx(x.device.index)
    ntau = 6 * (wmesh.last_index + 1) + 1 if nt is None or nt < 0 else int(nt)
    ncol = x[0].numel()
    out = torch.empty((ntau,) + tuple(x.shape[1:]), dtype=torch.complex128, device=x.device)
    mp = None
    if moments is not None:
        m = _dev(moments).reshape(3, ncol).contiguous()
        mp = _lib.ptr(m)
    else:
        _lib.tail_status_reset(c)
    _lib.call("tprf_fourier_wr_to_tr", c, _stat_code(wmesh), wmesh.beta, wmesh.n_max, ntau, ncol, _lib.ptr(x),
              _lib.ptr(out), mp)
    if moments is None:
        err, m0 = _lib.tail_status(c)
        if err >= 1e-3:  # fourier_matsubara.cpp:227-230
            raise _lib.TprfError(-4, f"ERROR: High frequency moments have an error greater than 1e-3.\n  Error = {err}")
        if err > 1e-6:  # :231-234
            warnings.warn(f"WARNING: High frequency moments have an error greater than 1e-6.\n Error = {err}")
        if m0 > 1e-6:  # :238-239
            warnings.warn(f"WARNING: High frequency tail is not zero: {m0}")
    tmesh = MeshImTime(wmesh.beta, wmesh.statistic, ntau)
    return Gf(MeshProduct(tmesh, rmesh), data=out)


def fourier_wr_to_tr(g_wr, nt=-1, moments=None):
    """G(tau, r) from G(i w_n, r); ``moments`` [3][...] = (m1, m2, m3) skips the tail fit
    (the explicit-moment overload of the reference's _fourier_impl)."""
    return _w_to_t(g_wr, nt, moments)


def chi_tr_from_chi_wr(chi_wr, ntau=-1):
    return _w_to_t(chi_wr, ntau)


# ---------------------------------------------------------------------------
# a5 imaginary time -> Matsubara -- chi_imtime.cpp:279-282, gf.cpp:257-260 ->
# fourier.hpp:132-163 -> fourier_matsubara.cpp:42-184

def _t_to_w(g, nw):
    tmesh, rmesh = g.mesh[0], g.mesh[1]
    if not isinstance(tmesh, MeshImTime):
        raise ValueError("expected an imaginary-time mesh first")
    x = _dev(g.data)
    c = _lib.ctx(x.device.index)
    L = tmesh.n_tau - 1
    nw = L // 6 if nw is None or nw < 0 else int(nw)
    wmesh = MeshImFreq(tmesh.beta, tmesh.statistic, nw)
    ncol = x[0].numel()
    out = torch.empty((wmesh.size,) + tuple(x.shape[1:]), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_fourier_tr_to_wr", c, _stat_code(tmesh), tmesh.beta, tmesh.n_tau, nw, ncol, _lib.ptr(x),
              _lib.ptr(out))
    return Gf(MeshProduct(wmesh, rmesh), data=out)


def fourier_tr_to_wr(g_tr, nw=-1):
    return _t_to_w(g_tr, nw)


def chi_wr_from_chi_tr(chi_tr, nw):
    """chi(i nu, r) from chi(tau, r) (chi_imtime.hpp:87)."""
    return _t_to_w(chi_tr, nw)


# ---------------------------------------------------------------------------
# a4 bubble -- chi_imtime.cpp:88-228

def _g_tr_args(g_tr):
    tmesh, rmesh = g_tr.mesh[0], g_tr.mesh[1]
    if not isinstance(tmesh, MeshImTime) or not isinstance(rmesh, MeshCycLat):
        raise ValueError("expected G(tau, r)")
    x = _dev(g_tr.data)
    return tmesh, rmesh, x, g_tr.target_shape[0]


def chi0_tr_from_grt_PH(g_tr):
    """chi0_abcd(tau, r) = G_da(tau, r) G_bc(beta - tau, -r) (chi_imtime.cpp:88-136)."""
    tmesh, rmesh, x, n = _g_tr_args(g_tr)
    c = _lib.ctx(x.device.index)
    out = torch.empty((tmesh.n_tau, rmesh.size, n, n, n, n), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_chi0_tr_from_grt_PH", c, tmesh.n_tau, _lib.dims3(rmesh.dims), n, _lib.ptr(x), _lib.ptr(out))
    return Gf(MeshProduct(MeshImTime(tmesh.beta, BOSON, tmesh.n_tau), rmesh), data=out)


def chi0_wr_from_grt_PH(g_tr, nw=1):
    """Fused product + tau -> nu (chi_imtime.cpp:139-182); chi0(tau, r) is never stored."""
    tmesh, rmesh, x, n = _g_tr_args(g_tr)
    c = _lib.ctx(x.device.index)
    wmesh = MeshImFreq(tmesh.beta, BOSON, nw)
    out = torch.empty((wmesh.size, rmesh.size, n, n, n, n), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_chi0_wr_from_grt_PH", c, tmesh.beta, tmesh.n_tau, _lib.dims3(rmesh.dims), n, int(nw),
              _lib.ptr(x), _lib.ptr(out))
    return Gf(MeshProduct(wmesh, rmesh), data=out)


def chi0_w0r_from_grt_PH(g_tr):
    """Static bubble by trapezoid integration in tau (chi_imtime.cpp:185-247)."""
    tmesh, rmesh, x, n = _g_tr_args(g_tr)
    c = _lib.ctx(x.device.index)
    out = torch.empty((1, rmesh.size, n, n, n, n), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_chi0_w0r_from_grt_PH", c, tmesh.beta, tmesh.n_tau, _lib.dims3(rmesh.dims), n, _lib.ptr(x),
              _lib.ptr(out))
    return Gf(MeshProduct(MeshImFreq(tmesh.beta, BOSON, 1), rmesh), data=out)


def chi_w0r_from_chi_tr(chi_tr):
    """Static chi(nu=0, r) by trapezoid integration (chi_imtime.cpp:250-277)."""
    tmesh, rmesh = chi_tr.mesh[0], chi_tr.mesh[1]
    x = _dev(chi_tr.data)
    n = chi_tr.target_shape[0]
    c = _lib.ctx(x.device.index)
    out = torch.empty((1,) + tuple(x.shape[1:]), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_chi_w0r_from_chi_tr", c, tmesh.beta, tmesh.n_tau, rmesh.size, n, _lib.ptr(x), _lib.ptr(out))
    return Gf(MeshProduct(MeshImFreq(tmesh.beta, BOSON, 1), rmesh), data=out)


# ---------------------------------------------------------------------------
# a7 RPA -- lattice_desc.py:547 -> rpa.cpp:29-68

def solve_rpa_PH(chi0, U):
    """chi = (1 - chi0 U)^-1 chi0 per mesh point, on the matrices M[(a,b),(c,d)]
    of rpa.cpp:40,54-55 (memory-order grouping of the row-major tensors)."""
    x = _dev(chi0.data)
    n = chi0.target_shape[0]
    u = _dev(U if not isinstance(U, Gf) else U.data)
    if tuple(u.shape) != (n, n, n, n):
        raise ValueError("U must be a rank-4 tensor matching chi0's target")
    c = _lib.ctx(x.device.index)
    out = torch.empty_like(x)
    npts = x.numel() // n ** 4
    _lib.tail_status_reset(c)
    _lib.call("tprf_solve_rpa_PH", c, npts, n, _lib.ptr(x), _lib.ptr(u), _lib.ptr(out))
    _lib.tail_status(c)
    return Gf(chi0.mesh, data=out)


def dynamical_screened_interaction_W(PI_wk, V_k):
    """Screened interaction W_abcd(w, k) from the bubble PI_wk and the bare
    interaction V (dynamical_screened_interaction.hpp:44,63,81,100; binding
    lattice_desc.py:892-970): M_W = M_V (1 - M_PI M_V)^-1 per (w, k) point, with
    M_X[(a,b),(c,d)] = X_abcd.  V_k is static (a Gf on the k mesh, or an
    [nk, n, n, n, n] array) or dynamic (a Gf on PI_wk's (w, k) mesh).  The
    frequency mesh may be Matsubara or real."""
    pi = _dev(PI_wk.data)
    if len(PI_wk.target_shape) != 4 or not isinstance(PI_wk.mesh, MeshProduct):
        raise ValueError("PI_wk must be a rank-4 Gf on a (frequency, k) mesh")
    n = PI_wk.target_shape[0]
    wmesh, kmesh = PI_wk.mesh[0], PI_wk.mesh[1]
    vdata = V_k.data if isinstance(V_k, Gf) else V_k
    dynamic = isinstance(V_k, Gf) and isinstance(V_k.mesh, MeshProduct)
    if isinstance(V_k, Gf):
        vk = V_k.mesh[1] if dynamic else V_k.mesh
        if getattr(vk, "dims", None) != getattr(kmesh, "dims", None):
            raise RuntimeError("dynamical_screened_interaction_W: k-space meshes are not the same")
    v = _dev(vdata)
    nw, nk = pi.shape[0], pi.shape[1]
    want = (nw, nk, n, n, n, n) if dynamic else (nk, n, n, n, n)
    if tuple(v.shape) != want:
        raise RuntimeError("dynamical_screened_interaction_W: k-space meshes are not the same")
    c = _lib.ctx(pi.device.index)
    out = torch.empty_like(pi)
    _lib.tail_status_reset(c)
    _lib.call("tprf_dynamical_screened_interaction_W", c, nw, nk, n, _lib.ptr(pi), _lib.ptr(v), 1 if dynamic else 0,
              _lib.ptr(out))
    _lib.tail_status(c)
    return Gf(PI_wk.mesh, data=out)


def lindhard_chi00(e_k, mesh, *args, mu=None, beta=None, delta=0.0):
    """Analytic particle-hole bubble chi00_abcd(w, q) (lindhard_chi00.cpp:34-108).

    ``lindhard_chi00(e_k, mesh, mu)`` on a bosonic Matsubara mesh
    (lattice_desc.py:411; the frequencies are i nu_m), or
    ``lindhard_chi00(e_k, mesh, beta, mu, delta)`` on a real-frequency mesh
    (lattice_desc.py:495; omega + i delta).  Returns a Gf on (mesh, k mesh) with
    an n^4 target.  O(N_k^2) work."""
    if isinstance(mesh, MeshImFreq):
        if args:
            (mu,) = args
        if mesh.statistic != BOSON:
            raise _lib.TprfError(-1, "lindhard_chi00: statistic is incorrect.")
        beta = mesh.beta
        z = mesh.values()
    elif isinstance(mesh, MeshReFreq):
        if args:
            beta, mu, *rest = args
            if rest:
                (delta,) = rest
        z = mesh.values() + 1j * float(delta)
    else:
        raise TypeError("lindhard_chi00: mesh must be MeshImFreq or MeshReFreq")
    if mu is None or beta is None:
        raise TypeError("lindhard_chi00: mu (and beta for a real-frequency mesh) required")
    kmesh = e_k.mesh
    n = e_k.target_shape[0]
    e = _dev(e_k.data)
    zd = torch.from_numpy(np.ascontiguousarray(z, dtype=np.complex128)).to(e.device)
    c = _lib.ctx(e.device.index)
    out = torch.empty((len(z), kmesh.size, n, n, n, n), dtype=torch.complex128, device=e.device)
    _lib.call("tprf_lindhard_chi00", c, float(beta), float(mu), len(z), _lib.ptr(zd), _lib.dims3(kmesh.dims), n,
              _lib.ptr(e), _lib.ptr(out))
    return Gf(MeshProduct(mesh, MeshBrZone(kmesh.dims)), data=out)


# ---------------------------------------------------------------------------
# f4 consumers of chi: GW -- gw.hpp:36-281 / gw.cpp, lattice_utility.cpp:57-101

def _kmesh_of(g):
    return g.mesh[1] if isinstance(g.mesh, MeshProduct) else g.mesh


def split_into_dynamic_wk_and_constant_k(chi_wk):
    """(chi_dyn_wk, chi_const_k): the zeroth tail-fit coefficient of every
    (k, target) column is the constant part (lattice_utility.cpp:57-84)."""
    wmesh, kmesh = chi_wk.mesh[0], chi_wk.mesh[1]
    x = _dev(chi_wk.data)
    c = _lib.ctx(x.device.index)
    ncol = x[0].numel()
    dyn = torch.empty_like(x)
    cst = torch.empty(tuple(x.shape[1:]), dtype=torch.complex128, device=x.device)
    _lib.call("tprf_split_into_dynamic_wk_and_constant_k", c, _stat_code(wmesh), wmesh.beta, wmesh.n_max, ncol,
              _lib.ptr(x), _lib.ptr(dyn), _lib.ptr(cst))
    return Gf(chi_wk.mesh, data=dyn), Gf(kmesh, data=cst)


def add_dynamic_and_static(g_dyn_wk, g_stat_k):
    """g_dyn_wk + g_stat_k at every frequency (lattice_utility.cpp:86-101)."""
    x = _dev(g_dyn_wk.data).clone()
    st = _dev(g_stat_k.data)
    c = _lib.ctx(x.device.index)
    _lib.call("tprf_add_dynamic_and_static", c, x.shape[0], x[0].numel(), _lib.ptr(x), _lib.ptr(st))
    return Gf(g_dyn_wk.mesh, data=x)


def rho_k_from_g_wk(g_wk):
    """Density matrix rho_ab(k) of G(i w, k) per k (gw.cpp:36-52): the tail-fit
    regularised Matsubara sum of every (k, a, b) column."""
    wmesh, kmesh = g_wk.mesh[0], g_wk.mesh[1]
    x = _dev(g_wk.data)
    c = _lib.ctx(x.device.index)
    out = torch.empty(tuple(x.shape[1:]), dtype=torch.complex128, device=x.device)
    _lib.tail_status_reset(c)
    _lib.call("tprf_density", c, wmesh.beta, wmesh.n_max, x[0].numel(), _lib.ptr(x), _lib.ptr(out))
    _lib.tail_status(c)
    return Gf(kmesh, data=out)


def gw_dynamic_sigma(W_tr, g_tr):
    """sigma_ab(tau, r) = -sum_cd W_acdb(tau, r) G_cd(tau, r) (gw.cpp:66-102)."""
    Wtm, gtm = W_tr.mesh[0], g_tr.mesh[0]
    if Wtm.size != gtm.size or Wtm.beta != gtm.beta:
        raise _lib.TprfError(-1, "gw_sigma_tr: tau meshes are not the same.")
    if Wtm.statistic != BOSON or gtm.statistic != FERMION:
        raise _lib.TprfError(-1, "gw_sigma_tr: statistics are incorrect.")
    if W_tr.mesh[1] != g_tr.mesh[1]:
        raise _lib.TprfError(-1, "gw_sigma_tr: real-space meshes are not the same.")
    w = _dev(W_tr.data)
    g = _dev(g_tr.data)
    n = g_tr.target_shape[0]
    c = _lib.ctx(g.device.index)
    out = torch.empty_like(g)
    _lib.call("tprf_gw_dynamic_sigma", c, g.numel() // (n * n), n, _lib.ptr(w), _lib.ptr(g), _lib.ptr(out))
    return Gf(g_tr.mesh, data=out)


def _fourier_k_r(x, mesh, to_r):
    """One-frequency lattice transform of a k (or r) function (make_gf_from_fourier)."""
    d = _dev(x)
    c = _lib.ctx(d.device.index)
    out = torch.empty_like(d)
    name = "tprf_fourier_wk_to_wr" if to_r else "tprf_fourier_wr_to_wk"
    _lib.call(name, c, 1, _lib.dims3(mesh.dims), d.numel() // mesh.size, _lib.ptr(d), _lib.ptr(out))
    return out


def _static_sigma(v_k, g_wk, fock):
    kmesh = g_wk.mesh[1]
    if v_k.mesh != kmesh:
        what = "fock_sigma" if fock else "hartree_sigma"
        raise _lib.TprfError(-1, f"{what}: k-space meshes are not the same.")
    rho = rho_k_from_g_wk(g_wk).data
    v = _dev(v_k.data)
    n = g_wk.target_shape[0]
    c = _lib.ctx(v.device.index)
    out = torch.empty_like(rho)
    _lib.call("tprf_static_sigma_k", c, 1 if fock else 0, _lib.dims3(kmesh.dims), n, _lib.ptr(v), _lib.ptr(rho),
              _lib.ptr(out))
    return Gf(kmesh, data=out)


def hartree_sigma(v_k, g):
    """hartree_sigma(v_k, g_wk): sigma_ab(k) = (1/N) sum_q V_abcd(q) rho_cd(k + q)
    (gw.cpp:124-147); hartree_sigma(v_k, rho_r): sigma_ab(r = 0) = V_abcd(k = 0)
    rho_cd(r = 0) (gw.cpp:113-122)."""
    if isinstance(g.mesh, MeshProduct):
        return _static_sigma(v_k, g, fock=False)
    v = _dev(v_k.data)
    rho = _dev(g.data)
    n = g.target_shape[0]
    c = _lib.ctx(v.device.index)
    out = torch.empty_like(rho)
    _lib.call("tprf_hartree_sigma_r", c, g.mesh.size, n, _lib.ptr(v), _lib.ptr(rho), _lib.ptr(out))
    return Gf(g.mesh, data=out)


def fock_sigma(v, g):
    """fock_sigma(v_k, g_wk): sigma_ab(k) = -(1/N) sum_q V_acdb(q) rho_dc(k + q)
    (gw.cpp:169-195); fock_sigma(v_r, rho_r): sigma_ab(r) = -V_acdb(r) rho_dc(r)
    (gw.cpp:149-167)."""
    if isinstance(g.mesh, MeshProduct):
        return _static_sigma(v, g, fock=True)
    if v.mesh != g.mesh:
        raise _lib.TprfError(-1, "fock_sigma: r-space meshes are not the same.")
    vd = _dev(v.data)
    rho = _dev(g.data)
    n = g.target_shape[0]
    c = _lib.ctx(vd.device.index)
    out = torch.empty_like(rho)
    _lib.call("tprf_fock_sigma_r", c, g.mesh.size, n, _lib.ptr(vd), _lib.ptr(rho), _lib.ptr(out))
    return Gf(g.mesh, data=out)


def gw_sigma(W, g_wk):
    """GW self-energy (gw.cpp:204-244).  gw_sigma(W_wk, g_wk): W is split into a
    dynamic part and its constant tail; the dynamic part gives
    sigma(tau, r) = -W_dyn(tau, r) G(tau, r), transformed back to (i w, k); the
    constant part gives the Fock term -V(r) rho(r) of the density rho(k) of
    g_wk, transformed to k and added at every frequency.  gw_sigma(v_k, g_wk)
    with a static v_k is the Fock self-energy (gw.cpp:198-201)."""
    if not isinstance(W.mesh, MeshProduct):
        return fock_sigma(W, g_wk)
    Wwm, gwm = W.mesh[0], g_wk.mesh[0]
    if Wwm.beta != gwm.beta:
        raise _lib.TprfError(-1, "gw_sigma: inverse temperatures are not the same.")
    if Wwm.statistic != BOSON or gwm.statistic != FERMION:
        raise _lib.TprfError(-1, "gw_sigma: statistics are incorrect.")
    if W.mesh[1] != g_wk.mesh[1]:
        raise _lib.TprfError(-1, "gw_sigma: k-space meshes are not the same.")
    kmesh = g_wk.mesh[1]
    W_dyn_wk, W_const_k = split_into_dynamic_wk_and_constant_k(W)
    g_tr = fourier_wr_to_tr(fourier_wk_to_wr(g_wk))
    W_dyn_tr = chi_tr_from_chi_wr(chi_wr_from_chi_wk(W_dyn_wk))
    sigma_dyn_tr = gw_dynamic_sigma(W_dyn_tr, g_tr)
    rmesh = MeshCycLat(kmesh.dims)
    W_const_r = Gf(rmesh, data=_fourier_k_r(W_const_k.data, kmesh, True))
    rho_r = Gf(rmesh, data=_fourier_k_r(rho_k_from_g_wk(g_wk).data, kmesh, True))
    sigma_fock_r = fock_sigma(W_const_r, rho_r)
    sigma_fock_k = Gf(kmesh, data=_fourier_k_r(sigma_fock_r.data, kmesh, False))
    sigma_wk = fourier_wr_to_wk(fourier_tr_to_wr(sigma_dyn_tr, nw=gwm.n_max))
    return add_dynamic_and_static(sigma_wk, sigma_fock_k)


# ---------------------------------------------------------------------------
# f4 consumers of chi: linearised Eliashberg -- eliashberg.hpp:52-199 / eliashberg.cpp

def eliashberg_g_delta_g_product(g_wk, delta_wk):
    """F_dc(w, k) = sum_ef G_cf(w, k) conj(G_ed(w, -k)) Delta_ef(w, k) on Delta's
    mesh (eliashberg.cpp:31-58)."""
    g = _dev(g_wk.data)
    dl = _dev(delta_wk.data)
    n = g_wk.target_shape[0]
    kmesh = delta_wk.mesh[1]
    c = _lib.ctx(g.device.index)
    out = torch.empty_like(dl)
    _lib.call("tprf_eliashberg_g_delta_g_product", c, g_wk.mesh[0].n_max, delta_wk.mesh[0].n_max,
              _lib.dims3(kmesh.dims), n, _lib.ptr(g), _lib.ptr(dl), _lib.ptr(out))
    return Gf(delta_wk.mesh, data=out)


def eliashberg_product(Gamma_pp, g_wk, delta_wk):
    """Linearised Eliashberg product by summation (eliashberg.cpp:92-120):
    Delta_ab(w, k) = -1/(2 N beta) sum_{n,q} Gamma_cadb(w - n, k - q) F_dc(n, q)."""
    gm = _dev(Gamma_pp.data)
    g = _dev(g_wk.data)
    dl = _dev(delta_wk.data)
    n = g_wk.target_shape[0]
    wmesh, kmesh = delta_wk.mesh[0], delta_wk.mesh[1]
    c = _lib.ctx(g.device.index)
    out = torch.empty_like(dl)
    _lib.call("tprf_eliashberg_product", c, wmesh.beta, Gamma_pp.mesh[0].n_max, g_wk.mesh[0].n_max, wmesh.n_max,
              _lib.dims3(kmesh.dims), n, _lib.ptr(gm), _lib.ptr(g), _lib.ptr(dl), _lib.ptr(out))
    return Gf(delta_wk.mesh, data=out)


def dynamic_and_constant_to_tr(Gamma_pp_dyn_wk, Gamma_pp_const_k):
    """(Gamma_dyn(tau, r), Gamma_const(r)) from (Gamma_dyn(i nu, k), Gamma_const(k))
    (eliashberg.cpp:122-134)."""
    kmesh = Gamma_pp_const_k.mesh
    dyn_tr = chi_tr_from_chi_wr(chi_wr_from_chi_wk(Gamma_pp_dyn_wk))
    const_r = Gf(MeshCycLat(kmesh.dims), data=_fourier_k_r(Gamma_pp_const_k.data, kmesh, True))
    return dyn_tr, const_r


def _gamma_f(gamma, F, mesh):
    gd = _dev(gamma)
    f = _dev(F)
    n = f.shape[-1]
    npts = f.numel() // (n * n)
    c = _lib.ctx(f.device.index)
    out = torch.empty_like(f)
    _lib.call("tprf_eliashberg_gamma_f_product", c, npts, n, _lib.ptr(gd), _lib.ptr(f), _lib.ptr(out))
    return Gf(mesh, data=out)


def eliashberg_dynamic_gamma_f_product(Gamma_pp_dyn_tr, F_tr):
    """Delta_ab(tau, r) = -1/2 sum_cd Gamma_cadb(tau, r) F_dc(tau, r) (eliashberg.cpp:214-240)."""
    if Gamma_pp_dyn_tr.mesh[0].size != F_tr.mesh[0].size:
        raise _lib.TprfError(-1, f"The size of the imaginary time mesh of Gamma ({Gamma_pp_dyn_tr.mesh[0].size}) "
                                 f"must be the size of the mesh of Delta ({F_tr.mesh[0].size}).")
    return _gamma_f(Gamma_pp_dyn_tr.data, F_tr.data, F_tr.mesh)


def eliashberg_constant_gamma_f_product(Gamma_pp_const_r, F_tr):
    """Delta_ab(r) = -1/2 sum_cd Gamma_cadb(r) F_dc(tau = 0, r) (eliashberg.cpp:175-186)."""
    return _gamma_f(Gamma_pp_const_r.data, _dev(F_tr.data)[0], F_tr.mesh[1])


def eliashberg_product_fft(Gamma_pp_dyn_tr, Gamma_pp_const_r, g_wk, delta_wk):
    """Linearised Eliashberg product through the convolution theorem
    (eliashberg.cpp:242-268): F = G G Delta to (tau, r); dynamic and constant
    Gamma products; back to (i w, k)."""
    F_wk = eliashberg_g_delta_g_product(g_wk, delta_wk)
    F_tr = fourier_wr_to_tr(fourier_wk_to_wr(F_wk))
    delta_tr = eliashberg_dynamic_gamma_f_product(Gamma_pp_dyn_tr, F_tr)
    delta_r = eliashberg_constant_gamma_f_product(Gamma_pp_const_r, F_tr)
    delta_wr = fourier_tr_to_wr(delta_tr)
    delta_wr = add_dynamic_and_static(delta_wr, delta_r)
    return fourier_wr_to_wk(delta_wr)


def eliashberg_product_fft_constant(Gamma_pp_const_r, g_wk, delta_wk):
    """The product for a Gamma constant in frequency (eliashberg.cpp:279-302)."""
    F_wk = eliashberg_g_delta_g_product(g_wk, delta_wk)
    F_tr = fourier_wr_to_tr(fourier_wk_to_wr(F_wk))
    delta_r = eliashberg_constant_gamma_f_product(Gamma_pp_const_r, F_tr)
    kmesh = delta_wk.mesh[1]
    delta_k = Gf(kmesh, data=_fourier_k_r(delta_r.data, kmesh, False))
    zero = Gf(F_wk.mesh, data=torch.zeros_like(_dev(F_wk.data)))
    return add_dynamic_and_static(zero, delta_k)


def construct_phi_wk(chi, U):
    """Phi = U chi U per (w, k) on the matrices M[(a,b),(d,c)] = X_abcd
    (eliashberg.cpp:336-360)."""
    x = _dev(chi.data)
    n = chi.target_shape[0]
    u = _dev(U.data if isinstance(U, Gf) else U).reshape(n, n, n, n).contiguous()
    c = _lib.ctx(x.device.index)
    out = torch.empty_like(x)
    _lib.call("tprf_construct_phi_wk", c, x.numel() // n ** 4, n, _lib.ptr(x), _lib.ptr(u), _lib.ptr(out))
    return Gf(chi.mesh, data=out)
