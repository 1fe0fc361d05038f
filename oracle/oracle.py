"""numpy/ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, the smoke check of the package's entry module and the
``cpu_baseline`` leg of bench.py use this module, always as the oracle the product is compared with.  It loads
oracle/libtprf_oracle.so (built by ``make -C oracle``) and exposes one
function per restated reference function, taking and returning complex128
numpy arrays in the reference's row-major layout.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtprf_oracle.so")
_lib = None

FERMION, BOSON = 0, 1


class OracleError(RuntimeError):
    pass


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d, i, l, p = ctypes.c_double, ctypes.c_int, ctypes.c_long, ctypes.c_void_p
        sig = {
            "or_lattice_dyson_g_wk": [d, i, d, i, i, p, i, p, p],
            "or_fourier_lattice": [l, p, l, p, p, i],
            "or_fourier_w_to_t": [i, d, i, i, l, p, p, p, p, p],
            "or_tail_moments": [i, d, i, l, p, p, p],
            "or_fourier_t_to_w": [i, d, i, i, l, p, p],
            "or_chi0_tr_from_grt_PH": [i, p, i, p, p],
            "or_chi0_wr_from_grt_PH": [d, i, p, i, i, p, p],
            "or_chi0_w0r_from_grt_PH": [d, i, p, i, p, p],
            "or_chi_w0r_from_chi_tr": [d, i, l, i, p, p],
            "or_solve_rpa_PH": [l, i, p, p, p],
            "or_screened_interaction_W": [l, l, i, i, p, p, p],
            "or_lindhard_chi00": [d, d, i, p, i, p, p],
            "or_eigh": [i, p, p, p],
            "or_fft1d": [i, i, p, p],
            "or_vandermonde_inverse8": [p],
            "or_fermi": [d],
            "or_num_threads": [],
            "or_set_num_threads": [i],
            "or_set_fit_check": [i],
        }
        for name, args in sig.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib.or_fermi.restype = ctypes.c_double
        _lib.or_vandermonde_inverse8.restype = None
        _lib.or_set_num_threads.restype = None
        _lib.or_set_fit_check.restype = None
    return _lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dims(dims):
    dims = tuple(int(x) for x in dims) + (1,) * (3 - len(dims))
    return (ctypes.c_int * 3)(*dims), dims


def _check(st, what):
    if st != 0:
        raise OracleError(f"{what}: oracle status {st}")


def set_num_threads(n):
    lib().or_set_num_threads(int(n))


def num_threads():
    return lib().or_num_threads()


def fermi(e):
    return lib().or_fermi(float(e))


def fft1d(x, sign):
    x = _c(x)
    out = np.empty_like(x)
    _check(lib().or_fft1d(len(x), int(sign), _p(x), _p(out)), "fft1d")
    return out


def vandermonde_inverse8():
    out = np.empty(64, dtype=np.float64)
    lib().or_vandermonde_inverse8(_p(out))
    return out.reshape(8, 8)


def eigh(a):
    a = _c(a)
    n = a.shape[0]
    ev = np.empty(n, dtype=np.float64)
    vec = np.empty((n, n), dtype=np.complex128)
    _check(lib().or_eigh(n, _p(a), _p(ev), _p(vec)), "eigh")
    return ev, vec


def lattice_dyson_g_wk(beta, n_iw, mu, e_k, sigma=None):
    """e_k[nk, n, n] -> g_wk[2 n_iw, nk, n, n]; sigma None, [2n_iw,n,n] or [2n_iw,nk,n,n]."""
    e_k = _c(e_k)
    nk, n = e_k.shape[0], e_k.shape[1]
    kind, sp = 0, None
    if sigma is not None:
        sigma = _c(sigma)
        kind = 1 if sigma.ndim == 3 else 2
        sp = _p(sigma)
    out = np.empty((2 * n_iw, nk, n, n), dtype=np.complex128)
    _check(lib().or_lattice_dyson_g_wk(beta, n_iw, mu, nk, n, _p(e_k), kind, sp, _p(out)), "dyson")
    return out


def lattice_dyson_g_zk(z, mu, e_k, sigma=None):
    """G[z, k] = inverse((z + mu) 1 - e_k - Sigma) on given frequencies z (real
    frequencies with their i delta: lattice_dyson_g0_fk / g_fk, gf.cpp:71-86,
    92-123, 142-148); sigma None, [nz, n, n] or [nz, nk, n, n].  Restated with
    numpy's inverse, point by point."""
    z = np.asarray(z, dtype=np.complex128)
    e_k = np.asarray(e_k, dtype=np.complex128)
    nk, n = e_k.shape[0], e_k.shape[1]
    M = (z[:, None, None, None] + mu) * np.eye(n)[None, None] - e_k[None]
    if sigma is not None:
        sigma = np.asarray(sigma, dtype=np.complex128)
        M = M - (sigma[:, None] if sigma.ndim == 3 else sigma)
    return np.linalg.inv(M)


def lattice_dyson_g_z(z, mu, e_k, sigma=None):
    """The k average (1/N_k) sum_k G[z, k], summed in k order as lattice_dyson_g_X
    accumulates (gf.cpp:153-204: lattice_dyson_g_w, lattice_dyson_g_f)."""
    g = lattice_dyson_g_zk(z, mu, e_k, sigma)
    acc = np.zeros((g.shape[0],) + g.shape[2:], dtype=np.complex128)
    for k in range(g.shape[1]):
        acc += g[:, k]
    return acc / g.shape[1]


def fourier_lattice(x, dims, direction):
    """x[nouter, nk, *target]; direction -1 = k->r (scaled 1/N_k), +1 = r->k."""
    x = _c(x)
    cd, dims = _dims(dims)
    nouter, nk = x.shape[0], x.shape[1]
    assert nk == dims[0] * dims[1] * dims[2]
    ntarget = int(np.prod(x.shape[2:])) if x.ndim > 2 else 1
    out = np.empty_like(x)
    _check(lib().or_fourier_lattice(nouter, cd, ntarget, _p(x), _p(out), int(direction)), "fourier_lattice")
    return out


def fourier_w_to_t(gw, beta, statistic=FERMION, ntau=None, moments=None, return_moments=False):
    """gw[nmesh, *cols] -> gt[ntau, *cols] (ntau default 6*n_max+1)."""
    gw = _c(gw)
    nmesh = gw.shape[0]
    n_max = nmesh // 2 if statistic == FERMION else (nmesh + 1) // 2
    if ntau is None:
        ntau = 6 * n_max + 1
    ncol = int(np.prod(gw.shape[1:])) if gw.ndim > 1 else 1
    out = np.empty((ntau,) + gw.shape[1:], dtype=np.complex128)
    mp = None
    if moments is not None:
        moments = _c(moments).reshape(3, ncol)
        mp = _p(moments)
    mout = np.empty((3, ncol), dtype=np.complex128)
    err = ctypes.c_double(0.0)
    st = lib().or_fourier_w_to_t(statistic, beta, n_max, ntau, ncol, _p(gw), _p(out), mp, _p(mout),
                                 ctypes.cast(ctypes.pointer(err), ctypes.c_void_p))
    _check(st, "fourier_w_to_t")
    if return_moments:
        return out, mout.reshape((3,) + gw.shape[1:]), err.value
    return out


def fourier_t_to_w(gt, beta, nw, statistic=BOSON):
    gt = _c(gt)
    ntau = gt.shape[0]
    ncol = int(np.prod(gt.shape[1:])) if gt.ndim > 1 else 1
    nout = 2 * nw if statistic == FERMION else 2 * nw - 1
    out = np.empty((nout,) + gt.shape[1:], dtype=np.complex128)
    _check(lib().or_fourier_t_to_w(statistic, beta, ntau, nw, ncol, _p(gt), _p(out)), "fourier_t_to_w")
    return out


def w_to_t_matrix(beta, n_max, statistic=FERMION, ntau=None):
    """Matrix T[ntau, nmesh] of the omega->tau transform with fitted tail moments.

    The transform (fourier_matsubara.cpp:218-308) is complex-linear in the
    Matsubara data once the least-squares tail fit is linear, so T is read off
    column by column from unit inputs (the fit-error assertion is switched off
    for that, since a unit input is no smooth Green's function).
    """
    nmesh = 2 * n_max if statistic == FERMION else 2 * n_max - 1
    lib().or_set_fit_check(0)
    try:
        return fourier_w_to_t(np.eye(nmesh, dtype=np.complex128), beta, statistic, ntau)
    finally:
        lib().or_set_fit_check(1)


def w_to_t_moments_matrix(beta, n_max, statistic=FERMION, ntau=None):
    """Matrix Mop[3, nmesh] of the fitted moments (m1, m2, m3) as a linear map of
    the Matsubara data (the tail fit of fourier_matsubara.cpp:225-246)."""
    nmesh = 2 * n_max if statistic == FERMION else 2 * n_max - 1
    lib().or_set_fit_check(0)
    try:
        _, m, _ = fourier_w_to_t(np.eye(nmesh, dtype=np.complex128), beta, statistic, ntau, return_moments=True)
    finally:
        lib().or_set_fit_check(1)
    return m


def t_to_w_matrix(beta, ntau, nw, statistic=BOSON, chunk=1024):
    """Matrix W[nmesh(nw), ntau] of the tau->omega transform (fourier_matsubara.cpp:42-184)."""
    cols = []
    for j0 in range(0, ntau, chunk):
        e = np.zeros((ntau, min(chunk, ntau - j0)), dtype=np.complex128)
        e[j0 + np.arange(e.shape[1]), np.arange(e.shape[1])] = 1.0
        cols.append(fourier_t_to_w(e, beta, nw, statistic))
    return np.concatenate(cols, axis=1)


def chi0_tr_from_grt_PH(g_tr, dims):
    g_tr = _c(g_tr)
    ntau, nr, n = g_tr.shape[0], g_tr.shape[1], g_tr.shape[2]
    cd, _ = _dims(dims)
    out = np.empty((ntau, nr, n, n, n, n), dtype=np.complex128)
    _check(lib().or_chi0_tr_from_grt_PH(ntau, cd, n, _p(g_tr), _p(out)), "chi0_tr")
    return out


def chi0_wr_from_grt_PH(g_tr, dims, beta, nw):
    g_tr = _c(g_tr)
    ntau, nr, n = g_tr.shape[0], g_tr.shape[1], g_tr.shape[2]
    cd, _ = _dims(dims)
    out = np.empty((2 * nw - 1, nr, n, n, n, n), dtype=np.complex128)
    _check(lib().or_chi0_wr_from_grt_PH(beta, ntau, cd, n, nw, _p(g_tr), _p(out)), "chi0_wr")
    return out


def chi0_w0r_from_grt_PH(g_tr, dims, beta):
    g_tr = _c(g_tr)
    ntau, nr, n = g_tr.shape[0], g_tr.shape[1], g_tr.shape[2]
    cd, _ = _dims(dims)
    out = np.empty((1, nr, n, n, n, n), dtype=np.complex128)
    _check(lib().or_chi0_w0r_from_grt_PH(beta, ntau, cd, n, _p(g_tr), _p(out)), "chi0_w0r")
    return out


def chi_w0r_from_chi_tr(chi_tr, beta):
    chi_tr = _c(chi_tr)
    ntau, nr, n = chi_tr.shape[0], chi_tr.shape[1], chi_tr.shape[2]
    out = np.empty((1, nr, n, n, n, n), dtype=np.complex128)
    _check(lib().or_chi_w0r_from_chi_tr(beta, ntau, nr, n, _p(chi_tr), _p(out)), "chi_w0r")
    return out


def solve_rpa_PH(chi0, U):
    chi0 = _c(chi0)
    U = _c(U)
    n = U.shape[0]
    npts = chi0.size // n ** 4
    out = np.empty_like(chi0)
    _check(lib().or_solve_rpa_PH(npts, n, _p(chi0), _p(U), _p(out)), "solve_rpa")
    return out


def screened_interaction_W(PI, V, dynamic=False):
    """dynamical_screened_interaction_W (dynamical_screened_interaction.cpp):
    PI[..., nk, n, n, n, n] with V[nk, n, n, n, n] (static, :79-116) or V of
    PI's shape (dynamic, :33-74)."""
    PI = _c(PI)
    V = _c(V)
    n = PI.shape[-1]
    nk = V.shape[0] if not dynamic else 1
    npts = PI.size // n ** 4
    out = np.empty_like(PI)
    _check(lib().or_screened_interaction_W(npts, nk, n, 1 if dynamic else 0, _p(PI), _p(V), _p(out)),
           "screened_interaction_W")
    return out


def lindhard_chi00(e_k, dims, beta, mu, nw):
    e_k = _c(e_k)
    n = e_k.shape[1]
    cd, dims = _dims(dims)
    nk = e_k.shape[0]
    out = np.empty((2 * nw - 1, nk, n, n, n, n), dtype=np.complex128)
    _check(lib().or_lindhard_chi00(beta, mu, nw, cd, n, _p(e_k), _p(out)), "lindhard")
    return out


# ---------------------------------------------------------------------------
# f4 consumers of chi (GW: c++/triqs_tprf/lattice/gw.cpp; Eliashberg:
# c++/triqs_tprf/lattice/eliashberg.cpp).  The pointwise contractions are numpy
# einsums over the reference's index patterns; the transforms are the oracle's.

def tail_moments(xw, beta, statistic):
    """Fitted moments t_0..t_3 [4, *cols] of every column of xw [mesh, *cols]."""
    xw = _c(xw)
    nmesh = xw.shape[0]
    n_max = nmesh // 2 if statistic == FERMION else (nmesh + 1) // 2
    ncol = int(np.prod(xw.shape[1:])) if xw.ndim > 1 else 1
    t = np.empty((4, ncol), dtype=np.complex128)
    err = ctypes.c_double(0.0)
    _check(lib().or_tail_moments(statistic, beta, n_max, ncol, _p(xw), _p(t),
                                 ctypes.cast(ctypes.pointer(err), ctypes.c_void_p)), "tail_moments")
    return t.reshape((4,) + xw.shape[1:])


def split_into_dynamic_wk_and_constant_k(chi_wk, beta, statistic=BOSON):
    """lattice_utility.cpp:57-84: constant = fitted t_0 per column."""
    c = tail_moments(chi_wk, beta, statistic)[0]
    return chi_wk - c[None], c


def density(g_w, beta):
    """TRIQS density() of a fermionic Matsubara function per column (used by
    rho_k_from_g_wk, gw.cpp:36-52): the Matsubara sum regularised by the fitted
    moments t1..t3, whose own sums are 1/2, -beta/4 and 0."""
    g_w = _c(g_w)
    n_max = g_w.shape[0] // 2
    t = tail_moments(g_w, beta, FERMION)
    iw = 1j * np.pi * (2 * np.arange(-n_max, n_max) + 1) / beta
    shp = (-1,) + (1,) * (g_w.ndim - 1)
    iw = iw.reshape(shp)
    rem = g_w - t[1][None] / iw - t[2][None] / iw ** 2 - t[3][None] / iw ** 3
    return rem.sum(0) / beta + t[1] / 2 - t[2] * beta / 4


def gw_dynamic_sigma(W_tr, g_tr):
    """gw.cpp:88-102: sigma_ab = -sum_cd W_acdb g_cd per (tau, r)."""
    return -np.einsum("...acdb,...cd->...ab", W_tr, g_tr)


def fock_sigma_r(v_r, rho_r):
    """gw.cpp:149-167: sigma_ab(r) = -sum_cd V_acdb(r) rho_dc(r)."""
    return -np.einsum("...acdb,...dc->...ab", v_r, rho_r)


def hartree_sigma_r(v_k, rho_r):
    """gw.cpp:113-122: only r = 0, V(k = 0) rho(r = 0)."""
    out = np.zeros_like(rho_r)
    out[0] = np.einsum("abcd,cd->ab", v_k[0], rho_r[0])
    return out


def _dims3(dims):
    return tuple(int(x) for x in dims) + (1,) * (3 - len(dims))


def _kq_index(dims, sign):
    """idx[k, q] = k + sign q on the lattice."""
    g = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1).reshape(-1, 3)
    j = (g[:, None, :] + sign * g[None, :, :]) % np.array(dims)
    return (j[..., 0] * dims[1] + j[..., 1]) * dims[2] + j[..., 2]


def static_sigma_k(v_k, rho_k, dims, fock):
    """hartree_sigma / fock_sigma (v_k, g_wk) by the direct q sum (gw.cpp:124-147,
    169-195); rho_k from ``density``."""
    dims = _dims3(dims)
    nk = v_k.shape[0]
    rho_kq = rho_k[_kq_index(dims, +1)]  # [k, q, n, n]
    if fock:
        return -np.einsum("qacdb,kqdc->kab", v_k, rho_kq) / nk
    return np.einsum("qabcd,kqcd->kab", v_k, rho_kq) / nk


def gw_sigma(W_wk, g_wk, beta, dims):
    """gw.cpp:204-244 composed from the oracle's transforms: W_wk [2n-1, nk, n^4]
    bosonic with n_max equal to g_wk's fermionic n_max."""
    dims = _dims3(dims)
    n_max = g_wk.shape[0] // 2
    W_dyn, W_const = split_into_dynamic_wk_and_constant_k(W_wk, beta, BOSON)
    g_tr = fourier_w_to_t(fourier_lattice(g_wk, dims, -1), beta, FERMION)
    W_tr = fourier_w_to_t(fourier_lattice(W_dyn, dims, -1), beta, BOSON)
    s_tr = gw_dynamic_sigma(W_tr, g_tr)
    W_const_r = fourier_lattice(W_const[None], dims, -1)[0]
    rho_r = fourier_lattice(density(g_wk, beta)[None], dims, -1)[0]
    fock_k = fourier_lattice(fock_sigma_r(W_const_r, rho_r)[None], dims, +1)[0]
    s_wk = fourier_lattice(fourier_t_to_w(s_tr, beta, n_max, FERMION), dims, +1)
    return s_wk + fock_k[None]


def eliashberg_g_delta_g_product(g_wk, delta_wk, dims):
    """eliashberg.cpp:31-58: F_dc(w, k) = sum_ef G_cf(w, k) conj(G_ed(w, -k)) Delta_ef(w, k)."""
    dims = _dims3(dims)
    nwg, nwd = g_wk.shape[0], delta_wk.shape[0]
    off = (nwg - nwd) // 2
    g = g_wk[off:off + nwd]
    mk = _kq_index(dims, -1)[0]  # 0 - k
    return np.einsum("wkcf,wked,wkef->wkdc", g, np.conj(g[:, mk]), delta_wk)


def eliashberg_product(gamma_wk, g_wk, delta_wk, beta, dims):
    """eliashberg.cpp:92-120 by brute force: Delta_ab(w, k) = -1/(2 N beta)
    sum_{n, q} Gamma_cadb(w - n, k - q) F_dc(n, q)."""
    dims = _dims3(dims)
    F = eliashberg_g_delta_g_product(g_wk, delta_wk, dims)
    nw, nk = delta_wk.shape[0], delta_wk.shape[1]
    ng = (gamma_wk.shape[0] + 1) // 2
    kmq = _kq_index(dims, -1)  # [k, q] -> k - q
    w = np.arange(nw)
    gidx = (w[:, None] - w[None, :]) + (ng - 1)  # [w, n]
    G = gamma_wk[gidx][:, :, kmq]  # [w, n, k, q, c, a, d, b]
    out = np.einsum("wnkqcadb,nqdc->wkab", G, F, optimize=True)
    return -out / (2 * nk * beta)


def eliashberg_gamma_f(gamma, F):
    """eliashberg.cpp:175-186, 214-240: Delta_ab = -1/2 sum_cd Gamma_cadb F_dc."""
    return -0.5 * np.einsum("...cadb,...dc->...ab", gamma, F)


def eliashberg_product_fft(gamma_dyn_tr, gamma_const_r, g_wk, delta_wk, beta, dims):
    """eliashberg.cpp:242-268 composed from the oracle's transforms."""
    dims = _dims3(dims)
    n_max = delta_wk.shape[0] // 2
    F = eliashberg_g_delta_g_product(g_wk, delta_wk, dims)
    F_tr = fourier_w_to_t(fourier_lattice(F, dims, -1), beta, FERMION)
    d_tr = eliashberg_gamma_f(gamma_dyn_tr, F_tr)
    d_r = eliashberg_gamma_f(gamma_const_r, F_tr[0])
    d_wr = fourier_t_to_w(d_tr, beta, n_max, FERMION) + d_r[None]
    return fourier_lattice(d_wr, dims, +1)


def preprocess_gamma_for_fft(gamma_wk, beta, dims):
    """eliashberg.py:196-254 with a fitted constant: (Gamma_dyn(tau, r), Gamma_const(r))."""
    dims = _dims3(dims)
    dyn, cst = split_into_dynamic_wk_and_constant_k(gamma_wk, beta, BOSON)
    dyn_tr = fourier_w_to_t(fourier_lattice(dyn, dims, -1), beta, BOSON)
    return dyn_tr, fourier_lattice(cst[None], dims, -1)[0]


def construct_phi_wk(chi, U):
    """eliashberg.cpp:336-360: Phi = U chi U on M[(a,b),(d,c)] = X_abcd."""
    n = U.shape[0]
    def grp(x):
        return np.swapaxes(x, -1, -2).reshape(x.shape[:-4] + (n * n, n * n))
    def ungrp(m):
        return np.swapaxes(m.reshape(m.shape[:-2] + (n, n, n, n)), -1, -2)
    Um = grp(U)
    return ungrp(Um @ grp(chi) @ Um)
