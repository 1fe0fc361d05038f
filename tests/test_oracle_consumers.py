"""Pins of the oracle's GW and Eliashberg restatements (consumers of chi:
c++/triqs_tprf/lattice/gw.cpp, eliashberg.cpp, lattice_utility.cpp) against
closed forms and the reference's own cross-checks.  CPU only.

* tail moments / split: exact pole sums have moments sum a b^(k-1); a
  constant plus poles splits into that constant;
* density: G0 of a Hermitian dispersion has rho = f(e - mu) (Fermi function);
* gw_sigma: the single-k electron-phonon self-energy of the reference's
  gw_singlekpoint_exactresult.py has a closed form;
* gw_sigma = dynamic part + Fock part (the reference's gw.py check), and the
  Fock part with the exact density is the g0w_sigma formula (Fermi function);
* Eliashberg: the summation product and the FFT product agree (the
  reference's product_summation_vs_fft.py), and a planted mistake in either
  breaks the agreement;
* construct_phi_wk: the identity in the grouping M[(a,b),(d,c)] returns chi.
"""
import numpy as np
import pytest

from conftest import relerr


def _nF(w, beta):
    return 0.5 - 0.5 * np.tanh(0.5 * beta * w)


def _nB(w, beta):
    return 1.0 / np.expm1(beta * w)


def _iw(beta, n_max):
    return 1j * np.pi * (2 * np.arange(-n_max, n_max) + 1) / beta


def _inu(beta, n_max):
    return 1j * 2 * np.pi * np.arange(-(n_max - 1), n_max) / beta


def test_tail_moments_of_pole_sums(oracle):
    """t_k = sum_i a_i b_i^(k-1) (k >= 1), t_0 = 0 for G = sum a/(iw - b);
    the bosonic split of V + poles returns V.  Bounds: the fit's own accuracy
    at this mesh (t3 is the least determined)."""
    beta, n = 10.0, 256
    a, b = np.array([0.3, 0.5, 0.2]), np.array([-1.2, 0.4, 2.0])
    g = (a[None] / (_iw(beta, n)[:, None] - b[None])).sum(1)[:, None]
    t = oracle.tail_moments(g, beta, oracle.FERMION)[:, 0]
    exact = np.array([0.0, a.sum(), (a * b).sum(), (a * b * b).sum()])
    assert np.abs(t - exact).max() <= 1e-8
    assert abs(t[1] - exact[1]) <= 1e-13
    V = 0.7 - 0.2j
    w = V + (a[None] / (_inu(beta, n)[:, None] - b[None])).sum(1)[:, None]
    dyn, c = oracle.split_into_dynamic_wk_and_constant_k(w, beta, oracle.BOSON)
    assert abs(c[0] - V) <= 1e-13
    assert np.abs(dyn + c[None] - w).max() <= 1e-15


@pytest.mark.parametrize("beta,n_max,tol", [(10.0, 256, 1e-6), (40.0, 512, 5e-6)])
def test_density_is_the_fermi_function(oracle, beta, n_max, tol):
    """rho of G0 = (iw + mu - e)^-1 is f(e - mu); the three-moment regularised
    sum leaves a truncation error of order t4 / w_max^3 (measured 3e-7 and
    2.5e-6 at these meshes)."""
    e, mu = np.linspace(-2, 2, 7), 0.3
    g = 1.0 / (_iw(beta, n_max)[:, None] + mu - e[None, :])
    rho = oracle.density(g, beta)
    assert np.abs(rho - 1 / (np.exp(beta * (e - mu)) + 1)).max() <= tol
    # a matrix-valued G0: rho = U f(eps) U^H
    rng = np.random.default_rng(4)
    h = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    h = (h + h.conj().T) / 2
    eps, U = np.linalg.eigh(h)
    G = np.linalg.inv((_iw(beta, n_max)[:, None, None] + mu) * np.eye(3) - h[None])
    ref = (U * (1 / (np.exp(beta * (eps - mu)) + 1))) @ U.conj().T
    assert np.abs(oracle.density(G, beta) - ref).max() <= tol


def test_gw_sigma_single_k_electron_phonon_exact(oracle):
    """gw_singlekpoint_exactresult.py: one k point, e = 0, mu = 0.5, W the
    phonon propagator g2 2 wD / ((i nu)^2 - wD^2); Sigma(iw) = g2 [(nB + nF)/(iw
    + wD - E) + (nB + 1 - nF)/(iw - wD - E)], E = -mu.  The tau grid of 3001
    points leaves 6.7e-6 (1e-5 relative); bound 2e-5."""
    mu, g2, wD, beta, nw = 0.5, 0.4, 0.1, 300.0, 500
    iw = _iw(beta, nw)
    g_wk = (1.0 / (iw + mu)).reshape(-1, 1, 1, 1)
    W = (g2 * 2 * wD / (_inu(beta, nw) ** 2 - wD ** 2)).reshape(-1, 1, 1, 1, 1, 1)
    s = oracle.gw_sigma(W, g_wk, beta, (1, 1, 1))[:, 0, 0, 0]
    E = -mu
    ex = g2 * ((_nB(wD, beta) + _nF(E, beta)) / (iw + wD - E) + (_nB(wD, beta) + 1 - _nF(E, beta)) / (iw - wD - E))
    assert np.abs(s - ex).max() <= 2e-5 * np.abs(ex).max()


def _chain_e_k(nk, t=-1.0):
    k = 2 * np.pi * np.arange(nk) / nk
    return (2 * t * np.cos(k)).reshape(nk, 1, 1).astype(complex)


def test_gw_sigma_is_dynamic_plus_fock_and_fock_is_g0w(oracle):
    """gw.py (reference test): gw_sigma(W) = gw_sigma(W_dyn) + fock(V_const),
    and the Fock part of G0 equals the g0w_sigma closed form
    -(1/N) sum_q V(q) f(e_{k+q} - mu) (single band) to the density's accuracy."""
    beta, nw, nk, mu, V = 10.0, 256, 8, 0.0, 1.0
    e = _chain_e_k(nk)
    iw = _iw(beta, nw)
    g0 = 1.0 / (iw[:, None, None, None] + mu - e[None])
    # a W with a constant tail and a dynamic part (a plasmon-like pole pair in every k)
    wk = 1.0 + 0.5 * np.cos(2 * np.pi * np.arange(nk) / nk)
    inu = _inu(beta, nw)
    Wd = (0.3 * 2 * wk[None, :] / (inu[:, None] ** 2 - wk[None, :] ** 2)).reshape(-1, nk, 1, 1, 1, 1)
    Wfull = Wd + V
    dims = (nk, 1, 1)
    s_full = oracle.gw_sigma(Wfull, g0, beta, dims)
    s_dyn = oracle.gw_sigma(Wd + 0.0, g0, beta, dims)
    Vk = np.full((nk, 1, 1, 1, 1), V, dtype=complex)
    rho = oracle.density(g0, beta)
    s_stat = oracle.static_sigma_k(Vk, rho, dims, fock=True)
    assert relerr(s_full, s_dyn + s_stat[None]) <= 1e-12
    f = 1 / (np.exp(beta * (e[:, 0, 0].real - mu)) + 1)
    g0w = -V * f.mean()
    assert np.abs(s_stat[:, 0, 0] - g0w).max() <= 1e-6


def _eliashberg_inputs(seed=0):
    """Two-band chain of 3 k points at beta 1 with n_max 30 (the reference's
    product_summation_vs_fft.py sizes), a vertex = constant + boson exchange,
    and a gap random in k decaying as 1/(1 + w^2)."""
    rng = np.random.default_rng(seed)
    beta, nw, nk = 1.0, 30, 3
    k = 2 * np.pi * np.arange(nk) / nk
    e = np.zeros((nk, 2, 2), complex)
    e[:, 0, 0] = -2 * np.cos(k)
    e[:, 1, 1] = -np.cos(k)
    e[:, 0, 1] = e[:, 1, 0] = -0.2 * np.cos(k)
    g = np.linalg.inv(_iw(beta, nw)[:, None, None, None] * np.eye(2) - e[None])
    C = rng.normal(size=(nk, 2, 2, 2, 2)) * 0.3
    D = rng.normal(size=(nk, 2, 2, 2, 2)) * 0.5
    O = rng.normal(size=(nk, 2, 2, 2, 2)) * 0.2

    def gamma(n_max):
        # an even boson exchange and an odd part, so Gamma(nu) != Gamma(-nu)
        inu = _inu(beta, n_max).reshape(-1, 1, 1, 1, 1, 1)
        return C[None] + D[None] * 3.0 / (inu ** 2 - 2.25) + O[None] * inu / (inu ** 2 - 1.0)

    R = rng.normal(size=(nk, 2, 2)) + 1j * rng.normal(size=(nk, 2, 2))
    delta = R[None] / (1 + np.abs(_iw(beta, nw)).reshape(-1, 1, 1, 1) ** 2)
    return beta, nw, (nk, 1, 1), g, gamma, delta


def test_eliashberg_summation_equals_fft(oracle, monkeypatch):
    """product_summation_vs_fft.py: the O(N_w^2 N_k^2) sum with Gamma on twice
    the mesh and the convolution-theorem product agree (reference atol 1e-8;
    measured 2e-11 against a maximum of 9e-3).  Planting k + q for k - q, or
    n - w for w - n, in the summation breaks the agreement."""
    beta, nw, dims, g, gamma, delta = _eliashberg_inputs()
    ds = oracle.eliashberg_product(gamma(2 * nw + 1), g, delta, beta, dims)
    gdyn_tr, gc_r = oracle.preprocess_gamma_for_fft(gamma(nw), beta, dims)
    df = oracle.eliashberg_product_fft(gdyn_tr, gc_r, g, delta, beta, dims)
    assert np.abs(ds - df).max() <= 1e-9
    assert np.abs(ds).max() > 1e-3
    # planted: k + q instead of k - q
    real = oracle._kq_index
    monkeypatch.setattr(oracle, "_kq_index", lambda d, s: real(d, -s))
    bad = oracle.eliashberg_product(gamma(2 * nw + 1), g, delta, beta, dims)
    monkeypatch.undo()
    assert np.abs(bad - df).max() > 1e-4
    # planted: Gamma(n - w) instead of Gamma(w - n) (the vertex has an odd part)
    flipped = oracle.eliashberg_product(gamma(2 * nw + 1)[::-1].copy(), g, delta, beta, dims)
    assert np.abs(flipped - df).max() > 1e-4


def test_eliashberg_fft_constant_gamma_equals_full(oracle):
    """fft_product_constant_vs_full.py: for a Gamma constant in frequency the
    full FFT product equals the constant-only product (the split returns the
    constant and a zero dynamic part)."""
    beta, nw, dims, g, gamma, delta = _eliashberg_inputs(2)
    rng = np.random.default_rng(9)
    C = rng.normal(size=(dims[0], 2, 2, 2, 2)) + 0j
    gw = np.broadcast_to(C[None], (2 * nw - 1,) + C.shape).copy()
    gdyn_tr, gc_r = oracle.preprocess_gamma_for_fft(gw, beta, dims)
    assert np.abs(gdyn_tr).max() <= 1e-12
    full = oracle.eliashberg_product_fft(gdyn_tr, gc_r, g, delta, beta, dims)
    F = oracle.eliashberg_g_delta_g_product(g, delta, dims)
    F_tr = oracle.fourier_w_to_t(oracle.fourier_lattice(F, dims, -1), beta, oracle.FERMION)
    d_k = oracle.fourier_lattice(oracle.eliashberg_gamma_f(gc_r, F_tr[0])[None], dims, +1)[0]
    assert np.abs(full - d_k[None]).max() <= 1e-12 * max(1.0, np.abs(d_k).max())


def test_construct_phi_identity_and_hubbard(oracle):
    """U = identity in the grouping M[(a,b),(d,c)] (U_abcd = d_ad d_bc) gives
    Phi = chi; a single band gives U^2 chi."""
    rng = np.random.default_rng(5)
    n = 2
    chi = rng.normal(size=(4, 3, n, n, n, n)) + 1j * rng.normal(size=(4, 3, n, n, n, n))
    I = np.einsum("ad,bc->abcd", np.eye(n), np.eye(n)).astype(complex)
    assert relerr(oracle.construct_phi_wk(chi, I), chi) <= 1e-15
    c1 = rng.normal(size=(5, 1, 1, 1, 1)) + 0j
    assert relerr(oracle.construct_phi_wk(c1, np.full((1, 1, 1, 1), 2.5 + 0j)), 6.25 * c1) <= 1e-15
