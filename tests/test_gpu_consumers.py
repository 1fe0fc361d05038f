"""GPU parity of the consumers of chi (GW: gw.cpp, Eliashberg: eliashberg.cpp,
lattice_utility.cpp) against the oracle on the same seeded inputs, through
the C ABI and the Python drop-ins of ``tprf_amd.lattice``.

Pointwise pieces (contractions, G Delta G, the summation product, Phi = U chi
U, the tail split and the density) agree with the oracle to 1e-13..1e-12
relative max-norm; the composed gw_sigma and FFT Eliashberg products, whose
stages are the chain's transforms, to 1e-11; and the physics pins of the
oracle (the single-k electron-phonon closed form, summation = FFT) hold for
the GPU results with the oracle's own bounds.
"""
import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return torch


def rnd(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def host(t):
    return t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)


def _iw(beta, n_max):
    return 1j * np.pi * (2 * np.arange(-n_max, n_max) + 1) / beta


def _inu(beta, n_max):
    return 1j * 2 * np.pi * np.arange(-(n_max - 1), n_max) / beta


def _g0(e, beta, n_max, mu=0.0):
    n = e.shape[-1]
    return np.linalg.inv((_iw(beta, n_max)[:, None, None, None] + mu) * np.eye(n) - e[None])


def _herm_e_k(rng, nk, n):
    h = rnd(rng, nk, n, n) * 0.5
    return (h + np.conj(np.swapaxes(h, 1, 2))) / 2


def _gf(mesh, data):
    from tprf_amd.gf import Gf
    return Gf(mesh, data=np.ascontiguousarray(data))


def _meshes(beta, n_f, n_b, dims, ntau=None):
    from tprf_amd.gf import MeshBrZone, MeshImFreq, MeshProduct
    k = MeshBrZone(dims)
    return (MeshProduct(MeshImFreq(beta, "Fermion", n_f), k), MeshProduct(MeshImFreq(beta, "Boson", n_b), k), k)


# ---------------------------------------------------------------- pointwise pieces
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8])
def test_vertex_contractions(T, oracle, n):
    """gw_dynamic_sigma, fock_sigma (r), hartree_sigma (r), the Eliashberg
    gamma-F products: the reference's index patterns against the oracle."""
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshCycLat, MeshImTime, MeshProduct
    rng = np.random.default_rng(n)
    dims, ntau, beta = (3, 2, 2), 7, 5.0
    nr = 12
    W = rnd(rng, ntau, nr, n, n, n, n)
    g = rnd(rng, ntau, nr, n, n)
    rm = MeshCycLat(dims)
    Wtr = _gf(MeshProduct(MeshImTime(beta, "Boson", ntau), rm), W)
    gtr = _gf(MeshProduct(MeshImTime(beta, "Fermion", ntau), rm), g)
    got = host(L.gw_dynamic_sigma(Wtr, gtr).data)
    assert relerr(got, oracle.gw_dynamic_sigma(W, g)) <= 1e-14
    v, rho = rnd(rng, nr, n, n, n, n), rnd(rng, nr, n, n)
    assert relerr(host(L.fock_sigma(_gf(rm, v), _gf(rm, rho)).data), oracle.fock_sigma_r(v, rho)) <= 1e-14
    from tprf_amd.gf import MeshBrZone
    vk = rnd(rng, nr, n, n, n, n)
    got = host(L.hartree_sigma(_gf(MeshBrZone(dims), vk), _gf(rm, rho)).data)
    assert relerr(got, oracle.hartree_sigma_r(vk, rho)) <= 1e-14
    Ftr = _gf(MeshProduct(MeshImTime(beta, "Fermion", ntau), rm), g)
    got = host(L.eliashberg_dynamic_gamma_f_product(Wtr, Ftr).data)
    assert relerr(got, oracle.eliashberg_gamma_f(W, g)) <= 1e-14
    got = host(L.eliashberg_constant_gamma_f_product(_gf(rm, v), Ftr).data)
    assert relerr(got, oracle.eliashberg_gamma_f(v, g[0])) <= 1e-14


@pytest.mark.parametrize("n,dims", [(1, (8, 1, 1)), (2, (4, 3, 1)), (3, (3, 3, 2)), (5, (2, 2, 1))])
def test_static_sigma_k_direct_sums(T, oracle, n, dims):
    """hartree_sigma(v_k, g_wk) and fock_sigma(v_k, g_wk): density per k, then the
    direct q sums, against the oracle."""
    from tprf_amd import lattice as L
    rng = np.random.default_rng(10 + n)
    beta, n_f = 8.0, 64
    nk = int(np.prod(dims))
    e = _herm_e_k(rng, nk, n)
    g = _g0(e, beta, n_f, 0.1)
    wk_mesh, _, kmesh = _meshes(beta, n_f, n_f, dims)
    v = rnd(rng, nk, n, n, n, n)
    rho = oracle.density(g, beta)
    assert relerr(host(L.rho_k_from_g_wk(_gf(wk_mesh, g)).data), rho) <= 1e-12
    for fock, fn in ((False, L.hartree_sigma), (True, L.fock_sigma)):
        got = host(fn(_gf(kmesh, v), _gf(wk_mesh, g)).data)
        assert relerr(got, oracle.static_sigma_k(v, rho, dims, fock)) <= 1e-12, fock


def test_split_into_dynamic_and_constant(T, oracle):
    """The tail split against the oracle: 1e-12 (the device applies the fit
    operator rounded to double, whose entries exceed the data by ~1e2; the
    oracle solves the least squares in long double: ~1e-13 apart)."""
    from tprf_amd import lattice as L
    rng = np.random.default_rng(3)
    beta, nmax, dims, n = 10.0, 128, (4, 2, 1), 2
    nk = 8
    a, b = rnd(rng, 3, nk, n, n, n, n) * 0.3, rng.normal(size=3)
    inu = _inu(beta, nmax).reshape(-1, 1, 1, 1, 1, 1)
    V = rnd(rng, nk, n, n, n, n)
    W = V[None] + sum(a[i][None] / (inu - b[i]) for i in range(3))
    _, wb_mesh, _ = _meshes(beta, nmax, nmax, dims)
    dyn, cst = L.split_into_dynamic_wk_and_constant_k(_gf(wb_mesh, W))
    odyn, ocst = oracle.split_into_dynamic_wk_and_constant_k(W, beta, oracle.BOSON)
    assert relerr(host(cst.data), ocst) <= 1e-12
    assert relerr(host(dyn.data), odyn) <= 1e-12
    assert relerr(host(cst.data), V) <= 1e-10


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_g_delta_g_summation_product_and_phi(T, oracle, n):
    """eliashberg_g_delta_g_product (G on a larger mesh than Delta), the
    summation product and construct_phi_wk (n >= 5: one workgroup per point)
    against the oracle."""
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshBrZone, MeshImFreq, MeshProduct
    rng = np.random.default_rng(20 + n)
    beta, nd, ng_f, dims = 2.0, 6, 9, (3, 2, 1)
    nk = 6
    e = _herm_e_k(rng, nk, n)
    g = _g0(e, beta, ng_f, 0.2)
    delta = rnd(rng, 2 * nd, nk, n, n)
    km = MeshBrZone(dims)
    gG = _gf(MeshProduct(MeshImFreq(beta, "Fermion", ng_f), km), g)
    dG = _gf(MeshProduct(MeshImFreq(beta, "Fermion", nd), km), delta)
    F = host(L.eliashberg_g_delta_g_product(gG, dG).data)
    assert relerr(F, oracle.eliashberg_g_delta_g_product(g, delta, dims)) <= 1e-13
    n_gam = 2 * nd + 1
    gam = rnd(rng, 2 * n_gam - 1, nk, n, n, n, n)
    gamG = _gf(MeshProduct(MeshImFreq(beta, "Boson", n_gam), km), gam)
    got = host(L.eliashberg_product(gamG, gG, dG).data)
    assert relerr(got, oracle.eliashberg_product(gam, g, delta, beta, dims)) <= 1e-13
    U = rnd(rng, n, n, n, n)
    chi = rnd(rng, 5, nk, n, n, n, n)
    phi = host(L.construct_phi_wk(_gf(MeshProduct(MeshImFreq(beta, "Boson", 3), km), chi), U).data)
    assert relerr(phi, oracle.construct_phi_wk(chi, U)) <= 1e-13


def test_eliashberg_summation_too_short_gamma_mesh(T):
    from tprf_amd import _lib
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshBrZone, MeshImFreq, MeshProduct
    km = MeshBrZone((2, 1, 1))
    g = _gf(MeshProduct(MeshImFreq(1.0, "Fermion", 4), km), np.ones((8, 2, 1, 1), complex))
    gam = _gf(MeshProduct(MeshImFreq(1.0, "Boson", 4), km), np.ones((7, 2, 1, 1, 1, 1), complex))
    with pytest.raises(_lib.TprfError, match="TWICE"):
        L.eliashberg_product(gam, g, g)


# ---------------------------------------------------------------- composed products
def test_gw_sigma_two_band_vs_oracle(T, oracle):
    """gw_sigma(W_wk, g_wk) on a two-band 4x4 lattice (W with a constant tail
    and a k-dependent boson pole) against the oracle's composition, 1e-11."""
    from tprf_amd import lattice as L
    rng = np.random.default_rng(31)
    beta, nmax, dims, n = 5.0, 64, (4, 4, 1), 2
    nk = 16
    e = _herm_e_k(rng, nk, n)
    g = _g0(e, beta, nmax, 0.3)
    wq = 1.0 + 0.3 * rng.random(nk)
    V = rnd(rng, nk, n, n, n, n) * 0.2
    D = rnd(rng, nk, n, n, n, n) * 0.2
    inu = _inu(beta, nmax)
    W = V[None] + D[None] * (2 * wq / (inu[:, None] ** 2 - wq[None] ** 2))[..., None, None, None, None]
    wf, wb, _ = _meshes(beta, nmax, nmax, dims)
    got = host(L.gw_sigma(_gf(wb, W), _gf(wf, g)).data)
    ref = oracle.gw_sigma(W, g, beta, dims)
    print(f"gw_sigma two-band vs oracle {relerr(got, ref):.2e}", flush=True)
    assert relerr(got, ref) <= 1e-11


def test_gw_sigma_single_k_electron_phonon_exact(T):
    """The reference's gw_singlekpoint_exactresult.py on the GPU: the closed form
    within the oracle's bound (2e-5 relative; the tau grid)."""
    from tprf_amd import lattice as L
    mu, g2, wD, beta, nw = 0.5, 0.4, 0.1, 300.0, 500
    iw = _iw(beta, nw)
    wf, wb, _ = _meshes(beta, nw, nw, (1, 1, 1))
    g = (1.0 / (iw + mu)).reshape(-1, 1, 1, 1)
    W = (g2 * 2 * wD / (_inu(beta, nw) ** 2 - wD ** 2)).reshape(-1, 1, 1, 1, 1, 1)
    s = host(L.gw_sigma(_gf(wb, W), _gf(wf, g)).data)[:, 0, 0, 0]
    nF = 0.5 - 0.5 * np.tanh(0.5 * beta * (-mu))
    nB = 1.0 / np.expm1(beta * wD)
    ex = g2 * ((nB + nF) / (iw + wD + mu) + (nB + 1 - nF) / (iw - wD + mu))
    err = np.abs(s - ex).max() / np.abs(ex).max()
    print(f"0D GW vs closed form {err:.2e}", flush=True)
    assert err <= 2e-5


def test_eliashberg_fft_vs_oracle_and_summation(T, oracle):
    """eliashberg_product_fft (with preprocess_gamma_for_fft) against the oracle
    (1e-11) and against the GPU summation product (the reference's
    product_summation_vs_fft.py check, 1e-9); eliashberg_product_fft_constant
    for a frequency-constant Gamma equals the full FFT product."""
    from tprf_amd import eliashberg as E
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshBrZone, MeshImFreq, MeshProduct
    from test_oracle_consumers import _eliashberg_inputs
    beta, nw, dims, g, gamma, delta = _eliashberg_inputs()
    km = MeshBrZone(dims)
    fm = MeshProduct(MeshImFreq(beta, "Fermion", nw), km)
    gG, dG = _gf(fm, g), _gf(fm, delta)
    gam_small = _gf(MeshProduct(MeshImFreq(beta, "Boson", nw), km), gamma(nw))
    dyn_tr, const_r = E.preprocess_gamma_for_fft(gam_small)
    got = host(L.eliashberg_product_fft(dyn_tr, const_r, gG, dG).data)
    odyn, oconst = oracle.preprocess_gamma_for_fft(gamma(nw), beta, dims)
    ref = oracle.eliashberg_product_fft(odyn, oconst, g, delta, beta, dims)
    print(f"eliashberg fft vs oracle {relerr(got, ref):.2e}", flush=True)
    assert relerr(got, ref) <= 1e-11
    gam_big = _gf(MeshProduct(MeshImFreq(beta, "Boson", 2 * nw + 1), km), gamma(2 * nw + 1))
    summ = host(L.eliashberg_product(gam_big, gG, dG).data)
    assert np.abs(summ - got).max() <= 1e-9
    C = np.random.default_rng(9).normal(size=(dims[0], 2, 2, 2, 2)) + 0j
    gc = _gf(MeshProduct(MeshImFreq(beta, "Boson", nw), km), np.broadcast_to(C[None], (2 * nw - 1,) + C.shape))
    dyn_tr, const_r = E.preprocess_gamma_for_fft(gc)
    full = host(L.eliashberg_product_fft(dyn_tr, const_r, gG, dG).data)
    const = host(L.eliashberg_product_fft_constant(const_r, gG, dG).data)
    assert relerr(const, full) <= 1e-12
