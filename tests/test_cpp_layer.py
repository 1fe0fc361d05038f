"""The C++ signature layer (include/tprf_hip.hpp) over the C ABI.

CPU: the header declares the reference's hot-path functions and compiles and
links with the host compiler against libtprf_hip.so.
GPU: tests/cpp/chain_cpp, a C++ caller that uses only that header, runs the
chain (lattice_dyson_g0_wk ... solve_rpa_PH, the driver imtime_bubble_chi0_wk,
and forward_batched) on a 3-orbital 4^3 cubic lattice; its results match the
oracle."""
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import relerr, rpa_residual

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HPP = os.path.join(ROOT, "include", "tprf_hip.hpp")
BIN = os.path.join(ROOT, "tests", "cpp", "chain_cpp")

# c++/triqs_tprf/lattice/gf.hpp:44,83,123,276,286,296,306; chi_imtime.hpp:38,40,54,70,87,105,123,141;
# rpa.hpp:43; lindhard_chi00.hpp:97; python/triqs_tprf/lattice_utils.py:191
REFERENCE_FUNCTIONS = ["lattice_dyson_g0_wk", "lattice_dyson_g_wk", "fourier_wk_to_wr", "fourier_wr_to_wk",
                       "fourier_wr_to_tr", "fourier_tr_to_wr", "chi0_tr_from_grt_PH", "chi0_wr_from_grt_PH",
                       "chi0_w0r_from_grt_PH", "chi_w0r_from_chi_tr", "chi_wr_from_chi_tr", "chi_tr_from_chi_wr",
                       "chi_wk_from_chi_wr", "chi_wr_from_chi_wk", "solve_rpa_PH", "lindhard_chi00",
                       "imtime_bubble_chi0_wk", "split_into_dynamic_wk_and_constant_k", "rho_k_from_g_wk",
                       "gw_dynamic_sigma", "hartree_sigma", "fock_sigma", "gw_sigma", "eliashberg_g_delta_g_product",
                       "eliashberg_product", "dynamic_and_constant_to_tr", "eliashberg_dynamic_gamma_f_product",
                       "eliashberg_constant_gamma_f_product", "eliashberg_product_fft",
                       "eliashberg_product_fft_constant", "construct_phi_wk"]


def test_header_declares_reference_signatures():
    text = open(HPP).read()
    for fn in REFERENCE_FUNCTIONS:
        assert re.search(r"\b" + fn + r"\(", text), fn
    # the reference's argument lists: (mu, e_k, mesh | sigma), (g, nt = -1), (chi, nw), (chi0, U)
    assert "lattice_dyson_g0_wk(double mu, e_k_cvt e_k, mesh::imfreq mesh)" in text
    assert "lattice_dyson_g_wk(double mu, e_k_cvt e_k, g_w_cvt sigma_w)" in text
    assert "lattice_dyson_g_wk(double mu, e_k_cvt e_k, g_wk_cvt sigma_wk)" in text
    assert "fourier_wr_to_tr(g_wr_cvt g_wr, int nt = -1)" in text
    assert "chi0_wr_from_grt_PH(g_tr_cvt g_tr, int nw)" in text
    assert "solve_rpa_PH(chi_wk_vt chi0, array4_view U)" in text
    assert "gw_sigma(chi_wk_cvt W_wk, g_wk_cvt g_wk)" in text
    assert "gw_dynamic_sigma(chi_tr_cvt W_tr, g_tr_cvt g_tr)" in text
    assert "eliashberg_product_fft(chi_tr_vt Gamma_pp_dyn_tr, chi_r_vt Gamma_pp_const_r, g_wk_vt g_wk," in text
    assert "construct_phi_wk(chi_wk_vt chi, array4_view U)" in text


@pytest.mark.parametrize("prog", ["chain_cpp", "consumers_cpp", "dyson_cpp"])
def test_header_compiles_and_links_with_host_compiler(prog):
    lib = os.path.join(ROOT, "tprf_amd", "libtprf_hip.so")
    assert os.path.exists(lib), "build first"
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, prog)
        cmd = ["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", prog + ".cpp"), "-o", out,
               "-L" + os.path.join(ROOT, "tprf_amd"), "-ltprf_hip", "-L/opt/rocm/lib", "-lamdhip64"]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        assert os.path.getsize(out) > 0


def _case():
    from tprf_amd import workloads
    wl = workloads.cubic_3orb(seed=2, L=4, beta=10.0, n_iw=64, nb=8)
    return wl


@pytest.mark.gpu
def test_cpp_caller_chain_matches_oracle(oracle):
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    assert os.path.exists(BIN), "build first: tests/cpp/chain_cpp"
    wl = _case()
    n, nk, nnu = wl.norb, wl.nk, wl.n_nu
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array(list(wl.dims) + [n, wl.n_iw, wl.nb, 3, 0], dtype=np.int32).tofile(f)
            np.array([wl.beta, wl.mu], dtype=np.float64).tofile(f)
            np.ascontiguousarray(wl.e_k, dtype=np.complex128).tofile(f)
            np.ascontiguousarray(wl.U, dtype=np.complex128).tofile(f)
        print("running tests/cpp/chain_cpp (3 orbitals, 4^3 k, 64 freqs, nb 8)", flush=True)
        r = subprocess.run([BIN, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        size = nnu * nk * n ** 4
        res = np.fromfile(fout, dtype=np.complex128).reshape(5, nnu, nk, n, n, n, n)
    g_tr = oracle.fourier_w_to_t(oracle.fourier_lattice(oracle.lattice_dyson_g_wk(wl.beta, wl.n_iw, wl.mu, wl.e_k),
                                                        wl.dims, -1), wl.beta)
    chi0 = oracle.fourier_lattice(oracle.chi0_wr_from_grt_PH(g_tr, wl.dims, wl.beta, wl.nb), wl.dims, +1)
    chi0_stored = oracle.fourier_lattice(oracle.fourier_t_to_w(oracle.chi0_tr_from_grt_PH(g_tr, wl.dims), wl.beta,
                                                               wl.nb), wl.dims, +1)
    chi = oracle.solve_rpa_PH(chi0, wl.U)
    assert relerr(res[0], chi0) <= 1e-10
    assert relerr(res[1], chi0_stored) <= 1e-10
    for i in (2, 3, 4):
        assert rpa_residual(chi0, wl.U, res[i]).max() <= 1e-10, i
    kappa_ok = relerr(res[2], chi) <= 1e-10
    print(f"C++ caller: chi0 {relerr(res[0], chi0):.2e}, stored {relerr(res[1], chi0_stored):.2e}, chi "
          f"{relerr(res[2], chi):.2e}, batch {relerr(res[3], chi):.2e} / {relerr(res[4], chi):.2e}", flush=True)
    assert kappa_ok
    _ = size


@pytest.mark.gpu
def test_cpp_caller_gw_and_eliashberg_match_oracle(oracle):
    """tests/cpp/consumers_cpp (gw_sigma of W and of a static v, construct_phi_wk,
    the summation and FFT Eliashberg products through the C++ signatures)
    against the oracle: 1e-11 (gw_sigma, FFT product), 1e-13 (the pointwise
    ones); the summation and FFT products agree to 1e-9."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    binp = os.path.join(ROOT, "tests", "cpp", "consumers_cpp")
    assert os.path.exists(binp), "build first: tests/cpp/consumers_cpp"
    from test_oracle_consumers import _eliashberg_inputs
    beta, nw, dims, g, gamma, delta = _eliashberg_inputs(3)
    n, nk = 2, int(np.prod(dims))
    rng = np.random.default_rng(8)
    inu = 1j * 2 * np.pi * np.arange(-(nw - 1), nw) / beta
    V = (rng.normal(size=(nk, n, n, n, n)) + 1j * rng.normal(size=(nk, n, n, n, n))) * 0.2
    W = V[None] + 0.3 * (2.0 / (inu ** 2 - 1.0))[:, None, None, None, None, None] * np.ones_like(V)[None]
    U = rng.normal(size=(n, n, n, n)) + 0j
    nbig = 2 * nw + 1
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array(list(dims) + [n, nw, nbig, 0, 0], dtype=np.int32).tofile(f)
            np.array([beta], dtype=np.float64).tofile(f)
            for a in (g, W, V, delta, gamma(nw), gamma(nbig), U):
                np.ascontiguousarray(a, dtype=np.complex128).tofile(f)
        print("running tests/cpp/consumers_cpp (2 orbitals, 3 k, n_max 30)", flush=True)
        r = subprocess.run([binp, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out = np.fromfile(fout, dtype=np.complex128)
    sizes = [g.size, nk * n * n, gamma(nw).size, delta.size, delta.size]
    parts, o = [], 0
    for s_ in sizes:
        parts.append(out[o:o + s_])
        o += s_
    assert o == out.size
    sig = parts[0].reshape(g.shape)
    assert relerr(sig, oracle.gw_sigma(W, g, beta, dims)) <= 1e-11
    sk = parts[1].reshape(nk, n, n)
    assert relerr(sk, oracle.static_sigma_k(V, oracle.density(g, beta), dims, fock=True)) <= 1e-12
    phi = parts[2].reshape(gamma(nw).shape)
    assert relerr(phi, oracle.construct_phi_wk(gamma(nw), U)) <= 1e-13
    ds = parts[3].reshape(delta.shape)
    assert relerr(ds, oracle.eliashberg_product(gamma(nbig), g, delta, beta, dims)) <= 1e-13
    df = parts[4].reshape(delta.shape)
    odyn, oconst = oracle.preprocess_gamma_for_fft(gamma(nw), beta, dims)
    assert relerr(df, oracle.eliashberg_product_fft(odyn, oconst, g, delta, beta, dims)) <= 1e-11
    assert np.abs(ds - df).max() <= 1e-9


@pytest.mark.gpu
def test_cpp_caller_real_frequency_dyson_matches_oracle(oracle):
    """tests/cpp/dyson_cpp (lattice_dyson_g0_fk, both lattice_dyson_g_fk
    overloads, lattice_dyson_g_f, lattice_dyson_g_w and the fk <-> fr round
    trip through the C++ signatures) against the oracle, 1e-13."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    binp = os.path.join(ROOT, "tests", "cpp", "dyson_cpp")
    assert os.path.exists(binp), "build first: tests/cpp/dyson_cpp"
    rng = np.random.default_rng(12)
    dims, n, nf, n_max = (3, 2, 1), 2, 21, 8
    wmin, wmax, delta, beta, mu = -4.0, 4.0, 0.02, 6.0, 0.15
    nk = int(np.prod(dims))
    a = rng.normal(size=(nk, n, n)) + 1j * rng.normal(size=(nk, n, n))
    e = 0.5 * (a + np.conj(np.swapaxes(a, 1, 2)))
    s_fk = (rng.normal(size=(nf, nk, n, n)) + 1j * rng.normal(size=(nf, nk, n, n))) * 0.2
    s_f = (rng.normal(size=(nf, n, n)) + 1j * rng.normal(size=(nf, n, n))) * 0.2
    s_w = (rng.normal(size=(2 * n_max, n, n)) + 1j * rng.normal(size=(2 * n_max, n, n))) * 0.2
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
        with open(fin, "wb") as f:
            np.array(list(dims) + [n, nf, n_max, 0, 0], dtype=np.int32).tofile(f)
            np.array([wmin, wmax, delta, beta, mu], dtype=np.float64).tofile(f)
            for x in (e, s_fk, s_f, s_w):
                np.ascontiguousarray(x, dtype=np.complex128).tofile(f)
        print(f"running tests/cpp/dyson_cpp ({n} orbitals, {nk} k, {nf} real frequencies)", flush=True)
        r = subprocess.run([binp, fin, fout], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        out = np.fromfile(fout, dtype=np.complex128)
    z = np.linspace(wmin, wmax, nf) + 1j * delta
    iw = 1j * np.pi * (2 * np.arange(-n_max, n_max) + 1) / beta
    refs = [oracle.lattice_dyson_g_zk(z, mu, e), oracle.lattice_dyson_g_zk(z, mu, e, s_fk),
            oracle.lattice_dyson_g_zk(z, mu, e, s_f), oracle.lattice_dyson_g_z(z, mu, e, s_f),
            oracle.lattice_dyson_g_z(iw, mu, e, s_w)]
    refs.append(refs[0])
    off = 0
    for i, ref in enumerate(refs):
        got = out[off:off + ref.size].reshape(ref.shape)
        off += ref.size
        assert relerr(got, ref) <= 1e-13, i
    assert off == out.size
