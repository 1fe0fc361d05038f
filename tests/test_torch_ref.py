"""CPU pins of the torch FP64 reference of chi_RPA(Q, 0) and of its autograd gradient.

The reference (oracle/torch_ref.py) must equal the oracle's own chain
(Dyson -> lattice FFT -> omega->tau -> nu = 0 bubble -> one momentum -> RPA)
and its autograd gradient must equal central finite differences of itself.
"""
import numpy as np
import pytest
import torch

from conftest import relerr

pytestmark = pytest.mark.filterwarnings("ignore::UserWarning")


def _e_k(t, tp, dims):
    d0, d1 = dims[0], dims[1]
    kx = 2 * torch.pi * torch.arange(d0, dtype=torch.float64) / d0
    ky = 2 * torch.pi * torch.arange(d1, dtype=torch.float64) / d1
    cx, cy = torch.cos(kx)[:, None], torch.cos(ky)[None, :]
    return (-2 * t * (cx + cy) - 4 * tp * cx * cy).reshape(-1, 1, 1).to(torch.complex128)


def _oracle_chain(O, e_k, mu, U, beta, n_iw, dims, q):
    g_wk = O.lattice_dyson_g_wk(beta, n_iw, mu, e_k)
    g_wr = O.fourier_lattice(g_wk, dims, -1)
    g_tr = O.fourier_w_to_t(g_wr, beta, O.FERMION)
    chi0_wr = O.chi0_wr_from_grt_PH(g_tr, dims, beta, 1)
    chi0_wk = O.fourier_lattice(chi0_wr, dims, +1)
    d = tuple(dims) + (1,) * (3 - len(dims))
    iq = (q[0] * d[1] + q[1]) * d[2] + (q[2] if len(q) > 2 else 0)
    chi0_q = chi0_wk[0, iq][None]
    return O.solve_rpa_PH(chi0_q, U)[0]


@pytest.fixture(scope="module")
def ref_small(oracle):
    from oracle.torch_ref import StaticChainRef
    return StaticChainRef(beta=5.0, n_iw=32, dims=(8, 8), q=(4, 4))


def test_reference_forward_equals_oracle_chain(oracle, ref_small):
    dims, q = (8, 8, 1), (4, 4, 0)
    t, tp, mu, U = 1.0, -0.3, -0.5, 2.0
    e_k = _e_k(torch.tensor(t), torch.tensor(tp), dims)
    Ut = torch.full((1, 1, 1, 1), U, dtype=torch.complex128)
    got = ref_small.chi(e_k, torch.tensor(mu, dtype=torch.float64), Ut).numpy()
    want = _oracle_chain(oracle, e_k.numpy(), mu, Ut.numpy(), 5.0, 32, dims, q)
    assert relerr(got, want) < 1e-12


def test_reference_forward_two_orbital(oracle):
    from oracle.torch_ref import StaticChainRef
    dims, q = (6, 6, 1), (3, 0, 0)
    rng = np.random.default_rng(3)
    A = rng.normal(0, 0.3, (2, 2)) + 1j * rng.normal(0, 0.3, (2, 2))
    h = (A + A.conj().T) / 2
    kx, ky = np.meshgrid(2 * np.pi * np.arange(6) / 6, 2 * np.pi * np.arange(6) / 6, indexing="ij")
    e_k = (-2 * (np.cos(kx) + np.cos(ky))).ravel()[:, None, None] * np.eye(2) + \
        (np.cos(kx) - np.cos(ky)).ravel()[:, None, None] * h
    U = (rng.normal(0, 0.2, (2, 2, 2, 2)) + 0j)
    ref = StaticChainRef(beta=4.0, n_iw=32, dims=dims, q=q)
    got = ref.chi(torch.from_numpy(e_k), torch.tensor(0.2, dtype=torch.float64), torch.from_numpy(U)).numpy()
    want = _oracle_chain(oracle, e_k, 0.2, U, 4.0, 32, dims, q)
    assert relerr(got, want) < 1e-12


def test_reference_gradient_matches_finite_differences(ref_small):
    dims = (8, 8, 1)
    x0 = {"t": 1.0, "tp": -0.3, "mu": -0.5, "U": 2.0}

    def L(v):
        e_k = _e_k(v["t"], v["tp"], dims)
        Ut = v["U"].reshape(1, 1, 1, 1).to(torch.complex128)
        return ref_small.chi(e_k, v["mu"], Ut)[0, 0, 0, 0].real

    leaves = {k: torch.tensor(x, dtype=torch.float64, requires_grad=True) for k, x in x0.items()}
    L(leaves).backward()
    h = 1e-6
    for k in x0:
        with torch.no_grad():
            vp = {kk: torch.tensor(x0[kk] + (h if kk == k else 0.0), dtype=torch.float64) for kk in x0}
            vm = {kk: torch.tensor(x0[kk] - (h if kk == k else 0.0), dtype=torch.float64) for kk in x0}
            fd = float((L(vp) - L(vm)) / (2 * h))
        g = float(leaves[k].grad)
        assert abs(g - fd) <= 1e-6 * max(1.0, abs(fd)), (k, g, fd)
