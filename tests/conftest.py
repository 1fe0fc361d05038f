import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X) and libtprf_hip.so")


def relerr(got, ref):
    """Relative max-norm ||got - ref||_inf / ||ref||_inf."""
    got, ref = np.asarray(got), np.asarray(ref)
    assert got.shape == ref.shape, (got.shape, ref.shape)
    den = np.abs(ref).max()
    return float(np.abs(got - ref).max() / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


def rpa_matrices(x, n):
    """[..., n, n, n, n] -> [npts, n^2, n^2] in the PH grouping M[(a,b),(c,d)] =
    X_abcd, the memory-order grouping of rpa.cpp:40,54-55."""
    return np.asarray(x).reshape(-1, n * n, n * n)


def rpa_residual(chi0, U, chi, floor=1e-3):
    """Per (nu, q) point the relative residual of (1 - chi0 U) chi = chi0
    (rpa.cpp:57): ||M_chi - M_chi0 - M_chi0 M_U M_chi||_F /
    (s_p (1 + ||M_U||_F ||M_chi||_F)) with s_p = max(||M_chi0||_F, floor x the
    largest ||M_chi0||_F).  Independent of the condition of 1 - chi0 U: a
    backward-stable solve gives ~1e-15 at every point.  The floor: the lattice
    and Matsubara transforms leave every chi0 entry with an absolute rounding
    error of ~1e-16 of the largest entry, so points whose chi0 is many orders
    below the largest (ones that vanish by symmetry) carry no relative
    accuracy of their own."""
    n = np.asarray(U).shape[0]
    X, Y, MU = rpa_matrices(chi0, n), rpa_matrices(chi, n), rpa_matrices(U, n)[0]
    R = Y - X - X @ (MU @ Y)
    num = np.linalg.norm(R, axis=(1, 2))
    nx = np.linalg.norm(X, axis=(1, 2))
    s = np.maximum(nx, floor * nx.max())
    den = s * (1.0 + np.linalg.norm(MU) * np.linalg.norm(Y, axis=(1, 2)))
    return num / np.where(den > 0, den, 1.0)


def rpa_kappa(chi0, U):
    """Per point the 2-norm condition number of 1 - M_chi0 M_U."""
    n = np.asarray(U).shape[0]
    X, MU = rpa_matrices(chi0, n), rpa_matrices(U, n)[0]
    s = np.linalg.svd(np.eye(n * n)[None] - X @ MU, compute_uv=False)
    return s[:, 0] / s[:, -1]


def mirror_nu(x, n):
    """S M(nu)^H S on every point of x[n_nu, nk, n, n, n, n], frequency-reversed:
    out_abcd(-nu, q) = conj(x_dcba(nu, q)).  The identity chi0 and chi_RPA obey
    for a Hermitian e_k and U_abcd = conj(U_dcba)."""
    return np.conj(np.asarray(x)[::-1].transpose(0, 1, 5, 4, 3, 2))
