"""Seeded synthetic inputs of the bubble + RPA workloads.

plain_square:  single-band 2-D square lattice, e_k = -2t(cos kx + cos ky), t = 1,
               mu = 0, beta = 10, 32x32 k, n_iw = 64, nb = 64, onsite U = 1.0
               (1x1 spin-channel tensor).
cubic_3orb:    3-orbital simple-cubic lattice, 16^3 k, beta = 40, n_iw = 256,
               nb = 64; per direction d a complex Hermitian 3x3 hopping
               T_d = (A_d + A_d^H)/2 - 0.26 * 1 with A_d entries N(0, 0.25^2)
               (real and imaginary parts) from the seed, e(k) = sum_d 2 cos(k_d) T_d;
               mu by bisection to filling 1/3; Kanamori spin-channel U with
               U = 2.3, J = 0.4, U' = U - 2J, J' = J (9x9 solves).
"""
import dataclasses

import numpy as np

from .rpa_tensor import kanamori_charge_and_spin_quartic_interaction_tensors


@dataclasses.dataclass
class Workload:
    name: str
    dims: tuple
    norb: int
    beta: float
    n_iw: int
    nb: int
    mu: float
    e_k: np.ndarray  # [nk, n, n] complex128
    U: np.ndarray    # [n, n, n, n] complex128

    @property
    def nk(self):
        return int(np.prod(self.dims))

    @property
    def ntau(self):
        return 6 * self.n_iw + 1

    @property
    def n_nu(self):
        return 2 * self.nb - 1


def kgrid(dims):
    g = np.meshgrid(*[2 * np.pi * np.arange(d) / d for d in dims], indexing="ij")
    return [x.ravel() for x in g]


def fermi(x):
    """1/(e^x + 1), evaluated without overflow."""
    out = np.empty_like(x)
    neg = x < 0
    out[neg] = 1.0 / (np.exp(x[neg]) + 1.0)
    e = np.exp(-x[~neg])
    out[~neg] = e / (1.0 + e)
    return out


def mu_for_filling(e_k, beta, filling, tol=1e-13):
    """Bisection on mu so that (1/(n N_k)) sum_{k,i} f(beta (e_ki - mu)) = filling."""
    ev = np.linalg.eigvalsh(e_k).ravel()
    lo, hi = ev.min() - 10.0, ev.max() + 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if fermi(beta * (ev - mid)).mean() > filling:
            hi = mid
        else:
            lo = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)


def plain_square(n=32, beta=10.0, n_iw=64, nb=64, U=1.0, t=1.0, mu=0.0):
    kx, ky, _ = kgrid((n, n, 1))
    e = (-2 * t * (np.cos(kx) + np.cos(ky))).astype(np.complex128).reshape(-1, 1, 1)
    Um = np.full((1, 1, 1, 1), U, dtype=np.complex128)
    return Workload("plain_square", (n, n, 1), 1, beta, n_iw, nb, mu, e, Um)


def cubic_hoppings(seed, norb=3, sigma=0.25, diag=-0.26):
    rng = np.random.default_rng(seed)
    T = []
    for _ in range(3):
        A = rng.normal(0.0, sigma, (norb, norb)) + 1j * rng.normal(0.0, sigma, (norb, norb))
        T.append(0.5 * (A + A.conj().T) + diag * np.eye(norb))
    return T


def cubic_3orb(seed=0, L=16, beta=40.0, n_iw=256, nb=64, U=2.3, J=0.4, norb=3):
    T = cubic_hoppings(seed, norb)
    k = kgrid((L, L, L))
    e = np.zeros((L ** 3, norb, norb), dtype=np.complex128)
    for d in range(3):
        e += 2 * np.cos(k[d])[:, None, None] * T[d][None]
    mu = mu_for_filling(e, beta, 1.0 / 3.0)
    _, U_s = kanamori_charge_and_spin_quartic_interaction_tensors(norb, U, U - 2 * J, J, J)
    return Workload("cubic_3orb", (L, L, L), norb, beta, n_iw, nb, mu, e, U_s)


def uneven_batch(seed=0, count=48, nb=64, U=2.3, J=0.4):
    """48 uneven lattices drawn from the seed: simple-cubic L^3 k with
    L in {8, 12, 16}, norb in {1, 2, 3}, n_iw in {128, 256, 512}, beta in
    {10, 20, 40}, mu ~ U[-1, 1]; hoppings as in cubic_3orb (per direction
    (A + A^H)/2 - 0.26 * 1, A entries N(0, 0.25^2)); nb bosonic frequencies;
    Kanamori spin-channel U (U, U' = U - 2J, J, J' = J; just U for one orbital)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        L = int(rng.choice([8, 12, 16]))
        norb = int(rng.choice([1, 2, 3]))
        n_iw = int(rng.choice([128, 256, 512]))
        beta = float(rng.choice([10.0, 20.0, 40.0]))
        mu = float(rng.uniform(-1.0, 1.0))
        T = cubic_hoppings(int(rng.integers(1 << 30)), norb)
        k = kgrid((L, L, L))
        e = np.zeros((L ** 3, norb, norb), dtype=np.complex128)
        for d in range(3):
            e += 2 * np.cos(k[d])[:, None, None] * T[d][None]
        _, U_s = kanamori_charge_and_spin_quartic_interaction_tensors(norb, U, U - 2 * J, J, J)
        out.append(Workload(f"uneven_{i}", (L, L, L), norb, beta, n_iw, nb, mu, e, U_s))
    return out


WORKLOADS = {"plain_square": plain_square, "cubic_3orb": cubic_3orb}
