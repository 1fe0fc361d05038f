"""Writes the golden fixtures of this directory from the oracle (run from the repo root):

    python tests/golden/make_golden.py

chain_square8.npz: the plain path on an 8x8 square lattice (e_k = -2(cos kx + cos ky),
mu = 0.1, beta = 10, 32 positive fermionic frequencies, ntau = 193, 16 bosonic
frequencies, onsite U = 1.0): G(tau, r) at every 8th tau (g_tr8 = g_tr[::8], both
endpoints included), chi0(i nu, q) and chi_RPA(i nu, q).
chain_cubic2orb.npz: a 2-orbital 4^3 cubic lattice (hoppings of the cubic workload,
seed 3; beta 8, 32 freqs, 16 bosonic freqs, Kanamori U = 1.2, J = 0.2).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def chain(O, e_k, dims, beta, n_iw, nb, mu, U):
    g_wk = O.lattice_dyson_g_wk(beta, n_iw, mu, e_k)
    g_wr = O.fourier_lattice(g_wk, dims, -1)
    g_tr = O.fourier_w_to_t(g_wr, beta)
    chi0_wr = O.chi0_wr_from_grt_PH(g_tr, dims, beta, nb)
    chi0_wk = O.fourier_lattice(chi0_wr, dims, +1)
    chi_wk = O.solve_rpa_PH(chi0_wk, U)
    return g_tr, chi0_wk, chi_wk


def main():
    from oracle import oracle as O
    from tprf_amd import workloads
    from tprf_amd.rpa_tensor import kanamori_charge_and_spin_quartic_interaction_tensors
    kx, ky, _ = workloads.kgrid((8, 8, 1))
    e = (-2 * (np.cos(kx) + np.cos(ky))).astype(np.complex128).reshape(-1, 1, 1)
    U = np.full((1, 1, 1, 1), 1.0, dtype=np.complex128)
    g_tr, chi0, chi = chain(O, e, (8, 8, 1), 10.0, 32, 16, 0.1, U)
    np.savez_compressed(os.path.join(HERE, "chain_square8.npz"), e_k=e, U=U, g_tr8=g_tr[::8], chi0_wk=chi0, chi_wk=chi,
                        beta=10.0, n_iw=32, nb=16, mu=0.1, dims=np.array([8, 8, 1]))
    T = workloads.cubic_hoppings(3, 2)
    k = workloads.kgrid((4, 4, 4))
    e = np.zeros((64, 2, 2), dtype=np.complex128)
    for d in range(3):
        e += 2 * np.cos(k[d])[:, None, None] * T[d][None]
    _, Us = kanamori_charge_and_spin_quartic_interaction_tensors(2, 1.2, 1.2 - 0.4, 0.2, 0.2)
    g_tr, chi0, chi = chain(O, e, (4, 4, 4), 8.0, 32, 16, 0.0, Us)
    np.savez_compressed(os.path.join(HERE, "chain_cubic2orb.npz"), e_k=e, U=Us, g_tr8=g_tr[::8], chi0_wk=chi0,
                        chi_wk=chi, beta=8.0, n_iw=32, nb=16, mu=0.0, dims=np.array([4, 4, 4]))


if __name__ == "__main__":
    main()
