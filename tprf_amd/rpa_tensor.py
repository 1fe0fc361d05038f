"""Interaction tensors for solve_rpa_PH, in the reference's vertex order
[a, abar, b, bbar] (python/triqs_tprf/rpa_tensor.py)."""
import itertools

import numpy as np


def kanamori_charge_and_spin_quartic_interaction_tensors(norb, U, Up, J, Jp):
    """Kanamori charge/spin tensors, rpa_tensor.py:87-120:
    all four indices equal -> U; a = bbar != b = abar -> (-U'+2J, U');
    a = abar != b = bbar -> (2U'-J, J); a = b != abar = bbar -> (J', J')."""
    shape = (norb,) * 4
    U_c, U_s = np.zeros(shape, dtype=complex), np.zeros(shape, dtype=complex)
    for a, abar, b, bbar in itertools.product(range(norb), repeat=4):
        sc = ss = 0.0
        if a == abar == b == bbar:
            sc = ss = U
        if a == bbar and a != b and abar == b:
            sc, ss = -Up + 2 * J, Up
        if a == abar and a != b and b == bbar:
            sc, ss = 2 * Up - J, J
        if a == b and a != abar and abar == bbar:
            sc = ss = Jp
        U_c[a, abar, b, bbar] = sc
        U_s[a, abar, b, bbar] = ss
    return U_c, U_s


def spinful_hubbard_tensor(U):
    """Single-band H = U n_up n_dn as the 4x4x4x4 tensor get_rpa_tensor builds
    (rpa_tensor.py:58-70, spin the only index): U[0,0,1,1] = U[1,1,0,0] = -U,
    U[0,1,1,0] = U[1,0,0,1] = +U."""
    T = np.zeros((2, 2, 2, 2), dtype=complex)
    T[0, 0, 1, 1] = T[1, 1, 0, 0] = -U
    T[0, 1, 1, 0] = T[1, 0, 0, 1] = U
    return T
