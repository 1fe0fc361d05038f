"""Driver of the imaginary-time bubble (python/triqs_tprf/lattice_utils.py:191-269)
and the model builders the workloads use."""
import numpy as np

from . import lattice as _L
from .gf import Gf, MeshBrZone


def imtime_bubble_chi0_wk(g_wk, nw=1, save_memory=False, verbose=False):
    """chi0(i nu, q) from G(i w, k) through the tau route, in the reference's stage
    order and branches (lattice_utils.py:241-266): k->r, w->tau, then the static
    trapezoid (nw == 1), the stored product + tau->nu, or the fused product
    (save_memory), and finally r->q."""
    g_wr = _L.fourier_wk_to_wr(g_wk)
    g_tr = _L.fourier_wr_to_tr(g_wr)
    del g_wr
    if nw == 1:
        chi0_wr = _L.chi0_w0r_from_grt_PH(g_tr)
    elif not save_memory:
        chi0_tr = _L.chi0_tr_from_grt_PH(g_tr)
        del g_tr
        chi0_wr = _L.chi_wr_from_chi_tr(chi0_tr, nw=nw)
        del chi0_tr
    else:
        chi0_wr = _L.chi0_wr_from_grt_PH(g_tr, nw=nw)
    return _L.chi_wk_from_chi_wr(chi0_wr)


def tb_dispersion(hoppings, dims, orbital_count):
    """e(k) = sum_R T_R exp(i k.R) on the lattice mesh ``dims``.

    ``hoppings`` maps integer displacement tuples R to n x n matrices (the
    TBLattice hopping dictionary of python/triqs_tprf/tight_binding.py; the
    phase convention exp(+i k.R) is the one of TRIQS TBLattice.fourier).
    Returns a Gf on MeshBrZone(dims) with numpy data.
    """
    mesh = MeshBrZone(dims)
    k = mesh.values()
    n = orbital_count
    e = np.zeros((mesh.size, n, n), dtype=np.complex128)
    for R, T in hoppings.items():
        R = tuple(R) + (0,) * (3 - len(R))
        ph = np.exp(1j * (k @ np.asarray(R, dtype=float)))
        e += ph[:, None, None] * np.asarray(T, dtype=np.complex128)[None]
    return Gf(mesh, data=e, device="cpu")
