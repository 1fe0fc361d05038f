"""Drop-in for the GW entry points of ``triqs_tprf.gw`` (python/triqs_tprf/gw.py,
which re-exports the C++ functions of c++/triqs_tprf/lattice/gw.hpp):
the self-energy pieces run on the GPU through the C ABI (see ``lattice``)."""
from .lattice import (dynamical_screened_interaction_W, fock_sigma, gw_dynamic_sigma, gw_sigma, hartree_sigma,
                      rho_k_from_g_wk, split_into_dynamic_wk_and_constant_k)

__all__ = ["dynamical_screened_interaction_W", "gw_sigma", "gw_dynamic_sigma", "hartree_sigma", "fock_sigma",
           "rho_k_from_g_wk", "split_into_dynamic_wk_and_constant_k"]
