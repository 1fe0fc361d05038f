"""Drop-in for the Eliashberg helpers of ``triqs_tprf.eliashberg`` that feed
the products (python/triqs_tprf/eliashberg.py): the products themselves are
in ``lattice`` (eliashberg.hpp)."""
from .gf import Gf
from .lattice import (construct_phi_wk, dynamic_and_constant_to_tr, eliashberg_constant_gamma_f_product,
                      eliashberg_dynamic_gamma_f_product, eliashberg_g_delta_g_product, eliashberg_product,
                      eliashberg_product_fft, eliashberg_product_fft_constant, split_into_dynamic_wk_and_constant_k)

__all__ = ["preprocess_gamma_for_fft", "construct_phi_wk", "dynamic_and_constant_to_tr", "eliashberg_product",
           "eliashberg_product_fft", "eliashberg_product_fft_constant", "eliashberg_g_delta_g_product",
           "eliashberg_dynamic_gamma_f_product", "eliashberg_constant_gamma_f_product"]


def preprocess_gamma_for_fft(Gamma_pp_wk, Gamma_pp_const_k=None):
    """(Gamma_dyn(tau, r), Gamma_const(r)) for eliashberg_product_fft
    (eliashberg.py:196-270): the constant part is fitted from the tail when not
    given (a number, an array over the target, or a Gf on the k mesh), and
    the dynamic part Gamma - Gamma_const goes to (tau, r)."""
    import torch
    if Gamma_pp_const_k is None:
        dyn_wk, const_k = split_into_dynamic_wk_and_constant_k(Gamma_pp_wk)
        return dynamic_and_constant_to_tr(dyn_wk, const_k)
    kmesh = Gamma_pp_wk.mesh[1]
    x = Gamma_pp_wk.data
    if isinstance(Gamma_pp_const_k, Gf):
        c = Gamma_pp_const_k.data.to(x.device, torch.complex128)
    else:
        c = torch.as_tensor(Gamma_pp_const_k, dtype=torch.complex128, device=x.device)
        c = torch.broadcast_to(c, tuple(x.shape[1:])).contiguous()
    const_k = Gf(kmesh, data=c)
    dyn_wk = Gf(Gamma_pp_wk.mesh, data=(x - c[None]).contiguous())
    return dynamic_and_constant_to_tr(dyn_wk, const_k)
