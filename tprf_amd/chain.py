"""The forward chain e_k -> G(w,k) -> G(w,r) -> G(tau,r) -> chi0(nu,r) -> chi0(nu,q)
-> chi_RPA(nu,q) on preallocated device buffers (one "eval").

Stages and the reference functions they replace:
  lattice_dyson_g_wk (gf.cpp:126), fourier_wk_to_wr (gf.cpp:209),
  fourier_wr_to_tr (gf.cpp:262), chi0_wr_from_grt_PH (chi_imtime.cpp:139, the
  fused product + tau->nu of chi_imtime.cpp:279), chi_wk_from_chi_wr
  (chi_imtime.cpp:289), solve_rpa_PH (rpa.cpp:66).
Every stage is one asynchronous C-ABI call on torch's current stream; the
tail-fit status is read once, after the chain.
"""
import numpy as np
import torch

from . import _lib


class BubbleRPA:
    """fused=True (default): the first two stages run as one call,
    tprf_lattice_dyson_g_wr (G(iw, k) is then not stored for a 16^3 lattice),
    and, when e_k is Hermitian (so G_ab(tau, r)^* = G_ba(tau, -r)), omega -> tau
    runs as tprf_fourier_wr_to_tr_herm (columns (r, a, b) with r < -r, or r = -r
    and a <= b, transformed; G(tau, r) is then stored for r <= -r only)
    and the bubble as tprf_chi0_wr_from_grt_PH_herm (orbital pairs a <= b
    transformed, the others formed from the symmetry); if also U_abcd =
    conj(U_dcba) and U_abcd = U_badc, chi0_abcd(-nu, q) = conj(chi0_dcba(nu, q))
    and the same holds for chi_RPA, so
    the bubble writes and r -> q transforms the frequencies nu >= 0 only and
    tprf_solve_rpa_PH_herm solves those and writes chi(-nu) as the conjugate
    transposes (chi0_wr and chi0_wk then hold nu >= 0 only).  fused=False: one
    call per reference function."""
    API_STAGES = ("dyson", "wk_to_wr", "wr_to_tr", "chi0_wr", "wr_to_wk", "rpa")
    FUSED_STAGES = ("dyson_wr", "wr_to_tr", "chi0_wr", "wr_to_wk", "rpa")

    def __init__(self, wl, device="cuda", save_intermediates=True, fused=True):
        self.wl = wl
        self.STAGES = self.FUSED_STAGES if fused else self.API_STAGES
        e = wl.e_k
        self.herm = bool(fused) and bool(np.abs(e - np.conj(np.swapaxes(e, 1, 2))).max() <= 1e-14 * max(1.0, np.abs(e).max()))
        u = wl.U
        tol_u = 1e-14 * max(1.0, np.abs(u).max())
        self.herm_nu = (self.herm and bool(np.abs(u - np.conj(np.einsum("abcd->dcba", u))).max() <= tol_u)
                        and bool(np.abs(u - np.einsum("abcd->badc", u)).max() <= tol_u))
        dev = torch.device(device)
        self.dev = dev
        n, nk = wl.norb, wl.nk
        c128 = torch.complex128
        self.e_k = torch.from_numpy(wl.e_k).to(dev).contiguous()
        self.U = torch.from_numpy(wl.U).to(dev).contiguous()
        # G(iw, k) is stored only by the one-call-per-function chain
        self.g_wk = None if fused else torch.empty((2 * wl.n_iw, nk, n, n), dtype=c128, device=dev)
        self.g_wr = torch.empty((2 * wl.n_iw, nk, n, n), dtype=c128, device=dev)
        self.g_tr = torch.empty((wl.ntau, nk, n, n), dtype=c128, device=dev)
        self.chi0_wr = torch.empty((wl.n_nu, nk, n, n, n, n), dtype=c128, device=dev)
        self.chi0_wk = torch.empty_like(self.chi0_wr)
        self.chi_wk = torch.empty_like(self.chi0_wr)
        self.dims = _lib.dims3(wl.dims)
        self.ctx = _lib.ctx(dev.index if dev.index is not None else torch.cuda.current_device())

    def stage(self, name):
        wl, c, p = self.wl, self.ctx, _lib.ptr
        n, nk = wl.norb, wl.nk
        if name == "dyson":
            _lib.call("tprf_lattice_dyson_g_wk", c, wl.beta, wl.n_iw, wl.mu, nk, n, p(self.e_k), _lib.SIGMA_NONE, None,
                      p(self.g_wk))
        elif name == "dyson_wr":
            if self.herm:  # only r <= -r is written: the symmetric omega -> tau reads no other column
                _lib.call("tprf_lattice_dyson_g_wr_herm", c, wl.beta, wl.n_iw, wl.mu, self.dims, n, p(self.e_k),
                          p(self.g_wr), 0)
            else:
                _lib.call("tprf_lattice_dyson_g_wr", c, wl.beta, wl.n_iw, wl.mu, self.dims, n, p(self.e_k),
                          p(self.g_wr))
        elif name == "wk_to_wr":
            _lib.call("tprf_fourier_wk_to_wr", c, 2 * wl.n_iw, self.dims, n * n, p(self.g_wk), p(self.g_wr))
        elif name == "wr_to_tr":
            if self.herm:  # only r <= -r is written: the Hermitian bubble reads no other column
                _lib.call("tprf_fourier_wr_to_tr_herm", c, wl.beta, wl.n_iw, wl.ntau, self.dims, n, p(self.g_wr),
                          p(self.g_tr), 0)
            else:
                _lib.call("tprf_fourier_wr_to_tr", c, _lib.FERMION, wl.beta, wl.n_iw, wl.ntau, nk * n * n,
                          p(self.g_wr), p(self.g_tr), None)
        elif name == "chi0_wr":
            if self.herm:  # with herm_nu only nu >= 0 is written: r -> q and RPA read no other rows
                _lib.call("tprf_chi0_wr_from_grt_PH_herm", c, wl.beta, wl.ntau, self.dims, n, wl.nb, p(self.g_tr),
                          p(self.chi0_wr), 0 if self.herm_nu else 1)
            else:
                _lib.call("tprf_chi0_wr_from_grt_PH", c, wl.beta, wl.ntau, self.dims, n, wl.nb, p(self.g_tr),
                          p(self.chi0_wr))
        elif name == "wr_to_wk":
            if self.herm_nu:  # nu >= 0 only: rows nb - 1 .. 2 nb - 2
                _lib.call("tprf_fourier_wr_to_wk", c, wl.nb, self.dims, n ** 4, p(self.chi0_wr[wl.nb - 1:]),
                          p(self.chi0_wk[wl.nb - 1:]))
            else:
                _lib.call("tprf_fourier_wr_to_wk", c, wl.n_nu, self.dims, n ** 4, p(self.chi0_wr), p(self.chi0_wk))
        elif name == "rpa":
            if self.herm_nu:
                _lib.call("tprf_solve_rpa_PH_herm", c, wl.nb, nk, n, p(self.chi0_wk), p(self.U), p(self.chi_wk))
            else:
                _lib.call("tprf_solve_rpa_PH", c, wl.n_nu * nk, n, p(self.chi0_wk), p(self.U), p(self.chi_wk))
        else:
            raise KeyError(name)

    def stage_reset(self):
        """Clear the tail-fit / singularity status words (one small memset)."""
        _lib.tail_status_reset(self.ctx)

    def run(self):
        self.stage_reset()
        for s in self.STAGES:
            self.stage(s)

    def check_status(self):
        """Synchronise and return (max tail-fit residual, max |m0|); raises on singular inversions."""
        return _lib.tail_status(self.ctx)
