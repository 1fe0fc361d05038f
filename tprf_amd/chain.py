"""The forward chain e_k -> G(w,k) -> G(w,r) -> G(tau,r) -> chi0(nu,r) -> chi0(nu,q)
-> chi_RPA(nu,q) on preallocated device buffers (one "eval").

Stages and the reference functions they replace:
  lattice_dyson_g_wk (gf.cpp:126), fourier_wk_to_wr (gf.cpp:209),
  fourier_wr_to_tr (gf.cpp:262), chi0_wr_from_grt_PH (chi_imtime.cpp:139, the
  fused product + tau->nu of chi_imtime.cpp:279), chi_wk_from_chi_wr
  (chi_imtime.cpp:289), solve_rpa_PH (rpa.cpp:66).
Every stage is one asynchronous C-ABI call on torch's current stream (the
library context is rebound to it at every call); the tail-fit status is read
once, after the chain.
"""
import numpy as np
import torch

from . import _lib


def w2t_pruned_fits(n_iw, ntau):
    """True when fourier_wr_to_tr runs its input-pruned kernel for this mesh (and
    tprf_fourier_wr_to_tr_herm accepts it): L = ntau - 1 a multiple of the
    smallest power of two M >= 2 n_iw with 64 <= M <= 1024, and L / M <= 64."""
    L = int(ntau) - 1
    M = 64
    while M < 2 * n_iw:
        M *= 2
    return M <= 1024 and L % M == 0 and L // M <= 64


class BubbleRPA:
    """fused=True (default): the first two stages run as one call,
    tprf_lattice_dyson_g_wr (G(iw, k) is then not stored for a 16^3 lattice),
    and, when e_k is Hermitian (so G_ab(tau, r)^* = G_ba(tau, -r)), omega -> tau
    runs as tprf_fourier_wr_to_tr_herm (columns (r, a, b) with r < -r, or r = -r
    and a <= b, transformed; G(tau, r) is then stored for r <= -r only) where
    its pruned kernel fits the mesh (``w2t_pruned_fits``; otherwise the plain
    transform of every column), and the bubble as
    tprf_chi0_wr_from_grt_PH_herm (orbital pairs a <= b transformed, the others
    formed from the symmetry).  If also U_abcd = conj(U_dcba) (every real
    Hermitian interaction, e.g. the Kanamori tensors), chi0_abcd(-nu, q) =
    conj(chi0_dcba(nu, q)) and the same holds for chi_RPA, so the bubble writes
    and r -> q transforms the frequencies nu >= 0 only and
    tprf_solve_rpa_PH_herm solves those and writes chi(-nu) as the conjugate
    transposes.  fused=False: one call per reference function.

    Results: ``g_wr``, ``g_tr`` (with the symmetric omega -> tau only the columns
    of r <= -r hold data), ``chi0_wr`` and ``chi0_wk`` (the bosonic
    frequencies from index ``nu_lo`` on: all of them, or nu >= 0 on the
    symmetric path), ``chi_wk`` (every frequency)."""
    API_STAGES = ("dyson", "wk_to_wr", "wr_to_tr", "chi0_wr", "wr_to_wk", "rpa")
    FUSED_STAGES = ("dyson_wr", "wr_to_tr", "chi0_wr", "wr_to_wk", "rpa")

    def __init__(self, wl, device="cuda", fused=True):
        self.wl = wl
        self.STAGES = self.FUSED_STAGES if fused else self.API_STAGES
        e = wl.e_k
        self.herm = bool(fused) and bool(np.abs(e - np.conj(np.swapaxes(e, 1, 2))).max() <= 1e-14 * max(1.0, np.abs(e).max()))
        # symmetric omega -> tau only where its pruned kernel takes the mesh
        self.herm_w2t = self.herm and w2t_pruned_fits(wl.n_iw, wl.ntau)
        u = wl.U
        tol_u = 1e-14 * max(1.0, np.abs(u).max())
        self.herm_nu = self.herm and bool(np.abs(u - np.conj(np.einsum("abcd->dcba", u))).max() <= tol_u)
        dev = torch.device(device)
        if dev.index is None:
            dev = torch.device(dev.type, torch.cuda.current_device())
        self.dev = dev
        n, nk = wl.norb, wl.nk
        c128 = torch.complex128
        self.e_k = torch.from_numpy(wl.e_k).to(dev).contiguous()
        self.U = torch.from_numpy(wl.U).to(dev).contiguous()
        # G(iw, k) is stored only by the one-call-per-function chain
        self.g_wk = None if fused else torch.empty((2 * wl.n_iw, nk, n, n), dtype=c128, device=dev)
        self.g_wr = torch.empty((2 * wl.n_iw, nk, n, n), dtype=c128, device=dev)
        self.g_tr = torch.empty((wl.ntau, nk, n, n), dtype=c128, device=dev)
        # chi0 in the reference's [2 nb - 1] frequency frame; on the symmetric path
        # the rows nu < 0 are never written (filled with NaN once, never exposed)
        self._chi0_wr = torch.empty((wl.n_nu, nk, n, n, n, n), dtype=c128, device=dev)
        self._chi0_wk = torch.empty_like(self._chi0_wr)
        self.nu_lo = wl.nb - 1 if self.herm_nu else 0
        if self.nu_lo:
            self._chi0_wr[: self.nu_lo].fill_(complex("nan"))
            self._chi0_wk[: self.nu_lo].fill_(complex("nan"))
        self.chi_wk = torch.empty_like(self._chi0_wr)
        self.dims = _lib.dims3(wl.dims)

    @property
    def chi0_wr(self):
        """chi0(i nu_m, r) for m >= nu_lo - (nb - 1) (all written rows)."""
        return self._chi0_wr[self.nu_lo:]

    @property
    def chi0_wk(self):
        """chi0(i nu_m, q) for the rows that hold data (as chi0_wr)."""
        return self._chi0_wk[self.nu_lo:]

    def _ctx(self):
        """The device's library context, bound to torch's current stream."""
        return _lib.ctx(self.dev.index)

    def stage(self, name):
        wl, c, p = self.wl, self._ctx(), _lib.ptr
        n, nk = wl.norb, wl.nk
        if name == "dyson":
            _lib.call("tprf_lattice_dyson_g_wk", c, wl.beta, wl.n_iw, wl.mu, nk, n, p(self.e_k), _lib.SIGMA_NONE, None,
                      p(self.g_wk))
        elif name == "dyson_wr":
            if self.herm:
                # full = 0 (only r <= -r written) when the symmetric omega -> tau,
                # which reads no other column, follows
                _lib.call("tprf_lattice_dyson_g_wr_herm", c, wl.beta, wl.n_iw, wl.mu, self.dims, n, p(self.e_k),
                          p(self.g_wr), 0 if self.herm_w2t else 1)
            else:
                _lib.call("tprf_lattice_dyson_g_wr", c, wl.beta, wl.n_iw, wl.mu, self.dims, n, p(self.e_k),
                          p(self.g_wr))
        elif name == "wk_to_wr":
            _lib.call("tprf_fourier_wk_to_wr", c, 2 * wl.n_iw, self.dims, n * n, p(self.g_wk), p(self.g_wr))
        elif name == "wr_to_tr":
            if self.herm_w2t:  # only r <= -r is written: the Hermitian bubble reads no other column
                _lib.call("tprf_fourier_wr_to_tr_herm", c, wl.beta, wl.n_iw, wl.ntau, self.dims, n, p(self.g_wr),
                          p(self.g_tr), 0)
            else:
                _lib.call("tprf_fourier_wr_to_tr", c, _lib.FERMION, wl.beta, wl.n_iw, wl.ntau, nk * n * n,
                          p(self.g_wr), p(self.g_tr), None)
        elif name == "chi0_wr":
            if self.herm:  # with herm_nu only nu >= 0 is written: r -> q and RPA read no other rows
                _lib.call("tprf_chi0_wr_from_grt_PH_herm", c, wl.beta, wl.ntau, self.dims, n, wl.nb, p(self.g_tr),
                          p(self._chi0_wr), 0 if self.herm_nu else 1)
            else:
                _lib.call("tprf_chi0_wr_from_grt_PH", c, wl.beta, wl.ntau, self.dims, n, wl.nb, p(self.g_tr),
                          p(self._chi0_wr))
        elif name == "wr_to_wk":
            nrow = wl.n_nu - self.nu_lo  # nu >= 0 only on the symmetric path
            _lib.call("tprf_fourier_wr_to_wk", c, nrow, self.dims, n ** 4, p(self.chi0_wr), p(self.chi0_wk))
        elif name == "rpa":
            if self.herm_nu:
                _lib.call("tprf_solve_rpa_PH_herm", c, wl.nb, nk, n, p(self._chi0_wk), p(self.U), p(self.chi_wk))
            else:
                _lib.call("tprf_solve_rpa_PH", c, wl.n_nu * nk, n, p(self._chi0_wk), p(self.U), p(self.chi_wk))
        else:
            raise KeyError(name)

    def stage_reset(self):
        """Clear the tail-fit / singularity status words (one small fill kernel)."""
        _lib.tail_status_reset(self._ctx())

    def run(self):
        self.stage_reset()
        for s in self.STAGES:
            self.stage(s)

    def capture(self):
        """Capture run() (status reset and every stage) into one HIP graph on a
        private stream; replay() then runs the whole eval as one graph launch.
        The first run() settles the tables, plans and kernel variants, so the
        capture allocates nothing."""
        self.run()
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=torch.cuda.Stream(device=self.dev)):
            self.run()
        self._ctx()  # the context back on torch's current stream
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()

    def check_status(self):
        """Synchronise and return (max tail-fit residual, max |m0|); raises on singular inversions."""
        return _lib.tail_status(self._ctx())
