"""Differentiable static susceptibility chi_RPA(Q, i nu = 0) of one lattice.

Forward (each step one C-ABI call on torch's current stream):
  e_k, mu --lattice_dyson_g_wk--> G(iw, k) --fourier_wk_to_wr--> G(iw, r)
  --fourier_wr_to_tr--> G(tau, r) --nu = 0 row of chi0_wr_from_grt_PH--> chi0(0, r)
  --one momentum of chi_wk_from_chi_wr--> chi0(Q, 0) --solve_rpa_PH--> chi_RPA(Q, 0)
(gf.cpp:91-132, fourier_lattice.cpp:49, fourier_matsubara.cpp:218-308,
chi_imtime.cpp:139-182, fourier_lattice.cpp:55-58, rpa.cpp:29-64).

Backward runs the hand-written adjoint kernels in reverse order.  Only
G(iw, k) and chi0(Q, 0) are kept from the forward pass: G(iw, r) and G(tau, r)
are recomputed by the two transforms instead of being stored.

Gradients follow torch's complex convention (dL/dRe z + i dL/dIm z), so
e_k, mu and U may come out of any torch expression (hoppings, Kanamori
parameters) and autograd carries the gradient on to those leaves.
"""
import torch

from . import _lib


def _dims3(dims):
    d = tuple(int(x) for x in dims)
    return d + (1,) * (3 - len(d))


class StaticRPA:
    """chi_RPA(Q, 0) as a function of (e_k[nk, n, n], mu, U[n, n, n, n]) for a fixed
    mesh: beta, n_iw fermionic frequencies (ntau = 6 n_iw + 1 by default),
    lattice dims and integer momentum q (Q = 2 pi q / dims)."""

    def __init__(self, beta, n_iw, dims, q, norb, ntau=None, device=None):
        self.beta = float(beta)
        self.n_iw = int(n_iw)
        self.dims = _dims3(dims)
        self.q = _dims3(q)
        self.n = int(norb)
        self.nk = self.dims[0] * self.dims[1] * self.dims[2]
        self.nw = 2 * self.n_iw
        self.ntau = 6 * self.n_iw + 1 if (ntau is None or ntau <= 0) else int(ntau)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dev = dev
        n, nk, c128 = self.n, self.nk, torch.complex128
        self._g_wr = torch.empty((self.nw, nk, n, n), dtype=c128, device=dev)
        self._g_tr = torch.empty((self.ntau, nk, n, n), dtype=c128, device=dev)
        self._x_tr = torch.empty((self.ntau, nk, n, n), dtype=c128, device=dev)
        self._chi_r = torch.empty((nk, n, n, n, n), dtype=c128, device=dev)
        self._cdims = _lib.dims3(self.dims)
        self._cq = _lib.dims3(self.q)

    # -- pieces ---------------------------------------------------------------
    def _ctx(self):
        return _lib.ctx(self.dev.index)

    def _check_input(self, e_k, U):
        n, nk = self.n, self.nk
        if tuple(e_k.shape) != (nk, n, n) or tuple(U.shape) != (n, n, n, n):
            raise ValueError(f"expected e_k[{nk},{n},{n}] and U[{n},{n},{n},{n}]")

    def _g_tr_from_g_wk(self, c, g_wk):
        p = _lib.ptr
        n, nk = self.n, self.nk
        _lib.call("tprf_fourier_wk_to_wr", c, self.nw, self._cdims, n * n, p(g_wk), p(self._g_wr))
        _lib.call("tprf_fourier_wr_to_tr", c, _lib.FERMION, self.beta, self.n_iw, self.ntau, nk * n * n,
                  p(self._g_wr), p(self._g_tr), None)

    def forward(self, e_k, mu, U, check_tail=True):
        """Returns (chi[n,n,n,n], chi0_q[n,n,n,n], g_wk)."""
        self._check_input(e_k, U)
        c, p = self._ctx(), _lib.ptr
        n, nk = self.n, self.nk
        e_k = e_k.detach().to(self.dev, torch.complex128).contiguous()
        U = U.detach().to(self.dev, torch.complex128).contiguous()
        g_wk = torch.empty((self.nw, nk, n, n), dtype=torch.complex128, device=self.dev)
        chi0_q = torch.empty((n, n, n, n), dtype=torch.complex128, device=self.dev)
        chi = torch.empty_like(chi0_q)
        if check_tail:
            _lib.tail_status_reset(c)
        _lib.call("tprf_lattice_dyson_g_wk", c, self.beta, self.n_iw, float(mu), nk, n, p(e_k), _lib.SIGMA_NONE,
                  None, p(g_wk))
        self._g_tr_from_g_wk(c, g_wk)
        _lib.call("tprf_chi0_static_from_grt_PH", c, self.beta, self.ntau, self._cdims, n, p(self._g_tr),
                  p(self._chi_r))
        _lib.call("tprf_fourier_r_to_q", c, self._cdims, self._cq, n ** 4, p(self._chi_r), p(chi0_q))
        _lib.call("tprf_solve_rpa_PH_adjoint", c, 1, n, p(chi0_q), p(U), None, p(chi), None, None)
        if check_tail:
            err, _ = _lib.tail_status(c)
            if err >= 1e-3:
                raise _lib.TprfError(-4, f"tail fit error {err:.3g} >= 1e-3")
        return chi, chi0_q, g_wk

    def backward(self, g_wk, chi0_q, U, chibar):
        """Returns (ebar_k[nk,n,n], mubar (0-dim float64), Ubar[n,n,n,n])."""
        c, p = self._ctx(), _lib.ptr
        n, nk = self.n, self.nk
        U = U.detach().to(self.dev, torch.complex128).contiguous()
        chibar = chibar.detach().to(self.dev, torch.complex128).contiguous()
        chi = torch.empty_like(chi0_q)
        chi0bar = torch.empty_like(chi0_q)
        Ubar = torch.empty_like(chi0_q)
        ebar = torch.empty((nk, n, n), dtype=torch.complex128, device=self.dev)
        mubar = torch.empty((), dtype=torch.float64, device=self.dev)
        # RPA and momentum projection
        _lib.call("tprf_solve_rpa_PH_adjoint", c, 1, n, p(chi0_q), p(U), p(chibar), p(chi), p(chi0bar), p(Ubar))
        chibar_r = self._chi_r
        _lib.call("tprf_fourier_r_to_q_adjoint", c, self._cdims, self._cq, n ** 4, p(chi0bar), p(chibar_r))
        # recompute G(tau, r) from the kept G(iw, k)
        self._g_tr_from_g_wk(c, g_wk)
        gbar_tr = self._x_tr
        _lib.call("tprf_chi0_static_from_grt_PH_adjoint", c, self.beta, self.ntau, self._cdims, n, p(self._g_tr),
                  p(chibar_r), p(gbar_tr))
        gbar_wr = self._g_wr
        _lib.call("tprf_fourier_wr_to_tr_adjoint", c, _lib.FERMION, self.beta, self.n_iw, self.ntau, nk * n * n,
                  p(gbar_tr), p(gbar_wr), 1)
        gbar_wk = self._g_tr.view(-1)[: self.nw * nk * n * n].view(self.nw, nk, n, n)
        _lib.call("tprf_fourier_wk_to_wr_adjoint", c, self.nw, self._cdims, n * n, p(gbar_wr), p(gbar_wk))
        _lib.call("tprf_lattice_dyson_g_wk_adjoint", c, self.n_iw, nk, n, p(g_wk), p(gbar_wk), p(ebar),
                  _lib.dptr(mubar))
        return ebar, mubar, Ubar


class _StaticRPAFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, e_k, mu, U, spec):
        mu_val = float(mu)
        chi, chi0_q, g_wk = spec.forward(e_k, mu_val, U)
        fctx.spec = spec
        fctx.mu_like = (mu.device, mu.dtype, mu.shape) if torch.is_tensor(mu) else None
        fctx.e_meta = (e_k.device, e_k.dtype)
        fctx.u_meta = (U.device, U.dtype)
        fctx.save_for_backward(g_wk, chi0_q, U.detach())
        return chi

    @staticmethod
    def backward(fctx, chibar):
        g_wk, chi0_q, U = fctx.saved_tensors
        ebar, mubar, Ubar = fctx.spec.backward(g_wk, chi0_q, U, chibar)
        ge = ebar.to(*fctx.e_meta) if fctx.needs_input_grad[0] else None
        gm = None
        if fctx.needs_input_grad[1] and fctx.mu_like is not None:
            dev, dt, shape = fctx.mu_like
            gm = mubar.to(dev, dt).reshape(shape)
        gU = Ubar.to(*fctx.u_meta) if fctx.needs_input_grad[2] else None
        return ge, gm, gU, None


def chi_rpa_static(e_k, mu, U, spec):
    """chi_RPA(Q, 0)[a,b,c,d] for the mesh of `spec` (a StaticRPA); differentiable in
    e_k (complex [nk,n,n]), mu (float64 tensor) and U (complex [n,n,n,n])."""
    return _StaticRPAFunction.apply(e_k, mu, U, spec)


def trace_loss(chi):
    """L = Re sum_{a,b} chi_{aabb}."""
    n = chi.shape[0]
    idx = torch.arange(n, device=chi.device)
    return chi[idx[:, None], idx[:, None], idx[None, :], idx[None, :]].sum().real


def square_e_k(t, tp, dims):
    """Single-band square lattice e(k) = -2t(cos kx + cos ky) - 4t' cos kx cos ky as a
    differentiable function of the hopping tensors t, tp; returns complex [nk, 1, 1]."""
    d0, d1 = int(dims[0]), int(dims[1])
    dev = t.device
    kx = 2 * torch.pi * torch.arange(d0, dtype=torch.float64, device=dev) / d0
    ky = 2 * torch.pi * torch.arange(d1, dtype=torch.float64, device=dev) / d1
    cx, cy = torch.cos(kx)[:, None], torch.cos(ky)[None, :]
    e = -2 * t * (cx + cy) - 4 * tp * cx * cy
    return e.reshape(-1, 1, 1).to(torch.complex128)


def hubbard_U(U):
    """1x1x1x1 onsite interaction tensor from a real U tensor (single band)."""
    return U.reshape(1, 1, 1, 1).to(torch.complex128)
