"""Differentiable static susceptibility chi_RPA(Q, i nu = 0) of one lattice.

Forward (each step one C-ABI call on torch's current stream):
  e_k, mu --lattice_dyson_g_wk--> G(iw, k) --fourier_wk_to_wr--> G(iw, r)
  --fourier_wr_to_tr--> G(tau, r) --nu = 0 row of chi0_wr_from_grt_PH--> chi0(0, r)
  --one momentum of chi_wk_from_chi_wr--> chi0(Q, 0) --solve_rpa_PH--> chi_RPA(Q, 0)
(gf.cpp:91-132, fourier_lattice.cpp:49, fourier_matsubara.cpp:218-308,
chi_imtime.cpp:139-182, fourier_lattice.cpp:55-58, rpa.cpp:29-64).

Backward runs the hand-written adjoint kernels in reverse order.

route="fused" (default) runs the first five stages as one call,
tprf_chi0_static_q_from_ek: the nu = 0 bubble of the fitted omega -> tau
transform is a bilinear form in G(iw) (the Matsubara sum plus a rank-~30
correction from the tail model and the boundary fit), and at one momentum Q
it is (1/nk) sum_k B(G(k), G(k - Q)), so G(iw, k) is evaluated in registers
and neither G(iw, k), G(iw, r) nor G(tau, r) is stored.  Its adjoint
recomputes G the same way; only e_k, mu, the per-k projections of the tail
fit and chi0(Q, 0) are kept between the passes.
route="matsubara" stores G(iw, k), transforms it to G(iw, r) and evaluates
the same bilinear form per r (tprf_chi0_static_from_gwr_PH), then projects
r -> Q; only G(iw, k) and chi0(Q, 0) are kept, G(iw, r) is recomputed.
route="tau" runs fourier_wr_to_tr and the tau-space bubble, the same map
evaluated the reference's way (and recomputes G(tau, r) in backward).
The fused and Matsubara routes take n <= 3 orbitals, the tau route n <= 4;
route=None picks "fused" for n <= 3 and "tau" for n = 4.

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
    """chi_RPA(Q, 0) as a function of (e_k, mu, U) for a fixed mesh: beta, n_iw
    fermionic frequencies (ntau = 6 n_iw + 1 by default), lattice dims and integer
    momentum q (Q = 2 pi q / dims).

    With nbatch > 1 the object evaluates nbatch lattices of that mesh at once
    (e_k[nbatch, nk, n, n], mu[nbatch], U[nbatch, n, n, n, n]), each stage one
    launch over the whole batch; ``graphed_grad`` captures forward + backward of
    L = sum_i Re tr chi_i in one HIP graph."""

    def __init__(self, beta, n_iw, dims, q, norb, ntau=None, nbatch=1, device=None, route=None):
        self.beta = float(beta)
        self.n_iw = int(n_iw)
        self.dims = _dims3(dims)
        self.q = _dims3(q)
        self.n = int(norb)
        self.nb = int(nbatch)
        self.nk = self.dims[0] * self.dims[1] * self.dims[2]
        self.nw = 2 * self.n_iw
        self.ntau = 6 * self.n_iw + 1 if (ntau is None or ntau <= 0) else int(ntau)
        dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.dev = dev
        n, nk, B, c128 = self.n, self.nk, self.nb, torch.complex128
        if route is None:
            route = "fused" if n <= 3 else "tau"
        if route not in ("fused", "matsubara", "tau"):
            raise ValueError("route must be 'fused', 'matsubara' or 'tau'")
        self.route = route
        if route != "fused":
            self._g_wr = torch.empty((self.nw, B, nk, n, n), dtype=c128, device=dev)
            self._chi_r = torch.empty((B, nk, n, n, n, n), dtype=c128, device=dev)
        if route == "tau":
            self._g_tr = torch.empty((self.ntau, B, nk, n, n), dtype=c128, device=dev)
            self._x_tr = torch.empty((self.ntau, B, nk, n, n), dtype=c128, device=dev)
        elif route == "matsubara":
            self._gbar_wr = torch.empty((self.nw, B, nk, n, n), dtype=c128, device=dev)
        self._cdims = _lib.dims3(self.dims)
        self._cq = _lib.dims3(self.q)
        self._graph = None

    # -- pieces ---------------------------------------------------------------
    def _ctx(self):
        return _lib.ctx(self.dev.index)

    def _inputs(self, e_k, mu, U):
        n, nk, B = self.n, self.nk, self.nb
        e_k = e_k.detach().to(self.dev, torch.complex128).reshape(B, nk, n, n).contiguous()
        U = U.detach().to(self.dev, torch.complex128).reshape(B, n, n, n, n).contiguous()
        mu = torch.as_tensor(mu, dtype=torch.float64).detach().to(self.dev).reshape(B).contiguous()
        return e_k, mu, U

    def _g_tr_from_g_wk(self, c, g_wk):
        p = _lib.ptr
        n, nk, B = self.n, self.nk, self.nb
        _lib.call("tprf_fourier_wk_to_wr", c, self.nw * B, self._cdims, n * n, p(g_wk), p(self._g_wr))
        _lib.call("tprf_fourier_wr_to_tr", c, _lib.FERMION, self.beta, self.n_iw, self.ntau, B * nk * n * n,
                  p(self._g_wr), p(self._g_tr), None)

    def _keep(self, e_k, mu):
        """What backward needs from forward: (G(i w, k),) on the matsubara and tau
        routes; (e_k, mu, per-k projections [29, nbatch, nk, n, n]) on the fused one."""
        n, nk, B = self.n, self.nk, self.nb
        if self.route == "fused":
            return (e_k, mu, self._new(29, B, nk, n, n))
        return (self._new(self.nw, B, nk, n, n),)

    def _forward_into(self, c, e_k, mu, U, keep, chi0_q, chi):
        p = _lib.ptr
        n, nk, B = self.n, self.nk, self.nb
        if self.route == "fused":
            _lib.call("tprf_chi0_static_q_from_ek", c, self.beta, self.n_iw, self.ntau, self._cdims, self._cq, n, B,
                      _lib.dptr(mu), p(e_k), p(chi0_q), p(keep[2]))
            _lib.call("tprf_solve_rpa_PH_adjoint", c, B, n, p(chi0_q), p(U), None, p(chi), None, None)
            return
        g_wk = keep[0]
        _lib.call("tprf_lattice_dyson_g_wk_batched", c, self.beta, self.n_iw, _lib.dptr(mu), B, nk, n, p(e_k),
                  p(g_wk))
        if self.route == "tau":
            self._g_tr_from_g_wk(c, g_wk)
            _lib.call("tprf_chi0_static_from_grt_PH", c, self.beta, self.ntau, self._cdims, n, B, p(self._g_tr),
                      p(self._chi_r))
        else:
            _lib.call("tprf_fourier_wk_to_wr", c, self.nw * B, self._cdims, n * n, p(g_wk), p(self._g_wr))
            _lib.call("tprf_chi0_static_from_gwr_PH", c, self.beta, self.n_iw, self.ntau, self._cdims, n, B,
                      p(self._g_wr), p(self._chi_r))
        _lib.call("tprf_fourier_r_to_q", c, self._cdims, self._cq, n ** 4, B, p(self._chi_r), p(chi0_q))
        _lib.call("tprf_solve_rpa_PH_adjoint", c, B, n, p(chi0_q), p(U), None, p(chi), None, None)

    def _backward_into(self, c, keep, chi0_q, U, chibar, chi, chi0bar, Ubar, ebar, mubar, recompute=True):
        p = _lib.ptr
        n, nk, B = self.n, self.nk, self.nb
        _lib.call("tprf_solve_rpa_PH_adjoint", c, B, n, p(chi0_q), p(U), p(chibar), p(chi), p(chi0bar), p(Ubar))
        if self.route == "fused":
            e_k, mu, proj = keep
            _lib.call("tprf_chi0_static_q_from_ek_adjoint", c, self.beta, self.n_iw, self.ntau, self._cdims,
                      self._cq, n, B, _lib.dptr(mu), p(e_k), p(chi0bar), p(proj), p(ebar), _lib.dptr(mubar))
            return
        g_wk = keep[0]
        chibar_r = self._chi_r
        _lib.call("tprf_fourier_r_to_q_adjoint", c, self._cdims, self._cq, n ** 4, B, p(chi0bar), p(chibar_r))
        if self.route == "matsubara":
            if recompute:  # G(i w, r) from the kept G(i w, k)
                _lib.call("tprf_fourier_wk_to_wr", c, self.nw * B, self._cdims, n * n, p(g_wk), p(self._g_wr))
            gbar_wr = self._gbar_wr
            _lib.call("tprf_chi0_static_from_gwr_PH_adjoint", c, self.beta, self.n_iw, self.ntau, self._cdims, n, B,
                      p(self._g_wr), p(chibar_r), p(gbar_wr))
            # k -> r adjoint in place: gbar(i w, k) overwrites gbar(i w, r)
            _lib.call("tprf_fourier_wk_to_wr_adjoint", c, self.nw * B, self._cdims, n * n, p(gbar_wr), p(gbar_wr))
            _lib.call("tprf_lattice_dyson_g_wk_adjoint", c, self.n_iw, nk, n, B, p(g_wk), p(gbar_wr), p(ebar),
                      _lib.dptr(mubar))
            return
        if recompute:
            # recompute G(tau, r) from the kept G(iw, k)
            self._g_tr_from_g_wk(c, g_wk)
        gbar_tr = self._x_tr
        _lib.call("tprf_chi0_static_from_grt_PH_adjoint", c, self.beta, self.ntau, self._cdims, n, B,
                  p(self._g_tr), p(chibar_r), p(gbar_tr))
        gbar_wr = self._g_wr
        _lib.call("tprf_fourier_wr_to_tr_adjoint", c, _lib.FERMION, self.beta, self.n_iw, self.ntau, B * nk * n * n,
                  p(gbar_tr), p(gbar_wr), 1)
        gbar_wk = self._g_tr.view(-1)[: self.nw * B * nk * n * n].view(self.nw, B, nk, n, n)
        _lib.call("tprf_fourier_wk_to_wr_adjoint", c, self.nw * B, self._cdims, n * n, p(gbar_wr), p(gbar_wk))
        _lib.call("tprf_lattice_dyson_g_wk_adjoint", c, self.n_iw, nk, n, B, p(g_wk), p(gbar_wk), p(ebar),
                  _lib.dptr(mubar))

    def _new(self, *shape, dtype=torch.complex128):
        return torch.empty(shape, dtype=dtype, device=self.dev)

    # -- eager API --------------------------------------------------------------
    def forward(self, e_k, mu, U, check_tail=True):
        """Returns (chi[nbatch, n,n,n,n], chi0_q, keep); keep (a tuple of tensors) is
        what backward needs."""
        n, nk, B = self.n, self.nk, self.nb
        e_k, mu, U = self._inputs(e_k, mu, U)
        c = self._ctx()
        keep = self._keep(e_k, mu)
        chi0_q, chi = self._new(B, n, n, n, n), self._new(B, n, n, n, n)
        if check_tail:
            _lib.tail_status_reset(c)
        self._forward_into(c, e_k, mu, U, keep, chi0_q, chi)
        if check_tail:
            err, _ = _lib.tail_status(c)
            if err >= 1e-3:
                raise _lib.TprfError(-4, f"tail fit error {err:.3g} >= 1e-3")
        return chi, chi0_q, keep

    def backward(self, keep, chi0_q, U, chibar):
        """Returns (ebar_k[nbatch, nk, n, n], mubar[nbatch], Ubar[nbatch, n, n, n, n])."""
        n, nk, B = self.n, self.nk, self.nb
        U = U.detach().to(self.dev, torch.complex128).reshape(B, n, n, n, n).contiguous()
        chibar = chibar.detach().to(self.dev, torch.complex128).reshape(B, n, n, n, n).contiguous()
        chi, chi0bar, Ubar = (self._new(B, n, n, n, n) for _ in range(3))
        ebar = self._new(B, nk, n, n)
        mubar = self._new(B, dtype=torch.float64)
        self._backward_into(self._ctx(), keep, chi0_q, U, chibar, chi, chi0bar, Ubar, ebar, mubar)
        return ebar, mubar, Ubar

    # -- one HIP graph: forward + backward of L = sum_i Re tr chi_i ------------
    def graphed_grad(self, e_k, mu, U):
        """Copies the inputs into static buffers and replays (capturing on the first
        call) forward + backward for the whole batch.  Returns (chi, ebar, mubar, Ubar)
        as views of static buffers, valid until the next call.  Forward and backward
        run back to back in one launch, so the backward uses G(tau, r) from the
        forward's transient work buffer instead of recomputing it (the per-call
        autograd path recomputes it, since it keeps only G(i w, k) between calls):
        G(i w, r) on the Matsubara route, G(tau, r) on the tau route."""
        n, nk, B = self.n, self.nk, self.nb
        e_k, mu, U = self._inputs(e_k, mu, U)
        if self._graph is None:
            st = {"e_k": e_k.clone(), "mu": mu.clone(), "U": U.clone(),
                  "chi0_q": self._new(B, n, n, n, n), "chi": self._new(B, n, n, n, n),
                  "chi2": self._new(B, n, n, n, n), "chi0bar": self._new(B, n, n, n, n),
                  "Ubar": self._new(B, n, n, n, n), "ebar": self._new(B, nk, n, n),
                  "mubar": self._new(B, dtype=torch.float64)}
            cb = torch.zeros((B, n, n, n, n), dtype=torch.complex128)
            for a in range(n):
                for b in range(n):
                    cb[:, a, a, b, b] = 1.0  # d(Re sum_ab chi_aabb) in the complex-gradient convention
            st["chibar"] = cb.to(self.dev)
            st["keep"] = self._keep(st["e_k"], st["mu"])

            def run():
                c = self._ctx()
                self._forward_into(c, st["e_k"], st["mu"], st["U"], st["keep"], st["chi0_q"], st["chi"])
                self._backward_into(c, st["keep"], st["chi0_q"], st["U"], st["chibar"], st["chi2"],
                                    st["chi0bar"], st["Ubar"], st["ebar"], st["mubar"], recompute=False)

            # warm-up outside the graph: plans and tables are made here, and the
            # scratch the calls need is measured
            c = self._ctx()
            _lib.workspace_size(c, reset=True)
            side = torch.cuda.Stream(device=self.dev)
            side.wait_stream(torch.cuda.current_stream(self.dev))
            with torch.cuda.stream(side):
                run()
            torch.cuda.current_stream(self.dev).wait_stream(side)
            torch.cuda.synchronize(self.dev)
            # the graph gets a workspace of its own (kept with its static buffers),
            # so no later call that grows the context's scratch, and no other
            # graph, touches the memory this graph's kernels use
            st["work"] = torch.empty(max(_lib.workspace_size(c, reset=True), 256), dtype=torch.uint8,
                                     device=self.dev)
            g = torch.cuda.CUDAGraph()
            _lib.set_workspace(c, st["work"])
            try:
                with torch.cuda.graph(g):
                    run()
            finally:
                _lib.set_workspace(c, None)
            self._graph, self._static = g, st
        st = self._static
        st["e_k"].copy_(e_k)
        st["mu"].copy_(mu)
        st["U"].copy_(U)
        self._graph.replay()
        return st["chi"], st["ebar"], st["mubar"], st["Ubar"]


class _StaticRPAFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, e_k, mu, U, spec):
        chi, chi0_q, keep = spec.forward(e_k, mu, U)
        fctx.spec = spec
        fctx.mu_like = (mu.device, mu.dtype, mu.shape) if torch.is_tensor(mu) else None
        fctx.e_meta = (e_k.device, e_k.dtype, e_k.shape)
        fctx.u_meta = (U.device, U.dtype, U.shape)
        fctx.nkeep = len(keep)
        fctx.save_for_backward(*keep, chi0_q, U.detach())
        return chi.reshape(U.shape)

    @staticmethod
    def backward(fctx, chibar):
        saved = fctx.saved_tensors
        keep, chi0_q, U = saved[: fctx.nkeep], saved[-2], saved[-1]
        ebar, mubar, Ubar = fctx.spec.backward(keep, chi0_q, U, chibar)
        ge = ebar.to(*fctx.e_meta[:2]).reshape(fctx.e_meta[2]) if fctx.needs_input_grad[0] else None
        gm = None
        if fctx.needs_input_grad[1] and fctx.mu_like is not None:
            dev, dt, shape = fctx.mu_like
            gm = mubar.to(dev, dt).reshape(shape)
        gU = Ubar.to(*fctx.u_meta[:2]).reshape(fctx.u_meta[2]) if fctx.needs_input_grad[2] else None
        return ge, gm, gU, None


def chi_rpa_static(e_k, mu, U, spec):
    """chi_RPA(Q, 0)[..., a,b,c,d] for the mesh of `spec` (a StaticRPA), shaped like U;
    differentiable in e_k (complex [(nbatch,) nk, n, n]), mu (float64, 0-dim or
    [nbatch]) and U (complex [(nbatch,) n, n, n, n])."""
    return _StaticRPAFunction.apply(e_k, mu, U, spec)


def trace_loss(chi):
    """L = Re sum_{a,b} chi_{aabb} (summed over a leading batch axis, if any)."""
    n = chi.shape[-1]
    idx = torch.arange(n, device=chi.device)
    return chi[..., idx[:, None], idx[:, None], idx[None, :], idx[None, :]].sum().real


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


def two_orbital_square_e_k(tp, v, dims, t=1.0):
    """Two-orbital square lattice: both orbitals e(k) = -2t(cos kx + cos ky) - 4t' cos kx cos ky,
    inter-orbital nearest-neighbour hopping v: e_12(k) = -2v(cos kx + cos ky).
    tp, v: real tensors of shape [B]; returns complex [B, nk, 2, 2]."""
    d0, d1 = int(dims[0]), int(dims[1])
    dev = tp.device
    kx = 2 * torch.pi * torch.arange(d0, dtype=torch.float64, device=dev) / d0
    ky = 2 * torch.pi * torch.arange(d1, dtype=torch.float64, device=dev) / d1
    cx, cy = torch.cos(kx)[:, None].expand(d0, d1).reshape(-1), torch.cos(ky)[None, :].expand(d0, d1).reshape(-1)
    eps = -2 * t * (cx + cy)[None, :] - 4 * tp[:, None] * (cx * cy)[None, :]
    off = -2 * v[:, None] * (cx + cy)[None, :]
    row0 = torch.stack([eps, off], dim=-1)
    row1 = torch.stack([off, eps], dim=-1)
    return torch.stack([row0, row1], dim=-2).to(torch.complex128)


def kanamori_spin_U(U, J):
    """Spin-channel Kanamori tensor of two orbitals (U' = U - 2J, J' = J) in TPRF's
    [a, a', b, b'] order (rpa_tensor.py:94-118) as a differentiable function of the
    real tensors U, J of shape [B]; returns complex [B, 2, 2, 2, 2]."""
    B = U.shape[0]
    out = torch.zeros((B, 2, 2, 2, 2), dtype=torch.float64, device=U.device)
    Up = U - 2 * J
    for a in range(2):
        for b in range(2):
            if a == b:
                out[:, a, a, a, a] = U
            else:
                out[:, a, b, b, a] = Up      # a = b' != b = a'
                out[:, a, a, b, b] = J       # a = a' != b = b'
                out[:, a, b, a, b] = J       # a = b != a' = b'
    return out.to(torch.complex128)
