"""FP64 reference of the static chi_RPA(Q, i nu = 0) chain in plain torch on the CPU.

TEST INFRASTRUCTURE ONLY: torch autograd through this module gives the
reference gradient that the hand-written adjoint kernels are checked against.
The product library never imports it.

The linear operators of the two Matsubara transforms that carry the
third-party tail fit and the boundary-derivative fit are read off the oracle's
C restatement column by column (oracle.w_to_t_matrix / t_to_w_matrix with
moments), so nothing here reuses the product's kernels or tables:
  G(iw,k)   = inverse((iw + mu) 1 - e_k)                       gf.cpp:91-132
  G(iw,r)   = (1/N_k) sum_k exp(-i k.r) G(iw,k)                fourier_lattice.cpp:39,49
  G(tau,r)  = three-pole tail model + FFT of the remainder     fourier_matsubara.cpp:249-307
              with moments m = Mop G(iw,r) (the oracle's fit)
  chi0(0,r) = nu = 0 of the bosonic tau -> nu transform of      chi_imtime.cpp:139-182,
              G_da(tau_j,r) G_bc(beta-tau_j,-r)                fourier_matsubara.cpp:42-184
  chi0(Q,0) = sum_r exp(+i Q.r) chi0(0,r)                      fourier_lattice.cpp:40,55-58
  chi       = inverse(1 - chi0 U) chi0 in the PH grouping      rpa.cpp:29-64
"""
import math

import numpy as np
import torch

from oracle import oracle as O


def _dims3(dims):
    d = tuple(int(x) for x in dims)
    return d + (1,) * (3 - len(d))


class StaticChainRef:
    def __init__(self, beta, n_iw, dims, q, ntau=None):
        self.beta, self.n_iw = float(beta), int(n_iw)
        self.dims, self.q = _dims3(dims), _dims3(q)
        self.nk = int(np.prod(self.dims))
        nw = 2 * self.n_iw
        self.ntau = 6 * self.n_iw + 1 if ntau is None else int(ntau)
        L = self.ntau - 1
        self.Mop = torch.from_numpy(O.w_to_t_moments_matrix(self.beta, self.n_iw, O.FERMION, self.ntau))
        n_w = np.arange(-self.n_iw, self.n_iw)
        w = np.pi * (2 * n_w + 1) / self.beta
        b = (0.0, 1.0, -1.0)
        self.pole = torch.from_numpy(np.stack([1.0 / (1j * w - bq) for bq in b]))  # [3, nw]
        self.pos = torch.from_numpy((n_w + L) % L)
        tau = np.arange(L) * (self.beta / L)

        def one_fermion(bq):  # oneFermion with a = 1, fourier_matsubara.cpp:31-33
            if bq >= 0:
                return -np.exp(-bq * tau) / (1 + math.exp(-self.beta * bq))
            return -np.exp(bq * (self.beta - tau)) / (1 + math.exp(self.beta * bq))

        self.psi = torch.from_numpy(np.stack([one_fermion(bq) for bq in b]) + 0j)  # [3, L]
        self.tphase = torch.from_numpy(np.exp(-1j * np.pi * np.arange(L) / L))
        bb = (-0.5, -1.0, 1.0)

        def one_boson(bq):  # oneBoson with a = 1, fourier_matsubara.cpp:35-37
            if bq >= 0:
                return np.exp(-bq * tau) / (math.exp(-self.beta * bq) - 1)
            return np.exp(bq * (self.beta - tau)) / (1 - math.exp(bq * self.beta))

        self.psi_b = torch.from_numpy(np.stack([one_boson(bq) for bq in bb]) + 0j)  # [3, L]
        self.inv_b = [1.0 / (-bq) for bq in bb]
        self.vinv = torch.from_numpy(O.vandermonde_inverse8().reshape(8, 8) + 0j)
        n = np.arange(-self.n_iw, self.n_iw)
        self.iw = torch.from_numpy(1j * np.pi * (2 * n + 1) / self.beta)
        i0, i1, i2 = np.meshgrid(*[np.arange(d) for d in self.dims], indexing="ij")
        d0, d1, d2 = self.dims
        mr = (((-i0) % d0) * d1 + (-i1) % d1) * d2 + (-i2) % d2
        self.mr = torch.from_numpy(mr.ravel())
        ph = 2 * np.pi * (self.q[0] * i0 / d0 + self.q[1] * i1 / d1 + self.q[2] * i2 / d2)
        self.phase = torch.from_numpy(np.exp(1j * ph).ravel())

    def w_to_t(self, gw):
        """Fermionic omega -> tau of gw[nw, ncol] (fourier_matsubara.cpp:249-307)."""
        L = self.ntau - 1
        m1, m2, m3 = self.Mop @ gw
        a = (m1 - m3, (m2 + m3) / 2, (m3 - m2) / 2)
        model = sum(a[q][None, :] * self.pole[q][:, None] for q in range(3))
        gin = torch.zeros((L, gw.shape[1]), dtype=torch.complex128)
        gin = gin.index_copy(0, self.pos, (gw - model) / self.beta)
        y = torch.fft.fft(gin, dim=0) * self.tphase[:, None]
        y = y + sum(self.psi[q][:, None] * a[q][None, :] for q in range(3))
        return torch.cat([y, (-(y[0] + m1))[None]], dim=0)

    def t_to_w0(self, x):
        """Bosonic tau -> nu at nu = 0 of x[ntau, ...] (fourier_matsubara.cpp:42-184)."""
        L = self.ntau - 1
        dtau = self.beta / L
        tm = (torch.arange(1, 9, dtype=torch.float64) * dtau).reshape((8,) + (1,) * (x.dim() - 1))
        dl = (x[1:9] - x[0][None]) / tm
        dr = (x[L][None] - x[L - 8:L].flip(0)) / tm
        v0 = self.vinv[0].reshape(tm.shape)
        v1 = self.vinv[1].reshape(tm.shape)
        gl0, gl1 = (v0 * dl).sum(0), (v1 * dl).sum(0)
        gr0, gr1 = (v0 * dr).sum(0), (v1 * dr).sum(0)
        m1 = -(x[0] - x[L])
        m2 = gl0 - gr0
        m3 = -(gl1 + gr1) * 2.0 / dtau
        a = (4 * (m1 - m3) / 3, m3 - (m1 + m2) / 2, m1 / 6 + m2 / 2 + m3 / 3)
        shp = (L,) + (1,) * (x.dim() - 1)
        model = sum(a[q][None] * self.psi_b[q].reshape(shp) for q in range(3))
        y0 = (self.beta / L) * (x[:L] - model).sum(0)
        return y0 + sum(a[q] * self.inv_b[q] for q in range(3))

    def chi(self, e_k, mu, U):
        """e_k complex [nk,n,n], mu real 0-dim, U complex [n,n,n,n] (CPU tensors)."""
        n, nk, nw = e_k.shape[-1], self.nk, 2 * self.n_iw
        eye = torch.eye(n, dtype=torch.complex128)
        M = (self.iw[:, None, None, None] + mu) * eye - e_k[None]
        G = torch.linalg.inv(M)
        Gk = G.reshape(nw, *self.dims, n, n)
        Gr = torch.fft.fftn(Gk, dim=(1, 2, 3)) / nk
        Gt = self.w_to_t(Gr.reshape(nw, nk * n * n)).reshape(self.ntau, nk, n, n)
        Gm = Gt.flip(0)[:, self.mr]  # G(tau_{L-j}, -r)
        x = torch.einsum("jrda,jrbc->jrabcd", Gt, Gm)
        chi0_r = self.t_to_w0(x)
        chi0_q = torch.einsum("r,rabcd->abcd", self.phase, chi0_r)
        X = chi0_q.reshape(n * n, n * n)
        MU = U.reshape(n * n, n * n)
        chim = torch.linalg.solve(torch.eye(n * n, dtype=torch.complex128) - X @ MU, X)
        return chim.reshape(n, n, n, n)
