"""GPU tests of the benchmark configurations end to end, the reference driver's
branches, the meshes the symmetric kernels do not take, and the error paths.

* The 3-orbital 16^3 k chain against the oracle on every (nu, q) point: chi0
  at 1e-10, and chi_RPA through the residual of the reference's RPA equation
  (1 - chi0 U) chi = chi0 (rpa.cpp:57) with the oracle's chi0, which bounds
  every point however ill-conditioned 1 - chi0 U is there; the forward error
  per point is bounded by the condition number as well.
* The 48-lattice uneven batch in one graph: every lattice against the same
  lattice's chain of one call per reference function, and a reduced uneven
  batch against the oracle.
* The 48 two-orbital 48x48 parameter sets of the batched forward + adjoint in
  one graph: every set against the per-lattice autograd call, two sets against
  the FP64 torch reference.
* imtime_bubble_chi0_wk (lattice_utils.py:191-269) in its three branches
  against the oracle, and the reference's chi0_save_memory.py:26-30 check.
* Error paths of fourier_matsubara.cpp:227-239 and the singular inversions,
  and every size the kernels refuse.
"""
import numpy as np
import pytest

from conftest import mirror_nu, relerr, rpa_kappa, rpa_residual

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    return torch


def host(t):
    return t.detach().cpu().numpy()


def oracle_chain(O, wl):
    g_wk = O.lattice_dyson_g_wk(wl.beta, wl.n_iw, wl.mu, wl.e_k)
    g_wr = O.fourier_lattice(g_wk, wl.dims, -1)
    del g_wk
    g_tr = O.fourier_w_to_t(g_wr, wl.beta, ntau=wl.ntau)
    del g_wr
    chi0_wr = O.chi0_wr_from_grt_PH(g_tr, wl.dims, wl.beta, wl.nb)
    chi0_wk = O.fourier_lattice(chi0_wr, wl.dims, +1)
    del chi0_wr
    return dict(g_tr=g_tr, chi0_wk=chi0_wk, chi_wk=O.solve_rpa_PH(chi0_wk, wl.U))


def canonical_r(dims):
    from tprf_amd.gf import MeshCycLat
    lat = MeshCycLat(dims)
    return np.array([r for r in range(lat.size) if r <= lat.minus(r)])


def check_chain_against(ch, chi0_ref, chi_ref, wl, label, chi_tol=TOL):
    """chi0 (the rows the chain stores) at 1e-10; chi_RPA on every point through
    the residual with the reference's chi0; the chain's own solve backward
    stable; on the symmetric path chi(-nu) the mirror of chi(nu)."""
    n = wl.norb
    got0 = host(ch.chi0_wk)
    e0 = relerr(got0, chi0_ref[ch.nu_lo:])
    chi = host(ch.chi_wk)
    res_ref = rpa_residual(chi0_ref, wl.U, chi).max()
    # the chain's own chi0 and chi (rows nu >= 0 on the symmetric path)
    res_own = rpa_residual(got0, wl.U, chi[ch.nu_lo:]).max()
    print(f"{label}: chi0 rel {e0:.2e}, chi residual vs reference chi0 {res_ref:.2e}, own {res_own:.2e}", flush=True)
    assert e0 <= TOL, label
    assert res_ref <= chi_tol, label
    assert res_own <= 1e-12, label
    if ch.nu_lo:
        assert np.array_equal(chi[: ch.nu_lo], mirror_nu(chi[ch.nu_lo + 1:], n)), label
    return chi


# ---------------------------------------------------------------- cfg2, every point
def test_cubic_3orb_full_size_every_point_vs_oracle(T, oracle):
    """3 orbitals, 16^3 k, beta 40, 256 freqs, 127 nu, Kanamori U = 2.3, J = 0.4:
    the default chain against the oracle.  1 - chi0 U is nearly singular at
    some points (cond up to ~4e5), so the forward error of chi_RPA there is
    cond x rounding; the residual bound holds at every point regardless, and
    the forward error is bounded per point by 1e-10 + 1e-11 cond."""
    from tprf_amd import chain, workloads
    wl = workloads.cubic_3orb(seed=0)
    ch = chain.BubbleRPA(wl)
    ch.run()
    err, _ = ch.check_status()
    T.cuda.synchronize()
    assert err < 1e-6 and ch.herm_nu
    print("cfg2: oracle chain ...", flush=True)
    ref = oracle_chain(oracle, wl)
    canon = canonical_r(wl.dims)
    assert relerr(host(ch.g_tr)[:, canon], ref["g_tr"][:, canon]) <= TOL
    chi = check_chain_against(ch, ref["chi0_wk"], ref["chi_wk"], wl, "cfg2")
    n4 = wl.norb ** 4
    kappa = rpa_kappa(ref["chi0_wk"], wl.U)
    d = np.abs(chi - ref["chi_wk"]).reshape(-1, n4).max(1) / np.abs(ref["chi_wk"]).reshape(-1, n4).max(1)
    ratio = d / (1e-10 + 1e-11 * kappa)
    print(f"cfg2: max cond {kappa.max():.3e}, max per-point forward error {d.max():.2e}, "
          f"max error / (1e-10 + 1e-11 cond) {ratio.max():.3f}", flush=True)
    assert ratio.max() <= 1.0


def test_cubic_3orb_graph_replay_matches_eager(T):
    """The default chain captured as one HIP graph (status reset + five stage
    launches) and replayed four times, with the outputs overwritten by NaN and
    unrelated eager work between the replays: every replay reproduces the
    eager chain bit for bit, and the captured status reset clears the words."""
    from tprf_amd import chain, workloads
    wl = workloads.cubic_3orb(seed=0)
    ch = chain.BubbleRPA(wl)
    ch.run()
    err, _ = ch.check_status()
    assert err < 1e-6
    ref0, ref = ch.chi0_wk.clone(), ch.chi_wk.clone()
    ch.capture()
    for i in range(4):
        ch.chi0_wk.fill_(float("nan"))
        ch.chi_wk.fill_(float("nan"))
        x = T.randn(1 << 20, device="cuda:0", dtype=T.float64)
        assert float((x * x).sum()) > 0  # eager kernels and a synchronisation between replays
        ch.replay()
        err_i, _ = ch.check_status()
        T.cuda.synchronize()
        print(f"replay {i}: tail residual {err_i:.2e}", flush=True)
        assert err_i == err
        assert T.equal(ch.chi0_wk, ref0) and T.equal(ch.chi_wk, ref), f"replay {i}"
    ch.run()  # eager again after the replays
    T.cuda.synchronize()
    assert T.equal(ch.chi_wk, ref)


# ---------------------------------------------------------------- cfg3 uneven batch
def test_uneven_batch_48_every_lattice(T):
    """The 48 uneven lattices (8^3 / 12^3 / 16^3 k, 1-3 orbitals, 128/256/512
    freqs, beta 10/20/40, 127 nu) in one graph over 4 streams, every lattice
    against its own chain of one call per reference function."""
    from tprf_amd import chain, workloads
    from tprf_amd.batched import BatchedChain
    wls = workloads.uneven_batch(seed=0)
    bc = BatchedChain(wls)
    bc.capture(4)
    bc.replay()
    err, _ = bc.check_status()
    T.cuda.synchronize()
    assert err < 1e-6
    for i, wl in enumerate(wls):
        ref = chain.BubbleRPA(wl, fused=False)
        ref.run()
        ref.check_status()
        T.cuda.synchronize()
        label = f"lattice {i}: {wl.dims[0]}^3 k, n {wl.norb}, n_iw {wl.n_iw}, beta {wl.beta}"
        check_chain_against(bc.chains[i], host(ref.chi0_wk), host(ref.chi_wk), wl, label)
        del ref


def test_uneven_batch_reduced_vs_oracle(T, oracle):
    """Six uneven lattices (4^3 / 6^3 k, 1-3 orbitals, 32/64/128 freqs) batched
    in one graph against the oracle chain of each."""
    from tprf_amd import workloads
    from tprf_amd.batched import BatchedChain
    from tprf_amd.rpa_tensor import kanamori_charge_and_spin_quartic_interaction_tensors
    specs = [(4, 1, 64, 10.0), (6, 2, 32, 5.0), (4, 3, 128, 20.0), (6, 3, 64, 10.0), (4, 2, 128, 40.0),
             (6, 1, 32, 20.0)]
    rng = np.random.default_rng(11)
    wls = []
    for i, (Lk, n, n_iw, beta) in enumerate(specs):
        Td = workloads.cubic_hoppings(int(rng.integers(1 << 30)), n)
        k = workloads.kgrid((Lk, Lk, Lk))
        e = sum(2 * np.cos(k[d])[:, None, None] * Td[d][None] for d in range(3))
        _, U_s = kanamori_charge_and_spin_quartic_interaction_tensors(n, 1.2, 0.8, 0.2, 0.2)
        wls.append(workloads.Workload(f"r{i}", (Lk, Lk, Lk), n, beta, n_iw, 12, float(rng.uniform(-1, 1)),
                                      e.astype(np.complex128), U_s))
    bc = BatchedChain(wls)
    bc.capture(2)
    bc.replay()
    err, _ = bc.check_status()
    T.cuda.synchronize()
    assert err < 1e-6
    for i, wl in enumerate(wls):
        ref = oracle_chain(oracle, wl)
        check_chain_against(bc.chains[i], ref["chi0_wk"], ref["chi_wk"], wl, f"reduced lattice {i}")


# ---------------------------------------------------------------- cfg5 batched forward + adjoint
def test_batched_grad_48_sets_full_size(T):
    """48 two-orbital square lattices, 48x48 k, 256 freqs, beta 20, Q = (pi, pi):
    forward + adjoint of L = sum Re tr chi_RPA(Q, 0) in one graph, every set
    against the per-lattice autograd call (1e-12), two sets against the FP64
    torch reference (1e-9)."""
    from oracle.torch_ref import StaticChainRef
    from tprf_amd import autograd as A
    B, beta, n_iw, dims, q = 48, 20.0, 256, (48, 48), (24, 24, 0)
    g = np.random.default_rng(5)
    tp = T.tensor(g.uniform(-0.4, 0.0, B), dtype=T.float64, device="cuda")
    mu = T.tensor(g.uniform(-1.0, 1.0, B), dtype=T.float64, device="cuda")
    v = T.tensor(g.uniform(0.0, 0.3, B), dtype=T.float64, device="cuda")
    U = T.tensor(g.uniform(1.0, 3.0, B), dtype=T.float64, device="cuda")
    e_k = A.two_orbital_square_e_k(tp, v, dims)
    Uk = A.kanamori_spin_U(U, U / 6)
    batch = A.StaticRPA(beta, n_iw, dims, q, norb=2, nbatch=B)
    chi, ebar, mubar, Ubar = (x.clone() for x in batch.graphed_grad(e_k, mu, Uk))
    T.cuda.synchronize()
    one = A.StaticRPA(beta, n_iw, dims, q, norb=2)
    worst = 0.0
    for i in range(B):
        e1 = e_k[i].clone().requires_grad_(True)
        m1 = mu[i].clone().requires_grad_(True)
        U1 = Uk[i].clone().requires_grad_(True)
        c1 = A.chi_rpa_static(e1, m1, U1, one)
        A.trace_loss(c1).backward()
        d = max(relerr(host(chi[i]), host(c1)), relerr(host(ebar[i]), host(e1.grad)),
                relerr(host(Ubar[i]), host(U1.grad)), abs(float(mubar[i] - m1.grad)) / abs(float(m1.grad)))
        worst = max(worst, d)
        assert d <= 1e-12, i
    print(f"cfg5: 48 sets batched vs per lattice max rel {worst:.2e}", flush=True)
    ref = StaticChainRef(beta, n_iw, dims, q)
    for i in (0, 31):
        e_r = e_k[i].detach().cpu().clone().requires_grad_(True)
        mu_r = mu[i].detach().cpu().clone().requires_grad_(True)
        U_r = Uk[i].detach().cpu().clone().requires_grad_(True)
        c_r = ref.chi(e_r, mu_r, U_r)
        A.trace_loss(c_r).backward()
        print(f"cfg5 set {i}: chi rel {relerr(host(chi[i]), c_r.detach().numpy()):.2e}", flush=True)
        assert relerr(host(chi[i]), c_r.detach().numpy()) <= 1e-10
        assert relerr(host(ebar[i]), e_r.grad.numpy()) <= 1e-9
        assert relerr(host(Ubar[i]), U_r.grad.numpy()) <= 1e-9
        assert abs(float(mubar[i]) - float(mu_r.grad)) <= 1e-9 * abs(float(mu_r.grad))


def test_graph_keeps_its_workspace_when_scratch_grows(T):
    """A captured forward + backward graph owns its scratch: a later call that
    needs more scratch than the context had (a larger static bubble) does not
    change what the graph's replays compute."""
    from tprf_amd import autograd as A
    B, beta, n_iw, dims, q = 3, 10.0, 32, (8, 8, 1), (4, 4, 0)
    g = np.random.default_rng(3)
    tp = T.tensor(g.uniform(-0.4, 0.0, B), dtype=T.float64, device="cuda")
    mu = T.tensor(g.uniform(-1.0, 1.0, B), dtype=T.float64, device="cuda")
    v = T.tensor(g.uniform(0.0, 0.3, B), dtype=T.float64, device="cuda")
    U = T.tensor(g.uniform(1.0, 3.0, B), dtype=T.float64, device="cuda")
    e_k, Uk = A.two_orbital_square_e_k(tp, v, dims), A.kanamori_spin_U(U, U / 6)
    batch = A.StaticRPA(beta, n_iw, dims, q, norb=2, nbatch=B)
    first = [x.clone() for x in batch.graphed_grad(e_k, mu, Uk)]
    T.cuda.synchronize()
    big_dims = (32, 32)
    big = A.StaticRPA(beta, n_iw, big_dims, (16, 16, 0), norb=2, nbatch=8)
    tp8 = T.full((8,), -0.2, dtype=T.float64, device="cuda")
    e8 = A.two_orbital_square_e_k(tp8, tp8 * -1.0, big_dims)
    for rep in range(3):
        _ = A.chi_rpa_static(e8, T.zeros(8, dtype=T.float64, device="cuda"),
                             A.kanamori_spin_U(tp8 * -5, tp8 * -1), big)
        T.cuda.synchronize()
        again = batch.graphed_grad(e_k, mu, Uk)
        T.cuda.synchronize()
        for a, b in zip(again, first):
            assert T.equal(a, b), rep


# ---------------------------------------------------------------- driver (row a8 / f1)
def _square_2orb(nk, beta, n_iw):
    """The two-orbital square lattice of the reference's chi00_square_lattice.py:
    h_loc = [[-0.3, -0.5], [-0.5, 0.4]], nearest-neighbour T = -[[1, 0.23],
    [0.23, 0.5]], built with tb_dispersion."""
    from tprf_amd.lattice_utils import tb_dispersion
    h_loc = np.array([[-0.3, -0.5], [-0.5, 0.4]])
    Tn = -np.array([[1.0, 0.23], [0.23, 0.5]])
    hop = {(0, 0): h_loc, (0, 1): Tn, (0, -1): Tn, (1, 0): Tn, (-1, 0): Tn}
    return tb_dispersion(hop, (nk, nk, 1), 2)


@pytest.mark.parametrize("nw,save_memory", [(1, False), (8, False), (8, True)])
def test_imtime_bubble_chi0_wk_branches(T, oracle, nw, save_memory):
    """imtime_bubble_chi0_wk (lattice_utils.py:241-266): the static trapezoid
    (nw == 1), the stored product + tau -> nu, and the fused product
    (save_memory), each against the oracle's restatement of that branch."""
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshImFreq
    from tprf_amd.lattice_utils import imtime_bubble_chi0_wk
    beta, n_iw, nk = 5.0, 64, 6
    e = _square_2orb(nk, beta, n_iw)
    g_wk = L.lattice_dyson_g0_wk(mu=0.1, e_k=e, mesh=MeshImFreq(beta, "Fermion", n_iw))
    chi0 = imtime_bubble_chi0_wk(g_wk, nw=nw, save_memory=save_memory)
    dims = (nk, nk, 1)
    g_tr = oracle.fourier_w_to_t(oracle.fourier_lattice(oracle.lattice_dyson_g_wk(beta, n_iw, 0.1, e.numpy()), dims, -1),
                                 beta)
    if nw == 1:
        chi0_wr = oracle.chi0_w0r_from_grt_PH(g_tr, dims, beta)
    elif not save_memory:
        chi0_wr = oracle.fourier_t_to_w(oracle.chi0_tr_from_grt_PH(g_tr, dims), beta, nw)
    else:
        chi0_wr = oracle.chi0_wr_from_grt_PH(g_tr, dims, beta, nw)
    ref = oracle.fourier_lattice(chi0_wr, dims, +1)
    assert chi0.mesh[0].size == (1 if nw == 1 else 2 * nw - 1)
    assert relerr(host(chi0.data), ref) <= TOL


def test_chi0_save_memory_matches_unfused(T):
    """The reference's chi0_save_memory.py:26-30 case: two orbitals, 2x2 k,
    beta 1, 20 fermionic and 10 bosonic freqs (generic kernels: these meshes
    fit no pruned shape); fused and stored-product branches agree."""
    from tprf_amd import lattice as L
    from tprf_amd.gf import MeshImFreq
    from tprf_amd.lattice_utils import imtime_bubble_chi0_wk
    e = _square_2orb(2, 1.0, 20)
    g0 = L.lattice_dyson_g0_wk(mu=0.0, e_k=e, mesh=MeshImFreq(1.0, "Fermion", 20))
    a = host(imtime_bubble_chi0_wk(g0, nw=10, save_memory=False).data)
    b = host(imtime_bubble_chi0_wk(g0, nw=10, save_memory=True).data)
    assert np.allclose(a, b)
    assert relerr(b, a) <= 1e-12


@pytest.mark.parametrize("n_iw", [500, 100])
def test_default_chain_on_meshes_without_pruned_kernels(T, oracle, n_iw):
    """The default (symmetric) chain at n_iw = 500 (the mesh of
    chi00_square_lattice.py:44) and 100: L = 6 n_iw is no multiple of a power of
    two >= 2 n_iw, so omega -> tau runs the plain transform of every column;
    the results match the oracle."""
    from tprf_amd import chain, workloads
    from tprf_amd.rpa_tensor import kanamori_charge_and_spin_quartic_interaction_tensors
    e = _square_2orb(2, 20.0, n_iw)
    _, U_s = kanamori_charge_and_spin_quartic_interaction_tensors(2, 0.8, 0.5, 0.15, 0.15)
    wl = workloads.Workload("square2", (2, 2, 1), 2, 20.0, n_iw, 16, 0.0, e.numpy(), U_s)
    ch = chain.BubbleRPA(wl)
    assert ch.herm and ch.herm_nu and not ch.herm_w2t
    ch.run()
    err, _ = ch.check_status()
    T.cuda.synchronize()
    assert err < 1e-6
    ref = oracle_chain(oracle, wl)
    assert relerr(host(ch.g_tr), ref["g_tr"]) <= TOL
    check_chain_against(ch, ref["chi0_wk"], ref["chi_wk"], wl, f"n_iw {n_iw}")


# ---------------------------------------------------------------- error paths and limits
def _g_wr(values, beta, n_iw):
    import torch
    from tprf_amd.gf import Gf, MeshCycLat, MeshImFreq, MeshProduct
    d = torch.from_numpy(np.ascontiguousarray(values.reshape(2 * n_iw, 1, 1, 1))).cuda()
    return Gf(MeshProduct(MeshImFreq(beta, "Fermion", n_iw), MeshCycLat((1, 1, 1))), data=d)


def test_tail_fit_error_warning_and_m0(T):
    """fourier_matsubara.cpp:227-239: a tail-fit error >= 1e-3 raises (status
    -4), one above 1e-6 warns, and a nonzero zeroth moment warns."""
    from tprf_amd import lattice as L
    from tprf_amd._lib import TprfError
    beta, n_iw = 10.0, 64
    iw = 1j * np.pi * (2 * np.arange(-n_iw, n_iw) + 1) / beta
    rng = np.random.default_rng(0)
    noise = rng.normal(size=iw.shape) + 1j * rng.normal(size=iw.shape)
    g = 1.0 / (iw - 0.3)
    with pytest.raises(TprfError) as e:
        L.fourier_wr_to_tr(_g_wr(g + 1e-2 * noise, beta, n_iw))
    assert e.value.code == -4 and "1e-3" in str(e.value)
    with pytest.warns(UserWarning, match="greater than 1e-6"):
        L.fourier_wr_to_tr(_g_wr(g + 1e-5 * noise, beta, n_iw))
    with pytest.warns(UserWarning, match="tail is not zero"):
        L.fourier_wr_to_tr(_g_wr(g + 1e-3, beta, n_iw))
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        L.fourier_wr_to_tr(_g_wr(g, beta, n_iw))  # a clean pole: neither error nor warning


def test_singular_dyson_and_rpa(T):
    """A matrix (i w + mu) 1 - e_k - Sigma that is exactly zero, and an RPA
    point with chi0 U = 1, report the singular status (-6)."""
    import torch
    from tprf_amd import lattice as L
    from tprf_amd._lib import TprfError
    from tprf_amd.gf import Gf, MeshBrZone, MeshImFreq, MeshProduct
    beta, n_iw, mu = 4.0, 8, 0.25
    e = np.full((1, 1, 1), 0.5 + 0j)
    wn = np.pi * (2 * np.arange(-n_iw, n_iw) + 1) / beta
    sigma = ((mu - 0.5) + 1j * wn).reshape(-1, 1, 1)
    with pytest.raises(TprfError) as ex:
        L.lattice_dyson_g_wk(mu, Gf(MeshBrZone((1, 1, 1)), data=e, device="cpu"),
                             sigma_w=Gf(MeshImFreq(beta, "Fermion", n_iw), data=sigma, device="cpu"))
    assert ex.value.code == -6
    chi0 = torch.full((3, 1, 1, 1, 1, 1), 0.5, dtype=torch.complex128, device="cuda")
    mesh = MeshProduct(MeshImFreq(beta, "Boson", 2), MeshBrZone((1, 1, 1)))
    with pytest.raises(TprfError) as ex:
        L.solve_rpa_PH(Gf(mesh, data=chi0), np.full((1, 1, 1, 1), 2.0 + 0j))
    assert ex.value.code == -6
    # and a regular one passes
    L.solve_rpa_PH(Gf(mesh, data=chi0), np.full((1, 1, 1, 1), 1.0 + 0j))


def test_unsupported_sizes_are_refused(T):
    """Every size outside the kernels' limits returns TPRF_ERR_UNSUPPORTED (-9)
    from the host checks, before any launch: orbitals above the limits of
    Dyson, RPA, W, Lindhard and the fused bubble (8), the RPA and Dyson
    adjoints (4), and Matsubara-axis meshes with no pruned kernel and ntau
    beyond one in-place column in LDS."""
    import torch
    from tprf_amd import _lib
    c = _lib.ctx()
    x = torch.zeros(16, dtype=torch.complex128, device="cuda")
    p = _lib.ptr(x)
    d = _lib.dims3((2, 1, 1))
    lib = _lib.lib()
    cases = {
        "dyson n=9": lambda: lib.tprf_lattice_dyson_g_wk(c, 1.0, 4, 0.0, 1, 9, p, 0, None, p),
        "rpa n=9": lambda: lib.tprf_solve_rpa_PH(c, 1, 9, p, p, p),
        "rpa_herm n=9": lambda: lib.tprf_solve_rpa_PH_herm(c, 1, 1, 9, p, p, p),
        "W n=9": lambda: lib.tprf_dynamical_screened_interaction_W(c, 1, 1, 9, p, p, 0, p),
        "chi0_wr n=9": lambda: lib.tprf_chi0_wr_from_grt_PH(c, 100.0, 61, d, 9, 4, p, p),
        "lindhard n=9": lambda: lib.tprf_lindhard_chi00(c, 1.0, 0.0, 1, p, d, 9, p, p),
        "rpa adjoint n=5": lambda: lib.tprf_solve_rpa_PH_adjoint(c, 1, 5, p, p, None, p, None, None),
        "dyson adjoint n=5": lambda: lib.tprf_lattice_dyson_g_wk_adjoint(c, 4, 1, 5, 1, p, p, p,
                                                                          _lib.dptr(x.real.contiguous())),
        # n_max 1700: no pruned kernel (M = 4096 > 1024), and ntau = 10201 needs
        # 10201 x 16 bytes > 152 KiB for even one in-place column
        "w2t n_max=1700": lambda: lib.tprf_fourier_wr_to_tr(c, 0, 100.0, 1700, -1, 1, p, p, p),
        "t2w ntau=10201": lambda: lib.tprf_fourier_tr_to_wr(c, 0, 100.0, 10201, 100, 1, p, p),
        "chi0_wr ntau=10201": lambda: lib.tprf_chi0_wr_from_grt_PH(c, 100.0, 10201, d, 1, 16, p, p),
    }
    for name, f in cases.items():
        st = f()
        msg = lib.tprf_ctx_last_error(c).decode()
        print(f"{name}: {st} {msg}", flush=True)
        assert st == -9, name


def test_rpa_herm_without_exchange_symmetry(T, oracle):
    """tprf_solve_rpa_PH_herm needs only U_abcd = conj(U_dcba): a Hermitian U
    with U_abcd != U_badc still gives the full solve's chi at every frequency."""
    from tprf_amd import _lib
    import torch
    rng = np.random.default_rng(9)
    n, nb, nk = 2, 3, 5
    a = rng.normal(size=(n,) * 4) + 1j * rng.normal(size=(n,) * 4)
    U = 0.5 * (a + np.conj(a.transpose(3, 2, 1, 0))) * 0.4
    assert np.abs(U - U.transpose(1, 0, 3, 2)).max() > 1e-2
    half = (rng.normal(size=(nb, nk) + (n,) * 4) + 1j * rng.normal(size=(nb, nk) + (n,) * 4)) * 0.1
    half[0] = 0.5 * (half[0] + mirror_nu(half[:1], n)[0])  # nu = 0 is its own mirror
    chi0 = np.concatenate([mirror_nu(half[1:], n), half], axis=0)
    assert relerr(mirror_nu(chi0, n), chi0) <= 1e-15
    ref = oracle.solve_rpa_PH(chi0, U)
    x = torch.from_numpy(chi0).cuda()
    u = torch.from_numpy(U).cuda()
    out = torch.empty_like(x)
    c = _lib.ctx()
    _lib.call("tprf_solve_rpa_PH_herm", c, nb, nk, n, _lib.ptr(x), _lib.ptr(u), _lib.ptr(out))
    _lib.tail_status(c)
    assert relerr(host(out), ref) <= 1e-12


def test_roctx_trace_option(T):
    """TPRF_OPT_TRACE = 1 opens a roctx range per entry point (the rocprofiler
    SDK's roctx library is part of the ROCm image); the traced chain gives
    bit-identical results, and the option switches off again."""
    import torch
    from tprf_amd import _lib, chain, workloads
    wl = workloads.plain_square(n=8, n_iw=16, nb=8)
    ch = chain.BubbleRPA(wl)
    ch.run()
    torch.cuda.synchronize()
    ref = ch.chi_wk.clone()
    c = _lib.ctx()
    _lib.set_trace(c, True)
    try:
        ch.run()
        torch.cuda.synchronize()
    finally:
        _lib.set_trace(c, False)
    assert torch.equal(ch.chi_wk, ref)
