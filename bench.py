"""Benchmark of the lattice bubble + RPA forward chain on one MI355X.

One step = one eval: e_k -> G(iw,k) -> G(iw,r) -> G(tau,r) -> chi0(i nu,r) (fused
product + tau->nu) -> chi0(i nu,q) -> chi_RPA(i nu,q) for the 3-orbital
simple-cubic workload (16^3 k, beta 40, 256 positive fermionic frequencies,
ntau 1537, 127 bosonic frequencies, Kanamori U = 2.3, J = 0.4), all complex
FP64, inputs resident in HBM.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Per-stage durations come from HIP events recorded on the stream the kernels
run on (torch's current stream, which the library context uses).  The CPU
baseline times the oracle (C/OpenMP restatement of the reference algorithm)
on one eval of the same workload on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "χ0+RPA evals/s (3-orb 16³k, 256 iω) at 1e-10 rel; grad/fwd time; batched/plain"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# HBM traffic of the dominant kernel per launch, from two rocprofv3 --pmc runs
# (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch) of `bench.py --only calibration`
# on the MI355X.  The same run's 1 GiB device copy reads FETCH_SIZE = 0.5 GiB and
# its 1 GiB fill WRITE_SIZE = 1 GiB, so fetches are scaled by 2.0 and writes by 1.0.
# The kernel is the fused bubble (Hermitian form, the default chain's) in the
# workgroup shape the autotuner kept there.
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
MEASURED_TRAFFIC = {"chi0_wr": {"kernel": "k_bubble_wg<3>", "fetch_kib": 235979, "write_kib": 335953}}


def bubble_executed_flops(wl):
    """FP64 FLOPs the fused bubble kernel executes (Hermitian form, one
    workgroup per canonical (r, -r) pair): for each canonical column (orbital
    pairs a < b with every (c, d), a == b with c <= d) and tau step j1 < L / 128,
    128 complex products (6 FLOPs each), one length-128 DFT (5 N log2 N) and 128
    twiddled accumulations (complex multiply-add, 8 FLOPs)."""
    import numpy as np
    n = wl.norb
    ncol = n * n * n * (n - 1) // 2 + n * n * (n + 1) // 2
    d = np.array(wl.dims)
    idx = np.indices(d).reshape(3, -1)
    minus = ((-idx) % d[:, None])
    lin = lambda v: (v[0] * d[1] + v[1]) * d[2] + v[2]
    ncanon = int(np.count_nonzero(lin(idx) <= lin(minus)))
    steps = (wl.ntau - 1) // 128
    return ncanon * ncol * steps * (6 * 128 + 5 * 128 * 7 + 8 * 128)


def stage_model(wl):
    """Algorithmic bytes (each materialised API tensor written once and read once,
    chi0(tau, r) never stored) and FLOPs (5 N log2 N per complex FFT, 6 per complex
    multiply) per eval and stage."""
    import math
    n, nk = wl.norb, wl.nk
    c = 16
    nw = 2 * wl.n_iw
    L = wl.ntau - 1
    nnu = wl.n_nu
    g_wk = nw * nk * n * n * c
    g_tr = wl.ntau * nk * n * n * c
    chi = nnu * nk * n ** 4 * c
    ek = nk * n * n * c
    fft3 = 5 * nk * math.log2(nk)
    return {
        "dyson": (ek + g_wk, 0.0),
        "wk_to_wr": (2 * g_wk, nw * n * n * fft3),
        # Dyson fused into k -> r: the two stages' algorithmic bytes (G(iw, k)
        # counted as written and read, as for the unfused pair)
        "dyson_wr": (ek + 3 * g_wk, nw * n * n * fft3),
        "wr_to_tr": (g_wk + g_tr, nk * n * n * 5 * L * math.log2(L)),
        "chi0_wr": (g_tr + chi, nk * n ** 4 * (5 * L * math.log2(L) + 6 * wl.ntau)),
        "wr_to_wk": (2 * chi, nnu * n ** 4 * fft3),
        "rpa": (2 * chi, nnu * nk * 8 * ((n * n) ** 3 * 2 + (n * n) ** 3 // 3)),
    }


def cpu_threads():
    """Host threads for the CPU baseline: every core this process may run on
    (its affinity mask), capped at the 16-core host share of one GPU on the
    MI355X boxes, which OMP_NUM_THREADS states there."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(avail, env) if env > 0 else avail


def cpu_baseline(wl, chain_result, reps=3):
    """``reps`` evals of the oracle (the C/OpenMP restatement of the reference
    algorithm, one call per stage as the reference decomposes it) on the host,
    per-stage wall times of the median eval, and the GPU result's relative
    error against it."""
    import numpy as np
    from oracle import oracle as O
    threads = cpu_threads()
    O.set_num_threads(threads)
    runs = []
    for _ in range(reps):
        st = {}
        t = time.perf_counter()
        g_wk = O.lattice_dyson_g_wk(wl.beta, wl.n_iw, wl.mu, wl.e_k)
        st["dyson"] = time.perf_counter() - t
        t = time.perf_counter()
        g_wr = O.fourier_lattice(g_wk, wl.dims, -1)
        st["wk_to_wr"] = time.perf_counter() - t
        del g_wk
        t = time.perf_counter()
        g_tr = O.fourier_w_to_t(g_wr, wl.beta)
        st["wr_to_tr"] = time.perf_counter() - t
        del g_wr
        t = time.perf_counter()
        chi0_wr = O.chi0_wr_from_grt_PH(g_tr, wl.dims, wl.beta, wl.nb)
        st["chi0_wr"] = time.perf_counter() - t
        del g_tr
        t = time.perf_counter()
        chi0_wk = O.fourier_lattice(chi0_wr, wl.dims, +1)
        st["wr_to_wk"] = time.perf_counter() - t
        del chi0_wr
        t = time.perf_counter()
        chi = O.solve_rpa_PH(chi0_wk, wl.U)
        st["rpa"] = time.perf_counter() - t
        runs.append(st)
    totals = sorted((sum(r.values()), i) for i, r in enumerate(runs))
    dt, med = totals[len(totals) // 2]
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    n = wl.norb
    X = chi0_wk.reshape(-1, n * n, n * n)
    MU = wl.U.reshape(n * n, n * n)
    s = np.linalg.svd(np.eye(n * n)[None] - X @ MU, compute_uv=False)
    kappa = s[:, 0] / s[:, -1]
    good = kappa <= 100
    n4 = n ** 4
    got_chi0, got_chi, lo = chain_result
    # relative residual of (1 - chi0 U) chi = chi0 per point with the oracle's chi0
    # (the per-point scale floored at 1e-3 of the largest, the transforms' rounding floor)
    Y = got_chi.reshape(-1, n * n, n * n)
    R = Y - X - X @ (MU @ Y)
    nx = np.linalg.norm(X, axis=(1, 2))
    den = np.maximum(nx, 1e-3 * nx.max()) * (1 + np.linalg.norm(MU) * np.linalg.norm(Y, axis=(1, 2)))
    parity = {
        "chi0_wk_rel": rel(got_chi0, chi0_wk[lo:]),
        "chi0_wk_frequencies": "nu >= 0 (the symmetric chain forms nu < 0 of chi_RPA only)" if lo else "all",
        "chi_rpa_residual_max_every_point": float((np.linalg.norm(R, axis=(1, 2)) / den).max()),
        "chi_rpa_rel_all": rel(got_chi, chi),
        "chi_rpa_rel_cond_le_100": rel(got_chi.reshape(-1, n4)[good], chi.reshape(-1, n4)[good]),
        "rpa_min_sigma": float(s[:, -1].min()),
        "rpa_max_cond": float(kappa.max()),
    }
    try:
        cpu = open("/proc/cpuinfo").read().split("model name")[1].split("\n")[0].strip(" :\t")
    except Exception:
        cpu = "unknown"
    base = {"value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "cpu_model": cpu, "nproc": os.cpu_count(),
            "stage_s": {k: round(v, 4) for k, v in runs[med].items()},
            "sample": f"{reps} full evals of the same workload (oracle: C/OpenMP restatement of the reference "
                      f"algorithm, {threads} threads on {cpu}), median {dt:.2f} s per eval"}
    return base, parity


def _timed(f, reps, warm=2):
    """Mean wall time (ms) of f() over reps calls, the stream drained on both sides."""
    import torch
    for _ in range(warm):
        f()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        f()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) * 1e3 / reps


def grad_leg(reps=20, route="fused"):
    """grad/fwd at the differentiation workload: single-band square lattice 64x64 k,
    beta 20, 512 fermionic frequencies, t = 1, t' = -0.3, mu = -0.5, U = 2.0,
    L = Re tr chi_RPA(Q = (pi, pi), i nu = 0); wall(forward + backward) / wall(forward)."""
    import torch
    from tprf_amd import autograd as A
    dims, q = (64, 64), (32, 32, 0)
    spec = A.StaticRPA(20.0, 512, dims, q, norb=1, route=route)
    x0 = {"t": 1.0, "tp": -0.3, "mu": -0.5, "U": 2.0}
    leaves = {k: torch.tensor(v, dtype=torch.float64, device="cuda", requires_grad=True) for k, v in x0.items()}

    def chi():
        return A.chi_rpa_static(A.square_e_k(leaves["t"], leaves["tp"], dims), leaves["mu"],
                                A.hubbard_U(leaves["U"]), spec)

    def fwd():
        with torch.no_grad():
            return A.trace_loss(chi())

    def fwd_bwd():
        for v in leaves.values():
            v.grad = None
        A.trace_loss(chi()).backward()

    tf = _timed(fwd, reps)
    tfb = _timed(fwd_bwd, reps)
    return {"workload": "square 64x64 k, beta 20, 512 fermionic freqs, t 1, t' -0.3, mu -0.5, U 2.0, "
                        "L = Re tr chi_RPA((pi,pi), 0)", "route": route,
            "fwd_ms": tf, "fwd_bwd_ms": tfb, "grad_over_fwd": tfb / tf,
            "grad": {k: float(v.grad) for k, v in leaves.items()},
            "L": float(fwd())}


def batched_leg(reps=3, route="fused"):
    """batched/plain at the batched forward+adjoint workload: 48 two-orbital square
    lattices, 48x48 k, 256 fermionic frequencies, beta 20, t' ~ U[-0.4, 0],
    mu ~ U[-1, 1], inter-orbital hopping ~ U[0, 0.3], U ~ U[1, 3], J = U/6,
    chi_RPA(Q = (pi, pi), 0) and dL/d(e_k, mu, U) for every set.  Plain: a loop of
    the differentiable per-lattice call; batched: one HIP-graph replay."""
    import numpy as np
    import torch
    from tprf_amd import autograd as A
    B, dims, q = 48, (48, 48), (24, 24, 0)
    g = np.random.default_rng(5)
    tp = torch.tensor(g.uniform(-0.4, 0.0, B), dtype=torch.float64, device="cuda")
    mu = torch.tensor(g.uniform(-1.0, 1.0, B), dtype=torch.float64, device="cuda")
    v = torch.tensor(g.uniform(0.0, 0.3, B), dtype=torch.float64, device="cuda")
    U = torch.tensor(g.uniform(1.0, 3.0, B), dtype=torch.float64, device="cuda")
    e_k = A.two_orbital_square_e_k(tp, v, dims)
    Uk = A.kanamori_spin_U(U, U / 6)
    one = A.StaticRPA(20.0, 256, dims, q, norb=2, route=route)
    res = []

    def plain():
        res.clear()
        for i in range(B):
            e1 = e_k[i].clone().requires_grad_(True)
            m1 = mu[i].clone().requires_grad_(True)
            U1 = Uk[i].clone().requires_grad_(True)
            c = A.chi_rpa_static(e1, m1, U1, one)
            A.trace_loss(c).backward()
            res.append((c.detach(), e1.grad, m1.grad, U1.grad))

    batch = A.StaticRPA(20.0, 256, dims, q, norb=2, nbatch=B, route=route)
    tp_ms = _timed(plain, reps, warm=1)
    tb_ms = _timed(lambda: batch.graphed_grad(e_k, mu, Uk), reps + 2, warm=1)
    chi, ebar, mubar, Ubar = batch.graphed_grad(e_k, mu, Uk)
    torch.cuda.synchronize()
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max())
    diff = max(max(rel(chi[i], res[i][0]), rel(ebar[i], res[i][1]), rel(Ubar[i], res[i][3]),
                   abs(float(mubar[i] - res[i][2])) / abs(float(res[i][2]))) for i in range(B))
    del batch
    return {"workload": "48 two-orbital square lattices 48x48 k, 256 fermionic freqs, beta 20, forward + adjoint "
                        "of chi_RPA((pi,pi), 0)", "route": route,
            "plain_ms": tp_ms, "batched_ms": tb_ms, "batched_over_plain": tp_ms / tb_ms,
            "max_rel_diff_batched_vs_plain": diff}


def screened_w_leg(reps=20):
    """dynamical_screened_interaction_W at the chain's size (3 orbitals, 16^3 k,
    127 bosonic frequencies, static V_k): mean kernel time over `reps` calls by
    HIP events, and its HBM rate on the bytes it must move (PI and W once,
    V_k once)."""
    import numpy as np
    import torch
    from tprf_amd import _lib
    n, nk, nw = 3, 16 ** 3, 127
    rng = np.random.default_rng(3)
    shape = (nw, nk, n, n, n, n)
    pi = torch.from_numpy((rng.normal(size=shape) + 1j * rng.normal(size=shape)) * 0.05).cuda()
    v = torch.from_numpy((rng.normal(size=shape[1:]) + 1j * rng.normal(size=shape[1:])) * 0.3).cuda()
    w = torch.empty_like(pi)
    c = _lib.ctx()
    call = lambda: _lib.call("tprf_dynamical_screened_interaction_W", c, nw, nk, n, _lib.ptr(pi), _lib.ptr(v), 0,
                             _lib.ptr(w))
    for _ in range(3):
        call()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        call()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    nbytes = 2 * pi.numel() * 16 + v.numel() * 16
    _lib.tail_status(c)
    return {"workload": "3 orbitals, 16^3 k, 127 bosonic freqs, static V_k", "ms": ms,
            "points_per_s": nw * nk / (ms * 1e-3), "hbm_gbs": nbytes / (ms * 1e-3) / 1e9,
            "hbm_frac": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def _rpa_residual_gpu(chi0, U, chi):
    """Largest per-point relative residual of (1 - chi0 U) chi = chi0 (PH grouping
    M[(a,b),(c,d)] = X_abcd; the per-point scale floored at 1e-3 of the largest),
    on the GPU with torch matrix products of the results."""
    import torch
    n = U.shape[0]
    X = chi0.reshape(-1, n * n, n * n)
    Y = chi.reshape(-1, n * n, n * n)
    MU = U.reshape(n * n, n * n)
    R = Y - X - torch.matmul(X, torch.matmul(MU, Y))
    nx = torch.linalg.matrix_norm(X)
    den = torch.clamp(nx, min=1e-3 * float(nx.max())) * (1 + float(torch.linalg.matrix_norm(MU)) *
                                                        torch.linalg.matrix_norm(Y))
    return float((torch.linalg.matrix_norm(R) / den).max())


def uneven_leg(reps=3):
    """batched/plain at the uneven batch: 48 lattices drawn from the seed with
    L^3 k (L in {8, 12, 16}), norb in {1, 2, 3}, n_iw in {128, 256, 512}, beta in
    {10, 20, 40}, mu ~ U[-1, 1], 64 bosonic frequencies, full forward chi0 + RPA.
    Plain: the drop-in call sequence per lattice (lattice_dyson_g0_wk,
    fourier_wk_to_wr, fourier_wr_to_tr, chi0_wr_from_grt_PH, chi_wk_from_chi_wr,
    solve_rpa_PH), each allocating its output; also the per-lattice loop of the
    batch's own chains (same kernels, no graph).  Batched: one HIP-graph replay
    of every lattice's resident chain, stream count autotuned.  Every lattice's
    chi_RPA is checked against its drop-in result."""
    import torch
    from tprf_amd import workloads, lattice as Lt
    from tprf_amd.batched import BatchedChain
    from tprf_amd.gf import Gf, MeshBrZone, MeshImFreq
    wls = workloads.uneven_batch(seed=0)
    ins = [(Gf(MeshBrZone(w.dims), data=torch.from_numpy(w.e_k).cuda()), torch.from_numpy(w.U).cuda()) for w in wls]

    def plain_one(w, e_k, U, keep=False):
        g_wk = Lt.lattice_dyson_g0_wk(mu=w.mu, e_k=e_k, mesh=MeshImFreq(w.beta, "Fermion", w.n_iw))
        g_wr = Lt.fourier_wk_to_wr(g_wk)
        del g_wk
        g_tr = Lt.fourier_wr_to_tr(g_wr)
        del g_wr
        chi0_wr = Lt.chi0_wr_from_grt_PH(g_tr, nw=w.nb)
        del g_tr
        chi0_wk = Lt.chi_wk_from_chi_wr(chi0_wr)
        del chi0_wr
        chi = Lt.solve_rpa_PH(chi0_wk, U)
        return (chi0_wk.data, chi.data) if keep else chi

    def plain():
        for w, (e_k, U) in zip(wls, ins):
            plain_one(w, e_k, U)

    tp_ms = _timed(plain, reps, warm=1)
    bc = BatchedChain(wls)
    per_streams = bc.autotune((1, 2, 4, 8))
    tb_ms = _timed(bc.replay, reps, warm=1)
    te_ms = _timed(bc.run_eager, reps, warm=1)
    rel, res = 0.0, 0.0
    for w, (e_k, U), ch in zip(wls, ins, bc.chains):
        c0, c = plain_one(w, e_k, U, keep=True)
        rel = max(rel, float((ch.chi_wk - c).abs().max() / c.abs().max()))
        res = max(res, _rpa_residual_gpu(c0, U, ch.chi_wk))
        del c0, c
    out = {"workload": "48 uneven cubic lattices (8^3/12^3/16^3 k, 1-3 orbitals, 128/256/512 freqs, beta 10/20/40), "
                       "forward chi0 + RPA over 127 bosonic freqs",
           "plain_ms": tp_ms, "batched_ms": tb_ms, "batched_over_plain": tp_ms / tb_ms,
           "same_chain_loop_ms": te_ms, "batched_over_same_chain_loop": te_ms / tb_ms,
           "graph_ms_by_streams": per_streams, "streams": bc.nstreams,
           "max_rel_diff_all_lattices": rel, "max_rpa_residual_all_lattices": res}
    del bc
    return out


def per_function_leg(wl, steps, warmup):
    """The same eval as one call per reference function (no fusion, no
    symmetry shortcuts): the drop-in path a caller of the six functions gets."""
    import torch
    from tprf_amd import chain
    ch = chain.BubbleRPA(wl, device="cuda:0", fused=False)
    for _ in range(warmup):
        ch.run()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(steps):
        ch.run()
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    del ch
    return {"stages": list(chain.BubbleRPA.API_STAGES), "ms_per_step": ms, "evals_per_s": 1e3 / ms}


def calibration_run():
    """Two chain evals and one plain device copy of exactly 1 GiB (the reference
    figure for the FETCH_SIZE / WRITE_SIZE counters)."""
    import torch
    from tprf_amd import chain, workloads
    ch = chain.BubbleRPA(workloads.cubic_3orb(seed=0))
    for _ in range(2):
        ch.run()
    a = torch.empty(64 * 1024 * 1024, dtype=torch.complex128, device="cuda")
    a.fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-extra", action="store_true", help="skip the grad/fwd and batched/plain legs")
    ap.add_argument("--roctx", action="store_true",
                    help="open a roctx range per library entry point (for rocprofv3 --marker-trace)")
    ap.add_argument("--only", choices=["grad", "batched", "uneven", "screened_w", "calibration", "grad_matsubara",
                                       "batched_matsubara"],
                    help="run one leg alone and print its JSON (for profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the chain would shard "
              "over GPUs by all-to-all transposes between its stages, but is built and measured on one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from tprf_amd import chain, workloads

    torch.cuda.set_device(0)
    if args.roctx:
        from tprf_amd import _lib
        _lib.set_trace(_lib.ctx(0))
    if args.only:
        if args.only == "calibration":
            calibration_run()
            print(json.dumps({"calibration": "done"}))
        else:
            print(json.dumps({"grad": grad_leg, "batched": batched_leg, "uneven": uneven_leg,
                              "screened_w": screened_w_leg,
                              "grad_matsubara": lambda: grad_leg(route="matsubara"),
                              "batched_matsubara": lambda: batched_leg(route="matsubara")}[args.only]()))
        return
    wl = workloads.cubic_3orb(seed=args.seed)
    ch = chain.BubbleRPA(wl, device="cuda:0")
    for _ in range(args.warmup):
        ch.run()
    fit_err, _ = ch.check_status()
    torch.cuda.synchronize()

    stages = ch.STAGES
    ev = {s: [] for s in stages}
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for _ in range(args.steps):
        ch.stage_reset()
        for s in stages:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            ch.stage(s)
            b.record()
            ev[s].append((a, b))
    t1.record()
    torch.cuda.synchronize()
    total_ms = t0.elapsed_time(t1)
    ms_per_step = total_ms / args.steps
    stage_ms = {s: sum(a.elapsed_time(b) for a, b in ev[s]) / args.steps for s in stages}
    # per-eval durations (first stage's start to last stage's end), for the median
    step_ms = sorted(ev[stages[0]][i][0].elapsed_time(ev[stages[-1]][i][1]) for i in range(args.steps))
    median_ms = step_ms[len(step_ms) // 2] if len(step_ms) % 2 else 0.5 * (step_ms[len(step_ms) // 2 - 1] +
                                                                            step_ms[len(step_ms) // 2])
    fit_err, _ = ch.check_status()
    eager_ms = ms_per_step

    # The headline: the same eval captured once as a HIP graph (status reset and
    # the five stage launches) and replayed, i.e. without host launch gaps; the
    # per-stage times above come from the eager loop with events between stages.
    ch.capture()
    for _ in range(args.warmup):
        ch.replay()
    torch.cuda.synchronize()
    g0 = torch.cuda.Event(enable_timing=True)
    g1 = torch.cuda.Event(enable_timing=True)
    g0.record()
    for _ in range(args.steps):
        ch.replay()
    g1.record()
    torch.cuda.synchronize()
    ms_per_step = g0.elapsed_time(g1) / args.steps
    graph_fit_err, _ = ch.check_status()
    fit_err = max(fit_err, graph_fit_err)

    model = stage_model(wl)
    dom = max(stage_ms, key=stage_ms.get)
    dom_bytes, dom_flops = model[dom]
    achieved = dom_bytes / (stage_ms[dom] * 1e-3) / 1e9
    meas = MEASURED_TRAFFIC.get(dom)
    traffic = None
    if meas:
        traffic = 1024.0 * (FETCH_SCALE * meas["fetch_kib"] + WRITE_SCALE * meas["write_kib"])
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "algorithmic_bytes": dom_bytes, "avg_ms": stage_ms[dom],
                "traffic_note": (f"rocprofv3 FETCH_SIZE x{FETCH_SCALE} + WRITE_SIZE x{WRITE_SCALE} of "
                                 f"{meas['kernel']}, profiles/pmc_write_fetch_chain_cubic_3orb.md") if meas else None}
    # FP64 rate of the dominant kernel from the FLOPs it executes (for the
    # bubble: the pruned transform over the canonical columns); the algorithmic
    # count of the stage (full transforms of every column) is given beside it
    exec_flops = bubble_executed_flops(wl) if dom == "chi0_wr" else dom_flops
    fp64 = {"kernel": dom, "achieved": exec_flops / (stage_ms[dom] * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "flops": exec_flops,
            "flops_note": ("executed: canonical columns x tau steps x (128 products + DFT-128 + 128 twiddled sums)"
                           if dom == "chi0_wr" else "algorithmic"),
            "algorithmic_flops": dom_flops,
            "algorithmic_achieved": dom_flops / (stage_ms[dom] * 1e-3) / 1e12}
    fp64["frac"] = fp64["achieved"] / FP64_PEAK_TFLOPS
    # per-eval totals over the reference's six functions (the fused stage's
    # entry repeats the first two)
    api = [model[k] for k in chain.BubbleRPA.API_STAGES]
    chain_bytes = sum(v[0] for v in api)
    floor_ms = sum(max(b / (HBM_PEAK_GBS * 1e9), f / (FP64_PEAK_TFLOPS * 1e12)) for b, f in api) * 1e3
    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_step,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "timing": "one captured HIP graph replay per eval (status reset + 5 stage launches)",
        "ms_per_step_eager": eager_ms,
        "ms_per_step_median": median_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded hoppings, no external data)",
        "config": {"workload": "cubic_3orb: 3-orbital simple cubic, 16^3 k, beta 40, 256 fermionic freqs "
                               "(512 stored), ntau 1537, 127 bosonic freqs, Kanamori U 2.3 J 0.4 (9x9 RPA solves)",
                   "seed": args.seed, "mu": wl.mu, "tail_fit_residual": fit_err},
        "stage_ms": stage_ms,
        "chain_hbm_frac": chain_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        # per stage the larger of its HBM time (algorithmic bytes at 8 TB/s) and its
        # FP64 time (algorithmic FLOPs at 78.6 TFLOP/s), summed over the chain
        "chain_floor": {"floor_ms": floor_ms, "frac": floor_ms / ms_per_step},
        "roofline": roofline,
        "roofline_fp64": fp64,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        torch.cuda.synchronize()
        # the symmetric chain stores chi0(nu, q) at nu >= 0 only (rows nb - 1 ..)
        got = (ch.chi0_wk.cpu().numpy(), ch.chi_wk.cpu().numpy(), ch.nu_lo)
        base, parity = cpu_baseline(wl, got)
        out["cpu_baseline"] = base
        out["parity_vs_oracle"] = parity
    if not args.no_extra:
        del ch
        torch.cuda.empty_cache()
        out["per_function_chain"] = per_function_leg(wl, args.steps, args.warmup)
        torch.cuda.empty_cache()
        out["grad_fwd"] = grad_leg()
        torch.cuda.empty_cache()
        out["batched"] = batched_leg()
        torch.cuda.empty_cache()
        out["batched_uneven"] = uneven_leg()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
