"""Committed golden fixtures (tests/golden, written by make_golden.py from the
oracle): the oracle must keep reproducing them (CPU), and the HIP chain must
match them at 1e-10 (GPU)."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import relerr

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _make_golden():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod
CASES = ["chain_square8.npz", "chain_cubic2orb.npz"]


def _load(name):
    return dict(np.load(os.path.join(HERE, name), allow_pickle=False))


def _workload(f):
    from tprf_amd.workloads import Workload
    return Workload("golden", tuple(int(x) for x in f["dims"]), f["e_k"].shape[-1], float(f["beta"]), int(f["n_iw"]),
                    int(f["nb"]), float(f["mu"]), f["e_k"], f["U"])


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(oracle, name):
    chain = _make_golden().chain
    f = _load(name)
    dims = tuple(int(x) for x in f["dims"])
    g_tr, chi0, chi = chain(oracle, f["e_k"], dims, float(f["beta"]), int(f["n_iw"]), int(f["nb"]), float(f["mu"]),
                            f["U"])
    assert relerr(g_tr[::8], f["g_tr8"]) <= 1e-13
    assert relerr(chi0, f["chi0_wk"]) <= 1e-13
    assert relerr(chi, f["chi_wk"]) <= 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
@pytest.mark.parametrize("name", CASES)
def test_gpu_chain_matches_golden(name, fused):
    """The chain of one call per reference function, and the fused chain
    (which for these Hermitian e_k and U transforms chi0 at nu >= 0 only and
    forms chi_RPA at nu < 0 from the conjugate-transpose identity)."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    from tprf_amd import chain
    f = _load(name)
    ch = chain.BubbleRPA(_workload(f), fused=fused)
    ch.run()
    err, _ = ch.check_status()
    torch.cuda.synchronize()
    assert err < 1e-6
    # the fused chain stores G(tau, r) for r <= -r only
    from tprf_amd.gf import MeshCycLat
    lat = MeshCycLat(ch.wl.dims)
    cols = [r for r in range(ch.wl.nk) if r <= lat.minus(r)] if fused else list(range(ch.wl.nk))
    assert relerr(ch.g_tr.cpu().numpy()[::8][:, cols], f["g_tr8"][:, cols]) <= 1e-10
    assert fused == ch.herm_nu
    assert ch.nu_lo == (int(f["nb"]) - 1 if fused else 0)
    assert relerr(ch.chi0_wk.cpu().numpy(), f["chi0_wk"][ch.nu_lo:]) <= 1e-10
    assert relerr(ch.chi_wk.cpu().numpy(), f["chi_wk"]) <= 1e-10
