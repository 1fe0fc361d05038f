"""CPU checks of the C ABI: the library loads, exports every function that
include/tprf_hip.h declares, the ctypes binding covers exactly those, and the
product path refuses to run without a GPU (no CPU fallback)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "tprf_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tprf_[a-zA-Z0-9_]+)\s*\(", text)))


def test_header_declares_hot_path():
    names = declared()
    for fn in ("tprf_lattice_dyson_g_wk", "tprf_fourier_wk_to_wr", "tprf_fourier_wr_to_wk", "tprf_fourier_wr_to_tr",
               "tprf_fourier_tr_to_wr", "tprf_chi0_tr_from_grt_PH", "tprf_chi0_wr_from_grt_PH",
               "tprf_chi0_w0r_from_grt_PH", "tprf_chi_w0r_from_chi_tr", "tprf_solve_rpa_PH"):
        assert fn in names


def test_library_exports_every_declared_symbol():
    import torch  # noqa: F401  (load torch's HIP runtime first, as the package does)
    path = os.path.join(ROOT, "tprf_amd", "libtprf_hip.so")
    assert os.path.exists(path), "build first: make -C tprf_amd/csrc"
    lib = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_binding_matches_header():
    from tprf_amd import _lib
    assert sorted(_lib._SIGNATURES) == declared()


def test_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU visible")
    from tprf_amd import lattice, workloads
    from tprf_amd.gf import Gf, MeshBrZone, MeshImFreq
    wl = workloads.plain_square(n=4, n_iw=8)
    with pytest.raises(RuntimeError):
        lattice.lattice_dyson_g0_wk(0.0, Gf(MeshBrZone(wl.dims), data=wl.e_k, device="cpu"),
                                    MeshImFreq(wl.beta, "Fermion", wl.n_iw))


def test_product_never_imports_oracle():
    """The product package neither imports nor loads anything under oracle/."""
    pkg = os.path.join(ROOT, "tprf_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in src.lower(), os.path.join(dirpath, f)
