"""tprf_lindhard_chi00_noalloc (no counterpart in the reference): the
Lindhard bubble with its workspace in the library's fixed file-scope device
array instead of allocated scratch, called through the C ABI and checked
against tprf_lindhard_chi00 on the same inputs, twice in a row (the second
call smaller, so it runs over what the first left in the array)."""
import ctypes

import numpy as np
import pytest

from conftest import relerr

pytestmark = pytest.mark.gpu


def _case(seed, dims, n, nz):
    rng = np.random.default_rng(seed)
    nk = int(np.prod(dims))
    a = (rng.normal(size=(nk, n, n)) + 1j * rng.normal(size=(nk, n, n))) * 0.7
    e_k = 0.5 * (a + np.conj(np.swapaxes(a, 1, 2)))
    beta = 7.0
    z = 1j * 2 * np.pi * np.arange(nz) / beta
    return beta, 0.2, e_k, z


def _run(entry, beta, mu, e_k, z, dims, n):
    import torch
    from tprf_amd import _lib
    c = _lib.ctx()
    nk = int(np.prod(dims))
    ek = torch.from_numpy(np.ascontiguousarray(e_k)).cuda()
    zz = torch.from_numpy(np.ascontiguousarray(z)).cuda()
    out = torch.full((len(z), nk, n, n, n, n), complex("nan"), dtype=torch.complex128, device="cuda")
    _lib.call(entry, c, beta, mu, len(z), _lib.ptr(zz), _lib.dims3(dims), n, _lib.ptr(ek), _lib.ptr(out))
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_lindhard_matches_allocating_entry_twice():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    for seed, dims, n, nz in [(1, (8, 8, 8), 3, 4), (2, (4, 4, 2), 2, 3)]:
        beta, mu, e_k, z = _case(seed, dims, n, nz)
        print(f"noalloc lindhard: dims {dims} n {n} nz {nz}", flush=True)
        ref = _run("tprf_lindhard_chi00", beta, mu, e_k, z, dims, n)
        got = _run("tprf_lindhard_chi00_noalloc", beta, mu, e_k, z, dims, n)
        assert np.isfinite(ref).all()
        # the same kernels on the same inputs: equal to rounding (bitwise in practice)
        assert relerr(got, ref) <= 1e-13, (dims, n)


def test_noalloc_lindhard_rejects_sizes_beyond_the_array():
    """nk (8 n + 16 n^3) bytes above the 32 MiB array: a clear error from the host
    side, before anything is launched."""
    import torch
    from tprf_amd import _lib
    if not torch.cuda.is_available():
        pytest.fail("no GPU visible")
    c = _lib.ctx()
    dims = (48, 48, 48)  # 110592 k-points x 1056 bytes (n = 4) = 117 MB
    dummy = torch.zeros(1, dtype=torch.complex128, device="cuda")
    st = _lib.lib().tprf_lindhard_chi00_noalloc(c, 5.0, 0.0, 1, _lib.ptr(dummy), _lib.dims3(dims), 4,
                                                 _lib.ptr(dummy), _lib.ptr(dummy))
    assert st == -9
    msg = _lib.lib().tprf_ctx_last_error(c).decode()
    assert "fixed maximum" in msg
    _ = ctypes
