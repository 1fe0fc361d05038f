"""ctypes binding of libtprf_hip.so, the C-ABI library declared in include/tprf_hip.h.

The library is the only compute path: when it is missing or no GPU is
visible, every call raises.  There is no CPU fallback.
"""
import ctypes
import os
import threading

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtprf_hip.so")

FERMION, BOSON = 0, 1
SIGMA_NONE, SIGMA_W, SIGMA_WK = 0, 1, 2

ERRORS = {
    -1: "bad argument",
    -2: "positive-only mesh",
    -3: "time mesh too short",
    -4: "tail fit error >= 1e-3",
    -5: "tail expansion order too small",
    -6: "singular matrix",
    -7: "HIP runtime error",
    -8: "out of memory",
    -9: "unsupported size",
}

TPRF_CHAIN_HERMITIAN_EK, TPRF_CHAIN_HERMITIAN_U = 1, 2


class ChainDesc(ctypes.Structure):
    """tprf_chain_desc: one instance of tprf_forward_batched."""
    _fields_ = [("beta", ctypes.c_double), ("n_iw", ctypes.c_int), ("ntau", ctypes.c_int), ("nb", ctypes.c_int),
                ("dims", ctypes.c_int * 3), ("norb", ctypes.c_int), ("mu", ctypes.c_double), ("flags", ctypes.c_int)]


class ChainPtrs(ctypes.Structure):
    """tprf_chain_ptrs: device pointers of one instance (inputs, workspace, outputs)."""
    _fields_ = [(f, ctypes.c_void_p) for f in ("e_k", "U", "g_wr", "g_tr", "chi0_wr", "chi0_wk", "chi_wk")]


_lib = None
_lock = threading.Lock()
_ctx = {}

_p, _i, _l, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
_SIGNATURES = {
    "tprf_ctx_create": [ctypes.POINTER(ctypes.c_void_p), _i],
    "tprf_ctx_destroy": [_p],
    "tprf_ctx_set_stream": [_p, _p],
    "tprf_ctx_set_option": [_p, _i, _i],
    "tprf_ctx_last_error": [_p],
    "tprf_ctx_tail_status": [_p, ctypes.POINTER(_d), ctypes.POINTER(_d)],
    "tprf_ctx_tail_status_reset": [_p],
    "tprf_ctx_set_workspace": [_p, _p, ctypes.c_size_t],
    "tprf_ctx_workspace_size": [_p, ctypes.POINTER(ctypes.c_size_t), _i],
    "tprf_lattice_dyson_g_wk": [_p, _d, _i, _d, _l, _i, _p, _i, _p, _p],
    "tprf_lattice_dyson_g_zk": [_p, _l, _p, _d, _l, _i, _p, _i, _p, _p],
    "tprf_lattice_dyson_g_z": [_p, _l, _p, _d, _l, _i, _p, _i, _p, _p],
    "tprf_lattice_dyson_g_wr": [_p, _d, _i, _d, _p, _i, _p, _p],
    "tprf_lattice_dyson_g_wr_herm": [_p, _d, _i, _d, _p, _i, _p, _p, _i],
    "tprf_fourier_wk_to_wr": [_p, _l, _p, _l, _p, _p],
    "tprf_fourier_wr_to_wk": [_p, _l, _p, _l, _p, _p],
    "tprf_fourier_wr_to_tr": [_p, _i, _d, _i, _i, _l, _p, _p, _p],
    "tprf_fourier_wr_to_tr_herm": [_p, _d, _i, _i, _p, _i, _p, _p, _i],
    "tprf_fourier_tr_to_wr": [_p, _i, _d, _i, _i, _l, _p, _p],
    "tprf_chi0_tr_from_grt_PH": [_p, _i, _p, _i, _p, _p],
    "tprf_chi0_wr_from_grt_PH": [_p, _d, _i, _p, _i, _i, _p, _p],
    "tprf_chi0_wr_from_grt_PH_herm": [_p, _d, _i, _p, _i, _i, _p, _p, _i],
    "tprf_chi0_w0r_from_grt_PH": [_p, _d, _i, _p, _i, _p, _p],
    "tprf_chi_w0r_from_chi_tr": [_p, _d, _i, _l, _i, _p, _p],
    "tprf_solve_rpa_PH": [_p, _l, _i, _p, _p, _p],
    "tprf_solve_rpa_PH_herm": [_p, _i, _l, _i, _p, _p, _p],
    "tprf_dynamical_screened_interaction_W": [_p, _l, _l, _i, _p, _p, _i, _p],
    "tprf_split_into_dynamic_wk_and_constant_k": [_p, _i, _d, _i, _l, _p, _p, _p],
    "tprf_density": [_p, _d, _i, _l, _p, _p],
    "tprf_gw_dynamic_sigma": [_p, _l, _i, _p, _p, _p],
    "tprf_fock_sigma_r": [_p, _l, _i, _p, _p, _p],
    "tprf_hartree_sigma_r": [_p, _l, _i, _p, _p, _p],
    "tprf_static_sigma_k": [_p, _i, _p, _i, _p, _p, _p],
    "tprf_add_dynamic_and_static": [_p, _l, _l, _p, _p],
    "tprf_eliashberg_g_delta_g_product": [_p, _i, _i, _p, _i, _p, _p, _p],
    "tprf_eliashberg_product": [_p, _d, _i, _i, _i, _p, _i, _p, _p, _p, _p],
    "tprf_eliashberg_gamma_f_product": [_p, _l, _i, _p, _p, _p],
    "tprf_construct_phi_wk": [_p, _l, _i, _p, _p, _p],
    "tprf_lindhard_chi00": [_p, _d, _d, _l, _p, _p, _i, _p, _p],
    "tprf_lindhard_chi00_noalloc": [_p, _d, _d, _l, _p, _p, _i, _p, _p],
    "tprf_lattice_dyson_g_wk_batched": [_p, _d, _i, _p, _l, _l, _i, _p, _p],
    "tprf_lattice_dyson_g_wk_adjoint": [_p, _i, _l, _i, _l, _p, _p, _p, _p],
    "tprf_fourier_wk_to_wr_adjoint": [_p, _l, _p, _l, _p, _p],
    "tprf_fourier_wr_to_tr_adjoint": [_p, _i, _d, _i, _i, _l, _p, _p, _i],
    "tprf_chi0_static_from_grt_PH": [_p, _d, _i, _p, _i, _l, _p, _p],
    "tprf_chi0_static_from_grt_PH_adjoint": [_p, _d, _i, _p, _i, _l, _p, _p, _p],
    "tprf_chi0_static_from_gwr_PH": [_p, _d, _i, _i, _p, _i, _l, _p, _p],
    "tprf_chi0_static_from_gwr_PH_adjoint": [_p, _d, _i, _i, _p, _i, _l, _p, _p, _p],
    "tprf_chi0_static_q_from_ek": [_p, _d, _i, _i, _p, _p, _i, _l, _p, _p, _p, _p],
    "tprf_chi0_static_q_from_ek_adjoint": [_p, _d, _i, _i, _p, _p, _i, _l, _p, _p, _p, _p, _p, _p],
    "tprf_fourier_r_to_q": [_p, _p, _p, _l, _l, _p, _p],
    "tprf_fourier_r_to_q_adjoint": [_p, _p, _p, _l, _l, _p, _p],
    "tprf_solve_rpa_PH_adjoint": [_p, _l, _i, _p, _p, _p, _p, _p, _p],
    "tprf_chain_buffer_sizes": [_p, _p],
    "tprf_forward_batched": [_p, _i, _p, _p, _i],
}


class TprfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{msg} (tprf status {code}: {ERRORS.get(code, 'unknown')})")
        self.code = code


def lib():
    """Load libtprf_hip.so (torch is imported first so both share one HIP runtime)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C tprf_amd/csrc`")
                l = ctypes.CDLL(LIB_PATH)
                for name, args in _SIGNATURES.items():
                    f = getattr(l, name)
                    f.argtypes = args
                    f.restype = ctypes.c_int
                l.tprf_ctx_last_error.restype = ctypes.c_char_p
                _lib = l
    return _lib


def ctx(device=None):
    """Per-device context, bound to torch's current stream of that device."""
    if not torch.cuda.is_available():
        raise RuntimeError("tprf_amd needs a HIP GPU: torch.cuda.is_available() is False")
    dev = torch.cuda.current_device() if device is None else int(device)
    L = lib()
    c = _ctx.get(dev)
    if c is None:
        h = ctypes.c_void_p()
        with torch.cuda.device(dev):
            st = L.tprf_ctx_create(ctypes.byref(h), dev)
        if st != 0:
            raise TprfError(st, "tprf_ctx_create failed")
        _ctx[dev] = c = h
    L.tprf_ctx_set_stream(c, ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    return c


def check(st, c, what):
    if st != 0:
        msg = lib().tprf_ctx_last_error(c)
        raise TprfError(st, f"{what}: {msg.decode() if msg else ''}")


def call(name, c, *args):
    check(getattr(lib(), name)(c, *args), c, name)


def tail_status(c):
    err, m0 = ctypes.c_double(0.0), ctypes.c_double(0.0)
    st = lib().tprf_ctx_tail_status(c, ctypes.byref(err), ctypes.byref(m0))
    check(st, c, "tail status")
    return err.value, m0.value


OPT_TRACE = 5


def set_trace(c, on=True):
    """roctx ranges per entry point on context c (rocprofv3 --marker-trace)."""
    check(lib().tprf_ctx_set_option(c, OPT_TRACE, 1 if on else 0), c, "trace option")


def tail_status_reset(c):
    check(lib().tprf_ctx_tail_status_reset(c), c, "tail status reset")


def workspace_size(c, reset=False):
    """Largest scratch (bytes) any call on context c asked for since the last reset."""
    b = ctypes.c_size_t(0)
    check(lib().tprf_ctx_workspace_size(c, ctypes.byref(b), 1 if reset else 0), c, "workspace size")
    return b.value


def set_workspace(c, t):
    """Bind the device memory of tensor t (None unbinds) as context c's scratch."""
    if t is None:
        check(lib().tprf_ctx_set_workspace(c, None, 0), c, "set workspace")
    else:
        check(lib().tprf_ctx_set_workspace(c, ctypes.c_void_p(t.data_ptr()), t.numel() * t.element_size()), c,
              "set workspace")


def dims3(dims):
    d = tuple(int(x) for x in dims) + (1,) * (3 - len(dims))
    return (ctypes.c_int * 3)(*d)


def ptr(t):
    """Device pointer of a contiguous complex128 CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not (t.is_cuda and t.dtype == torch.complex128 and t.is_contiguous()):
        raise ValueError("expected a contiguous complex128 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())


def dptr(t):
    """Device pointer of a contiguous float64 CUDA tensor."""
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError("expected a contiguous float64 CUDA tensor")
    return ctypes.c_void_p(t.data_ptr())
