"""tprf_amd -- MI355X-native lattice bubble + RPA chain of TRIQS/tprf.

``tprf_amd.lattice`` mirrors the hot-path functions of ``triqs_tprf.lattice``;
the work runs in hand-written HIP kernels of libtprf_hip.so through its C ABI
(include/tprf_hip.h).
"""
from . import gf, lattice, lattice_utils, rpa_tensor  # noqa: F401
from .gf import Gf, MeshBrZone, MeshCycLat, MeshImFreq, MeshImTime, MeshProduct  # noqa: F401

__version__ = "0.1.0"
