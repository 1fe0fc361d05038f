"""Minimal mesh and Green's function containers (the TRIQS gf shapes the
reference's functions take and return; TRIQS itself is not available).

A ``Gf`` holds ``data`` in the reference's row-major layout
[mesh0][mesh1][target...] as a complex128 torch tensor, on the GPU for
everything the library computes.  Meshes carry exactly the parameters the
reference's kernels read: beta/statistic/size for Matsubara and
imaginary-time meshes, dims for the lattice meshes.
"""
import numpy as np
import torch

FERMION, BOSON = "Fermion", "Boson"


def _stat(S):
    S = {"f": FERMION, "fermion": FERMION, "b": BOSON, "boson": BOSON}.get(str(S).lower(), S)
    if S not in (FERMION, BOSON):
        raise ValueError(f"unknown statistic {S!r}")
    return S


class MeshImFreq:
    """Matsubara mesh with n_max = N: fermions n = -N..N-1, bosons m = -(N-1)..N-1."""

    def __init__(self, beta, S, n_max):
        self.beta, self.statistic, self.n_max = float(beta), _stat(S), int(n_max)
        if self.n_max < 1:
            raise ValueError("n_max must be >= 1")

    @property
    def size(self):
        return 2 * self.n_max if self.statistic == FERMION else 2 * self.n_max - 1

    @property
    def first_index(self):
        return -self.n_max if self.statistic == FERMION else -(self.n_max - 1)

    @property
    def last_index(self):
        return self.n_max - 1

    def index_to_linear(self, n):
        return int(n) - self.first_index

    def values(self):
        n = np.arange(self.first_index, self.last_index + 1)
        odd = 1 if self.statistic == FERMION else 0
        return 1j * np.pi * (2 * n + odd) / self.beta

    def __len__(self):
        return self.size

    def __eq__(self, o):
        return isinstance(o, MeshImFreq) and (o.beta, o.statistic, o.n_max) == (self.beta, self.statistic, self.n_max)

    def __repr__(self):
        return f"MeshImFreq(beta={self.beta}, S={self.statistic!r}, n_max={self.n_max})"


class MeshReFreq:
    """Real-frequency mesh of n_w points from omega_min to omega_max, both included."""

    def __init__(self, omega_min, omega_max, n_w):
        self.omega_min, self.omega_max, self.n_w = float(omega_min), float(omega_max), int(n_w)
        if self.n_w < 1:
            raise ValueError("n_w must be >= 1")

    @property
    def size(self):
        return self.n_w

    def values(self):
        return np.linspace(self.omega_min, self.omega_max, self.n_w)

    def __len__(self):
        return self.size

    def __eq__(self, o):
        return isinstance(o, MeshReFreq) and (o.omega_min, o.omega_max, o.n_w) == (self.omega_min, self.omega_max,
                                                                                    self.n_w)

    def __repr__(self):
        return f"MeshReFreq(omega_min={self.omega_min}, omega_max={self.omega_max}, n_w={self.n_w})"


class MeshImTime:
    """tau_j = j beta / (n_tau - 1), j = 0..n_tau-1 (both ends included)."""

    def __init__(self, beta, S, n_tau):
        self.beta, self.statistic, self.n_tau = float(beta), _stat(S), int(n_tau)

    @property
    def size(self):
        return self.n_tau

    def values(self):
        return np.arange(self.n_tau) * (self.beta / (self.n_tau - 1))

    def __len__(self):
        return self.n_tau

    def __eq__(self, o):
        return isinstance(o, MeshImTime) and (o.beta, o.statistic, o.n_tau) == (self.beta, self.statistic, self.n_tau)

    def __repr__(self):
        return f"MeshImTime(beta={self.beta}, S={self.statistic!r}, n_tau={self.n_tau})"


class _Lattice:
    def __init__(self, dims):
        d = tuple(int(x) for x in dims)
        self.dims = d + (1,) * (3 - len(d))
        if any(x < 1 for x in self.dims):
            raise ValueError("lattice dims must be >= 1")

    @property
    def size(self):
        return int(np.prod(self.dims))

    def __len__(self):
        return self.size

    def index(self, i0, i1=0, i2=0):
        return (i0 * self.dims[1] + i1) * self.dims[2] + i2

    def __eq__(self, o):
        return type(o) is type(self) and o.dims == self.dims


class MeshBrZone(_Lattice):
    """Reciprocal-space mesh k = 2 pi (i0/d0, i1/d1, i2/d2) in units of the primitive vectors."""

    def values(self):
        g = np.meshgrid(*[np.arange(d) / d for d in self.dims], indexing="ij")
        return 2 * np.pi * np.stack([x.ravel() for x in g], axis=-1)

    def __repr__(self):
        return f"MeshBrZone(dims={self.dims})"


class MeshCycLat(_Lattice):
    """Real-space cyclic lattice; r = -r is (-i mod d) per dimension."""

    def minus(self, r):
        i2 = r % self.dims[2]
        i1 = (r // self.dims[2]) % self.dims[1]
        i0 = r // (self.dims[1] * self.dims[2])
        return self.index((-i0) % self.dims[0], (-i1) % self.dims[1], (-i2) % self.dims[2])

    def __repr__(self):
        return f"MeshCycLat(dims={self.dims})"


class MeshProduct:
    def __init__(self, *components):
        self.components = tuple(components)

    def __getitem__(self, i):
        return self.components[i]

    @property
    def size(self):
        return int(np.prod([m.size for m in self.components]))

    def __eq__(self, o):
        return isinstance(o, MeshProduct) and o.components == self.components

    def __repr__(self):
        return "MeshProduct(" + ", ".join(repr(m) for m in self.components) + ")"


def _mesh_shape(mesh):
    if isinstance(mesh, MeshProduct):
        return tuple(m.size for m in mesh.components)
    return (mesh.size,)


class Gf:
    """Green's function / susceptibility: ``data[mesh...][target...]`` (complex128 torch tensor)."""

    def __init__(self, mesh, target_shape=None, data=None, device=None):
        self.mesh = mesh
        mshape = _mesh_shape(mesh)
        if data is None:
            if target_shape is None:
                raise ValueError("need target_shape or data")
            dev = torch.device("cuda") if device is None else torch.device(device)
            data = torch.zeros(mshape + tuple(target_shape), dtype=torch.complex128, device=dev)
        else:
            if isinstance(data, np.ndarray):
                data = torch.from_numpy(np.ascontiguousarray(data, dtype=np.complex128))
            if device is not None:
                data = data.to(device)
            if data.dtype != torch.complex128:
                data = data.to(torch.complex128)
            data = data.contiguous()
            if tuple(data.shape[: len(mshape)]) != mshape:
                raise ValueError(f"data shape {tuple(data.shape)} does not match mesh shape {mshape}")
        self.data = data
        self.target_shape = tuple(data.shape[len(mshape):])

    @property
    def target_rank(self):
        return len(self.target_shape)

    def numpy(self):
        return self.data.detach().cpu().numpy()

    def copy(self):
        return Gf(self.mesh, data=self.data.clone())

    def __repr__(self):
        return f"Gf(mesh={self.mesh!r}, target_shape={self.target_shape})"
