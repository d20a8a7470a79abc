"""Synthetic workloads of the pressure-correction solve path.

Wraps ``libsimall_workload.so`` (host-only C++; see workload/workload.h) and
defines the five solver workloads with their seeds:

  cavity2d   128x128x1 slab, unit conductances, pinned Neumann Poisson
  poisson3d  N^3 perturbed (+-20 % h) and permuted hex mesh, log-uniform
             conductances 10^U[-2,2], pinned Neumann Poisson (N = 128: 2.1M cells)
  momentum3d N^3 on [0,2pi]^3, Taylor-Green upwind convection + diffusion with
             a pseudo-transient diagonal (N = 96: 885k cells)
  transport  N^3 perturbed/permuted, 32 scalar equations sharing one pattern
  adjoint    J = sum w_i x_i^2 on the poisson3d system

Seeds: 0x51A11000 + 0x100 * workload_index + quantity (1 nodes, 2 permutation,
3 conductance, 4 rhs, 5 weights, 6 time step, 7 faces, 8+e per equation).
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimall_workload.so")

SEED_BASE = 0x51A11000
Q_NODES, Q_PERM, Q_GAMMA, Q_RHS, Q_WEIGHTS, Q_DT, Q_FACES, Q_EQ = 1, 2, 3, 4, 5, 6, 7, 8

# boundary treatment per zone (oracle.h / mesh entry points share these values)
BC_WALL, BC_PRESSURE, BC_SYMMETRY, BC_VELOCITY_INLET, BC_OUTFLOW, BC_PERIODIC = 0, 1, 2, 3, 4, 5


def seed(workload_index, quantity):
    return SEED_BASE + 0x100 * workload_index + quantity


class WlMesh(C.Structure):
    _fields_ = [
        ("n_cells", C.c_int64),
        ("n_faces", C.c_int64),
        ("n_nodes", C.c_int64),
        ("nodeX", C.POINTER(C.c_double)),
        ("nodeY", C.POINTER(C.c_double)),
        ("nodeZ", C.POINTER(C.c_double)),
        ("owner", C.POINTER(C.c_uint64)),
        ("neighbor", C.POINTER(C.c_uint64)),
        ("areaX", C.POINTER(C.c_double)),
        ("areaY", C.POINTER(C.c_double)),
        ("areaZ", C.POINTER(C.c_double)),
        ("faceCx", C.POINTER(C.c_double)),
        ("faceCy", C.POINTER(C.c_double)),
        ("faceCz", C.POINTER(C.c_double)),
        ("boundaryZone", C.POINTER(C.c_uint32)),
        ("faceNodeOffsets", C.POINTER(C.c_int32)),
        ("faceNodeIndices", C.POINTER(C.c_uint64)),
        ("volume", C.POINTER(C.c_double)),
        ("cellCx", C.POINTER(C.c_double)),
        ("cellCy", C.POINTER(C.c_double)),
        ("cellCz", C.POINTER(C.c_double)),
        ("cellFaceOffsets", C.POINTER(C.c_int32)),
        ("cellFaceIndices", C.POINTER(C.c_uint64)),
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        lib = C.CDLL(LIB_PATH)
        lib.wl_mesh_cartesian.restype = C.c_int
        lib.wl_mesh_cartesian.argtypes = [C.c_int] * 3 + [C.c_double] * 7 + [C.c_uint64, C.c_uint64, C.POINTER(WlMesh)]
        lib.wl_mesh_free.argtypes = [C.POINTER(WlMesh)]
        lib.wl_mesh_free.restype = None
        for name, args in {
            "wl_fill_uniform": [C.c_uint64, C.c_double, C.c_double, C.c_void_p, C.c_int64],
            "wl_fill_normal": [C.c_uint64, C.c_void_p, C.c_int64],
            "wl_fill_loguniform": [C.c_uint64, C.c_double, C.c_double, C.c_void_p, C.c_int64],
            "wl_remove_mean": [C.c_void_p, C.c_int64],
            "wl_taylor_green": [C.POINTER(WlMesh), C.c_void_p, C.c_void_p, C.c_void_p],
            "wl_permutation": [C.c_uint64, C.c_int64, C.c_void_p],
        }.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = None
        _lib = lib
    return _lib


class Mesh:
    """SoA hex mesh (numpy views over the C arrays; freed with the object)."""

    FACE = ["owner", "neighbor", "areaX", "areaY", "areaZ", "faceCx", "faceCy", "faceCz", "boundaryZone"]
    CELL = ["volume", "cellCx", "cellCy", "cellCz"]

    def __init__(self, Nx, Ny, Nz, origin=(0.0, 0.0, 0.0), extent=(1.0, 1.0, 1.0), perturb=0.0,
                 node_seed=0, perm_seed=0):
        lib = load()
        self._m = WlMesh()
        rc = lib.wl_mesh_cartesian(Nx, Ny, Nz, *origin, *extent, perturb, node_seed, perm_seed, C.byref(self._m))
        if rc != 0:
            raise RuntimeError(f"wl_mesh_cartesian failed ({rc})")
        self.dims = (Nx, Ny, Nz)
        self.origin, self.extent = origin, extent
        m = self._m
        nF, nC, nN = m.n_faces, m.n_cells, m.n_nodes
        self.n_cells, self.n_faces, self.n_nodes = nC, nF, nN
        view = np.ctypeslib.as_array
        for k in self.FACE:
            setattr(self, k, view(getattr(m, k), (nF,)))
        for k in self.CELL:
            setattr(self, k, view(getattr(m, k), (nC,)))
        self.cellFaceOffsets = view(m.cellFaceOffsets, (nC + 1,))
        self.cellFaceIndices = view(m.cellFaceIndices, (int(self.cellFaceOffsets[-1]),))
        self.faceNodeOffsets = view(m.faceNodeOffsets, (nF + 1,))
        self.faceNodeIndices = view(m.faceNodeIndices, (int(self.faceNodeOffsets[-1]),))
        self.nodeX, self.nodeY, self.nodeZ = (view(getattr(m, k), (nN,)) for k in ("nodeX", "nodeY", "nodeZ"))

    @property
    def struct(self):
        return self._m

    def interior_faces(self):
        return np.nonzero(self.neighbor != np.uint64(0xFFFFFFFFFFFFFFFF))[0]

    def __del__(self):
        try:
            load().wl_mesh_free(C.byref(self._m))
        except Exception:
            pass


def uniform(seed_, lo, hi, n):
    out = np.empty(n, np.float64)
    load().wl_fill_uniform(seed_, lo, hi, out.ctypes.data, n)
    return out


def normal(seed_, n):
    out = np.empty(n, np.float64)
    load().wl_fill_normal(seed_, out.ctypes.data, n)
    return out


def loguniform(seed_, lo_exp, hi_exp, n):
    out = np.empty(n, np.float64)
    load().wl_fill_loguniform(seed_, lo_exp, hi_exp, out.ctypes.data, n)
    return out


def remove_mean(v):
    load().wl_remove_mean(v.ctypes.data, v.size)
    return v


def taylor_green(mesh):
    n = mesh.n_cells
    U = [np.empty(n, np.float64) for _ in range(3)]
    load().wl_taylor_green(C.byref(mesh.struct), *(u.ctypes.data for u in U))
    return U


def permutation(seed_, n):
    out = np.empty(n, np.int64)
    load().wl_permutation(seed_, n, out.ctypes.data)
    return out


# ----------------------------------------------------------------- workloads
class Workload:
    """Mesh + per-face / per-cell inputs of one solver workload."""

    def __init__(self, name, index, mesh, bc_of_zone, **fields):
        self.name = name
        self.index = index
        self.mesh = mesh
        self.bc_of_zone = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
        for k, v in fields.items():
            setattr(self, k, v)


def cavity2d(N=128):
    """128x128x1 on [0,1]^2 x [0,0.1]; unit conductance; b = U[-1,1] - mean."""
    mesh = Mesh(N, N, 1, extent=(1.0, 1.0, 0.1))
    b = remove_mean(uniform(seed(1, Q_RHS), -1.0, 1.0, mesh.n_cells))
    return Workload("cavity2d", 1, mesh, [BC_WALL] * 7, gamma=None, b=b, rho=1.0,
                    tol=1e-8, max_iter=10000)


def poisson3d(N=128):
    """N^3 on [0,1]^3, interior nodes +-20 % h, cells permuted, gamma = 10^U[-2,2]."""
    mesh = Mesh(N, N, N, perturb=0.2, node_seed=seed(2, Q_NODES), perm_seed=seed(2, Q_PERM))
    gamma = loguniform(seed(2, Q_GAMMA), -2.0, 2.0, mesh.n_faces)
    b = remove_mean(normal(seed(2, Q_RHS), mesh.n_cells))
    return Workload("poisson3d", 2, mesh, [BC_WALL] * 7, gamma=gamma, b=b, rho=1.0,
                    tol=1e-8, max_iter=10000)


def momentum3d(N=96):
    """Taylor-Green upwind + diffusion on [0,2pi]^3, Gamma = h/50, dt = 10^U[-1,1]."""
    L = 2.0 * math.pi
    mesh = Mesh(N, N, N, extent=(L, L, L))
    Ux, Uy, Uz = taylor_green(mesh)
    h = L / N
    dt = loguniform(seed(3, Q_DT), -1.0, 1.0, mesh.n_cells)
    b = normal(seed(3, Q_RHS), mesh.n_cells)
    return Workload("momentum3d", 3, mesh, [BC_WALL] * 7, U=(Ux, Uy, Uz), mu=h / 50.0, rho=1.0,
                    dt=dt, b=b, urf=1.0, tol=1e-8, max_iter=2000)


def transport(N=100, n_eq=32):
    """n_eq SPD systems A_e = V/dt I + Gamma_e L on one perturbed, permuted mesh."""
    mesh = Mesh(N, N, N, perturb=0.2, node_seed=seed(4, Q_NODES), perm_seed=seed(4, Q_PERM))
    gammas = loguniform(seed(4, Q_GAMMA), -6.0, -3.0, n_eq)
    bs = [normal(seed(4, Q_EQ + e), mesh.n_cells) for e in range(n_eq)]
    return Workload("transport", 4, mesh, [BC_SYMMETRY] * 7, gammas=gammas, bs=bs, dt=1.0,
                    tol=1e-8, max_iter=2000)
