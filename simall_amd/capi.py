"""ctypes binding of the C ABI declared in include/simall_amd.h.

The library is the in-tree ``simall_amd/libsimall_amd.so`` built by
``__graft_entry__.build()``.  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsimall_amd.so")

MI_OK = 0
MI_ERR_NOT_CONVERGED = 1
MI_ERR_BREAKDOWN = 2
MI_ERR_BAD_ARG = 3
MI_ERR_DEVICE = 4
MI_ERR_NO_DEVICE = 5
MI_ERR_UNSUPPORTED = 6
MI_CHECKPOINT_AMG = 1
MI_ERR_ZERO_PIVOT = 7

MI_SOLVER_GMRES, MI_SOLVER_BICGSTAB, MI_SOLVER_CG, MI_SOLVER_TFQMR = 0, 1, 2, 3
MI_PC_JACOBI, MI_PC_ILU, MI_PC_SSOR, MI_PC_AMG, MI_PC_NONE = 0, 1, 2, 3, 4
MI_PRECISION_FP64, MI_PRECISION_MIXED = 0, 1


class SolverCfg(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("preconditioner", C.c_int),
        ("tolerance", C.c_double),
        ("max_iterations", C.c_int),
        ("restart", C.c_int),
        ("precision", C.c_int),
        ("warm_start", C.c_int),
        ("true_residual", C.c_int),
        ("inner_tolerance", C.c_double),
        ("max_outer", C.c_int),
        ("reorder", C.c_int),
        ("check_interval", C.c_int),
        ("amg_max_levels", C.c_int),
        ("amg_coarse_size", C.c_int),
        ("amg_strength", C.c_double),
        ("amg_jacobi_omega", C.c_double),
        ("amg_smoother", C.c_int),
        ("amg_sweeps", C.c_int),
        ("amg_coarse_sweeps", C.c_int),
        ("keep_preconditioner", C.c_int),
        ("use_graph", C.c_int),
        ("amg_resetup", C.c_int),
        ("amg_truncation", C.c_double),
    ]


MI_FLOW_SIMPLE, MI_FLOW_PISO, MI_FLOW_SIMPLEC = 0, 1, 2
MI_TIME_EULER, MI_TIME_BDF2 = 0, 1


class FlowCfg(C.Structure):
    _fields_ = [
        ("algorithm", C.c_int),
        ("rho", C.c_double),
        ("mu", C.c_double),
        ("urfU", C.c_double),
        ("urfP", C.c_double),
        ("dt", C.c_double),
        ("time_scheme", C.c_int),
        ("n_correctors", C.c_int),
        ("n_nonorth", C.c_int),
        ("momentum", SolverCfg),
        ("pressure", SolverCfg),
        ("momentum_batch", C.c_int),
    ]


class MeshNodes(C.Structure):
    _fields_ = [("n_nodes", C.c_int64)] + [
        (k, C.c_void_p) for k in ("x", "y", "z", "faceNodeOffsets", "faceNodeIndices")
    ]


class CheckpointMeta(C.Structure):
    _fields_ = [("iteration", C.c_int32), ("time", C.c_double), ("settings_hash", C.c_uint64)]


class NamedField(C.Structure):
    _fields_ = [("name", C.c_char_p), ("data", C.c_void_p)]


class NamedVector(C.Structure):
    _fields_ = [("name", C.c_char_p), ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p)]


class FlowResiduals(C.Structure):
    _fields_ = [("mom", C.c_double * 3), ("cont", C.c_double)]

    def as_dict(self):
        return {"mom": list(self.mom), "cont": self.cont}


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int),
        ("outer_iterations", C.c_int),
        ("converged", C.c_int),
        ("status", C.c_int),
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("relative_residual", C.c_double),
        ("true_relative_residual", C.c_double),
        ("setup_seconds", C.c_double),
        ("solve_seconds", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_vp = C.c_void_p
_i = C.c_int
_i64 = C.c_int64
_sz = C.c_size_t
_d = C.c_double
_pd = C.POINTER(C.c_double)

# name -> (restype, argtypes); the list of every exported entry point.
PROTOTYPES = {
    "mi_get_abi_version": (_i, []),
    "mi_init": (_i, [_i]),
    "mi_device_count": (_i, [C.POINTER(_i)]),
    "mi_last_error": (C.c_char_p, []),
    "mi_device_synchronize": (_i, []),
    "mi_device_alloc": (_i, [_sz, C.POINTER(_vp)]),
    "mi_device_free": (_i, [_vp]),
    "mi_memcpy_h2d": (_i, [_vp, _vp, _sz]),
    "mi_memcpy_d2h": (_i, [_vp, _vp, _sz]),
    "mi_memcpy_d2d": (_i, [_vp, _vp, _sz]),
    "mi_memset_device": (_i, [_vp, _i, _sz]),
    "mi_axpy": (_i, [_d, _vp, _vp, _sz, _vp]),
    "mi_scal": (_i, [_d, _vp, _sz, _vp]),
    "mi_copy": (_i, [_vp, _vp, _sz, _vp]),
    "mi_dot": (_i, [_vp, _vp, _sz, _pd, _vp]),
    "mi_nrm2": (_i, [_vp, _sz, _pd, _vp]),
    "mi_axpby": (_i, [_d, _vp, _d, _vp, _sz, _vp]),
    "mi_xpby": (_i, [_vp, _d, _vp, _sz, _vp]),
    "mi_spmv_csr": (_i, [_sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_jacobi_apply": (_i, [_vp, _vp, _vp, _sz, _vp]),
    "mi_ilu0_factor_inplace": (_i, [_sz, _vp, _vp, _vp, C.POINTER(_i)]),
    "mi_ilu0_apply": (_i, [_sz, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_csr_create": (_i, [_i64, _i64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "mi_csr_update_values": (_i, [_vp, _vp]),
    "mi_csr_update_values_device": (_i, [_vp, _vp]),
    "mi_csr_get_values": (_i, [_vp, _vp]),
    "mi_csr_get_pattern": (_i, [_vp, _vp, _vp]),
    "mi_csr_info": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i64)]),
    "mi_csr_values_device": (_i, [_vp, C.POINTER(_vp)]),
    "mi_csr_spmv": (_i, [_vp, _vp, _vp]),
    "mi_csr_spmv_device": (_i, [_vp, _i, _vp, _vp, _vp]),
    "mi_csr_time_spmv": (_i, [_vp, _i, _vp, _vp, _i, _pd]),
    "mi_csr_time_spmv_cold": (_i, [_vp, _i, _vp, _vp, _i, _i64, _pd]),
    "mi_csr_transpose": (_i, [_vp, C.POINTER(_vp)]),
    "mi_csr_transpose_values": (_i, [_vp, _vp]),
    "mi_csr_destroy": (_i, [_vp]),
    "mi_solver_cfg_default": (None, [C.POINTER(SolverCfg)]),
    "mi_solver_create": (_i, [C.POINTER(SolverCfg), C.POINTER(_vp)]),
    "mi_solver_setup": (_i, [_vp, _vp]),
    "mi_solver_solve": (_i, [_vp, _vp, _vp, C.POINTER(Stats)]),
    "mi_solver_solve_device": (_i, [_vp, _vp, _vp, C.POINTER(Stats)]),
    "mi_adjoint_solve": (_i, [_vp, _vp, _vp, C.POINTER(Stats)]),
    "mi_solver_last_residual": (_d, [_vp]),
    "mi_solver_precond_apply": (_i, [_vp, _vp, _vp]),
    "mi_solver_time_spmv": (_i, [_vp, _i, _i, _pd, _pd]),
    "mi_solver_time_spmv_cold": (_i, [_vp, _i, _i, _i64, _pd, _pd]),
    "mi_solver_info": (_i, [_vp, C.POINTER(_i), _pd, C.POINTER(_i), C.POINTER(_i)]),
    "mi_solver_history": (_i, [_vp, _pd, _i, C.POINTER(_i)]),
    "mi_solver_destroy": (_i, [_vp]),
    "mi_mesh_create": (_i, [_vp, _vp, _i, C.POINTER(_vp)]),
    "mi_mesh_info": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "mi_mesh_create_matrix": (_i, [_vp, C.POINTER(_vp)]),
    "mi_assemble_pressure": (_i, [_vp, _d, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i)]),
    "mi_assemble_momentum": (_i, [_vp, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    "mi_face_flux": (_i, [_vp, _d, _vp, _vp, _vp, _vp, _vp]),
    "mi_face_flux_noalloc": (_i, [_vp, _d, _vp, _vp, _vp, _vp, _vp]),
    "mi_mesh_create_twins": (_i, [_vp, _vp, _i, _vp, C.POINTER(_vp)]),
    "mi_simplec_denominators": (_i, [_vp, _vp, _vp, _vp]),
    "mi_face_gradient": (_i, [_vp, _d, _vp, _i, _vp, _vp, _vp]),
    "mi_batch_create": (_i, [_vp, _i, C.POINTER(_vp)]),
    "mi_batch_create_ex": (_i, [_vp, _i, _i, C.POINTER(_vp)]),
    "mi_batch_set_equation": (_i, [_vp, _i, _vp]),
    "mi_batch_set_matrix": (_i, [_vp, _vp]),
    "mi_batch_apply_preconditioner": (_i, [_vp, _i, _vp, _vp]),
    "mi_batch_set_shape": (_i, [_vp, _i, _i, _i]),
    "mi_batch_tuning_shape": (_i, [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                                   C.POINTER(_i), C.POINTER(_d)]),
    "mi_time_grid_barrier": (_i, [_i, _i, C.POINTER(_d), C.POINTER(_d)]),
    "mi_color_permutation": (_i, [_i64, _vp, _vp, _vp, C.POINTER(C.c_int32)]),
    "mi_batch_set_values": (_i, [_vp, _vp]),
    "mi_batch_solve": (_i, [_vp, C.POINTER(SolverCfg), _vp, _vp, _vp, _vp]),
    "mi_batch_solve_device": (_i, [_vp, C.POINTER(SolverCfg), _vp, _vp, _vp, _vp]),
    "mi_batch_tuning": (_i, [_vp, C.POINTER(_i), C.POINTER(_d), C.POINTER(_d)]),
    "mi_batch_destroy": (_i, [_vp]),
    "mi_mesh_destroy": (_i, [_vp]),
    "mi_ls_gradient": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "mi_correct_fields": (_i, [_vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mi_flow_cfg_default": (None, [_i, C.POINTER(FlowCfg)]),
    "mi_flow_create": (_i, [_vp, _vp, C.POINTER(FlowCfg), C.POINTER(_vp)]),
    "mi_flow_fields": (_i, [_vp] + [C.POINTER(_vp)] * 5),
    "mi_flow_iterate": (_i, [_vp, C.POINTER(FlowResiduals)]),
    "mi_flow_advance_time_step": (_i, [_vp, _i, C.POINTER(FlowResiduals)]),
    "mi_flow_destroy": (_i, [_vp]),
    "mi_checkpoint_write": (_i, [C.c_char_p, _vp, _vp, _i, _vp, _i, _vp, _vp]),
    "mi_checkpoint_open": (_i, [C.c_char_p, C.POINTER(_vp)]),
    "mi_checkpoint_info": (_i, [_vp, _vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "mi_checkpoint_scalar": (_i, [_vp, C.c_char_p, _vp, _i64, C.POINTER(_i64)]),
    "mi_checkpoint_vector": (_i, [_vp, C.c_char_p, _vp, _vp, _vp, _i64, C.POINTER(_i64)]),
    "mi_checkpoint_close": (_i, [_vp]),
    "mi_checkpoint_amg_levels": (_i, [_vp, C.POINTER(_i)]),
    "mi_checkpoint_append_amg": (_i, [C.c_char_p, _vp]),
    "mi_flow_checkpoint_write": (_i, [_vp, C.c_char_p, _vp, _vp, _vp, _i]),
    "mi_flow_checkpoint_read": (_i, [_vp, C.c_char_p, _vp]),
    "mi_amg_host_build": (_i, [_i64, _vp, _vp, _vp, C.POINTER(SolverCfg), C.POINTER(_vp)]),
    "mi_solver_amg_host": (_i, [_vp, C.POINTER(_vp)]),
    "mi_amg_host_levels": (_i, [_vp, C.POINTER(_i), _pd, _pd]),
    "mi_amg_host_level_info": (_i, [_vp, _i, _i, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), _pd]),
    "mi_amg_host_level_get": (_i, [_vp, _i, _i, _vp, _vp, _vp]),
    "mi_amg_host_destroy": (_i, [_vp]),
    "mi_rcm_permutation": (_i, [_i64, _vp, _vp, _vp, C.POINTER(C.c_int32)]),
}

_lib = None


def load():
    """Loads libsimall_amd.so and binds every prototype (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mi status {code}: {msg}")
        self.code = code


def check(rc, allowed=(MI_OK,)):
    if rc not in allowed:
        raise MiError(rc, load().mi_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Host pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


# --------------------------------------------------------------- device arrays
class DeviceArray:
    """A device allocation owned by Python (freed on close / garbage collection)."""

    def __init__(self, n, dtype=np.float64):
        self.lib = load()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(self.lib.mi_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(self.lib.mi_memcpy_h2d(self.ptr, ptr(a), self.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        check(self.lib.mi_memcpy_d2h(ptr(out), self.ptr, self.nbytes))
        return out

    def zero(self):
        check(self.lib.mi_memset_device(self.ptr, 0, self.nbytes))

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.mi_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------ handles
class Csr:
    """mi_csr handle over host CSR arrays (int32 rowPtr/colIdx, float64 values)."""

    def __init__(self, rowPtr, colIdx, values):
        self.lib = load()
        self.rowPtr = np.ascontiguousarray(rowPtr, dtype=np.int32)
        self.colIdx = np.ascontiguousarray(colIdx, dtype=np.int32)
        vals = np.ascontiguousarray(values, dtype=np.float64)
        self.n = self.rowPtr.size - 1
        self.nnz = self.colIdx.size
        h = C.c_void_p()
        check(self.lib.mi_csr_create(self.n, self.nnz, ptr(self.rowPtr), ptr(self.colIdx), ptr(vals), C.byref(h)))
        self.h = h.value

    def update_values(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        check(self.lib.mi_csr_update_values(self.h, ptr(v)))

    def values(self):
        out = np.empty(self.nnz, np.float64)
        check(self.lib.mi_csr_get_values(self.h, ptr(out)))
        return out

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n, np.float64)
        check(self.lib.mi_csr_spmv(self.h, ptr(x), ptr(y)))
        return y

    def spmv_device(self, dx, dy, precision=MI_PRECISION_FP64):
        check(self.lib.mi_csr_spmv_device(self.h, precision, dx.ptr, dy.ptr, None))

    def time_spmv(self, dx, dy, precision=MI_PRECISION_FP64, reps=20):
        us = C.c_double()
        check(self.lib.mi_csr_time_spmv(self.h, precision, dx.ptr, dy.ptr, reps, C.byref(us)))
        return us.value

    def transpose(self):
        """Explicit A^T as a new handle (keeps this matrix alive)."""
        h = C.c_void_p()
        check(self.lib.mi_csr_transpose(self.h, C.byref(h)))
        At = CsrView(h.value)
        At._src = self
        return At

    def refresh_transpose(self, At):
        check(self.lib.mi_csr_transpose_values(self.h, At.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_cfg(**kw):
    cfg = SolverCfg()
    load().mi_solver_cfg_default(C.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise AttributeError(k)
        setattr(cfg, k, v)
    return cfg


class Solver:
    """mi_solver handle (ILinearSolver semantics: solve returns iterations)."""

    def __init__(self, cfg=None, **kw):
        self.lib = load()
        self.cfg = cfg if cfg is not None else default_cfg(**kw)
        h = C.c_void_p()
        check(self.lib.mi_solver_create(C.byref(self.cfg), C.byref(h)))
        self.h = h.value
        self.A = None

    def setup(self, A):
        self.A = A
        check(self.lib.mi_solver_setup(self.h, A.h))

    def solve(self, b, x=None):
        """Returns (x, stats, status)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b) if x is None else np.array(x, dtype=np.float64, copy=True)
        st = Stats()
        rc = self.lib.mi_solver_solve(self.h, ptr(b), ptr(x), C.byref(st))
        check(rc, allowed=(MI_OK, MI_ERR_NOT_CONVERGED, MI_ERR_BREAKDOWN))
        return x, st, rc

    def solve_device(self, db, dx):
        st = Stats()
        rc = self.lib.mi_solver_solve_device(self.h, db.ptr, dx.ptr, C.byref(st))
        check(rc, allowed=(MI_OK, MI_ERR_NOT_CONVERGED, MI_ERR_BREAKDOWN))
        return st, rc

    def adjoint_solve_device(self, dg, dlam):
        """A^T lambda = g on the matrix bound by setup (device vectors)."""
        st = Stats()
        rc = self.lib.mi_adjoint_solve(self.h, dg.ptr, dlam.ptr, C.byref(st))
        check(rc, allowed=(MI_OK, MI_ERR_NOT_CONVERGED, MI_ERR_BREAKDOWN))
        return st, rc

    def last_residual(self):
        return self.lib.mi_solver_last_residual(self.h)

    def precond_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        check(self.lib.mi_solver_precond_apply(self.h, ptr(r), ptr(z)))
        return z

    def time_spmv(self, precision=MI_PRECISION_FP64, reps=50):
        us, by = C.c_double(), C.c_double()
        check(self.lib.mi_solver_time_spmv(self.h, precision, reps, C.byref(us), C.byref(by)))
        return us.value, by.value

    def time_spmv_cold(self, precision=MI_PRECISION_FP64, reps=20, flush_bytes=1 << 30):
        """Average of `reps` SpMV launches rotating over copies of the matrix
        (more than flush_bytes in all), so none finds its streams cached."""
        us, by = C.c_double(), C.c_double()
        check(self.lib.mi_solver_time_spmv_cold(self.h, precision, reps, flush_bytes, C.byref(us),
                                                C.byref(by)))
        return us.value, by.value

    def info(self):
        lv, oc, b0, b1 = C.c_int(), C.c_double(), C.c_int(), C.c_int()
        check(self.lib.mi_solver_info(self.h, C.byref(lv), C.byref(oc), C.byref(b0), C.byref(b1)))
        return {"levels": lv.value, "operator_complexity": oc.value, "bandwidth_before": b0.value,
                "bandwidth_after": b1.value}

    def history(self):
        cnt = C.c_int()
        check(self.lib.mi_solver_history(self.h, None, 0, C.byref(cnt)))
        out = np.empty(cnt.value, np.float64)
        if cnt.value:
            check(self.lib.mi_solver_history(self.h, out.ctypes.data_as(_pd), cnt.value, C.byref(cnt)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MeshSoa(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_faces", C.c_int64)] + [
        (k, C.c_void_p)
        for k in ("owner", "neighbor", "areaX", "areaY", "areaZ", "faceCx", "faceCy", "faceCz",
                  "boundaryZone", "cellCx", "cellCy", "cellCz", "volume", "cellFaceOffsets",
                  "cellFaceIndices")
    ]


class CsrView:
    """An mi_csr created by the library (e.g. on a mesh pattern)."""

    def __init__(self, handle):
        self.lib = load()
        self.h = handle
        n, nnz = C.c_int64(), C.c_int64()
        check(self.lib.mi_csr_info(self.h, C.byref(n), C.byref(nnz)))
        self.n, self.nnz = n.value, nnz.value

    values = Csr.values
    update_values = Csr.update_values
    transpose = Csr.transpose
    refresh_transpose = Csr.refresh_transpose
    spmv = Csr.spmv
    spmv_device = Csr.spmv_device
    time_spmv = Csr.time_spmv
    close = Csr.close
    __del__ = Csr.__del__

    def pattern(self):
        rp = np.empty(self.n + 1, np.int32)
        ci = np.empty(self.nnz, np.int32)
        check(self.lib.mi_csr_get_pattern(self.h, ptr(rp), ptr(ci)))
        return rp, ci


def _dptr(d):
    return None if d is None else d.ptr


def mesh_soa(mesh):
    """MeshSoa over the host arrays of a simall_amd.workload.Mesh (the mesh
    must outlive the struct)."""
    s = MeshSoa()
    s.n_cells, s.n_faces = mesh.n_cells, mesh.n_faces
    for k, _ in MeshSoa._fields_[2:]:
        setattr(s, k, getattr(mesh, k).ctypes.data)
    return s


def mesh_nodes(mesh):
    """MeshNodes over the node arrays of a simall_amd.workload.Mesh."""
    s = MeshNodes()
    s.n_nodes = mesh.nodeX.size
    s.x, s.y, s.z = (getattr(mesh, k).ctypes.data for k in ("nodeX", "nodeY", "nodeZ"))
    s.faceNodeOffsets = mesh.faceNodeOffsets.ctypes.data
    s.faceNodeIndices = mesh.faceNodeIndices.ctypes.data
    return s


def checkpoint_write(path, mesh, scalars=None, vectors=None, iteration=0, time=0.0,
                     settings_hash=0):
    """Checkpoint::write of a workload Mesh with named host fields:
    scalars {name: array[n_cells]}, vectors {name: (x, y, z)}."""
    lib = load()
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data

    sc = list((scalars or {}).items())
    vc = list((vectors or {}).items())
    S = (NamedField * max(len(sc), 1))()
    V = (NamedVector * max(len(vc), 1))()
    for i, (name, a) in enumerate(sc):
        S[i].name, S[i].data = name.encode(), arr(a)
    for i, (name, (x, y, z)) in enumerate(vc):
        V[i].name, V[i].x, V[i].y, V[i].z = name.encode(), arr(x), arr(y), arr(z)
    meta = CheckpointMeta(iteration, time, settings_hash)
    soa, nodes = mesh_soa(mesh), mesh_nodes(mesh)
    check(lib.mi_checkpoint_write(os.fsencode(path), C.byref(soa), C.byref(nodes), len(sc), S,
                                  len(vc), V, C.byref(meta)))


def checkpoint_append_amg(path, amg_host):
    """Appends an AmgHost hierarchy to a checkpoint file (the AMG trailer)."""
    check(load().mi_checkpoint_append_amg(os.fsencode(path), amg_host.h))


class Checkpoint:
    """A checkpoint file read into host memory (mi_checkpoint_open)."""

    def __init__(self, path):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.mi_checkpoint_open(os.fsencode(path), C.byref(h)))
        self.h = h.value
        meta = CheckpointMeta()
        nc, nf, nn = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.mi_checkpoint_info(self.h, C.byref(meta), C.byref(nc), C.byref(nf),
                                          C.byref(nn)))
        self.iteration, self.time, self.settings_hash = meta.iteration, meta.time, meta.settings_hash
        self.n_cells, self.n_faces, self.n_nodes = nc.value, nf.value, nn.value

    def scalar(self, name):
        cnt = C.c_int64()
        check(self.lib.mi_checkpoint_scalar(self.h, name.encode(), None, 0, C.byref(cnt)))
        out = np.empty(cnt.value)
        check(self.lib.mi_checkpoint_scalar(self.h, name.encode(), ptr(out), out.size, C.byref(cnt)))
        return out

    @property
    def amg_levels(self):
        """Levels of the AMG trailer (0: none)."""
        lv = C.c_int()
        check(self.lib.mi_checkpoint_amg_levels(self.h, C.byref(lv)))
        return lv.value

    def vector(self, name):
        cnt = C.c_int64()
        check(self.lib.mi_checkpoint_vector(self.h, name.encode(), None, None, None, 0,
                                            C.byref(cnt)))
        out = [np.empty(cnt.value) for _ in range(3)]
        check(self.lib.mi_checkpoint_vector(self.h, name.encode(), *(ptr(a) for a in out),
                                            cnt.value, C.byref(cnt)))
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_checkpoint_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """mi_mesh handle built from a simall_amd.workload.Mesh (host SoA arrays)."""

    def __init__(self, mesh, bc_of_zone, twin=None):
        self.lib = load()
        s = mesh_soa(mesh)
        self.host = mesh
        self.bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
        h = C.c_void_p()
        if twin is None:
            check(self.lib.mi_mesh_create(C.byref(s), ptr(self.bc), self.bc.size, C.byref(h)))
        else:
            tw = np.ascontiguousarray(twin, dtype=np.int32)
            check(self.lib.mi_mesh_create_twins(C.byref(s), ptr(self.bc), self.bc.size, ptr(tw),
                                                C.byref(h)))
        self.h = h.value
        nc, nf, nnz = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.mi_mesh_info(self.h, C.byref(nc), C.byref(nf), C.byref(nnz)))
        self.n_cells, self.n_faces, self.nnz = nc.value, nf.value, nnz.value

    def create_matrix(self):
        h = C.c_void_p()
        check(self.lib.mi_mesh_create_matrix(self.h, C.byref(h)))
        return CsrView(h.value)

    def assemble_pressure(self, A, d_rhs, rho=1.0, d_aP=None, d_gamma=None, d_phi=None):
        pinned = C.c_int()
        check(self.lib.mi_assemble_pressure(self.h, rho, _dptr(d_aP), _dptr(d_gamma), _dptr(d_phi),
                                            A.h, d_rhs.ptr, C.byref(pinned)))
        return bool(pinned.value)

    def assemble_momentum(self, A, d_rhs, comp, mu, rho, d_phi, d_pgradV=None, d_dt=None,
                          d_Uold=None, d_Ucomp=None, urf=1.0, bc_value=None):
        bv = None if bc_value is None else np.ascontiguousarray(bc_value, dtype=np.float64)
        check(self.lib.mi_assemble_momentum(self.h, comp, mu, rho, d_phi.ptr, _dptr(d_pgradV),
                                            _dptr(d_dt), _dptr(d_Uold), _dptr(d_Ucomp), urf,
                                            ptr(bv), A.h, d_rhs.ptr))

    def simplec_denominators(self, d_aP, A_mom, d_den):
        check(self.lib.mi_simplec_denominators(self.h, d_aP.ptr, A_mom.h, d_den.ptr))

    def face_gradient(self, d_x, d_lam, d_grad, pinned, rho=1.0, d_aP=None):
        """d_grad[f] = dJ/dgamma_f = -lambda^T dA/dgamma_f x (pressure matrix)."""
        check(self.lib.mi_face_gradient(self.h, rho, _dptr(d_aP), int(bool(pinned)), d_x.ptr,
                                        d_lam.ptr, d_grad.ptr))

    def face_flux(self, d_Ux, d_Uy, d_Uz, d_phi, rho=1.0, bc_value=None):
        bv = None if bc_value is None else np.ascontiguousarray(bc_value, dtype=np.float64)
        check(self.lib.mi_face_flux(self.h, rho, d_Ux.ptr, d_Uy.ptr, d_Uz.ptr, ptr(bv), d_phi.ptr))

    def ls_gradient(self, d_phi, d_gx, d_gy, d_gz):
        """LeastSquaresGradient::evaluate of a cell field (device arrays)."""
        check(self.lib.mi_ls_gradient(self.h, d_phi.ptr, d_gx.ptr, d_gy.ptr, d_gz.ptr))

    def correct_fields(self, urfP, rho, d_aP, d_pp, d_p, d_U, d_phi=None):
        """SimpleAlgorithm::correct_fields on device arrays (d_U: three arrays)."""
        check(self.lib.mi_correct_fields(self.h, urfP, rho, d_aP.ptr, d_pp.ptr, d_p.ptr, d_U[0].ptr,
                                         d_U[1].ptr, d_U[2].ptr, _dptr(d_phi)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AmgHost:
    """The SA-AMG hierarchy built on the host (inspection; no device needed),
    or (from_solver) a copy of the one an AMG solver applies."""

    A, P, R, COARSE_INV, AGG, STRONG = 0, 1, 2, 3, 4, 5

    def __init__(self, rowPtr=None, colIdx=None, values=None, cfg=None, handle=None):
        self.lib = load()
        if handle is None:
            rp = np.ascontiguousarray(rowPtr, dtype=np.int32)
            ci = np.ascontiguousarray(colIdx, dtype=np.int32)
            va = np.ascontiguousarray(values, dtype=np.float64)
            h = C.c_void_p()
            cfg = cfg if cfg is not None else default_cfg()
            check(self.lib.mi_amg_host_build(rp.size - 1, ptr(rp), ptr(ci), ptr(va), C.byref(cfg), C.byref(h)))
            handle = h.value
        self.h = handle
        nl, oc, gc = C.c_int(), C.c_double(), C.c_double()
        check(self.lib.mi_amg_host_levels(self.h, C.byref(nl), C.byref(oc), C.byref(gc)))
        self.levels, self.operator_complexity, self.grid_complexity = nl.value, oc.value, gc.value

    def info(self, level, which):
        nr, nc, nnz, lam = C.c_int64(), C.c_int64(), C.c_int64(), C.c_double()
        check(self.lib.mi_amg_host_level_info(self.h, level, which, C.byref(nr), C.byref(nc), C.byref(nnz), C.byref(lam)))
        return nr.value, nc.value, nnz.value, lam.value

    @classmethod
    def from_solver(cls, solver):
        h = C.c_void_p()
        check(load().mi_solver_amg_host(solver.h, C.byref(h)))
        return cls(handle=h.value)

    def aggregates(self, level):
        """(aggregate of every row of A_l, aggregate count)."""
        nr, nc, nnz, _ = self.info(level, self.AGG)
        out = np.empty(nnz, np.int32)
        check(self.lib.mi_amg_host_level_get(self.h, level, self.AGG, None, ptr(out), None))
        return out, nc

    def strong(self, level):
        """1 per entry of A_l that is a strong coupling of the filtered smoothing."""
        nr, nc, nnz, _ = self.info(level, self.STRONG)
        out = np.empty(nnz, np.int32)
        check(self.lib.mi_amg_host_level_get(self.h, level, self.STRONG, None, ptr(out), None))
        return out

    def matrix(self, level, which):
        """(rowPtr, colIdx, values, ncols) of A_l / P_l / R_l, or the dense coarse inverse."""
        nr, nc, nnz, _ = self.info(level, which)
        if which == self.COARSE_INV:
            v = np.empty(nnz)
            check(self.lib.mi_amg_host_level_get(self.h, level, which, None, None, ptr(v)))
            return v.reshape(nr, nc)
        rp = np.empty(nr + 1, np.int32)
        ci = np.empty(nnz, np.int32)
        v = np.empty(nnz)
        check(self.lib.mi_amg_host_level_get(self.h, level, which, ptr(rp), ptr(ci), ptr(v)))
        return rp, ci, v, nc

    def scipy(self, level, which):
        import scipy.sparse as sp
        rp, ci, v, nc = self.matrix(level, which)
        return sp.csr_matrix((v, ci, rp), shape=(rp.size - 1, nc))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_amg_host_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rcm_permutation(rowPtr, colIdx):
    rp = np.ascontiguousarray(rowPtr, dtype=np.int32)
    ci = np.ascontiguousarray(colIdx, dtype=np.int32)
    perm = np.empty(rp.size - 1, np.int32)
    bw = C.c_int32()
    check(load().mi_rcm_permutation(rp.size - 1, ptr(rp), ptr(ci), ptr(perm), C.byref(bw)))
    return perm, bw.value


MI_BATCH_SHARED_VALUES = 1
MI_BATCH_ILU0 = 2


def time_grid_barrier(workgroups, iterations=200):
    """(us per cooperative grid.sync(), us per kernel boundary) on this device."""
    a, b = C.c_double(), C.c_double()
    check(load().mi_time_grid_barrier(int(workgroups), int(iterations), C.byref(a), C.byref(b)))
    return a.value, b.value


def color_permutation(rowPtr, colIdx):
    """perm[new] = old of the multicolour ordering, and the colour count."""
    rp = np.ascontiguousarray(rowPtr, dtype=np.int32)
    ci = np.ascontiguousarray(colIdx, dtype=np.int32)
    perm = np.empty(rp.size - 1, np.int32)
    nc = C.c_int32()
    check(load().mi_color_permutation(rp.size - 1, ptr(rp), ptr(ci), ptr(perm), C.byref(nc)))
    return perm, nc.value


class Batch:
    """mi_batch: n_eq equations sharing the pattern of `A` (Csr or CsrView);
    flags: MI_BATCH_SHARED_VALUES (one matrix for every equation) and / or
    MI_BATCH_ILU0 (two-colour ordering, ILU(0) available)."""

    def __init__(self, A, n_eq, flags=0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.mi_batch_create_ex(A.h, int(n_eq), int(flags), C.byref(h)))
        self.h = h.value
        self.A = A
        self.flags = int(flags)
        self.n_eq, self.n, self.nnz = int(n_eq), A.n, A.nnz

    def set_equation(self, e, A):
        check(self.lib.mi_batch_set_equation(self.h, int(e), A.h))

    def set_matrix(self, A):
        check(self.lib.mi_batch_set_matrix(self.h, A.h))

    def set_values(self, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(self.n_eq, self.nnz)
        check(self.lib.mi_batch_set_values(self.h, ptr(v)))

    def _cfg(self, cfg, kw):
        c = cfg if cfg is not None else default_cfg(kind=MI_SOLVER_CG, preconditioner=MI_PC_JACOBI)
        for k, v in kw.items():
            setattr(c, k, v)
        return c

    def solve(self, b, cfg=None, **kw):
        """b: (n_eq, n) -> (x, iterations, rel_residual, status)."""
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(self.n_eq, self.n)
        x = np.empty_like(b)
        its = np.empty(self.n_eq, np.int32)
        rel = np.empty(self.n_eq, np.float64)
        c = self._cfg(cfg, kw)
        rc = self.lib.mi_batch_solve(self.h, C.byref(c), ptr(b), ptr(x), ptr(its), ptr(rel))
        check(rc, allowed=(MI_OK, MI_ERR_NOT_CONVERGED))
        return x, its, rel, rc

    def solve_device(self, db, dx, cfg=None, **kw):
        its = np.empty(self.n_eq, np.int32)
        rel = np.empty(self.n_eq, np.float64)
        c = self._cfg(cfg, kw)
        rc = self.lib.mi_batch_solve_device(self.h, C.byref(c), db.ptr, dx.ptr, ptr(its), ptr(rel))
        check(rc, allowed=(MI_OK, MI_ERR_NOT_CONVERGED))
        return its, rel, rc

    def apply_preconditioner(self, r, preconditioner):
        """z = M^-1 r per equation; r: (n_eq, n) host array."""
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.n_eq, self.n)
        dr = DeviceArray.from_host(r.ravel())
        dz = DeviceArray(r.size)
        check(self.lib.mi_batch_apply_preconditioner(self.h, int(preconditioner), dr.ptr, dz.ptr))
        return dz.download().reshape(self.n_eq, self.n)

    def tuning(self):
        rb, us, by = C.c_int(), C.c_double(), C.c_double()
        check(self.lib.mi_batch_tuning(self.h, C.byref(rb), C.byref(us), C.byref(by)))
        v, g, bd, gv, vus = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_double()
        check(self.lib.mi_batch_tuning_shape(self.h, C.byref(v), C.byref(g), C.byref(bd),
                                             C.byref(gv), C.byref(vus)))
        return {"rows_per_block": rb.value, "spmm_us": us.value, "spmm_bytes": by.value,
                "slots_per_thread_max": v.value, "spmm_workgroups": g.value,
                "banded": bd.value, "vector_workgroups": gv.value, "vector_us": vus.value}

    def set_shape(self, slots_per_thread_max, spmm_workgroups, banded):
        """Pins the SpMM shape (mi_batch_set_shape): same shape, same bits."""
        check(self.lib.mi_batch_set_shape(self.h, int(slots_per_thread_max),
                                          int(spmm_workgroups), int(banded)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def flow_cfg(algorithm=MI_FLOW_SIMPLE, momentum=None, pressure=None, **kw):
    """mi_flow_cfg with the defaults of the algorithm; momentum / pressure are
    dicts of SolverCfg fields."""
    c = FlowCfg()
    load().mi_flow_cfg_default(algorithm, C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    for sub, fields in (("momentum", momentum), ("pressure", pressure)):
        for k, v in (fields or {}).items():
            if not hasattr(getattr(c, sub), k):
                raise AttributeError(f"{sub}.{k}")
            setattr(getattr(c, sub), k, v)
    return c


class Flow:
    """mi_flow: device-resident SIMPLE / PISO iteration on a DeviceMesh."""

    def __init__(self, dmesh, cfg=None, bc_value=None):
        self.lib = load()
        self.mesh = dmesh
        self.cfg = cfg if cfg is not None else flow_cfg()
        self.bv = None if bc_value is None else np.ascontiguousarray(bc_value, dtype=np.float64)
        h = C.c_void_p()
        check(self.lib.mi_flow_create(dmesh.h, ptr(self.bv), C.byref(self.cfg), C.byref(h)))
        self.h = h.value
        p = [C.c_void_p() for _ in range(5)]
        check(self.lib.mi_flow_fields(self.h, *(C.byref(x) for x in p)))
        self._ptrs = [x.value for x in p]

    def _get(self, k, n):
        out = np.empty(n)
        check(self.lib.mi_memcpy_d2h(ptr(out), self._ptrs[k], out.nbytes))
        return out

    def _set(self, k, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(self.lib.mi_memcpy_h2d(self._ptrs[k], ptr(a), a.nbytes))

    @property
    def U(self):
        return [self._get(k, self.mesh.n_cells) for k in range(3)]

    @U.setter
    def U(self, U):
        for k in range(3):
            self._set(k, U[k])

    @property
    def p(self):
        return self._get(3, self.mesh.n_cells)

    @p.setter
    def p(self, v):
        self._set(3, v)

    @property
    def phi(self):
        return self._get(4, self.mesh.n_faces)

    def iterate(self):
        r = FlowResiduals()
        check(self.lib.mi_flow_iterate(self.h, C.byref(r)))
        return r.as_dict()

    def advance_time_step(self, n_inner=1):
        r = FlowResiduals()
        check(self.lib.mi_flow_advance_time_step(self.h, n_inner, C.byref(r)))
        return r.as_dict()

    def checkpoint_write(self, path, iteration=0, time=0.0, settings_hash=0, amg=False):
        """Writes p and U with the DeviceMesh's host mesh (mi_flow_checkpoint_write);
        amg=True appends the pressure solver's AMG hierarchy."""
        m = self.mesh.host
        soa, nodes = mesh_soa(m), mesh_nodes(m)
        meta = CheckpointMeta(iteration, time, settings_hash)
        check(self.lib.mi_flow_checkpoint_write(self.h, os.fsencode(path), C.byref(soa),
                                                C.byref(nodes), C.byref(meta),
                                                MI_CHECKPOINT_AMG if amg else 0))

    def checkpoint_read(self, path):
        """Loads p and U; returns (iteration, time, settings_hash)."""
        meta = CheckpointMeta()
        check(self.lib.mi_flow_checkpoint_read(self.h, os.fsencode(path), C.byref(meta)))
        return meta.iteration, meta.time, meta.settings_hash

    def close(self):
        if getattr(self, "h", None):
            self.lib.mi_flow_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

