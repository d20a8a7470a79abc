"""CPU oracle of the SimAll pressure-correction solve path (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker: the product (libsimall_amd.so) never
calls it.  ``liboracle.so`` is built from oracle.c by oracle/Makefile.

Each function restates one reference routine (file:line in oracle.h/oracle.c).
How each is pinned (tests/test_oracle.py):

  spmv               reference KAT y=[2,4,6,13] (tests/unit/gpu/test_gpu_models.cpp:144-178),
                     dense brute force, A.1 = 0 on Neumann rows
  dot/norm2/axpy     KAT dot(1..5,1..5)=55, nrm2=sqrt(55) (test_gpu_models.cpp:117-141),
                     compensated sums
  jacobi             closed form 1/a_ii, missing/zero diagonal -> 1
  cg                 KAT tridiag(-1,4,-1) n=10 (<50 its, ||r||<1e-8); exact answer vs
                     numpy.linalg.solve; finite termination in <= n steps; Thomas algorithm
  bicgstab + ilu0    KAT n=8 nonsymmetric (||r||<1e-7, test_gpu_models.cpp:238-287); ILU(0) of a
                     tridiagonal is its exact LU (scipy.linalg.lu); (LU)_ij = a_ij on the pattern
  ssor               identity -> r (tests/unit/solver/test_ssor_preconditioner.cpp:44-63);
                     CG+SSOR 1-D Laplacian n=50 (:65-88)
  gmres              exact solution vs numpy on random nonsymmetric systems
  tfqmr              reference KATs: Jacobi on the 1-D Laplacian n=30 (||r||<1e-5) and ILU on
                     upwind convection-diffusion n=40, Pe=5 (||r||<1e-4)
                     (tests/unit/solver/test_tfqmr_solver.cpp:66-111); dense solves
  gmres_transpose    KAT psi=(0.5,1/6) (tests/unit/longtail/test_longtail_modules.cpp:32-54)
  build_sparsity     brute-force neighbour sets from the face lists
  assemble_pressure  uniform grid, unit gamma: diagonal = #interior neighbours, off-diag -1,
                     zero row sums before the pin, symmetry; coefficient formula restated
                     independently in numpy
  assemble_momentum  zero-flux case equals pressure-style diffusion; upwind row sums equal net
                     outflow; M-matrix sign pattern
  face_flux          closed form for a uniform velocity field (phi = U.S)
  face_gradient      central finite differences of J = sum w x^2 through assemble_pressure and a
                     dense solve (pinned and pressure-outlet cases); closed form
                     -G (lam_o - lam_n)(x_o - x_n) on unpinned faces (tests/test_adjoint_oracle.py)
  ls_gradient        exact gradient of a linear field on cells without boundary faces;
                     bit-exact vs LeastSquaresGradient compiled from the reference
  correct_fields     constant p' leaves U and phi unchanged; bit-exact vs correct_fields
                     compiled from the reference
  Flow               SIMPLE / PISO iterations vs SimpleAlgorithm / PisoAlgorithm compiled from
                     the reference with a dense direct solve as their linear solver
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

PC_JACOBI, PC_ILU, PC_SSOR, PC_AMG, PC_NONE = 0, 1, 2, 3, 4


class OrcCsr(C.Structure):
    _fields_ = [("n", C.c_int), ("rowPtr", C.c_void_p), ("colIdx", C.c_void_p), ("values", C.c_void_p)]


class OrcMesh(C.Structure):
    _fields_ = [("n_cells", C.c_int64), ("n_faces", C.c_int64)] + [
        (k, C.c_void_p)
        for k in ("owner", "neighbor", "areaX", "areaY", "areaZ", "faceCx", "faceCy", "faceCz",
                  "boundaryZone", "cellCx", "cellCy", "cellCz", "volume", "cellFaceOffsets",
                  "cellFaceIndices")
    ]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = C.CDLL(LIB_PATH)
        P = C.POINTER
        vp, i, d, i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        sig = {
            "orc_spmv": (None, [P(OrcCsr), vp, vp]),
            "orc_spmv_transpose": (None, [P(OrcCsr), vp, vp]),
            "orc_dot": (d, [vp, vp, i64]),
            "orc_norm2": (d, [vp, i64]),
            "orc_axpy": (None, [d, vp, vp, i64]),
            "orc_scal": (None, [d, vp, i64]),
            "orc_jacobi_setup": (None, [P(OrcCsr), vp]),
            "orc_ilu0_setup": (None, [P(OrcCsr), vp, vp]),
            "orc_ilu0_apply": (None, [P(OrcCsr), vp, vp, vp, vp]),
            "orc_ssor_apply": (None, [P(OrcCsr), vp, d, vp, vp]),
            "orc_cg_solve": (i, [P(OrcCsr), i, d, i, vp, vp, P(d), vp, P(i)]),
            "orc_bicgstab_solve": (i, [P(OrcCsr), i, d, i, vp, vp, P(d), vp, P(i)]),
            "orc_gmres_solve": (i, [P(OrcCsr), i, d, i, i, vp, vp, P(d)]),
            "orc_tfqmr_solve": (i, [P(OrcCsr), i, d, i, vp, vp, P(d)]),
            "orc_gmres_transpose": (i, [P(OrcCsr), vp, vp, i, i, d, P(i), P(d)]),
            "orc_build_sparsity": (i64, [P(OrcMesh), vp, vp]),
            "orc_assemble_pressure": (i, [P(OrcMesh), vp, i, d, vp, vp, vp, P(OrcCsr), vp, vp]),
            "orc_assemble_momentum": (None, [P(OrcMesh), vp, i, vp, i, d, d, vp, vp, vp, vp, vp,
                                             vp, d, P(OrcCsr), vp, vp]),
            "orc_ls_gradient": (None, [P(OrcMesh), vp, vp, vp, vp]),
            "orc_correct_fields": (None, [P(OrcMesh), d, d, vp, vp, vp, vp, vp, vp, vp]),
            "orc_face_flux": (None, [P(OrcMesh), vp, i, vp, d, vp, vp, vp, vp]),
            "orc_face_gradient": (None, [P(OrcMesh), vp, i, d, vp, i, vp, vp, vp]),
            "orc_periodic_twins": (None, [P(OrcMesh), vp, i, vp]),
            "orc_simplec_denominators": (None, [P(OrcCsr), vp, vp]),
            "orc_build_sparsity_p": (i64, [P(OrcMesh), vp, vp, vp]),
            "orc_assemble_pressure_p": (i, [P(OrcMesh), vp, i, vp, d, vp, vp, vp, P(OrcCsr), vp,
                                            vp]),
            "orc_assemble_momentum_p": (None, [P(OrcMesh), vp, i, vp, vp, i, d, d, vp, vp, vp, vp,
                                               vp, vp, d, P(OrcCsr), vp, vp]),
            "orc_face_flux_p": (None, [P(OrcMesh), vp, i, vp, vp, d, vp, vp, vp, vp]),
            "orc_face_gradient_p": (None, [P(OrcMesh), vp, i, vp, d, vp, i, vp, vp, vp]),
            "orc_splitmix64": (C.c_uint64, [P(C.c_uint64)]),
            "orc_uniform": (d, [P(C.c_uint64)]),
            "orc_normal": (d, [P(C.c_uint64)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class Matrix:
    """CSR triple kept alive for the C structs."""

    def __init__(self, rowPtr, colIdx, values):
        self.rowPtr = np.ascontiguousarray(rowPtr, dtype=np.int32)
        self.colIdx = np.ascontiguousarray(colIdx, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.n = self.rowPtr.size - 1
        self.s = OrcCsr(self.n, _p(self.rowPtr), _p(self.colIdx), _p(self.values))

    @classmethod
    def from_dense(cls, A):
        A = np.asarray(A, dtype=np.float64)
        rp, ci, va = [0], [], []
        for i in range(A.shape[0]):
            for j in range(A.shape[1]):
                if A[i, j] != 0.0 or i == j:
                    ci.append(j)
                    va.append(A[i, j])
            rp.append(len(ci))
        return cls(rp, ci, va)

    def diag_positions(self):
        """Position of a_ii in each row (every row must hold it)."""
        rows = np.repeat(np.arange(self.n), np.diff(self.rowPtr))
        pos = np.flatnonzero(self.colIdx == rows)
        assert pos.size == self.n, "a row without its diagonal"
        return pos

    def dense(self):
        D = np.zeros((self.n, self.n))
        for i in range(self.n):
            for k in range(self.rowPtr[i], self.rowPtr[i + 1]):
                D[i, self.colIdx[k]] += self.values[k]
        return D


def _vec(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def spmv(A, x):
    x = _vec(x)
    y = np.empty(A.n)
    load().orc_spmv(C.byref(A.s), _p(x), _p(y))
    return y


def spmv_transpose(A, x):
    x = _vec(x)
    y = np.empty(A.n)
    load().orc_spmv_transpose(C.byref(A.s), _p(x), _p(y))
    return y


def dot(a, b):
    a, b = _vec(a), _vec(b)
    return load().orc_dot(_p(a), _p(b), a.size)


def norm2(a):
    a = _vec(a)
    return load().orc_norm2(_p(a), a.size)


def axpy(alpha, x, y):
    x = _vec(x)
    y = np.array(y, dtype=np.float64, copy=True)
    load().orc_axpy(alpha, _p(x), _p(y), x.size)
    return y


def scal(a, x):
    x = np.array(x, dtype=np.float64, copy=True)
    load().orc_scal(a, _p(x), x.size)
    return x


def jacobi_inv_diag(A):
    out = np.empty(A.n)
    load().orc_jacobi_setup(C.byref(A.s), _p(out))
    return out


def ilu0(A):
    lu = np.empty(A.values.size)
    dp = np.empty(A.n, np.int32)
    load().orc_ilu0_setup(C.byref(A.s), _p(lu), _p(dp))
    return lu, dp


def ilu0_apply(A, lu, dp, r):
    r = _vec(r)
    z = np.empty(A.n)
    load().orc_ilu0_apply(C.byref(A.s), _p(lu), _p(dp), _p(r), _p(z))
    return z


def ssor_apply(A, r, omega=1.0):
    dp = np.full(A.n, -1, np.int32)
    for i in range(A.n):
        for k in range(A.rowPtr[i], A.rowPtr[i + 1]):
            if A.colIdx[k] == i:
                dp[i] = k
                break
    r = _vec(r)
    z = np.empty(A.n)
    load().orc_ssor_apply(C.byref(A.s), _p(dp), omega, _p(r), _p(z))
    return z


def _krylov(fn, A, b, pc, tol, max_iter):
    b = _vec(b)
    x = np.zeros(A.n)
    last = C.c_double()
    hist = np.zeros(max(max_iter, 1) + 1)
    nh = C.c_int()
    it = fn(C.byref(A.s), pc, tol, max_iter, _p(b), _p(x), C.byref(last), _p(hist), C.byref(nh))
    if it < 0:
        raise NotImplementedError("preconditioner not restated in the oracle")
    return x, it, last.value, hist[: nh.value].copy()


def cg(A, b, pc=PC_JACOBI, tol=1e-8, max_iter=1000):
    """CGSolver::solve -> (x, iterations, last_residual, history)."""
    return _krylov(load().orc_cg_solve, A, b, pc, tol, max_iter)


def bicgstab(A, b, pc=PC_ILU, tol=1e-8, max_iter=1000):
    """BiCGSTABSolver::solve -> (x, iterations, last_residual, history)."""
    return _krylov(load().orc_bicgstab_solve, A, b, pc, tol, max_iter)


def gmres(A, b, pc=PC_ILU, tol=1e-8, max_iter=1000, restart=30):
    b = _vec(b)
    x = np.zeros(A.n)
    last = C.c_double()
    it = load().orc_gmres_solve(C.byref(A.s), pc, tol, max_iter, restart, _p(b), _p(x), C.byref(last))
    return x, it, last.value


def tfqmr(A, b, pc=PC_JACOBI, tol=1e-8, max_iter=1000):
    """TfqmrSolver::solve (LinearSolvers.cpp:497-603) -> (x, iterations, last)."""
    b = _vec(b)
    x = np.zeros(A.n)
    last = C.c_double()
    it = load().orc_tfqmr_solve(C.byref(A.s), pc, tol, max_iter, _p(b), _p(x), C.byref(last))
    return x, it, last.value


def gmres_transpose(A, b, max_iter=500, restart=30, tol=1e-10):
    """gmres_transpose (DiscreteAdjoint.cpp:54-146) -> (x, ok, iterations, residual)."""
    b = _vec(b)
    x = np.zeros(A.n)
    its = C.c_int()
    res = C.c_double()
    ok = load().orc_gmres_transpose(C.byref(A.s), _p(b), _p(x), max_iter, restart, tol, C.byref(its), C.byref(res))
    return x, bool(ok), its.value, res.value


# ------------------------------------------------------------- assembly
def mesh_struct(mesh):
    """OrcMesh over a simall_amd.workload.Mesh (arrays must stay alive)."""
    s = OrcMesh()
    s.n_cells, s.n_faces = mesh.n_cells, mesh.n_faces
    for k in ("owner", "neighbor", "areaX", "areaY", "areaZ", "faceCx", "faceCy", "faceCz",
              "boundaryZone", "cellCx", "cellCy", "cellCz", "volume", "cellFaceOffsets",
              "cellFaceIndices"):
        setattr(s, k, getattr(mesh, k).ctypes.data)
    return s


BC_PERIODIC = 5


def _twin(twin):
    return None if twin is None else np.ascontiguousarray(twin, dtype=np.int32)


def periodic_twins(mesh, bc_of_zone):
    """SimpleAlgorithm::build_periodic_pairs (SimpleAlgorithm.cpp:101-186):
    twin[f] = the paired face of a periodic boundary face, else -1."""
    ms = mesh_struct(mesh)
    bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
    twin = np.empty(mesh.n_faces, np.int32)
    load().orc_periodic_twins(C.byref(ms), _p(bc), bc.size, _p(twin))
    return twin


def simplec_denominators(pattern, amom_values, aP):
    """max(aP_c - sum_{k != c} |A_mom[c, k]|, 1e-30) (SimpleAlgorithm.cpp:378-387)."""
    A = Matrix(pattern.rowPtr, pattern.colIdx, amom_values)
    den = np.empty(pattern.n)
    load().orc_simplec_denominators(C.byref(A.s), _p(_vec(aP)), _p(den))
    return den


def build_sparsity(mesh, twin=None):
    ms = mesh_struct(mesh)
    tw = _twin(twin)
    rp = np.empty(mesh.n_cells + 1, np.int32)
    nnz = load().orc_build_sparsity_p(C.byref(ms), _p(tw), _p(rp), None)
    ci = np.empty(nnz, np.int32)
    load().orc_build_sparsity_p(C.byref(ms), _p(tw), _p(rp), _p(ci))
    return rp, ci


def assemble_pressure(mesh, pattern, bc_of_zone, rho=1.0, aP=None, gamma=None, phi=None,
                      twin=None):
    """-> (values, rhs, pinned)."""
    ms = mesh_struct(mesh)
    bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
    tw = _twin(twin)
    vals = np.empty(pattern.colIdx.size)
    rhs = np.empty(mesh.n_cells)
    pinned = load().orc_assemble_pressure_p(C.byref(ms), _p(bc), bc.size, _p(tw), rho, _p(aP),
                                            _p(gamma), _p(phi), C.byref(pattern.s), _p(vals),
                                            _p(rhs))
    return vals, rhs, bool(pinned)


def face_gradient(mesh, bc_of_zone, x, lam, pinned, rho=1.0, aP=None, twin=None):
    """dJ/dgamma_f = -lambda^T (dA/dgamma_f) x for every face (pressure matrix)."""
    ms = mesh_struct(mesh)
    bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
    tw = _twin(twin)
    x = _vec(x)
    lam = _vec(lam)
    aP = None if aP is None else _vec(aP)
    grad = np.empty(mesh.n_faces)
    load().orc_face_gradient_p(C.byref(ms), _p(bc), bc.size, _p(tw), rho, _p(aP), int(pinned),
                               _p(x), _p(lam), _p(grad))
    return grad


def assemble_momentum(mesh, pattern, bc_of_zone, comp, mu, rho, phi, dt=None, Uold=None,
                      Ucomp=None, urf=1.0, bc_value=None, pgradV=None, Unm1=None, twin=None):
    ms = mesh_struct(mesh)
    bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
    tw = _twin(twin)
    vals = np.empty(pattern.colIdx.size)
    rhs = np.empty(mesh.n_cells)
    bv = None if bc_value is None else np.ascontiguousarray(bc_value, dtype=np.float64)
    opt = [None if a is None else _vec(a) for a in (pgradV, dt, Uold, Unm1, Ucomp)]
    load().orc_assemble_momentum_p(C.byref(ms), _p(bc), bc.size, _p(tw), _p(bv), comp, mu, rho,
                                   _p(_vec(phi)), *(_p(a) for a in opt), urf, C.byref(pattern.s),
                                   _p(vals), _p(rhs))
    return vals, rhs


def ls_gradient(mesh, phi):
    """LeastSquaresGradient::evaluate (Gradient.cpp:93-136) -> (gx, gy, gz)."""
    ms = mesh_struct(mesh)
    g = np.empty((3, mesh.n_cells))
    load().orc_ls_gradient(C.byref(ms), _p(_vec(phi)), _p(g[0]), _p(g[1]), _p(g[2]))
    return g[0].copy(), g[1].copy(), g[2].copy()


def correct_fields(mesh, urfP, rho, aP, pp, p, U, phi):
    """SimpleAlgorithm::correct_fields (SimpleAlgorithm.cpp:450-489) on copies
    -> (p, (Ux, Uy, Uz), phi)."""
    ms = mesh_struct(mesh)
    p = np.array(p, dtype=np.float64)
    U = [np.array(u, dtype=np.float64) for u in U]
    phi = None if phi is None else np.array(phi, dtype=np.float64)
    load().orc_correct_fields(C.byref(ms), urfP, rho, _p(_vec(aP)), _p(_vec(pp)), _p(p), _p(U[0]),
                              _p(U[1]), _p(U[2]), _p(phi))
    return p, tuple(U), phi


class Flow:
    """SimpleAlgorithm / PisoAlgorithm state and outer iteration
    (SimpleAlgorithm.cpp:49-98, 491-575; PisoAlgorithm.cpp:27-70) over the
    oracle's assembly, gradient and correction, with exact sparse solves in
    place of the linear solvers."""

    def __init__(self, mesh, bc_of_zone, bc_value=None, algorithm="simple", rho=1.0, mu=1e-3,
                 urfU=None, urfP=None, dt=None, bdf2=None, n_correctors=2, n_nonorth=0):
        piso = algorithm == "piso"
        self.simplec = algorithm == "simplec"
        self.mesh, self.bc, self.bc_value = mesh, list(bc_of_zone), bc_value
        self.piso, self.rho, self.mu = piso, rho, mu
        # SimpleOptions / PisoOptions / SimplecOptions defaults
        self.urfU = (1.0 if piso else 0.9 if self.simplec else 0.7) if urfU is None else urfU
        self.urfP = (1.0 if piso or self.simplec else 0.3) if urfP is None else urfP
        self.dt = (1e-3 if piso else 0.0) if dt is None else dt
        self.bdf2 = piso if bdf2 is None else bdf2
        self.n_correctors, self.n_nonorth = max(1, n_correctors), n_nonorth
        n = mesh.n_cells
        self.twin = periodic_twins(mesh, self.bc) if BC_PERIODIC in self.bc else None
        rp, ci = build_sparsity(mesh, self.twin)
        self.pattern = Matrix(rp, ci, np.zeros(ci.size))
        self.U = [np.zeros(n) for _ in range(3)]
        self.Un = [np.zeros(n) for _ in range(3)]
        self.Unm1 = [np.zeros(n) for _ in range(3)]
        self.p = np.zeros(n)
        self.aP = np.zeros(n)
        self.phi = np.zeros(mesh.n_faces)
        self.step = 0

    def _solve(self, vals, b):
        import scipy.sparse as sp
        import scipy.sparse.linalg as spl
        P = self.pattern
        A = sp.csr_matrix((vals, P.colIdx, P.rowPtr), shape=(P.n, P.n)).tocsc()
        return spl.spsolve(A, b), A

    def _momentum(self, res):
        m = self.mesh
        self.phi = face_flux(m, self.bc, self.rho, self.U, self.bc_value, self.twin)
        g = ls_gradient(m, self.p)
        dtc = np.full(m.n_cells, self.dt) if self.dt > 0 else None
        use_bdf2 = self.dt > 0 and self.bdf2 and self.step >= 1
        for k in range(3):
            vals, rhs = assemble_momentum(m, self.pattern, self.bc, k, self.mu, self.rho, self.phi,
                                          dt=dtc, Uold=self.Un[k] if dtc is not None else None,
                                          Ucomp=self.U[k], urf=self.urfU, bc_value=self.bc_value,
                                          pgradV=g[k] * m.volume,
                                          Unm1=self.Unm1[k] if use_bdf2 else None, twin=self.twin)
            if k == 0:
                self.aP = vals[self.pattern.diag_positions()].copy()
            self.amom = vals
            x, A = self._solve(vals, rhs)
            self.U[k] = x
            res["mom"][k] = float(np.linalg.norm(A @ x - rhs))

    def _pressure(self, res):
        m = self.mesh
        den = simplec_denominators(self.pattern, self.amom, self.aP) if self.simplec else self.aP
        vals, rhs, _ = assemble_pressure(m, self.pattern, self.bc, self.rho, aP=den, phi=self.phi,
                                         twin=self.twin)
        pp, _ = self._solve(vals, rhs)
        res["cont"] = float(np.linalg.norm(rhs))
        self.p, U, self.phi = correct_fields(m, self.urfP, self.rho, self.aP, pp, self.p, self.U,
                                             self.phi)
        self.U = list(U)

    def iterate(self):
        res = {"mom": [0.0, 0.0, 0.0], "cont": 0.0}
        self._momentum(res)
        if self.piso:
            for _ in range(self.n_correctors):
                self.phi = face_flux(self.mesh, self.bc, self.rho, self.U, self.bc_value,
                                     self.twin)
                for _ in range(self.n_nonorth + 1):
                    self._pressure(res)
        else:
            self.phi = face_flux(self.mesh, self.bc, self.rho, self.U, self.bc_value, self.twin)
            self._pressure(res)
        return res

    def advance_time_step(self, n_inner=1):
        if self.step >= 1:
            self.Unm1 = [u.copy() for u in self.Un]
        self.Un = [u.copy() for u in self.U]
        res = None
        for _ in range(max(1, n_inner)):
            res = self.iterate()
        self.step += 1
        return res


def face_flux(mesh, bc_of_zone, rho, U, bc_value=None, twin=None):
    ms = mesh_struct(mesh)
    bc = np.ascontiguousarray(bc_of_zone, dtype=np.int32)
    tw = _twin(twin)
    phi = np.empty(mesh.n_faces)
    bv = None if bc_value is None else np.ascontiguousarray(bc_value, dtype=np.float64)
    Ux, Uy, Uz = (_vec(u) for u in U)
    load().orc_face_flux_p(C.byref(ms), _p(bc), bc.size, _p(tw), _p(bv), rho, _p(Ux), _p(Uy),
                           _p(Uz), _p(phi))
    return phi
