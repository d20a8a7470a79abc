"""The FP64 GPU path against the FP64 oracle, config by config, with both
sides solved tightly: the solutions agree to 1e-10 relative.

Both sides stop on their own FP64 residual (the GPU on its true residual at
1e-12, the oracle's Jacobi-PCG / ILU(0)-BiCGStab at 1e-13); the systems are
the configs' generators at sizes the oracle finishes in seconds (config 1 at
its full 128x128 size).  On these systems the oracle at 1e-12 lies within
4e-13 (cavity), 4e-12 (32^3) and 1.1e-11 (48^3) of a sparse direct solve, so
1e-10 is attainable with a margin of 10 or more; 1e-13 is not asked of the
GPU, whose true-residual floor is about 1e-13 at these sizes.

The mixed form (FP32 inner iterations, FP64 refinement) meets the same bound
when its FP64 refinement is driven to the same tight tolerance.  The last
test runs the headline mixed solve at the full 128^3 size: FP64 true residual
<= 1e-8 and its distance to a tight FP64 GPU solve.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from simall_amd import capi
from simall_amd import workload as wl

pytestmark = pytest.mark.gpu

MATCH = 1e-10


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def pressure_system(w):
    """Oracle assembly of a pressure workload, pinned as the reference pins."""
    m = w.mesh
    rp, ci = oracle.build_sparsity(m)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, pinned = oracle.assemble_pressure(m, P, w.bc_of_zone, gamma=w.gamma)
    assert pinned
    b = w.b.copy()
    b[0] = 0.0
    return oracle.Matrix(rp, ci, vals), b


def gpu_solve(Ao, b, **kw):
    cfg = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-12,
               max_iterations=20000, true_residual=1, max_outer=30)
    cfg.update(kw)
    s = capi.Solver(**cfg)
    s.setup(capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values))
    x, st, rc = s.solve(b)
    assert rc == capi.MI_OK, (cfg, st.true_relative_residual)
    return x, st


@pytest.fixture(scope="module")
def cavity():
    Ao, b = pressure_system(wl.cavity2d(128))
    xo, ito, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-13, 20000)
    return Ao, b, xo


@pytest.mark.parametrize("pc", [capi.MI_PC_JACOBI, capi.MI_PC_AMG])
def test_config1_cavity_fp64_matches_oracle(gpu, cavity, pc):
    Ao, b, xo = cavity
    x, st = gpu_solve(Ao, b, preconditioner=pc)
    e = rel(x, xo)
    print(f"config 1 (128^2) pc={pc}: {st.iterations} its, true rel res "
          f"{st.true_relative_residual:.2e}, |x - x_oracle|/|x_oracle| = {e:.2e}")
    assert e <= MATCH


@pytest.fixture(scope="module")
def poisson48():
    Ao, b = pressure_system(wl.poisson3d(48))
    xo, ito, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-13, 50000)
    return Ao, b, xo


@pytest.mark.parametrize("prec", [capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED])
@pytest.mark.parametrize("pc", [capi.MI_PC_AMG, capi.MI_PC_JACOBI])
def test_config2_poisson48_matches_oracle(gpu, poisson48, pc, prec):
    """The config-2 generator (perturbed nodes, permuted cells, log-uniform
    conductances over 4 decades) at 48^3 = 110,592 rows; the solver reorders
    it (RCM) internally and returns x in the caller's numbering."""
    Ao, b, xo = poisson48
    x, st = gpu_solve(Ao, b, preconditioner=pc, precision=prec)
    e = rel(x, xo)
    print(f"config 2 (48^3) pc={pc} prec={prec}: {st.iterations} its in {st.outer_iterations} "
          f"sweeps, true rel res {st.true_relative_residual:.2e}, rel err {e:.2e}")
    assert e <= MATCH


def test_config3_momentum_matches_oracle(gpu):
    """The config-3 system (Taylor-Green upwind convection + diffusion,
    pseudo-transient diagonal) at 32^3, ILU(0)-BiCGStab in the caller's order
    (the reference's factorization bit for bit) and under the multicolour
    ordering, FP64 and mixed."""
    w = wl.momentum3d(32)
    m = w.mesh
    phi = oracle.face_flux(m, w.bc_of_zone, 1.0, w.U)
    rp, ci = oracle.build_sparsity(m)
    pat = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _ = oracle.assemble_momentum(m, pat, w.bc_of_zone, 0, w.mu, 1.0, phi, dt=w.dt)
    Ao = oracle.Matrix(rp, ci, vals)
    xo, ito, lasto, _ = oracle.bicgstab(Ao, w.b, oracle.PC_ILU, 1e-13, 5000)
    S = sp.csr_matrix((vals, ci, rp), shape=(m.n_cells,) * 2)
    assert np.linalg.norm(w.b - S @ xo) <= 1e-12 * np.linalg.norm(w.b)
    for reorder, prec in ((0, capi.MI_PRECISION_FP64), (2, capi.MI_PRECISION_FP64),
                          (2, capi.MI_PRECISION_MIXED)):
        x, st = gpu_solve(Ao, w.b, kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU,
                          reorder=reorder, precision=prec)
        e = rel(x, xo)
        print(f"config 3 (32^3) reorder={reorder} prec={prec}: {st.iterations} its "
              f"(oracle {ito}), rel err {e:.2e}")
        assert e <= MATCH


def test_config4_batched_transport_matches_oracle(gpu):
    """Four equations of the config-4 generator (V/dt I + Gamma_e L, Gamma_e
    log-uniform over [1e-6, 1e-3]) at 32^3 in one batched solve driven to a
    1e-12 FP64 true residual: each solution against the oracle's own solve."""
    w = wl.transport(32, 4)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    phi = capi.DeviceArray.from_host(np.zeros(m.n_faces))
    dt = capi.DeviceArray.from_host(np.full(m.n_cells, w.dt))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    rp, ci = A.pattern()
    vals = []
    for e in range(4):
        dm.assemble_momentum(A, rhs, 0, float(w.gammas[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        vals.append(A.values())
    B = capi.Batch(A, 4)
    B.set_values(np.stack(vals))
    bs = np.stack(w.bs[:4])
    x, its, relres, rc = B.solve(bs, tolerance=1e-12, max_iterations=5000, max_outer=40)
    assert rc == capi.MI_OK and np.all(relres <= 1e-12)
    for e in range(4):
        xo, ito, _, _ = oracle.cg(oracle.Matrix(rp, ci, vals[e]), bs[e], oracle.PC_JACOBI, 1e-13,
                                  20000)
        err = rel(x[e], xo)
        print(f"config 4 (32^3) equation {e}: {its[e]} its (oracle {ito}), rel err {err:.2e}")
        assert err <= MATCH


def test_config5_adjoint_matches_oracle(gpu):
    """Config 5 at 24^3: the forward solve, the adjoint A^T lambda = 2 w x and
    the per-face gradient against the oracle.  The pinned pressure matrix is
    symmetric (the pin scales only a_00), so the oracle's tight Jacobi-PCG on
    A^T gives the reference adjoint; the gradient is the oracle's face loop on
    the oracle's x and lambda."""
    w = wl.poisson3d(24)
    m = w.mesh
    Ao, b = pressure_system(w)
    S = sp.csr_matrix((Ao.values, Ao.colIdx, Ao.rowPtr), shape=(Ao.n,) * 2)
    assert (S != S.T).nnz == 0
    wts = wl.uniform(wl.seed(5, wl.Q_WEIGHTS), 0.0, 1.0, m.n_cells)
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    d_rhs = capi.DeviceArray(m.n_cells)
    pinned = dm.assemble_pressure(A, d_rhs, d_gamma=capi.DeviceArray.from_host(w.gamma))
    assert pinned and np.array_equal(A.values(), Ao.values)
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-12,
                    max_iterations=5000, true_residual=1, max_outer=30)
    s.setup(A)
    db, dx = capi.DeviceArray.from_host(b), capi.DeviceArray(m.n_cells)
    st, rc = s.solve_device(db, dx)
    assert rc == capi.MI_OK
    x = dx.download()
    dg = capi.DeviceArray.from_host(2.0 * wts * x)
    dlam = capi.DeviceArray(m.n_cells)
    # the adjoint's FP64 true-residual floor is ~5e-11 here (a sparse direct
    # solve reaches 1.4e-11): lambda carries the pinned row's 1e6 scale, and
    # 2 w x is not zero on the pinned cell as b is
    sa = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-10,
                     max_iterations=5000, true_residual=1, max_outer=30)
    sa.setup(A)
    st, rc = sa.adjoint_solve_device(dg, dlam)
    assert rc == capi.MI_OK, st.true_relative_residual
    lam = dlam.download()
    dgrad = capi.DeviceArray(m.n_faces)
    dm.face_gradient(dx, dlam, dgrad, pinned)
    grad = dgrad.download()
    xo, _, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-13, 50000)
    At = oracle.Matrix(S.T.tocsr().indptr.astype(np.int32), S.T.tocsr().indices.astype(np.int32),
                       S.T.tocsr().data.copy())
    lo, _, _, _ = oracle.cg(At, 2.0 * wts * xo, oracle.PC_JACOBI, 1e-13, 50000)
    go = oracle.face_gradient(m, w.bc_of_zone, xo, lo, pinned)
    ex, el = rel(x, xo), rel(lam, lo)
    eg = np.abs(grad - go).max() / np.abs(go).max()
    print(f"config 5 (24^3): rel err x {ex:.2e}, lambda {el:.2e}, max gradient err / max "
          f"|gradient| {eg:.2e}")
    assert ex <= MATCH and el <= MATCH and eg <= MATCH


def test_headline_mixed_solve_full_size(gpu):
    """Config 2 at its full 128^3 size (2,097,152 rows, 14,581,760 nonzeros),
    assembled on the GPU: the mixed SA-AMG-PCG solve meets the FP64 true
    residual 1e-8 (checked on the host with the oracle's SpMV), and lies within
    the distance the conditioning allows of a tight (1e-12) FP64 GPU solve."""
    w = wl.poisson3d(128)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    d_rhs = capi.DeviceArray(m.n_cells)
    pinned = dm.assemble_pressure(A, d_rhs, d_gamma=capi.DeviceArray.from_host(w.gamma))
    assert pinned and A.n == 2097152 and A.nnz == 14581760
    b = w.b.copy()
    b[0] = 0.0
    rp, ci = A.pattern()
    Ao = oracle.Matrix(rp, ci, A.values())
    out = {}
    for name, prec, tol in (("mixed", capi.MI_PRECISION_MIXED, 1e-8),
                            ("fp64", capi.MI_PRECISION_FP64, 1e-8),
                            ("tight", capi.MI_PRECISION_FP64, 1e-12)):
        s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=tol,
                        max_iterations=2000, precision=prec, true_residual=1, max_outer=30)
        s.setup(A)
        x, st, rc = s.solve(b)
        r = np.linalg.norm(b - oracle.spmv(Ao, x)) / np.linalg.norm(b)
        print(f"128^3 {name}: {st.iterations} its in {st.outer_iterations} sweeps, host-checked "
              f"true rel res {r:.2e}", flush=True)
        assert rc == capi.MI_OK and r <= tol
        out[name] = x
    em, ef = rel(out["mixed"], out["tight"]), rel(out["fp64"], out["tight"])
    print(f"128^3: |x_mixed - x_tight|/|x_tight| = {em:.2e}, |x_fp64 - x_tight| = {ef:.2e}")
    # a residual of 1e-8 bounds the error by cond(A) * 1e-8 in the worst case;
    # measured errors sit near 1e-7 on this 10^4-contrast, pinned system
    assert em <= 2e-6 and ef <= 2e-6
