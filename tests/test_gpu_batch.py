"""Batched multi-RHS solve of transport equations sharing one pattern.

Systems: A_e = V/dt I + Gamma_e L on a perturbed, permuted hex mesh with
symmetry boundaries, assembled on the GPU by the momentum assembly with zero
flux (transient + diffusion).  Each batched solution is checked against the
oracle's FP64 Jacobi-PCG solve of the same equation (relative error within
1e-6; the FP64 true residual of every equation <= 1e-8 by construction and
re-checked on the host with scipy).  Graph replays are checked bit-exact
against eager launches and against a second replayed solve.  The batch is
compared with the single-equation solver on the same arithmetic (FP32 Jacobi
inner iterations, FP64 refinement): same per-equation convergence.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from simall_amd import capi
from simall_amd import workload as wl

pytestmark = pytest.mark.gpu


def _systems(N, n_eq):
    w = wl.transport(N, n_eq)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    phi = capi.DeviceArray.from_host(np.zeros(m.n_faces))
    dt = capi.DeviceArray.from_host(np.full(m.n_cells, w.dt))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    rp, ci = A.pattern()
    vals = []
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(w.gammas[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        vals.append(A.values())
    return w, dm, A, rp, ci, vals


@pytest.fixture(scope="module")
def sys8(gpu):
    return _systems(16, 8)


def test_batch_matches_oracle_per_equation(sys8):
    w, dm, A, rp, ci, vals = sys8
    n_eq = len(vals)
    B = capi.Batch(A, n_eq)
    B.set_values(np.stack(vals))
    b = np.stack(w.bs)
    x, its, rel, rc = B.solve(b, tolerance=1e-8, max_iterations=2000)
    assert rc == capi.MI_OK
    assert np.all(rel <= 1e-8)
    assert len(set(its.tolist())) > 1  # uneven iteration counts across equations
    for e in range(n_eq):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        r = b[e] - M @ x[e]
        assert np.linalg.norm(r) <= 1.01e-8 * np.linalg.norm(b[e])
        xo, _, _, _ = oracle.cg(oracle.Matrix(rp, ci, vals[e]), b[e], oracle.PC_JACOBI, 1e-12, 5000)
        assert np.linalg.norm(x[e] - xo) <= 1e-6 * np.linalg.norm(xo)
    # the operator is V/dt I + Gamma_e L: the mass term makes the systems SPD
    # and unequal in conditioning
    assert its.min() < its.max()


def test_batch_odd_count_and_zero_rhs(sys8):
    w, dm, A, rp, ci, vals = sys8
    B = capi.Batch(A, 5)
    B.set_values(np.stack(vals[:5]))
    b = np.stack(w.bs[:5])
    b[2] = 0.0
    x, its, rel, rc = B.solve(b, tolerance=1e-8)
    assert rc == capi.MI_OK
    assert its[2] == 0 and rel[2] == 0.0 and np.all(x[2] == 0.0)
    assert np.all(its[[0, 1, 3, 4]] > 0)
    for e in (0, 1, 3, 4):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])


def test_batch_graph_replays_match_eager_launches(sys8):
    w, dm, A, rp, ci, vals = sys8
    B = capi.Batch(A, len(vals))
    B.set_values(np.stack(vals))
    b = np.stack(w.bs)
    xe, itse, rele, _ = B.solve(b, tolerance=1e-8, use_graph=0, check_interval=4)
    xg, itsg, relg, _ = B.solve(b, tolerance=1e-8, use_graph=1, check_interval=4)
    xg2, itsg2, _, _ = B.solve(b, tolerance=1e-8, use_graph=1, check_interval=4)
    assert np.array_equal(itse, itsg) and np.array_equal(itsg, itsg2)
    assert np.array_equal(xe, xg) and np.array_equal(xg, xg2)
    assert np.array_equal(rele, relg)


def test_batch_iteration_limit_reports_not_converged(sys8):
    w, dm, A, rp, ci, vals = sys8
    B = capi.Batch(A, 4)
    B.set_values(np.stack(vals[-4:]))
    x, its, rel, rc = B.solve(np.stack(w.bs[-4:]), tolerance=1e-12, max_iterations=3, max_outer=1)
    assert rc == capi.MI_ERR_NOT_CONVERGED
    assert np.all(its <= 3)


def test_batch_tuning_is_reported(sys8):
    w, dm, A, rp, ci, vals = sys8
    B = capi.Batch(A, 32)
    t = B.tuning()
    print(f"tuning: {t}")
    assert t["rows_per_block"] > 0 and t["spmm_us"] > 0 and t["spmm_bytes"] > 0
    assert t["slots_per_thread_max"] in (1, 2, 4)
    assert t["spmm_workgroups"] in (256, 512, 1024, 2048, 4096)
    assert 8 <= t["vector_workgroups"] <= 4096 and t["vector_us"] > 0
    # whatever shapes were picked, an odd width (one slot per thread) solves too
    B3 = capi.Batch(A, 3)
    B3.set_values(np.stack(vals[:3]))
    x, its, rel, rc = B3.solve(np.stack(w.bs[:3]), tolerance=1e-8, max_iterations=2000)
    assert rc == capi.MI_OK and np.all(rel <= 1e-8)


def test_batch_set_shape_pins_the_bits(sys8):
    """Two batches given the same SpMM shape return the same bits; another
    shape the same solution to the tolerance; bad shapes are refused."""
    w, dm, A, rp, ci, vals = sys8
    b = np.stack(w.bs[:4])
    out = []
    for shape in ((4, 512, 1), (1, 256, 0)):
        xs = []
        for _ in range(2):
            B = capi.Batch(A, 4)
            B.set_values(np.stack(vals[:4]))
            B.set_shape(*shape)
            t = B.tuning()
            assert (t["slots_per_thread_max"], t["spmm_workgroups"], t["banded"]) == shape
            x, its, rel, rc = B.solve(b, tolerance=1e-8, max_iterations=2000)
            assert rc == capi.MI_OK
            xs.append(x)
        assert np.array_equal(xs[0], xs[1])
        out.append(xs[0])
    assert np.linalg.norm(out[0] - out[1]) <= 1e-6 * np.linalg.norm(out[1])
    B = capi.Batch(A, 4)
    for bad in ((3, 512, 0), (4, 100, 0), (4, 8192, 0), (4, 512, 2)):
        with pytest.raises(capi.MiError):
            B.set_shape(*bad)


def test_grid_barrier_timing(gpu):
    """mi_time_grid_barrier reports a cooperative grid.sync() and a kernel
    boundary; a grid larger than can be resident is refused."""
    sync_us, launch_us = capi.time_grid_barrier(256, 50)
    print(f"256 workgroups: grid.sync {sync_us:.2f} us, kernel boundary {launch_us:.2f} us")
    assert sync_us > 0 and launch_us > 0
    lib = capi.load()
    a, b = capi.C.c_double(), capi.C.c_double()
    assert lib.mi_time_grid_barrier(1 << 24, 10, capi.C.byref(a), capi.C.byref(b)) == capi.MI_ERR_BAD_ARG
    assert lib.mi_time_grid_barrier(0, 10, capi.C.byref(a), capi.C.byref(b)) == capi.MI_ERR_BAD_ARG


def test_batch_matches_single_equation_solver(sys8):
    w, dm, A, rp, ci, vals = sys8
    n_eq = len(vals)
    B = capi.Batch(A, n_eq)
    B.set_values(np.stack(vals))
    b = np.stack(w.bs)
    x, its, rel, rc = B.solve(b, tolerance=1e-8)
    assert rc == capi.MI_OK
    for e in range(n_eq):
        s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                        precision=capi.MI_PRECISION_MIXED, true_residual=1)
        s.setup(capi.Csr(rp, ci, vals[e]))
        xs, st, rc1 = s.solve(b[e])
        assert rc1 == capi.MI_OK
        assert np.linalg.norm(x[e] - xs) <= 1e-7 * np.linalg.norm(xs)
        assert abs(int(its[e]) - st.iterations) <= max(2, st.iterations // 5)


def test_batch_reordered_path(gpu):
    """n >= 100,000: the batch reorders the shared pattern (RCM) like the
    single-equation solver; results come back in the caller's numbering."""
    w, dm, A, rp, ci, vals = _systems(48, 4)
    B = capi.Batch(A, 4)
    for e in range(4):
        dm.assemble_momentum(A, capi.DeviceArray(A.n), 0, float(w.gammas[e]), 1.0,
                             capi.DeviceArray.from_host(np.zeros(w.mesh.n_faces)),
                             d_dt=capi.DeviceArray.from_host(np.full(A.n, w.dt)),
                             d_Uold=capi.DeviceArray.from_host(np.zeros(A.n)))
        B.set_equation(e, A)
    b = np.stack(w.bs)
    x, its, rel, rc = B.solve(b, tolerance=1e-8)
    assert rc == capi.MI_OK and np.all(rel <= 1e-8)
    for e in range(4):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])


def _convective_systems(N, n_eq, seed=3):
    """Convection-diffusion transport: the momentum assembly with the
    Taylor-Green mass flux, a pseudo-transient diagonal and a diffusivity per
    equation: nonsymmetric systems on one pattern."""
    L = 2 * np.pi
    m = wl.Mesh(N, N, N, extent=(L, L, L), perturb=0.1, node_seed=seed)
    bc = [wl.BC_WALL] * 7
    dm = capi.DeviceMesh(m, bc)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    U = wl.taylor_green(m)
    phi = capi.DeviceArray.from_host(oracle.face_flux(m, bc, 1.0, U))
    dt = capi.DeviceArray.from_host(wl.loguniform(seed, -1, 0, m.n_cells))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    rp, ci = A.pattern()
    gam = wl.loguniform(seed + 1, -3, -1, n_eq)
    vals = []
    for e in range(n_eq):
        dm.assemble_momentum(A, rhs, 0, float(gam[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        vals.append(A.values())
    bs = np.stack([wl.normal(seed + 10 + e, m.n_cells) for e in range(n_eq)])
    return A, rp, ci, vals, bs


def test_batch_bicgstab_nonsymmetric_matches_oracle(gpu):
    A, rp, ci, vals, b = _convective_systems(16, 6)
    M0 = sp.csr_matrix((vals[0], ci, rp), shape=(A.n,) * 2)
    assert abs(M0 - M0.T).max() > 1e-3  # nonsymmetric
    B = capi.Batch(A, len(vals))
    B.set_values(np.stack(vals))
    b[3] = 0.0
    x, its, rel, rc = B.solve(b, kind=capi.MI_SOLVER_BICGSTAB, tolerance=1e-8, max_iterations=2000)
    print(f"batched BiCGStab iterations {its.tolist()}")
    assert rc == capi.MI_OK and np.all(rel <= 1e-8)
    assert its[3] == 0 and np.all(x[3] == 0.0)
    for e in range(len(vals)):
        if e == 3:
            continue
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])
        xo, ito, _, _ = oracle.bicgstab(oracle.Matrix(rp, ci, vals[e]), b[e], oracle.PC_JACOBI,
                                        1e-12, 5000)
        assert np.linalg.norm(x[e] - xo) <= 1e-6 * np.linalg.norm(xo)


def test_batch_bicgstab_graph_replays_match_eager(gpu):
    A, rp, ci, vals, b = _convective_systems(12, 4, seed=5)
    B = capi.Batch(A, len(vals))
    B.set_values(np.stack(vals))
    kw = dict(kind=capi.MI_SOLVER_BICGSTAB, tolerance=1e-8, check_interval=3)
    xe, itse, _, _ = B.solve(b, use_graph=0, **kw)
    xg, itsg, _, _ = B.solve(b, use_graph=1, **kw)
    xg2, itsg2, _, _ = B.solve(b, use_graph=1, **kw)
    assert np.array_equal(itse, itsg) and np.array_equal(itsg, itsg2)
    assert np.array_equal(xe, xg) and np.array_equal(xg, xg2)
    # the same batch then solved with CG keeps working (graph rebuilt per kind)
    x, its, rel, rc = B.solve(b, kind=capi.MI_SOLVER_CG, tolerance=1e-6, max_iterations=3000)
    assert rc in (capi.MI_OK, capi.MI_ERR_NOT_CONVERGED)


def test_batch_bicgstab_reordered_path(gpu):
    """48^3 = 110592 rows: the batch reorders its pattern (RCM)."""
    A, rp, ci, vals, b = _convective_systems(48, 3, seed=7)
    B = capi.Batch(A, len(vals))
    B.set_values(np.stack(vals))
    x, its, rel, rc = B.solve(b, kind=capi.MI_SOLVER_BICGSTAB, tolerance=1e-8, max_iterations=3000)
    print(f"48^3 batched BiCGStab iterations {its.tolist()}")
    assert rc == capi.MI_OK
    for e in range(len(vals)):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])


def test_batch_config4_full_size(gpu):
    """The 32-equation transport batch at its full size (100^3, 1M rows, 6.94M
    nonzeros): every equation's FP64 true residual <= 1e-8, re-checked on the
    host with scipy, uneven iteration counts, and three equations against the
    single-equation solver (FP64 Jacobi-PCG solved tightly) to 1e-6."""
    w = wl.transport(100, 32)
    m = w.mesh
    dm = capi.DeviceMesh(m, w.bc_of_zone)
    A = dm.create_matrix()
    rhs = capi.DeviceArray(m.n_cells)
    phi = capi.DeviceArray.from_host(np.zeros(m.n_faces))
    dt = capi.DeviceArray.from_host(np.full(m.n_cells, w.dt))
    uold = capi.DeviceArray.from_host(np.zeros(m.n_cells))
    rp, ci = A.pattern()
    B = capi.Batch(A, 32)
    vals = []
    for e in range(32):
        dm.assemble_momentum(A, rhs, 0, float(w.gammas[e]), 1.0, phi, d_dt=dt, d_Uold=uold)
        B.set_equation(e, A)
        vals.append(A.values())
    b = np.stack(w.bs)
    x, its, rel, rc = B.solve(b, tolerance=1e-8, max_iterations=2000, inner_tolerance=1e-5)
    print(f"config 4 full size: iterations {its.min()}..{its.max()}, max rel {rel.max():.2e}")
    assert rc == capi.MI_OK and np.all(rel <= 1e-8)
    assert its.max() > 2 * its.min()
    for e in range(32):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e]), e
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-11,
                    max_iterations=20000, true_residual=1)
    for e in (0, 15, 31):
        s.setup(capi.Csr(rp, ci, vals[e]))
        xe, st, _ = s.solve(b[e])
        assert st.converged
        err = np.linalg.norm(x[e] - xe) / np.linalg.norm(xe)
        print(f"equation {e}: {its[e]} iterations, |x - x_tight| / |x_tight| = {err:.2e}")
        assert err <= 1e-6, e
