"""Batched solves with ILU(0) and with one matrix shared by every equation.

A batch created with MI_BATCH_ILU0 orders its rows by two colours (the
ordering of mi_color_permutation) and factorizes on the device.  The factors
are pinned by one preconditioner apply against the oracle's ILU(0)
(ILU0Preconditioner::setup / apply, LinearSolvers.cpp:81-151, restated in
oracle.c) run on the same colour-permuted matrix: the device works in FP32,
so the bar is 2e-6 of |z| (a dropped pivot update, a wrong sign or a
transposed coupling moves z by 1e-3 or more on these systems).  Solutions
are checked against the oracle's tight FP64 solves (1e-6) and by their FP64
true residual (<= 1e-8); iteration counts against the oracle's BiCGStab / CG
with ILU(0) on the permuted matrix (FP32 inner iterations with FP64
refinement: within 3).  A batch with shared values must give bit for bit
what a batch holding one copy of the matrix per equation gives.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from simall_amd import capi
from simall_amd import workload as wl

pytestmark = pytest.mark.gpu

ILU = capi.MI_BATCH_ILU0
SHARED = capi.MI_BATCH_SHARED_VALUES


def _convective(N, n_eq, seed=3):
    """Nonsymmetric convection-diffusion systems on one pattern (Taylor-Green
    mass flux, pseudo-transient diagonal, a diffusivity per equation)."""
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
    return dm, A, rp, ci, vals, bs


def _symmetric(N, n_eq):
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
    return dm, A, rp, ci, vals, np.stack(w.bs)


def _permuted(rp, ci, vals, perm):
    n = rp.size - 1
    M = sp.csr_matrix((vals, ci, rp), shape=(n, n))
    Mp = M[perm][:, perm].tocsr()
    Mp.sort_indices()
    return oracle.Matrix(Mp.indptr.astype(np.int32), Mp.indices.astype(np.int32),
                         Mp.data.astype(np.float64))


@pytest.fixture(scope="module")
def conv4(gpu):
    return _convective(16, 4)


def test_two_colour_permutation_of_a_hex_pattern(conv4):
    dm, A, rp, ci, vals, bs = conv4
    perm, nc = capi.color_permutation(rp, ci)
    assert nc == 2
    assert np.array_equal(np.sort(perm), np.arange(A.n))
    P = _permuted(rp, ci, vals[0], perm)
    n0 = A.n // 2
    # no coupling inside a colour: every off-diagonal of a first-colour row lies in the second
    for i in range(0, A.n, 97):
        cols = P.colIdx[P.rowPtr[i]:P.rowPtr[i + 1]]
        off = cols[cols != i]
        assert np.all((off >= n0) if i < n0 else (off < n0)), i


def test_batch_ilu_apply_matches_oracle_factorization(conv4):
    dm, A, rp, ci, vals, bs = conv4
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    r = np.stack([wl.normal(40 + e, A.n) for e in range(len(vals))])
    z = B.apply_preconditioner(r, capi.MI_PC_ILU)
    perm, _ = capi.color_permutation(rp, ci)
    for e in range(len(vals)):
        P = _permuted(rp, ci, vals[e], perm)
        lu, dp = oracle.ilu0(P)
        zp = oracle.ilu0_apply(P, lu, dp, r[e][perm])
        zo = np.empty_like(zp)
        zo[perm] = zp
        err = np.linalg.norm(z[e] - zo) / np.linalg.norm(zo)
        print(f"eq {e}: |z - z_oracle| / |z_oracle| = {err:.2e}")
        assert err < 2e-6
        # and the factorization is not the Jacobi scaling
        assert np.linalg.norm(r[e] / sp.csr_matrix((vals[e], ci, rp)).diagonal() - zo) > 1e-2 * np.linalg.norm(zo)
    zj = B.apply_preconditioner(r, capi.MI_PC_JACOBI)
    for e in range(len(vals)):
        d = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2).diagonal()
        assert np.allclose(zj[e], r[e] / d, rtol=2e-7, atol=0)


def test_batch_ilu_apply_with_shared_values(conv4):
    dm, A, rp, ci, vals, bs = conv4
    Bs = capi.Batch(A, 3, flags=ILU | SHARED)
    Bs.set_matrix(A)  # the matrix last assembled: vals[-1]
    r = np.stack([wl.normal(50 + e, A.n) for e in range(3)])
    z = Bs.apply_preconditioner(r, capi.MI_PC_ILU)
    perm, _ = capi.color_permutation(rp, ci)
    P = _permuted(rp, ci, vals[-1], perm)
    lu, dp = oracle.ilu0(P)
    for e in range(3):
        zp = oracle.ilu0_apply(P, lu, dp, r[e][perm])
        zo = np.empty_like(zp)
        zo[perm] = zp
        assert np.linalg.norm(z[e] - zo) < 2e-6 * np.linalg.norm(zo)


def test_batch_bicgstab_ilu_matches_oracle(conv4):
    dm, A, rp, ci, vals, bs = conv4
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    b = bs.copy()
    b[2] = 0.0
    x, its, rel, rc = B.solve(b, kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU,
                              tolerance=1e-8, max_iterations=2000)
    xj, itsj, _, rcj = B.solve(b, kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_JACOBI,
                               tolerance=1e-8, max_iterations=2000)
    print(f"batched BiCGStab-ILU iterations {its.tolist()}, Jacobi {itsj.tolist()}")
    assert rc == capi.MI_OK and rcj == capi.MI_OK
    assert its[2] == 0 and np.all(x[2] == 0.0)
    perm, _ = capi.color_permutation(rp, ci)
    for e in range(len(vals)):
        if e == 2:
            continue
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])
        xo, _, _, _ = oracle.bicgstab(oracle.Matrix(rp, ci, vals[e]), b[e], oracle.PC_ILU, 1e-12,
                                      5000)
        assert np.linalg.norm(x[e] - xo) <= 1e-6 * np.linalg.norm(xo)
        _, ito, _, _ = oracle.bicgstab(_permuted(rp, ci, vals[e], perm), b[e][perm],
                                       oracle.PC_ILU, 1e-8, 5000)
        assert abs(int(its[e]) - ito) <= 3, (e, its[e], ito)
        assert its[e] < itsj[e]


def test_batch_cg_ilu_symmetric_systems(gpu):
    dm, A, rp, ci, vals, bs = _symmetric(16, 6)
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    x, its, rel, rc = B.solve(bs, kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_ILU,
                              tolerance=1e-8, max_iterations=3000)
    xj, itsj, _, _ = B.solve(bs, kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI,
                             tolerance=1e-8, max_iterations=3000)
    print(f"batched CG-ILU iterations {its.tolist()}, Jacobi {itsj.tolist()}")
    assert rc == capi.MI_OK
    perm, _ = capi.color_permutation(rp, ci)
    for e in range(len(vals)):
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(bs[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(bs[e])
        assert np.linalg.norm(x[e] - xj[e]) <= 1e-6 * np.linalg.norm(xj[e])
        _, ito, _, _ = oracle.cg(_permuted(rp, ci, vals[e], perm), bs[e][perm], oracle.PC_ILU,
                                 1e-8, 5000)
        assert abs(int(its[e]) - ito) <= 3, (e, its[e], ito)
        assert its[e] <= itsj[e]


@pytest.mark.parametrize("kind,pc", [(capi.MI_SOLVER_CG, capi.MI_PC_JACOBI),
                                     (capi.MI_SOLVER_BICGSTAB, capi.MI_PC_JACOBI),
                                     (capi.MI_SOLVER_BICGSTAB, capi.MI_PC_ILU),
                                     (capi.MI_SOLVER_BICGSTAB, capi.MI_PC_NONE)])
def test_shared_values_equal_per_equation_copies(conv4, kind, pc):
    # CG on a symmetric system, BiCGStab on the convective one
    dm, A, rp, ci, vals, bs = _symmetric(12, 3) if kind == capi.MI_SOLVER_CG else conv4
    n_eq = 3
    Bs = capi.Batch(A, n_eq, flags=ILU | SHARED)
    Bp = capi.Batch(A, n_eq, flags=ILU)
    Bs.set_matrix(A)
    Bp.set_matrix(A)
    # the timed SpMM shapes may differ; the shape fixes the order of the dot
    # products' partial sums, so both batches take the first one's
    ts, tp = Bs.tuning(), Bp.tuning()
    print(f"tuned shapes: shared {ts}, per-equation {tp}")
    for B in (Bs, Bp):
        B.set_shape(ts["slots_per_thread_max"], ts["spmm_workgroups"], ts["banded"])
    kw = dict(kind=kind, preconditioner=pc, tolerance=1e-8, max_iterations=4000)
    xs, itss, rels, rcs = Bs.solve(bs[:n_eq], **kw)
    xp, itsp, relp, rcp = Bp.solve(bs[:n_eq], **kw)
    print(f"kind {kind} pc {pc}: shared {itss.tolist()} per-equation {itsp.tolist()}")
    assert rcs == rcp == capi.MI_OK
    assert np.array_equal(itss, itsp)
    assert np.array_equal(xs, xp)
    assert np.array_equal(rels, relp)


def test_batch_ilu_graph_replays_match_eager(gpu):
    dm, A, rp, ci, vals, b = _convective(12, 4, seed=5)
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    kw = dict(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU, tolerance=1e-8,
              check_interval=3)
    xe, itse, _, _ = B.solve(b, use_graph=0, **kw)
    xg, itsg, _, _ = B.solve(b, use_graph=1, **kw)
    xg2, itsg2, _, _ = B.solve(b, use_graph=1, **kw)
    assert np.array_equal(itse, itsg) and np.array_equal(itsg, itsg2)
    assert np.array_equal(xe, xg) and np.array_equal(xg, xg2)
    kc = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_ILU, tolerance=1e-6,
              max_iterations=3000, check_interval=3)
    ce, _, _, _ = B.solve(b, use_graph=0, **kc)
    cg, _, _, _ = B.solve(b, use_graph=1, **kc)
    assert np.array_equal(ce, cg)


def test_batch_ilu_new_values_refactorize(conv4):
    """Values set after a solve are factorized again before the next ILU(0) solve."""
    dm, A, rp, ci, vals, bs = conv4
    B = capi.Batch(A, 2, flags=ILU)
    B.set_values(np.stack(vals[:2]))
    r = np.stack([wl.normal(60 + e, A.n) for e in range(2)])
    z0 = B.apply_preconditioner(r, capi.MI_PC_ILU)
    B.set_values(np.stack(vals[2:4]))
    z1 = B.apply_preconditioner(r, capi.MI_PC_ILU)
    B2 = capi.Batch(A, 2, flags=ILU)
    B2.set_values(np.stack(vals[2:4]))
    assert np.array_equal(z1, B2.apply_preconditioner(r, capi.MI_PC_ILU))
    assert not np.array_equal(z0, z1)


def test_batch_ilu_large_shared_reordered(gpu):
    """48^3 = 110592 rows, three right-hand sides on one momentum matrix."""
    dm, A, rp, ci, vals, b = _convective(48, 1, seed=9)
    bs = np.stack([wl.normal(70 + e, A.n) for e in range(3)])
    B = capi.Batch(A, 3, flags=ILU | SHARED)
    B.set_matrix(A)
    x, its, rel, rc = B.solve(bs, kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU,
                              tolerance=1e-8, max_iterations=3000)
    print(f"48^3 shared BiCGStab-ILU iterations {its.tolist()}")
    assert rc == capi.MI_OK
    M = sp.csr_matrix((vals[0], ci, rp), shape=(A.n,) * 2)
    for e in range(3):
        assert np.linalg.norm(bs[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(bs[e])


def test_batch_ilu_refusals(conv4):
    dm, A, rp, ci, vals, bs = conv4
    lib = capi.load()
    B = capi.Batch(A, 2)
    B.set_values(np.stack(vals[:2]))
    with pytest.raises(capi.MiError) as ei:
        B.solve(bs[:2], kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU)
    assert ei.value.code == capi.MI_ERR_UNSUPPORTED
    Bs = capi.Batch(A, 2, flags=SHARED)
    assert lib.mi_batch_set_equation(Bs.h, 0, A.h) == capi.MI_ERR_BAD_ARG
    v = np.stack(vals[:2])
    assert lib.mi_batch_set_values(Bs.h, capi.ptr(v)) == capi.MI_ERR_BAD_ARG
    # a pattern with a triangle of couplings needs three colours
    rp3 = np.array([0, 3, 6, 9], np.int32)
    ci3 = np.array([0, 1, 2, 0, 1, 2, 0, 1, 2], np.int32)
    T = capi.Csr(rp3, ci3, np.array([4.0, -1, -1, -1, 4, -1, -1, -1, 4]))
    h = capi.C.c_void_p()
    assert lib.mi_batch_create_ex(T.h, 1, ILU, capi.C.byref(h)) == capi.MI_ERR_UNSUPPORTED
    assert lib.mi_batch_create_ex(T.h, 1, 4, capi.C.byref(h)) == capi.MI_ERR_BAD_ARG


@pytest.mark.parametrize("pc", [capi.MI_PC_ILU, capi.MI_PC_JACOBI])
def test_batch_gmres_matches_oracle(conv4, pc):
    """Batched GMRES(20) (one cycle per FP64 refinement sweep, classical
    Gram-Schmidt twice): every FP64 true residual <= 1e-8, solutions within
    1e-6 of the oracle's GMRES (GMRESmSolver::solve) solved tightly, a zero
    right-hand side returns 0 in 0 iterations, and the single-equation mixed
    GMRES with the same preconditioner needs about as many Arnoldi steps."""
    dm, A, rp, ci, vals, bs = conv4
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    b = bs.copy()
    b[1] = 0.0
    kw = dict(kind=capi.MI_SOLVER_GMRES, preconditioner=pc, restart=20, tolerance=1e-8,
              max_iterations=2000)
    x, its, rel, rc = B.solve(b, **kw)
    print(f"batched GMRES(20) pc {pc}: steps {its.tolist()} rel {rel.max():.2e}")
    assert rc == capi.MI_OK and np.all(rel <= 1e-8)
    assert its[1] == 0 and np.all(x[1] == 0.0)
    s = capi.Solver(kind=capi.MI_SOLVER_GMRES, preconditioner=capi.MI_PC_JACOBI, restart=20,
                    tolerance=1e-8, max_iterations=2000, precision=capi.MI_PRECISION_MIXED,
                    true_residual=1)
    for e in range(len(vals)):
        if e == 1:
            continue
        M = sp.csr_matrix((vals[e], ci, rp), shape=(A.n,) * 2)
        assert np.linalg.norm(b[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(b[e])
        xo, _, _ = oracle.gmres(oracle.Matrix(rp, ci, vals[e]), b[e], oracle.PC_JACOBI, 1e-13,
                                5000, 30)
        assert np.linalg.norm(x[e] - xo) <= 1e-6 * np.linalg.norm(xo), e
        if pc == capi.MI_PC_JACOBI:
            s.setup(capi.Csr(rp, ci, vals[e]))
            _, st, _ = s.solve(b[e])
            print(f"  eq {e}: batched {its[e]} steps, single-equation mixed {st.iterations}")
            assert abs(int(its[e]) - st.iterations) <= max(3, st.iterations // 5)


def test_batch_gmres_graph_replays_and_limits(gpu):
    dm, A, rp, ci, vals, b = _convective(12, 3, seed=11)
    B = capi.Batch(A, len(vals), flags=ILU)
    B.set_values(np.stack(vals))
    kw = dict(kind=capi.MI_SOLVER_GMRES, preconditioner=capi.MI_PC_ILU, restart=7,
              tolerance=1e-8, max_iterations=500)
    xe, itse, _, _ = B.solve(b, use_graph=0, **kw)
    xg, itsg, _, _ = B.solve(b, use_graph=1, **kw)
    xg2, itsg2, _, _ = B.solve(b, use_graph=1, **kw)
    assert np.array_equal(itse, itsg) and np.array_equal(itsg, itsg2)
    assert np.array_equal(xe, xg) and np.array_equal(xg, xg2)
    # restart 1 still converges; an iteration limit of 3 reports not converged
    x1, its1, rel1, rc1 = B.solve(b, **dict(kw, restart=1, max_iterations=3000))
    assert rc1 == capi.MI_OK and np.all(rel1 <= 1e-8)
    x3, its3, rel3, rc3 = B.solve(b, **dict(kw, max_iterations=3))
    assert rc3 == capi.MI_ERR_NOT_CONVERGED and np.all(its3 <= 3)
    lib = capi.load()
    c = capi.default_cfg(kind=capi.MI_SOLVER_GMRES, restart=40)
    db = capi.DeviceArray.from_host(b.ravel())
    dx = capi.DeviceArray(b.size)
    its = np.zeros(len(vals), np.int32)
    rel = np.zeros(len(vals))
    assert lib.mi_batch_solve_device(B.h, capi.C.byref(c), db.ptr, dx.ptr, capi.ptr(its),
                                     capi.ptr(rel)) == capi.MI_ERR_UNSUPPORTED


def test_batch_gmres_shared_matrix_large(gpu):
    """48^3, three right-hand sides on one momentum matrix: GMRES(30)-ILU(0) as
    the flow's momentum configuration, batched on the shared matrix."""
    dm, A, rp, ci, vals, b = _convective(48, 1, seed=9)
    bs = np.stack([wl.normal(80 + e, A.n) for e in range(3)])
    B = capi.Batch(A, 3, flags=ILU | SHARED)
    B.set_matrix(A)
    x, its, rel, rc = B.solve(bs, kind=capi.MI_SOLVER_GMRES, preconditioner=capi.MI_PC_ILU,
                              restart=30, tolerance=1e-8, max_iterations=3000)
    print(f"48^3 shared GMRES(30)-ILU steps {its.tolist()}")
    assert rc == capi.MI_OK
    M = sp.csr_matrix((vals[0], ci, rp), shape=(A.n,) * 2)
    for e in range(3):
        assert np.linalg.norm(bs[e] - M @ x[e]) <= 1.01e-8 * np.linalg.norm(bs[e])
