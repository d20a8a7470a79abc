"""GPU ILU(0) and BiCGStab against the oracle.

ILU(0): the host factorization and the level-scheduled device sweeps perform
the reference's operations in its order (contraction off), so z = M^-1 r is
compared for exact equality.  BiCGStab: iteration counts within +-2 of the
oracle and the FP64 true residual at or below the tolerance.
"""
import math

import numpy as np
import pytest

import oracle
from simall_amd import capi
from simall_amd import workload as wl

pytestmark = pytest.mark.gpu


def tridiag_csr(n, lo, d, up):
    rp, ci, va = [0], [], []
    for i in range(n):
        if i > 0:
            ci.append(i - 1)
            va.append(lo)
        ci.append(i)
        va.append(d)
        if i + 1 < n:
            ci.append(i + 1)
            va.append(up)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(va)


@pytest.fixture(scope="module")
def momentum():
    """Taylor-Green upwind convection-diffusion, 16^3, pseudo-transient diagonal."""
    N = 16
    L = 2 * math.pi
    m = wl.Mesh(N, N, N, extent=(L, L, L))
    U = wl.taylor_green(m)
    bc = [wl.BC_WALL] * 7
    phi = oracle.face_flux(m, bc, 1.0, U)
    rp, ci = oracle.build_sparsity(m)
    pat = oracle.Matrix(rp, ci, np.zeros(ci.size))
    dt = wl.loguniform(6, -1, 1, m.n_cells)
    vals, _ = oracle.assemble_momentum(m, pat, bc, 0, (L / N) / 50, 1.0, phi, dt=dt)
    b = wl.normal(7, m.n_cells)
    return oracle.Matrix(rp, ci, vals), b


def solver(A, **kw):
    cfg = dict(kind=capi.MI_SOLVER_BICGSTAB, preconditioner=capi.MI_PC_ILU, tolerance=1e-8,
               max_iterations=1000, reorder=0)
    cfg.update(kw)
    s = capi.Solver(**cfg)
    s.setup(capi.Csr(A.rowPtr, A.colIdx, A.values))
    return s


def test_bicgstab_ilu0_known_answer(gpu):
    # tests/unit/gpu/test_gpu_models.cpp:238-287
    rp, ci, va = tridiag_csr(8, -0.5, 3.0, 1.0)
    A = oracle.Matrix(rp, ci, va)
    s = solver(A, tolerance=1e-10, max_iterations=200)
    x, st, rc = s.solve(np.ones(8))
    assert rc == capi.MI_OK
    assert np.linalg.norm(np.ones(8) - A.dense() @ x) < 1e-7
    xo, ito, _, _ = oracle.bicgstab(A, np.ones(8), oracle.PC_ILU, 1e-10, 200)
    assert st.iterations == ito == 1


def test_ilu0_apply_bit_exact(gpu, momentum):
    A, b = momentum
    s = solver(A)
    r = np.random.default_rng(0).standard_normal(A.n)
    z = s.precond_apply(r)
    lu, dp = oracle.ilu0(A)
    assert np.array_equal(z, oracle.ilu0_apply(A, lu, dp, r))


def test_ilu0_facade_bit_exact(gpu, momentum):
    A, b = momentum
    drp, dci = capi.DeviceArray.from_host(A.rowPtr), capi.DeviceArray.from_host(A.colIdx)
    dv = capi.DeviceArray.from_host(A.values)
    ok = capi.C.c_int()
    capi.check(gpu.mi_ilu0_factor_inplace(A.n, drp.ptr, dci.ptr, dv.ptr, capi.C.byref(ok)))
    lu, dp = oracle.ilu0(A)
    assert ok.value == 1 and np.array_equal(dv.download(), lu)
    r = np.random.default_rng(1).standard_normal(A.n)
    dr, dz = capi.DeviceArray.from_host(r), capi.DeviceArray(A.n)
    capi.check(gpu.mi_ilu0_apply(A.n, drp.ptr, dci.ptr, dv.ptr, dr.ptr, dz.ptr, None))
    assert np.array_equal(dz.download(), oracle.ilu0_apply(A, lu, dp, r))


@pytest.mark.parametrize("pc", [capi.MI_PC_ILU, capi.MI_PC_JACOBI])
def test_bicgstab_parity(gpu, momentum, pc):
    A, b = momentum
    s = solver(A, preconditioner=pc)
    x, st, rc = s.solve(b)
    xo, ito, lasto, _ = oracle.bicgstab(A, b, oracle.PC_ILU if pc == capi.MI_PC_ILU else oracle.PC_JACOBI,
                                        1e-8, 1000)
    print(f"BiCGStab pc={pc}: GPU {st.iterations} its, oracle {ito}")
    assert rc == capi.MI_OK and abs(st.iterations - ito) <= 2
    rel = np.linalg.norm(b - oracle.spmv(A, x)) / np.linalg.norm(b)
    assert rel < 1e-7  # recurrence residual < 1e-8; true residual close to it
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6


def test_bicgstab_mixed_and_true_residual(gpu, momentum):
    A, b = momentum
    for prec in (capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED):
        s = solver(A, precision=prec, true_residual=1, reorder=1)
        x, st, rc = s.solve(b)
        rel = np.linalg.norm(b - oracle.spmv(A, x)) / np.linalg.norm(b)
        print(f"BiCGStab-ILU precision {prec}: {st.iterations} its, {st.outer_iterations} sweeps, rel {rel:.2e}")
        assert rc == capi.MI_OK and rel <= 1e-8


@pytest.mark.parametrize("prec", [capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED])
def test_bicgstab_ilu_multicolour_ordering(gpu, momentum, prec):
    """reorder = 2: ILU(0) of the multicolour-ordered matrix (one sweep level
    per colour); the solution comes back in the caller's numbering."""
    A, b = momentum
    s = solver(A, precision=prec, true_residual=1, reorder=2)
    x, st, rc = s.solve(b)
    info = s.info()
    rel = np.linalg.norm(b - oracle.spmv(A, x)) / np.linalg.norm(b)
    print(f"multicolour ILU-BiCGStab prec {prec}: {st.iterations} its, levels {info['levels']}, rel {rel:.2e}")
    assert rc == capi.MI_OK and rel <= 1e-8
    assert info["levels"] <= 8  # colours of a 7-point stencil under greedy colouring


def test_ilu0_run_sweeps_match_oracle(gpu, momentum):
    """A matrix already in red-black order has two runs of independent rows,
    the first without L and the second without U entries, so the device
    applies M^-1 in two CSR-stream launches (one per run); z = M^-1 r agrees
    with the oracle's serial sweeps to rounding."""
    import scipy.sparse as sp

    A, b = momentum
    N = 16
    i = np.arange(A.n)
    parity = (i % N + (i // N) % N + i // (N * N)) % 2
    perm = np.concatenate([np.flatnonzero(parity == 0), np.flatnonzero(parity == 1)])
    S = sp.csr_matrix((A.values, A.colIdx, A.rowPtr), shape=(A.n, A.n))[perm][:, perm].tocsr()
    S.sort_indices()
    rows = np.repeat(np.arange(A.n), np.diff(S.indptr))
    same = (parity[perm][rows] == parity[perm][S.indices]) & (rows != S.indices)
    assert not same.any()  # two-coloured: no coupling inside a colour
    B = oracle.Matrix(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy())
    for prec in (capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED):
        s = solver(B, precision=prec)
        assert s.info()["levels"] == 1  # one launch per run: two per apply
        r = np.random.default_rng(3).standard_normal(B.n)
        z = s.precond_apply(r)
        lu, dp = oracle.ilu0(B)
        zo = oracle.ilu0_apply(B, lu, dp, r)
        tol = 1e-13 if prec == capi.MI_PRECISION_FP64 else 1e-5
        assert np.linalg.norm(z - zo) <= tol * np.linalg.norm(zo), prec
        x, st, rc = s.solve(b[perm])
        xo, ito, _, _ = oracle.bicgstab(B, b[perm], oracle.PC_ILU, 1e-8, 1000)
        assert rc == capi.MI_OK and abs(st.iterations - ito) <= 2



def test_ssor_apply_bit_exact(gpu, momentum):
    """SSOR(1) sweeps in the caller's order against the oracle's serial
    SsorPreconditioner (LinearSolvers.cpp:163-236): exact equality."""
    A, b = momentum
    s = solver(A, preconditioner=capi.MI_PC_SSOR)
    r = np.random.default_rng(4).standard_normal(A.n)
    z = s.precond_apply(r)
    assert np.array_equal(z, oracle.ssor_apply(A, r, 1.0))
    # identity: SSOR(1) returns r (tests/unit/solver/test_ssor_preconditioner.cpp:44-63)
    I = oracle.Matrix.from_dense(np.eye(5))
    si = solver(I, preconditioner=capi.MI_PC_SSOR)
    assert np.array_equal(si.precond_apply(np.arange(1.0, 6.0)), np.arange(1.0, 6.0))


def test_cg_ssor_reference_case(gpu):
    # tests/unit/solver/test_ssor_preconditioner.cpp:65-88
    rp, ci, va = tridiag_csr(50, -1.0, 2.0, -1.0)
    A = oracle.Matrix(rp, ci, va)
    s = solver(A, kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_SSOR, tolerance=1e-10,
               max_iterations=200)
    x, st, rc = s.solve(np.ones(50))
    xo, ito, _, _ = oracle.cg(A, np.ones(50), oracle.PC_SSOR, 1e-10, 200)
    assert rc == capi.MI_OK and st.iterations > 0 and abs(st.iterations - ito) <= 1
    assert np.linalg.norm(np.ones(50) - A.dense() @ x) < 1e-7


@pytest.mark.parametrize("prec", [capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED])
def test_bicgstab_ssor_parity(gpu, momentum, prec):
    A, b = momentum
    s = solver(A, preconditioner=capi.MI_PC_SSOR, precision=prec, true_residual=1)
    x, st, rc = s.solve(b)
    xo, ito, _, _ = oracle.bicgstab(A, b, oracle.PC_SSOR, 1e-8, 1000)
    rel = np.linalg.norm(b - oracle.spmv(A, x)) / np.linalg.norm(b)
    print(f"SSOR-BiCGStab prec {prec}: GPU {st.iterations} its, oracle {ito}, rel {rel:.2e}")
    assert rc == capi.MI_OK and rel <= 1e-8
    if prec == capi.MI_PRECISION_FP64:
        assert abs(st.iterations - ito) <= 2


def test_ssor_run_sweeps_red_black(gpu, momentum):
    """A red-black ordered matrix: SSOR swept run by run on the CSR-stream
    engine agrees with the oracle's serial sweeps to rounding."""
    import scipy.sparse as sp

    A, b = momentum
    N = 16
    i = np.arange(A.n)
    parity = (i % N + (i // N) % N + i // (N * N)) % 2
    perm = np.concatenate([np.flatnonzero(parity == 0), np.flatnonzero(parity == 1)])
    S = sp.csr_matrix((A.values, A.colIdx, A.rowPtr), shape=(A.n, A.n))[perm][:, perm].tocsr()
    S.sort_indices()
    B = oracle.Matrix(S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data.copy())
    s = solver(B, preconditioner=capi.MI_PC_SSOR)
    assert s.info()["levels"] == 2
    r = np.random.default_rng(5).standard_normal(B.n)
    z = s.precond_apply(r)
    zo = oracle.ssor_apply(B, r, 1.0)
    assert np.linalg.norm(z - zo) <= 1e-13 * np.linalg.norm(zo)


# A 3x3 matrix whose row 1 stores no diagonal (columns 0 and 2 only).
KAT_RP = np.array([0, 2, 4, 5], np.int32)
KAT_CI = np.array([0, 2, 0, 2, 2], np.int32)
KAT_VA = np.array([2.0, 1.0, 1.0, 3.0, 4.0])


def test_ilu0_missing_diagonal_known_answer(gpu):
    """ILU0Preconditioner (src/solver/LinearSolvers.cpp:81-151) on KAT_*, by
    hand: setup eliminates row 1 with row 0 (m = 1/2, a_12 = 3 - 1/2 = 2.5), so
    lu = [2, 1, 0.5, 2.5, 4].  For r = (1, 2, 8): y = (1, 1.5, 8); z_2 = 8/4 = 2;
    row 1 has diagPtr_ = -1, so its loop runs from position 0 through the end
    of row 1, subtracting 1*z_2 (row 0's entry) and 2.5*z_2 with no division:
    z_1 = 1.5 - 2 - 5 = -5.5; z_0 = (1 - 1*2)/2 = -0.5.  The device level
    sweeps (caller's order) and the oracle both give exactly that."""
    A = oracle.Matrix(KAT_RP, KAT_CI, KAT_VA)
    lu, dp = oracle.ilu0(A)
    assert np.array_equal(lu, [2.0, 1.0, 0.5, 2.5, 4.0]) and dp[1] == -1
    r = np.array([1.0, 2.0, 8.0])
    expect = np.array([-0.5, -5.5, 2.0])
    assert np.array_equal(oracle.ilu0_apply(A, lu, dp, r), expect)
    s = solver(A)
    assert np.array_equal(s.precond_apply(r), expect)


def test_ilu0_facade_missing_diagonal_and_zero_pivot(gpu):
    """The GpuKernels facade (src/gpu/GpuKernels.cpp:345-477): the factorization
    refuses a matrix with a row lacking its diagonal and leaves the values as
    they were; it stops at the first zero pivot with the rows before it
    factored; its apply sweeps a row without a diagonal over the row's own
    upper entries and divides by 1."""
    drp, dci = capi.DeviceArray.from_host(KAT_RP), capi.DeviceArray.from_host(KAT_CI)
    dv = capi.DeviceArray.from_host(KAT_VA)
    ok = capi.C.c_int(7)
    capi.check(gpu.mi_ilu0_factor_inplace(3, drp.ptr, dci.ptr, dv.ptr, capi.C.byref(ok)))
    assert ok.value == 0 and np.array_equal(dv.download(), KAT_VA)
    # apply on the factor [2, 1, 0.5, 2.5, 4]: y = (1, 1.5, 8), z_2 = 2,
    # z_1 = 1.5 - 2.5*2 = -3.5 (no division), z_0 = (1 - 2)/2 = -0.5
    dlu = capi.DeviceArray.from_host(np.array([2.0, 1.0, 0.5, 2.5, 4.0]))
    dr, dz = capi.DeviceArray.from_host(np.array([1.0, 2.0, 8.0])), capi.DeviceArray(3)
    capi.check(gpu.mi_ilu0_apply(3, drp.ptr, dci.ptr, dlu.ptr, dr.ptr, dz.ptr, None))
    assert np.array_equal(dz.download(), [-0.5, -3.5, 2.0])
    # [[1,1,0],[1,1,0],[0,0,1]]: row 1 becomes (m = 1, a_11 = 1 - 1 = 0) and the
    # facade returns false there, row 2 untouched
    rp = np.array([0, 2, 4, 5], np.int32)
    ci = np.array([0, 1, 0, 1, 2], np.int32)
    dv2 = capi.DeviceArray.from_host(np.array([1.0, 1.0, 1.0, 1.0, 1.0]))
    d2rp, d2ci = capi.DeviceArray.from_host(rp), capi.DeviceArray.from_host(ci)
    capi.check(gpu.mi_ilu0_factor_inplace(3, d2rp.ptr, d2ci.ptr, dv2.ptr, capi.C.byref(ok)))
    assert ok.value == 0 and np.array_equal(dv2.download(), [1.0, 1.0, 1.0, 0.0, 1.0])
    # a stored zero on row 0's diagonal is met as the pivot of row 1: nothing changes
    dv3 = capi.DeviceArray.from_host(np.array([0.0, 1.0, 1.0, 1.0, 1.0]))
    capi.check(gpu.mi_ilu0_factor_inplace(3, d2rp.ptr, d2ci.ptr, dv3.ptr, capi.C.byref(ok)))
    assert ok.value == 0 and np.array_equal(dv3.download(), [0.0, 1.0, 1.0, 1.0, 1.0])


def _missing_diagonal_matrix(n=3000, seed=4):
    """Random nonsymmetric, diagonally dominant rows with sorted columns; every
    37th row stores no diagonal and every 53rd a zero one."""
    rng = np.random.default_rng(seed)
    rp, ci, va = [0], [], []
    for i in range(n):
        cols = set((i + rng.integers(-40, 41, 6)).clip(0, n - 1).tolist())
        cols.add(i)
        if i % 37 == 5:
            cols.discard(i)
        for j in sorted(cols):
            ci.append(j)
            if j == i:
                va.append(0.0 if i % 53 == 7 else 8.0 + rng.random())
            else:
                va.append(rng.uniform(-1, 1))
        rp.append(len(ci))
    return oracle.Matrix(np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))


def test_ilu0_and_ssor_rows_without_diagonal_bit_exact(gpu):
    """Rows without a stored diagonal and rows with a zero one: the device
    ILU(0) (host factorization, level sweeps with the reference's loop for such
    rows) and SSOR(1) sweeps equal the oracle bit for bit."""
    A = _missing_diagonal_matrix()
    r = np.random.default_rng(9).standard_normal(A.n)
    lu, dp = oracle.ilu0(A)
    assert (dp < 0).sum() > 50
    s = solver(A)
    assert np.array_equal(s.precond_apply(r), oracle.ilu0_apply(A, lu, dp, r))
    s = solver(A, preconditioner=capi.MI_PC_SSOR)
    assert np.array_equal(s.precond_apply(r), oracle.ssor_apply(A, r, 1.0))
