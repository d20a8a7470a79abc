"""GPU parity of the vector kernels, SpMV and Jacobi-PCG against the oracle.

Tolerances: SpMV and dots are FP64 with a different summation order than the
serial reference (lane-strided partial sums), so they agree to a few ulps of
the row magnitude (rtol 1e-13 against sums of |a_ij x_j|).  Krylov iterates
agree to the iteration count +-2 and to the solution within 1e-6 relative at a
1e-8 residual tolerance (both runs stop below 1e-8; the difference is bounded
by the conditioning of the pinned system).
"""
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


def random_csr(n, seed, min_len=0, max_len=9, with_diag=True):
    rng = np.random.default_rng(seed)
    rp, ci = [0], []
    for i in range(n):
        k = int(rng.integers(min_len, max_len + 1))
        cols = set(rng.choice(n, size=min(k, n), replace=False).tolist())
        if with_diag:
            cols.add(i)
        ci.extend(sorted(cols))
        rp.append(len(ci))
    va = rng.standard_normal(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), va


def test_blas1_known_answers(gpu):
    # tests/unit/gpu/test_gpu_models.cpp:117-141 on device pointers
    x = capi.DeviceArray.from_host(np.arange(1.0, 6.0))
    y = capi.DeviceArray.from_host(np.ones(5))
    capi.check(gpu.mi_axpy(2.0, x.ptr, y.ptr, 5, None))
    gpu.mi_device_synchronize()
    assert np.allclose(y.download(), 1.0 + 2.0 * np.arange(1.0, 6.0), atol=1e-12)
    capi.check(gpu.mi_scal(0.5, y.ptr, 5, None))
    gpu.mi_device_synchronize()
    assert abs(y.download()[0] - 1.5) < 1e-12
    d = capi.C.c_double()
    capi.check(gpu.mi_dot(x.ptr, x.ptr, 5, capi.C.byref(d), None))
    assert d.value == 55.0
    capi.check(gpu.mi_nrm2(x.ptr, 5, capi.C.byref(d), None))
    assert abs(d.value - np.sqrt(55.0)) < 1e-12 * np.sqrt(55.0)
    # axpby / xpby / copy / jacobi
    capi.check(gpu.mi_axpby(2.0, x.ptr, -1.0, y.ptr, 5, None))
    gpu.mi_device_synchronize()
    yy = 2.0 * np.arange(1.0, 6.0) - 0.5 * (1.0 + 2.0 * np.arange(1.0, 6.0))
    assert np.allclose(y.download(), yy, atol=1e-12)
    capi.check(gpu.mi_xpby(x.ptr, 3.0, y.ptr, 5, None))
    gpu.mi_device_synchronize()
    assert np.allclose(y.download(), np.arange(1.0, 6.0) + 3.0 * yy, atol=1e-12)
    z = capi.DeviceArray(5)
    capi.check(gpu.mi_jacobi_apply(x.ptr, y.ptr, z.ptr, 5, None))
    gpu.mi_device_synchronize()
    assert np.allclose(z.download(), y.download() / np.arange(1.0, 6.0), rtol=1e-15)


@pytest.mark.parametrize("n", [1, 1000, 100003, 3000000])
def test_dot_large_deterministic(gpu, n):
    rng = np.random.default_rng(n)
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    da, db = capi.DeviceArray.from_host(a), capi.DeviceArray.from_host(b)
    d1, d2 = capi.C.c_double(), capi.C.c_double()
    capi.check(gpu.mi_dot(da.ptr, db.ptr, n, capi.C.byref(d1), None))
    capi.check(gpu.mi_dot(da.ptr, db.ptr, n, capi.C.byref(d2), None))
    assert d1.value == d2.value  # fixed reduction order
    exact = float(np.sum(np.multiply(a, b, dtype=np.longdouble)))
    assert abs(d1.value - exact) <= 1e-13 * np.abs(a * b).sum()


def test_spmv_known_answer(gpu):
    # tests/unit/gpu/test_gpu_models.cpp:144-178, through both entry points
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    va = np.array([4, -1, -1, 4, -1, -1, 4, -1, -1, 4], np.float64)
    A = capi.Csr(rp, ci, va)
    assert np.array_equal(A.spmv([1, 2, 3, 4]), [2, 4, 6, 13])
    drp, dci, dva = (capi.DeviceArray.from_host(v) for v in (rp, ci, va))
    dx = capi.DeviceArray.from_host(np.arange(1.0, 5.0))
    dy = capi.DeviceArray(4)
    capi.check(gpu.mi_spmv_csr(4, drp.ptr, dci.ptr, dva.ptr, dx.ptr, dy.ptr, None))
    gpu.mi_device_synchronize()
    assert np.array_equal(dy.download(), [2, 4, 6, 13])


@pytest.mark.parametrize("max_len", [0, 3, 7, 15, 30, 60, 200])
def test_spmv_random_row_lengths(gpu, max_len):
    n = 5000
    rp, ci, va = random_csr(n, max_len, 0, max_len, with_diag=max_len > 0)
    A = capi.Csr(rp, ci, va)
    x = np.random.default_rng(1).standard_normal(n)
    y = A.spmv(x)
    Ao = oracle.Matrix(rp, ci, va)
    yo = oracle.spmv(Ao, x)
    # the CSR-stream engine rounds each product and adds in column order, as
    # CSRMatrix::spmv does: identical results
    assert np.array_equal(y, yo)
    scale = oracle.spmv(oracle.Matrix(rp, ci, np.abs(va)), np.abs(x)) + 1e-300
    # the raw-pointer entry point (vector-CSR, lane-strided sums): a few ulps
    drp, dci, dva = (capi.DeviceArray.from_host(v) for v in (rp, ci, va))
    dxx, dyy = capi.DeviceArray.from_host(x), capi.DeviceArray(n)
    capi.check(gpu.mi_spmv_csr(n, drp.ptr, dci.ptr, dva.ptr, dxx.ptr, dyy.ptr, None))
    gpu.mi_device_synchronize()
    assert np.all(np.abs(dyy.download() - yo) <= 1e-13 * scale)
    # FP32 values and vectors
    dx = capi.DeviceArray.from_host(x.astype(np.float32))
    dy = capi.DeviceArray(n, np.float32)
    A.spmv_device(dx, dy, capi.MI_PRECISION_MIXED)
    y32 = dy.download().astype(np.float64)
    assert np.all(np.abs(y32 - yo) <= 2e-6 * scale)


def test_spmv_long_rows(gpu):
    # rows longer than one LDS row block take the block-reduction path
    rng = np.random.default_rng(9)
    n = 6000
    rp, ci = [0], []
    for i in range(n):
        k = 5000 if i in (3, 4000) else int(rng.integers(1, 9))
        ci.extend(sorted(rng.choice(n, size=k, replace=False).tolist()))
        rp.append(len(ci))
    rp, ci = np.array(rp, np.int32), np.array(ci, np.int32)
    va = rng.standard_normal(ci.size)
    x = rng.standard_normal(n)
    y = capi.Csr(rp, ci, va).spmv(x)
    yo = oracle.spmv(oracle.Matrix(rp, ci, va), x)
    long_rows = np.array([3, 4000])
    short = np.setdiff1d(np.arange(n), long_rows)
    assert np.array_equal(y[short], yo[short])
    scale = oracle.spmv(oracle.Matrix(rp, ci, np.abs(va)), np.abs(x))
    assert np.all(np.abs(y - yo) <= 1e-13 * scale)


@pytest.mark.parametrize("far", [False, True])
def test_spmv_column_ordered_blocks(gpu, far):
    """Row blocks are stored in column order when each block's columns span
    fewer than 2^21 - 1, and in CSR order otherwise (far=True: one entry of
    row 5 reaches the last column of a 2.1M-row matrix).  Either way the
    products are summed in CSR order: identical results.  Row 9 holds 24
    entries, more than the SELL-64 form stores, so the row blocks run."""
    rng = np.random.default_rng(21)
    n = (1 << 21) + 4096
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    for off in (-1, 1, -700, 700, -40000, 40000):
        i = np.arange(max(0, -off), min(n, n - off))
        rows.append(i)
        cols.append(i + off)
    wide = np.setdiff1d(np.arange(100, 2500, 100), [9 - 1, 9 + 1, 9 + 700])[:17]
    rows.append(np.full(wide.size, 9))
    cols.append(wide)
    if far:
        rows.append(np.array([5]))
        cols.append(np.array([n - 1]))
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    ci = c.astype(np.int32)
    va = rng.standard_normal(ci.size)
    x = rng.standard_normal(n)
    A = capi.Csr(rp, ci, va)
    y = A.spmv(x)
    yo = oracle.spmv(oracle.Matrix(rp, ci, va), x)
    assert np.array_equal(y, yo)
    dx = capi.DeviceArray.from_host(x.astype(np.float32))
    dy = capi.DeviceArray(n, np.float32)
    A.spmv_device(dx, dy, capi.MI_PRECISION_MIXED)
    scale = oracle.spmv(oracle.Matrix(rp, ci, np.abs(va)), np.abs(x)) + 1e-300
    assert np.all(np.abs(dy.download().astype(np.float64) - yo) <= 2e-6 * scale)


@pytest.mark.parametrize("n,width", [(64 * 900 + 37, 7), (64 * 700 + 5, 13), (50, 3)])
def test_spmv_sell_form(gpu, n, width):
    """Matrices whose rows are all short (at most 16 entries, padding within a
    quarter of nnz) run on the SELL-64 form: one lane per row summing its CSR
    entries in order.  Ragged rows (some empty, some shorter), n not a multiple
    of 64, and rows longer than the 8 slots loaded ahead (width 13): FP64
    results identical to the oracle's; the FP32 product within rounding."""
    rng = np.random.default_rng(n)
    lens = np.full(n, width)
    lens[rng.choice(n, n // 10, replace=False)] = rng.integers(0, width, n // 10)
    lens[: min(n, 3)] = 0
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    ci = np.empty(rp[-1], np.int32)
    for i in range(n):
        if lens[i]:
            lo = max(0, min(i - 3 * width, n - 6 * width))
            pool = np.arange(lo, min(n, lo + 6 * width))
            ci[rp[i]:rp[i + 1]] = np.sort(rng.choice(pool, lens[i], replace=False))
    rp = rp.astype(np.int32)
    va = rng.standard_normal(ci.size)
    va[rng.random(ci.size) < 0.01] = 0.0
    x = rng.standard_normal(n)
    A = capi.Csr(rp, ci, va)
    y = A.spmv(x)
    yo = oracle.spmv(oracle.Matrix(rp, ci, va), x)
    assert np.array_equal(y, yo)
    dx = capi.DeviceArray.from_host(x)
    dy = capi.DeviceArray(n, np.float64)
    A.spmv_device(dx, dy, capi.MI_PRECISION_FP64)
    assert np.array_equal(dy.download(), yo)
    dx32 = capi.DeviceArray.from_host(x.astype(np.float32))
    dy32 = capi.DeviceArray(n, np.float32)
    A.spmv_device(dx32, dy32, capi.MI_PRECISION_MIXED)
    scale = oracle.spmv(oracle.Matrix(rp, ci, np.abs(va)), np.abs(x)) + 1e-300
    assert np.all(np.abs(dy32.download().astype(np.float64) - yo) <= 2e-6 * scale)


def test_spmv_empty_matrix(gpu):
    A = capi.Csr(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert A.spmv(np.zeros(0)).size == 0


def test_csr_rejects_bad_pattern(gpu):
    with pytest.raises(capi.MiError) as e:
        capi.Csr([0, 2, 1], [0, 1], [1.0, 2.0])
    assert e.value.code == capi.MI_ERR_BAD_ARG
    with pytest.raises(capi.MiError):
        capi.Csr([0, 1], [5], [1.0])


def test_cg_tridiagonal_known_answer(gpu):
    # tests/unit/gpu/test_gpu_models.cpp:181-235
    rp, ci, va = tridiag_csr(10, -1.0, 4.0, -1.0)
    b = np.zeros(10)
    b[0] = b[-1] = 1.0
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-10,
                    max_iterations=200)
    A = capi.Csr(rp, ci, va)
    s.setup(A)
    x, st, rc = s.solve(b)
    assert rc == capi.MI_OK and st.converged and st.iterations < 50
    D = oracle.Matrix(rp, ci, va).dense()
    assert np.linalg.norm(b - D @ x) < 1e-8
    xo, ito, lasto, ho = oracle.cg(oracle.Matrix(rp, ci, va), b, oracle.PC_JACOBI, 1e-10, 200)
    assert st.iterations == ito
    # identical residual history down to the rounding floor (~1e-16 of ||b||)
    assert np.allclose(s.history(), ho, rtol=1e-10, atol=1e-14 * ho[0])


def test_cg_zero_rhs_and_limits(gpu):
    rp, ci, va = tridiag_csr(50, -1.0, 2.0, -1.0)
    A = capi.Csr(rp, ci, va)
    s = capi.Solver(kind=capi.MI_SOLVER_CG, tolerance=1e-8, max_iterations=100)
    s.setup(A)
    x, st, rc = s.solve(np.zeros(50))
    # reference: pAp == 0 on the first iteration -> returns 0, last = 0
    assert st.iterations == 0 and not x.any() and rc == capi.MI_OK
    s2 = capi.Solver(kind=capi.MI_SOLVER_CG, tolerance=1e-14, max_iterations=5)
    s2.setup(A)
    x, st, rc = s2.solve(np.ones(50))
    xo, ito, lasto, _ = oracle.cg(oracle.Matrix(rp, ci, va), np.ones(50), oracle.PC_JACOBI, 1e-14, 5)
    assert rc == capi.MI_ERR_NOT_CONVERGED and st.iterations == ito == 5
    assert abs(s2.last_residual() - lasto) < 1e-12 * lasto
    assert np.allclose(x, xo, rtol=1e-12, atol=1e-14)


@pytest.fixture(scope="module")
def cavity():
    w = wl.cavity2d(128)
    rp, ci = oracle.build_sparsity(w.mesh)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, pinned = oracle.assemble_pressure(w.mesh, P, w.bc_of_zone)
    b = w.b.copy()
    b[0] = 0.0
    return w, oracle.Matrix(rp, ci, vals), b


def test_cavity_jacobi_pcg_parity(gpu, cavity):
    w, Ao, b = cavity
    xo, ito, lasto, ho = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-8, 10000)
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                    max_iterations=10000)
    A = capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values)
    s.setup(A)
    x, st, rc = s.solve(b)
    print(f"cavity 128^2 Jacobi-PCG: GPU {st.iterations} its, oracle {ito}; "
          f"true rel res {st.true_relative_residual:.3e}")
    assert rc == capi.MI_OK and abs(st.iterations - ito) <= 2
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6
    h = s.history()
    k = min(len(h), len(ho)) // 2
    assert np.allclose(h[:k], ho[:k], rtol=1e-6)


def test_cavity_mixed_precision_true_residual(gpu, cavity):
    w, Ao, b = cavity
    xt, _, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-13, 20000)
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                    max_iterations=20000, precision=capi.MI_PRECISION_MIXED)
    s.setup(capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values))
    x, st, rc = s.solve(b)
    r = b - oracle.spmv(Ao, x)
    rel = np.linalg.norm(r) / np.linalg.norm(b)
    print(f"cavity mixed: {st.iterations} inner its, {st.outer_iterations} sweeps, rel res {rel:.3e}")
    assert rc == capi.MI_OK and rel <= 1e-8
    assert abs(rel - st.true_relative_residual) <= 1e-3 * rel + 1e-15


def _ragged_spd(seed=5):
    """A 27x29x25 7-point operator with random conductances, plus 3 long
    symmetric couplings on each of 20 rows (up to 10 entries): short, ragged rows,
    n = 19575 (not a multiple of 64), strictly diagonally dominant."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    nx, ny, nz = 27, 29, 25
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    pairs = [(idx[:, :, :-1], idx[:, :, 1:]), (idx[:, :-1, :], idx[:, 1:, :]),
             (idx[:-1, :, :], idx[1:, :, :])]
    r = np.concatenate([a.ravel() for a, _ in pairs])
    c = np.concatenate([b.ravel() for _, b in pairs])
    far_r = np.repeat(rng.choice(n, 20, replace=False), 3)
    far_c = (far_r + rng.integers(500, 5000, far_r.size)) % n
    keep = far_r != far_c
    r = np.concatenate([r, far_r[keep]])
    c = np.concatenate([c, far_c[keep]])
    w = 10.0 ** rng.uniform(-1, 1, r.size)
    A = sp.coo_matrix((-w, (r, c)), shape=(n, n))
    A = (A + A.T).tocsr()
    A.sum_duplicates()
    d = -np.asarray(A.sum(axis=1)).ravel() + 0.05
    A = (A + sp.diags(d)).tocsr()
    A.sort_indices()
    rp, ci = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    lens = np.diff(rp)
    ns = (n + 63) // 64
    L = np.zeros(ns * 64, np.int64)
    L[:n] = lens
    padded = 64 * L.reshape(ns, 64).max(1).sum()
    assert lens.max() <= 16 and padded <= 1.25 * ci.size and lens.max() > 8
    return oracle.Matrix(rp, ci, A.data.copy()), rng.standard_normal(n)


def test_krylov_on_sell_form_ragged_rows(gpu):
    """CG and BiCGStab whose operator runs on the SELL-64 form (short ragged
    rows, some longer than the 8 slots loaded ahead): FP64 Jacobi-PCG follows
    the oracle's iteration; the mixed solve (FP32 difference form, diagonal
    slot flagged in the slot columns) meets the FP64 true-residual tolerance."""
    Ao, b = _ragged_spd()
    A = capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values)
    xo, ito, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-10, 5000)
    s = capi.Solver(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-10,
                    max_iterations=5000)
    s.setup(A)
    x, st, rc = s.solve(b)
    assert rc == capi.MI_OK and abs(st.iterations - ito) <= 2
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    for kind in (capi.MI_SOLVER_CG, capi.MI_SOLVER_BICGSTAB):
        for prec in (capi.MI_PRECISION_FP64, capi.MI_PRECISION_MIXED):
            s = capi.Solver(kind=kind, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-9,
                            max_iterations=5000, precision=prec)
            s.setup(A)
            x, st, rc = s.solve(b)
            rel = np.linalg.norm(b - oracle.spmv(Ao, x)) / np.linalg.norm(b)
            assert rc == capi.MI_OK and rel <= 1.5e-9, (kind, prec, rel)


def test_fp64_true_residual_mode(gpu, cavity):
    w, Ao, b = cavity
    s = capi.Solver(kind=capi.MI_SOLVER_CG, tolerance=1e-8, max_iterations=20000, true_residual=1)
    s.setup(capi.Csr(Ao.rowPtr, Ao.colIdx, Ao.values))
    x, st, rc = s.solve(b)
    rel = np.linalg.norm(b - oracle.spmv(Ao, x)) / np.linalg.norm(b)
    assert rc == capi.MI_OK and rel < 1e-8


@pytest.mark.parametrize("kind", [capi.MI_SOLVER_CG, capi.MI_SOLVER_BICGSTAB])
def test_mixed_refinement_never_converges_on_nan(gpu, kind):
    """A NaN in b makes the FP64 true residual NaN; the mixed refinement loop
    must report failure, as the reference's `last_ < tol` never holds for NaN
    (src/solver/LinearSolvers.cpp:323, :440-443)."""
    rp, ci, va = tridiag_csr(300, -1.0, 4.0, -1.0)
    A = capi.Csr(rp, ci, va)
    b = np.ones(300)
    b[17] = np.nan
    s = capi.Solver(kind=kind, preconditioner=capi.MI_PC_JACOBI, tolerance=1e-8,
                    max_iterations=100, precision=capi.MI_PRECISION_MIXED)
    s.setup(A)
    x, st, rc = s.solve(b)
    assert rc == capi.MI_ERR_BREAKDOWN and not st.converged
    assert not np.isfinite(s.last_residual()) or s.last_residual() >= 1e-8
    # the same solver still solves a finite system afterwards
    b[17] = 1.0
    x, st, rc = s.solve(b)
    assert rc == capi.MI_OK and st.converged
    assert np.linalg.norm(b - oracle.spmv(oracle.Matrix(rp, ci, va), x)) <= 1e-8 * np.linalg.norm(b)


def test_values_written_through_the_device_pointer(gpu):
    """Writes through mi_csr_values_device take effect once the caller marks
    them with mi_csr_update_values_device(A, ptr) (no copy: same pointer);
    the SELL-64 / column-ordered value copies and the solver follow."""
    import ctypes as C
    rp, ci, va = tridiag_csr(5000, -1.0, 4.0, -1.0)
    A = capi.Csr(rp, ci, va)
    x = np.random.default_rng(2).standard_normal(5000)
    y0 = A.spmv(x)
    p = C.c_void_p()
    capi.check(gpu.mi_csr_values_device(A.h, C.byref(p)))
    new = np.ascontiguousarray(va * 2.0)
    capi.check(gpu.mi_memcpy_h2d(p, new.ctypes.data, new.nbytes))
    capi.check(gpu.mi_csr_update_values_device(A.h, p))
    assert np.array_equal(A.values(), new)
    assert np.array_equal(A.spmv(x), 2.0 * y0)  # scaling by 2 is exact
    s = capi.Solver(kind=capi.MI_SOLVER_CG, tolerance=1e-10, max_iterations=500)
    s.setup(A)
    xs, st, rc = s.solve(y0)
    assert rc == capi.MI_OK and np.allclose(xs, 0.5 * x, rtol=0, atol=1e-8)
    # a second in-place write, again marked, is seen by the same solver
    capi.check(gpu.mi_memcpy_h2d(p, va.ctypes.data, va.nbytes))
    capi.check(gpu.mi_csr_update_values_device(A.h, p))
    xs, st, rc = s.solve(y0)
    assert rc == capi.MI_OK and np.allclose(xs, x, rtol=0, atol=1e-8)


def _fp32_rowsum(rp, ci, va, x):
    """y_i = sum over the row's entries in CSR order of fl32(a * x), each
    addition rounded to FP32: the FP32 SELL-64 kernels' arithmetic."""
    n = rp.size - 1
    lens = np.diff(rp)
    a32, x32 = va.astype(np.float32), x.astype(np.float32)
    s = np.zeros(n, np.float32)
    for k in range(int(lens.max(initial=0))):
        rows = np.nonzero(lens > k)[0]
        idx = rp[rows] + k
        s[rows] = (s[rows] + a32[idx] * x32[ci[idx]]).astype(np.float32)
    return s


@pytest.mark.parametrize("band", ["narrow", "wide"])
def test_spmv_sell_fp32_column_forms(gpu, band):
    """A square SELL-64 operator whose entries lie within 32767 of the
    diagonal streams 16-bit column offsets in its FP32 kernels (6 instead of
    8 bytes per slot), a wider one int32 columns; both sum the same FP32
    products in CSR order, equal to an FP32 emulation bit for bit."""
    rng = np.random.default_rng(11)
    n = 70000
    reach = 2000 if band == "narrow" else 60000
    offs = rng.integers(-reach, reach + 1, size=(n, 7))
    lens0 = rng.integers(0, 8, n)
    rows = []
    for i in range(n):
        c = np.clip(i + offs[i, :lens0[i]], 0, n - 1)
        rows.append(np.unique(np.concatenate([[i], c])))  # one diagonal entry per row
    lens = np.array([r.size for r in rows])
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate(rows).astype(np.int32)
    rp = rp.astype(np.int32)
    far = np.abs(ci.astype(np.int64) - np.repeat(np.arange(n), lens)).max()
    assert (far < 32767) == (band == "narrow")
    va = rng.standard_normal(ci.size)
    x = rng.standard_normal(n)
    A = capi.Csr(rp, ci, va)
    dx32 = capi.DeviceArray.from_host(x.astype(np.float32))
    dy32 = capi.DeviceArray(n, np.float32)
    A.spmv_device(dx32, dy32, capi.MI_PRECISION_MIXED)
    assert np.array_equal(dy32.download(), _fp32_rowsum(rp, ci, va, x))
