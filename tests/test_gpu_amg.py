"""GPU SA-AMG and internal reordering against the numpy model and the oracle.

Tolerances: the device V-cycle in FP64 matches the numpy model of the same
hierarchy to 1e-11 relative (different summation order in SpMV rows); the
FP32 hierarchy to 1e-4 relative.  Solves are checked on the FP64 true
residual (<= 1e-8) and against the oracle's tight (1e-13) Jacobi-PCG solution.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from simall_amd import capi
from simall_amd import workload as wl

from amg_model import VCycle, pcg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys24():
    w = wl.poisson3d(24)
    m = w.mesh
    rp, ci = oracle.build_sparsity(m)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, _ = oracle.assemble_pressure(m, P, w.bc_of_zone, gamma=w.gamma)
    b = w.b.copy()
    b[0] = 0.0
    A = sp.csr_matrix((vals, ci, rp), shape=(m.n_cells,) * 2)
    Ao = oracle.Matrix(rp, ci, vals)
    xt, _, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-13, 20000)
    return A, Ao, b, xt


def solver(A, **kw):
    cfg = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-8,
               max_iterations=500)
    cfg.update(kw)
    s = capi.Solver(**cfg)
    s.setup(capi.Csr(A.indptr, A.indices, A.data))
    return s


@pytest.mark.parametrize("reorder", [0, 1])
def test_vcycle_matches_model_fp64(gpu, sys24, reorder):
    A, Ao, b, xt = sys24
    s = solver(A, reorder=reorder)
    r = np.random.default_rng(3).standard_normal(A.shape[0])
    z = s.precond_apply(r)
    if reorder:
        perm, _ = capi.rcm_permutation(A.indptr, A.indices)
        B = A[perm][:, perm].tocsr()
        B.sort_indices()
    else:
        perm, B = np.arange(A.shape[0]), A
    H = capi.AmgHost(B.indptr, B.indices, B.data)
    zm = np.empty_like(r)
    zm[perm] = VCycle(H, 2)(r[perm])
    assert s.info()["levels"] == H.levels
    assert np.linalg.norm(z - zm) <= 1e-11 * np.linalg.norm(zm)


def test_vcycle_matches_model_fp32(gpu, sys24):
    A, Ao, b, xt = sys24
    s = solver(A, reorder=0, precision=capi.MI_PRECISION_MIXED)
    r = np.random.default_rng(4).standard_normal(A.shape[0])
    z = s.precond_apply(r)
    H = capi.AmgHost(A.indptr, A.indices, A.data)
    zm = VCycle(H, 2)(r)
    assert np.linalg.norm(z - zm) <= 1e-4 * np.linalg.norm(zm)


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_amg_pcg_fp64_parity(gpu, sys24, degree):
    A, Ao, b, xt = sys24
    s = solver(A, reorder=0, amg_sweeps=degree)
    x, st, rc = s.solve(b)
    H = capi.AmgHost(A.indptr, A.indices, A.data)
    xm, itm = pcg(A, b, VCycle(H, degree), tol=1e-8)
    print(f"AMG-PCG degree {degree}: GPU {st.iterations} its, model {itm}")
    assert rc == capi.MI_OK and abs(st.iterations - itm) <= 1
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel < 1e-8
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-5


def test_amg_pcg_reordered_and_true_residual(gpu, sys24):
    A, Ao, b, xt = sys24
    s = solver(A, reorder=1, true_residual=1)
    x, st, rc = s.solve(b)
    info = s.info()
    assert info["bandwidth_after"] < info["bandwidth_before"] / 4
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rc == capi.MI_OK and rel < 1e-8
    assert abs(st.true_relative_residual - rel) <= 1e-3 * rel


def test_amg_mixed_precision(gpu, sys24):
    A, Ao, b, xt = sys24
    s = solver(A, reorder=1, precision=capi.MI_PRECISION_MIXED)
    x, st, rc = s.solve(b)
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    print(f"mixed AMG: {st.iterations} inner its in {st.outer_iterations} sweeps, rel {rel:.2e}")
    assert rc == capi.MI_OK and rel <= 1e-8
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) < 1e-5


def test_jacobi_with_and_without_reordering(gpu):
    w = wl.poisson3d(14)
    m = w.mesh
    rp, ci = oracle.build_sparsity(m)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, _ = oracle.assemble_pressure(m, P, w.bc_of_zone, gamma=w.gamma)
    b = w.b.copy()
    b[0] = 0.0
    Ao = oracle.Matrix(rp, ci, vals)
    xo, ito, _, _ = oracle.cg(Ao, b, oracle.PC_JACOBI, 1e-8, 20000)
    A = capi.Csr(rp, ci, vals)
    res = []
    for reorder in (0, 1):
        s = capi.Solver(kind=capi.MI_SOLVER_CG, tolerance=1e-8, max_iterations=20000, reorder=reorder)
        s.setup(A)
        x, st, rc = s.solve(b)
        assert rc == capi.MI_OK
        res.append((x, st.iterations))
    assert abs(res[0][1] - ito) <= 2 and abs(res[1][1] - ito) <= 2
    assert np.linalg.norm(res[1][0] - xo) / np.linalg.norm(xo) < 1e-6


def test_spmv_timing_entry_points(gpu, sys24):
    """mi_solver_time_spmv (back to back) and _cold (a read sweep of at least
    512 MiB before every launch) time the working matrix and report its
    algorithmic bytes; a smaller flush is refused."""
    A, Ao, b, xt = sys24
    s = solver(A, reorder=1)
    n, nnz = A.shape[0], A.nnz
    for prec, vb in ((capi.MI_PRECISION_FP64, 8), (capi.MI_PRECISION_MIXED, 4)):
        us, by = s.time_spmv(prec, reps=5)
        usc, byc = s.time_spmv_cold(prec, reps=3)
        assert us > 0 and usc > 0
        assert by == byc == nnz * (vb + 4) + (n + 1) * 4 + 2 * n * vb
    with pytest.raises(capi.MiError):
        s.time_spmv_cold(capi.MI_PRECISION_FP64, reps=1, flush_bytes=1 << 20)
