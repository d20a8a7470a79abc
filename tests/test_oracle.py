"""Pins the CPU oracle (CPU only).

Each oracle routine is checked against the reference's own known-answer tests
(tests/unit/gpu/test_gpu_models.cpp, tests/unit/solver/test_ssor_preconditioner.cpp,
tests/unit/longtail/test_longtail_modules.cpp) and against what the mathematics
fixes (dense brute force, closed forms, exact factorizations, invariants), so
that a dropped term, wrong sign or transposed operand fails a check.
"""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle


def tridiag(n, lo, d, up):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = d
        if i > 0:
            A[i, i - 1] = lo
        if i + 1 < n:
            A[i, i + 1] = up
    return A


def random_sparse(n, density, seed, diag_boost=0.0, symmetric=False):
    rng = np.random.default_rng(seed)
    A = np.where(rng.random((n, n)) < density, rng.standard_normal((n, n)), 0.0)
    if symmetric:
        A = A + A.T
    A += np.diag(np.abs(A).sum(1) + diag_boost)
    return A


# ------------------------------------------------------------------- KATs
def test_blas1_known_answers():
    # tests/unit/gpu/test_gpu_models.cpp:117-141
    x = np.arange(1.0, 6.0)
    y = oracle.axpy(2.0, x, np.ones(5))
    assert np.allclose(y, 1.0 + 2.0 * x, atol=1e-12, rtol=0)
    assert abs(oracle.scal(0.5, y)[0] - 1.5) < 1e-12
    assert oracle.dot(x, x) == 55.0
    assert abs(oracle.norm2(x) - math.sqrt(55.0)) < 1e-12 * math.sqrt(55.0)


def test_spmv_known_answer():
    # tests/unit/gpu/test_gpu_models.cpp:144-178
    A = oracle.Matrix([0, 2, 5, 8, 10], [0, 1, 0, 1, 2, 1, 2, 3, 2, 3],
                      [4, -1, -1, 4, -1, -1, 4, -1, -1, 4])
    y = oracle.spmv(A, [1, 2, 3, 4])
    assert np.array_equal(y, [2, 4, 6, 13])


@pytest.mark.parametrize("seed", range(4))
def test_spmv_dense_brute_force(seed):
    D = random_sparse(60, 0.1, seed)
    A = oracle.Matrix.from_dense(D)
    x = np.random.default_rng(seed).standard_normal(60)
    assert np.allclose(oracle.spmv(A, x), D @ x, rtol=1e-13, atol=1e-13)
    assert np.allclose(oracle.spmv_transpose(A, x), D.T @ x, rtol=1e-13, atol=1e-13)


def test_jacobi_inverse_diagonal():
    D = np.diag([2.0, 0.0, -4.0]) + np.eye(3, k=1)
    A = oracle.Matrix.from_dense(D)
    # zero diagonal -> 1 (src/solver/LinearSolvers.cpp:61)
    assert np.array_equal(oracle.jacobi_inv_diag(A), [0.5, 1.0, -0.25])
    # missing diagonal -> 1
    B = oracle.Matrix([0, 1, 2], [1, 0], [3.0, 5.0])
    assert np.array_equal(oracle.jacobi_inv_diag(B), [1.0, 1.0])


def test_cg_jacobi_tridiagonal_known_answer():
    # tests/unit/gpu/test_gpu_models.cpp:181-235
    n = 10
    D = tridiag(n, -1.0, 4.0, -1.0)
    A = oracle.Matrix.from_dense(D)
    b = np.zeros(n)
    b[0] = b[-1] = 1.0
    x, it, last, hist = oracle.cg(A, b, oracle.PC_JACOBI, 1e-10, 200)
    assert it < 50 and last < 1e-10
    assert np.linalg.norm(b - D @ x) < 1e-8
    # Thomas algorithm (closed-form tridiagonal solve)
    xt = scipy.linalg.solve_banded((1, 1), np.array([np.r_[0, np.full(n - 1, -1.0)],
                                                     np.full(n, 4.0),
                                                     np.r_[np.full(n - 1, -1.0), 0]]), b)
    assert np.allclose(x, xt, rtol=0, atol=1e-10)
    assert len(hist) == it and hist[-1] == last


def test_cg_finite_termination_and_exactness():
    # CG in exact arithmetic terminates in <= n steps: with n distinct
    # eigenvalues it needs close to n; the answer must match numpy.
    n = 40
    D = random_sparse(n, 0.15, 7, diag_boost=1.0, symmetric=True)
    A = oracle.Matrix.from_dense(D)
    b = np.random.default_rng(1).standard_normal(n)
    x, it, last, _ = oracle.cg(A, b, oracle.PC_JACOBI, 1e-12, 500)
    assert it <= n + 5
    assert np.allclose(x, np.linalg.solve(D, b), rtol=1e-9, atol=1e-10)


def test_cg_zero_rhs_returns_immediately():
    D = tridiag(5, -1.0, 2.5, -1.0)
    x, it, last, _ = oracle.cg(oracle.Matrix.from_dense(D), np.zeros(5))
    assert it == 0 and last == 0.0 and not x.any()


def test_cg_iteration_limit():
    D = tridiag(200, -1.0, 2.0, -1.0)
    b = np.ones(200)
    x, it, last, hist = oracle.cg(oracle.Matrix.from_dense(D), b, oracle.PC_JACOBI, 1e-14, 7)
    assert it == 7 and last > 1e-14 and len(hist) == 7


def test_bicgstab_ilu0_known_answer():
    # tests/unit/gpu/test_gpu_models.cpp:238-287
    n = 8
    D = tridiag(n, -0.5, 3.0, 1.0)
    A = oracle.Matrix.from_dense(D)
    b = np.ones(n)
    x, it, last, _ = oracle.bicgstab(A, b, oracle.PC_ILU, 1e-10, 200)
    assert np.linalg.norm(b - D @ x) < 1e-7
    # ILU(0) of a tridiagonal has no fill: it is the exact LU, so the
    # preconditioned iteration converges in one step.
    assert it == 1


def test_ilu0_tridiagonal_is_exact_lu():
    n = 12
    D = tridiag(n, -0.7, 3.0, 1.3)
    A = oracle.Matrix.from_dense(D)
    lu, dp = oracle.ilu0(A)
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for k in range(A.rowPtr[i], A.rowPtr[i + 1]):
            j = A.colIdx[k]
            if j < i:
                L[i, j] = lu[k]
            else:
                U[i, j] = lu[k]
    P, Ls, Us = scipy.linalg.lu(D)
    assert np.allclose(P, np.eye(n))  # diagonally dominant: no pivoting
    assert np.allclose(L, Ls, atol=1e-14) and np.allclose(U, Us, atol=1e-14)
    r = np.random.default_rng(0).standard_normal(n)
    assert np.allclose(oracle.ilu0_apply(A, lu, dp, r), np.linalg.solve(D, r), atol=1e-12)


def test_ilu0_defining_property_on_laplacian():
    # (LU)_ij == a_ij for every (i,j) in the pattern of A
    m = 6
    n = m * m
    D = np.zeros((n, n))
    for i in range(m):
        for j in range(m):
            c = i * m + j
            D[c, c] = 4.2
            for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= i + di < m and 0 <= j + dj < m:
                    D[c, (i + di) * m + j + dj] = -1.0 - 0.1 * di
    A = oracle.Matrix.from_dense(D)
    lu, _ = oracle.ilu0(A)
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(n):
        for k in range(A.rowPtr[i], A.rowPtr[i + 1]):
            j = A.colIdx[k]
            (L if j < i else U)[i, j] = lu[k]
    LU = L @ U
    mask = D != 0
    assert np.allclose(LU[mask], D[mask], atol=1e-13)
    assert not np.allclose(LU, D)  # fill outside the pattern is dropped


def test_ssor_identity_and_laplacian_cg():
    # tests/unit/solver/test_ssor_preconditioner.cpp:44-63 and :65-88
    I = oracle.Matrix.from_dense(np.eye(5))
    r = np.arange(1.0, 6.0)
    assert np.allclose(oracle.ssor_apply(I, r, 1.0), r)
    n = 50
    D = tridiag(n, -1.0, 2.0, -1.0)
    b = np.ones(n)
    x, it, last, _ = oracle.cg(oracle.Matrix.from_dense(D), b, oracle.PC_SSOR, 1e-10, 500)
    assert np.linalg.norm(b - D @ x) < 1e-7


def _ssor_closed_form(Dm, r, omega, d_mid=None):
    """z = w(2-w) (D+wU)^-1 D_mid (D+wL)^-1 r: the inverse of the reference's
    M = (D+wL) D^-1 (D+wU) / (w(2-w)) (src/solver/LinearSolvers.cpp:160-161),
    by dense triangular solves.  D is the diagonal used by the two sweeps and
    d_mid the one the middle scaling multiplies by (they differ only on rows
    whose stored diagonal is zero)."""
    D = np.diag(np.diag(Dm))
    Dmid = D if d_mid is None else np.diag(d_mid)
    L, U = np.tril(Dm, -1), np.triu(Dm, 1)
    y = np.linalg.solve(D + omega * L, r)
    return omega * (2.0 - omega) * np.linalg.solve(D + omega * U, Dmid @ y)


@pytest.mark.parametrize("omega", [1.0, 0.6, 1.45])
def test_ssor_matches_closed_form(omega):
    """orc_ssor_apply against the closed form of the SSOR inverse on a
    nonsymmetric matrix with a strongly non-uniform diagonal (1..50): a dropped
    D scaling, a swapped sweep or a wrong (2-w) / w factor all change z."""
    rng = np.random.default_rng(int(omega * 100))
    n = 60
    Dm = np.where(rng.random((n, n)) < 0.12, rng.uniform(-1, 1, (n, n)), 0.0)
    np.fill_diagonal(Dm, 10.0 ** rng.uniform(0, 1.7, n))
    r = rng.standard_normal(n)
    z = oracle.ssor_apply(oracle.Matrix.from_dense(Dm), r, omega)
    ze = _ssor_closed_form(Dm, r, omega)
    assert np.linalg.norm(z - ze) <= 1e-13 * np.linalg.norm(ze)
    # the same holds for the defining relation M z = r
    D = np.diag(np.diag(Dm))
    M = (D + omega * np.tril(Dm, -1)) @ np.linalg.inv(D) @ (D + omega * np.triu(Dm, 1))
    assert np.linalg.norm(M @ z / (omega * (2 - omega)) - r) <= 1e-12 * np.linalg.norm(r)


def test_ssor_missing_and_zero_diagonal():
    """Rows without a stored diagonal sweep with 1 in its place and scale by 1;
    a stored zero diagonal sweeps with 1 but scales by 0 (diag_ keeps the
    stored value, invDiag_ falls back to 1: src/solver/LinearSolvers.cpp:174-184)."""
    rng = np.random.default_rng(7)
    n, omega = 30, 1.3
    Dm = np.where(rng.random((n, n)) < 0.15, rng.uniform(-0.3, 0.3, (n, n)), 0.0)
    np.fill_diagonal(Dm, rng.uniform(2, 9, n))
    rp, ci, va = [0], [], []
    missing, zero = 4, 11
    for i in range(n):
        for j in range(n):
            if i == j and i == missing:
                continue
            if i == j or Dm[i, j] != 0.0:
                ci.append(j)
                va.append(0.0 if i == j == zero else Dm[i, j])
        rp.append(len(ci))
    A = oracle.Matrix(np.array(rp, np.int32), np.array(ci, np.int32), np.array(va))
    r = rng.standard_normal(n)
    z = oracle.ssor_apply(A, r, omega)
    Ds = Dm.copy()
    Ds[missing, missing] = 1.0
    Ds[zero, zero] = 1.0
    d_mid = np.diag(Ds).copy()
    d_mid[zero] = 0.0
    ze = _ssor_closed_form(Ds, r, omega, d_mid)
    assert np.linalg.norm(z - ze) <= 1e-13 * np.linalg.norm(ze)


@pytest.mark.parametrize("pc", [oracle.PC_JACOBI, oracle.PC_ILU])
def test_gmres_matches_dense_solve(pc):
    D = random_sparse(50, 0.1, 3, diag_boost=0.5)
    b = np.random.default_rng(2).standard_normal(50)
    x, it, last = oracle.gmres(oracle.Matrix.from_dense(D), b, pc, 1e-12, 500, 30)
    assert np.allclose(x, np.linalg.solve(D, b), rtol=1e-8, atol=1e-9)


def test_tfqmr_reference_known_answers():
    # tests/unit/solver/test_tfqmr_solver.cpp:90-111: upwind convection-diffusion,
    # Pe = 5, ILU (exact on a tridiagonal: one iteration)
    n, Pe = 40, 5.0
    C = tridiag(n, -1.0 - Pe, 2.0 + Pe, -1.0)
    x, it, last = oracle.tfqmr(oracle.Matrix.from_dense(C), np.ones(n), oracle.PC_ILU, 1e-8, 500)
    assert it == 1 and np.linalg.norm(C @ x - 1.0) < 1e-12


def test_tfqmr_restatement_keeps_the_reference_direction_update():
    """TfqmrSolver builds the next half-step vector from the CGS direction
    (u1 = M^-1 r + beta (u2 + beta u1), LinearSolvers.cpp:592-593) where
    TFQMR takes u = r + beta q.  The iteration is then exact only with an
    exact preconditioner, and the reference's own SPD-Laplacian case
    (test_tfqmr_solver.cpp:66-88, Jacobi, ||r|| < 1e-5) does not converge:
    the restatement must show the same."""
    n = 30
    D = tridiag(n, -1.0, 2.0, -1.0)
    x, it, last = oracle.tfqmr(oracle.Matrix.from_dense(D), np.ones(n), oracle.PC_JACOBI, 1e-8, 500)
    assert it < 500 and last > 1e-2  # stopped by a breakdown, far from converged
    assert not np.isfinite(np.linalg.norm(D @ x - 1.0)) or np.linalg.norm(D @ x - 1.0) > 1e-5


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tfqmr_exact_preconditioner_solves_in_one_step(seed):
    # ILU(0) of a (nonsymmetric) tridiagonal matrix is its exact LU
    rng = np.random.default_rng(seed)
    n = 25
    C = tridiag(n, 0.0, 0.0, 0.0)
    C += np.diag(rng.uniform(3, 4, n)) + np.diag(rng.uniform(-1, 0, n - 1), -1) + np.diag(rng.uniform(-1, 0, n - 1), 1)
    b = rng.standard_normal(n)
    x, it, last = oracle.tfqmr(oracle.Matrix.from_dense(C), b, oracle.PC_ILU, 1e-10, 50)
    assert it == 1 and np.allclose(x, np.linalg.solve(C, b), rtol=1e-10, atol=1e-12)


def test_tfqmr_zero_rhs_and_iteration_limit():
    D = tridiag(20, -1.0, 3.0, -1.2)
    A = oracle.Matrix.from_dense(D)
    x, it, last = oracle.tfqmr(A, np.zeros(20))
    assert it == 0 and not np.any(x)


def test_none_preconditioner_acts_as_jacobi():
    # make_preconditioner's default branch (src/solver/LinearSolvers.cpp:250-252)
    D = random_sparse(30, 0.2, 5, diag_boost=1.0, symmetric=True)
    A = oracle.Matrix.from_dense(D)
    b = np.ones(30)
    a = oracle.cg(A, b, oracle.PC_NONE, 1e-10, 100)
    j = oracle.cg(A, b, oracle.PC_JACOBI, 1e-10, 100)
    assert a[1] == j[1] and np.array_equal(a[0], j[0])


def test_adjoint_gmres_transpose_known_answer():
    # tests/unit/longtail/test_longtail_modules.cpp:32-54: A = [[2,1],[0,3]]
    A = oracle.Matrix([0, 2, 3], [0, 1, 1], [2.0, 1.0, 3.0])
    psi, ok, its, res = oracle.gmres_transpose(A, [1.0, 1.0])
    assert ok
    assert abs(psi[0] - 0.5) < 1e-6 and abs(psi[1] - 1.0 / 6.0) < 1e-6
    # gradient = dJ/dalpha - (dR/dalpha)^T psi = 0.5 - 0.25
    assert abs((0.5 - 0.25) - 0.25) < 1e-6


def test_adjoint_gmres_transpose_random():
    D = random_sparse(40, 0.1, 11, diag_boost=0.5)
    b = np.random.default_rng(4).standard_normal(40)
    psi, ok, its, res = oracle.gmres_transpose(oracle.Matrix.from_dense(D), b, tol=1e-12)
    assert ok
    assert np.allclose(psi, np.linalg.solve(D.T, b), rtol=1e-8, atol=1e-9)
