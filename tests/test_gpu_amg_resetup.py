"""Device numeric re-setup of the SA-AMG hierarchy and the blocked dense
coarsest inverse.

When the bound matrix keeps its pattern, a setup for new values recomputes
the hierarchy on the device with the aggregates, strong couplings and
patterns of the last full setup (amg_resetup = 1, the default).  Checked here
against the numpy model of that numeric phase (tests/amg_model.py, itself
pinned to the host build in test_amg_host.py):
  - patterns unchanged, level-0 values = the new matrix;
  - P_l = (I - omega D_F^-1 A_F) T on P's pattern, rtol 1e-12 (the model's
    sparse products sum in another order); with the default truncation the
    rows of that pattern scaled to the complete rows' sums (the model with
    pattern=), without it (amg_truncation = 0) the complete rows;
  - R_l = P_l^T exactly; A_{l+1} = R_l A_l P_l, rtol 1e-12;
  - lambda_max by the same power iteration, rtol 1e-10 (dot products summed
    in another order);
  - the dense coarsest inverse: after a full setup identical (bit for bit)
    to the host dense_inverse of the same matrix; after a re-setup
    ||inv A - I|| small;
  - the V-cycle equals the numpy V-cycle of the downloaded hierarchy (1e-11),
    and the solves meet the FP64 true-residual tolerance and agree with a
    solver set up from scratch on the new values.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from simall_amd import capi
from simall_amd import workload as wl

from amg_model import VCycle, smoothed_prolongator, restrict_to_pattern, lambda_max_estimate

pytestmark = pytest.mark.gpu


def poisson(N):
    w = wl.poisson3d(N)
    m = w.mesh
    rp, ci = oracle.build_sparsity(m)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, _ = oracle.assemble_pressure(m, P, w.bc_of_zone, gamma=w.gamma)
    b = w.b.copy()
    b[0] = 0.0
    return sp.csr_matrix((vals, ci, rp), shape=(m.n_cells,) * 2), b


def perturbed(A, seed):
    """Same pattern, new values: every symmetric off-diagonal pair scaled by a
    factor in [0.5, 2], the diagonal keeping its excess over the row's
    off-diagonal sum (an M-matrix like the pressure operator)."""
    A = A.tocsr().copy()
    rng = np.random.default_rng(seed)
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices
    U = sp.triu(A, 1).tocoo()
    f = sp.csr_matrix((rng.uniform(0.5, 2.0, U.nnz), (U.row, U.col)), shape=A.shape)
    f = f + f.T
    off = rows != cols
    fac = np.asarray(f[rows[off], cols[off]]).ravel()
    d_excess = A.diagonal() + np.bincount(rows[off], weights=A.data[off], minlength=n)
    A.data[off] *= fac
    offsum = np.bincount(rows[off], weights=A.data[off], minlength=n)
    diag = ~off
    A.data[diag] = (d_excess - offsum)[rows[diag]] * rng.uniform(1.0, 1.2, n)[rows[diag]]
    return A


CFG = dict(kind=capi.MI_SOLVER_CG, preconditioner=capi.MI_PC_AMG, tolerance=1e-10,
           max_iterations=500, reorder=0, true_residual=1, amg_coarse_size=200)


@pytest.fixture(scope="module")
def case():
    A1, b = poisson(24)
    A2 = perturbed(A1, 7)
    return A1, A2, b


def test_full_setup_coarse_inverse_is_the_host_inverse(gpu, case):
    A1, A2, b = case
    s = capi.Solver(**CFG)
    s.setup(capi.Csr(A1.indptr, A1.indices, A1.data))
    x, st, _ = s.solve(b)
    assert st.converged
    Hd = capi.AmgHost.from_solver(s)
    cfg = capi.default_cfg()  # the solver's truncation default included
    cfg.amg_coarse_size = 200
    Hh = capi.AmgHost(A1.indptr, A1.indices, A1.data, cfg=cfg)
    assert Hd.levels == Hh.levels >= 3
    inv_d = Hd.matrix(Hd.levels - 1, Hd.COARSE_INV)
    inv_h = Hh.matrix(Hh.levels - 1, Hh.COARSE_INV)
    print(f"coarsest {inv_d.shape[0]} rows; blocked device inverse == host inverse:",
          np.array_equal(inv_d, inv_h))
    assert np.array_equal(inv_d, inv_h)


@pytest.mark.parametrize("precision,truncation", [(capi.MI_PRECISION_FP64, 0.1),
                                                  (capi.MI_PRECISION_MIXED, 0.1),
                                                  (capi.MI_PRECISION_FP64, 0.0)])
def test_resetup_matches_the_numeric_model(gpu, case, precision, truncation):
    A1, A2, b = case
    s = capi.Solver(**dict(CFG, precision=precision, amg_truncation=truncation))
    M = capi.Csr(A1.indptr, A1.indices, A1.data)
    s.setup(M)
    x1, st1, _ = s.solve(b)
    assert st1.converged
    H1 = capi.AmgHost.from_solver(s)
    M.update_values(A2.data)
    x2, st2, _ = s.solve(b)
    assert st2.converged and st2.true_relative_residual <= 1e-10
    H2 = capi.AmgHost.from_solver(s)
    assert H2.levels == H1.levels
    assert abs(H2.scipy(0, H2.A) - A2).max() == 0.0
    for l in range(H2.levels - 1):
        A_l, P, R = H2.scipy(l, H2.A), H2.scipy(l, H2.P), H2.scipy(l, H2.R)
        P1 = H1.scipy(l, H1.P)
        assert np.array_equal(P.indptr, P1.indptr) and np.array_equal(P.indices, P1.indices)
        agg, nagg = H1.aggregates(l)
        strong = H1.strong(l)
        lmax = H2.info(l, H2.A)[3]
        lam = lambda_max_estimate(A_l)
        assert abs(lam * 1.05 - lmax) <= 1e-10 * lmax, (l, lam * 1.05, lmax)
        Pm = restrict_to_pattern(smoothed_prolongator(A_l, agg, nagg, strong, lmax,
                                                      pattern=P if truncation > 0 else None), P)
        err_p = abs(Pm - P.data).max() / abs(P.data).max()
        assert err_p <= 1e-12, (l, err_p)
        assert abs(P.T - R).max() == 0.0
        Ac = H2.scipy(l + 1, H2.A)
        Ac1 = H1.scipy(l + 1, H1.A)
        assert np.array_equal(Ac.indptr, Ac1.indptr) and np.array_equal(Ac.indices, Ac1.indices)
        ref = restrict_to_pattern(R @ A_l @ P, Ac)
        err_c = abs(ref - Ac.data).max() / abs(Ac.data).max()
        print(f"level {l}: n={A_l.shape[0]} lmax {lmax:.6f} |P - model| {err_p:.1e} "
              f"|A_c - R A P| {err_c:.1e}")
        assert err_c <= 1e-12, (l, err_c)
    # the coarsest inverse of the recomputed operator
    C = H2.scipy(H2.levels - 1, H2.A).toarray()
    inv = H2.matrix(H2.levels - 1, H2.COARSE_INV)
    assert abs(inv @ C - np.eye(C.shape[0])).max() <= 1e-9
    # the values changed, so the hierarchy did
    assert abs(H2.scipy(1, H2.A) - H1.scipy(1, H1.A)).max() > 0.0
    if precision == capi.MI_PRECISION_FP64:
        r = np.random.default_rng(5).standard_normal(A2.shape[0])
        z = s.precond_apply(r)
        zm = VCycle(H2, 2)(r)
        assert np.linalg.norm(z - zm) <= 1e-11 * np.linalg.norm(zm)


def test_resetup_solution_agrees_with_a_full_setup(gpu, case):
    A1, A2, b = case
    kw = dict(CFG, tolerance=1e-12)
    s = capi.Solver(**kw)
    M = capi.Csr(A1.indptr, A1.indices, A1.data)
    s.setup(M)
    s.solve(b)
    M.update_values(A2.data)
    xr, str_, _ = s.solve(b)
    f = capi.Solver(**dict(kw, amg_resetup=0))
    M2 = capi.Csr(A1.indptr, A1.indices, A1.data)
    f.setup(M2)
    f.solve(b)
    M2.update_values(A2.data)
    xf, stf, _ = f.solve(b)
    assert str_.converged and stf.converged
    rel = np.linalg.norm(xr - xf) / np.linalg.norm(xf)
    print(f"re-setup {str_.iterations} its, full setup {stf.iterations} its, rel diff {rel:.1e}")
    assert rel <= 1e-9
    assert str_.iterations <= stf.iterations + 10
    # amg_resetup = 0 applies the hierarchy a host build gives for the new values
    Hf = capi.AmgHost.from_solver(f)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 200
    Hh = capi.AmgHost(A2.indptr, A2.indices, A2.data, cfg=cfg)
    assert Hf.levels == Hh.levels
    for l in range(Hh.levels):
        assert abs(Hf.scipy(l, Hf.A) - Hh.scipy(l, Hh.A)).max() == 0.0


def test_repeated_resetups_follow_the_values(gpu, case):
    """Three value sets in a row on one solver, and back to the first: the
    last hierarchy equals the re-setup of the first values (a re-setup keeps
    no state from the values before)."""
    A1, A2, b = case
    A3 = perturbed(A1, 11)
    s = capi.Solver(**CFG)
    M = capi.Csr(A1.indptr, A1.indices, A1.data)
    s.setup(M)
    s.solve(b)
    M.update_values(A2.data)
    s.solve(b)
    Ha = capi.AmgHost.from_solver(s)
    M.update_values(A3.data)
    x3, st3, _ = s.solve(b)
    assert st3.converged and st3.true_relative_residual <= 1e-10
    M.update_values(A2.data)
    s.solve(b)
    Hb = capi.AmgHost.from_solver(s)
    for l in range(Ha.levels):
        for which in (Ha.A, Ha.P, Ha.R) if l + 1 < Ha.levels else (Ha.A,):
            assert np.array_equal(Ha.matrix(l, which)[2], Hb.matrix(l, which)[2])
    assert np.array_equal(Ha.matrix(Ha.levels - 1, Ha.COARSE_INV),
                          Hb.matrix(Hb.levels - 1, Hb.COARSE_INV))
