"""Invariants of the SA-AMG hierarchy and the RCM ordering (CPU only).

The north-star SA-AMG has no reference counterpart (the reference's AMG is
Ruge-Stueben, src/solver/AMGPreconditioner.cpp); it is pinned by what the
construction must satisfy: R = P^T, A_c = P^T A P (scipy product), P 1 = 1 for
an operator with A 1 = 0, aggregates covering every row, the coarse inverse,
symmetry of the cycle, and h-robust PCG convergence.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from simall_amd import capi
from simall_amd import workload as wl

from amg_model import VCycle, pcg


def system(N, pin=True, permuted=True):
    w = wl.poisson3d(N)
    m = w.mesh
    rp, ci = oracle.build_sparsity(m)
    P = oracle.Matrix(rp, ci, np.zeros(ci.size))
    vals, _, _ = oracle.assemble_pressure(m, P, w.bc_of_zone, gamma=w.gamma)
    A = sp.csr_matrix((vals, ci, rp), shape=(m.n_cells, m.n_cells))
    if not pin:
        A = A.tolil()
        A[0, 0] = A[0, 0] / 1e6
        A = A.tocsr()
    b = w.b.copy()
    b[0] = 0.0
    return A, b


@pytest.fixture(scope="module")
def hier():
    A, b = system(16)
    return A, b, capi.AmgHost(A.indptr, A.indices, A.data)


def test_restriction_is_prolongation_transpose(hier):
    A, b, H = hier
    assert H.levels >= 2
    for l in range(H.levels - 1):
        P, R = H.scipy(l, H.P), H.scipy(l, H.R)
        assert (abs(P.T - R)).max() == 0.0


def test_galerkin_operator(hier):
    A, b, H = hier
    assert abs(H.scipy(0, H.A) - A).max() == 0.0
    for l in range(H.levels - 1):
        Al, P, Ac = H.scipy(l, H.A), H.scipy(l, H.P), H.scipy(l + 1, H.A)
        ref = (P.T @ Al @ P).toarray()
        assert np.allclose(Ac.toarray(), ref, rtol=1e-12, atol=1e-12 * abs(ref).max())


def test_coarse_inverse(hier):
    A, b, H = hier
    Ac = H.scipy(H.levels - 1, H.A).toarray()
    inv = H.matrix(H.levels - 1, H.COARSE_INV)
    assert np.allclose(inv @ Ac, np.eye(Ac.shape[0]), atol=1e-9)


def test_prolongator_preserves_constants_on_neumann_operator():
    # unpinned Neumann operator: A 1 = 0, so the smoothed prolongator keeps
    # the constant vector: P 1 = 1 (every row aggregated)
    A, _ = system(12, pin=False)
    assert abs(A @ np.ones(A.shape[0])).max() < 1e-12
    H = capi.AmgHost(A.indptr, A.indices, A.data)
    P = H.scipy(0, H.P)
    assert np.allclose(P @ np.ones(P.shape[1]), 1.0, atol=1e-12)
    # every fine row has a nonzero in P (aggregates cover all rows)
    assert (np.diff(P.indptr) > 0).all()


def test_lambda_max_estimate(hier):
    A, b, H = hier
    d = A.diagonal()
    S = sp.diags(1 / np.sqrt(d)) @ A @ sp.diags(1 / np.sqrt(d))
    lam = spla.eigsh(S, k=1, which="LA", return_eigenvectors=False)[0]
    est = H.info(0, H.A)[3]
    assert 0.95 * lam <= est <= 1.2 * lam


def test_vcycle_symmetric_and_spd(hier):
    A, b, H = hier
    M = VCycle(H, degree=2)
    rng = np.random.default_rng(0)
    u, v = rng.standard_normal((2, A.shape[0]))
    assert abs(u @ M(v) - v @ M(u)) <= 1e-12 * abs(u @ M(u))
    assert u @ M(u) > 0


@pytest.mark.parametrize("N", [12, 24])
def test_amg_pcg_converges_fast(N):
    A, b = system(N)
    H = capi.AmgHost(A.indptr, A.indices, A.data)
    x, its = pcg(A, b, VCycle(H, 2), tol=1e-8)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-8
    # Jacobi-PCG needs hundreds of iterations here
    assert its <= 30


def test_rcm_permutation():
    A, _ = system(10)
    perm, bw = capi.rcm_permutation(A.indptr, A.indices)
    assert np.array_equal(np.sort(perm), np.arange(A.shape[0]))
    B = A[perm][:, perm].tocoo()
    assert np.abs(B.row - B.col).max() == bw
    # a permuted 10^3 grid has bandwidth ~ n; RCM brings it near a level set
    Ac = A.tocoo()
    assert np.abs(Ac.row - Ac.col).max() > 5 * bw
    assert bw <= 3 * 10 * 10


def test_color_permutation():
    """Multicolour ordering of the ILU(0) solvers: a 7-point grid takes two
    colours, rows of one colour are never coupled, and each colour keeps the
    reverse Cuthill-McKee order; a 3-cycle needs three colours."""
    A, _ = system(10)
    perm, nc = capi.color_permutation(A.indptr, A.indices)
    assert nc == 2
    assert np.array_equal(np.sort(perm), np.arange(A.shape[0]))
    B = A[perm][:, perm].tocoo()
    off = B.row != B.col
    n0 = int(B.col[off & (B.col < B.row)].max()) + 1  # last first-colour row + 1
    assert np.all((B.row[off] < n0) != (B.col[off] < n0))
    rcm, _ = capi.rcm_permutation(A.indptr, A.indices)
    pos = np.empty_like(rcm)
    pos[rcm] = np.arange(rcm.size)
    assert np.all(np.diff(pos[perm[:n0]]) > 0) and np.all(np.diff(pos[perm[n0:]]) > 0)
    tri = np.array([0, 3, 6, 9], np.int32), np.array([0, 1, 2] * 3, np.int32)
    assert capi.color_permutation(*tri)[1] == 3


def test_host_csr_input_is_validated():
    """Malformed patterns are rejected before any host algorithm walks them."""
    A, _ = system(4)
    rp, ci = A.indptr.astype(np.int32), A.indices.astype(np.int32)
    bad_col = ci.copy()
    bad_col[5] = -1
    bad_col2 = ci.copy()
    bad_col2[7] = A.shape[0]
    bad_rp = rp.copy()
    bad_rp[3], bad_rp[4] = bad_rp[4], bad_rp[3] - 1
    for r, c in ((rp, bad_col), (rp, bad_col2), (bad_rp, ci)):
        with pytest.raises(capi.MiError):
            capi.rcm_permutation(r, c)
        with pytest.raises(capi.MiError):
            capi.AmgHost(r, c, np.ones(c.size))


def test_numeric_phase_model_reproduces_the_host_setup():
    """The numpy statement of the numeric phase (the smoothed prolongator from
    the recorded aggregates and strong couplings, lambda_max by the same power
    iteration, A_c = R A P) reproduces every level of a host build; the device
    re-setup is checked against this model on new values."""
    from amg_model import smoothed_prolongator, restrict_to_pattern, lambda_max_estimate
    A, b = system(16)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 64
    cfg.amg_truncation = 0.0  # complete rows (test_truncated_prolongator: truncated)
    H = capi.AmgHost(A.indptr, A.indices, A.data, cfg=cfg)
    assert H.levels >= 3
    for l in range(H.levels - 1):
        Al, P = H.scipy(l, H.A), H.scipy(l, H.P)
        agg, nagg = H.aggregates(l)
        strong = H.strong(l)
        assert agg.size == Al.shape[0] and strong.size == Al.nnz
        assert nagg == P.shape[1] and agg.max() == nagg - 1 and agg.min() >= -1
        lmax = H.info(l, H.A)[3]
        lam = lambda_max_estimate(Al)
        assert abs(lam * 1.05 - lmax) <= 1e-10 * lmax
        Pm = smoothed_prolongator(Al, agg, nagg, strong, lmax)
        vals = restrict_to_pattern(Pm, P)
        assert np.allclose(vals, P.data, rtol=1e-13, atol=1e-15)
        # the host dropped only exact zeros
        outside = Pm - sp.csr_matrix((vals, P.indices, P.indptr), shape=P.shape)
        assert abs(outside).max() <= 1e-15 * abs(P).max()
        Ac = H.scipy(l + 1, H.A)
        ref = restrict_to_pattern(P.T @ Al @ P, Ac)
        assert np.allclose(ref, Ac.data, rtol=1e-12, atol=1e-13 * abs(Ac.data).max())


def test_rows_with_only_weak_couplings_are_not_smoothed():
    """A row whose off-diagonal couplings are all weak has a filtered diagonal
    d_F = a_ii + sum of them = its row sum, zero on the Neumann rows of the
    pressure operator: D_F^-1 A_F is undefined there and the row keeps its
    tentative entry 1 (not 1 - omega from a rounding residue of d_F)."""
    A, b = system(32)
    H = capi.AmgHost(A.indptr, A.indices, A.data, cfg=capi.default_cfg())
    Al, P = H.scipy(0, H.A), H.scipy(0, H.P)
    agg, _ = H.aggregates(0)
    strong = H.strong(0)
    rows = np.repeat(np.arange(Al.shape[0]), np.diff(Al.indptr))
    n_strong = np.bincount(rows, weights=strong.astype(float), minlength=Al.shape[0])
    rsum = np.asarray(Al.sum(axis=1)).ravel()
    isolated = np.flatnonzero((n_strong == 0) & (np.abs(rsum) <= 1e-12 * Al.diagonal()))
    assert isolated.size > 0
    for i in isolated:
        cols = P.indices[P.indptr[i]:P.indptr[i + 1]]
        assert list(cols) == [agg[i]] and P.data[P.indptr[i]] == 1.0, i


def test_truncated_prolongator():
    """amg_truncation = 0.1: every row of every prolongator keeps exactly the
    entries of the complete smoothed row whose magnitude reaches 0.1 x the
    row's largest, scaled so that the row sum is the complete row's; the
    Galerkin operators are R A P of the truncated prolongators; the numpy
    statement restricted to the kept entries reproduces them."""
    from amg_model import smoothed_prolongator, restrict_to_pattern
    A, b = system(16)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 64
    cfg.amg_truncation = 0.0
    full = capi.AmgHost(A.indptr, A.indices, A.data, cfg=cfg)
    cfg.amg_truncation = 0.1
    H = capi.AmgHost(A.indptr, A.indices, A.data, cfg=cfg)
    assert H.levels >= 3
    assert H.scipy(0, H.P).nnz < full.scipy(0, full.P).nnz
    assert H.scipy(1, H.A).nnz < full.scipy(1, full.A).nnz
    for l in range(H.levels - 1):
        Al, P = H.scipy(l, H.A), H.scipy(l, H.P)
        agg, nagg = H.aggregates(l)
        lmax = H.info(l, H.A)[3]
        Pc = smoothed_prolongator(Al, agg, nagg, H.strong(l), lmax)  # complete rows
        Pc.eliminate_zeros()
        rmax = np.maximum.reduceat(np.abs(Pc.data), Pc.indptr[:-1])
        rows = np.repeat(np.arange(Pc.shape[0]), np.diff(Pc.indptr))
        keep = np.abs(Pc.data) >= 0.1 * rmax[rows]
        K = sp.csr_matrix((Pc.data * keep, Pc.indices.copy(), Pc.indptr.copy()), shape=Pc.shape)
        K.eliminate_zeros()
        K.sort_indices()
        assert np.array_equal(K.indptr, P.indptr) and np.array_equal(K.indices, P.indices), l
        assert np.allclose(np.asarray(P.sum(axis=1)).ravel(), np.asarray(Pc.sum(axis=1)).ravel(),
                           rtol=1e-12, atol=1e-14)
        Pm = smoothed_prolongator(Al, agg, nagg, H.strong(l), lmax, pattern=P)
        assert np.allclose(restrict_to_pattern(Pm, P), P.data, rtol=1e-12, atol=1e-15)
        Ac = H.scipy(l + 1, H.A)
        ref = restrict_to_pattern(P.T @ Al @ P, Ac)
        assert np.allclose(ref, Ac.data, rtol=1e-12, atol=1e-13 * abs(Ac.data).max())


def test_strong_mask_marks_the_strength_test():
    """Entry a_ij (i != j) is strong iff |a_ij| >= theta sqrt(|a_ii a_jj|) and
    a_ij != 0 (theta 0.02), which the prolongator's filtering keeps."""
    A, b = system(12)
    cfg = capi.default_cfg()
    cfg.amg_coarse_size = 64
    H = capi.AmgHost(A.indptr, A.indices, A.data, cfg=cfg)
    strong = H.strong(0)
    A = A.tocsr()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    d = A.diagonal()
    a = np.abs(A.data)
    want = (rows != A.indices) & (a > 0) & (a >= 0.02 * np.sqrt(np.abs(d[rows] * d[A.indices])))
    assert np.array_equal(strong.astype(bool), want)
