"""numpy model of the device SA-AMG V-cycle, for tests.

Applies, with scipy sparse matrices, the same cycle the device runs on a
hierarchy obtained from ``capi.AmgHost``: Chebyshev smoothing of degree k on
[lmin, lmax] of D^-1 A (k = 1: lmin = lmax/2, damped Jacobi; `degree` on the
finest level, `coarse_degree` below it, as the solver's amg_sweeps and
amg_coarse_sweeps with their defaults 2 and 1), restriction by R = P^T,
prolongation by P, dense inverse on the coarsest level.
"""
import numpy as np


def chebyshev_coefficients(lmax, degree, ratio=0.3):
    lmin = lmax * (0.5 if degree == 1 else ratio)
    theta, delta = 0.5 * (lmax + lmin), 0.5 * (lmax - lmin)
    sigma = theta / delta
    alpha, beta = [1.0 / theta], [0.0]
    rho = 1.0 / sigma
    for _ in range(1, degree):
        rn = 1.0 / (2.0 * sigma - rho)
        beta.append(rn * rho)
        alpha.append(2.0 * rn / delta)
        rho = rn
    return alpha, beta


class VCycle:
    def __init__(self, host, degree=2, coarse_degree=1):
        self.levels = []
        self.degree = degree
        for l in range(host.levels):
            A = host.scipy(l, host.A)
            _, _, _, lam = host.info(l, host.A)
            last = l + 1 == host.levels
            P = None if last else host.scipy(l, host.P)
            R = None if last else host.scipy(l, host.R)
            d = A.diagonal()
            inv = np.where(d != 0, 1.0 / np.where(d != 0, d, 1.0), 1.0)
            k = degree if l == 0 else coarse_degree
            self.levels.append((A, P, R, inv, k, *chebyshev_coefficients(lam, k)))
        self.coarse_inverse = host.matrix(host.levels - 1, host.COARSE_INV)

    def __call__(self, b, level=0):
        A, P, R, inv, k, al, be = self.levels[level]
        if level == len(self.levels) - 1:
            return self.coarse_inverse @ b
        r = b.copy()
        d = al[0] * inv * b
        x = d.copy()
        for s in range(1, k):
            r -= A @ d
            d = be[s] * d + al[s] * inv * r
            x += d
        r -= A @ d
        e = P @ self(R @ r, level + 1)
        x += e
        r -= A @ e
        d = al[0] * inv * r
        x += d
        for s in range(1, k):
            r -= A @ d
            d = be[s] * d + al[s] * inv * r
            x += d
        return x


def pcg(A, b, M, tol=1e-8, maxit=1000):
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    bn = max(np.linalg.norm(b), 1e-30)
    for it in range(maxit):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x += alpha * p
        r -= alpha * Ap
        if np.linalg.norm(r) / bn < tol:
            return x, it + 1
        z = M(r)
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return x, maxit


def smoothed_prolongator(A, agg, nagg, strong, lmax, omega_scale=4.0 / 3.0, pattern=None):
    """P = (I - omega D_F^-1 A_F) T for fixed aggregates and strong couplings
    (every entry, no pattern restriction): T[i, agg[i]] = 1 for aggregated
    rows; A_F keeps A's strong off-diagonal entries and has the filtered
    diagonal D_F = diag(A) + the row's weak off-diagonal entries, taken as 0
    within 1e-10 |a_ii| of 0 (its couplings, all weak, cancel the diagonal:
    the row is not smoothed); omega = omega_scale / lmax.  `strong` is per
    entry of A in CSR order.  With `pattern` (the prolongator of a truncated
    setup), the rows are restricted to its entries and scaled to the
    complete rows' sums."""
    import scipy.sparse as sp
    A = A.tocsr()
    n = A.shape[0]
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    cols = A.indices
    diag = rows == cols
    d = np.bincount(rows[diag], weights=A.data[diag], minlength=n)
    weak = (~diag) & (np.asarray(strong) == 0)
    dF = d + np.bincount(rows[weak], weights=A.data[weak], minlength=n)
    dF = np.where(np.abs(dF) <= 1e-10 * np.abs(d), 0.0, dF)
    keep = (~diag) & (np.asarray(strong) != 0)
    AF = sp.csr_matrix((A.data[keep], (rows[keep], cols[keep])), shape=(n, n)) + sp.diags(dF)
    has = agg >= 0
    T = sp.csr_matrix((np.ones(has.sum()), (np.arange(n)[has], agg[has])), shape=(n, nagg))
    omega = omega_scale / lmax
    w = np.where(dF != 0, omega / np.where(dF != 0, dF, 1.0), 0.0)
    P = (T - sp.diags(w) @ (AF @ T)).tocsr()
    if pattern is None:
        return P
    full = np.asarray(P.sum(axis=1)).ravel()
    pattern = pattern.tocsr()
    kept = restrict_to_pattern(P, pattern)
    Pk = sp.csr_matrix((kept, pattern.indices, pattern.indptr), shape=P.shape)
    ks = np.asarray(Pk.sum(axis=1)).ravel()
    scale = np.where(ks != 0, full / np.where(ks != 0, ks, 1.0), 1.0)
    return (sp.diags(scale) @ Pk).tocsr()


def restrict_to_pattern(M, pattern):
    """M's values at the (row, column) positions of the sparse matrix
    `pattern`, in its CSR order (positions M does not hold give 0)."""
    M = M.tocsr()
    n = pattern.shape[0]
    rows = np.repeat(np.arange(n), np.diff(pattern.indptr))
    return np.asarray(M[rows, pattern.indices]).ravel()


def lambda_start(n):
    """The power iteration's start vector (amg_lambda_start): xorshift64."""
    out = np.empty(n)
    st = 0x9E3779B97F4A7C15
    mask = (1 << 64) - 1
    for i in range(n):
        st ^= (st << 13) & mask
        st ^= st >> 7
        st ^= (st << 17) & mask
        out[i] = float(st >> 11) * (1.0 / 9007199254740992.0) - 0.5
    return out


def lambda_max_estimate(A, iters=20):
    """amg_lambda_max: power iteration on D^-1/2 A D^-1/2 from lambda_start,
    the last Rayleigh quotient (without the 1.05 margin)."""
    A = A.tocsr()
    d = A.diagonal()
    s = np.where(d != 0, 1.0 / np.sqrt(np.abs(np.where(d != 0, d, 1.0))), 1.0)
    x = lambda_start(A.shape[0])
    lam = 1.0
    for _ in range(iters):
        y = s * (A @ (s * x))
        xy, xx, yy = x @ y, x @ x, y @ y
        lam = abs(xy / xx) if xx > 0 else 1.0
        x = y * (1.0 / np.sqrt(yy) if yy > 0 else 1.0)
    return lam
