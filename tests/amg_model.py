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
