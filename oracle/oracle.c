/*
 * oracle.c — CPU restatement of the SimAll pressure-correction solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP backend, never shipped or
 * measured as the product.  Compiled with gcc -O2 -ffp-contract=off so that
 * each multiply and add rounds separately, as the reference's serial loops do
 * on x86-64 without FMA.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ==================================================================== RNG */
uint64_t orc_splitmix64(uint64_t* s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_uniform(uint64_t* s)
{
    return (double)(orc_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

double orc_normal(uint64_t* s)
{
    const double u1 = 1.0 - orc_uniform(s); /* (0, 1] */
    const double u2 = orc_uniform(s);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

/* ==================================================================== CSR */
/* src/solver/CSRMatrix.hpp:25-35 */
void orc_spmv(const orc_csr* A, const double* x, double* y)
{
    for (int i = 0; i < A->n; ++i) {
        double s = 0.0;
        for (int k = A->rowPtr[i]; k < A->rowPtr[i + 1]; ++k)
            s += A->values[k] * x[A->colIdx[k]];
        y[i] = s;
    }
}

/* src/adjoint/DiscreteAdjoint.cpp:13-22 */
void orc_spmv_transpose(const orc_csr* A, const double* x, double* y)
{
    memset(y, 0, sizeof(double) * (size_t)A->n);
    for (int i = 0; i < A->n; ++i) {
        const double xi = x[i];
        for (int k = A->rowPtr[i]; k < A->rowPtr[i + 1]; ++k)
            y[A->colIdx[k]] += A->values[k] * xi;
    }
}

/* src/solver/LinearSolvers.cpp:25-48 */
double orc_dot(const double* a, const double* b, int64_t n)
{
    double s = 0;
    for (int64_t i = 0; i < n; ++i)
        s += a[i] * b[i];
    return s;
}

double orc_norm2(const double* a, int64_t n) { return sqrt(orc_dot(a, a, n)); }

void orc_axpy(double alpha, const double* x, double* y, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        y[i] += alpha * x[i];
}

void orc_scal(double a, double* x, int64_t n)
{
    for (int64_t i = 0; i < n; ++i)
        x[i] *= a;
}

/* ======================================================== preconditioners */
/* src/solver/LinearSolvers.cpp:54-65 */
void orc_jacobi_setup(const orc_csr* A, double* invDiag)
{
    for (int i = 0; i < A->n; ++i) {
        invDiag[i] = 1.0;
        for (int k = A->rowPtr[i]; k < A->rowPtr[i + 1]; ++k)
            if (A->colIdx[k] == i) {
                invDiag[i] = (A->values[k] != 0.0) ? 1.0 / A->values[k] : 1.0;
                break;
            }
    }
}

/* src/solver/LinearSolvers.cpp:81-126 */
void orc_ilu0_setup(const orc_csr* A, double* lu, int* diagPtr)
{
    const int n = A->n;
    const int* rp = A->rowPtr;
    const int* ci = A->colIdx;
    memcpy(lu, A->values, sizeof(double) * (size_t)rp[n]);
    for (int i = 0; i < n; ++i) {
        diagPtr[i] = -1;
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                diagPtr[i] = k;
                break;
            }
    }
    for (int i = 1; i < n; ++i) {
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int k = ci[kp];
            if (k >= i)
                break;
            if (diagPtr[k] < 0)
                continue;
            const double piv = lu[diagPtr[k]];
            if (piv == 0.0)
                continue;
            const double m = lu[kp] / piv;
            lu[kp] = m;
            for (int jp = diagPtr[k] + 1; jp < rp[k + 1]; ++jp) {
                const int j = ci[jp];
                for (int qp = kp + 1; qp < rp[i + 1]; ++qp) {
                    if (ci[qp] == j) {
                        lu[qp] -= m * lu[jp];
                        break;
                    }
                    if (ci[qp] > j)
                        break;
                }
            }
        }
    }
}

/* src/solver/LinearSolvers.cpp:127-151 */
void orc_ilu0_apply(const orc_csr* A, const double* lu, const int* diagPtr,
                    const double* r, double* z)
{
    const int n = A->n;
    const int* rp = A->rowPtr;
    const int* ci = A->colIdx;
    double* y = (double*)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i) {
        double s = r[i];
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int j = ci[kp];
            if (j >= i)
                break;
            s -= lu[kp] * y[j];
        }
        y[i] = s;
    }
    for (int i = 0; i < n; ++i)
        z[i] = 0.0;
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        const int kd = diagPtr[i];
        for (int kp = kd + 1; kp < rp[i + 1]; ++kp)
            s -= lu[kp] * z[ci[kp]];
        z[i] = (kd >= 0 && lu[kd] != 0.0) ? s / lu[kd] : s;
    }
    free(y);
}

/* src/solver/LinearSolvers.cpp:163-236 (setup :168-187, apply :188-230) */
void orc_ssor_apply(const orc_csr* A, const int* diagPtr, double omega,
                    const double* r, double* z)
{
    const int n = A->n;
    const int* rp = A->rowPtr;
    const int* ci = A->colIdx;
    const double* va = A->values;
    double* y = (double*)calloc((size_t)n, sizeof(double));
    double* v = (double*)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < n; ++i) {
        const double invD = (diagPtr[i] >= 0 && va[diagPtr[i]] != 0.0) ? 1.0 / va[diagPtr[i]] : 1.0;
        double s = omega * r[i];
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int j = ci[kp];
            if (j >= i)
                break;
            s -= omega * va[kp] * y[j];
        }
        y[i] = s * invD;
    }
    const double scale = 2.0 - omega;
    for (int i = 0; i < n; ++i) {
        const double d = diagPtr[i] >= 0 ? va[diagPtr[i]] : 1.0;
        v[i] = d * y[i] / omega;
    }
    for (int i = 0; i < n; ++i)
        z[i] = 0.0;
    for (int i = n - 1; i >= 0; --i) {
        const double invD = (diagPtr[i] >= 0 && va[diagPtr[i]] != 0.0) ? 1.0 / va[diagPtr[i]] : 1.0;
        double s = v[i];
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int j = ci[kp];
            if (j <= i)
                continue;
            s -= va[kp] * z[j];
        }
        z[i] = s * omega * invD;
    }
    for (int i = 0; i < n; ++i)
        z[i] *= scale;
    free(y);
    free(v);
}

/* A preconditioner instance, set up once per solve as the reference does
 * (src/solver/LinearSolvers.cpp:422, :362, :269).  make_preconditioner maps
 * None and unknown kinds to Jacobi (src/solver/LinearSolvers.cpp:240-254). */
typedef struct {
    int kind;
    const orc_csr* A;
    double* invDiag;
    double* lu;
    int* diagPtr;
} orc_pc;

static int pc_setup(orc_pc* P, const orc_csr* A, int kind)
{
    memset(P, 0, sizeof(*P));
    if (kind == ORC_PC_AMG)
        return -1; /* the reference RS-AMG is not restated here */
    if (kind != ORC_PC_ILU && kind != ORC_PC_SSOR)
        kind = ORC_PC_JACOBI;
    P->kind = kind;
    P->A = A;
    const int n = A->n;
    if (kind == ORC_PC_JACOBI) {
        P->invDiag = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        orc_jacobi_setup(A, P->invDiag);
    } else if (kind == ORC_PC_ILU) {
        P->lu = (double*)malloc(sizeof(double) * (size_t)(A->rowPtr[n] > 0 ? A->rowPtr[n] : 1));
        P->diagPtr = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        orc_ilu0_setup(A, P->lu, P->diagPtr);
    } else {
        P->diagPtr = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) {
            P->diagPtr[i] = -1;
            for (int k = A->rowPtr[i]; k < A->rowPtr[i + 1]; ++k)
                if (A->colIdx[k] == i) {
                    P->diagPtr[i] = k;
                    break;
                }
        }
    }
    return 0;
}

static void pc_apply(const orc_pc* P, const double* r, double* z)
{
    const int n = P->A->n;
    if (P->kind == ORC_PC_JACOBI) {
        for (int i = 0; i < n; ++i)
            z[i] = P->invDiag[i] * r[i];
    } else if (P->kind == ORC_PC_ILU) {
        orc_ilu0_apply(P->A, P->lu, P->diagPtr, r, z);
    } else {
        orc_ssor_apply(P->A, P->diagPtr, 1.0, r, z);
    }
}

static void pc_free(orc_pc* P)
{
    free(P->invDiag);
    free(P->lu);
    free(P->diagPtr);
}

static void* xcalloc(int n)
{
    return calloc((size_t)(n > 0 ? n : 1), sizeof(double));
}

/* ================================================================ Krylov */
/* CGSolver::solve, src/solver/LinearSolvers.cpp:420-452 */
int orc_cg_solve(const orc_csr* A, int pc, double tol, int maxIter,
                 const double* b, double* x, double* last_residual,
                 double* history, int* n_history)
{
    orc_pc P;
    if (pc_setup(&P, A, pc) != 0)
        return -1;
    const int n = A->n;
    double* r = (double*)xcalloc(n);
    double* z = (double*)xcalloc(n);
    double* p = (double*)xcalloc(n);
    double* Ap = (double*)xcalloc(n);
    int nh = 0;
    int ret = maxIter;
    double last = 0.0;
    for (int i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
    }
    pc_apply(&P, r, z);
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = orc_dot(r, z, n);
    double bnorm = orc_norm2(b, n);
    if (bnorm < 1e-30)
        bnorm = 1e-30;
    for (int it = 0; it < maxIter; ++it) {
        orc_spmv(A, p, Ap);
        const double pAp = orc_dot(p, Ap, n);
        if (pAp == 0) {
            last = sqrt(rz) / bnorm;
            ret = it;
            goto done;
        }
        const double alpha = rz / pAp;
        orc_axpy(alpha, p, x, n);
        orc_axpy(-alpha, Ap, r, n);
        const double rn = orc_norm2(r, n);
        last = rn / bnorm;
        if (history)
            history[nh] = last;
        ++nh;
        if (last < tol) {
            ret = it + 1;
            goto done;
        }
        pc_apply(&P, r, z);
        const double rzNew = orc_dot(r, z, n);
        const double beta = rzNew / rz;
        for (int i = 0; i < n; ++i)
            p[i] = z[i] + beta * p[i];
        rz = rzNew;
    }
done:
    if (last_residual)
        *last_residual = last;
    if (n_history)
        *n_history = nh;
    free(r);
    free(z);
    free(p);
    free(Ap);
    pc_free(&P);
    return ret;
}

/* BiCGSTABSolver::solve, src/solver/LinearSolvers.cpp:360-404 */
int orc_bicgstab_solve(const orc_csr* A, int pc, double tol, int maxIter,
                       const double* b, double* x, double* last_residual,
                       double* history, int* n_history)
{
    orc_pc P;
    if (pc_setup(&P, A, pc) != 0)
        return -1;
    const int n = A->n;
    double* r = (double*)xcalloc(n);
    double* rh = (double*)xcalloc(n);
    double* p = (double*)xcalloc(n);
    double* v = (double*)xcalloc(n);
    double* s = (double*)xcalloc(n);
    double* t = (double*)xcalloc(n);
    double* y = (double*)xcalloc(n);
    double* z = (double*)xcalloc(n);
    int nh = 0;
    int ret = maxIter;
    double last = 0.0;
    for (int i = 0; i < n; ++i)
        x[i] = 0.0;
    orc_spmv(A, x, v);
    for (int i = 0; i < n; ++i)
        r[i] = b[i] - v[i];
    memcpy(rh, r, sizeof(double) * (size_t)n);
    double rho = 1.0, alpha = 1.0, omega = 1.0;
    double bnorm = orc_norm2(b, n);
    if (bnorm < 1e-30)
        bnorm = 1e-30;
    for (int i = 0; i < n; ++i)
        v[i] = 0.0;
    for (int it = 0; it < maxIter; ++it) {
        const double rhoNew = orc_dot(rh, r, n);
        if (rhoNew == 0.0) {
            last = orc_norm2(r, n) / bnorm;
            ret = it;
            goto done;
        }
        const double beta = (rhoNew / rho) * (alpha / omega);
        for (int i = 0; i < n; ++i)
            p[i] = r[i] + beta * (p[i] - omega * v[i]);
        pc_apply(&P, p, y);
        orc_spmv(A, y, v);
        alpha = rhoNew / orc_dot(rh, v, n);
        for (int i = 0; i < n; ++i)
            s[i] = r[i] - alpha * v[i];
        if (orc_norm2(s, n) < tol * bnorm) {
            orc_axpy(alpha, y, x, n);
            last = orc_norm2(s, n) / bnorm;
            if (history)
                history[nh] = last;
            ++nh;
            ret = it + 1;
            goto done;
        }
        pc_apply(&P, s, z);
        orc_spmv(A, z, t);
        double tt = orc_dot(t, t, n);
        if (tt < 1e-30)
            tt = 1e-30;
        omega = orc_dot(t, s, n) / tt;
        for (int i = 0; i < n; ++i) {
            x[i] += alpha * y[i] + omega * z[i];
            r[i] = s[i] - omega * t[i];
        }
        rho = rhoNew;
        last = orc_norm2(r, n) / bnorm;
        if (history)
            history[nh] = last;
        ++nh;
        if (last < tol) {
            ret = it + 1;
            goto done;
        }
    }
done:
    if (last_residual)
        *last_residual = last;
    if (n_history)
        *n_history = nh;
    free(r);
    free(rh);
    free(p);
    free(v);
    free(s);
    free(t);
    free(y);
    free(z);
    pc_free(&P);
    return ret;
}

/* GMRESmSolver::solve, src/solver/LinearSolvers.cpp:267-342 */
/* TfqmrSolver::solve, src/solver/LinearSolvers.cpp:497-603: x0 = 0; the
 * double-step quasi-minimal residual loop with the residual bound
 * tau*sqrt(2k+3) and a true-residual confirmation.  Returns the iteration
 * count as the reference: it + 1 on convergence, it on a breakdown
 * (|rh.v| or |rho| < 1e-30), maxIter at the limit. */
int orc_tfqmr_solve(const orc_csr* A, int pc, double tol, int maxIter, const double* b,
                    double* x, double* last_residual)
{
    orc_pc P;
    if (pc_setup(&P, A, pc) != 0)
        return -1;
    const int n = A->n;
    double* r = (double*)xcalloc(n);
    double* rh = (double*)xcalloc(n);
    double* w = (double*)xcalloc(n);
    double* d = (double*)xcalloc(n);
    double* u1 = (double*)xcalloc(n);
    double* u2 = (double*)xcalloc(n);
    double* v = (double*)xcalloc(n);
    double* Au1 = (double*)xcalloc(n);
    double* Au2 = (double*)xcalloc(n);
    double* z = (double*)xcalloc(n);
    int ret = maxIter;
    double last = 0.0;
    for (int i = 0; i < n; ++i) {
        x[i] = 0.0;
        r[i] = b[i];
        rh[i] = r[i];
        w[i] = r[i];
    }
    pc_apply(&P, r, u1);
    orc_spmv(A, u1, v);
    memcpy(Au1, v, sizeof(double) * (size_t)n);
    double tau = orc_norm2(r, n);
    double bnorm = orc_norm2(b, n);
    if (bnorm < 1.0e-30)
        bnorm = 1.0e-30;
    if (tau / bnorm < tol) {
        last = tau / bnorm;
        ret = 0;
        goto done;
    }
    {
        double rho = orc_dot(rh, r, n);
        double theta = 0.0, eta = 0.0, alpha = 0.0;
        for (int it = 0; it < maxIter; ++it) {
            const double sigma = orc_dot(rh, v, n);
            if (fabs(sigma) < 1.0e-30) {
                last = tau / bnorm;
                ret = it;
                goto done;
            }
            alpha = rho / sigma;
            pc_apply(&P, v, z);
            for (int i = 0; i < n; ++i)
                u2[i] = u1[i] - alpha * z[i];
            orc_spmv(A, u2, Au2);
            for (int half = 0; half < 2; ++half) {
                const double* Au = half == 0 ? Au1 : Au2;
                const double* u = half == 0 ? u1 : u2;
                for (int i = 0; i < n; ++i)
                    w[i] -= alpha * Au[i];
                const double theta_new = orc_norm2(w, n) / (tau > 1.0e-30 ? tau : 1.0e-30);
                const double c = 1.0 / sqrt(1.0 + theta_new * theta_new);
                const double tau_new = tau * theta_new * c;
                const double eta_new = c * c * alpha;
                const double coef = theta * theta * eta / (alpha > 1.0e-30 ? alpha : 1.0e-30);
                for (int i = 0; i < n; ++i) {
                    d[i] = u[i] + coef * d[i];
                    x[i] += eta_new * d[i];
                }
                tau = tau_new;
                theta = theta_new;
                eta = eta_new;
            }
            const double resBound = tau * sqrt(2.0 * (it + 1) + 1.0);
            last = resBound / bnorm;
            if (last < tol) {
                orc_spmv(A, x, z);
                for (int i = 0; i < n; ++i)
                    z[i] = b[i] - z[i];
                last = orc_norm2(z, n) / bnorm;
                if (last < tol) {
                    ret = it + 1;
                    goto done;
                }
            }
            for (int i = 0; i < n; ++i)
                r[i] -= alpha * (Au1[i] + Au2[i]);
            const double rhoNew = orc_dot(rh, r, n);
            if (fabs(rho) < 1.0e-30) {
                last = tau / bnorm;
                ret = it;
                goto done;
            }
            const double beta = rhoNew / rho;
            pc_apply(&P, r, z);
            for (int i = 0; i < n; ++i)
                u1[i] = z[i] + beta * (u2[i] + beta * u1[i]);
            orc_spmv(A, u1, Au1);
            for (int i = 0; i < n; ++i)
                v[i] = Au1[i] + beta * (Au2[i] + beta * v[i]);
            rho = rhoNew;
        }
    }
done:
    if (last_residual)
        *last_residual = last;
    free(r);
    free(rh);
    free(w);
    free(d);
    free(u1);
    free(u2);
    free(v);
    free(Au1);
    free(Au2);
    free(z);
    pc_free(&P);
    return ret;
}

int orc_gmres_solve(const orc_csr* A, int pc, double tol, int maxIter, int restart,
                    const double* b, double* x, double* last_residual)
{
    orc_pc P;
    if (pc_setup(&P, A, pc) != 0)
        return -1;
    const int n = A->n;
    const int m = restart > 1 ? restart : 1;
    double* r = (double*)xcalloc(n);
    double* w = (double*)xcalloc(n);
    double* z = (double*)xcalloc(n);
    double* V = (double*)calloc((size_t)(m + 1) * (size_t)(n > 0 ? n : 1), sizeof(double));
    double* H = (double*)calloc((size_t)(m + 1) * (size_t)m, sizeof(double));
    double* cs = (double*)calloc((size_t)m, sizeof(double));
    double* sn = (double*)calloc((size_t)m, sizeof(double));
    double* s = (double*)calloc((size_t)(m + 1), sizeof(double));
    double* yv = (double*)calloc((size_t)(m + 1), sizeof(double));
#define HH(i, j) H[(size_t)(i) * (size_t)m + (size_t)(j)]
#define VV(i) (V + (size_t)(i) * (size_t)n)
    double last = 0.0;
    int ret;
    for (int i = 0; i < n; ++i)
        x[i] = 0.0;
    double bnorm = orc_norm2(b, n);
    if (bnorm < 1e-30)
        bnorm = 1e-30;
    int iter = 0;
    for (int outer = 0; outer < maxIter; ++outer) {
        orc_spmv(A, x, w);
        for (int i = 0; i < n; ++i)
            r[i] = b[i] - w[i];
        pc_apply(&P, r, z);
        const double beta = orc_norm2(z, n);
        if (beta < tol * bnorm) {
            last = beta / bnorm;
            ret = iter;
            goto done;
        }
        memcpy(VV(0), z, sizeof(double) * (size_t)n);
        orc_scal(1.0 / beta, VV(0), n);
        for (int i = 0; i <= m; ++i)
            s[i] = 0.0;
        s[0] = beta;
        int j = 0;
        for (j = 0; j < m && iter < maxIter; ++j, ++iter) {
            orc_spmv(A, VV(j), w);
            pc_apply(&P, w, z);
            for (int i = 0; i <= j; ++i) {
                HH(i, j) = orc_dot(VV(i), z, n);
                orc_axpy(-HH(i, j), VV(i), z, n);
            }
            HH(j + 1, j) = orc_norm2(z, n);
            if (HH(j + 1, j) != 0.0) {
                memcpy(VV(j + 1), z, sizeof(double) * (size_t)n);
                orc_scal(1.0 / HH(j + 1, j), VV(j + 1), n);
            }
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                HH(i, j) = t;
            }
            const double dd = hypot(HH(j, j), HH(j + 1, j));
            cs[j] = HH(j, j) / dd;
            sn[j] = HH(j + 1, j) / dd;
            HH(j, j) = dd;
            HH(j + 1, j) = 0.0;
            const double t = cs[j] * s[j];
            s[j + 1] = -sn[j] * s[j];
            s[j] = t;
            last = fabs(s[j + 1]) / bnorm;
            if (last < tol) {
                ++j;
                break;
            }
        }
        for (int i = j - 1; i >= 0; --i) {
            double sum = s[i];
            for (int k = i + 1; k < j; ++k)
                sum -= HH(i, k) * yv[k];
            yv[i] = sum / HH(i, i);
        }
        for (int i = 0; i < j; ++i)
            orc_axpy(yv[i], VV(i), x, n);
        if (last < tol) {
            ret = iter;
            goto done;
        }
    }
    ret = iter;
done:
#undef HH
#undef VV
    if (last_residual)
        *last_residual = last;
    free(r);
    free(w);
    free(z);
    free(V);
    free(H);
    free(cs);
    free(sn);
    free(s);
    free(yv);
    pc_free(&P);
    return ret;
}

/* gmres_transpose, src/adjoint/DiscreteAdjoint.cpp:39-51 (Jacobi), :54-146 */
int orc_gmres_transpose(const orc_csr* A, const double* b, double* x, int maxIter,
                        int restart, double tol, int* itersOut, double* resOut)
{
    const int n = A->n;
    const int m = restart;
    double* invD = (double*)xcalloc(n);
    double* r = (double*)xcalloc(n);
    double* w = (double*)xcalloc(n);
    double* V = (double*)calloc((size_t)(m + 1) * (size_t)(n > 0 ? n : 1), sizeof(double));
    double* H = (double*)calloc((size_t)(m + 1) * (size_t)(m > 0 ? m : 1), sizeof(double));
    double* g = (double*)calloc((size_t)(m + 1), sizeof(double));
    double* cs = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    double* sn = (double*)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    double* y = (double*)calloc((size_t)(m + 1), sizeof(double));
#define HH(i, j) H[(size_t)(i) * (size_t)m + (size_t)(j)]
#define VV(i) (V + (size_t)(i) * (size_t)n)
    int ok = 0;
    double res = 0.0;
    for (int i = 0; i < n; ++i) {
        x[i] = 0.0;
        invD[i] = 1.0;
        for (int k = A->rowPtr[i]; k < A->rowPtr[i + 1]; ++k)
            if (A->colIdx[k] == i) {
                const double d = A->values[k];
                invD[i] = (fabs(d) > 1e-300) ? 1.0 / d : 1.0;
                break;
            }
    }
    double bnorm = orc_norm2(b, n);
    if (bnorm < 1e-300)
        bnorm = 1e-300;
    int iters = 0;
    while (iters < maxIter) {
        orc_spmv_transpose(A, x, r);
        for (int i = 0; i < n; ++i)
            r[i] = invD[i] * (b[i] - r[i]);
        const double beta = orc_norm2(r, n);
        res = beta;
        if (beta / bnorm < tol) {
            ok = 1;
            goto done;
        }
        memset(V, 0, sizeof(double) * (size_t)(m + 1) * (size_t)n);
        memset(H, 0, sizeof(double) * (size_t)(m + 1) * (size_t)m);
        for (int i = 0; i <= m; ++i)
            g[i] = 0.0;
        for (int i = 0; i < m; ++i)
            cs[i] = sn[i] = 0.0;
        g[0] = beta;
        for (int i = 0; i < n; ++i)
            VV(0)[i] = r[i] / beta;
        int inner = 0;
        for (int j = 0; j < m; ++j) {
            orc_spmv_transpose(A, VV(j), w);
            for (int i = 0; i < n; ++i)
                w[i] *= invD[i];
            for (int i = 0; i <= j; ++i) {
                HH(i, j) = orc_dot(w, VV(i), n);
                for (int k = 0; k < n; ++k)
                    w[k] -= HH(i, j) * VV(i)[k];
            }
            HH(j + 1, j) = orc_norm2(w, n);
            if (HH(j + 1, j) > 1e-300)
                for (int k = 0; k < n; ++k)
                    VV(j + 1)[k] = w[k] / HH(j + 1, j);
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
                HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
                HH(i, j) = t;
            }
            const double denom = hypot(HH(j, j), HH(j + 1, j));
            cs[j] = HH(j, j) / denom;
            sn[j] = HH(j + 1, j) / denom;
            HH(j, j) = denom;
            HH(j + 1, j) = 0.0;
            const double t = cs[j] * g[j];
            g[j + 1] = -sn[j] * g[j];
            g[j] = t;
            inner = j + 1;
            ++iters;
            if (fabs(g[j + 1]) / bnorm < tol)
                break;
            if (iters >= maxIter)
                break;
        }
        for (int i = inner - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < inner; ++k)
                s -= HH(i, k) * y[k];
            y[i] = (fabs(HH(i, i)) > 1e-300) ? s / HH(i, i) : 0.0;
        }
        for (int k = 0; k < inner; ++k)
            for (int i = 0; i < n; ++i)
                x[i] += y[k] * VV(k)[i];
        res = fabs(g[inner]);
        if (res / bnorm < tol) {
            ok = 1;
            goto done;
        }
    }
    ok = (res / (bnorm > 1e-300 ? bnorm : 1e-300)) < tol;
done:
#undef HH
#undef VV
    if (itersOut)
        *itersOut = iters;
    if (resOut)
        *resOut = res;
    free(invD);
    free(r);
    free(w);
    free(V);
    free(H);
    free(g);
    free(cs);
    free(sn);
    free(y);
    return ok;
}

/* ============================================================== assembly */
static int zone_bc(const int* bcOfZone, int nZones, uint32_t zone)
{
    /* A zone with no BoundarySpec: find_bc returns null and every face loop
     * takes its fall-through branch (src/solver/SimpleAlgorithm.cpp:189-197). */
    if (!bcOfZone || (int64_t)zone >= (int64_t)nZones)
        return ORC_BC_OUTFLOW;
    return bcOfZone[zone];
}

static int cmp_int(const void* a, const void* b)
{
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* SimpleAlgorithm::build_periodic_pairs, src/solver/SimpleAlgorithm.cpp:101-186:
 * periodic zones (ascending zone id, the order of a BoundarySpec list in zone
 * order) paired two by two; translation t = mean centroid of the second
 * zone's faces - mean of the first's; every face fa of the first zone is
 * paired with the first face fb of the second zone (face order) of least
 * |c_fb - (c_fa + t)|^2, and twin[fb] = fa (a later fa overwrites). */
void orc_periodic_twins(const orc_mesh* m, const int* bcOfZone, int nZones, int* twin)
{
    for (int64_t f = 0; f < m->n_faces; ++f)
        twin[f] = -1;
    int zones[64];
    int nz = 0;
    for (int z = 0; z < nZones && nz < 64; ++z)
        if (bcOfZone[z] == ORC_BC_PERIODIC)
            zones[nz++] = z;
    if (nz < 2)
        return;
    int64_t* list = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m->n_faces > 0 ? m->n_faces : 1));
    int64_t* listB = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m->n_faces > 0 ? m->n_faces : 1));
    for (int zi = 0; zi + 1 < nz; zi += 2) {
        int64_t na = 0, nb = 0;
        for (int64_t f = 0; f < m->n_faces; ++f) {
            if (m->neighbor[f] != UINT64_MAX)
                continue;
            if ((int)m->boundaryZone[f] == zones[zi])
                list[na++] = f;
            else if ((int)m->boundaryZone[f] == zones[zi + 1])
                listB[nb++] = f;
        }
        if (na == 0 || nb == 0)
            continue;
        double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
        for (int64_t i = 0; i < na; ++i) {
            ax += m->faceCx[list[i]];
            ay += m->faceCy[list[i]];
            az += m->faceCz[list[i]];
        }
        for (int64_t i = 0; i < nb; ++i) {
            bx += m->faceCx[listB[i]];
            by += m->faceCy[listB[i]];
            bz += m->faceCz[listB[i]];
        }
        const double ia = 1.0 / (double)na, ib = 1.0 / (double)nb;
        const double tx = bx * ib - ax * ia, ty = by * ib - ay * ia, tz = bz * ib - az * ia;
        for (int64_t i = 0; i < na; ++i) {
            const int64_t fa = list[i];
            int64_t best = -1;
            double bestd2 = DBL_MAX;
            const double px = m->faceCx[fa] + tx, py = m->faceCy[fa] + ty,
                         pz = m->faceCz[fa] + tz;
            for (int64_t j = 0; j < nb; ++j) {
                const int64_t fb = listB[j];
                const double d2 = (m->faceCx[fb] - px) * (m->faceCx[fb] - px)
                                  + (m->faceCy[fb] - py) * (m->faceCy[fb] - py)
                                  + (m->faceCz[fb] - pz) * (m->faceCz[fb] - pz);
                if (d2 < bestd2) {
                    bestd2 = d2;
                    best = fb;
                }
            }
            if (best >= 0) {
                twin[fa] = (int)best;
                twin[best] = (int)fa;
            }
        }
    }
    free(list);
    free(listB);
}

/* The cell across boundary face f's periodic twin, or UINT64_MAX. */
static uint64_t twin_cell(const orc_mesh* m, const int* twin, int64_t f)
{
    if (!twin || twin[f] < 0)
        return UINT64_MAX;
    return m->owner[twin[f]];
}

/* SimpleAlgorithm::build_sparsity, src/solver/SimpleAlgorithm.cpp:49-98 */
int64_t orc_build_sparsity(const orc_mesh* m, int* rowPtr, int* colIdx)
{
    return orc_build_sparsity_p(m, NULL, rowPtr, colIdx);
}

int64_t orc_build_sparsity_p(const orc_mesh* m, const int* twin, int* rowPtr, int* colIdx)
{
    const int64_t nC = m->n_cells;
    int buf[64];
    int64_t nnz = 0;
    if (rowPtr)
        rowPtr[0] = 0;
    for (int64_t c = 0; c < nC; ++c) {
        int cnt = 0;
        buf[cnt++] = (int)c;
        for (int k = m->cellFaceOffsets[c]; k < m->cellFaceOffsets[c + 1]; ++k) {
            const uint64_t f = m->cellFaceIndices[k];
            const uint64_t oth = (m->owner[f] == (uint64_t)c) ? m->neighbor[f] : m->owner[f];
            if (oth != UINT64_MAX && cnt < 64)
                buf[cnt++] = (int)oth;
            else if (oth == UINT64_MAX && cnt < 64 && twin_cell(m, twin, (int64_t)f) != UINT64_MAX)
                buf[cnt++] = (int)twin_cell(m, twin, (int64_t)f);
        }
        qsort(buf, (size_t)cnt, sizeof(int), cmp_int);
        int u = 0;
        for (int i = 0; i < cnt; ++i)
            if (u == 0 || buf[i] != buf[u - 1])
                buf[u++] = buf[i];
        if (colIdx)
            for (int i = 0; i < u; ++i)
                colIdx[nnz + i] = buf[i];
        nnz += u;
        if (rowPtr)
            rowPtr[c + 1] = (int)nnz;
    }
    return nnz;
}

/* find_col / add_at, src/solver/SimpleAlgorithm.cpp:23-35 */
static int find_col(const orc_csr* A, int row, int col)
{
    for (int k = A->rowPtr[row]; k < A->rowPtr[row + 1]; ++k)
        if (A->colIdx[k] == col)
            return k;
    return -1;
}

static void add_at(const orc_csr* A, double* values, int row, int col, double v)
{
    const int k = find_col(A, row, col);
    if (k >= 0)
        values[k] += v;
}

/* SimpleAlgorithm::assemble_pressure_correction, SIMPLE branch,
 * src/solver/SimpleAlgorithm.cpp:358-447. */
int orc_assemble_pressure(const orc_mesh* m, const int* bcOfZone, int nZones,
                          double rho, const double* aP, const double* gamma,
                          const double* phi, const orc_csr* pat,
                          double* values, double* rhs)
{
    return orc_assemble_pressure_p(m, bcOfZone, nZones, NULL, rho, aP, gamma, phi, pat, values,
                                   rhs);
}

/* The SIMPLEC denominator of every cell (src/solver/SimpleAlgorithm.cpp:378-387,
 * 400-409): max(aP_c - sum over row c's off-diagonal entries of |a|, 1e-30),
 * from the momentum matrix on the pattern. */
void orc_simplec_denominators(const orc_csr* amom, const double* aP, double* den)
{
    for (int64_t c = 0; c < amom->n; ++c) {
        double sumOff = 0.0;
        for (int k = amom->rowPtr[c]; k < amom->rowPtr[c + 1]; ++k)
            if (amom->colIdx[k] != (int)c)
                sumOff += fabs(amom->values[k]);
        const double d = aP[c] - sumOff;
        den[c] = d > 1e-30 ? d : 1e-30;
    }
}

int orc_assemble_pressure_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                            double rho, const double* aP, const double* gamma,
                            const double* phi, const orc_csr* pat,
                            double* values, double* rhs)
{
    const int64_t nC = m->n_cells;
    const int64_t nnz = pat->rowPtr[nC];
    int anyDirichlet = 0;
    for (int64_t k = 0; k < nnz; ++k)
        values[k] = 0.0;
    for (int64_t c = 0; c < nC; ++c)
        rhs[c] = 0.0;
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f];
        const uint64_t nb = m->neighbor[f];
        const double Ax = m->areaX[f], Ay = m->areaY[f], Az = m->areaZ[f];
        const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
        const double g = gamma ? gamma[f] : 1.0;
        if (nb != UINT64_MAX) {
            const double dx = m->cellCx[nb] - m->cellCx[o];
            const double dy = m->cellCy[nb] - m->cellCy[o];
            const double dz = m->cellCz[nb] - m->cellCz[o];
            const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
            double aPf = 1.0;
            if (aP) {
                const double dO = aP[o] > 1e-30 ? aP[o] : 1e-30;
                const double dN = aP[nb] > 1e-30 ? aP[nb] : 1e-30;
                aPf = 0.5 * (dO + dN);
            }
            const double den = aPf * ddotA;
            const double coef = rho * g * Amag2 / (den > 1e-30 ? den : 1e-30);
            add_at(pat, values, (int)o, (int)o, coef);
            add_at(pat, values, (int)o, (int)nb, -coef);
            add_at(pat, values, (int)nb, (int)nb, coef);
            add_at(pat, values, (int)nb, (int)o, -coef);
        } else {
            const int bc = zone_bc(bcOfZone, nZones, m->boundaryZone[f]);
            const uint64_t tw = twin_cell(m, twin, f);
            if (tw != UINT64_MAX || bc == ORC_BC_PRESSURE) {
                /* periodic (:411-419) or Dirichlet p' = 0 (:420-431) */
                const double dx = m->faceCx[f] - m->cellCx[o];
                const double dy = m->faceCy[f] - m->cellCy[o];
                const double dz = m->faceCz[f] - m->cellCz[o];
                const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
                const double dO = aP ? (aP[o] > 1e-30 ? aP[o] : 1e-30) : 1.0;
                const double den = dO * ddotA;
                const double coef = rho * g * Amag2 / (den > 1e-30 ? den : 1e-30);
                add_at(pat, values, (int)o, (int)o, coef);
                if (tw != UINT64_MAX)
                    add_at(pat, values, (int)o, (int)tw, -coef);
                else
                    anyDirichlet = 1;
            }
        }
        if (phi) {
            rhs[o] -= phi[f];
            if (nb != UINT64_MAX)
                rhs[nb] += phi[f];
        }
    }
    if (!anyDirichlet && nC > 0) {
        const int kd = find_col(pat, 0, 0);
        if (kd >= 0)
            values[kd] *= 1.0e6;
        rhs[0] = 0.0;
        return 1;
    }
    return 0;
}

/* Gradient step of the discrete adjoint for the pressure matrix:
 * grad[f] = -lambda^T (dA/dgamma_f) x, the term -(dR/dalpha)^T psi of
 * adjoint::solve_discrete_adjoint (src/adjoint/DiscreteAdjoint.cpp:170-182)
 * with R = A(gamma) x - b and alpha = gamma.  dA/dgamma_f is the face's
 * contribution in orc_assemble_pressure with gamma_f = 1 (the reference's
 * coefficient, src/solver/SimpleAlgorithm.cpp:373-397, 420-431), entry by
 * entry, and the pinned diagonal scaled by 1e6 (:441-446). */
void orc_face_gradient(const orc_mesh* m, const int* bcOfZone, int nZones, double rho,
                       const double* aP, int pinned, const double* x, const double* lam,
                       double* grad)
{
    orc_face_gradient_p(m, bcOfZone, nZones, NULL, rho, aP, pinned, x, lam, grad);
}

void orc_face_gradient_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                         double rho, const double* aP, int pinned, const double* x,
                         const double* lam, double* grad)
{
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f];
        const uint64_t nb = m->neighbor[f];
        const double Ax = m->areaX[f], Ay = m->areaY[f], Az = m->areaZ[f];
        const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
        /* entries (row, col, d value / d gamma_f) */
        int rows[4], cols[4];
        double dv[4];
        int ne = 0;
        if (nb != UINT64_MAX) {
            const double dx = m->cellCx[nb] - m->cellCx[o];
            const double dy = m->cellCy[nb] - m->cellCy[o];
            const double dz = m->cellCz[nb] - m->cellCz[o];
            const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
            double aPf = 1.0;
            if (aP) {
                const double dO = aP[o] > 1e-30 ? aP[o] : 1e-30;
                const double dN = aP[nb] > 1e-30 ? aP[nb] : 1e-30;
                aPf = 0.5 * (dO + dN);
            }
            const double den = aPf * ddotA;
            const double dcoef = rho * Amag2 / (den > 1e-30 ? den : 1e-30);
            rows[0] = (int)o;  cols[0] = (int)o;  dv[0] = dcoef;
            rows[1] = (int)o;  cols[1] = (int)nb; dv[1] = -dcoef;
            rows[2] = (int)nb; cols[2] = (int)nb; dv[2] = dcoef;
            rows[3] = (int)nb; cols[3] = (int)o;  dv[3] = -dcoef;
            ne = 4;
        } else if (twin_cell(m, twin, f) != UINT64_MAX
                   || zone_bc(bcOfZone, nZones, m->boundaryZone[f]) == ORC_BC_PRESSURE) {
            const double dx = m->faceCx[f] - m->cellCx[o];
            const double dy = m->faceCy[f] - m->cellCy[o];
            const double dz = m->faceCz[f] - m->cellCz[o];
            const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
            const double dO = aP ? (aP[o] > 1e-30 ? aP[o] : 1e-30) : 1.0;
            const double den = dO * ddotA;
            rows[0] = (int)o; cols[0] = (int)o;
            dv[0] = rho * Amag2 / (den > 1e-30 ? den : 1e-30);
            ne = 1;
            const uint64_t tw = twin_cell(m, twin, f);
            if (tw != UINT64_MAX) {
                rows[1] = (int)o; cols[1] = (int)tw; dv[1] = -dv[0];
                ne = 2;
            }
        }
        double t = 0.0;
        for (int e = 0; e < ne; ++e) {
            double v = dv[e];
            if (pinned && rows[e] == 0 && cols[e] == 0)
                v *= 1.0e6;
            t += lam[rows[e]] * v * x[cols[e]];
        }
        grad[f] = -t;
    }
}

/* SimpleAlgorithm::assemble_momentum, src/solver/SimpleAlgorithm.cpp:240-355,
 * face loop :253-303, transient :320-337 (implicit Euler), URF :340-348. */
void orc_assemble_momentum(const orc_mesh* m, const int* bcOfZone, int nZones,
                           const double* bcValue, int comp, double mu, double rho,
                           const double* phi, const double* pgradV, const double* dtCell,
                           const double* Uold, const double* Unm1, const double* Ucomp,
                           double urfU, const orc_csr* pat, double* values, double* rhs)
{
    orc_assemble_momentum_p(m, bcOfZone, nZones, NULL, bcValue, comp, mu, rho, phi, pgradV,
                            dtCell, Uold, Unm1, Ucomp, urfU, pat, values, rhs);
}

void orc_assemble_momentum_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                             const double* bcValue, int comp, double mu, double rho,
                             const double* phi, const double* pgradV, const double* dtCell,
                             const double* Uold, const double* Unm1, const double* Ucomp,
                             double urfU, const orc_csr* pat, double* values, double* rhs)
{
    const int64_t nC = m->n_cells;
    const int64_t nnz = pat->rowPtr[nC];
    for (int64_t k = 0; k < nnz; ++k)
        values[k] = 0.0;
    for (int64_t c = 0; c < nC; ++c)
        rhs[c] = 0.0;
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f];
        const uint64_t nb = m->neighbor[f];
        const double Ax = m->areaX[f], Ay = m->areaY[f], Az = m->areaZ[f];
        const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
        const double Ff = phi[f];
        const double ConvO = Ff > 0.0 ? Ff : 0.0;
        const double ConvN = -Ff > 0.0 ? -Ff : 0.0;
        if (nb != UINT64_MAX) {
            const double dx = m->cellCx[nb] - m->cellCx[o];
            const double dy = m->cellCy[nb] - m->cellCy[o];
            const double dz = m->cellCz[nb] - m->cellCz[o];
            const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
            const double D = mu * Amag2 / (ddotA > 1e-30 ? ddotA : 1e-30);
            add_at(pat, values, (int)o, (int)o, ConvO + D);
            add_at(pat, values, (int)o, (int)nb, -(ConvN + D));
            add_at(pat, values, (int)nb, (int)nb, ConvN + D);
            add_at(pat, values, (int)nb, (int)o, -(ConvO + D));
        } else {
            const int bc = zone_bc(bcOfZone, nZones, m->boundaryZone[f]);
            const double dx = m->faceCx[f] - m->cellCx[o];
            const double dy = m->faceCy[f] - m->cellCy[o];
            const double dz = m->faceCz[f] - m->cellCz[o];
            const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
            const double D = mu * Amag2 / (ddotA > 1e-30 ? ddotA : 1e-30);
            const uint64_t tw = twin_cell(m, twin, f);
            if (tw != UINT64_MAX) {
                /* periodic: the owner's half of an interior coupling (:279-285) */
                add_at(pat, values, (int)o, (int)o, ConvO + D);
                add_at(pat, values, (int)o, (int)tw, -(ConvN + D));
            } else if (bc == ORC_BC_WALL || bc == ORC_BC_VELOCITY_INLET) {
                const double Ub = (bc == ORC_BC_WALL)
                                      ? 0.0
                                      : (bcValue ? bcValue[(size_t)m->boundaryZone[f] * 3 + (size_t)comp] : 0.0);
                add_at(pat, values, (int)o, (int)o, ConvO + D);
                rhs[o] += (ConvN + D) * Ub;
            } else if (bc == ORC_BC_SYMMETRY) {
                /* no contribution */
            } else {
                add_at(pat, values, (int)o, (int)o, ConvO);
            }
        }
    }
    /* pressure-gradient source, :306-315 (pgradV = dp/dx_comp * V) */
    if (pgradV)
        for (int64_t c = 0; c < nC; ++c)
            rhs[c] -= pgradV[c];
    /* time term, :317-336: BDF2 when U^{n-1} is given, else implicit Euler */
    if (dtCell) {
        for (int64_t c = 0; c < nC; ++c) {
            const double Vrho_dt = rho * m->volume[c] / dtCell[c];
            const int kd = find_col(pat, (int)c, (int)c);
            if (kd < 0)
                continue;
            const double un = Uold ? Uold[c] : 0.0;
            if (Unm1) {
                values[kd] += 1.5 * Vrho_dt;
                rhs[c] += (2.0 * un - 0.5 * Unm1[c]) * Vrho_dt;
            } else {
                values[kd] += Vrho_dt;
                rhs[c] += un * Vrho_dt;
            }
        }
    }
    for (int64_t c = 0; c < nC; ++c) {
        const int kd = find_col(pat, (int)c, (int)c);
        if (kd < 0)
            continue;
        const double aDiag = values[kd];
        const double scaled = aDiag / urfU;
        values[kd] = scaled;
        rhs[c] += (scaled - aDiag) * (Ucomp ? Ucomp[c] : 0.0);
    }
}

/* SimpleAlgorithm::compute_face_fluxes, src/solver/SimpleAlgorithm.cpp:201-237 */
void orc_face_flux(const orc_mesh* m, const int* bcOfZone, int nZones,
                   const double* bcValue, double rho, const double* Ux,
                   const double* Uy, const double* Uz, double* phi)
{
    orc_face_flux_p(m, bcOfZone, nZones, NULL, bcValue, rho, Ux, Uy, Uz, phi);
}

void orc_face_flux_p(const orc_mesh* m, const int* bcOfZone, int nZones, const int* twin,
                     const double* bcValue, double rho, const double* Ux, const double* Uy,
                     const double* Uz, double* phi)
{
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f];
        const uint64_t nb = m->neighbor[f];
        double ux, uy, uz;
        if (nb != UINT64_MAX) {
            ux = 0.5 * (Ux[o] + Ux[nb]);
            uy = 0.5 * (Uy[o] + Uy[nb]);
            uz = 0.5 * (Uz[o] + Uz[nb]);
        } else if (twin_cell(m, twin, f) != UINT64_MAX) {
            const uint64_t tw = twin_cell(m, twin, f); /* :214-218 */
            ux = 0.5 * (Ux[o] + Ux[tw]);
            uy = 0.5 * (Uy[o] + Uy[tw]);
            uz = 0.5 * (Uz[o] + Uz[tw]);
        } else {
            const uint32_t z = m->boundaryZone[f];
            const int bc = zone_bc(bcOfZone, nZones, z);
            if (bc == ORC_BC_VELOCITY_INLET) {
                ux = bcValue ? bcValue[(size_t)z * 3 + 0] : 0.0;
                uy = bcValue ? bcValue[(size_t)z * 3 + 1] : 0.0;
                uz = bcValue ? bcValue[(size_t)z * 3 + 2] : 0.0;
            } else if (bc == ORC_BC_WALL || bc == ORC_BC_SYMMETRY) {
                ux = uy = uz = 0.0;
            } else {
                ux = Ux[o];
                uy = Uy[o];
                uz = Uz[o];
            }
        }
        phi[f] = rho * (ux * m->areaX[f] + uy * m->areaY[f] + uz * m->areaZ[f]);
    }
}

/* LeastSquaresGradient (src/solver/Gradient.cpp:42-136): per cell the
 * symmetric moment matrix sum w d d^T over its faces, w = 1/(|d|^2 + 1e-30),
 * d the offset to the neighbour centroid (boundary faces: to the face
 * centroid), inverted by its adjugate (zero when |det| < 1e-30); then
 * grad = M^-1 sum w d (phi_n - phi_o) over interior faces (boundary faces
 * contribute w d * 0).  Face loops in face order, as the reference. */
static void ls_acc(double* a, double dx, double dy, double dz)
{
    const double w2 = 1.0 / (dx * dx + dy * dy + dz * dz + 1e-30);
    a[0] += w2 * dx * dx;
    a[1] += w2 * dx * dy;
    a[2] += w2 * dx * dz;
    a[3] += w2 * dy * dy;
    a[4] += w2 * dy * dz;
    a[5] += w2 * dz * dz;
}

void orc_ls_gradient(const orc_mesh* m, const double* phi, double* gx, double* gy, double* gz)
{
    const int64_t nC = m->n_cells;
    double* A = (double*)calloc((size_t)(nC > 0 ? nC : 1) * 6, sizeof(double));
    double* B = (double*)calloc((size_t)(nC > 0 ? nC : 1) * 3, sizeof(double));
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f], nb = m->neighbor[f];
        if (nb == UINT64_MAX) {
            ls_acc(A + o * 6, m->faceCx[f] - m->cellCx[o], m->faceCy[f] - m->cellCy[o],
                   m->faceCz[f] - m->cellCz[o]);
        } else {
            const double dx = m->cellCx[nb] - m->cellCx[o];
            const double dy = m->cellCy[nb] - m->cellCy[o];
            const double dz = m->cellCz[nb] - m->cellCz[o];
            ls_acc(A + o * 6, dx, dy, dz);
            ls_acc(A + nb * 6, -dx, -dy, -dz);
        }
    }
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f], nb = m->neighbor[f];
        double dx, dy, dz, dphi;
        if (nb == UINT64_MAX) {
            dx = m->faceCx[f] - m->cellCx[o];
            dy = m->faceCy[f] - m->cellCy[o];
            dz = m->faceCz[f] - m->cellCz[o];
            dphi = 0.0;
        } else {
            dx = m->cellCx[nb] - m->cellCx[o];
            dy = m->cellCy[nb] - m->cellCy[o];
            dz = m->cellCz[nb] - m->cellCz[o];
            dphi = phi[nb] - phi[o];
        }
        const double w2 = 1.0 / (dx * dx + dy * dy + dz * dz + 1e-30);
        B[o * 3 + 0] += w2 * dx * dphi;
        B[o * 3 + 1] += w2 * dy * dphi;
        B[o * 3 + 2] += w2 * dz * dphi;
        if (nb != UINT64_MAX) {
            B[nb * 3 + 0] += w2 * dx * -dphi;
            B[nb * 3 + 1] += w2 * dy * -dphi;
            B[nb * 3 + 2] += w2 * dz * -dphi;
        }
    }
    for (int64_t c = 0; c < nC; ++c) {
        const double* a = A + c * 6;
        const double xx = a[0], xy = a[1], xz = a[2], yy = a[3], yz = a[4], zz = a[5];
        const double det =
            xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * xz) + xz * (xy * yz - yy * xz);
        double mv[6] = {0, 0, 0, 0, 0, 0};
        if (fabs(det) >= 1e-30) {
            const double inv = 1.0 / det;
            mv[0] = (yy * zz - yz * yz) * inv;
            mv[1] = -(xy * zz - yz * xz) * inv;
            mv[2] = (xy * yz - yy * xz) * inv;
            mv[3] = (xx * zz - xz * xz) * inv;
            mv[4] = -(xx * yz - xy * xz) * inv;
            mv[5] = (xx * yy - xy * xy) * inv;
        }
        const double bx = B[c * 3], by = B[c * 3 + 1], bz = B[c * 3 + 2];
        gx[c] = mv[0] * bx + mv[1] * by + mv[2] * bz;
        gy[c] = mv[1] * bx + mv[3] * by + mv[4] * bz;
        gz[c] = mv[2] * bx + mv[4] * by + mv[5] * bz;
    }
    free(A);
    free(B);
}

/* SimpleAlgorithm::correct_fields, src/solver/SimpleAlgorithm.cpp:450-489 */
void orc_correct_fields(const orc_mesh* m, double urfP, double rho, const double* aP,
                        const double* pp, double* p, double* Ux, double* Uy, double* Uz,
                        double* phi)
{
    const int64_t nC = m->n_cells;
    for (int64_t c = 0; c < nC; ++c)
        p[c] += urfP * pp[c];
    double* g = (double*)malloc(sizeof(double) * (size_t)(nC > 0 ? nC : 1) * 3);
    orc_ls_gradient(m, pp, g, g + nC, g + 2 * nC);
    for (int64_t c = 0; c < nC; ++c) {
        const double w = m->volume[c] / (aP[c] > 1e-30 ? aP[c] : 1e-30);
        Ux[c] -= w * g[c];
        Uy[c] -= w * g[nC + c];
        Uz[c] -= w * g[2 * nC + c];
    }
    free(g);
    if (!phi)
        return;
    for (int64_t f = 0; f < m->n_faces; ++f) {
        const uint64_t o = m->owner[f], nb = m->neighbor[f];
        if (nb == UINT64_MAX)
            continue;
        const double Ax = m->areaX[f], Ay = m->areaY[f], Az = m->areaZ[f];
        const double Amag2 = Ax * Ax + Ay * Ay + Az * Az;
        const double dx = m->cellCx[nb] - m->cellCx[o];
        const double dy = m->cellCy[nb] - m->cellCy[o];
        const double dz = m->cellCz[nb] - m->cellCz[o];
        const double ddotA = fabs(dx * Ax + dy * Ay + dz * Az);
        const double aPf = 0.5 * (aP[o] + aP[nb]);
        const double den = aPf * ddotA;
        const double coef = rho * Amag2 / (den > 1e-30 ? den : 1e-30);
        phi[f] += coef * (pp[o] - pp[nb]);
    }
}

