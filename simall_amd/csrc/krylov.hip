// krylov.hip — fused preconditioned CG and BiCGStab kernels.
//
// One CG iteration of CGSolver::solve (src/solver/LinearSolvers.cpp:430-450):
//   Ap = A p ; pAp = p.Ap ; alpha = rz / pAp           k_cg_spmv_dot + fin ALPHA
//   x += alpha p ; r -= alpha Ap ; ||r|| ; stop test     k_cg_update[_jacobi] + fin
//   z = M r ; rzNew = r.z ; beta = rzNew / rz            (fused for Jacobi)
//   p = z + beta p                                       k_cg_p[_jacobi]
// and of BiCGSTABSolver::solve (src/solver/LinearSolvers.cpp:360-404).
// Scalars live in KState on the device.  Every reduction is a producing
// kernel that stores block partials and a one-block finalize kernel (k_fin)
// that folds them in a fixed order and advances the scalar recurrence, so an
// iteration needs no host round trip and results are run-to-run identical;
// the host only polls `done` every few iterations.
#include "krylov.hpp"
#include "stream.hpp"

namespace mi
{

unsigned range_grid(int64_t rows, int min_rows_per_block)
{
    int64_t g = (rows + min_rows_per_block - 1) / min_rows_per_block;
    g = (g + 7) / 8 * 8;
    if (g < 8)
        g = 8;
    if (g > kMaxReduceBlocks)
        g = kMaxReduceBlocks;
    return unsigned(g);
}

// ------------------------------------------------------------- finalize
enum FinOp {
    FIN_CG_INIT,       // tot: r.z, b.b
    FIN_CG_ALPHA,      // tot: p.Ap
    FIN_CG_UPDATE_JAC, // tot: r.r, r.z
    FIN_CG_UPDATE,     // tot: r.r
    FIN_CG_RZ,         // tot: r.z
    FIN_BI_INIT,       // tot: r.r, b.b
    FIN_BI_ALPHA,      // tot: rh.v
    FIN_BI_OMEGA,      // tot: t.s, t.t
    FIN_BI_S,          // tot: s.s
    FIN_BI_UPDATE      // tot: r.r, rh.r
};

struct FinArgs
{
    KState* st;
    double* history;
    double tol;
    int max_iter;
};

__device__ __forceinline__ void cg_stop_test(KState* st, double rr, double* history)
{
    st->last = sqrt(rr) / st->bnorm;
    const int it = st->iter + 1;
    st->iter = it;
    if (history)
        history[it - 1] = st->last;
    if (st->last < st->tol)
        st->done = 1;
    else if (it >= st->max_iter)
        st->done = 2;
}

template <int NV>
__global__ __launch_bounds__(kBlock) void k_fin(const double* __restrict__ partials, int G, int op,
                                                FinArgs f)
{
    KState* st = f.st;
    const bool init = op == FIN_CG_INIT || op == FIN_BI_INIT;
    if (!init && st->done)
        return;
    double t[NV];
    fold_partials<NV>(partials, G, t);
    if (threadIdx.x != 0)
        return;
    double tot[2] = {t[0], NV > 1 ? t[NV - 1] : 0.0};
    switch (op) {
    case FIN_CG_INIT: {
        st->rz = tot[0];
        const double bn = sqrt(tot[1]);
        st->bnorm = bn > 1e-30 ? bn : 1e-30;
        st->tol = f.tol;
        st->max_iter = f.max_iter;
        st->iter = 0;
        st->done = f.max_iter > 0 ? 0 : 2;
        st->last = 0.0;
        st->alpha = st->beta = 0.0;
        break;
    }
    case FIN_CG_ALPHA:
        if (tot[0] == 0.0) {
            // src/solver/LinearSolvers.cpp:433-436
            st->last = sqrt(st->rz) / st->bnorm;
            st->done = 3;
        } else {
            st->alpha = st->rz / tot[0];
        }
        break;
    case FIN_CG_UPDATE_JAC:
        cg_stop_test(st, tot[0], f.history);
        if (!st->done) {
            st->beta = tot[1] / st->rz;
            st->rz = tot[1];
        }
        break;
    case FIN_CG_UPDATE:
        cg_stop_test(st, tot[0], f.history);
        break;
    case FIN_CG_RZ:
        st->beta = tot[0] / st->rz;
        st->rz = tot[0];
        break;
    case FIN_BI_INIT: {
        const double bn = sqrt(tot[1]);
        st->bnorm = bn > 1e-30 ? bn : 1e-30;
        st->tol = f.tol;
        st->max_iter = f.max_iter;
        st->iter = 0;
        st->rz = tot[0];      // rho_new = rh.r with rh = r
        st->scratch[0] = 1.0; // rho
        st->alpha = 1.0;
        st->omega = 1.0;
        st->beta = tot[0];    // (rho_new / rho) * (alpha / omega)
        st->last = sqrt(tot[0]) / st->bnorm;
        st->done = f.max_iter > 0 ? 0 : 2;
        if (tot[0] == 0.0 && f.max_iter > 0)
            st->done = 3; // rho_new == 0 at it = 0 (LinearSolvers.cpp:364-367)
        break;
    }
    case FIN_BI_ALPHA:
        st->alpha = st->rz / tot[0];
        break;
    case FIN_BI_OMEGA:
        st->omega = tot[0] / (tot[1] > 1e-30 ? tot[1] : 1e-30);
        break;
    case FIN_BI_S: {
        const double sn = sqrt(tot[0]);
        if (sn < st->tol * st->bnorm) {
            // early exit on ||s|| (LinearSolvers.cpp:386-391); x += alpha y was applied
            st->last = sn / st->bnorm;
            st->iter += 1;
            if (f.history)
                f.history[st->iter - 1] = st->last;
            st->done = 1;
        }
        break;
    }
    case FIN_BI_UPDATE: {
        const double rho = st->rz;
        st->scratch[0] = rho;
        st->rz = tot[1];
        st->last = sqrt(tot[0]) / st->bnorm;
        const int it = st->iter + 1;
        st->iter = it;
        if (f.history)
            f.history[it - 1] = st->last;
        if (st->last < st->tol)
            st->done = 1;
        else if (it >= st->max_iter)
            st->done = 2;
        else if (tot[1] == 0.0)
            st->done = 3;
        else
            st->beta = (tot[1] / rho) * (st->alpha / st->omega);
        break;
    }
    default:
        break;
    }
}

static int fin(int nv, const double* partials, unsigned G, int op, KState* st, double* history,
               double tol, int max_iter, hipStream_t s)
{
    FinArgs f{st, history, tol, max_iter};
    if (nv == 1)
        k_fin<1><<<1, kBlock, 0, s>>>(partials, int(G), op, f);
    else
        k_fin<2><<<1, kBlock, 0, s>>>(partials, int(G), op, f);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// ------------------------------------------------------------------- CG
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_jacobi(CgArgs<T> a, int warm)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T bi = a.b[i];
        T ri;
        if (warm) {
            ri = a.r[i];
        } else {
            ri = bi;
            if (!a.xacc)
                a.x[i] = T(0);
            a.r[i] = ri;
        }
        const T zi = a.invD[i] * ri;
        a.p[i] = zi;
        v[0] += double(ri) * double(zi);
        v[1] += double(bi) * double(bi);
    }
    block_partials<2>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_r(CgArgs<T> a)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        if (!a.xacc)
            a.x[i] = T(0);
        a.r[i] = a.b[i];
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_p(CgArgs<T> a)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T ri = a.r[i], zi = a.z[i], bi = a.b[i];
        a.p[i] = zi;
        v[0] += double(ri) * double(zi);
        v[1] += double(bi) * double(bi);
    }
    block_partials<2>(v, a.partials);
}

template <class T> struct EpiCgAp
{
    const T* p;
    T* Ap;
    double acc = 0.0;
    T pi = 0;
    __device__ __forceinline__ void pre(int i) { pi = p[i]; }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T ap = T(s);
        Ap[i] = ap;
        acc += double(pi) * double(ap);
    }
};

// Ap = A p ; partials of p.Ap
template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dot(CgArgs<T> a)
{
    if (a.st->done)
        return;
    StreamMat<T> A{a.n, a.nrb, a.blk, a.rp, a.ci, a.val, a.diag, 1, a.cw, a.hdr, a.soff, a.eci, a.nsl};
    A.ecd = a.ecd;
    EpiCgAp<T> e{a.p, a.Ap};
    rows_of<kSell>(A, a.p, e);
    double v[1] = {e.acc};
    block_partials<1>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update_jacobi(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const T al = T(a.st->alpha);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    const double al64 = a.st->alpha;
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        if (a.xacc)
            a.xacc[i] += al64 * double(a.p[i]);
        else
            a.x[i] = a.x[i] + al * a.p[i];
        const T ri = a.r[i] - al * a.Ap[i];
        a.r[i] = ri;
        const T zi = a.invD[i] * ri;
        v[0] += double(ri) * double(ri);
        v[1] += double(ri) * double(zi);
    }
    block_partials<2>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_p_jacobi(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const T be = T(a.st->beta);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        a.p[i] = a.invD[i] * a.r[i] + be * a.p[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const T al = T(a.st->alpha);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[1] = {0.0};
    const double al64 = a.st->alpha;
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        if (a.xacc)
            a.xacc[i] += al64 * double(a.p[i]);
        else
            a.x[i] = a.x[i] + al * a.p[i];
        const T ri = a.r[i] - al * a.Ap[i];
        a.r[i] = ri;
        v[0] += double(ri) * double(ri);
    }
    block_partials<1>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_rz(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[1] = {0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        v[0] += double(a.r[i]) * double(a.z[i]);
    block_partials<1>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_p(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const T be = T(a.st->beta);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        a.p[i] = a.z[i] + be * a.p[i];
}

template <class T> int cg_init_jacobi(const CgArgs<T>& a, bool warm, double tol, int max_iter,
                                       hipStream_t s)
{
    k_cg_init_jacobi<T><<<a.grid, kBlock, 0, s>>>(a, warm ? 1 : 0);
    MI_HIP(hipGetLastError());
    return fin(2, a.partials, a.grid, FIN_CG_INIT, a.st, a.history, tol, max_iter, s);
}

template <class T> int cg_init_r(const CgArgs<T>& a, bool warm, hipStream_t s)
{
    if (!warm) {
        k_cg_init_r<T><<<a.grid, kBlock, 0, s>>>(a);
        MI_HIP(hipGetLastError());
    }
    return MI_OK;
}

template <class T> int cg_init_p(const CgArgs<T>& a, double tol, int max_iter, hipStream_t s)
{
    k_cg_init_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return fin(2, a.partials, a.grid, FIN_CG_INIT, a.st, a.history, tol, max_iter, s);
}

template <class T> static int spmv_dot(const CgArgs<T>& a, hipStream_t s)
{
    const unsigned G = a.soff ? sell_grid(a.nsl) : stream_grid(a.nrb);
    (a.soff ? k_cg_spmv_dot<T, true> : k_cg_spmv_dot<T, false>)<<<G, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return fin(1, a.partials, G, FIN_CG_ALPHA, a.st, a.history, 0, 0, s);
}

template <class T> int cg_iteration_jacobi(const CgArgs<T>& a, hipStream_t s)
{
    MI_TRY(spmv_dot(a, s));
    k_cg_update_jacobi<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    MI_TRY(fin(2, a.partials, a.grid, FIN_CG_UPDATE_JAC, a.st, a.history, 0, 0, s));
    k_cg_p_jacobi<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int cg_iteration_part1(const CgArgs<T>& a, hipStream_t s)
{
    MI_TRY(spmv_dot(a, s));
    k_cg_update<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return fin(1, a.partials, a.grid, FIN_CG_UPDATE, a.st, a.history, 0, 0, s);
}

template <class T> int cg_iteration_spmv(const CgArgs<T>& a, hipStream_t s)
{
    return spmv_dot(a, s);
}

// The fused iteration's finalize: FIN_CG_UPDATE on r.r followed by FIN_CG_RZ
// on r.z, each folded from its own producing launch.
__global__ __launch_bounds__(kBlock) void k_fin_cg_fused(const double* __restrict__ prr, int grr,
                                                         const double* __restrict__ prz, int grz,
                                                         FinArgs f)
{
    KState* st = f.st;
    if (st->done)
        return;
    double t[1], u[1];
    fold_partials<1>(prr, grr, t);
    fold_partials<1>(prz, grz, u);
    if (threadIdx.x != 0)
        return;
    cg_stop_test(st, t[0], f.history);
    if (!st->done) {
        st->beta = u[0] / st->rz;
        st->rz = u[0];
    }
}

template <class T> int cg_iteration_fused_tail(const CgArgs<T>& a, const double* prr, unsigned grr,
                                               const double* prz, unsigned grz, hipStream_t s)
{
    FinArgs f{a.st, a.history, 0, 0};
    k_fin_cg_fused<<<1, kBlock, 0, s>>>(prr, int(grr), prz, int(grz), f);
    MI_HIP(hipGetLastError());
    k_cg_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int cg_iteration_part2(const CgArgs<T>& a, hipStream_t s)
{
    k_cg_rz<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    MI_TRY(fin(1, a.partials, a.grid, FIN_CG_RZ, a.st, a.history, 0, 0, s));
    k_cg_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// --------------------------------------------------------------- BiCGStab
template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_init(BiArgs<T> a, int warm)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T bi = a.b[i];
        T ri;
        if (warm) {
            ri = a.r[i];
        } else {
            ri = bi;
            if (!a.xacc)
                a.x[i] = T(0);
            a.r[i] = ri;
        }
        a.rh[i] = ri;
        a.p[i] = T(0);
        a.v[i] = T(0);
        v[0] += double(ri) * double(ri);
        v[1] += double(bi) * double(bi);
    }
    block_partials<2>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_p(BiArgs<T> a)
{
    if (a.st->done)
        return;
    const T be = T(a.st->beta), om = T(a.st->omega);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock)
        a.p[i] = a.r[i] + be * (a.p[i] - om * a.v[i]);
}

template <class T> struct EpiBi
{
    const T* w1; // rh (mode 0) or s (mode 1)
    T* out;
    int mode;
    double acc0 = 0.0, acc1 = 0.0;
    T wi = 0;
    __device__ __forceinline__ void pre(int i) { wi = w1[i]; }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type sres)
    {
        const T o = T(sres);
        out[i] = o;
        acc0 += double(wi) * double(o);
        if (mode == 1)
            acc1 += double(o) * double(o);
    }
};

// mode 0: v = A y, partials of rh.v ; mode 1: t = A z, partials of t.s, t.t
template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_bi_spmv(BiArgs<T> a, int mode)
{
    if (a.st->done)
        return;
    StreamMat<T> A{a.n, a.nrb, a.blk, a.rp, a.ci, a.val, a.diag, 1, a.cw, a.hdr, a.soff, a.eci, a.nsl};
    A.ecd = a.ecd;
    EpiBi<T> e{mode == 0 ? a.rh : a.s, mode == 0 ? a.v : a.t, mode};
    rows_of<kSell>(A, mode == 0 ? a.y : a.z, e);
    double v[2] = {e.acc0, e.acc1};
    block_partials<2>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_s(BiArgs<T> a)
{
    if (a.st->done)
        return;
    const T al = T(a.st->alpha);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[1] = {0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T si = a.r[i] - al * a.v[i];
        a.s[i] = si;
        if (a.xacc)
            a.xacc[i] += a.st->alpha * double(a.y[i]);
        else
            a.x[i] += al * a.y[i];
        v[0] += double(si) * double(si);
    }
    block_partials<1>(v, a.partials);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_update(BiArgs<T> a)
{
    if (a.st->done)
        return;
    const T om = T(a.st->omega);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        if (a.xacc)
            a.xacc[i] += a.st->omega * double(a.z[i]);
        else
            a.x[i] += om * a.z[i];
        const T ri = a.s[i] - om * a.t[i];
        a.r[i] = ri;
        v[0] += double(ri) * double(ri);
        v[1] += double(a.rh[i]) * double(ri);
    }
    block_partials<2>(v, a.partials);
}

template <class T> int bicg_init(const BiArgs<T>& a, bool warm, double tol, int max_iter,
                                  hipStream_t s)
{
    k_bi_init<T><<<a.grid, kBlock, 0, s>>>(a, warm ? 1 : 0);
    MI_HIP(hipGetLastError());
    return fin(2, a.partials, a.grid, FIN_BI_INIT, a.st, a.history, tol, max_iter, s);
}

template <class T> int bicg_stage1(const BiArgs<T>& a, hipStream_t s)
{
    k_bi_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int bicg_stage2(const BiArgs<T>& a, hipStream_t s)
{
    const unsigned G = a.soff ? sell_grid(a.nsl) : stream_grid(a.nrb);
    (a.soff ? k_bi_spmv<T, true> : k_bi_spmv<T, false>)<<<G, kBlock, 0, s>>>(a, 0);
    MI_HIP(hipGetLastError());
    MI_TRY(fin(2, a.partials, G, FIN_BI_ALPHA, a.st, a.history, 0, 0, s));
    k_bi_s<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return fin(1, a.partials, a.grid, FIN_BI_S, a.st, a.history, 0, 0, s);
}

template <class T> int bicg_stage3(const BiArgs<T>& a, hipStream_t s)
{
    const unsigned G = a.soff ? sell_grid(a.nsl) : stream_grid(a.nrb);
    (a.soff ? k_bi_spmv<T, true> : k_bi_spmv<T, false>)<<<G, kBlock, 0, s>>>(a, 1);
    MI_HIP(hipGetLastError());
    MI_TRY(fin(2, a.partials, G, FIN_BI_OMEGA, a.st, a.history, 0, 0, s));
    k_bi_update<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return fin(2, a.partials, a.grid, FIN_BI_UPDATE, a.st, a.history, 0, 0, s);
}

#define MI_INSTANTIATE(T)                                                                        \
    template int cg_init_jacobi<T>(const CgArgs<T>&, bool, double, int, hipStream_t);            \
    template int cg_init_r<T>(const CgArgs<T>&, bool, hipStream_t);                              \
    template int cg_init_p<T>(const CgArgs<T>&, double, int, hipStream_t);                       \
    template int cg_iteration_jacobi<T>(const CgArgs<T>&, hipStream_t);                          \
    template int cg_iteration_part1<T>(const CgArgs<T>&, hipStream_t);                           \
    template int cg_iteration_part2<T>(const CgArgs<T>&, hipStream_t);                           \
    template int cg_iteration_spmv<T>(const CgArgs<T>&, hipStream_t);                            \
    template int cg_iteration_fused_tail<T>(const CgArgs<T>&, const double*, unsigned,           \
                                            const double*, unsigned, hipStream_t);               \
    template int bicg_init<T>(const BiArgs<T>&, bool, double, int, hipStream_t);                 \
    template int bicg_stage1<T>(const BiArgs<T>&, hipStream_t);                                  \
    template int bicg_stage2<T>(const BiArgs<T>&, hipStream_t);                                  \
    template int bicg_stage3<T>(const BiArgs<T>&, hipStream_t);

MI_INSTANTIATE(double)
MI_INSTANTIATE(float)

// ------------------------------------------------------------ plain sums
__global__ __launch_bounds__(kBlock) void k_fin_sum(const double* __restrict__ partials, int G,
                                                    int take_sqrt, double* __restrict__ out)
{
    double tot[1];
    fold_partials<1>(partials, G, tot);
    if (threadIdx.x == 0)
        out[0] = take_sqrt ? sqrt(tot[0]) : tot[0];
}

int launch_fin_sum(const double* partials, unsigned G, bool take_sqrt, double* out, hipStream_t s)
{
    k_fin_sum<<<1, kBlock, 0, s>>>(partials, int(G), take_sqrt ? 1 : 0, out);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

} // namespace mi
