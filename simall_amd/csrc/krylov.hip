// krylov.hip — fused preconditioned CG kernels.
//
// One CG iteration of CGSolver::solve (src/solver/LinearSolvers.cpp:430-450):
//   Ap = A p ; pAp = p.Ap ; alpha = rz / pAp           k_cg_spmv_dot
//   x += alpha p ; r -= alpha Ap ; ||r|| ; stop test     k_cg_update[_jacobi]
//   z = M r ; rzNew = r.z ; beta = rzNew / rz            (fused for Jacobi)
//   p = z + beta p                                       k_cg_p[_jacobi]
// Scalars live in KState on the device; each reduction is finished by the last
// block of its launch, so the host only polls `done` every few iterations.
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

// ------------------------------------------------------------------- init
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_jacobi(CgArgs<T> a, int warm, double tol,
                                                           int max_iter)
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
            a.x[i] = T(0);
            a.r[i] = ri;
        }
        const T zi = a.invD[i] * ri;
        a.p[i] = zi;
        v[0] += double(ri) * double(zi);
        v[1] += double(bi) * double(bi);
    }
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        st->rz = tot[0];
        const double bn = sqrt(tot[1]);
        st->bnorm = bn > 1e-30 ? bn : 1e-30;
        st->tol = tol;
        st->max_iter = max_iter;
        st->iter = 0;
        st->done = max_iter > 0 ? 0 : 2;
        st->last = 0.0;
        st->alpha = st->beta = 0.0;
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_r(CgArgs<T> a)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        a.x[i] = T(0);
        a.r[i] = a.b[i];
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_init_p(CgArgs<T> a, double tol, int max_iter)
{
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T ri = a.r[i], zi = a.z[i], bi = a.b[i];
        a.p[i] = zi;
        v[0] += double(ri) * double(zi);
        v[1] += double(bi) * double(bi);
    }
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        st->rz = tot[0];
        const double bn = sqrt(tot[1]);
        st->bnorm = bn > 1e-30 ? bn : 1e-30;
        st->tol = tol;
        st->max_iter = max_iter;
        st->iter = 0;
        st->done = max_iter > 0 ? 0 : 2;
        st->last = 0.0;
        st->alpha = st->beta = 0.0;
    }
}

// -------------------------------------------------------------- iteration
template <class T> struct EpiCgAp
{
    const T* p;
    T* Ap;
    double acc = 0.0;
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T ap = T(s);
        Ap[i] = ap;
        acc += double(p[i]) * double(ap);
    }
};

// Ap = A p ; alpha = rz / p.Ap
template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dot(CgArgs<T> a)
{
    if (a.st->done)
        return;
    StreamMat<T> A{a.n, a.nrb, a.blk, a.rp, a.ci, a.val};
    EpiCgAp<T> e{a.p, a.Ap};
    stream_rows(A, a.p, e);
    double v[1] = {e.acc};
    double tot[1];
    if (grid_sum<1>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        const double pAp = tot[0];
        if (pAp == 0.0) {
            // src/solver/LinearSolvers.cpp:433-436
            st->last = sqrt(st->rz) / st->bnorm;
            st->done = 3;
        } else {
            st->alpha = st->rz / pAp;
        }
    }
}

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

template <class T>
__global__ __launch_bounds__(kBlock) void k_cg_update_jacobi(CgArgs<T> a)
{
    if (a.st->done)
        return;
    const T al = T(a.st->alpha);
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    double v[2] = {0.0, 0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T xi = a.x[i] + al * a.p[i];
        const T ri = a.r[i] - al * a.Ap[i];
        a.x[i] = xi;
        a.r[i] = ri;
        const T zi = a.invD[i] * ri;
        v[0] += double(ri) * double(ri);
        v[1] += double(ri) * double(zi);
    }
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        cg_stop_test(st, tot[0], a.history);
        if (!st->done) {
            st->beta = tot[1] / st->rz;
            st->rz = tot[1];
        }
    }
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
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        const T xi = a.x[i] + al * a.p[i];
        const T ri = a.r[i] - al * a.Ap[i];
        a.x[i] = xi;
        a.r[i] = ri;
        v[0] += double(ri) * double(ri);
    }
    double tot[1];
    if (grid_sum<1>(v, a.partials, a.counter, tot) && threadIdx.x == 0)
        cg_stop_test(a.st, tot[0], a.history);
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
    double tot[1];
    if (grid_sum<1>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        st->beta = tot[0] / st->rz;
        st->rz = tot[0];
    }
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

// -------------------------------------------------------------- launchers
template <class T> int cg_init_jacobi(const CgArgs<T>& a, bool warm, double tol, int max_iter,
                                       hipStream_t s)
{
    k_cg_init_jacobi<T><<<a.grid, kBlock, 0, s>>>(a, warm ? 1 : 0, tol, max_iter);
    MI_HIP(hipGetLastError());
    return MI_OK;
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
    k_cg_init_p<T><<<a.grid, kBlock, 0, s>>>(a, tol, max_iter);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> static int launch_spmv_dot(const CgArgs<T>& a, hipStream_t s)
{
    k_cg_spmv_dot<T><<<stream_grid(a.nrb), kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int cg_iteration_jacobi(const CgArgs<T>& a, hipStream_t s)
{
    MI_TRY(launch_spmv_dot(a, s));
    k_cg_update_jacobi<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    k_cg_p_jacobi<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int cg_iteration_part1(const CgArgs<T>& a, hipStream_t s)
{
    MI_TRY(launch_spmv_dot(a, s));
    k_cg_update<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int cg_iteration_part2(const CgArgs<T>& a, hipStream_t s)
{
    k_cg_rz<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    k_cg_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// --------------------------------------------------------------- BiCGStab
template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_init(BiArgs<T> a, int warm, double tol, int max_iter)
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
            a.x[i] = T(0);
            a.r[i] = ri;
        }
        a.rh[i] = ri;
        a.p[i] = T(0);
        a.v[i] = T(0);
        v[0] += double(ri) * double(ri);
        v[1] += double(bi) * double(bi);
    }
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        const double bn = sqrt(tot[1]);
        st->bnorm = bn > 1e-30 ? bn : 1e-30;
        st->tol = tol;
        st->max_iter = max_iter;
        st->iter = 0;
        st->rz = tot[0];          // rho_new = rh.r with rh = r
        st->scratch[0] = 1.0;     // rho
        st->alpha = 1.0;
        st->omega = 1.0;
        st->beta = tot[0];        // (rho_new / rho) * (alpha / omega)
        st->last = sqrt(tot[0]) / st->bnorm;
        st->done = max_iter > 0 ? 0 : 2;
        if (tot[0] == 0.0 && max_iter > 0)
            st->done = 3; // rho_new == 0 at it = 0 (LinearSolvers.cpp:364-367)
    }
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
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type sres)
    {
        const T o = T(sres);
        out[i] = o;
        if (mode == 0) {
            acc0 += double(w1[i]) * double(o);
        } else {
            acc0 += double(o) * double(w1[i]);
            acc1 += double(o) * double(o);
        }
    }
};

// mode 0: v = A y, alpha = rho_new / rh.v ; mode 1: t = A z, omega = t.s / max(t.t, 1e-30)
template <class T>
__global__ __launch_bounds__(kBlock) void k_bi_spmv(BiArgs<T> a, int mode)
{
    if (a.st->done)
        return;
    StreamMat<T> A{a.n, a.nrb, a.blk, a.rp, a.ci, a.val};
    EpiBi<T> e{mode == 0 ? a.rh : a.s, mode == 0 ? a.v : a.t, mode};
    stream_rows(A, mode == 0 ? a.y : a.z, e);
    double v[2] = {e.acc0, e.acc1};
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        if (mode == 0)
            st->alpha = st->rz / tot[0];
        else
            st->omega = tot[0] / (tot[1] > 1e-30 ? tot[1] : 1e-30);
    }
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
        a.x[i] += al * a.y[i];
        v[0] += double(si) * double(si);
    }
    double tot[1];
    if (grid_sum<1>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        const double sn = sqrt(tot[0]);
        if (sn < st->tol * st->bnorm) {
            // early exit on ||s|| (LinearSolvers.cpp:386-391)
            st->last = sn / st->bnorm;
            st->iter += 1;
            if (a.history)
                a.history[st->iter - 1] = st->last;
            st->done = 1;
        }
    }
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
        a.x[i] += om * a.z[i];
        const T ri = a.s[i] - om * a.t[i];
        a.r[i] = ri;
        v[0] += double(ri) * double(ri);
        v[1] += double(a.rh[i]) * double(ri);
    }
    double tot[2];
    if (grid_sum<2>(v, a.partials, a.counter, tot) && threadIdx.x == 0) {
        KState* st = a.st;
        const double rho = st->rz;
        st->scratch[0] = rho;
        st->rz = tot[1];
        st->last = sqrt(tot[0]) / st->bnorm;
        const int it = st->iter + 1;
        st->iter = it;
        if (a.history)
            a.history[it - 1] = st->last;
        if (st->last < st->tol)
            st->done = 1;
        else if (it >= st->max_iter)
            st->done = 2;
        else if (tot[1] == 0.0)
            st->done = 3;
        else
            st->beta = (tot[1] / rho) * (st->alpha / st->omega);
    }
}

template <class T> int bicg_init(const BiArgs<T>& a, bool warm, double tol, int max_iter,
                                  hipStream_t s)
{
    k_bi_init<T><<<a.grid, kBlock, 0, s>>>(a, warm ? 1 : 0, tol, max_iter);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int bicg_stage1(const BiArgs<T>& a, hipStream_t s)
{
    k_bi_p<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> static int bi_spmv(const BiArgs<T>& a, int mode, hipStream_t s)
{
    k_bi_spmv<T><<<stream_grid(a.nrb), kBlock, 0, s>>>(a, mode);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int bicg_stage2(const BiArgs<T>& a, hipStream_t s)
{
    MI_TRY(bi_spmv(a, 0, s));
    k_bi_s<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int bicg_stage3(const BiArgs<T>& a, hipStream_t s)
{
    MI_TRY(bi_spmv(a, 1, s));
    k_bi_update<T><<<a.grid, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

#define MI_INSTANTIATE_BI(T)                                                                     \
    template int bicg_init<T>(const BiArgs<T>&, bool, double, int, hipStream_t);                 \
    template int bicg_stage1<T>(const BiArgs<T>&, hipStream_t);                                  \
    template int bicg_stage2<T>(const BiArgs<T>&, hipStream_t);                                  \
    template int bicg_stage3<T>(const BiArgs<T>&, hipStream_t);

MI_INSTANTIATE_BI(double)
MI_INSTANTIATE_BI(float)

#define MI_INSTANTIATE_CG(T)                                                                     \
    template int cg_init_jacobi<T>(const CgArgs<T>&, bool, double, int, hipStream_t);            \
    template int cg_init_r<T>(const CgArgs<T>&, bool, hipStream_t);                              \
    template int cg_init_p<T>(const CgArgs<T>&, double, int, hipStream_t);                       \
    template int cg_iteration_jacobi<T>(const CgArgs<T>&, hipStream_t);                          \
    template int cg_iteration_part1<T>(const CgArgs<T>&, hipStream_t);                           \
    template int cg_iteration_part2<T>(const CgArgs<T>&, hipStream_t);

MI_INSTANTIATE_CG(double)
MI_INSTANTIATE_CG(float)

} // namespace mi
