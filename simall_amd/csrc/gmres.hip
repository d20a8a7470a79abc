// gmres.hip — restarted GMRES(m) kernels.
//
// One Arnoldi step of GMRESmSolver::solve (src/solver/LinearSolvers.cpp:293-326):
//   w = A v_j ; z = M^-1 w (written to v_{j+1})          k_gm_spmv + preconditioner
//   h1 = V^T z                                           k_gm_dots + k_gm_fold
//   z -= V h1 ; h2 = V^T z                               k_gm_orth + k_gm_fold
//   z -= V h2 ; ||z||                                     k_gm_orth
//   H(:,j) = h1 + h2, h_{j+1,j} = ||z||, Givens, |s_{j+1}| k_gm_step (one block)
//   v_{j+1} = z / h_{j+1,j}                               k_gm_scale
// The reference orthogonalises with modified Gram-Schmidt, one dot and one
// axpy per basis vector.  Here classical Gram-Schmidt applied twice takes its
// place: the same projection with orthogonality to working precision, in a
// fixed number of passes over the basis whatever j is.  Reductions are block
// partials folded in a fixed order, so results are run-to-run identical.
#include "gmres.hpp"

#include <cmath>

namespace mi
{

unsigned gm_grid(int64_t n)
{
    constexpr int64_t per_block = int64_t(kGmRowsPerThread) * kBlock;
    int64_t g = (n + per_block - 1) / per_block;
    // small systems: spread over more blocks with fewer rows each
    const int64_t spread = (n + kBlock - 1) / kBlock;
    g = std::max(g, std::min<int64_t>(spread, 512));
    g = (g + 7) / 8 * 8;
    if (g < 8)
        g = 8;
    return unsigned(g);
}

namespace
{

__device__ __forceinline__ bool gm_idle(const GmState* g, bool in_cycle)
{
    return g->k.done != 0 || (in_cycle && g->cycle_end != 0);
}

template <class T> struct EpiGmRes
{
    const double* b;
    T* out;
    const T* invD;
    double bi = 0.0;
    T di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        bi = b[i];
        if (invD)
            di = invD[i];
    }
    __device__ __forceinline__ void row(int i, double s)
    {
        const T r = T(bi - s);
        out[i] = invD ? T(di * r) : r;
    }
};

template <class T> struct EpiGmStep
{
    T* out;
    const T* invD;
    T di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        if (invD)
            di = invD[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        out[i] = invD ? T(di * T(s)) : T(s);
    }
};

// w = b - A64 x (or v_0 = D^-1 (b - A64 x) with Jacobi)
template <class T, bool kSell> __global__ __launch_bounds__(kBlock) void k_gm_residual(GmArgs<T> a)
{
    if (gm_idle(a.g, false))
        return;
    EpiGmRes<T> e{a.b, a.invD ? a.V : a.w, a.invD};
    rows_of<kSell>(a.A64, a.x, e);
}

// w = A v_j (or v_{j+1} = D^-1 A v_j with Jacobi)
template <class T, bool kSell> __global__ __launch_bounds__(kBlock) void k_gm_spmv(GmArgs<T> a, int j)
{
    if (gm_idle(a.g, true))
        return;
    EpiGmStep<T> e{a.invD ? a.V + size_t(j + 1) * a.ld : a.w, a.invD};
    rows_of<kSell>(a.A, a.V + size_t(j) * a.ld, e);
}

// Per-wave sums of nv values, then one partial per value and block:
// partials[i * G + block].
__device__ __forceinline__ void gm_store_partials(double* lds, int nv, double* partials)
{
    __syncthreads();
    constexpr int W = kBlock / kWave;
    for (int i = threadIdx.x; i < nv; i += kBlock) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w)
            s += lds[i * W + w];
        partials[size_t(i) * gridDim.x + blockIdx.x] = s;
    }
}

constexpr int R = kGmRowsPerThread;

// This thread's rows of the block's range: row k_r = beg + tid + r * kBlock.
// Loads use the clamped index kk[r] (always inside the vector), so the basis
// loops are straight-line code whose loads the compiler can issue together;
// rows outside the range carry z = 0 and are never stored.
struct GmRows
{
    int kk[R];
    bool in[R];
    __device__ __forceinline__ GmRows(const RowRange& rr, int n)
    {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int k = rr.beg + int(threadIdx.x) + r * kBlock;
            in[r] = k < rr.end;
            kk[r] = min(k, n - 1);
        }
    }
};

// lds[i][wave] = this wave's share of v_i . z for NG vectors from i0
template <class T, int NG>
__device__ __forceinline__ void gm_dot_group(const GmArgs<T>& a, int i0, const GmRows& rw,
                                             const double (&zr)[R], double* lds)
{
    T v[NG][R];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[g][r] = a.V[size_t(i0 + g) * a.ld + rw.kk[r]];
    double s[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        s[g] = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            s[g] += double(v[g][r]) * zr[r];
    }
#pragma unroll
    for (int g = 0; g < NG; ++g)
        s[g] = wave_sum(s[g]);
    if ((threadIdx.x & (kWave - 1)) == 0)
#pragma unroll
        for (int g = 0; g < NG; ++g)
            lds[(i0 + g) * (kBlock / kWave) + threadIdx.x / kWave] = s[g];
}

template <class T>
__device__ __forceinline__ void gm_dots(const GmArgs<T>& a, int nv, const GmRows& rw,
                                        const double (&zr)[R], double* lds)
{
    int i = 0;
    for (; i + 4 <= nv; i += 4)
        gm_dot_group<T, 4>(a, i, rw, zr, lds);
    for (; i < nv; ++i)
        gm_dot_group<T, 1>(a, i, rw, zr, lds);
    gm_store_partials(lds, nv, a.partials);
}

// z -= h_i v_i for NG vectors from i0, in ascending i per row
template <class T, int NG>
__device__ __forceinline__ void gm_update_group(const GmArgs<T>& a, int i0, const double* h,
                                                const GmRows& rw, double (&zr)[R])
{
    T v[NG][R];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[g][r] = a.V[size_t(i0 + g) * a.ld + rw.kk[r]];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const double hg = h[i0 + g];
#pragma unroll
        for (int r = 0; r < R; ++r)
            zr[r] -= hg * double(v[g][r]);
    }
}

// partials of v_i . z for i < nv, z = v_zslot.  in_cycle selects which
// flags make the launch a no-op.
template <class T>
__global__ __launch_bounds__(kBlock) void k_gm_dots(GmArgs<T> a, int nv, int zslot, int in_cycle)
{
    if (gm_idle(a.g, in_cycle != 0))
        return;
    __shared__ double lds[(kGmMaxRestart + 1) * (kBlock / kWave)];
    const GmRows rw(block_rows(a.n, gridDim.x, 1), a.n);
    const T* z = a.V + size_t(zslot) * a.ld;
    double zr[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        zr[r] = rw.in[r] ? double(z[rw.kk[r]]) : 0.0;
    gm_dots(a, nv, rw, zr, lds);
}

// z = v_zslot; z -= sum_i h_i v_i (i < nv), stored back; then either the
// partials of v_i . z (norm == 0) or of z . z (norm == 1).
template <class T>
__global__ __launch_bounds__(kBlock) void k_gm_orth(GmArgs<T> a, int nv, int zslot,
                                                    const double* __restrict__ h, int norm)
{
    if (gm_idle(a.g, true))
        return;
    __shared__ double lds[(kGmMaxRestart + 1) * (kBlock / kWave)];
    const GmRows rw(block_rows(a.n, gridDim.x, 1), a.n);
    T* z = a.V + size_t(zslot) * a.ld;
    double zr[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
        zr[r] = rw.in[r] ? double(z[rw.kk[r]]) : 0.0;
    int i = 0;
    for (; i + 4 <= nv; i += 4)
        gm_update_group<T, 4>(a, i, h, rw, zr);
    for (; i < nv; ++i)
        gm_update_group<T, 1>(a, i, h, rw, zr);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (rw.in[r]) {
            const T zt = T(zr[r]);
            z[rw.kk[r]] = zt;
            zr[r] = double(zt);
        } else {
            zr[r] = 0.0; // the update ran on the clamped row: drop it
        }
    }
    if (norm) {
        double s = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r)
            s += zr[r] * zr[r];
        s = wave_sum(s);
        if ((threadIdx.x & (kWave - 1)) == 0)
            lds[threadIdx.x / kWave] = s;
        gm_store_partials(lds, 1, a.partials);
        return;
    }
    gm_dots(a, nv, rw, zr, lds);
}

// out[i] = fold of partials[i * G ...], one block per value
__global__ __launch_bounds__(kBlock) void k_gm_fold(const GmState* g, const double* partials,
                                                    int G, double* out)
{
    if (gm_idle(g, true))
        return;
    double tot[1];
    fold_partials<1>(partials + size_t(blockIdx.x) * G, G, tot);
    if (threadIdx.x == 0)
        out[blockIdx.x] = tot[0];
}

__global__ void k_gm_init(GmState* g, const double* bnorm, double tol, int max_iter, int mixed)
{
    if (threadIdx.x != 0)
        return;
    GmState s{};
    s.k.bnorm = fmax(bnorm[0], 1e-30);
    s.k.tol = tol;
    s.k.max_iter = max_iter;
    s.k.done = max_iter <= 0 ? 2 : 0;
    s.cycle_end = 1;
    s.mixed = mixed;
    *g = s;
}

// beta = ||v_0||; the restart test of GMRESmSolver::solve
// (LinearSolvers.cpp:281-292): converged when beta < tol ||b||.
template <class T> __global__ __launch_bounds__(kBlock) void k_gm_cycle(GmArgs<T> a)
{
    GmState* g = a.g;
    if (gm_idle(g, false))
        return;
    double tot[1];
    fold_partials<1>(a.partials, a.G, tot);
    if (threadIdx.x != 0)
        return;
    const double beta = sqrt(tot[0]);
    g->j = 0;
    g->cycle_end = 1;
    if (beta < g->k.tol * g->k.bnorm) {
        g->k.last = beta / g->k.bnorm;
        g->k.done = 1;
        return;
    }
    if (!(beta == beta) || beta == 0.0) {
        g->k.done = 3; // NaN residual
        return;
    }
    if (g->k.iter >= g->k.max_iter) {
        g->k.done = 2;
        return;
    }
    a.sv[0] = beta;
    for (int i = 1; i <= a.m; ++i)
        a.sv[i] = 0.0;
    g->scale = 1.0 / beta;
    g->cycle_end = 0;
}

// v_slot *= scale
template <class T> __global__ __launch_bounds__(kBlock) void k_gm_scale(GmArgs<T> a, int slot)
{
    if (gm_idle(a.g, true))
        return;
    const double sc = a.g->scale;
    T* v = a.V + size_t(slot) * a.ld;
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int k = rr.beg + threadIdx.x; k < rr.end; k += kBlock)
        v[k] = T(double(v[k]) * sc);
}

// Hessenberg column j, Givens rotations and the residual estimate
// (LinearSolvers.cpp:303-326).  partials hold the block sums of ||z||^2.
template <class T> __global__ __launch_bounds__(kBlock) void k_gm_step(GmArgs<T> a, int j)
{
    GmState* g = a.g;
    if (gm_idle(g, true))
        return;
    double tot[1];
    fold_partials<1>(a.partials, a.G, tot);
    if (threadIdx.x != 0)
        return;
    double* Hj = a.H + size_t(j) * size_t(a.m + 1);
    for (int i = 0; i <= j; ++i)
        Hj[i] = a.h1[i] + a.h2[i];
    const double hn = sqrt(tot[0]);
    Hj[j + 1] = hn;
    g->scale = hn != 0.0 ? 1.0 / hn : 1.0;
    for (int i = 0; i < j; ++i) {
        const double t = a.cs[i] * Hj[i] + a.sn[i] * Hj[i + 1];
        Hj[i + 1] = -a.sn[i] * Hj[i] + a.cs[i] * Hj[i + 1];
        Hj[i] = t;
    }
    const double dd = hypot(Hj[j], Hj[j + 1]);
    if (!(dd > 0.0)) {
        // singular Hessenberg column: keep the steps before this one
        g->k.done = 3;
        g->cycle_end = 1;
        return;
    }
    a.cs[j] = Hj[j] / dd;
    a.sn[j] = Hj[j + 1] / dd;
    Hj[j] = dd;
    Hj[j + 1] = 0.0;
    const double t = a.cs[j] * a.sv[j];
    a.sv[j + 1] = -a.sn[j] * a.sv[j];
    a.sv[j] = t;
    const double last = fabs(a.sv[j + 1]) / g->k.bnorm;
    g->k.last = last;
    if (a.history && g->steps < a.hist_cap)
        a.history[g->steps] = last;
    ++g->steps;
    g->j = j + 1;
    if (last < g->k.tol) {
        // the reference leaves the step that met the tolerance out of its count
        g->cycle_end = 1;
        if (g->mixed)
            ++g->k.iter;
        else
            g->k.done = 1;
        return;
    }
    ++g->k.iter;
    if (j + 1 == a.m || g->k.iter >= g->k.max_iter)
        g->cycle_end = 1;
}

// y = H^-1 s over the cycle's steps (upper triangular after the rotations)
template <class T> __global__ void k_gm_backsolve(GmArgs<T> a)
{
    if (threadIdx.x != 0)
        return;
    const int jc = a.g->j;
    const size_t ldh = size_t(a.m + 1);
    for (int i = jc - 1; i >= 0; --i) {
        double sum = a.sv[i];
        for (int k = i + 1; k < jc; ++k)
            sum -= a.H[size_t(k) * ldh + i] * a.y[k];
        a.y[i] = sum / a.H[size_t(i) * ldh + i];
    }
}

// x += sum_{i < jc} y_i v_i, in the order of the reference's axpy loop
template <class T> __global__ __launch_bounds__(kBlock) void k_gm_xupdate(GmArgs<T> a)
{
    const int jc = a.g->j;
    if (jc == 0)
        return;
    const RowRange rr = block_rows(a.n, gridDim.x, 1);
    for (int k = rr.beg + threadIdx.x; k < rr.end; k += kBlock) {
        double xv = a.x[k];
        for (int i = 0; i < jc; ++i)
            xv += a.y[i] * double(a.V[size_t(i) * a.ld + k]);
        a.x[k] = xv;
    }
}

} // namespace

template <class T> int gm_init(const GmArgs<T>& a, const double* bnorm_dev, double tol,
                                int max_iter, bool mixed, hipStream_t s)
{
    if (a.m < 1 || a.m > kGmMaxRestart)
        return set_error(MI_ERR_BAD_ARG, "GMRES restart %d outside [1, %d]", a.m, kGmMaxRestart);
    k_gm_init<<<1, 1, 0, s>>>(a.g, bnorm_dev, tol, max_iter, mixed ? 1 : 0);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int gm_cycle_residual(const GmArgs<T>& a, hipStream_t s)
{
    (a.A64.soff ? k_gm_residual<T, true> : k_gm_residual<T, false>)<<<a.Gs, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int gm_cycle_begin(const GmArgs<T>& a, hipStream_t s)
{
    k_gm_dots<T><<<a.G, kBlock, 0, s>>>(a, 1, 0, 0);
    k_gm_cycle<T><<<1, kBlock, 0, s>>>(a);
    k_gm_scale<T><<<a.G, kBlock, 0, s>>>(a, 0);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int gm_step_spmv(const GmArgs<T>& a, int j, hipStream_t s)
{
    (a.A.soff ? k_gm_spmv<T, true> : k_gm_spmv<T, false>)<<<a.Gs, kBlock, 0, s>>>(a, j);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int gm_step_orth(const GmArgs<T>& a, int j, hipStream_t s)
{
    const int nv = j + 1, z = j + 1;
    k_gm_dots<T><<<a.G, kBlock, 0, s>>>(a, nv, z, 1);
    k_gm_fold<<<nv, kBlock, 0, s>>>(a.g, a.partials, a.G, a.h1);
    k_gm_orth<T><<<a.G, kBlock, 0, s>>>(a, nv, z, a.h1, 0);
    k_gm_fold<<<nv, kBlock, 0, s>>>(a.g, a.partials, a.G, a.h2);
    k_gm_orth<T><<<a.G, kBlock, 0, s>>>(a, nv, z, a.h2, 1);
    k_gm_step<T><<<1, kBlock, 0, s>>>(a, j);
    k_gm_scale<T><<<a.G, kBlock, 0, s>>>(a, z);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int gm_cycle_end(const GmArgs<T>& a, hipStream_t s)
{
    k_gm_backsolve<T><<<1, 64, 0, s>>>(a);
    k_gm_xupdate<T><<<a.G, kBlock, 0, s>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

#define MI_GM_INSTANTIATE(T)                                                                    \
    template int gm_init<T>(const GmArgs<T>&, const double*, double, int, bool, hipStream_t);   \
    template int gm_cycle_residual<T>(const GmArgs<T>&, hipStream_t);                           \
    template int gm_cycle_begin<T>(const GmArgs<T>&, hipStream_t);                              \
    template int gm_step_spmv<T>(const GmArgs<T>&, int, hipStream_t);                           \
    template int gm_step_orth<T>(const GmArgs<T>&, int, hipStream_t);                           \
    template int gm_cycle_end<T>(const GmArgs<T>&, hipStream_t);

MI_GM_INSTANTIATE(double)
MI_GM_INSTANTIATE(float)

} // namespace mi
