// amg.hip — smoothed-aggregation AMG V-cycle on the device.
//
// Solve-phase pieces of the hierarchy built by amg_setup.cpp, as gfx950
// kernels in T = double (FP64 solves) or float (inner iterations of the mixed
// solve; the coarsest solve stays FP64):
//   smoother     Chebyshev polynomial in D^-1 A of degree k (k = 1 is damped
//                Jacobi); each step is one fused kernel
//                r -= A d ; d' = beta d + alpha D^-1 r ; x += d'
//   restriction  b_c = R r        (R = P^T stored explicitly: gather, no atomics)
//   prolongation e = P x_c ; x += e
//   coarsest     x = A_c^-1 b with the dense FP64 inverse (one wavefront per row)
// The reference's V-cycle (src/solver/AMGPreconditioner.cpp:388-426) uses
// sequential symmetric Gauss-Seidel sweeps; the polynomial smoother is its
// data-parallel counterpart and keeps the cycle symmetric for CG.
#include "amg.hpp"
#include "krylov.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

namespace mi
{

// ------------------------------------------------------------------ kernels
template <class T>
__global__ __launch_bounds__(kBlock) void k_cheb_init(int n, const T* __restrict__ b,
                                                      const T* __restrict__ invD, T a0,
                                                      T* __restrict__ x, T* __restrict__ r,
                                                      T* __restrict__ d)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int i = lb * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const T bi = b[i];
    const T di = a0 * invD[i] * bi;
    r[i] = bi;
    d[i] = di;
    x[i] = di;
}

// mode 0: r -= A din ; dout = beta din + alpha D^-1 r ; x += dout
// mode 1: r -= A din
template <class T> struct EpiCheb
{
    const T* din;
    T* dout;
    T* r;
    const T* invD;
    T* x;
    T beta, alpha;
    int mode;
    T ri0 = 0, di = 0, id = 0, xi = 0;
    __device__ __forceinline__ void pre(int i)
    {
        ri0 = r[i];
        if (mode == 0) {
            di = din[i];
            id = invD[i];
            xi = x[i];
        }
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T ri = ri0 - T(s);
        r[i] = ri;
        if (mode == 0) {
            const T dn = beta * di + alpha * id * ri;
            dout[i] = dn;
            x[i] = xi + dn;
        }
    }
};

template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_cheb_step(StreamMat<T> A, const T* __restrict__ din,
                                                      T* __restrict__ dout, T* __restrict__ r,
                                                      const T* __restrict__ invD,
                                                      T* __restrict__ x, T beta, T alpha, int mode)
{
    EpiCheb<T> e{din, dout, r, invD, x, beta, alpha, mode};
    rows_of<kSell>(A, din, e);
}

// The first level-0 launch of a V-cycle fused with a CG iteration: the CG
// update of CGSolver::solve (src/solver/LinearSolvers.cpp:430-450; x +=
// alpha p, r -= alpha Ap, partials of r.r) and then, from the updated r,
// k_cheb_init's start (x0 = d = a0 D^-1 r, r0 = r).
template <class T>
__global__ __launch_bounds__(kBlock) void k_cheb_init_cg(int n, CgFuse<T> f,
                                                         const T* __restrict__ invD, T a0,
                                                         T* __restrict__ x0, T* __restrict__ r0,
                                                         T* __restrict__ d)
{
    if (f.st->done)
        return;
    const T al = T(f.st->alpha);
    const double al64 = f.st->alpha;
    const RowRange rr = block_rows(n, gridDim.x, 1);
    double v[1] = {0.0};
    for (int i = rr.beg + threadIdx.x; i < rr.end; i += kBlock) {
        if (f.xacc)
            f.xacc[i] += al64 * double(f.p[i]);
        else
            f.x[i] = f.x[i] + al * f.p[i];
        const T ri = f.r[i] - al * f.Ap[i];
        f.r[i] = ri;
        v[0] += double(ri) * double(ri);
        const T di = a0 * invD[i] * ri;
        r0[i] = ri;
        d[i] = di;
        x0[i] = di;
    }
    block_partials<1>(v, f.prr);
}

// The last level-0 launch of a fused V-cycle: a mode-0 Chebyshev step that
// also accumulates b0 . x (the CG's r.z, x being the cycle's result z).
template <class T> struct EpiChebDot : EpiCheb<T>
{
    const T* b0;
    T bi = 0;
    double acc = 0.0;
    __device__ __forceinline__ void pre(int i)
    {
        EpiCheb<T>::pre(i);
        bi = b0[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T ri = this->ri0 - T(s);
        this->r[i] = ri;
        const T dn = this->beta * this->di + this->alpha * this->id * ri;
        this->dout[i] = dn;
        const T xn = this->xi + dn;
        this->x[i] = xn;
        acc += double(bi) * double(xn);
    }
};

template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_cheb_step_dot(StreamMat<T> A, const T* __restrict__ din,
                                                          T* __restrict__ dout, T* __restrict__ r,
                                                          const T* __restrict__ invD,
                                                          T* __restrict__ x, T beta, T alpha,
                                                          const T* __restrict__ b0,
                                                          const KState* st, double* prz)
{
    if (st->done)
        return;
    EpiChebDot<T> e;
    e.din = din;
    e.dout = dout;
    e.r = r;
    e.invD = invD;
    e.x = x;
    e.beta = beta;
    e.alpha = alpha;
    e.mode = 0;
    e.b0 = b0;
    rows_of<kSell>(A, din, e);
    double v[1] = {e.acc};
    block_partials<1>(v, prz);
}

// e = P x_c ; x += e
template <class T> struct EpiProlong
{
    T* e;
    T* x;
    T xi = 0;
    __device__ __forceinline__ void pre(int i) { xi = x[i]; }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        e[i] = T(s);
        x[i] = xi + T(s);
    }
};

template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_prolong(StreamMat<T> P, const T* __restrict__ xc,
                                                    T* __restrict__ e, T* __restrict__ x)
{
    EpiProlong<T> ep{e, x};
    rows_of<kSell>(P, xc, ep);
}

// b_c = R r
template <class T> struct EpiStoreT
{
    T* y;
    __device__ __forceinline__ void pre(int) {}
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s) { y[i] = T(s); }
};

template <class T, bool kSell>
__global__ __launch_bounds__(kBlock) void k_restrict(StreamMat<T> R, const T* __restrict__ r,
                                                     T* __restrict__ bc)
{
    EpiStoreT<T> e{bc};
    rows_of<kSell>(R, r, e);
}

// x = Ainv b, dense FP64, one workgroup per row: the n_c workgroups keep
// four times the loads of one wavefront per row in flight, which the
// matrix (22 MB at n_c = 1661, out of cache within a V-cycle) needs.
template <class T>
__global__ __launch_bounds__(kBlock) void k_coarse_solve(int nc, const double* __restrict__ inv,
                                                         const T* __restrict__ b,
                                                         T* __restrict__ x)
{
    __shared__ double lds[kBlock / kWave];
    const int row = int(blockIdx.x);
    const double* a = inv + size_t(row) * size_t(nc);
    double s[1] = {0.0};
    for (int j = int(threadIdx.x); j < nc; j += kBlock)
        s[0] += a[j] * double(b[j]);
    block_sum<1>(s, lds);
    if (threadIdx.x == 0)
        x[row] = T(s[0]);
}

// ------------------------------------------------------------------ launch
// 1-D grid padded to a multiple of 8 blocks so that xcd_logical_block gives
// each XCD one contiguous slice of rows.
static unsigned rows_grid(int n, int lanes)
{
    unsigned g = grid_for(int64_t(n) * lanes, kBlock);
    return (g + 7) / 8 * 8;
}

template <class T> int DMat<T>::upload(const HCsr& h, hipStream_t s, bool diff_form)
{
    nr = h.nr;
    nc = h.nc;
    nnz = h.nnz();
    lanes = lanes_for(nr, nnz);
    MI_TRY(rp.upload(h.rp.data(), h.rp.size(), s));
    MI_TRY(ci.upload(h.ci.data(), h.ci.size(), s));
    std::vector<T> tv(h.v.begin(), h.v.end());
    std::vector<int> kd;
    if (diff_form && std::is_same<T, float>::value && nr == nc) {
        // off-diagonals rounded, diagonal slot = FP64 row sum rounded once
        kd.assign(static_cast<size_t>(nr), -1);
        for (int i = 0; i < nr; ++i) {
            double rs = 0.0;
            for (int k = h.rp[size_t(i)]; k < h.rp[size_t(i) + 1]; ++k) {
                rs += h.v[size_t(k)];
                if (h.ci[size_t(k)] == i && kd[size_t(i)] < 0)
                    kd[size_t(i)] = k;
            }
            if (kd[size_t(i)] < 0) {
                kd.clear(); // a row without its diagonal: keep the rounded values
                break;
            }
            tv[size_t(kd[size_t(i)])] = T(rs);
        }
        if (kd.empty())
            tv.assign(h.v.begin(), h.v.end());
    }
    if (!kd.empty())
        MI_TRY(dg.upload(kd.data(), kd.size(), s));
    MI_TRY(v.upload(tv.data(), tv.size(), s));
    // the hierarchy's operators need no column-order sums: long rows get lanes
    MI_TRY(rb.upload(nr, h.rp.data(), s, stream_lanes_per_row(nr, nnz, true))); // synchronizes
    MI_TRY(sell.build(nr, nnz, h.rp.data(), h.ci.data(), nr == nc, s));
    if (sell.built()) {
        // values kept only in SELL-64 slot order (sm() streams them)
        DevBuf<T> ev;
        MI_TRY(ev.alloc(size_t(sell.padded)));
        MI_TRY(launch_permute<T>(sell.padded, sell.eperm.p, v.p, ev.p, s));
        MI_HIP(hipStreamSynchronize(s));
        v = std::move(ev); // eperm stays for refresh
    } else if (nr > 0 && column_order_enabled()) {
        // values kept only in the column order of the row blocks (sm() streams them)
        MI_TRY(rb.sort_columns(rp.p, ci.p, nnz, s));
        if (rb.sorted()) {
            DevBuf<T> sv;
            MI_TRY(sv.alloc(size_t(nnz)));
            MI_TRY(launch_permute<T>(nnz, rb.sperm.p, v.p, sv.p, s));
            MI_HIP(hipStreamSynchronize(s));
            v = std::move(sv);
        }
    }
    return MI_OK;
}

// tmp = v with each row's diagonal slot (dg) holding the FP64 row sum: the
// difference form upload computes on the host.
__global__ __launch_bounds__(kBlock) void k_diff_rowsum(int n, const int* __restrict__ rp,
                                                        const double* __restrict__ v,
                                                        const int* __restrict__ dg,
                                                        double* __restrict__ tmp)
{
#pragma clang fp contract(off)
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double rs = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        rs += v[k];
        tmp[k] = v[k];
    }
    tmp[dg[i]] = rs;
}

// dst[m] = T(src[perm[m]]) (perm[m] = -1: a SELL padding slot, 0), or
// T(src[m]) without perm.
template <class T>
__global__ __launch_bounds__(kBlock) void k_permute_cvt(int64_t m, const int* __restrict__ perm,
                                                        const double* __restrict__ src,
                                                        T* __restrict__ dst)
{
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < m;
         t += int64_t(gridDim.x) * kBlock) {
        const int q = perm ? perm[t] : int(t);
        dst[t] = q >= 0 ? T(src[q]) : T(0);
    }
}

// invD_i = 1 / a_ii (the row's last nonzero diagonal entry), 1 without one.
template <class T>
__global__ __launch_bounds__(kBlock) void k_level_invd(int n, const int* __restrict__ rp,
                                                       const int* __restrict__ ci,
                                                       const double* __restrict__ v,
                                                       T* __restrict__ invD)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    T inv = T(1);
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i && v[k] != 0.0)
            inv = T(1.0 / v[k]);
    invD[i] = inv;
}

template <class T> int DMat<T>::refresh(const double* src, double* tmp, hipStream_t s)
{
    if (nnz == 0)
        return MI_OK;
    if (dg.n) {
        k_diff_rowsum<<<grid_for(nr, kBlock), kBlock, 0, s>>>(nr, rp.p, src, dg.p, tmp);
        src = tmp;
    }
    if (sell.built())
        k_permute_cvt<T><<<reduce_grid(sell.padded, 4), kBlock, 0, s>>>(sell.padded, sell.eperm.p,
                                                                        src, v.p);
    else if (rb.sorted())
        k_permute_cvt<T><<<reduce_grid(nnz, 4), kBlock, 0, s>>>(nnz, rb.sperm.p, src, v.p);
    else
        k_permute_cvt<T><<<reduce_grid(nnz, 4), kBlock, 0, s>>>(nnz, nullptr, src, v.p);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// Chebyshev step coefficients on [lmin, lmax] of D^-1 A, lmin = lmax * ratio
// (ratio 1/2 for degree 1: damped Jacobi).
void chebyshev(double lmax_est, int ldeg, double lmin_ratio, std::vector<double>& alpha,
               std::vector<double>& beta)
{
    const double lmax = lmax_est > 0 ? lmax_est : 2.0;
    const double ratio = ldeg == 1 ? 0.5 : lmin_ratio;
    const double lmin = lmax * ratio;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    alpha.assign(size_t(ldeg), 0.0);
    beta.assign(size_t(ldeg), 0.0);
    alpha[0] = 1.0 / theta;
    double rho = 1.0 / sigma;
    for (int k = 1; k < ldeg; ++k) {
        const double rn = 1.0 / (2.0 * sigma - rho);
        beta[size_t(k)] = rn * rho;
        alpha[size_t(k)] = 2.0 * rn / delta;
        rho = rn;
    }
}

template <class T> static int step(const DMat<T>& A, const T* din, T* dout, T* r, const T* invD,
                                   T* x, double beta, double alpha, int mode, Launcher& L)
{
    if (A.nr == 0)
        return MI_OK;
    return L(A.sell.built() ? k_cheb_step<T, true> : k_cheb_step<T, false>, A.grid(), kBlock,
             A.sm(), din, dout, r, invD, x, T(beta), T(alpha), mode);
}

template <class T>
static int step_dot(const DMat<T>& A, const T* din, T* dout, T* r, const T* invD, T* x,
                    double beta, double alpha, const CgFuse<T>& f, Launcher& L)
{
    return L(A.sell.built() ? k_cheb_step_dot<T, true> : k_cheb_step_dot<T, false>, A.grid(),
             kBlock, A.sm(), din, dout, r, invD, x, T(beta), T(alpha),
             static_cast<const T*>(f.r), static_cast<const KState*>(f.st), f.prz);
}

template <class T> int DHierarchy<T>::build(const AmgHierarchyHost& H, int degree,
                                            int coarse_degree, double lmin_ratio, hipStream_t s,
                                            const double* cinv_device)
{
    const int nl = int(H.levels.size());
    graphs.clear();
    cg_graphs.clear();
    L.clear();
    L.resize(size_t(nl));
    deg = std::max(degree, 1);
    this->lmin_ratio = lmin_ratio;
    for (int l = 0; l < nl; ++l) {
        const AmgLevelHost& h = H.levels[size_t(l)];
        DLevel<T>& d = L[size_t(l)];
        const int n = h.A.nr;
        const int ldeg = l == 0 || coarse_degree < 1 ? deg : coarse_degree;
        d.deg = ldeg;
        MI_TRY(d.A.upload(h.A, s, true));
        if (l + 1 < nl) {
            MI_TRY(d.P.upload(h.P, s));
            MI_TRY(d.R.upload(h.R, s));
        }
        std::vector<T> inv(size_t(n), T(1));
        for (int i = 0; i < n; ++i)
            for (int k = h.A.rp[size_t(i)]; k < h.A.rp[size_t(i) + 1]; ++k)
                if (h.A.ci[size_t(k)] == i && h.A.v[size_t(k)] != 0.0)
                    inv[size_t(i)] = T(1.0 / h.A.v[size_t(k)]);
        MI_TRY(d.invD.upload(inv.data(), inv.size(), s));
        MI_HIP(hipStreamSynchronize(s));
        for (DevBuf<T>* b : {&d.x, &d.b, &d.r, &d.d, &d.d2, &d.e})
            MI_TRY(b->alloc(size_t(std::max(n, 1))));
        chebyshev(h.lmax, ldeg, lmin_ratio, d.alpha, d.beta);
    }
    nc = H.coarse_n;
    if (cinv_device) {
        MI_TRY(cinv.alloc(size_t(nc) * size_t(nc)));
        if (nc > 0)
            MI_HIP(hipMemcpyAsync(cinv.p, cinv_device, size_t(nc) * size_t(nc) * sizeof(double),
                                  hipMemcpyDeviceToDevice, s));
    } else {
        MI_TRY(cinv.upload(H.coarse_inverse.data(), H.coarse_inverse.size(), s));
    }
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

// ------------------------------------------------------------- refresh
template <class T>
int DHierarchy<T>::refresh(const std::vector<LevelValues>& v, const double* cinv_device,
                           double* tmp, hipStream_t s)
{
    const int nl = int(L.size());
    if (int(v.size()) != nl)
        return set_error(MI_ERR_BAD_ARG, "refresh of %d levels for a hierarchy of %d",
                         int(v.size()), nl);
    graphs.clear(); // the captured cycles hold the old Chebyshev coefficients
    cg_graphs.clear();
    for (int l = 0; l < nl; ++l) {
        DLevel<T>& d = L[size_t(l)];
        const LevelValues& lv = v[size_t(l)];
        MI_TRY(d.A.refresh(lv.A, tmp, s));
        if (d.A.nr > 0)
            k_level_invd<T><<<grid_for(d.A.nr, kBlock), kBlock, 0, s>>>(d.A.nr, d.A.rp.p,
                                                                        d.A.ci.p, lv.A, d.invD.p);
        MI_HIP(hipGetLastError());
        if (l + 1 < nl) {
            MI_TRY(d.P.refresh(lv.P, tmp, s));
            MI_TRY(d.R.refresh(lv.R, tmp, s));
        }
        chebyshev(lv.lmax, d.deg, lmin_ratio, d.alpha, d.beta);
    }
    if (nc > 0)
        MI_HIP(hipMemcpyAsync(cinv.p, cinv_device, size_t(nc) * size_t(nc) * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    return MI_OK;
}

// ------------------------------------------------- dense coarsest inverse
// Blocked Gauss-Jordan on [A | I] with the host dense_inverse's per-entry
// operations.  Pivot step c of the host: the row piv >= c of largest |a_rc|
// (the first on ties, c unless another is strictly larger) is exchanged with
// row c; a pivot below 1e-13 * amax is replaced by +-1e-10 * amax; row c is
// scaled by ip = 1 / pivot; every row r != c with f_r = a_rc != 0 takes
// a_rk -= f_r a_ck (k >= c) and inv_rk -= f_r inv_ck (all k).
//
// Panel c0 .. c0 + nb - 1 (nb <= NB):
//   k_gj_panel   one workgroup holds columns c0.. of every row in registers
//                and performs the nb pivot steps on them, recording ip_j, the
//                exchanged row and f_r,j of every row (F, n x NB).  The rows
//                the panel touches (its own rows c0.. and the exchanged ones,
//                at most 2 NB: the set S) are listed in info.
//   k_gj_rows    per column right of the panel (and every inverse column):
//                the rows of S replay the nb steps (exchange, scale, update)
//                and the scaled pivot-row values u_j,k are stored (U).
//   k_gj_update  every other row r: a_rk -= f_r,j u_j,k for j = 0 .. nb-1,
//                skipping f_r,j = 0, as the host's nb steps do.
// Columns left of the panel are no longer read (the host leaves them at the
// values they reached); so are panel columns after their step.
constexpr int kGjThreads = 1024;

__global__ __launch_bounds__(kBlock) void k_gj_setup(int n, const double* __restrict__ a,
                                                     double* __restrict__ inv,
                                                     double* __restrict__ amax)
{
    // inv = I; amax = max |a_ii| (NaN diagonals ignored, as std::max does)
    const size_t N = size_t(n);
    for (size_t q = size_t(blockIdx.x) * kBlock + threadIdx.x; q < N * N;
         q += size_t(gridDim.x) * kBlock)
        inv[q] = (q / N == q % N) ? 1.0 : 0.0;
    if (blockIdx.x != 0)
        return;
    __shared__ double lds[kBlock];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += kBlock) {
        const double v = fabs(a[size_t(i) * N + size_t(i)]);
        m = (m < v) ? v : m;
    }
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double o = lds[threadIdx.x + w];
            lds[threadIdx.x] = (lds[threadIdx.x] < o) ? o : lds[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        amax[0] = lds[0];
}

template <int NB, int RPT>
__global__ __launch_bounds__(kGjThreads) void k_gj_panel(int n, int c0,
                                                         const double* __restrict__ amax_p,
                                                         const double* __restrict__ a,
                                                         double* __restrict__ F,
                                                         int* __restrict__ info,
                                                         double* __restrict__ ipv)
{
#pragma clang fp contract(off)
    constexpr int kWaves = kGjThreads / kWave;
    __shared__ double wbest[kWaves];
    __shared__ int wrow[kWaves];
    __shared__ double rowc[NB], rowp[NB];
    __shared__ double dabs;
    __shared__ int spiv[NB];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nb = min(NB, n - c0);
    const size_t N = size_t(n);
    const double amax = amax_p[0];
    double pv[RPT][NB];
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const int r = tid + q * kGjThreads;
#pragma unroll
        for (int j = 0; j < NB; ++j)
            pv[q][j] = (r < n && j < nb) ? a[size_t(r) * N + size_t(c0 + j)] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        if (j >= nb)
            break;
        const int c = c0 + j;
        // largest |a_rc| over r >= c, the first row on ties
        double best = -1.0;
        int bi = INT_MAX;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * kGjThreads;
            if (r >= c && r < n) {
                const double v = fabs(pv[q][j]);
                if (v > best) {
                    best = v;
                    bi = r;
                }
            }
        }
#pragma unroll
        for (int off = kWave / 2; off >= 1; off >>= 1) {
            const double ob = __shfl_xor(best, off, kWave);
            const int oi = __shfl_xor(bi, off, kWave);
            if (ob > best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        if (lane == 0) {
            wbest[wid] = best;
            wrow[wid] = bi;
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q)
            if (tid + q * kGjThreads == c)
                dabs = fabs(pv[q][j]);
        __syncthreads();
        double bb = wbest[0];
        int br = wrow[0];
        for (int w = 1; w < kWaves; ++w) {
            const double ob = wbest[w];
            const int oi = wrow[w];
            if (ob > bb || (ob == bb && oi < br)) {
                bb = ob;
                br = oi;
            }
        }
        const int piv = (br < n && bb > dabs) ? br : c;
        // exchange rows c and piv (panel columns >= j)
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * kGjThreads;
            if (r == c)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (k >= j)
                        rowc[k] = pv[q][k];
            if (r == piv)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (k >= j)
                        rowp[k] = pv[q][k];
        }
        __syncthreads();
        double p = rowp[j];
        if (fabs(p) <= 1e-13 * amax)
            p = (p >= 0 ? 1.0 : -1.0) * 1e-10 * (amax > 0 ? amax : 1.0);
        const double ip = 1.0 / p;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * kGjThreads;
            if (r == piv && piv != c)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (k >= j)
                        pv[q][k] = rowc[k];
            if (r == c)
#pragma unroll
                for (int k = 0; k < NB; ++k)
                    if (k >= j)
                        pv[q][k] = (k == j ? p : rowp[k]) * ip;
        }
        // factors of column c and the updates of the panel's later columns
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * kGjThreads;
            if (r < n) {
                const double f = r == c ? 0.0 : pv[q][j];
                F[size_t(r) * NB + size_t(j)] = f;
                if (f != 0.0)
#pragma unroll
                    for (int k = 0; k < NB; ++k)
                        if (k > j && k < nb)
                            pv[q][k] -= f * (rowp[k] * ip);
            }
        }
        if (tid == 0) {
            ipv[j] = ip;
            spiv[j] = piv;
        }
    }
    if (tid == 0) {
        // S: slots 0 .. nb-1 are rows c0 .. c0+nb-1, then the other exchanged rows
        int* pslot = info + 1;
        int* slist = info + 1 + NB;
        int cnt = nb;
        for (int j = 0; j < nb; ++j)
            slist[j] = c0 + j;
        for (int j = 0; j < nb; ++j) {
            const int pr = spiv[j];
            int sl = -1;
            if (pr < c0 + nb) {
                sl = pr - c0;
            } else {
                for (int t = nb; t < cnt; ++t)
                    if (slist[t] == pr)
                        sl = t;
                if (sl < 0) {
                    sl = cnt;
                    slist[cnt++] = pr;
                }
            }
            pslot[j] = sl;
        }
        info[0] = cnt;
    }
}

// Column q of the panel's right part: q < ca is column c0 + nb + q of a,
// q >= ca column q - ca of inv.
template <int NB>
__global__ __launch_bounds__(kBlock) void k_gj_rows(int n, int c0, const int* __restrict__ info,
                                                    const double* __restrict__ ipv,
                                                    const double* __restrict__ F,
                                                    double* __restrict__ a,
                                                    double* __restrict__ inv,
                                                    double* __restrict__ U)
{
#pragma clang fp contract(off)
    __shared__ double vals[2 * NB][kBlock];
    const int nb = min(NB, n - c0);
    const int ca = n - c0 - nb;
    const int ncol = ca + n;
    const int q = int(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= ncol)
        return;
    double* M = q < ca ? a : inv;
    const size_t k = q < ca ? size_t(c0 + nb + q) : size_t(q - ca);
    const size_t N = size_t(n);
    const int cnt = info[0];
    const int* pslot = info + 1;
    const int* slist = info + 1 + NB;
    double* v = &vals[0][threadIdx.x];
    for (int t = 0; t < cnt; ++t)
        v[t * kBlock] = M[size_t(slist[t]) * N + k];
    for (int j = 0; j < nb; ++j) {
        const int ps = pslot[j];
        if (ps != j) {
            const double x = v[j * kBlock];
            v[j * kBlock] = v[ps * kBlock];
            v[ps * kBlock] = x;
        }
        const double u = v[j * kBlock] * ipv[j];
        v[j * kBlock] = u;
        U[size_t(j) * size_t(ncol) + size_t(q)] = u;
        for (int t = 0; t < cnt; ++t) {
            if (t == j)
                continue;
            const double f = F[size_t(slist[t]) * NB + size_t(j)];
            if (f != 0.0)
                v[t * kBlock] -= f * u;
        }
    }
    for (int t = 0; t < cnt; ++t)
        M[size_t(slist[t]) * N + k] = v[t * kBlock];
}

// Rows outside S: blocks of 64 columns (one per lane) by kGjUpdateRows rows
// (the block's 4 wavefronts take every 4th row).
constexpr int kGjUpdateRows = 64;

template <int NB>
__global__ __launch_bounds__(kBlock) void k_gj_update(int n, int c0, const int* __restrict__ info,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ U,
                                                      double* __restrict__ a,
                                                      double* __restrict__ inv)
{
#pragma clang fp contract(off)
    const int nb = min(NB, n - c0);
    const int ca = n - c0 - nb;
    const int ncol = ca + n;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int q = int(blockIdx.x) * kWave + lane;
    if (q >= ncol)
        return;
    double* M = q < ca ? a : inv;
    const size_t k = q < ca ? size_t(c0 + nb + q) : size_t(q - ca);
    const size_t N = size_t(n);
    const int cnt = info[0];
    const int* slist = info + 1 + NB;
    double u[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
        u[j] = j < nb ? U[size_t(j) * size_t(ncol) + size_t(q)] : 0.0;
    const int r0 = int(blockIdx.y) * kGjUpdateRows;
    const int r1 = min(n, r0 + kGjUpdateRows);
    for (int r = r0 + wid; r < r1; r += kBlock / kWave) {
        bool in_s = r >= c0 && r < c0 + nb;
        for (int t = nb; t < cnt && !in_s; ++t)
            in_s = slist[t] == r;
        if (in_s)
            continue;
        const double* fr = F + size_t(r) * NB;
        double x = M[size_t(r) * N + k];
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j < nb) {
                const double f = fr[j];
                if (f != 0.0)
                    x -= f * u[j];
            }
        }
        M[size_t(r) * N + k] = x;
    }
}

template <int NB, int RPT>
static int gj_run(int n, double* a, double* inv, DenseInverse& w, hipStream_t s)
{
    MI_TRY(w.F.alloc(size_t(n) * NB));
    MI_TRY(w.U.alloc(size_t(NB) * size_t(2 * n)));
    MI_TRY(w.ipv.alloc(NB));
    MI_TRY(w.amax.alloc(1));
    MI_TRY(w.info.alloc(1 + 3 * NB));
    k_gj_setup<<<grid_for(int64_t(n) * n, kBlock), kBlock, 0, s>>>(n, a, inv, w.amax.p);
    for (int c0 = 0; c0 < n; c0 += NB) {
        const int nb = std::min(NB, n - c0);
        const int ncol = (n - c0 - nb) + n;
        k_gj_panel<NB, RPT><<<1, kGjThreads, 0, s>>>(n, c0, w.amax.p, a, w.F.p, w.info.p, w.ipv.p);
        k_gj_rows<NB><<<grid_for(ncol, kBlock), kBlock, 0, s>>>(n, c0, w.info.p, w.ipv.p, w.F.p, a,
                                                                inv, w.U.p);
        const dim3 g(grid_for(ncol, kWave), grid_for(n, kGjUpdateRows));
        k_gj_update<NB><<<g, kBlock, 0, s>>>(n, c0, w.info.p, w.F.p, w.U.p, a, inv);
    }
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int DenseInverse::run(int n, double* a, double* inv, hipStream_t s)
{
    if (n <= 0)
        return MI_OK;
    if (n > kAmgMaxDense)
        return set_error(MI_ERR_BAD_ARG, "dense inverse of %d rows (at most %d)", n, kAmgMaxDense);
    static_assert(kAmgMaxDense <= 4 * kGjThreads, "panel rows per thread");
    if (n <= kGjThreads)
        MI_TRY((gj_run<16, 1>(n, a, inv, *this, s)));
    else if (n <= 2 * kGjThreads)
        MI_TRY((gj_run<8, 2>(n, a, inv, *this, s)));
    else
        MI_TRY((gj_run<8, 4>(n, a, inv, *this, s)));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

template <class T> int DHierarchy<T>::vcycle(const T* b0, T* x0, hipStream_t s, bool use_graph)
{
    if (use_graph)
        return graphs.run(b0, x0, s, [&](Launcher& L) -> int { return cycle(b0, x0, L); });
    Launcher L{s};
    return cycle(b0, x0, L);
}

template <class T> int DHierarchy<T>::vcycle_cg(const CgFuse<T>& f, T* x0, hipStream_t s,
                                                bool use_graph, unsigned* grr, unsigned* grz)
{
    if (L.size() < 2 || L[0].A.nr == 0)
        return set_error(MI_ERR_UNSUPPORTED, "fused CG V-cycle needs two levels or more");
    *grr = range_grid(L[0].A.nr, 256);
    *grz = L[0].A.grid();
    if (use_graph) {
        const void* key[GraphCache::kKey] = {f.r, x0, f.p, f.Ap, f.x, f.xacc, f.st, f.prr, f.prz};
        return cg_graphs.run_keyed(key, s, [&](Launcher& Lc) -> int { return cycle(f.r, x0, Lc, &f); });
    }
    Launcher Lc{s};
    return cycle(f.r, x0, Lc, &f);
}

template <class T> int DHierarchy<T>::cycle(const T* b0, T* x0, Launcher& Lc, const CgFuse<T>* cf)
{
    const int nl = int(L.size());
    auto bptr = [&](int l) -> const T* { return l == 0 ? b0 : L[size_t(l)].b.p; };
    auto xptr = [&](int l) -> T* { return l == 0 ? x0 : L[size_t(l)].x.p; };
    // downward leg: pre-smooth from x = 0, restrict the residual
    for (int l = 0; l + 1 < nl; ++l) {
        DLevel<T>& d = L[size_t(l)];
        const int n = d.A.nr;
        const T* b = bptr(l);
        T* x = xptr(l);
        if (cf && l == 0)
            MI_TRY(Lc(k_cheb_init_cg<T>, range_grid(n, 256), kBlock, n, *cf,
                      static_cast<const T*>(d.invD.p), T(d.alpha[0]), x, d.r.p, d.d.p));
        else
            MI_TRY(Lc(k_cheb_init<T>, rows_grid(n, 1), kBlock, n, b, d.invD.p, T(d.alpha[0]), x,
                      d.r.p, d.d.p));
        T* din = d.d.p;
        T* dout = d.d2.p;
        for (int k = 1; k < d.deg; ++k) {
            MI_TRY(step<T>(d.A, din, dout, d.r.p, d.invD.p, x, d.beta[size_t(k)], d.alpha[size_t(k)],
                           0, Lc));
            std::swap(din, dout);
        }
        MI_TRY(step<T>(d.A, din, nullptr, d.r.p, d.invD.p, x, 0.0, 0.0, 1, Lc));
        if (d.R.nr > 0)
            MI_TRY(Lc(d.R.sell.built() ? k_restrict<T, true> : k_restrict<T, false>, d.R.grid(),
                      kBlock, d.R.sm(), static_cast<const T*>(d.r.p), L[size_t(l) + 1].b.p));
    }
    // coarsest level
    {
        const T* b = bptr(nl - 1);
        T* x = xptr(nl - 1);
        if (nc > 0)
            MI_TRY(Lc(k_coarse_solve<T>, unsigned(nc), kBlock, nc,
                      static_cast<const double*>(cinv.p), b, x));
    }
    // upward leg: prolongate, post-smooth
    for (int l = nl - 2; l >= 0; --l) {
        DLevel<T>& d = L[size_t(l)];
        const int n = d.A.nr;
        T* x = xptr(l);
        const T* xc = xptr(l + 1);
        if (n > 0)
            MI_TRY(Lc(d.P.sell.built() ? k_prolong<T, true> : k_prolong<T, false>, d.P.grid(),
                      kBlock, d.P.sm(), xc, d.e.p, x));
        // r -= A e ; d = alpha0 D^-1 r ; x += d, then the remaining steps; the
        // cycle's last launch (level 0, last step) of a fused CG iteration
        // also stores the r.z partials
        const bool dot = cf && l == 0;
        if (dot && d.deg == 1)
            MI_TRY(step_dot<T>(d.A, d.e.p, d.d.p, d.r.p, d.invD.p, x, 0.0, d.alpha[0], *cf, Lc));
        else
            MI_TRY(step<T>(d.A, d.e.p, d.d.p, d.r.p, d.invD.p, x, 0.0, d.alpha[0], 0, Lc));
        T* din = d.d.p;
        T* dout = d.d2.p;
        for (int k = 1; k < d.deg; ++k) {
            if (dot && k + 1 == d.deg)
                MI_TRY(step_dot<T>(d.A, din, dout, d.r.p, d.invD.p, x, d.beta[size_t(k)],
                                   d.alpha[size_t(k)], *cf, Lc));
            else
                MI_TRY(step<T>(d.A, din, dout, d.r.p, d.invD.p, x, d.beta[size_t(k)],
                               d.alpha[size_t(k)], 0, Lc));
            std::swap(din, dout);
        }
    }
    return MI_OK;
}

template struct DHierarchy<double>;
template struct DHierarchy<float>;
template struct DMat<double>;
template struct DMat<float>;

// --------------------------------------------------------------- precond
namespace
{

struct AmgPrecond final : Precond
{
    AmgParams params;
    int degree = 2;
    int coarse_degree = 0;
    double lmin_ratio = 0.3;
    // the hierarchy of the last full setup; after a numeric re-setup its
    // values are those of that setup until sync_host downloads the new ones
    mutable AmgHierarchyHost host;
    mutable bool host_stale = false;
    hipStream_t last_stream = nullptr;
    DevBuf<double> cinv; // inverse of the coarsest operator (copied into each hierarchy)
    DenseInverse gj;
    DHierarchy<double> h64;
    DHierarchy<float> h32;
    bool has64 = false, has32 = false;
    std::shared_ptr<Pattern> built_pat;   // pattern of the last full setup
    std::unique_ptr<AmgNumeric> num;      // re-setup data, recorded on first use
    DevBuf<double> tmp;                   // scratch of the refreshes
    bool num_failed = false;

    int setup(mi_csr* A, bool need64, bool need32, hipStream_t s) override
    {
        MI_TRY(A->sync_host());
        HCsr H;
        H.nr = H.nc = int(A->n());
        H.rp = A->pat->h_rp;
        H.ci = A->pat->h_ci;
        H.v = A->h_val;
        params.host_coarse_inverse = false;
        num.reset();
        num_failed = false;
        host_stale = false;
        built_pat.reset();
        last_stream = s;
        if (amg_build(H, params, host) != 0)
            return set_error(MI_ERR_UNSUPPORTED,
                             "AMG coarsening stalled above %d rows (no dense coarsest solve)",
                             kAmgMaxDense);
        MI_TRY(build_device(need64, need32, s));
        built_pat = A->pat;
        return MI_OK;
    }
    bool can_resetup(const mi_csr* A, bool need64, bool need32) const override
    {
        if (num_failed || !built_pat || A->pat != built_pat || need64 != has64 || need32 != has32)
            return false;
        for (size_t l = 0; l + 1 < host.levels.size(); ++l)
            if (host.levels[l].agg.empty() || host.levels[l].strong.empty())
                return false;
        return true;
    }
    int resetup(mi_csr* A, bool need64, bool need32, hipStream_t s) override
    {
        (void)need64;
        (void)need32;
        last_stream = s;
        if (!num) {
            auto nm = std::make_unique<AmgNumeric>();
            const int rc = nm->record(host, A->pat->rp.p, A->pat->ci.p, s);
            if (rc != MI_OK) {
                // a hierarchy the numeric phase cannot follow: full setups from now on
                clear_error();
                const int rc2 = setup(A, need64, need32, s);
                num_failed = true;
                return rc2;
            }
            MI_TRY(tmp.alloc(nm->max_nnz()));
            num = std::move(nm);
        }
        MI_TRY(num->run(A->val.p, params.omega_scale, params.truncation > 0.0, s));
        const int nc = host.coarse_n;
        MI_TRY(gj.run(nc, num->dense.p, cinv.p, s));
        std::vector<LevelValues> v(host.levels.size());
        for (size_t l = 0; l < v.size(); ++l)
            v[l] = num->values(int(l));
        if (has64)
            MI_TRY(h64.refresh(v, cinv.p, tmp.p, s));
        if (has32)
            MI_TRY(h32.refresh(v, cinv.p, tmp.p, s));
        MI_HIP(hipStreamSynchronize(s));
        host_stale = true;
        return MI_OK;
    }
    // Brings `host` up to the device's values after numeric re-setups.
    int sync_host() const
    {
        if (!host_stale)
            return MI_OK;
        MI_TRY(num->download(host, last_stream));
        host_stale = false;
        return MI_OK;
    }
    const AmgHierarchyHost* hierarchy() const override
    {
        if (!(has64 || has32) || sync_host() != MI_OK)
            return nullptr;
        return &host;
    }
    int host_copy(AmgHierarchyHost& out, hipStream_t s) override
    {
        if (!(has64 || has32))
            return set_error(MI_ERR_UNSUPPORTED, "no AMG hierarchy has been built");
        MI_TRY(sync_host());
        out = host;
        const size_t nn = size_t(host.coarse_n) * size_t(host.coarse_n);
        out.coarse_inverse.resize(nn);
        if (nn)
            MI_HIP(hipMemcpyAsync(out.coarse_inverse.data(), cinv.p, nn * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        return MI_OK;
    }
    int install(const AmgHierarchyHost& H, bool need64, bool need32, hipStream_t s) override
    {
        host = H;
        host.coarse_inverse.clear();
        num.reset();
        host_stale = false;
        built_pat.reset();
        last_stream = s;
        return build_device(need64, need32, s);
    }
    // Device hierarchies from `host`.
    int build_device(bool need64, bool need32, hipStream_t s)
    {
        // the coarsest inverse, once for both precisions, on the device
        const int nc = host.coarse_n;
        DevBuf<double> dense;
        MI_TRY(dense.upload(host.coarse_dense.data(), host.coarse_dense.size(), s));
        MI_TRY(cinv.alloc(size_t(nc) * size_t(nc)));
        MI_TRY(gj.run(nc, dense.p, cinv.p, s));
        has64 = has32 = false;
        if (need64) {
            MI_TRY(h64.build(host, degree, coarse_degree, lmin_ratio, s, cinv.p));
            has64 = true;
        }
        if (need32) {
            MI_TRY(h32.build(host, degree, coarse_degree, lmin_ratio, s, cinv.p));
            has32 = true;
        }
        return MI_OK;
    }
    int apply64(const double* r, double* z, hipStream_t s) override
    {
        if (!has64)
            return set_error(MI_ERR_UNSUPPORTED, "AMG hierarchy has no FP64 copy");
        return h64.vcycle(r, z, s, use_graph);
    }
    int apply32(const float* r, float* z, hipStream_t s) override
    {
        if (!has32)
            return set_error(MI_ERR_UNSUPPORTED, "AMG hierarchy has no FP32 copy");
        return h32.vcycle(r, z, s, use_graph);
    }
    bool cg_fusable(bool fp64) const override
    {
        const size_t nl = fp64 ? (has64 ? h64.L.size() : 0) : (has32 ? h32.L.size() : 0);
        return nl >= 2;
    }
    int apply_cg64(const CgFuse<double>& f, double* z, unsigned* grr, unsigned* grz,
                   hipStream_t s) override
    {
        if (!has64)
            return set_error(MI_ERR_UNSUPPORTED, "AMG hierarchy has no FP64 copy");
        return h64.vcycle_cg(f, z, s, use_graph, grr, grz);
    }
    int apply_cg32(const CgFuse<float>& f, float* z, unsigned* grr, unsigned* grz,
                   hipStream_t s) override
    {
        if (!has32)
            return set_error(MI_ERR_UNSUPPORTED, "AMG hierarchy has no FP32 copy");
        return h32.vcycle_cg(f, z, s, use_graph, grr, grz);
    }
    const char* name() const override { return "sa-amg"; }
    int levels() const override { return int(host.levels.size()); }
    double complexity() const override { return host.operator_complexity; }
};

} // namespace

std::unique_ptr<Precond> make_amg(const mi_solver_cfg& c)
{
    auto p = std::make_unique<AmgPrecond>();
    if (c.amg_max_levels > 0)
        p->params.max_levels = c.amg_max_levels;
    if (c.amg_coarse_size > 0)
        p->params.coarse_size = c.amg_coarse_size;
    if (c.amg_strength >= 0)
        p->params.theta = c.amg_strength;
    if (c.amg_jacobi_omega > 0)
        p->params.omega_scale = c.amg_jacobi_omega;
    if (c.amg_truncation >= 0)
        p->params.truncation = c.amg_truncation;
    p->degree = c.amg_sweeps > 0 ? c.amg_sweeps : 2;
    p->coarse_degree = c.amg_coarse_sweeps;
    return p;
}

} // namespace mi
