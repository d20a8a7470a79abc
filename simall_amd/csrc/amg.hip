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
        v = std::move(ev);
        sell.eperm.release();
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

template <class T> static int step(const DMat<T>& A, const T* din, T* dout, T* r, const T* invD,
                                   T* x, double beta, double alpha, int mode, Launcher& L)
{
    if (A.nr == 0)
        return MI_OK;
    return L(A.sell.built() ? k_cheb_step<T, true> : k_cheb_step<T, false>, A.grid(), kBlock,
             A.sm(), din, dout, r, invD, x, T(beta), T(alpha), mode);
}

template <class T> int DHierarchy<T>::build(const AmgHierarchyHost& H, int degree,
                                            int coarse_degree, double lmin_ratio, hipStream_t s,
                                            const double* cinv_device)
{
    const int nl = int(H.levels.size());
    graphs.clear();
    L.clear();
    L.resize(size_t(nl));
    deg = std::max(degree, 1);
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
        // Chebyshev coefficients on [lmin, lmax] of D^-1 A
        const double lmax = h.lmax > 0 ? h.lmax : 2.0;
        const double ratio = ldeg == 1 ? 0.5 : lmin_ratio;
        const double lmin = lmax * ratio;
        const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
        const double sigma = theta / delta;
        d.alpha.assign(size_t(ldeg), 0.0);
        d.beta.assign(size_t(ldeg), 0.0);
        d.alpha[0] = 1.0 / theta;
        double rho = 1.0 / sigma;
        for (int k = 1; k < ldeg; ++k) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            d.beta[size_t(k)] = rn * rho;
            d.alpha[size_t(k)] = 2.0 * rn / delta;
            rho = rn;
        }
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

// ------------------------------------------------- dense coarsest inverse
// Pivot step of column c (one workgroup): the row of largest |a_rc|, r >= c
// (the first on ties, c itself unless another is strictly larger), swapped
// into row c; a pivot below 1e-13 * amax replaced by +-1e-10 * amax; row c of
// a and inv scaled by 1 / pivot; column c saved in f (the factors).
__global__ __launch_bounds__(kBlock) void k_gj_pivot(int n, int c, double amax,
                                                     double* __restrict__ a,
                                                     double* __restrict__ inv,
                                                     double* __restrict__ f)
{
    __shared__ double sv[kBlock];
    __shared__ int si[kBlock];
    const int tid = threadIdx.x;
    const size_t N = size_t(n);
    double best = -1.0;
    int bi = n;
    for (int r = c + tid; r < n; r += kBlock) {
        const double v = fabs(a[size_t(r) * N + size_t(c)]);
        if (v > best) {
            best = v;
            bi = r;
        }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            const double v2 = sv[tid + w];
            const int i2 = si[tid + w];
            if (v2 > sv[tid] || (v2 == sv[tid] && i2 < si[tid])) {
                sv[tid] = v2;
                si[tid] = i2;
            }
        }
        __syncthreads();
    }
    int piv = si[0];
    if (piv >= n || !(sv[0] > fabs(a[size_t(c) * N + size_t(c)])))
        piv = c;
    __syncthreads();
    if (piv != c) {
        for (int k = tid; k < n; k += kBlock) {
            const size_t x = size_t(c) * N + size_t(k), y = size_t(piv) * N + size_t(k);
            const double t0 = a[x], t1 = inv[x];
            a[x] = a[y];
            inv[x] = inv[y];
            a[y] = t0;
            inv[y] = t1;
        }
    }
    __syncthreads();
    double p = a[size_t(c) * N + size_t(c)];
    if (fabs(p) <= 1e-13 * amax)
        p = (p >= 0 ? 1.0 : -1.0) * 1e-10 * (amax > 0 ? amax : 1.0);
    const double ip = 1.0 / p;
    __syncthreads();
    for (int k = tid; k < n; k += kBlock) {
        const size_t x = size_t(c) * N + size_t(k);
        a[x] = (k == c ? p : a[x]) * ip;
        inv[x] *= ip;
    }
    __syncthreads();
    for (int r = tid; r < n; r += kBlock)
        f[r] = a[size_t(r) * N + size_t(c)];
}

// Rows r != c with a nonzero factor: a[r][k] -= f_r a[c][k] (k >= c; the
// pivot row is zero left of c) and inv[r][k] -= f_r inv[c][k] (all k).
// Grid: x over the n - c + n columns, y over the rows.
__global__ __launch_bounds__(kBlock) void k_gj_eliminate(int n, int c,
                                                         const double* __restrict__ f,
                                                         double* __restrict__ a,
                                                         double* __restrict__ inv)
{
#pragma clang fp contract(off)
    const int r = int(blockIdx.y);
    const int ca = n - c;
    const int j = int(blockIdx.x) * kBlock + int(threadIdx.x);
    if (r == c || j >= ca + n)
        return;
    const double fr = f[r];
    if (fr == 0.0)
        return;
    const size_t N = size_t(n);
    if (j < ca) {
        const size_t k = size_t(c + j);
        a[size_t(r) * N + k] -= fr * a[size_t(c) * N + k];
    } else {
        const size_t k = size_t(j - ca);
        inv[size_t(r) * N + k] -= fr * inv[size_t(c) * N + k];
    }
}

__global__ __launch_bounds__(kBlock) void k_identity(int n, double* __restrict__ m)
{
    const size_t N = size_t(n);
    for (size_t q = size_t(blockIdx.x) * kBlock + threadIdx.x; q < N * N;
         q += size_t(gridDim.x) * kBlock)
        m[q] = (q / N == q % N) ? 1.0 : 0.0;
}

int dense_inverse_device(int n, double amax, double* a, double* inv, hipStream_t s)
{
    if (n <= 0)
        return MI_OK;
    if (n > 65535)
        return set_error(MI_ERR_BAD_ARG, "dense inverse of %d rows exceeds the grid", n);
    DevBuf<double> f;
    MI_TRY(f.alloc(size_t(n)));
    k_identity<<<grid_for(int64_t(n) * n, kBlock), kBlock, 0, s>>>(n, inv);
    MI_HIP(hipGetLastError());
    for (int c = 0; c < n; ++c) {
        k_gj_pivot<<<1, kBlock, 0, s>>>(n, c, amax, a, inv, f.p);
        const dim3 g(grid_for(int64_t(n - c) + n, kBlock), unsigned(n));
        k_gj_eliminate<<<g, kBlock, 0, s>>>(n, c, f.p, a, inv);
    }
    MI_HIP(hipGetLastError());
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

template <class T> int DHierarchy<T>::cycle(const T* b0, T* x0, Launcher& Lc)
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
        // r -= A e ; d = alpha0 D^-1 r ; x += d, then the remaining steps
        MI_TRY(step<T>(d.A, d.e.p, d.d.p, d.r.p, d.invD.p, x, 0.0, d.alpha[0], 0, Lc));
        T* din = d.d.p;
        T* dout = d.d2.p;
        for (int k = 1; k < d.deg; ++k) {
            MI_TRY(step<T>(d.A, din, dout, d.r.p, d.invD.p, x, d.beta[size_t(k)], d.alpha[size_t(k)],
                           0, Lc));
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
    AmgHierarchyHost host;
    DevBuf<double> cinv; // inverse of the coarsest operator (copied into each hierarchy)
    DHierarchy<double> h64;
    DHierarchy<float> h32;
    bool has64 = false, has32 = false;

    int setup(mi_csr* A, bool need64, bool need32, hipStream_t s) override
    {
        MI_TRY(A->sync_host());
        HCsr H;
        H.nr = H.nc = int(A->n());
        H.rp = A->pat->h_rp;
        H.ci = A->pat->h_ci;
        H.v = A->h_val;
        params.host_coarse_inverse = false;
        if (amg_build(H, params, host) != 0)
            return set_error(MI_ERR_UNSUPPORTED,
                             "AMG coarsening stalled above %d rows (no dense coarsest solve)",
                             kAmgMaxDense);
        return build_device(need64, need32, s);
    }
    const AmgHierarchyHost* hierarchy() const override
    {
        return (has64 || has32) ? &host : nullptr;
    }
    int install(const AmgHierarchyHost& H, bool need64, bool need32, hipStream_t s) override
    {
        host = H;
        host.coarse_inverse.clear();
        return build_device(need64, need32, s);
    }
    // Device hierarchies from `host`.
    int build_device(bool need64, bool need32, hipStream_t s)
    {
        // the coarsest inverse, once for both precisions, on the device
        const int nc = host.coarse_n;
        double amax = 0.0;
        for (int i = 0; i < nc; ++i)
            amax = std::max(amax, std::abs(host.coarse_dense[size_t(i) * size_t(nc) + size_t(i)]));
        DevBuf<double> dense;
        MI_TRY(dense.upload(host.coarse_dense.data(), host.coarse_dense.size(), s));
        MI_TRY(cinv.alloc(size_t(nc) * size_t(nc)));
        MI_TRY(dense_inverse_device(nc, amax, dense.p, cinv.p, s));
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
    p->degree = c.amg_sweeps > 0 ? c.amg_sweeps : 2;
    p->coarse_degree = c.amg_coarse_sweeps;
    return p;
}

} // namespace mi
