// ilu0.hip — ILU(0) preconditioner: host factorization on A's pattern and
// level-scheduled triangular solves on the device.
//
// Factorization: the IKJ elimination of ILU0Preconditioner::setup
// (src/solver/LinearSolvers.cpp:81-126; also GpuKernels' ilu0_factor_inplace,
// src/gpu/GpuKernels.cpp:345-388) — rows in order, updates restricted to the
// pattern, rows whose pivot is zero skipped — run once per value change on
// the host, then kept resident.
// Apply (src/solver/LinearSolvers.cpp:127-151): the unit-lower forward sweep
// and the upper backward sweep.  A row of L depends only on rows in earlier
// dependency levels, so each level is one launch with one thread per row;
// within a row the products are accumulated in column order with contraction
// off, which reproduces the serial sweeps' arithmetic exactly.  Orderings
// whose rows form a few long runs of mutually independent rows (multicolour)
// sweep run by run instead, each run one CSR-stream launch over the split
// strictly lower or upper factor; there a row's products are summed first and
// then subtracted, a different rounding of the same sweep.
#include "ilu0.hpp"
#include "stream.hpp"

#include <algorithm>
#include <cstring>

namespace mi
{

#pragma clang fp contract(off)

template <class T>
__global__ __launch_bounds__(kBlock) void k_lower_level(int cnt, const int* __restrict__ rows,
                                                        const int* __restrict__ rp,
                                                        const int* __restrict__ lend,
                                                        const int* __restrict__ ci,
                                                        const T* __restrict__ lu,
                                                        const T* __restrict__ r,
                                                        T* __restrict__ y)
{
    const int t = int(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= cnt)
        return;
    const int i = rows[t];
    T s = r[i];
    for (int k = rp[i]; k < lend[i]; ++k)
        s -= lu[k] * y[ci[k]];
    y[i] = s;
}

// A row without a stored diagonal is not divided.  With `prefix` its sum
// runs, as ILU0Preconditioner::apply's loop from diagPtr_[i] + 1 = 0 does
// (src/solver/LinearSolvers.cpp:146-150), over every position before the end
// of the row -- rows 0..i of the factor -- where the columns <= i meet the
// still-zero entries of the reference's z; without it (the GpuKernels facade,
// src/gpu/GpuKernels.cpp:466-477) over the row's own upper entries.
template <class T>
__global__ __launch_bounds__(kBlock) void k_upper_level(int cnt, const int* __restrict__ rows,
                                                        const int* __restrict__ rp,
                                                        const int* __restrict__ diag,
                                                        const int* __restrict__ lend,
                                                        const int* __restrict__ ci,
                                                        const T* __restrict__ lu,
                                                        const T* __restrict__ y,
                                                        T* __restrict__ z, int prefix)
{
    const int t = int(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= cnt)
        return;
    const int i = rows[t];
    const int kd = diag[i];
    T s = y[i];
    if (kd >= 0) {
        for (int k = kd + 1; k < rp[i + 1]; ++k)
            s -= lu[k] * z[ci[k]];
        const T d = lu[kd];
        z[i] = (d != T(0)) ? s / d : s;
        return;
    }
    for (int k = prefix ? 0 : lend[i]; k < rp[i + 1]; ++k) {
        const int j = ci[k];
        s -= lu[k] * (j > i ? z[j] : T(0));
    }
    z[i] = s;
}

// y_i = r_i - sum_{j < i} l_ij y_j over the rows of one run
template <class T> struct EpiIluLower
{
    const T* r;
    T* y;
    T ri = T(0);
    __device__ __forceinline__ void pre(int i) { ri = r[i]; }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s) { y[i] = ri - T(s); }
};

// z_i = (y_i - sum_{j > i} u_ij z_j) / u_ii over the rows of one run
template <class T> struct EpiIluUpper
{
    const T* y;
    const T* d;
    T* z;
    T yi = T(0), di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        yi = y[i];
        di = d[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T t = yi - T(s);
        z[i] = (di != T(0)) ? t / di : t;
    }
};

// z_i = (r_i - sum_j m_ij x_j) / u_ii: a row whose other triangle is empty
// (y_i = r_i - sum l_ij y_j with no U entries, or y_i = r_i with no L
// entries), in the operations of the two sweeps
template <class T> struct EpiIluRow
{
    const T* r;
    const T* d;
    T* z;
    T ri = T(0), di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        ri = r[i];
        di = d[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        const T t = ri - T(s);
        z[i] = (di != T(0)) ? t / di : t;
    }
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_ilu_two_run(StreamMat<T> M, const T* x,
                                                        const T* __restrict__ r,
                                                        const T* __restrict__ d, T* z)
{
    EpiIluRow<T> e{r, d, z};
    stream_rows(M, x, e);
}

// SSOR with omega = 1 (SsorPreconditioner, src/solver/LinearSolvers.cpp:163-236,
// as make_preconditioner builds it: :247-248): forward
// y_i = (r_i - sum_{j<i} a_ij y_j) / a_ii, then z_i = (a_ii y_i -
// sum_{j>i} a_ij z_j) / a_ii, each division a multiplication by the stored
// 1/a_ii (1 for a zero diagonal).  With omega = 1 the reference's scalings by
// omega and 2 - omega are exact and drop out.
template <class T>
__global__ __launch_bounds__(kBlock) void k_ssor_lower_level(int cnt, const int* __restrict__ rows,
                                                             const int* __restrict__ rp,
                                                             const int* __restrict__ lend,
                                                             const int* __restrict__ ci,
                                                             const T* __restrict__ a,
                                                             const T* __restrict__ invD,
                                                             const T* __restrict__ r,
                                                             T* __restrict__ y)
{
    const int t = int(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= cnt)
        return;
    const int i = rows[t];
    T s = r[i];
    for (int k = rp[i]; k < lend[i]; ++k)
        s -= a[k] * y[ci[k]];
    y[i] = s * invD[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ssor_upper_level(int cnt, const int* __restrict__ rows,
                                                             const int* __restrict__ rp,
                                                             const int* __restrict__ diag,
                                                             const int* __restrict__ lend,
                                                             const int* __restrict__ ci,
                                                             const T* __restrict__ a,
                                                             const T* __restrict__ D,
                                                             const T* __restrict__ invD,
                                                             const T* __restrict__ y,
                                                             T* __restrict__ z)
{
    const int t = int(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= cnt)
        return;
    const int i = rows[t];
    T s = D[i] * y[i];
    // the row's entries right of the diagonal (lend: first column >= i, the
    // start when no diagonal is stored)
    for (int k = diag[i] >= 0 ? diag[i] + 1 : lend[i]; k < rp[i + 1]; ++k)
        s -= a[k] * z[ci[k]];
    z[i] = s * invD[i];
}

template <class T> struct EpiSsorLower
{
    const T* r;
    const T* invD;
    T* y;
    T ri = T(0), di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        ri = r[i];
        di = invD[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s) { y[i] = (ri - T(s)) * di; }
};

template <class T> struct EpiSsorUpper
{
    const T* y;
    const T* D;
    const T* invD;
    T* z;
    T yi = T(0), Di = T(1), di = T(1);
    __device__ __forceinline__ void pre(int i)
    {
        yi = y[i];
        Di = D[i];
        di = invD[i];
    }
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s)
    {
        z[i] = (Di * yi - T(s)) * di;
    }
};

template <class T>
__global__ __launch_bounds__(kBlock) void k_ssor_lower_run(StreamMat<T> L, const T* __restrict__ r,
                                                           const T* __restrict__ invD, T* y)
{
    EpiSsorLower<T> e{r, invD, y};
    stream_rows(L, y, e);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ssor_upper_run(StreamMat<T> U, const T* __restrict__ y,
                                                           const T* __restrict__ D,
                                                           const T* __restrict__ invD, T* z)
{
    EpiSsorUpper<T> e{y, D, invD, z};
    stream_rows(U, z, e);
}

// D = a_ii, invD = 1 / a_ii (1 when zero); a row without a stored diagonal
// has D = invD = 1 (diag_ and invDiag_ keep their initial 1:
// src/solver/LinearSolvers.cpp:174-184)
template <class T>
__global__ __launch_bounds__(kBlock) void k_ssor_diag(int n, const int* __restrict__ diag,
                                                      const double* __restrict__ a,
                                                      T* __restrict__ D, T* __restrict__ invD)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const double d = diag[i] >= 0 ? a[diag[i]] : 1.0;
    D[i] = T(d);
    invD[i] = T(d != 0.0 ? 1.0 / d : 1.0);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_to_t(int64_t n, const double* __restrict__ a,
                                                 T* __restrict__ b)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        b[i] = T(a[i]);
}

// IKJ elimination of rows [r0, r1) of one run, in place on the FP64 factor:
// the operations of ilu0_factor_host for those rows.  The rows a row
// eliminates with lie in earlier runs, whose elimination has finished.
__global__ __launch_bounds__(kBlock) void k_ilu_factor_run(int r0, int r1,
                                                           const int* __restrict__ rp,
                                                           const int* __restrict__ ci,
                                                           const int* __restrict__ diag,
                                                           double* __restrict__ lu)
{
    const int i = r0 + int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= r1)
        return;
    const int e = rp[i + 1];
    for (int kp = rp[i]; kp < e; ++kp) {
        const int k = ci[kp];
        if (k >= i)
            break;
        const int dk = diag[k];
        const double piv = lu[dk];
        if (piv == 0.0)
            continue;
        const double m = lu[kp] / piv;
        lu[kp] = m;
        int qp = kp + 1;
        for (int jp = dk + 1; jp < rp[k + 1]; ++jp) {
            const int j = ci[jp];
            while (qp < e && ci[qp] < j)
                ++qp;
            if (qp >= e)
                break;
            if (ci[qp] == j)
                lu[qp] -= m * lu[jp];
        }
    }
}

// Split factors in the iteration precision: L and U values through their
// maps into the full factor, and the diagonal.
template <class T>
__global__ __launch_bounds__(kBlock) void k_ilu_split(int nl, const int* __restrict__ lmap, int nu,
                                                      const int* __restrict__ umap, int n,
                                                      const int* __restrict__ diag,
                                                      const double* __restrict__ lu,
                                                      T* __restrict__ lval, T* __restrict__ uval,
                                                      T* __restrict__ d)
{
    const int64_t total = int64_t(nl) + nu + n;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < total; t += stride) {
        if (t < nl)
            lval[t] = T(lu[lmap[t]]);
        else if (t < int64_t(nl) + nu)
            uval[t - nl] = T(lu[umap[t - nl]]);
        else
            d[t - nl - nu] = T(lu[diag[t - nl - nu]]);
    }
}

// The rows written by a run are never read by the same launch (a run has no
// dependency inside it), so y may be both the gathered vector and the output.
template <class T>
__global__ __launch_bounds__(kBlock) void k_ilu_lower_run(StreamMat<T> L, const T* __restrict__ r,
                                                          T* y)
{
    EpiIluLower<T> e{r, y};
    stream_rows(L, y, e);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_ilu_upper_run(StreamMat<T> U, const T* __restrict__ y,
                                                          const T* __restrict__ d, T* z)
{
    EpiIluUpper<T> e{y, d, z};
    stream_rows(U, z, e);
}

#pragma clang fp contract(on)

// ------------------------------------------------------------------- host
int ilu0_factor_host(int n, const int* rp, const int* ci, std::vector<double>& lu,
                     std::vector<int>& diag, int* zero_pivots)
{
    diag.assign(static_cast<size_t>(n), -1);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                diag[size_t(i)] = k;
                break;
            }
    int zeros = 0;
    for (int i = 1; i < n; ++i) {
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int k = ci[kp];
            if (k >= i)
                break;
            if (diag[size_t(k)] < 0)
                continue;
            const double piv = lu[size_t(diag[size_t(k)])];
            if (piv == 0.0) {
                ++zeros;
                continue;
            }
            const double m = lu[size_t(kp)] / piv;
            lu[size_t(kp)] = m;
            // row i -= m * row k on the pattern; both rows are sorted, so the
            // search for column j resumes where the previous one stopped
            int qp = kp + 1;
            for (int jp = diag[size_t(k)] + 1; jp < rp[k + 1]; ++jp) {
                const int j = ci[jp];
                while (qp < rp[i + 1] && ci[qp] < j)
                    ++qp;
                if (qp >= rp[i + 1])
                    break;
                if (ci[qp] == j)
                    lu[size_t(qp)] -= m * lu[size_t(jp)];
            }
        }
    }
    for (int i = 0; i < n; ++i)
        if (diag[size_t(i)] >= 0 && lu[size_t(diag[size_t(i)])] == 0.0)
            ++zeros;
    if (zero_pivots)
        *zero_pivots = zeros;
    return MI_OK;
}

bool ilu0_factor_facade(int n, const int* rp, const int* ci, std::vector<double>& lu)
{
    std::vector<int> diag(static_cast<size_t>(n), -1);
    for (int i = 0; i < n; ++i) {
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] == i) {
                diag[size_t(i)] = k;
                break;
            }
        if (diag[size_t(i)] < 0)
            return false; // nothing has been modified yet
    }
    for (int i = 1; i < n; ++i) {
        for (int kp = rp[i]; kp < rp[i + 1]; ++kp) {
            const int k = ci[kp];
            if (k >= i)
                break;
            const double piv = lu[size_t(diag[size_t(k)])];
            if (piv == 0.0)
                return false; // rows before i stay factored, as the facade leaves them
            const double m = lu[size_t(kp)] / piv;
            lu[size_t(kp)] = m;
            int qp = kp + 1;
            for (int jp = diag[size_t(k)] + 1; jp < rp[k + 1]; ++jp) {
                const int j = ci[jp];
                while (qp < rp[i + 1] && ci[qp] < j)
                    ++qp;
                if (qp >= rp[i + 1])
                    break;
                if (ci[qp] == j)
                    lu[size_t(qp)] -= m * lu[size_t(jp)];
            }
        }
        if (lu[size_t(diag[size_t(i)])] == 0.0)
            return false;
    }
    return true;
}

void ilu0_levels(int n, const int* rp, const int* ci, const std::vector<int>& lend,
                 std::vector<int>& lrows, std::vector<int>& loff, std::vector<int>& urows,
                 std::vector<int>& uoff, const std::vector<int>* prefix_diag)
{
    std::vector<int> lev(static_cast<size_t>(n), 0);
    int nl = 0;
    for (int i = 0; i < n; ++i) {
        int l = 0;
        for (int k = rp[i]; k < lend[size_t(i)]; ++k)
            l = std::max(l, lev[size_t(ci[k])] + 1);
        lev[size_t(i)] = l;
        nl = std::max(nl, l + 1);
    }
    auto bucket = [&](int levels, std::vector<int>& rows, std::vector<int>& off) {
        off.assign(size_t(levels) + 1, 0);
        for (int i = 0; i < n; ++i)
            ++off[size_t(lev[size_t(i)]) + 1];
        for (int l = 0; l < levels; ++l)
            off[size_t(l) + 1] += off[size_t(l)];
        rows.assign(static_cast<size_t>(n), 0);
        std::vector<int> pos(off.begin(), off.end() - 1);
        for (int i = 0; i < n; ++i)
            rows[size_t(pos[size_t(lev[size_t(i)])]++)] = i;
    };
    bucket(n > 0 ? nl : 0, lrows, loff);
    int nu = 0;
    for (int i = n - 1; i >= 0; --i) {
        int l = 0;
        // a row without a diagonal under the prefix rule reads z_j of every
        // column j > i in rows 0..i
        const bool pre = prefix_diag && (*prefix_diag)[size_t(i)] < 0;
        for (int k = pre ? 0 : lend[size_t(i)]; k < rp[i + 1]; ++k)
            if (ci[k] > i)
                l = std::max(l, lev[size_t(ci[k])] + 1);
        // lev[] of rows > i already hold their U levels (processed first)
        lev[size_t(i)] = l;
        nu = std::max(nu, l + 1);
    }
    bucket(n > 0 ? nu : 0, urows, uoff);
}

void ilu0_segments(int n, const int* rp, const int* ci, const std::vector<int>& lend,
                   const std::vector<int>& diag, std::vector<int>& lbeg, std::vector<int>& ubeg)
{
    lbeg.assign(1, 0);
    int start = 0;
    for (int i = 0; i < n; ++i) {
        bool dep = false;
        for (int k = rp[i]; k < lend[size_t(i)] && !dep; ++k)
            dep = ci[k] >= start;
        if (dep) {
            lbeg.push_back(i);
            start = i;
        }
    }
    lbeg.push_back(n);
    ubeg.assign(1, n);
    int end = n;
    for (int i = n - 1; i >= 0; --i) {
        bool dep = false;
        for (int k = diag[size_t(i)] + 1; k < rp[i + 1] && !dep; ++k)
            dep = ci[k] < end;
        if (dep) {
            ubeg.push_back(i + 1);
            end = i + 1;
        }
    }
    ubeg.push_back(0);
}

namespace
{
// Split pattern of one triangle: rows i keep the entries k in [kb(i), ke(i)).
template <class KB, class KE>
void split_triangle(int n, const int* rp, const int* ci, KB kb, KE ke, std::vector<int>& trp,
                    std::vector<int>& tci, std::vector<int>& tmap)
{
    trp.assign(size_t(n) + 1, 0);
    tci.clear();
    tmap.clear();
    for (int i = 0; i < n; ++i) {
        for (int k = kb(i); k < ke(i); ++k) {
            tci.push_back(ci[k]);
            tmap.push_back(k);
        }
        trp[size_t(i) + 1] = int(tci.size());
    }
    (void)rp;
}

// Row blocks of the runs [b[t], b[t+1]) (ascending) or [b[t+1], b[t])
// (descending): per run, its block boundaries appended to blk and the pair
// (offset, block count) to sb.
void run_blocks(const std::vector<int>& bounds, bool descending, const std::vector<int>& trp,
                std::vector<int>& blk, std::vector<int>& sb)
{
    blk.clear();
    sb.clear();
    std::vector<int> b;
    for (size_t t = 0; t + 1 < bounds.size(); ++t) {
        const int r0 = descending ? bounds[t + 1] : bounds[t];
        const int r1 = descending ? bounds[t] : bounds[t + 1];
        build_row_blocks(r1 - r0, trp.data() + r0, b);
        sb.push_back(int(blk.size()));
        sb.push_back(int(b.size()) - 1);
        for (int v : b)
            blk.push_back(v + r0);
    }
}
} // namespace

template <class T> int Ilu0<T>::prepare(const Pattern& P, hipStream_t s)
{
    if (level_built && levels_for == &P)
        return MI_OK;
    n = int(P.n);
    const int* rp = P.h_rp.data();
    const int* ci = P.h_ci.data();
    std::vector<int> diag(static_cast<size_t>(n), -1), lend(static_cast<size_t>(n));
    int missing = 0;
    for (int i = 0; i < n; ++i) {
        int k = rp[i];
        while (k < rp[i + 1] && ci[k] < i)
            ++k;
        lend[size_t(i)] = k;
        if (k < rp[i + 1] && ci[k] == i)
            diag[size_t(i)] = k;
        else
            ++missing;
    }
    std::vector<int> lrows, loff, urows, uoff;
    ilu0_levels(n, rp, ci, lend, lrows, loff, urows, uoff, &diag);
    MI_TRY(d_lrows.upload(lrows.data(), lrows.size(), s));
    MI_TRY(d_urows.upload(urows.data(), urows.size(), s));
    MI_TRY(d_lend.upload(lend.data(), lend.size(), s));
    MI_TRY(d_diag.upload(diag.data(), diag.size(), s));
    h_loff = loff;
    h_uoff = uoff;
    h_diag = diag;
    // run sweeps when every run holds many rows (each launch has the work of
    // a large SpMV); otherwise the dependency levels
    ilu0_segments(n, rp, ci, lend, diag, lbeg, ubeg);
    const int runs = int(std::max(lbeg.size(), ubeg.size())) - 1;
    // rows without a diagonal keep the level sweeps, which follow the
    // reference's treatment of them exactly
    seg = missing == 0 && n > 0 && int64_t(n) >= int64_t(2048) * runs;
    two = false;
    if (seg) {
        std::vector<int> trp, tci, blk;
        split_triangle(n, rp, ci, [&](int i) { return rp[i]; },
                       [&](int i) { return lend[size_t(i)]; }, trp, tci, Lmap);
        run_blocks(lbeg, false, trp, blk, Lsb);
        const bool first_without_l = lbeg.size() == 3 && trp[size_t(lbeg[1])] == 0;
        MI_TRY(Lrp.upload(trp.data(), trp.size(), s));
        MI_TRY(Lci.upload(tci.data(), tci.size(), s));
        MI_TRY(Lblk.upload(blk.data(), blk.size(), s));
        MI_TRY(d_Lmap.upload(Lmap.data(), Lmap.size(), s));
        MI_HIP(hipStreamSynchronize(s));
        split_triangle(n, rp, ci, [&](int i) { return diag[size_t(i)] + 1; },
                       [&](int i) { return rp[i + 1]; }, trp, tci, Umap);
        run_blocks(ubeg, true, trp, blk, Usb);
        // two runs [0, b) and [b, n) where the first has no L entries and
        // the second no U entries (a red-black ordering): one launch per run
        two = first_without_l && ubeg.size() == 3 && lbeg[1] == ubeg[1] &&
              trp[size_t(n)] == trp[size_t(ubeg[1])];
        MI_TRY(Urp.upload(trp.data(), trp.size(), s));
        MI_TRY(Uci.upload(tci.data(), tci.size(), s));
        MI_TRY(Ublk.upload(blk.data(), blk.size(), s));
        MI_TRY(d_Umap.upload(Umap.data(), Umap.size(), s));
        MI_TRY(Lval.alloc(std::max<size_t>(Lmap.size(), 1)));
        MI_TRY(Uval.alloc(std::max<size_t>(Umap.size(), 1)));
        MI_TRY(Ud.alloc(static_cast<size_t>(std::max(n, 1))));
    }
    MI_TRY(d_y.alloc(static_cast<size_t>(std::max(n, 1))));
    rp_dev = P.rp.p;
    ci_dev = P.ci.p;
    MI_HIP(hipStreamSynchronize(s));
    level_built = true;
    levels_for = &P;
    return MI_OK;
}

template <class T> int Ilu0<T>::setup(const Pattern& P, const std::vector<double>& vals, hipStream_t s)
{
    MI_TRY(prepare(P, s));
    std::vector<double> lu(vals);
    std::vector<int> diag;
    MI_TRY(ilu0_factor_host(n, P.h_rp.data(), P.h_ci.data(), lu, diag, &zero_pivots));
    if (seg) {
        std::vector<T> lv(Lmap.size()), uv(Umap.size()), dv(static_cast<size_t>(n));
        for (size_t j = 0; j < Lmap.size(); ++j)
            lv[j] = T(lu[size_t(Lmap[j])]);
        for (size_t j = 0; j < Umap.size(); ++j)
            uv[j] = T(lu[size_t(Umap[j])]);
        for (int i = 0; i < n; ++i)
            dv[size_t(i)] = T(lu[size_t(diag[size_t(i)])]);
        MI_TRY(Lval.upload(lv.data(), lv.size(), s));
        MI_TRY(Uval.upload(uv.data(), uv.size(), s));
        MI_TRY(Ud.upload(dv.data(), dv.size(), s));
    }
    std::vector<T> tv(lu.begin(), lu.end());
    MI_TRY(d_lu.upload(tv.data(), tv.size(), s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

template <class T> int Ilu0<T>::setup_device(const Pattern& P, const double* vals, hipStream_t s)
{
    MI_TRY(prepare(P, s));
    if (!seg)
        return set_error(MI_ERR_UNSUPPORTED, "ILU(0) device factorization needs a run ordering");
    const size_t nnz = size_t(P.nnz);
    MI_TRY(d_fac.alloc(std::max<size_t>(nnz, 1)));
    if (nnz)
        MI_HIP(hipMemcpyAsync(d_fac.p, vals, nnz * sizeof(double), hipMemcpyDeviceToDevice, s));
    for (size_t t = 0; t + 1 < lbeg.size(); ++t) {
        const int r0 = lbeg[t], r1 = lbeg[t + 1];
        if (r1 > r0)
            k_ilu_factor_run<<<grid_for(r1 - r0, kBlock), kBlock, 0, s>>>(r0, r1, P.rp.p, P.ci.p,
                                                                          d_diag.p, d_fac.p);
    }
    const int nl = int(Lmap.size()), nu = int(Umap.size());
    k_ilu_split<T><<<reduce_grid(int64_t(nl) + nu + n, 4), kBlock, 0, s>>>(
        nl, d_Lmap.p, nu, d_Umap.p, n, d_diag.p, d_fac.p, Lval.p, Uval.p, Ud.p);
    MI_HIP(hipGetLastError());
    zero_pivots = -1; // not counted on this path
    return MI_OK;
}

template <class T> int Ilu0<T>::setup_ssor(const Pattern& P, const double* vals, hipStream_t s)
{
    MI_TRY(prepare(P, s));
    ssor = true;
    const int64_t nnz = P.nnz;
    MI_TRY(d_invD.alloc(static_cast<size_t>(std::max(n, 1))));
    MI_TRY(Ud.alloc(static_cast<size_t>(std::max(n, 1))));
    if (n > 0) {
        k_ssor_diag<T><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, d_diag.p, vals, Ud.p, d_invD.p);
        MI_HIP(hipGetLastError());
    }
    if (seg) {
        const int nl = int(Lmap.size()), nu = int(Umap.size());
        // the diagonal slot of k_ilu_split is rewritten with the same a_ii
        k_ilu_split<T><<<reduce_grid(int64_t(nl) + nu + n, 4), kBlock, 0, s>>>(
            nl, d_Lmap.p, nu, d_Umap.p, n, d_diag.p, vals, Lval.p, Uval.p, Ud.p);
    } else {
        MI_TRY(d_lu.alloc(static_cast<size_t>(std::max<int64_t>(nnz, 1))));
        if (nnz > 0)
            k_to_t<T><<<reduce_grid(nnz, 4), kBlock, 0, s>>>(nnz, vals, d_lu.p);
    }
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template <class T> int Ilu0<T>::apply(const T* r, T* z, hipStream_t s)
{
    // Launched directly: a graph of the ~600 level launches of a 3-D mesh in
    // natural order measured no faster (the cost of each tiny level launch is
    // on the device), and the run orderings need two to four launches.
    if (ssor) {
        if (seg) {
            for (size_t t = 0; t + 1 < Lsb.size(); t += 2) {
                const int nrb = Lsb[t + 1];
                if (nrb == 0)
                    continue;
                StreamMat<T> L{n, nrb, Lblk.p + Lsb[t], Lrp.p, Lci.p, Lval.p};
                k_ssor_lower_run<T><<<stream_grid(nrb), kBlock, 0, s>>>(L, r, d_invD.p, d_y.p);
            }
            for (size_t t = 0; t + 1 < Usb.size(); t += 2) {
                const int nrb = Usb[t + 1];
                if (nrb == 0)
                    continue;
                StreamMat<T> U{n, nrb, Ublk.p + Usb[t], Urp.p, Uci.p, Uval.p};
                k_ssor_upper_run<T><<<stream_grid(nrb), kBlock, 0, s>>>(U, d_y.p, Ud.p, d_invD.p, z);
            }
        } else {
            for (size_t l = 0; l + 1 < h_loff.size(); ++l) {
                const int beg = h_loff[l], cnt = h_loff[l + 1] - beg;
                k_ssor_lower_level<T><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(
                    cnt, d_lrows.p + beg, rp_dev, d_lend.p, ci_dev, d_lu.p, d_invD.p, r, d_y.p);
            }
            for (size_t l = 0; l + 1 < h_uoff.size(); ++l) {
                const int beg = h_uoff[l], cnt = h_uoff[l + 1] - beg;
                k_ssor_upper_level<T><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(
                    cnt, d_urows.p + beg, rp_dev, d_diag.p, d_lend.p, ci_dev, d_lu.p, Ud.p,
                    d_invD.p, d_y.p, z);
            }
        }
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    if (seg && two) {
        // rows [b, n): z = (r - L r) / d ; rows [0, b): z = (r - U z) / d
        const int nl = Lsb[3], nu = Usb[3];
        if (nl > 0) {
            StreamMat<T> L{n, nl, Lblk.p + Lsb[2], Lrp.p, Lci.p, Lval.p};
            k_ilu_two_run<T><<<stream_grid(nl), kBlock, 0, s>>>(L, r, r, Ud.p, z);
        }
        if (nu > 0) {
            StreamMat<T> U{n, nu, Ublk.p + Usb[2], Urp.p, Uci.p, Uval.p};
            k_ilu_two_run<T><<<stream_grid(nu), kBlock, 0, s>>>(U, z, r, Ud.p, z);
        }
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    if (seg) {
        for (size_t t = 0; t + 1 < Lsb.size(); t += 2) {
            const int nrb = Lsb[t + 1];
            if (nrb == 0)
                continue;
            StreamMat<T> L{n, nrb, Lblk.p + Lsb[t], Lrp.p, Lci.p, Lval.p};
            k_ilu_lower_run<T><<<stream_grid(nrb), kBlock, 0, s>>>(L, r, d_y.p);
        }
        for (size_t t = 0; t + 1 < Usb.size(); t += 2) {
            const int nrb = Usb[t + 1];
            if (nrb == 0)
                continue;
            StreamMat<T> U{n, nrb, Ublk.p + Usb[t], Urp.p, Uci.p, Uval.p};
            k_ilu_upper_run<T><<<stream_grid(nrb), kBlock, 0, s>>>(U, d_y.p, Ud.p, z);
        }
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    for (size_t l = 0; l + 1 < h_loff.size(); ++l) {
        const int beg = h_loff[l], cnt = h_loff[l + 1] - beg;
        k_lower_level<T><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(cnt, d_lrows.p + beg, rp_dev,
                                                                  d_lend.p, ci_dev, d_lu.p, r, d_y.p);
    }
    for (size_t l = 0; l + 1 < h_uoff.size(); ++l) {
        const int beg = h_uoff[l], cnt = h_uoff[l + 1] - beg;
        k_upper_level<T><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(
            cnt, d_urows.p + beg, rp_dev, d_diag.p, d_lend.p, ci_dev, d_lu.p, d_y.p, z, 1);
    }
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template struct Ilu0<double>;
template struct Ilu0<float>;

namespace
{
struct IluPrecond final : Precond
{
    Ilu0<double> f64;
    Ilu0<float> f32;
    bool has64 = false, has32 = false;
    bool ssor = false; // SSOR(1) sweeps on A itself instead of ILU(0) factors
    int setup(mi_csr* A, bool need64, bool need32, hipStream_t s) override
    {
        const Pattern& P = *A->pat;
        has64 = has32 = false;
        if (ssor) {
            if (need64)
                MI_TRY(f64.setup_ssor(P, A->val.p, s));
            if (need32)
                MI_TRY(f32.setup_ssor(P, A->val.p, s));
            MI_HIP(hipStreamSynchronize(s));
            has64 = need64;
            has32 = need32;
            return MI_OK;
        }
        if (need64)
            MI_TRY(f64.prepare(P, s));
        if (need32)
            MI_TRY(f32.prepare(P, s));
        // run orderings factor on the device from the device values; the
        // caller's order is factored on the host, the reference's way
        const bool dev = (!need64 || f64.seg) && (!need32 || f32.seg);
        if (dev) {
            if (need64)
                MI_TRY(f64.setup_device(P, A->val.p, s));
            if (need32)
                MI_TRY(f32.setup_device(P, A->val.p, s));
            MI_HIP(hipStreamSynchronize(s));
        } else {
            MI_TRY(A->sync_host());
            if (need64)
                MI_TRY(f64.setup(P, A->h_val, s));
            if (need32)
                MI_TRY(f32.setup(P, A->h_val, s));
        }
        has64 = need64;
        has32 = need32;
        return MI_OK;
    }
    int apply64(const double* r, double* z, hipStream_t s) override
    {
        return has64 ? f64.apply(r, z, s) : set_error(MI_ERR_UNSUPPORTED, "no FP64 ILU(0)");
    }
    int apply32(const float* r, float* z, hipStream_t s) override
    {
        return has32 ? f32.apply(r, z, s) : set_error(MI_ERR_UNSUPPORTED, "no FP32 ILU(0)");
    }
    const char* name() const override { return ssor ? "ssor" : "ilu0"; }
    int levels() const override { return has64 ? f64.sweep_launches() : f32.sweep_launches(); }
};
} // namespace

std::unique_ptr<Precond> make_ilu0() { return std::make_unique<IluPrecond>(); }

std::unique_ptr<Precond> make_ssor()
{
    auto p = std::make_unique<IluPrecond>();
    p->ssor = true;
    return p;
}

} // namespace mi

using namespace mi;

extern "C" {

MI_API int mi_ilu0_factor_inplace(size_t n, const void* rowPtr, const void* colIdx, void* values,
                                  int* ok)
{
    MI_TRY(ensure_device());
    if (ok)
        *ok = 1;
    if (n == 0)
        return MI_OK;
    if (!rowPtr || !colIdx || !values)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    hipStream_t s = default_stream();
    std::vector<int> rp(n + 1);
    MI_HIP(hipMemcpyAsync(rp.data(), rowPtr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const size_t nnz = size_t(rp[n]);
    std::vector<int> ci(nnz);
    std::vector<double> v(nnz);
    MI_HIP(hipMemcpyAsync(ci.data(), colIdx, nnz * sizeof(int), hipMemcpyDeviceToHost, s));
    MI_HIP(hipMemcpyAsync(v.data(), values, nnz * sizeof(double), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const bool done = ilu0_factor_facade(int(n), rp.data(), ci.data(), v);
    MI_HIP(hipMemcpyAsync(values, v.data(), nnz * sizeof(double), hipMemcpyHostToDevice, s));
    MI_HIP(hipStreamSynchronize(s));
    if (ok)
        *ok = done ? 1 : 0;
    return MI_OK;
}

MI_API int mi_ilu0_apply(size_t n, const void* rowPtr, const void* colIdx, const void* values,
                         const void* r, void* z, void* stream)
{
    MI_TRY(ensure_device());
    if (n == 0)
        return MI_OK;
    if (!rowPtr || !colIdx || !values || !r || !z)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : default_stream();
    // The façade has no handle to cache the level schedule in: it is rebuilt
    // from the device pattern on every call (use an mi_solver for repeated
    // applications).
    std::vector<int> rp(n + 1);
    MI_HIP(hipMemcpyAsync(rp.data(), rowPtr, (n + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const size_t nnz = size_t(rp[n]);
    std::vector<int> ci(nnz);
    MI_HIP(hipMemcpyAsync(ci.data(), colIdx, nnz * sizeof(int), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    std::vector<int> lend(n), diag(n, -1), lrows, loff, urows, uoff;
    for (size_t i = 0; i < n; ++i) {
        int k = rp[i];
        while (k < rp[i + 1] && ci[size_t(k)] < int(i))
            ++k;
        lend[i] = k;
        if (k < rp[i + 1] && ci[size_t(k)] == int(i))
            diag[i] = k; // absent: divide by 1 (src/gpu/GpuKernels.cpp:466-477)
    }
    ilu0_levels(int(n), rp.data(), ci.data(), lend, lrows, loff, urows, uoff, nullptr);
    DevBuf<int> dl, du, dlend, ddiag;
    DevBuf<double> y;
    MI_TRY(dl.upload(lrows.data(), lrows.size(), s));
    MI_TRY(du.upload(urows.data(), urows.size(), s));
    MI_TRY(dlend.upload(lend.data(), lend.size(), s));
    MI_TRY(ddiag.upload(diag.data(), diag.size(), s));
    MI_TRY(y.alloc(n));
    const int* drp = static_cast<const int*>(rowPtr);
    const int* dci = static_cast<const int*>(colIdx);
    const double* lu = static_cast<const double*>(values);
    for (size_t l = 0; l + 1 < loff.size(); ++l) {
        const int beg = loff[l], cnt = loff[l + 1] - beg;
        k_lower_level<double><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(
            cnt, dl.p + beg, drp, dlend.p, dci, lu, static_cast<const double*>(r), y.p);
    }
    for (size_t l = 0; l + 1 < uoff.size(); ++l) {
        const int beg = uoff[l], cnt = uoff[l + 1] - beg;
        k_upper_level<double><<<grid_for(cnt, kBlock), kBlock, 0, s>>>(
            cnt, du.p + beg, drp, ddiag.p, dlend.p, dci, lu, y.p, static_cast<double*>(z), 0);
    }
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

} // extern "C"
