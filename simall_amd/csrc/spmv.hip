// spmv.hip — CSR SpMV: the CSR-stream engine for matrices held in handles,
// and vector-CSR (LANES lanes per row, LANES = 64: one wavefront per row) for
// the raw device-pointer entry point, which has no host-side row partition.
#include "csr.hpp"
#include "spmv.hpp"
#include "stream.hpp"

#include "host_threads.hpp"

#include <climits>

namespace mi
{

template <int LANES, class VT, class XT, class YT>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, const int* __restrict__ rp,
                                                 const int* __restrict__ ci,
                                                 const VT* __restrict__ val,
                                                 const XT* __restrict__ x, YT* __restrict__ y)
{
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const int row = int(gid / LANES);
    const int lane = int(gid & (LANES - 1));
    const auto s = csr_row_dot<LANES>(row, n, lane, rp, ci, val, x);
    if (row < n && lane == 0)
        y[row] = YT(s);
}

template <int LANES>
__global__ __launch_bounds__(kBlock) void k_residual(int n, const int* __restrict__ rp,
                                                     const int* __restrict__ ci,
                                                     const double* __restrict__ val,
                                                     const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ r)
{
    const int64_t gid = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    const int row = int(gid / LANES);
    const int lane = int(gid & (LANES - 1));
    const double s = csr_row_dot<LANES>(row, n, lane, rp, ci, val, x);
    if (row < n && lane == 0)
        r[row] = b[row] - s;
}

void build_row_blocks(int n, const int* rp, std::vector<int>& blk, int max_rows)
{
    blk.clear();
    blk.push_back(0);
    int r = 0;
    while (r < n) {
        const int start = r;
        while (r < n && r - start < max_rows && rp[r + 1] - rp[start] <= kStreamNnz)
            ++r;
        if (r == start)
            ++r; // a row longer than kStreamNnz gets a block of its own
        blk.push_back(r);
    }
}

int stream_lanes_per_row(int64_t n, int64_t nnz, bool wide)
{
    if (!wide || n <= 0)
        return 1;
    const double m = double(nnz) / double(n);
    if (m < 10.0)
        return 1;
    int l = 1;
    while (l < 8 && 2 * l <= m / 3.0)
        l *= 2;
    return l;
}

// One workgroup per row block: the block's (column, slot) keys sorted in LDS
// by a bitonic network, then written as packed entries relative to the
// block's smallest column (written to the block header).  A block of one long row keeps its CSR order
// (its columns are already ascending).
__global__ __launch_bounds__(kBlock) void k_sort_blocks(const int* __restrict__ blk,
                                                        const int* __restrict__ rp,
                                                        const int* __restrict__ ci,
                                                        unsigned* __restrict__ cw,
                                                        int* __restrict__ sperm,
                                                        int* __restrict__ hdr,
                                                        int* __restrict__ bad)
{
    __shared__ unsigned key[kStreamNnz];
    __shared__ int cmin[kBlock];
    const int rb = blockIdx.x;
    const int tid = threadIdx.x;
    const int r0 = blk[rb], r1 = blk[rb + 1];
    const int q0 = rp[r0], q1 = rp[r1];
    const int cnt = q1 - q0;
    // smallest column of the block
    int m = INT_MAX;
    for (int k = q0 + tid; k < q1; k += kBlock)
        m = min(m, ci[k]);
    cmin[tid] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w)
            cmin[tid] = min(cmin[tid], cmin[tid + w]);
        __syncthreads();
    }
    const int c0 = cnt > 0 ? cmin[0] : 0;
    if (tid == 0)
        hdr[4 * rb + 2] = c0;
    // column field in [0, 2^21 - 1): a packed entry is never the -1 marker
    auto field = [&](int col) -> unsigned {
        const int64_t d = int64_t(col) - c0;
        if (d >= int64_t(kColMask))
            bad[0] = 1;
        return unsigned(d) & kColMask;
    };
    if (cnt > kStreamNnz) {
        for (int k = q0 + tid; k < q1; k += kBlock) {
            cw[k] = field(ci[k]);
            sperm[k] = k;
        }
        return;
    }
    for (int j = tid; j < kStreamNnz; j += kBlock)
        key[j] = j < cnt ? (field(ci[q0 + j]) << 11) | unsigned(j) : 0xFFFFFFFFu;
    for (int size = 2; size <= kStreamNnz; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (int t = tid; t < kStreamNnz / 2; t += kBlock) {
                const int lo = 2 * t - (t & (stride - 1));
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned a = key[lo], b = key[hi];
                if ((a > b) == up) {
                    key[lo] = b;
                    key[hi] = a;
                }
            }
        }
    }
    __syncthreads();
    for (int j = tid; j < cnt; j += kBlock) {
        const unsigned kv = key[j];
        const unsigned slot = kv & 2047u;
        cw[q0 + j] = (slot << kSlotShift) | (kv >> 11);
        sperm[q0 + j] = q0 + int(slot);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_permute(int64_t nnz, const int* __restrict__ perm,
                                                    const T* __restrict__ src, T* __restrict__ dst)
{
    for (int64_t m = int64_t(blockIdx.x) * kBlock + threadIdx.x; m < nnz;
         m += int64_t(gridDim.x) * kBlock)
    {
        const int q = perm[m];
        dst[m] = q >= 0 ? src[q] : T(0); // -1: a SELL padding slot
    }
}

template <class T>
int launch_permute(int64_t nnz, const int* perm, const T* src, T* dst, hipStream_t s)
{
    if (nnz <= 0)
        return MI_OK;
    k_permute<T><<<grid_for(nnz, kBlock), kBlock, 0, s>>>(nnz, perm, src, dst);
    MI_HIP(hipGetLastError());
    return MI_OK;
}
template int launch_permute<double>(int64_t, const int*, const double*, double*, hipStream_t);
template int launch_permute<float>(int64_t, const int*, const float*, float*, hipStream_t);

int SellForm::build(int n, int64_t nnz, const int* rp, const int* ci, bool mark_diag,
                    hipStream_t s)
{
    soff.release();
    eci.release();
    eperm.release();
    ecd.release();
    nsl = 0;
    padded = 0;
    if (n <= 0 || n >= kSellDiag || !sell_enabled())
        return MI_OK;
    const int ns = (n + kSellSlice - 1) / kSellSlice;
    std::vector<int> off(size_t(ns) + 1);
    int64_t tot = 0;
    for (int sl = 0; sl < ns; ++sl) {
        int w = 0;
        const int r1 = std::min(n, (sl + 1) * kSellSlice);
        for (int i = sl * kSellSlice; i < r1; ++i)
            w = std::max(w, rp[i + 1] - rp[i]);
        if (w > kSellMaxWidth)
            return MI_OK;
        off[size_t(sl)] = int(tot);
        tot += int64_t(kSellSlice) * w;
        if (tot > INT_MAX)
            return MI_OK;
    }
    off[size_t(ns)] = int(tot);
    if (tot > nnz + nnz / 4 + kSellSlice)
        return MI_OK;
    // columns share their word with the diagonal flag (bit 30): a rectangular
    // operator with 2^30 or more columns keeps row blocks
    for (int64_t k = 0; k < nnz; ++k)
        if (ci[k] < 0 || ci[k] >= kSellDiag)
            return MI_OK;
    std::vector<int> c(size_t(tot), -1), pm(size_t(tot), -1);
    host_parallel_for(ns, host_threads(), [&](int b, int e, int) {
        for (int sl = b; sl < e; ++sl) {
            const int r1 = std::min(n, (sl + 1) * kSellSlice);
            for (int i = sl * kSellSlice; i < r1; ++i) {
                bool seen = !mark_diag;
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    const size_t m = size_t(off[size_t(sl)]) +
                                     size_t(kSellSlice) * size_t(k - rp[i]) +
                                     size_t(i - sl * kSellSlice);
                    int col = ci[k];
                    if (!seen && col == i) {
                        col |= kSellDiag;
                        seen = true;
                    }
                    c[m] = col;
                    pm[m] = k;
                }
            }
        }
    }, 256);
    MI_TRY(soff.upload(off.data(), off.size(), s));
    MI_TRY(eci.upload(c.data(), c.size(), s));
    MI_TRY(eperm.upload(pm.data(), pm.size(), s));
    // 16-bit column offsets from the row, for square operators within reach
    std::vector<short> d;
    if (mark_diag && sell_delta_enabled()) {
        bool ok = true;
        d.assign(size_t(tot), kSellDelta);
        std::vector<int> ndiag(size_t(n), 0);
        for (int sl = 0; sl < ns && ok; ++sl) {
            const int r1 = std::min(n, (sl + 1) * kSellSlice);
            for (int i = sl * kSellSlice; i < r1 && ok; ++i)
                for (int k = rp[i]; k < rp[i + 1]; ++k) {
                    const int64_t off_c = int64_t(ci[k]) - i;
                    if (off_c < -32767 || off_c > 32767 || (off_c == 0 && ++ndiag[size_t(i)] > 1)) {
                        ok = false;
                        break;
                    }
                    d[size_t(off[size_t(sl)]) + size_t(kSellSlice) * size_t(k - rp[i]) +
                      size_t(i - sl * kSellSlice)] = short(off_c);
                }
        }
        if (ok)
            MI_TRY(ecd.upload(d.data(), d.size(), s));
    }
    MI_HIP(hipStreamSynchronize(s));
    nsl = ns;
    padded = tot;
    return MI_OK;
}

int RowBlocks::sort_columns(const int* rp_dev, const int* ci_dev, int64_t nnz, hipStream_t s)
{
    cw.release();
    sperm.release();
    if (nrb == 0 || nnz == 0)
        return MI_OK;
    DevBuf<unsigned> w;
    DevBuf<int> perm, bad;
    MI_TRY(w.alloc(size_t(nnz)));
    MI_TRY(perm.alloc(size_t(nnz)));
    MI_TRY(bad.alloc(1));
    MI_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    k_sort_blocks<<<unsigned(nrb), kBlock, 0, s>>>(blk.p, rp_dev, ci_dev, w.p, perm.p, hdr.p,
                                                   bad.p);
    MI_HIP(hipGetLastError());
    int h_bad = 0;
    MI_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (h_bad)
        return MI_OK; // a block whose columns span too far: CSR order stays
    cw = std::move(w);
    sperm = std::move(perm);
    return MI_OK;
}

int RowBlocks::upload(int n, const int* rp_host, hipStream_t s, int lanes_per_row)
{
    if (lanes_per_row < 1 || lanes_per_row > 64 || (lanes_per_row & (lanes_per_row - 1)))
        return set_error(MI_ERR_BAD_ARG, "lanes per row %d is not a power of two in [1, 64]",
                         lanes_per_row);
    lpr = lanes_per_row;
    std::vector<int> b;
    build_row_blocks(n, rp_host, b, kStreamRows / lpr);
    nrb = int(b.size()) - 1;
    MI_TRY(blk.upload(b.data(), b.size(), s));
    std::vector<int> h(4 * b.size(), 0);
    for (size_t r = 0; r < b.size(); ++r) {
        h[4 * r] = b[r];
        h[4 * r + 1] = rp_host[b[r]];
    }
    MI_TRY(hdr.upload(h.data(), h.size(), s));
    return hipStreamSynchronize(s) == hipSuccess ? MI_OK
                                                  : set_error(MI_ERR_DEVICE, "row block upload");
}

template <class T> struct EpiStore
{
    T* y;
    __device__ __forceinline__ void pre(int) {}
    __device__ __forceinline__ void row(int i, typename acc_type<T>::type s) { y[i] = T(s); }
};

// y = A x on the SELL-64 form (sell_rows: one lane per row, CSR order).
template <class T>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(StreamMat<T> A, const T* __restrict__ x,
                                                      T* __restrict__ y)
{
    EpiStore<T> e{y};
    sell_rows(A, x, e);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_stream_spmv(StreamMat<T> A, const T* __restrict__ x,
                                                        T* __restrict__ y)
{
    EpiStore<T> e{y};
    stream_rows(A, x, e);
}

// y = A x for one column-ordered row block per workgroup with one thread per
// row: the CSR-stream engine's arithmetic (products at their CSR slots, rows
// summed in column order, long rows by the same strided fold), without the
// epilogue, lane-group and persistent-range machinery the solver kernels need.
template <class T>
__global__ __launch_bounds__(kBlock) void k_spmv_cols(StreamMat<T> A, const T* __restrict__ x,
                                                      T* __restrict__ y)
{
#pragma clang fp contract(off)
    using Acc = typename acc_type<T>::type;
    Acc* prod = reinterpret_cast<Acc*>(stream_lds());
    const int rb = xcd_logical_block(blockIdx.x, gridDim.x);
    if (rb >= A.nrb)
        return;
    const int4 h0 = reinterpret_cast<const int4*>(A.hdr)[rb];
    const int4 h1 = reinterpret_cast<const int4*>(A.hdr)[rb + 1];
    const int r0 = h0.x, k0 = h0.y, c0 = h0.z, r1 = h1.x, k1 = h1.y;
    const int tid = threadIdx.x;
    if (k1 - k0 > kStreamNnz) {
        Acc s = 0;
        for (int k = k0 + tid; k < k1; k += kBlock)
            s += Acc(__builtin_nontemporal_load(&A.val[k])) *
                 Acc(x[c0 + int(__builtin_nontemporal_load(&A.cw[k]) & kColMask)]);
        prod[tid] = s;
        __syncthreads();
        if (tid == 0) {
            Acc t = 0;
            for (int j = 0; j < kBlock; ++j)
                t += prod[j];
            y[r0] = T(t);
        }
        return;
    }
    const int i = r0 + tid;
    int kb = 0, ke = 0;
    if (i < r1) {
        kb = A.rp[i] - k0;
        ke = A.rp[i + 1] - k0;
    }
    unsigned w[kLoadsPerThread];
    T a[kLoadsPerThread];
#pragma unroll
    for (int j = 0; j < kLoadsPerThread; ++j) {
        const int k = k0 + tid + j * kBlock;
        w[j] = k < k1 ? __builtin_nontemporal_load(&A.cw[k]) : 0xFFFFFFFFu;
        a[j] = k < k1 ? __builtin_nontemporal_load(&A.val[k]) : T(0);
    }
#pragma unroll
    for (int j = 0; j < kLoadsPerThread; ++j)
        if (w[j] != 0xFFFFFFFFu)
            prod[w[j] >> kSlotShift] = Acc(a[j]) * Acc(x[c0 + int(w[j] & kColMask)]);
    __syncthreads();
    if (i < r1) {
        Acc s = 0;
        for (int k = kb; k < ke; ++k)
            s += prod[k];
        y[i] = T(s);
    }
}

template <class T> int launch_spmv_stream(const Pattern& P, const T* val, const T* x, T* y,
                                          hipStream_t s)
{
    StreamMat<T> A{int(P.n), P.rb.nrb, P.rb.blk.p, P.rp.p, P.ci.p, val};
    A.hdr = P.rb.hdr.p;
    return launch_spmv_view(A, x, y, s);
}

template <class T> int launch_spmv_view(const StreamMat<T>& A, const T* x, T* y, hipStream_t s)
{
    if (A.n == 0)
        return MI_OK;
    if (A.soff) {
        k_sell_spmv<T><<<sell_grid(A.nsl), kBlock, 0, s>>>(A, x, y);
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    // FP64: one row block per workgroup (no reduction partials here), so the
    // dispatcher hands out blocks as workgroups retire; measured on the 128^3
    // pressure matrix from HBM 47.9 vs 48.9 us.  FP32 keeps persistent ranges,
    // whose next-block prefetch is worth more (39.7 vs 38.1 us).
    if (sizeof(T) == 8 && A.cw && A.hdr && A.lpr == 1) {
        k_spmv_cols<T><<<unsigned((int64_t(A.nrb) + 7) / 8 * 8), kBlock, 0, s>>>(A, x, y);
        MI_HIP(hipGetLastError());
        return MI_OK;
    }
    const unsigned G = sizeof(T) == 8 ? unsigned((int64_t(A.nrb) + 7) / 8 * 8)
                                      : stream_grid(A.nrb);
    k_stream_spmv<T><<<G, kBlock, 0, s>>>(A, x, y);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

template int launch_spmv_stream<double>(const Pattern&, const double*, const double*, double*,
                                        hipStream_t);
template int launch_spmv_stream<float>(const Pattern&, const float*, const float*, float*,
                                       hipStream_t);
template int launch_spmv_view<double>(const StreamMat<double>&, const double*, double*,
                                      hipStream_t);
template int launch_spmv_view<float>(const StreamMat<float>&, const float*, float*, hipStream_t);

template <class VT, class XT, class YT>
static int spmv_dispatch(int n, int64_t nnz, const int* rp, const int* ci, const VT* val,
                         const XT* x, YT* y, hipStream_t s, int lanes)
{
    if (n <= 0)
        return MI_OK;
    if (lanes <= 0)
        lanes = lanes_for(n, nnz);
    const unsigned grid = grid_for(int64_t(n) * lanes, kBlock);
    switch (lanes) {
    case 2: k_spmv<2><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    case 4: k_spmv<4><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    case 8: k_spmv<8><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    case 16: k_spmv<16><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    case 32: k_spmv<32><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    default: k_spmv<64><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, y); break;
    }
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int launch_spmv_f64(int n, int64_t nnz, const int* rp, const int* ci, const double* val,
                    const double* x, double* y, hipStream_t s, int lanes)
{
    return spmv_dispatch(n, nnz, rp, ci, val, x, y, s, lanes);
}

int launch_spmv_f32(int n, int64_t nnz, const int* rp, const int* ci, const float* val,
                    const float* x, float* y, hipStream_t s, int lanes)
{
    return spmv_dispatch(n, nnz, rp, ci, val, x, y, s, lanes);
}

int launch_residual_f64(int n, int64_t nnz, const int* rp, const int* ci, const double* val,
                        const double* x, const double* b, double* r, hipStream_t s, int lanes)
{
    if (n <= 0)
        return MI_OK;
    if (lanes <= 0)
        lanes = lanes_for(n, nnz);
    const unsigned grid = grid_for(int64_t(n) * lanes, kBlock);
    switch (lanes) {
    case 2: k_residual<2><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    case 4: k_residual<4><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    case 8: k_residual<8><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    case 16: k_residual<16><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    case 32: k_residual<32><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    default: k_residual<64><<<grid, kBlock, 0, s>>>(n, rp, ci, val, x, b, r); break;
    }
    MI_HIP(hipGetLastError());
    return MI_OK;
}

} // namespace mi
