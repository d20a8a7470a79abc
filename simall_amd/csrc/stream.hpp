// stream.hpp — the CSR-stream row engine shared by every SpMV-shaped kernel.
//
// Rows are grouped on the host into row blocks of at most kStreamRows rows
// and kStreamNnz nonzeros.  For one row block the workgroup
//   1. streams the block's contiguous span of colIdx/values with coalesced
//      loads, forms every product a_ij * x_j (x gathered through the cache)
//      and parks it in LDS;
//   2. gives each row to one thread, which adds its products in column order
//      and hands (row, sum) to an epilogue (store y, update a smoother or
//      Krylov vector, accumulate a dot product).  Matrices with long rows
//      that need no bit-compatibility (the AMG hierarchy's own operators)
//      give each row lpr lanes instead (StreamMat::lpr).
// Products are rounded and then added left to right — the arithmetic of
// CSRMatrix::spmv (src/solver/CSRMatrix.hpp:25-35) — so with contraction off
// the FP64 results are identical to the reference's serial loop.  A row
// longer than kStreamNnz is one block of its own, summed with a block-wide
// reduction (fixed order, deterministic).  FP32 matrices may instead carry
// the difference form (stream_rows_diff), selected by StreamMat::diag.
//
// Column-ordered blocks (StreamMat::cw): the gathers of x, not the streams of
// colIdx/values, bound this engine (a gather instruction whose 64 lanes hit
// scattered cache lines costs several times a coalesced one).  The nonzeros
// of every row block can therefore be stored in column order, each with the
// CSR slot it came from: a wavefront's gather then walks ascending, mostly
// adjacent columns, and the product still lands in LDS at its CSR slot, so
// the row sums and their order are unchanged.
//
// Workgroups are persistent over a contiguous range of row blocks; the range
// of hardware block b follows xcd_logical_block so each XCD walks one slice
// of rows and keeps its part of x in its own L2.
#pragma once

#include "krylov.hpp"

#include <type_traits>
#include <vector>

namespace mi
{

constexpr int kStreamRows = 256; // rows per row block (one per thread)
constexpr int kStreamNnz = 2048; // nonzeros per row block (LDS products)
constexpr int kLoadsPerThread = kStreamNnz / kBlock; // loads per thread per row block

// Host partition of rows into row blocks of at most max_rows rows and
// kStreamNnz nonzeros: rows blk[b] .. blk[b+1].
void build_row_blocks(int n, const int* rp, std::vector<int>& blk, int max_rows = kStreamRows);

// Lanes per row for the summation phase: 1 (column-order sums, the
// reference's arithmetic) unless `wide` and the mean row holds 10 or more
// nonzeros, then the largest power of two <= mean/3, at most 8 (measured
// best on the SA-AMG levels of a 2M-row 3D pressure system).
int stream_lanes_per_row(int64_t n, int64_t nnz, bool wide);

// Packed entry of a column-ordered row block: slot << 21 | (col - c0), c0 =
// the block's smallest column, slot = position in the block's CSR span (0
// for a block of one long row, whose column order is its CSR order).
constexpr int kSlotShift = 21;
constexpr unsigned kColMask = (1u << kSlotShift) - 1u;

struct RowBlocks
{
    DevBuf<int> blk;
    // {first row, first nonzero, smallest column, 0} of every block boundary
    // (nrb + 1 quadruples; the column base is set by sort_columns): one load
    // instead of blk -> rowPtr at the start of a block
    DevBuf<int> hdr;
    int nrb = 0;
    int lpr = 1; // lanes per row (a power of two); rows per block <= kStreamRows / lpr
    // Column order (sort_columns): cw[m] packed entries, sperm[m] = CSR
    // position of the m-th stored nonzero; each block's column base is in hdr.
    // Empty when not built or when a block's columns span 2^21 - 1 or more.
    DevBuf<unsigned> cw;
    DevBuf<int> sperm;
    bool sorted() const { return cw.n > 0; }
    int upload(int n, const int* rp_host, hipStream_t s, int lanes_per_row = 1);
    // Builds cw / sperm on the device (one workgroup per row block sorts its
    // columns in LDS); leaves them empty for out-of-range columns.
    int sort_columns(const int* rp_dev, const int* ci_dev, int64_t nnz, hipStream_t s);
};

// dst[m] = src[perm[m]]: values into a row-block column order.
template <class T>
int launch_permute(int64_t nnz, const int* perm, const T* src, T* dst, hipStream_t s);

// SELL-64 form of a matrix whose rows are all short (the 7-point pressure
// operator, the prolongator): slices of kSellSlice consecutive rows, one
// wavefront per slice and one lane per row.  Slot k of the slice's lane l
// sits at soff[s] + 64 k + l and holds the row's k-th CSR entry (column -1:
// padding), so every lane sums its row in CSR order with no LDS staging and
// no barrier, and every stream load is one coalesced 64-lane access.
// Measured on the 2M-row pressure matrix against column-ordered row blocks:
// FP64 35.8 vs 41.3 us, FP32 21.7 vs 28.4 us (results identical).
constexpr int kSellSlice = 64;
constexpr int kSellMaxWidth = 16;        // longest row the form stores
constexpr int kSellUnroll = 8;           // slots loaded ahead of the gathers
constexpr int kSellDiag = 1 << 30;       // marks a row's diagonal slot (square operators)
constexpr int kSellColMask = kSellDiag - 1;
constexpr short kSellDelta = -32768;     // padding of the 16-bit column form
constexpr int kMaxSellBlocks = 8192;     // workgroups (4 slices each) before slices per wave grow

struct SellForm
{
    DevBuf<int> soff;  // nsl + 1 slot offsets (multiples of 64)
    DevBuf<int> eci;   // column per slot, kSellDiag on the diagonal slot, -1 padding
    DevBuf<int> eperm; // CSR position per slot, -1 padding (value permutation)
    // Square operators whose entries all lie within 32767 of the diagonal
    // and whose rows hold one diagonal entry each: column - row per slot as
    // int16 (kSellDelta pads; 0 is the diagonal), streamed by the FP32
    // kernels in place of eci (6 instead of 8 bytes per slot).  Empty otherwise.
    DevBuf<short> ecd;
    int nsl = 0;
    int64_t padded = 0;
    bool built() const { return nsl > 0; }
    // Host build from CSR arrays; leaves the form empty (and returns MI_OK)
    // when a row holds more than kSellMaxWidth entries, padding would exceed
    // a quarter of nnz, or n or a column index reaches 2^30.  mark_diag: flag the first entry with
    // column == row in each row (the difference form's diagonal slot).
    int build(int n, int64_t nnz, const int* rp_host, const int* ci_host, bool mark_diag,
              hipStream_t s);
};

// Whether patterns get a SELL-64 form (default on; SIMALL_AMD_NO_SELL=1 keeps
// every matrix on row blocks, for comparisons).
bool sell_enabled();

// Workgroups of a SELL launch: one per 4 slices up to kMaxSellBlocks (a
// multiple of 8); kernels with reduction partials size them for that.
inline unsigned sell_grid(int nsl)
{
    int64_t g = (int64_t(nsl) + 3) / 4;
    if (g > kMaxSellBlocks)
        g = kMaxSellBlocks;
    g = (g + 7) / 8 * 8;
    if (g < 8)
        g = 8;
    return unsigned(g);
}

// Workgroups for a stream launch: one per row block up to kMaxStreamBlocks
// (a multiple of 8), beyond that each workgroup walks a contiguous range.
// Persistent workgroups measured faster than one block per workgroup on the
// 2M-row pressure solve (FP32 blocks prefetch the next block's streams).
// Kernels that write one reduction partial per workgroup size their buffers
// for kMaxStreamBlocks.
constexpr int kMaxStreamBlocks = 2048;
inline unsigned stream_grid(int nrb)
{
    int64_t g = nrb;
    if (g > kMaxStreamBlocks)
        g = kMaxStreamBlocks;
    g = (g + 7) / 8 * 8;
    if (g < 8)
        g = 8;
    return unsigned(g);
}

#if defined(__HIPCC__)

#pragma clang fp contract(off)

template <class T> struct StreamMat
{
    int n;
    int nrb;
    const int* blk;
    const int* rp;
    const int* ci;
    const T* val;
    // FP32 only: position of the diagonal in each row when val holds the
    // difference form (see stream_rows_diff); nullptr for the plain form.
    const int* diag = nullptr;
    // Lanes per row of the summation phase (power of two, the RowBlocks'
    // lpr).  1 sums each row in column order on one thread; more lanes take
    // every lpr-th product and combine with xor shuffles (fixed order).
    int lpr = 1;
    // Column-ordered blocks: packed entries (RowBlocks::cw), column bases in
    // hdr (required with cw); val is then in the same order.  nullptr: ci and
    // val in CSR order.
    const unsigned* cw = nullptr;
    // Block boundaries as {first row, first nonzero, column base, 0}
    // (RowBlocks::hdr); nullptr: read blk and rp.
    const int* hdr = nullptr;
    // SELL-64 form (SellForm): slot offsets and columns; val then holds the
    // padded values in slot order and the row-block fields are unused.
    const int* soff = nullptr;
    const int* eci = nullptr;
    int nsl = 0;
    // FP32 SELL-64 only: 16-bit column offsets from the row (SellForm::ecd),
    // read in place of eci when set.
    const short* ecd = nullptr;
};

// Workgroups for a launch over A in whichever form it carries.
template <class T> inline unsigned launch_grid(const StreamMat<T>& A)
{
    return A.soff ? sell_grid(A.nsl) : stream_grid(A.nrb);
}

// Rows r0..r1, nonzeros k0..k1 and column base c0 of row block rb.
template <class T>
__device__ __forceinline__ void stream_block(const StreamMat<T>& A, int rb, int& r0, int& r1,
                                             int& k0, int& k1, int& c0)
{
    if (A.hdr) {
        const int4 h0 = reinterpret_cast<const int4*>(A.hdr)[rb];
        const int4 h1 = reinterpret_cast<const int4*>(A.hdr)[rb + 1];
        r0 = h0.x;
        k0 = h0.y;
        c0 = h0.z;
        r1 = h1.x;
        k1 = h1.y;
    } else {
        r0 = A.blk[rb];
        r1 = A.blk[rb + 1];
        k0 = A.rp[r0];
        k1 = A.rp[r1];
        c0 = 0;
    }
}

// Column of a stream entry: the plain column, or decoded from a packed one
// relative to the block's column base c0.
template <class T>
__device__ __forceinline__ int stream_col(const StreamMat<T>& A, int e, int c0)
{
    return A.cw ? c0 + int(unsigned(e) & kColMask) : e;
}

// LDS slot of the j-th entry of thread tid.
template <class T>
__device__ __forceinline__ int stream_slot(const StreamMat<T>& A, int e, int tid, int j)
{
    return A.cw ? int(unsigned(e) >> kSlotShift) : tid + j * kBlock;
}

// Sum over the lpr lanes of a row group (aligned within the wavefront).
template <class Acc> __device__ __forceinline__ Acc lane_group_sum(Acc s, int lpr)
{
    for (int off = lpr >> 1; off >= 1; off >>= 1)
        s += __shfl_xor(s, off);
    return s;
}

// One 16 KB LDS buffer per workgroup shared by both row forms: 2048 FP64
// products, or 2048 FP32 values plus 2048 gathered FP32 x entries.
__device__ __forceinline__ double* stream_lds()
{
    __shared__ double buf[kStreamNnz];
    return buf;
}

// Column/value streams of one row block (nonzeros k0..k1, at most
// kStreamNnz) into registers: kLoadsPerThread independent loads per thread,
// c[j] = -1 past the end (never a packed entry, whose column field is below
// 2^21 - 1).
template <class T>
__device__ __forceinline__ void stream_load(const StreamMat<T>& A, int k0, int k1, int tid,
                                            int (&c)[kLoadsPerThread], T (&a)[kLoadsPerThread])
{
    const int* src = A.cw ? reinterpret_cast<const int*>(A.cw) : A.ci;
#pragma unroll
    for (int j = 0; j < kLoadsPerThread; ++j) {
        const int k = k0 + tid + j * kBlock;
        c[j] = k < k1 ? __builtin_nontemporal_load(&src[k]) : -1;
        a[j] = k < k1 ? __builtin_nontemporal_load(&A.val[k]) : T(0);
    }
}

// Prefetch of the block after rb (same workgroup range) when it is a regular
// block; returns whether c/a now hold it.  The loads stay in flight while
// the current block's rows are summed.
template <class T>
__device__ __forceinline__ bool stream_prefetch(const StreamMat<T>& A, int rb, int rb1, int k1,
                                                int tid, int (&c)[kLoadsPerThread],
                                                T (&a)[kLoadsPerThread])
{
    if (rb + 1 >= rb1)
        return false;
    const int q1 = A.hdr ? A.hdr[4 * (rb + 2) + 1] : A.rp[A.blk[rb + 2]];
    if (q1 - k1 > kStreamNnz)
        return false;
    stream_load(A, k1, q1, tid, c, a);
    return true;
}

// Difference form for FP32 operators: A x = sum_{j != i} a_ij (x_j - x_i)
// + d_i x_i, where the diagonal slot holds d_i = sum_j a_ij (the FP64 row
// sum, rounded once).  Rounding the coefficients then perturbs the
// conductances and the row sums separately instead of destroying the zero
// row sums of a Neumann Laplacian: the FP32 operator stays spectrally within
// FP32 rounding of the FP64 one, and smooth error is not amplified.
template <class Epi>
__device__ __forceinline__ void stream_rows_diff(const StreamMat<float>& A,
                                                 const float* __restrict__ x, Epi& epi)
{
    float* av = reinterpret_cast<float*>(stream_lds());
    float* xv = av + kStreamNnz;
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int per = (A.nrb + G - 1) / G;
    const int rb0 = lb * per;
    const int rb1 = min(A.nrb, rb0 + per);
    const int tid = threadIdx.x;
    int c[kLoadsPerThread];
    float a[kLoadsPerThread];
    bool have = false;
    for (int rb = rb0; rb < rb1; ++rb) {
        int r0, r1, k0, k1, c0;
        stream_block(A, rb, r0, r1, k0, k1, c0);
        if (k1 - k0 > kStreamNnz) {
            if (tid == 0)
                epi.pre(r0);
            const int kd = A.diag[r0];
            const float xi = x[r0];
            float s = 0.f;
            const int* src = A.cw ? reinterpret_cast<const int*>(A.cw) : A.ci;
            for (int k = k0 + tid; k < k1; k += kBlock) {
                const float v = __builtin_nontemporal_load(&A.val[k]);
                const float xj = x[stream_col(A, __builtin_nontemporal_load(&src[k]), c0)];
                s += (k == kd) ? v * xi : v * (xj - xi);
            }
            av[tid] = s;
            __syncthreads();
            if (tid == 0) {
                float t = 0.f;
                for (int j = 0; j < kBlock; ++j)
                    t += av[j];
                epi.row(r0, t);
            }
            __syncthreads();
            have = false;
            continue;
        }
        // this thread's row and its epilogue operands, loaded before the
        // products so their latency overlaps the gathers and the barrier
        const int lpr = A.lpr;
        const int sub = tid & (lpr - 1);
        const int i = r0 + tid / lpr;
        const bool mine = i < r1 && sub == 0;
        if (mine)
            epi.pre(i);
        // the row's extent and diagonal slot, ahead of the barrier
        int kb = 0, ke = 0, kd = 0;
        if (i < r1) {
            kb = A.rp[i] - k0;
            ke = A.rp[i + 1] - k0;
            kd = A.diag[i] - k0;
        }
        if (!have)
            stream_load(A, k0, k1, tid, c, a);
#pragma unroll
        for (int j = 0; j < kLoadsPerThread; ++j) {
            if (c[j] != -1) {
                const int q = stream_slot(A, c[j], tid, j);
                av[q] = a[j];
                xv[q] = x[stream_col(A, c[j], c0)];
            }
        }
        __syncthreads();
        have = stream_prefetch(A, rb, rb1, k1, tid, c, a);
        float s = 0.f;
        if (i < r1) {
            const float xi = xv[kd];
            for (int k = kb + sub; k < ke; k += lpr)
                s += (k == kd) ? av[k] * xi : av[k] * (xv[k] - xi);
        }
        if (lpr > 1)
            s = lane_group_sum(s, lpr);
        if (mine)
            epi.row(i, s);
        __syncthreads();
    }
}

// SELL-64 rows: wavefront w of workgroup b takes slice 4q + w for the
// workgroup's contiguous range of slice quads q (one quad per workgroup when
// the grid is sell_grid's).  Each lane sums its row in CSR order, in the
// arithmetic of the row-block paths: plain products a_ij x_j, or for an FP32
// difference-form operator (A.diag set) a_ij (x_j - x_i) off the diagonal and
// d_i x_i on it.  No barrier: lanes past the last slice simply return.
// The FP32 SELL-64 rows with 16-bit column offsets (A.ecd): the same FP32
// sums in the same order as sell_rows (identical results), the column of
// slot k being i + ecd[k] (0: the diagonal).
template <class X, class Epi>
__device__ __forceinline__ void sell_rows_d16(const StreamMat<float>& A, const X* __restrict__ x,
                                              Epi& epi)
{
    const bool diff = A.diag != nullptr;
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int nq = (A.nsl + 3) >> 2;
    const int per = (nq + G - 1) / G;
    const int q0 = lb * per;
    const int q1 = min(nq, q0 + per);
    const int wave = int(threadIdx.x) / kSellSlice;
    const int lane = int(threadIdx.x) % kSellSlice;
    for (int q = q0; q < q1; ++q) {
        const int sl = q * 4 + wave;
        if (sl >= A.nsl)
            return;
        const int i = sl * kSellSlice + lane;
        const bool mine = i < A.n;
        if (mine)
            epi.pre(i);
        const int o0 = A.soff[sl] + lane;
        const int w = (A.soff[sl + 1] - A.soff[sl]) / kSellSlice;
        int d[kSellUnroll];
        float a[kSellUnroll];
#pragma unroll
        for (int k = 0; k < kSellUnroll; ++k) {
            d[k] = k < w ? int(__builtin_nontemporal_load(&A.ecd[o0 + k * kSellSlice])) : kSellDelta;
            a[k] = k < w ? __builtin_nontemporal_load(&A.val[o0 + k * kSellSlice]) : 0.0f;
        }
        X xv[kSellUnroll];
#pragma unroll
        for (int k = 0; k < kSellUnroll; ++k)
            xv[k] = d[k] != kSellDelta ? x[i + d[k]] : X(0);
        float s = 0;
        if (diff) {
            const X xi = mine ? x[i] : X(0);
#pragma unroll
            for (int k = 0; k < kSellUnroll; ++k)
                if (d[k] != kSellDelta)
                    s += (d[k] == 0) ? float(a[k] * xi) : float(a[k] * (xv[k] - xi));
            for (int k = kSellUnroll; k < w; ++k) {
                const int dk = A.ecd[o0 + k * kSellSlice];
                if (dk != kSellDelta) {
                    const float ak = A.val[o0 + k * kSellSlice];
                    s += (dk == 0) ? float(ak * xi) : float(ak * (x[i + dk] - xi));
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kSellUnroll; ++k)
                if (d[k] != kSellDelta)
                    s += float(a[k]) * float(xv[k]);
            for (int k = kSellUnroll; k < w; ++k) {
                const int dk = A.ecd[o0 + k * kSellSlice];
                if (dk != kSellDelta)
                    s += float(A.val[o0 + k * kSellSlice]) * float(x[i + dk]);
            }
        }
        if (mine)
            epi.row(i, s);
    }
}

template <class T, class X, class Epi>
__device__ __forceinline__ void sell_rows(const StreamMat<T>& A, const X* __restrict__ x, Epi& epi)
{
    using Acc = typename acc_type<T>::type;
    constexpr bool kDiffCapable = std::is_same<T, float>::value && std::is_same<X, float>::value;
    if constexpr (kDiffCapable) {
        if (A.ecd) {
            sell_rows_d16(A, x, epi);
            return;
        }
    }
    const bool diff = kDiffCapable && A.diag != nullptr;
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int nq = (A.nsl + 3) >> 2;
    const int per = (nq + G - 1) / G;
    const int q0 = lb * per;
    const int q1 = min(nq, q0 + per);
    const int wave = int(threadIdx.x) / kSellSlice;
    const int lane = int(threadIdx.x) % kSellSlice;
    for (int q = q0; q < q1; ++q) {
        const int sl = q * 4 + wave;
        if (sl >= A.nsl)
            return;
        const int i = sl * kSellSlice + lane;
        const bool mine = i < A.n;
        if (mine)
            epi.pre(i);
        const int o0 = A.soff[sl] + lane;
        const int w = (A.soff[sl + 1] - A.soff[sl]) / kSellSlice;
        int c[kSellUnroll];
        T a[kSellUnroll];
#pragma unroll
        for (int k = 0; k < kSellUnroll; ++k) {
            c[k] = k < w ? __builtin_nontemporal_load(&A.eci[o0 + k * kSellSlice]) : -1;
            a[k] = k < w ? __builtin_nontemporal_load(&A.val[o0 + k * kSellSlice]) : T(0);
        }
        X xv[kSellUnroll];
#pragma unroll
        for (int k = 0; k < kSellUnroll; ++k)
            xv[k] = c[k] >= 0 ? x[c[k] & kSellColMask] : X(0);
        Acc s = 0;
        if (kDiffCapable && diff) {
            const X xi = mine ? x[i] : X(0);
#pragma unroll
            for (int k = 0; k < kSellUnroll; ++k)
                if (c[k] >= 0)
                    s += (c[k] & kSellDiag) ? Acc(a[k] * xi) : Acc(a[k] * (xv[k] - xi));
            for (int k = kSellUnroll; k < w; ++k) {
                const int ck = A.eci[o0 + k * kSellSlice];
                if (ck >= 0) {
                    const T ak = A.val[o0 + k * kSellSlice];
                    s += (ck & kSellDiag) ? Acc(ak * xi) : Acc(ak * (x[ck & kSellColMask] - xi));
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < kSellUnroll; ++k)
                if (c[k] >= 0)
                    s += Acc(a[k]) * Acc(xv[k]);
            for (int k = kSellUnroll; k < w; ++k) {
                const int ck = A.eci[o0 + k * kSellSlice];
                if (ck >= 0)
                    s += Acc(A.val[o0 + k * kSellSlice]) * Acc(x[ck & kSellColMask]);
            }
        }
        if (mine)
            epi.row(i, s);
    }
}

// Runs epi.row(i, sum_i) for every row of A handled by this workgroup.  The
// epilogue's pre(i) is called first, on the same thread, at the start of the
// row's block: it loads the per-row operands row(i, .) will use.  Row blocks
// only: kernels that also take SELL-64 views are instantiated once per form
// (rows_of below), so neither path's registers limit the other's occupancy
// (one kernel holding both ran the 30-nonzero-per-row FP32 level at 63
// instead of 40 us).
template <class T, class X, class Epi>
__device__ __forceinline__ void stream_rows(const StreamMat<T>& A, const X* __restrict__ x, Epi& epi)
{
    if constexpr (std::is_same<T, float>::value && std::is_same<X, float>::value) {
        if (A.diag) {
            stream_rows_diff(A, x, epi);
            return;
        }
    }
    using Acc = typename acc_type<T>::type;
    Acc* prod = reinterpret_cast<Acc*>(stream_lds());
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int per = (A.nrb + G - 1) / G;
    const int rb0 = lb * per;
    const int rb1 = min(A.nrb, rb0 + per);
    const int tid = threadIdx.x;
    int c[kLoadsPerThread];
    T a[kLoadsPerThread];
    bool have = false;
    for (int rb = rb0; rb < rb1; ++rb) {
        int r0, r1, k0, k1, c0;
        stream_block(A, rb, r0, r1, k0, k1, c0);
        if (k1 - k0 > kStreamNnz) {
            // one long row: strided partial sums, fixed-order block reduction
            if (tid == 0)
                epi.pre(r0);
            const int* src = A.cw ? reinterpret_cast<const int*>(A.cw) : A.ci;
            Acc s = 0;
            for (int k = k0 + tid; k < k1; k += kBlock)
                s += Acc(__builtin_nontemporal_load(&A.val[k])) *
                     Acc(x[stream_col(A, __builtin_nontemporal_load(&src[k]), c0)]);
            prod[tid] = s;
            __syncthreads();
            if (tid == 0) {
                Acc t = 0;
                for (int j = 0; j < kBlock; ++j)
                    t += prod[j];
                epi.row(r0, t);
            }
            __syncthreads();
            have = false;
            continue;
        }
        // this thread's row and its epilogue operands, loaded before the
        // products so their latency overlaps the gathers and the barrier
        const int lpr = A.lpr;
        const int sub = tid & (lpr - 1);
        const int i = r0 + tid / lpr;
        const bool mine = i < r1 && sub == 0;
        if (mine)
            epi.pre(i);
        // the row's extent, ahead of the barrier
        int kb = 0, ke = 0;
        if (i < r1) {
            kb = A.rp[i] - k0;
            ke = A.rp[i + 1] - k0;
        }
        if (!have)
            stream_load(A, k0, k1, tid, c, a);
#pragma unroll
        for (int j = 0; j < kLoadsPerThread; ++j)
            if (c[j] != -1)
                prod[stream_slot(A, c[j], tid, j)] = Acc(a[j]) * Acc(x[stream_col(A, c[j], c0)]);
        __syncthreads();
        // FP32 blocks are short on bytes per synchronisation, so the next
        // block's streams are prefetched; for FP64 the extra live registers
        // cost more occupancy than the overlap gains (measured)
        if constexpr (sizeof(T) == 4)
            have = stream_prefetch(A, rb, rb1, k1, tid, c, a);
        Acc s = 0;
        if (i < r1)
            for (int k = kb + sub; k < ke; k += lpr)
                s += prod[k];
        if (lpr > 1)
            s = lane_group_sum(s, lpr);
        if (mine)
            epi.row(i, s);
        __syncthreads();
    }
}

// The row engine of a kernel instantiated for one form: SELL-64 (kSell) or
// row blocks.  Launch sites pick the instance from StreamMat::soff.
template <bool kSell, class T, class X, class Epi>
__device__ __forceinline__ void rows_of(const StreamMat<T>& A, const X* __restrict__ x, Epi& epi)
{
    if constexpr (kSell)
        sell_rows(A, x, epi);
    else
        stream_rows(A, x, epi);
}

#pragma clang fp contract(on)

#endif

} // namespace mi
