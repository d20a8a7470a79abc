// stream.hpp — the CSR-stream row engine shared by every SpMV-shaped kernel.
//
// Rows are grouped on the host into row blocks of at most kStreamRows rows
// and kStreamNnz nonzeros.  For one row block the workgroup
//   1. streams the block's contiguous span of colIdx/values with coalesced
//      loads, forms every product a_ij * x_j (x gathered through the cache)
//      and parks it in LDS;
//   2. gives each row to one thread, which adds its products in column order
//      and hands (row, sum) to an epilogue (store y, update a smoother or
//      Krylov vector, accumulate a dot product).
// Products are rounded and then added left to right — the arithmetic of
// CSRMatrix::spmv (src/solver/CSRMatrix.hpp:25-35) — so with contraction off
// the FP64 results are identical to the reference's serial loop.  A row
// longer than kStreamNnz is one block of its own, summed with a block-wide
// reduction (fixed order, deterministic).  FP32 matrices may instead carry
// the difference form (stream_rows_diff), selected by StreamMat::diag.
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

// Host partition of rows into row blocks: rows blk[b] .. blk[b+1].
void build_row_blocks(int n, const int* rp, std::vector<int>& blk);

struct RowBlocks
{
    DevBuf<int> blk;
    int nrb = 0;
    int upload(int n, const int* rp_host, hipStream_t s);
};

// Workgroups for a stream launch: enough to cover the row blocks, a multiple
// of 8, at most kMaxReduceBlocks.
inline unsigned stream_grid(int nrb)
{
    int64_t g = nrb;
    if (g > kMaxReduceBlocks)
        g = kMaxReduceBlocks;
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
};

// One 16 KB LDS buffer per workgroup shared by both row forms: 2048 FP64
// products, or 2048 FP32 values plus 2048 gathered FP32 x entries.
__device__ __forceinline__ double* stream_lds()
{
    __shared__ double buf[kStreamNnz];
    return buf;
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
    for (int rb = rb0; rb < rb1; ++rb) {
        const int r0 = A.blk[rb], r1 = A.blk[rb + 1];
        const int k0 = A.rp[r0], k1 = A.rp[r1];
        if (k1 - k0 > kStreamNnz) {
            const int kd = A.diag[r0];
            const float xi = x[r0];
            float s = 0.f;
            for (int k = k0 + tid; k < k1; k += kBlock) {
                const float a = __builtin_nontemporal_load(&A.val[k]);
                const float xj = x[__builtin_nontemporal_load(&A.ci[k])];
                s += (k == kd) ? a * xi : a * (xj - xi);
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
            continue;
        }
        for (int k = k0 + tid; k < k1; k += kBlock) {
            av[k - k0] = __builtin_nontemporal_load(&A.val[k]);
            xv[k - k0] = x[__builtin_nontemporal_load(&A.ci[k])];
        }
        __syncthreads();
        const int i = r0 + tid;
        if (i < r1) {
            const int kd = A.diag[i] - k0;
            const float xi = xv[kd];
            float s = 0.f;
            const int e = A.rp[i + 1] - k0;
            for (int k = A.rp[i] - k0; k < e; ++k)
                s += (k == kd) ? av[k] * xi : av[k] * (xv[k] - xi);
            epi.row(i, s);
        }
        __syncthreads();
    }
}

// Runs epi.row(i, sum_i) for every row of A handled by this workgroup.
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
    for (int rb = rb0; rb < rb1; ++rb) {
        const int r0 = A.blk[rb], r1 = A.blk[rb + 1];
        const int k0 = A.rp[r0], k1 = A.rp[r1];
        if (k1 - k0 > kStreamNnz) {
            // one long row: strided partial sums, fixed-order block reduction
            Acc s = 0;
            for (int k = k0 + tid; k < k1; k += kBlock)
                s += Acc(__builtin_nontemporal_load(&A.val[k])) *
                     Acc(x[__builtin_nontemporal_load(&A.ci[k])]);
            prod[tid] = s;
            __syncthreads();
            if (tid == 0) {
                Acc t = 0;
                for (int j = 0; j < kBlock; ++j)
                    t += prod[j];
                epi.row(r0, t);
            }
            __syncthreads();
            continue;
        }
        for (int k = k0 + tid; k < k1; k += kBlock)
            prod[k - k0] = Acc(__builtin_nontemporal_load(&A.val[k])) *
                           Acc(x[__builtin_nontemporal_load(&A.ci[k])]);
        __syncthreads();
        const int i = r0 + tid;
        if (i < r1) {
            Acc s = 0;
            const int e = A.rp[i + 1] - k0;
            for (int k = A.rp[i] - k0; k < e; ++k)
                s += prod[k];
            epi.row(i, s);
        }
        __syncthreads();
    }
}

#pragma clang fp contract(on)

#endif

} // namespace mi
