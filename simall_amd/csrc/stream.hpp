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
// reduction (fixed order, deterministic).
//
// Workgroups are persistent over a contiguous range of row blocks; the range
// of hardware block b follows xcd_logical_block so each XCD walks one slice
// of rows and keeps its part of x in its own L2.
#pragma once

#include "krylov.hpp"

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
};

// Runs epi.row(i, sum_i) for every row of A handled by this workgroup.
template <class T, class X, class Epi>
__device__ __forceinline__ void stream_rows(const StreamMat<T>& A, const X* __restrict__ x, Epi& epi)
{
    using Acc = typename acc_type<T>::type;
    __shared__ Acc prod[kStreamNnz];
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
