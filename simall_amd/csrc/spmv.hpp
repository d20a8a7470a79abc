// spmv.hpp — CSR row kernels shared by SpMV and the fused Krylov kernels.
//
// Layout: int32 rowPtr/colIdx, values in T (double or float).  A "vector"
// row kernel gives each row LANES consecutive lanes of a 64-wide wavefront
// (LANES = 64 is one wavefront per row).  Lanes stride the row's entries so
// each load instruction of a wavefront reads a contiguous span of colIdx and
// values; the segmented sum then folds the LANES partial sums with xor
// shuffles.  Matrix streams are read with non-temporal loads so that the
// cache keeps the gathered x vector.
#pragma once

#include "mi_common.hpp"

namespace mi
{

template <class T> struct acc_type;
template <> struct acc_type<double>
{
    using type = double;
};
template <> struct acc_type<float>
{
    using type = float;
};

#if defined(__HIPCC__)

template <int LANES, class Acc>
__device__ __forceinline__ Acc segment_sum(Acc s)
{
#pragma unroll
    for (int off = LANES / 2; off >= 1; off >>= 1)
        s += __shfl_xor(s, off, LANES);
    return s;
}

// Sum over row `row` of values[k] * x[colIdx[k]], valid in every lane of the
// row's lane group.  `row` may be >= n (returns 0) so that whole wavefronts
// stay converged for the shuffles.
template <int LANES, class VT, class XT>
__device__ __forceinline__ typename acc_type<VT>::type
csr_row_dot(int row, int n, int lane, const int* __restrict__ rp, const int* __restrict__ ci,
            const VT* __restrict__ val, const XT* __restrict__ x)
{
    using Acc = typename acc_type<VT>::type;
    Acc s = 0;
    if (row < n) {
        const int beg = rp[row], end = rp[row + 1];
        for (int k = beg + lane; k < end; k += LANES) {
            const int c = __builtin_nontemporal_load(&ci[k]);
            const VT v = __builtin_nontemporal_load(&val[k]);
            s += Acc(v) * Acc(x[c]);
        }
    }
    return segment_sum<LANES, Acc>(s);
}

#endif

// Lanes per row from the mean row length: the smallest power of two >= the
// mean, clamped to [2, 64].
inline int lanes_for(int64_t n, int64_t nnz)
{
    const double mean = n > 0 ? double(nnz) / double(n) : 1.0;
    int l = 2;
    while (l < 64 && l < mean)
        l *= 2;
    return l;
}

// Host launchers (spmv.hip).  All are asynchronous on `s`.
int launch_spmv_f64(int n, int64_t nnz, const int* rp, const int* ci, const double* val,
                    const double* x, double* y, hipStream_t s, int lanes = 0);
int launch_spmv_f32(int n, int64_t nnz, const int* rp, const int* ci, const float* val,
                    const float* x, float* y, hipStream_t s, int lanes = 0);
// y = A x with the CSR-stream engine on a pattern's row blocks.
struct Pattern;
template <class T> int launch_spmv_stream(const Pattern& P, const T* val, const T* x, T* y,
                                          hipStream_t s);
#if defined(__HIPCC__)
// y = A x over a stream view (CSR or column-ordered blocks).
template <class T> struct StreamMat;
template <class T> int launch_spmv_view(const StreamMat<T>& A, const T* x, T* y, hipStream_t s);
#endif
// r = b - A x  (FP64)
int launch_residual_f64(int n, int64_t nnz, const int* rp, const int* ci, const double* val,
                        const double* x, const double* b, double* r, hipStream_t s, int lanes = 0);

} // namespace mi
