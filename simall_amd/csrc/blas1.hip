// blas1.hip — the GpuKernels free functions (src/gpu/GpuKernels.hpp:47-108)
// as gfx950 kernels on device pointers.  Reductions run as one deterministic
// two-level launch (block partials, last block folds them in fixed order).
#include "blas1.hpp"
#include "krylov.hpp"
#include "spmv.hpp"

namespace mi
{

__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a, const double* __restrict__ x,
                                                  double b, double* __restrict__ y, int mode)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
        switch (mode) {
        case 0: y[i] += a * x[i]; break;          // axpy
        case 1: y[i] *= a; break;                 // scal
        case 2: y[i] = x[i]; break;               // copy
        case 3: y[i] = a * x[i] + b * y[i]; break; // axpby
        default: y[i] = x[i] + b * y[i]; break;   // xpby
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_dot(int64_t n, const double* __restrict__ x,
                                                const double* __restrict__ y,
                                                double* __restrict__ partials)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    double v[1] = {0.0};
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        v[0] += x[i] * y[i];
    block_partials<1>(v, partials);
}

__global__ __launch_bounds__(kBlock) void k_jacobi(int64_t n, const double* __restrict__ diag,
                                                   const double* __restrict__ r,
                                                   double* __restrict__ z)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        z[i] = r[i] / diag[i];
}

int ReduceScratch::ensure(hipStream_t s)
{
    if (partials.p)
        return MI_OK;
    MI_TRY(partials.alloc(size_t(kMaxReduceBlocks) * kMaxReduceValues));
    MI_TRY(result.alloc(kMaxReduceValues));
    (void)s;
    return MI_OK;
}

int launch_vec(int mode, int64_t n, double a, const double* x, double b, double* y, hipStream_t s)
{
    if (n <= 0)
        return MI_OK;
    k_axpby<<<reduce_grid(n, 4), kBlock, 0, s>>>(n, a, x, b, y, mode);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int launch_dot(int64_t n, const double* x, const double* y, ReduceScratch& rs, double* out_dev,
               hipStream_t s)
{
    MI_TRY(rs.ensure(s));
    const unsigned G = reduce_grid(n, 8);
    k_dot<<<G, kBlock, 0, s>>>(n, x, y, rs.partials.p);
    MI_HIP(hipGetLastError());
    return launch_fin_sum(rs.partials.p, G, false, out_dev, s);
}

static ReduceScratch& facade_scratch()
{
    static ReduceScratch rs;
    return rs;
}

} // namespace mi

using namespace mi;

static hipStream_t as_stream(void* s) { return s ? static_cast<hipStream_t>(s) : default_stream(); }

extern "C" {

MI_API int mi_axpy(double alpha, const void* x, void* y, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && (!x || !y))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    return launch_vec(0, int64_t(n), alpha, static_cast<const double*>(x), 0.0,
                      static_cast<double*>(y), as_stream(stream));
}

MI_API int mi_scal(double alpha, void* x, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && !x)
        return set_error(MI_ERR_BAD_ARG, "null vector");
    return launch_vec(1, int64_t(n), alpha, nullptr, 0.0, static_cast<double*>(x),
                      as_stream(stream));
}

MI_API int mi_copy(const void* src, void* dst, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && (!src || !dst))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    return launch_vec(2, int64_t(n), 0.0, static_cast<const double*>(src), 0.0,
                      static_cast<double*>(dst), as_stream(stream));
}

MI_API int mi_axpby(double a, const void* x, double b, void* y, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && (!x || !y))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    return launch_vec(3, int64_t(n), a, static_cast<const double*>(x), b,
                      static_cast<double*>(y), as_stream(stream));
}

MI_API int mi_xpby(const void* x, double beta, void* y, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && (!x || !y))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    return launch_vec(4, int64_t(n), 0.0, static_cast<const double*>(x), beta,
                      static_cast<double*>(y), as_stream(stream));
}

MI_API int mi_dot(const void* x, const void* y, size_t n, double* result, void* stream)
{
    MI_TRY(ensure_device());
    if (!result || (n && (!x || !y)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    ReduceScratch& rs = facade_scratch();
    hipStream_t s = as_stream(stream);
    MI_TRY(rs.ensure(s));
    MI_TRY(launch_dot(int64_t(n), static_cast<const double*>(x), static_cast<const double*>(y), rs,
                      rs.result.p, s));
    MI_HIP(hipMemcpyAsync(result, rs.result.p, sizeof(double), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_nrm2(const void* x, size_t n, double* result, void* stream)
{
    MI_TRY(mi_dot(x, x, n, result, stream));
    *result = std::sqrt(*result);
    return MI_OK;
}

MI_API int mi_spmv_csr(size_t n, const void* rowPtr, const void* colIdx, const void* values,
                       const void* x, void* y, void* stream)
{
    MI_TRY(ensure_device());
    if (n == 0)
        return MI_OK;
    if (!rowPtr || !x || !y)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    if (n > size_t(INT32_MAX))
        return set_error(MI_ERR_BAD_ARG, "n exceeds int32 range");
    // The mean row length picks the lane count; without host access to rowPtr
    // the façade reads rowPtr[n] once.
    int nnz = 0;
    hipStream_t s = as_stream(stream);
    MI_HIP(hipMemcpyAsync(&nnz, static_cast<const int*>(rowPtr) + n, sizeof(int),
                          hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    if (nnz > 0 && (!colIdx || !values))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    return launch_spmv_f64(int(n), nnz, static_cast<const int*>(rowPtr),
                           static_cast<const int*>(colIdx), static_cast<const double*>(values),
                           static_cast<const double*>(x), static_cast<double*>(y), s);
}

MI_API int mi_jacobi_apply(const void* diag, const void* r, void* z, size_t n, void* stream)
{
    MI_TRY(ensure_device());
    if (n && (!diag || !r || !z))
        return set_error(MI_ERR_BAD_ARG, "null vector");
    if (!n)
        return MI_OK;
    k_jacobi<<<reduce_grid(int64_t(n), 4), kBlock, 0, as_stream(stream)>>>(
        int64_t(n), static_cast<const double*>(diag), static_cast<const double*>(r),
        static_cast<double*>(z));
    MI_HIP(hipGetLastError());
    return MI_OK;
}

} // extern "C"
