// spmv.hip — CSR SpMV kernels (vector-CSR with LANES lanes per row).
#include "spmv.hpp"

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
