// csr.hip — mi_csr handles: upload, value updates, SpMV entry points.
#include "csr.hpp"
#include "krylov.hpp"
#include "spmv.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace mi;

namespace mi
{

__global__ void k_diag_pos(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                           int* __restrict__ diag)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    int d = -1;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i) {
            d = k;
            break;
        }
    diag[i] = d;
}

template <class T>
__global__ void k_inv_diag(int n, const int* __restrict__ diag, const double* __restrict__ val,
                           T* __restrict__ invD)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int d = diag[i];
    double v = 1.0;
    if (d >= 0 && val[d] != 0.0)
        v = 1.0 / val[d];
    invD[i] = T(v);
}

// Off-diagonals rounded to FP32; the diagonal slot receives the FP64 row sum.
__global__ void k_diff_f32(int n, const int* __restrict__ rp, const int* __restrict__ diag,
                           const double* __restrict__ val, float* __restrict__ out)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k) {
        d += val[k];
        out[k] = float(val[k]);
    }
    if (diag[i] >= 0)
        out[diag[i]] = float(d);
}

__global__ void k_convert_f32(int64_t n, const double* __restrict__ src, float* __restrict__ dst)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride)
        dst[i] = float(src[i]);
}

int make_pattern(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                 std::shared_ptr<Pattern>& out, hipStream_t s)
{
    if (n < 0 || nnz < 0 || n > INT32_MAX - 1 || nnz > INT32_MAX)
        return set_error(MI_ERR_BAD_ARG, "sizes out of int32 range (n=%lld nnz=%lld)",
                         (long long)n, (long long)nnz);
    if ((n > 0 && !rp) || (nnz > 0 && !ci))
        return set_error(MI_ERR_BAD_ARG, "null CSR array");
    if (n > 0 && (rp[0] != 0 || rp[n] != nnz))
        return set_error(MI_ERR_BAD_ARG, "rowPtr[0]=%d rowPtr[n]=%d, expected 0 and nnz=%lld",
                         rp[0], rp[n], (long long)nnz);
    for (int64_t i = 0; i < n; ++i) {
        if (rp[i + 1] < rp[i])
            return set_error(MI_ERR_BAD_ARG, "rowPtr decreases at row %lld", (long long)i);
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] < 0 || ci[k] >= n)
                return set_error(MI_ERR_BAD_ARG, "column %d out of range in row %lld", ci[k],
                                 (long long)i);
    }
    auto p = std::make_shared<Pattern>();
    p->n = n;
    p->nnz = nnz;
    if (n > 0) {
        p->h_rp.assign(rp, rp + n + 1);
        p->h_ci.assign(ci, ci + nnz);
        p->all_diag = true;
        for (int64_t i = 0; i < n && p->all_diag; ++i) {
            bool has = false;
            for (int k = rp[i]; k < rp[i + 1] && !has; ++k)
                has = ci[k] == i;
            p->all_diag = has;
        }
        MI_TRY(p->rp.upload(p->h_rp.data(), size_t(n + 1), s));
        MI_TRY(p->ci.upload(p->h_ci.data(), size_t(nnz), s));
        MI_TRY(p->diag.alloc(size_t(n)));
        k_diag_pos<<<grid_for(n, kBlock), kBlock, 0, s>>>(int(n), p->rp.p, p->ci.p, p->diag.p);
        MI_HIP(hipGetLastError());
    } else {
        p->h_rp.assign(1, 0);
    }
    p->lanes = lanes_for(n, nnz);
    if (n > 0) {
        MI_TRY(p->rb.upload(int(n), p->h_rp.data(), s));
        if (column_order_enabled())
            MI_TRY(p->rb.sort_columns(p->rp.p, p->ci.p, nnz, s));
        MI_TRY(p->sell.build(int(n), nnz, p->h_rp.data(), p->h_ci.data(), true, s));
    }
    out = std::move(p);
    return MI_OK;
}

template <class T>
static int permuted_copy(const Pattern& P, const T* src, DevBuf<T>& dst, bool& ok, hipStream_t s)
{
    if (ok)
        return MI_OK;
    MI_TRY(dst.alloc(size_t(P.nnz)));
    MI_TRY(launch_permute<T>(P.nnz, P.rb.sperm.p, src, dst.p, s));
    ok = true;
    return MI_OK;
}

template <class T>
static int sell_copy(const Pattern& P, const T* src, DevBuf<T>& dst, bool& ok, hipStream_t s)
{
    if (ok)
        return MI_OK;
    if (dst.n < size_t(P.sell.padded))
        MI_TRY(dst.alloc(size_t(P.sell.padded)));
    MI_TRY(launch_permute<T>(P.sell.padded, P.sell.eperm.p, src, dst.p, s));
    ok = true;
    return MI_OK;
}

template <class T> static void sell_fields(const Pattern& P, StreamMat<T>& out)
{
    out.soff = P.sell.soff.p;
    out.eci = P.sell.eci.p;
    out.nsl = P.sell.nsl;
    out.cw = nullptr;
    if constexpr (std::is_same<T, float>::value)
        out.ecd = P.sell.ecd.n ? P.sell.ecd.p : nullptr;
}

int stream_view64(mi_csr* A, hipStream_t s, StreamMat<double>& out, bool sell)
{
    const Pattern& P = *A->pat;
    out = StreamMat<double>{int(P.n), P.rb.nrb, P.rb.blk.p, P.rp.p, P.ci.p, A->val.p, nullptr,
                            P.rb.lpr};
    out.hdr = P.rb.hdr.p;
    if (sell && P.sell.built()) {
        MI_TRY(sell_copy(P, A->val.p, A->e_val, A->e_val_ok, s));
        out.val = A->e_val.p;
        sell_fields(P, out);
        return MI_OK;
    }
    if (P.rb.sorted()) {
        MI_TRY(permuted_copy(P, A->val.p, A->s_val, A->s_val_ok, s));
        out.val = A->s_val.p;
        out.cw = P.rb.cw.p;
    }
    return MI_OK;
}

int stream_view32(mi_csr* A, hipStream_t s, bool op, StreamMat<float>& out, bool sell)
{
    const Pattern& P = *A->pat;
    if (op)
        MI_TRY(A->ensure_fp32_op(s));
    else
        MI_TRY(A->ensure_fp32(s));
    const float* v = op ? A->op32() : A->val32.p;
    out = StreamMat<float>{int(P.n), P.rb.nrb, P.rb.blk.p, P.rp.p, P.ci.p, v,
                           op ? A->diag32() : nullptr, P.rb.lpr};
    out.hdr = P.rb.hdr.p;
    if (sell && P.sell.built()) {
        if (op)
            MI_TRY(sell_copy(P, v, A->e_op32, A->e_op32_ok, s));
        else
            MI_TRY(sell_copy(P, v, A->e_val32, A->e_val32_ok, s));
        out.val = op ? A->e_op32.p : A->e_val32.p;
        sell_fields(P, out);
        return MI_OK;
    }
    if (P.rb.sorted()) {
        if (op)
            MI_TRY(permuted_copy(P, v, A->s_op32, A->s_op32_ok, s));
        else
            MI_TRY(permuted_copy(P, v, A->s_val32, A->s_val32_ok, s));
        out.val = op ? A->s_op32.p : A->s_val32.p;
        out.cw = P.rb.cw.p;
    }
    return MI_OK;
}

int make_csr_on_pattern(const std::shared_ptr<Pattern>& pat, mi_csr** out)
{
    auto* A = new (std::nothrow) mi_csr();
    if (!A)
        return set_error(MI_ERR_DEVICE, "out of host memory");
    A->pat = pat;
    int rc = A->val.alloc(size_t(pat->nnz));
    if (rc == MI_OK && pat->nnz) {
        hipStream_t s = default_stream();
        rc = (hipMemsetAsync(A->val.p, 0, size_t(pat->nnz) * sizeof(double), s) == hipSuccess &&
              hipStreamSynchronize(s) == hipSuccess)
                 ? MI_OK
                 : set_error(MI_ERR_DEVICE, "hipMemsetAsync failed");
    }
    if (rc != MI_OK) {
        delete A;
        return rc;
    }
    A->h_val.assign(size_t(pat->nnz), 0.0);
    A->host_valid = true;
    *out = A;
    return MI_OK;
}

int launch_inv_diag(const Pattern& p, const double* val, double* invD, hipStream_t s)
{
    if (p.n == 0)
        return MI_OK;
    k_inv_diag<double><<<grid_for(p.n, kBlock), kBlock, 0, s>>>(int(p.n), p.diag.p, val, invD);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int launch_inv_diag32(const Pattern& p, const double* val, float* invD, hipStream_t s)
{
    if (p.n == 0)
        return MI_OK;
    k_inv_diag<float><<<grid_for(p.n, kBlock), kBlock, 0, s>>>(int(p.n), p.diag.p, val, invD);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int launch_convert_f32(int64_t n, const double* src, float* dst, hipStream_t s)
{
    if (n <= 0)
        return MI_OK;
    k_convert_f32<<<reduce_grid(n, 4), kBlock, 0, s>>>(n, src, dst);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// Copies of the streams a view reads (row pointers, block boundaries,
// columns or packed entries, values), for timing launches whose data is
// not in any cache.
template <class T> struct StreamCopy
{
    DevBuf<int> rp, blk, hdr, ci, soff;
    DevBuf<short> cd;
    DevBuf<T> val;
    StreamMat<T> V;
    // padded: slots of the SELL-64 form when A carries it
    int make(const StreamMat<T>& A, int64_t nnz, int64_t padded, hipStream_t s)
    {
        auto dup = [&](auto& dst, const auto* src, size_t count) -> int {
            if (!src || count == 0)
                return MI_OK;
            MI_TRY(dst.alloc(count));
            MI_HIP(hipMemcpyAsync(dst.p, src, count * sizeof(*src), hipMemcpyDeviceToDevice, s));
            return MI_OK;
        };
        V = A;
        if (A.soff) {
            MI_TRY(dup(soff, A.soff, size_t(A.nsl) + 1));
            MI_TRY(dup(ci, A.eci, size_t(padded)));
            MI_TRY(dup(cd, A.ecd, size_t(padded)));
            MI_TRY(dup(val, A.val, size_t(padded)));
            V.soff = soff.p;
            V.eci = ci.p;
            V.ecd = A.ecd ? cd.p : nullptr;
            V.val = val.p;
            return MI_OK;
        }
        MI_TRY(dup(rp, A.rp, size_t(A.n) + 1));
        MI_TRY(dup(blk, A.blk, size_t(A.nrb) + 1));
        MI_TRY(dup(hdr, A.hdr, 4 * (size_t(A.nrb) + 1)));
        MI_TRY(dup(ci, A.cw ? reinterpret_cast<const int*>(A.cw) : A.ci, size_t(nnz)));
        MI_TRY(dup(val, A.val, size_t(nnz)));
        V.rp = rp.p;
        V.blk = blk.p;
        V.hdr = A.hdr ? hdr.p : nullptr;
        V.ci = A.cw ? A.ci : ci.p;
        V.cw = A.cw ? reinterpret_cast<const unsigned*>(ci.p) : nullptr;
        V.val = val.p;
        return MI_OK;
    }
};

template <class T>
static int time_cold(const StreamMat<T>& A, int64_t nnz, int64_t padded, const T* x, T* y,
                     int reps, int64_t flush_bytes, hipStream_t s, double* avg_us)
{
    // enough copies to exceed flush_bytes, at most 64 (a matrix smaller than
    // flush_bytes / 64 is timed with 64 copies)
    const double per = double(nnz) * double(4 + sizeof(T)) + double(A.n + 1) * 4.0;
    const int nc = int(std::min(64.0, std::ceil(double(flush_bytes) / per) + 1.0));
    std::vector<StreamCopy<T>> cp(static_cast<size_t>(nc));
    for (auto& c : cp)
        MI_TRY(c.make(A, nnz, padded, s));
    for (auto& c : cp) // first touch of every copy
        MI_TRY(launch_spmv_view<T>(c.V, x, y, s));
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    float ms = 0.f;
    int rc = MI_OK;
    if (hipEventRecord(e0, s) != hipSuccess)
        rc = set_error(MI_ERR_DEVICE, "event record");
    for (int i = 0; i < reps && rc == MI_OK; ++i)
        rc = launch_spmv_view<T>(cp[size_t(i % nc)].V, x, y, s);
    if (rc == MI_OK &&
        (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = set_error(MI_ERR_DEVICE, "event timing");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    MI_TRY(rc);
    *avg_us = double(ms) * 1000.0 / reps;
    return MI_OK;
}

} // namespace mi

int mi_csr::sync_host()
{
    if (host_valid)
        return MI_OK;
    h_val.resize(size_t(nnz()));
    if (nnz()) {
        hipStream_t s = default_stream();
        MI_HIP(hipMemcpyAsync(h_val.data(), val.p, size_t(nnz()) * sizeof(double),
                              hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
    }
    host_valid = true;
    return MI_OK;
}

int mi_csr::ensure_fp32(hipStream_t s)
{
    if (val32_valid)
        return MI_OK;
    MI_TRY(val32.alloc(size_t(nnz())));
    MI_TRY(launch_convert_f32(nnz(), val.p, val32.p, s));
    val32_valid = true;
    return MI_OK;
}

int mi_csr::ensure_fp32_diff(hipStream_t s)
{
    if (val32d_valid)
        return MI_OK;
    MI_TRY(val32d.alloc(size_t(nnz())));
    if (n() > 0) {
        k_diff_f32<<<grid_for(n(), kBlock), kBlock, 0, s>>>(int(n()), pat->rp.p, pat->diag.p, val.p,
                                                            val32d.p);
        MI_HIP(hipGetLastError());
    }
    val32d_valid = true;
    return MI_OK;
}

__global__ void k_gather_vals(int64_t nnz, const int* __restrict__ map,
                              const double* __restrict__ src, double* __restrict__ dst)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t k = int64_t(blockIdx.x) * kBlock + threadIdx.x; k < nnz; k += stride)
        dst[k] = src[map[k]];
}

int mi_csr::ensure_fp32_op(hipStream_t s)
{
    return pat->all_diag ? ensure_fp32_diff(s) : ensure_fp32(s);
}

extern "C" {

MI_API int mi_csr_create(int64_t n, int64_t nnz, const int32_t* rowPtr, const int32_t* colIdx,
                         const double* values, mi_csr** out)
{
    clear_error();
    if (!out)
        return set_error(MI_ERR_BAD_ARG, "null output handle");
    *out = nullptr;
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    std::shared_ptr<Pattern> pat;
    MI_TRY(make_pattern(n, nnz, rowPtr, colIdx, pat, s));
    mi_csr* A = nullptr;
    MI_TRY(make_csr_on_pattern(pat, &A));
    if (nnz > 0 && values) {
        A->h_val.assign(values, values + nnz);
        int rc = A->val.upload(A->h_val.data(), size_t(nnz), s);
        if (rc != MI_OK) {
            delete A;
            return rc;
        }
    }
    if (hipStreamSynchronize(s) != hipSuccess) {
        delete A;
        return set_error(MI_ERR_DEVICE, "upload failed");
    }
    *out = A;
    return MI_OK;
}

MI_API int mi_csr_update_values(mi_csr* A, const double* values)
{
    if (!A || (A->nnz() && !values))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    A->h_val.assign(values, values + A->nnz());
    MI_TRY(A->val.upload(A->h_val.data(), size_t(A->nnz()), s));
    MI_HIP(hipStreamSynchronize(s));
    A->values_changed(true);
    return MI_OK;
}

MI_API int mi_csr_update_values_device(mi_csr* A, const double* values_device)
{
    if (!A || (A->nnz() && !values_device))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    if (A->nnz() && values_device != A->val.p)
        MI_HIP(hipMemcpyAsync(A->val.p, values_device, size_t(A->nnz()) * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    MI_HIP(hipStreamSynchronize(s));
    A->values_changed(false);
    return MI_OK;
}

MI_API int mi_csr_get_values(const mi_csr* A, double* values_host)
{
    if (!A || (A->nnz() && !values_host))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    MI_TRY(const_cast<mi_csr*>(A)->sync_host());
    std::memcpy(values_host, A->h_val.data(), size_t(A->nnz()) * sizeof(double));
    return MI_OK;
}

MI_API int mi_csr_get_pattern(const mi_csr* A, int32_t* rowPtr_host, int32_t* colIdx_host)
{
    if (!A)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (rowPtr_host)
        std::memcpy(rowPtr_host, A->pat->h_rp.data(), size_t(A->n() + 1) * sizeof(int32_t));
    if (colIdx_host && A->nnz())
        std::memcpy(colIdx_host, A->pat->h_ci.data(), size_t(A->nnz()) * sizeof(int32_t));
    return MI_OK;
}

MI_API int mi_csr_info(const mi_csr* A, int64_t* n, int64_t* nnz)
{
    if (!A)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (n)
        *n = A->n();
    if (nnz)
        *nnz = A->nnz();
    return MI_OK;
}

MI_API int mi_csr_values_device(mi_csr* A, double** values_device)
{
    if (!A || !values_device)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    *values_device = A->val.p;
    // The caller may write through this pointer: host and FP32 copies go stale.
    A->values_changed(false);
    return MI_OK;
}

MI_API int mi_csr_spmv(mi_csr* A, const double* x, double* y)
{
    if (!A || (A->n() && (!x || !y)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    const int64_t n = A->n();
    if (n == 0)
        return MI_OK;
    hipStream_t s = default_stream();
    DevBuf<double> dx, dy;
    MI_TRY(dx.upload(x, size_t(n), s));
    MI_TRY(dy.alloc(size_t(n)));
    StreamMat<double> V;
    MI_TRY(stream_view64(A, s, V, true));
    MI_TRY(launch_spmv_view<double>(V, dx.p, dy.p, s));
    MI_HIP(hipMemcpyAsync(y, dy.p, size_t(n) * sizeof(double), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

static int spmv_device_impl(mi_csr* A, int precision, const void* x, void* y, hipStream_t s)
{
    const int64_t n = A->n();
    if (n == 0)
        return MI_OK;
    if (precision == MI_PRECISION_FP64) {
        StreamMat<double> V;
        MI_TRY(stream_view64(A, s, V, true));
        return launch_spmv_view<double>(V, static_cast<const double*>(x), static_cast<double*>(y),
                                        s);
    }
    if (precision == MI_PRECISION_MIXED) {
        StreamMat<float> V;
        MI_TRY(stream_view32(A, s, false, V, true));
        return launch_spmv_view<float>(V, static_cast<const float*>(x), static_cast<float*>(y), s);
    }
    return set_error(MI_ERR_BAD_ARG, "unknown precision %d", precision);
}

MI_API int mi_csr_spmv_device(mi_csr* A, int precision, const void* x, void* y, void* stream)
{
    if (!A || (A->n() && (!x || !y)))
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : default_stream();
    MI_TRY(spmv_device_impl(A, precision, x, y, s));
    if (!stream)
        MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_csr_time_spmv(mi_csr* A, int precision, const void* x, void* y, int reps,
                            double* avg_us)
{
    if (!A || !avg_us || reps < 1 || (A->n() && (!x || !y)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    MI_TRY(spmv_device_impl(A, precision, x, y, s)); // warm + FP32 copy outside the timing
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    MI_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
        MI_TRY(spmv_device_impl(A, precision, x, y, s));
    MI_HIP(hipEventRecord(e1, s));
    MI_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    MI_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_us = double(ms) * 1000.0 / reps;
    return MI_OK;
}

MI_API int mi_csr_time_spmv_cold(mi_csr* A, int precision, const void* x, void* y, int reps,
                                 int64_t flush_bytes, double* avg_us)
{
    if (!A || !avg_us || reps < 1 || (A->n() && (!x || !y)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (flush_bytes < (int64_t(512) << 20))
        return set_error(MI_ERR_BAD_ARG, "flush of %lld bytes is below 512 MiB",
                         (long long)flush_bytes);
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    if (A->n() == 0) {
        *avg_us = 0.0;
        return MI_OK;
    }
    if (precision == MI_PRECISION_FP64) {
        StreamMat<double> V;
        MI_TRY(stream_view64(A, s, V, true));
        return time_cold<double>(V, A->nnz(), A->pat->sell.padded, static_cast<const double*>(x),
                                 static_cast<double*>(y), reps, flush_bytes, s, avg_us);
    }
    if (precision == MI_PRECISION_MIXED) {
        StreamMat<float> V;
        MI_TRY(stream_view32(A, s, false, V, true));
        return time_cold<float>(V, A->nnz(), A->pat->sell.padded, static_cast<const float*>(x),
                                static_cast<float*>(y),
                                reps, flush_bytes, s, avg_us);
    }
    return set_error(MI_ERR_BAD_ARG, "unknown precision %d", precision);
}

MI_API int mi_csr_transpose(const mi_csr* A, mi_csr** out)
{
    if (!A || !out)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    const int64_t n = A->n(), nnz = A->nnz();
    const std::vector<int>& rp = A->pat->h_rp;
    const std::vector<int>& ci = A->pat->h_ci;
    // counting sort by column: rows of A^T come out with ascending columns
    std::vector<int32_t> trp(static_cast<size_t>(n + 1), 0), tci(static_cast<size_t>(nnz));
    std::vector<int> map(static_cast<size_t>(nnz));
    for (int64_t k = 0; k < nnz; ++k)
        ++trp[size_t(ci[size_t(k)]) + 1];
    for (int64_t i = 0; i < n; ++i)
        trp[size_t(i) + 1] += trp[size_t(i)];
    std::vector<int32_t> next(trp.begin(), trp.end() - 1);
    for (int64_t i = 0; i < n; ++i)
        for (int k = rp[size_t(i)]; k < rp[size_t(i) + 1]; ++k) {
            const int p = next[size_t(ci[size_t(k)])]++;
            tci[size_t(p)] = int32_t(i);
            map[size_t(p)] = k;
        }
    std::shared_ptr<Pattern> pat;
    MI_TRY(make_pattern(n, nnz, trp.data(), tci.data(), pat, s));
    mi_csr* T = nullptr;
    MI_TRY(make_csr_on_pattern(pat, &T));
    std::unique_ptr<mi_csr> guard(T);
    if (nnz) {
        MI_TRY(T->tmap.upload(map.data(), map.size(), s));
        MI_HIP(hipStreamSynchronize(s));
    }
    T->tsrc = A;
    *out = guard.release();
    return mi_csr_transpose_values(A, *out);
}

MI_API int mi_csr_transpose_values(const mi_csr* A, mi_csr* At)
{
    if (!A || !At)
        return set_error(MI_ERR_BAD_ARG, "null argument");
    if (At->tsrc != A || At->nnz() != A->nnz())
        return set_error(MI_ERR_BAD_ARG, "matrix is not a transpose created from this matrix");
    MI_TRY(ensure_device());
    hipStream_t s = default_stream();
    if (A->nnz()) {
        k_gather_vals<<<reduce_grid(A->nnz(), 4), kBlock, 0, s>>>(A->nnz(), At->tmap.p, A->val.p,
                                                                  At->val.p);
        MI_HIP(hipGetLastError());
    }
    MI_HIP(hipStreamSynchronize(s));
    At->values_changed(false);
    return MI_OK;
}

MI_API int mi_csr_destroy(mi_csr* A)
{
    delete A;
    return MI_OK;
}

} // extern "C"
