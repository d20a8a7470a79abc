// csr.hpp — the device-resident CSR matrix behind mi_csr.
#pragma once

#include "mi_common.hpp"
#include "stream.hpp"

#include <memory>
#include <vector>

// Sparsity pattern shared by matrices built on one mesh (momentum and
// pressure share it: src/solver/SimpleAlgorithm.cpp:78-82).
namespace mi
{
struct Pattern
{
    int64_t n = 0, nnz = 0;
    std::vector<int> h_rp, h_ci;
    DevBuf<int> rp, ci;
    DevBuf<int> diag; // position of a_ii in each row, -1 if absent
    bool all_diag = false; // every row stores its diagonal entry
    RowBlocks rb;     // CSR-stream row blocks
    SellForm sell;    // SELL-64 form when every row is short
    int lanes = 8;
};
} // namespace mi

struct mi_csr
{
    std::shared_ptr<mi::Pattern> pat;
    mi::DevBuf<double> val;
    mi::DevBuf<float> val32;  // FP32 copy for the inner iterations, built lazily
    mi::DevBuf<float> val32d; // FP32 difference form: diagonal slot = row sum
    bool val32d_valid = false;
    std::vector<double> h_val;
    bool host_valid = false;  // h_val mirrors val
    bool val32_valid = false;
    uint64_t version = 1;     // bumped on every value update

    // Values in the pattern's column order (RowBlocks::sort_columns), built on
    // demand by the stream views below.
    mi::DevBuf<double> s_val;
    mi::DevBuf<float> s_val32, s_op32;
    bool s_val_ok = false, s_val32_ok = false, s_op32_ok = false;
    // Values in the pattern's SELL-64 slot order (padding 0), on demand.
    mi::DevBuf<double> e_val;
    mi::DevBuf<float> e_val32, e_op32;
    bool e_val_ok = false, e_val32_ok = false, e_op32_ok = false;

    // Transposed copy (mi_csr_transpose): tmap[k] = position in the source
    // matrix of this matrix's k-th value.
    mi::DevBuf<int> tmap;
    const mi_csr* tsrc = nullptr;

    int64_t n() const { return pat->n; }
    // Values were replaced on the device (host_valid: h_val already holds
    // them): derived copies are stale and solvers rebuild on next use.
    void values_changed(bool host_valid_now)
    {
        host_valid = host_valid_now;
        val32_valid = false;
        val32d_valid = false;
        s_val_ok = s_val32_ok = s_op32_ok = false;
        e_val_ok = e_val32_ok = e_op32_ok = false;
        ++version;
    }
    int64_t nnz() const { return pat->nnz; }
    // Makes h_val current (device -> host copy when values were set on device).
    int sync_host();
    // Makes val32 current.
    int ensure_fp32(hipStream_t s);
    // Makes val32d current (requires every row to hold its diagonal).
    int ensure_fp32_diff(hipStream_t s);
    // The FP32 operator of the inner iterations: the difference form when
    // every row holds its diagonal, else the rounded values.  diag32() is
    // the matching StreamMat::diag (nullptr for the rounded values).
    int ensure_fp32_op(hipStream_t s);
    const float* op32() const { return pat->all_diag ? val32d.p : val32.p; }
    const int* diag32() const { return pat->all_diag ? pat->diag.p : nullptr; }
};

namespace mi
{
// Builds a pattern from host CSR arrays (validated) and uploads it.
int make_pattern(int64_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                 std::shared_ptr<Pattern>& out, hipStream_t s);
// Creates an mi_csr on an existing pattern with zero values.
int make_csr_on_pattern(const std::shared_ptr<Pattern>& pat, mi_csr** out);
// invD[i] = 1/a_ii, or 1 when the diagonal is missing or zero
// (src/solver/LinearSolvers.cpp:54-65).
int launch_inv_diag(const Pattern& p, const double* val, double* invD, hipStream_t s);
int launch_inv_diag32(const Pattern& p, const double* val, float* invD, hipStream_t s);
int launch_convert_f32(int64_t n, const double* src, float* dst, hipStream_t s);
#if defined(__HIPCC__)
// Stream views of a matrix for the CSR-stream kernels: column-ordered blocks
// with a permuted copy of the values when the pattern has them, else CSR.
// view32: op = the inner-iteration operator (op32 / diag32), else the
// rounded values.  sell: the SELL-64 form when the pattern has one (only for
// kernels that run stream_rows / sell_rows, launched with launch_grid).
int stream_view64(mi_csr* A, hipStream_t s, StreamMat<double>& out, bool sell = false);
int stream_view32(mi_csr* A, hipStream_t s, bool op, StreamMat<float>& out, bool sell = false);
#endif
} // namespace mi
