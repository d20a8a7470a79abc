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
    RowBlocks rb;     // CSR-stream row blocks
    int lanes = 8;
};
} // namespace mi

struct mi_csr
{
    std::shared_ptr<mi::Pattern> pat;
    mi::DevBuf<double> val;
    mi::DevBuf<float> val32;  // FP32 copy for the inner iterations, built lazily
    std::vector<double> h_val;
    bool host_valid = false;  // h_val mirrors val
    bool val32_valid = false;
    uint64_t version = 1;     // bumped on every value update

    int64_t n() const { return pat->n; }
    int64_t nnz() const { return pat->nnz; }
    // Makes h_val current (device -> host copy when values were set on device).
    int sync_host();
    // Makes val32 current.
    int ensure_fp32(hipStream_t s);
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
} // namespace mi
