// amg_setup.hpp — smoothed-aggregation AMG hierarchy construction (host, FP64).
#pragma once

#include <cstdint>
#include <vector>

namespace mi
{

// Host CSR with nr rows and nc columns.
struct HCsr
{
    int nr = 0, nc = 0;
    std::vector<int> rp{0};
    std::vector<int> ci;
    std::vector<double> v;
    int64_t nnz() const { return int64_t(ci.size()); }
};

// Largest coarsest level inverted densely (128 MiB of FP64).
constexpr int kAmgMaxDense = 4096;

struct AmgParams
{
    int max_levels = 10;
    int coarse_size = 2048; // stop coarsening at this many rows (<= kAmgMaxDense)
    bool renumber = true;   // RCM-renumber every coarse level (amg_renumber_coarse)
    double theta = 0.02;        // strength: |a_ij| >= theta * sqrt(|a_ii a_jj|)
    double omega_scale = 4.0 / 3.0; // prolongator smoothing weight * lambda_max(D^-1 A)
    int threads = 0;            // 0: hardware concurrency (at most 16)
    bool filter = true;         // smooth P with the filtered operator (weak entries lumped)
    double truncation = 0.0;    // drop P entries below this x the row's largest |entry|
    // Invert the coarsest operator on the host (coarse_inverse); otherwise
    // only coarse_dense is filled and the caller inverts it (the device
    // hierarchy does, with the same elimination: dense_inverse_device).
    bool host_coarse_inverse = true;
};

struct AmgLevelHost
{
    HCsr A;          // level operator
    HCsr P, R;       // prolongator to this level from the next, R = P^T (empty on the coarsest)
    double lmax = 0; // estimate of lambda_max(D^-1 A)
    int n_aggregates = 0;
    // What a numeric re-setup on new values of the same pattern reuses (empty
    // in hierarchies read from a checkpoint): the aggregate of every row in
    // the next level's numbering, and per entry of A whether it is a strong
    // coupling kept by the filtered smoothing operator.
    std::vector<int> agg;
    std::vector<uint8_t> strong;
};

struct AmgHierarchyHost
{
    std::vector<AmgLevelHost> levels;
    std::vector<double> coarse_dense;   // the coarsest A, dense row-major
    std::vector<double> coarse_inverse; // its inverse (AmgParams::host_coarse_inverse)
    int coarse_n = 0;
    double operator_complexity = 0.0;
    double grid_complexity = 0.0;
};

// Builds the hierarchy from a square FP64 matrix.  Returns 0 on success.
int amg_build(const HCsr& A, const AmgParams& p, AmgHierarchyHost& out);

// Pieces, exposed for tests of the construction.
void amg_strength(const HCsr& A, double theta, std::vector<int>& srp, std::vector<int>& sci);
int amg_aggregate(int n, const std::vector<int>& srp, const std::vector<int>& sci,
                  const HCsr& A, std::vector<int>& agg);
void amg_smoothed_prolongator(const HCsr& A, const std::vector<int>& agg, int nagg, double omega,
                              HCsr& P, const std::vector<int>* srp, const std::vector<int>* sci,
                              double truncation = 0.0);
// A filtered diagonal within this fraction of |a_ii| of zero is zero: the
// row's couplings are all weak and cancel its diagonal (a zero row sum), and
// the smoothing leaves the row's tentative entry alone.
constexpr double kAmgFilterCancel = 1e-10;
void csr_transpose(const HCsr& A, HCsr& T);
void csr_multiply(const HCsr& A, const HCsr& B, HCsr& C, int threads);
double amg_lambda_max(const HCsr& A, int iters);
// The start vector of amg_lambda_max's power iteration (n entries).
void amg_lambda_start(int n, std::vector<double>& x);
int dense_inverse(int n, std::vector<double>& a, std::vector<double>& inv, int threads = 1);

} // namespace mi
