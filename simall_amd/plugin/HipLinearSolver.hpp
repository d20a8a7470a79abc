// HipLinearSolver.hpp — solver::ILinearSolver on the MI355X backend.
//
// Drop-in for the objects make_linear_solver returns
// (src/solver/Solver.cpp:25-39) on the configurations the device supports.
// Semantics follow CGSolver::solve (src/solver/LinearSolvers.cpp:420-452):
// x0 = 0, stop on ||r||/max(||b||,1e-30) < tolerance, return the iteration
// count (maxIterations when not converged), last_residual() the final ratio,
// never throw.
#pragma once

#include "simall_amd.h"
#include "simall_iface.hpp"

#include <memory>
#include <vector>

namespace simall::hip
{

class HipLinearSolver final : public solver::ILinearSolver
{
public:
    // Returns nullptr for configurations without a device path (the caller
    // keeps its own solver) or when no device is present.
    static std::unique_ptr<HipLinearSolver> create(const solver::LinearSolverConfig& cfg,
                                                   int precision = MI_PRECISION_FP64);
    static bool supported(const solver::LinearSolverConfig& cfg);

    ~HipLinearSolver() override;
    int solve(const solver::CSRMatrix& A, const util::aligned_vector<double>& b,
              util::aligned_vector<double>& x) override;
    double last_residual() const override { return last_; }

    const mi_stats& stats() const { return stats_; }

private:
    explicit HipLinearSolver(const mi_solver_cfg& c, int max_iterations);
    int bind(const solver::CSRMatrix& A);

    mi_solver_cfg cfg_{};
    int maxIterations_ = 1000;
    mi_solver* solver_ = nullptr;
    mi_csr* A_ = nullptr;
    std::vector<int> rowPtr_, colIdx_; // pattern of A_ (re-used while unchanged)
    double last_ = 0.0;
    mi_stats stats_{};
};

} // namespace simall::hip
