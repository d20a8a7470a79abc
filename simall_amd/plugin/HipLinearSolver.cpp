// HipLinearSolver.cpp — ILinearSolver adapter over the C ABI.
#include "HipLinearSolver.hpp"

#include <cstring>

namespace simall::hip
{

bool HipLinearSolver::supported(const solver::LinearSolverConfig& cfg)
{
    // On device: CG, BiCGStab and GMRES(m) with Jacobi / None (Jacobi, as
    // make_preconditioner's default branch: src/solver/LinearSolvers.cpp:240-254),
    // ILU(0), SSOR or SA-AMG.  TFQMR stays on the host: its direction update
    // (LinearSolvers.cpp:592-593) diverges unless the preconditioner is exact,
    // so a device copy would only reproduce a divergent iteration.
    if (cfg.kind != solver::LinearSolverKind::CG && cfg.kind != solver::LinearSolverKind::BiCGSTAB &&
        cfg.kind != solver::LinearSolverKind::GMRES)
        return false;
    if (cfg.kind == solver::LinearSolverKind::GMRES && cfg.restart > 256)
        return false; // longest device restart cycle
    return cfg.preconditioner == solver::PreconditionerKind::Jacobi ||
           cfg.preconditioner == solver::PreconditionerKind::None ||
           cfg.preconditioner == solver::PreconditionerKind::ILU ||
           cfg.preconditioner == solver::PreconditionerKind::SSOR ||
           cfg.preconditioner == solver::PreconditionerKind::AMG;
}

std::unique_ptr<HipLinearSolver> HipLinearSolver::create(const solver::LinearSolverConfig& cfg,
                                                         int precision)
{
    if (!supported(cfg) || mi_init(0) != MI_OK)
        return nullptr;
    mi_solver_cfg c;
    mi_solver_cfg_default(&c);
    c.kind = static_cast<int>(cfg.kind);
    c.preconditioner = static_cast<int>(cfg.preconditioner);
    c.tolerance = cfg.tolerance;
    c.max_iterations = cfg.maxIterations;
    c.restart = cfg.restart;
    c.precision = precision;
    auto s = std::unique_ptr<HipLinearSolver>(new HipLinearSolver(c, cfg.maxIterations));
    if (!s->solver_)
        return nullptr;
    return s;
}

HipLinearSolver::HipLinearSolver(const mi_solver_cfg& c, int max_iterations)
    : cfg_(c), maxIterations_(max_iterations)
{
    if (mi_solver_create(&cfg_, &solver_) != MI_OK)
        solver_ = nullptr;
}

HipLinearSolver::~HipLinearSolver()
{
    if (solver_)
        mi_solver_destroy(solver_);
    if (A_)
        mi_csr_destroy(A_);
}

int HipLinearSolver::bind(const solver::CSRMatrix& A)
{
    const int64_t n = static_cast<int64_t>(A.rows());
    const int64_t nnz = static_cast<int64_t>(A.colIdx.size());
    const bool same = A_ && rowPtr_.size() == A.rowPtr.size() && colIdx_.size() == A.colIdx.size() &&
                      std::memcmp(rowPtr_.data(), A.rowPtr.data(), rowPtr_.size() * sizeof(int)) == 0 &&
                      std::memcmp(colIdx_.data(), A.colIdx.data(), colIdx_.size() * sizeof(int)) == 0;
    if (same)
        return mi_csr_update_values(A_, A.values.data());
    if (A_) {
        mi_csr_destroy(A_);
        A_ = nullptr;
    }
    rowPtr_.assign(A.rowPtr.begin(), A.rowPtr.end());
    colIdx_.assign(A.colIdx.begin(), A.colIdx.end());
    int rc = mi_csr_create(n, nnz, A.rowPtr.data(), A.colIdx.data(), A.values.data(), &A_);
    if (rc != MI_OK)
        return rc;
    return mi_solver_setup(solver_, A_);
}

int HipLinearSolver::solve(const solver::CSRMatrix& A, const util::aligned_vector<double>& b,
                           util::aligned_vector<double>& x)
{
    const std::size_t n = A.rows();
    x.assign(n, 0.0); // the CPU solvers start from zero (LinearSolvers.cpp:424)
    std::memset(&stats_, 0, sizeof(stats_));
    if (!solver_ || b.size() != n || bind(A) != MI_OK) {
        last_ = 1.0;
        return maxIterations_;
    }
    const int rc = mi_solver_solve(solver_, b.data(), x.data(), &stats_);
    last_ = mi_solver_last_residual(solver_);
    if (rc == MI_OK)
        return stats_.iterations;
    if (rc == MI_ERR_BREAKDOWN)
        return stats_.iterations; // CG p.Ap == 0 / BiCGStab rho == 0 return the current count
                                  // (LinearSolvers.cpp:375-378, 433-436)
    return maxIterations_;        // the iteration limit (GMRES counts up to it: :285-340)
}

} // namespace simall::hip
