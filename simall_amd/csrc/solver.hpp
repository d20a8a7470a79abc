// solver.hpp — mi_solver internals: preconditioner interface and workspace.
#pragma once

#include "csr.hpp"
#include "krylov.hpp"

#include <memory>
#include <vector>

namespace mi
{

// A preconditioner z = M^-1 r in FP64 and/or FP32.  inv_diag*() non-null
// means M is diagonal and the Krylov kernels fuse its application.
struct Precond
{
    virtual ~Precond() = default;
    virtual int setup(mi_csr* A, bool need64, bool need32, hipStream_t s) = 0;
    virtual const double* inv_diag64() const { return nullptr; }
    virtual const float* inv_diag32() const { return nullptr; }
    virtual int apply64(const double*, double*, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner has no FP64 apply");
    }
    virtual int apply32(const float*, float*, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner has no FP32 apply");
    }
    virtual const char* name() const = 0;
};

std::unique_ptr<Precond> make_jacobi();

} // namespace mi

struct mi_solver
{
    mi_solver_cfg cfg{};
    mi_csr* A = nullptr;
    uint64_t built_version = 0;
    bool built64 = false, built32 = false;
    std::unique_ptr<mi::Precond> pc;
    hipStream_t stream = nullptr;

    // FP64 vectors
    mi::DevBuf<double> b, x, r, z, p, Ap, hist, tmp;
    // FP32 vectors of the inner iterations
    mi::DevBuf<float> r32, z32, p32, Ap32, x32, b32;
    mi::DevBuf<mi::KState> st;
    mi::KState* h_st = nullptr; // pinned mirror of st
    mi::DevBuf<double> partials;
    mi::DevBuf<unsigned> counter;
    mi::DevBuf<double> dscal;   // device scalars for norms
    double* h_scal = nullptr;   // pinned
    std::vector<double> history;
    double last = 0.0;
    double setup_seconds = 0.0;

    ~mi_solver();
};
