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
struct AmgHierarchyHost;

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
    virtual int levels() const { return 1; }
    virtual double complexity() const { return 1.0; }
    // AMG only: the host hierarchy of the last setup, and building the device
    // hierarchies from a given one instead of from a matrix.
    virtual const AmgHierarchyHost* hierarchy() const { return nullptr; }
    virtual int install(const AmgHierarchyHost&, bool, bool, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner %s has no hierarchy", name());
    }
    // Whether new values of A can be taken by resetup (same pattern object as
    // the last full setup, the same precisions), and that numeric setup.
    virtual bool can_resetup(const mi_csr*, bool, bool) const { return false; }
    virtual int resetup(mi_csr*, bool, bool, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner %s has no numeric re-setup", name());
    }
    // AMG only: a host copy of the hierarchy applied now, with the device's
    // coarsest inverse.
    virtual int host_copy(AmgHierarchyHost&, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner %s has no hierarchy", name());
    }
    // CG fused with the apply (CgFuse): whether this preconditioner can run
    // the CG update and the r.z reduction inside its own launches for the
    // given precision, and that apply; *grr / *grz receive the workgroup
    // counts of the launches that wrote the r.r / r.z partials.
    virtual bool cg_fusable(bool /*fp64*/) const { return false; }
    virtual int apply_cg64(const CgFuse<double>&, double*, unsigned*, unsigned*, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner %s has no fused CG apply", name());
    }
    virtual int apply_cg32(const CgFuse<float>&, float*, unsigned*, unsigned*, hipStream_t)
    {
        return set_error(MI_ERR_UNSUPPORTED, "preconditioner %s has no fused CG apply", name());
    }
    bool use_graph = true; // replay applies from captured HIP graphs
};

std::unique_ptr<Precond> make_jacobi();

struct GmWork; // GMRES(m) workspace (gmres.hpp)

// Symmetric reordering of a pattern for the solver's working matrix.
struct Reordering
{
    std::shared_ptr<Pattern> source;  // pattern the permutation was computed for
    std::vector<int> perm;            // perm[new] = old
    std::shared_ptr<Pattern> pat;     // reordered pattern
    DevBuf<int> d_perm;
    DevBuf<int> d_vmap;               // reordered entry k comes from source entry vmap[k]
    std::vector<int> vmap;
    int bandwidth_before = 0, bandwidth_after = 0;
    int colors = 0; // multicolour ordering: colour count (0 for RCM)
};

// Makes the solver's next setup install H (an AMG hierarchy built for its
// working matrix's pattern, e.g. from a checkpoint) instead of building one;
// with keep_preconditioner the solves after it keep H.  Requires an AMG
// configuration with keep_preconditioner and a bound matrix.
int solver_adopt_amg(mi_solver* S, const AmgHierarchyHost& H);
// The AMG hierarchy the solver's preconditioner was last set up with, or
// nullptr.
const AmgHierarchyHost* solver_amg(const mi_solver* S);

} // namespace mi

struct mi_solver
{
    mi_solver_cfg cfg{};
    mi_csr* A = nullptr;                 // the caller's matrix
    mi_csr* M = nullptr;                 // working matrix: A, or Mown (reordered A)
    std::unique_ptr<mi_csr> Mown;
    std::unique_ptr<mi::Reordering> ro;
    uint64_t built_version = 0;
    std::unique_ptr<mi::Precond> pc;
    // A hierarchy to install at the next setup instead of building one
    // (solver_adopt_amg).
    std::unique_ptr<mi::AmgHierarchyHost> pending_amg;
    hipStream_t stream = nullptr;

    mi::DevBuf<double> b, x, hist; // host-API staging of b and x; residual history
    mi::DevBuf<double> bw, xw;     // b and x in the working ordering
    mi::DevBuf<double> w64[8];     // FP64 Krylov vectors
    mi::DevBuf<float> w32[10];     // FP32 Krylov vectors of the inner iterations
    mi::DevBuf<mi::KState> st;
    mi::KState* h_st = nullptr; // pinned mirror of st
    mi::DevBuf<double> partials;
    mi::DevBuf<double> dscal;   // device scalars for norms
    double* h_scal = nullptr;   // pinned
    std::unique_ptr<mi::GmWork> gm; // GMRES(m) basis and state, allocated on first use
    std::vector<double> history;
    double last = 0.0;
    double setup_seconds = 0.0;

    // Discrete adjoint: a solver with this configuration on the explicit
    // transpose of A, refreshed when A's values change (mi_adjoint_solve).
    std::unique_ptr<mi_csr> At;
    std::unique_ptr<mi_solver> adj;
    const mi_csr* at_src = nullptr;
    uint64_t at_version = 0;

    ~mi_solver();
};
