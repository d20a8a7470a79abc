// ilu0.hpp — ILU(0) factorization and level-scheduled device solves.
#pragma once

#include "solver.hpp"

#include <vector>

namespace mi
{

// In-place ILU(0) of CSR values (sorted columns) on the host; rows with a
// zero or missing pivot are skipped as in the reference.  zero_pivots counts
// zero pivots met (including zero final diagonals).
int ilu0_factor_host(int n, const int* rp, const int* ci, std::vector<double>& lu,
                     std::vector<int>& diag, int* zero_pivots);

// ILU(0) of the GPU facade (GpuKernels::ilu0_factor_inplace,
// src/gpu/GpuKernels.cpp:345-388): false, with nothing modified, when a row
// has no diagonal; false at the first zero pivot met, with the rows before it
// already factored; true otherwise.  Same elimination as ilu0_factor_host.
bool ilu0_factor_facade(int n, const int* rp, const int* ci, std::vector<double>& lu);

// Runs of consecutive rows with no dependency inside a run: for the L sweep,
// row i starts a new run when one of its lower entries lies in the current
// run; for the U sweep the same scanning from the last row.  lbeg / ubeg
// hold the run boundaries in sweep order (ubeg descending).
void ilu0_segments(int n, const int* rp, const int* ci, const std::vector<int>& lend,
                   const std::vector<int>& diag, std::vector<int>& lbeg, std::vector<int>& ubeg);

// Dependency levels of the L and U sweeps: rows of level l are
// rows[off[l] .. off[l+1]).  lend[i] = first position of row i with col >= i.
// With prefix_diag, a row whose (*prefix_diag)[i] < 0 (no stored diagonal)
// depends on every column > i of rows 0..i, as ILU0Preconditioner::apply
// sweeps such a row (src/solver/LinearSolvers.cpp:146-150).
void ilu0_levels(int n, const int* rp, const int* ci, const std::vector<int>& lend,
                 std::vector<int>& lrows, std::vector<int>& loff, std::vector<int>& urows,
                 std::vector<int>& uoff, const std::vector<int>* prefix_diag);

template <class T> struct Ilu0
{
    int n = 0;
    int zero_pivots = 0;
    const int* rp_dev = nullptr;
    const int* ci_dev = nullptr;
    DevBuf<T> d_lu, d_y;
    DevBuf<int> d_lrows, d_urows, d_lend, d_diag;
    std::vector<int> h_loff, h_uoff;
    bool level_built = false;
    const Pattern* levels_for = nullptr;

    // Segment sweeps: when the rows fall into a few long runs of consecutive
    // rows that do not depend on each other (a multicolour ordering: one run
    // per colour), each run is one CSR-stream launch over the split strictly
    // lower (L) or strictly upper (U) factor, with coalesced factor loads.
    bool seg = false;
    bool two = false; // two runs, the first without L and the second without U entries
    DevBuf<int> Lrp, Lci, Lblk, Urp, Uci, Ublk;
    DevBuf<T> Lval, Uval, Ud;
    std::vector<int> Lsb, Usb; // per run in sweep order: offset into L/Ublk, row blocks
    std::vector<int> Lmap, Umap; // position in the LU factor of each L / U entry
    std::vector<int> lbeg, ubeg; // run boundaries in sweep order
    std::vector<int> h_diag;
    DevBuf<int> d_Lmap, d_Umap;
    DevBuf<double> d_fac; // FP64 factor of the device factorization

    // Pattern analysis (levels, runs, split patterns), once per pattern.
    int prepare(const Pattern& P, hipStream_t s);
    // Host IKJ factorization of host values (every ordering).
    int setup(const Pattern& P, const std::vector<double>& vals, hipStream_t s);
    // Device factorization of device values, run by run (run orderings only:
    // seg after prepare); the same operations as the host factorization.
    int setup_device(const Pattern& P, const double* vals_device, hipStream_t s);
    // SSOR(1) on A's own values (device): no factorization.
    int setup_ssor(const Pattern& P, const double* vals_device, hipStream_t s);
    bool ssor = false;
    DevBuf<T> d_invD; // SSOR: 1 / a_ii (Ud holds a_ii)
    int apply(const T* r, T* z, hipStream_t s);
    // launches per sweep: runs (segment sweeps) or dependency levels
    int sweep_launches() const
    {
        if (seg)
            return (two && !ssor) ? 1 : int(Lsb.size() / 2);
        return h_loff.empty() ? 0 : int(h_loff.size()) - 1;
    }
};

std::unique_ptr<Precond> make_ilu0();
// SSOR with omega = 1 (make_preconditioner's SsorPreconditioner(1.0)).
std::unique_ptr<Precond> make_ssor();

} // namespace mi
