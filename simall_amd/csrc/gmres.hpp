// gmres.hpp — device state and launchers of restarted GMRES(m).
//
// Left-preconditioned GMRES(m) of GMRESmSolver::solve
// (src/solver/LinearSolvers.cpp:266-340).  The Krylov basis lives in HBM as
// m+1 vectors of stride `ld`; the Hessenberg matrix, the Givens rotations and
// the least-squares right-hand side are small FP64 arrays on the device, and
// one-block kernels advance them, so an Arnoldi step needs no host round
// trip.  The host only polls GmState between groups of steps.
#pragma once

#include "krylov.hpp"
#include "stream.hpp"

namespace mi
{

// Longest restart cycle the device kernels accept (LDS of the dot kernel).
constexpr int kGmMaxRestart = 256;
// Rows per thread of the basis kernels: one block covers at most
// kGmRowsPerThread * kBlock rows, which fixes the smallest grid.
constexpr int kGmRowsPerThread = 4;

struct GmState
{
    KState k;      // iter, done (1 converged, 2 limit, 3 breakdown), last, tol, bnorm, max_iter
    double scale;  // 1 / norm of the newest basis vector (scaled by k_gm_scale)
    int j;         // Arnoldi steps completed in the current cycle
    int cycle_end; // the current cycle has ended (estimate met, breakdown, restart length)
    int steps;     // Arnoldi steps executed over all cycles (history index)
    int mixed;     // 1: the in-cycle estimate only ends the cycle; the FP64
                   //    residual test at the next cycle start decides convergence
};

template <class T> struct GmArgs
{
    int n;
    int m;             // restart length
    size_t ld;         // stride between basis vectors
    T* V;              // basis v_0 .. v_m
    T* w;              // A v_j, or the cycle's residual, before the preconditioner
    const T* invD;     // Jacobi preconditioner fused into the producing kernels, else nullptr
    StreamMat<T> A;    // operator of the Arnoldi steps
    StreamMat<double> A64; // FP64 operator of the cycle-start residual
    const double* b;
    double* x;         // FP64 iterate
    GmState* g;
    double* H;         // (m+1) x m Hessenberg, column j at H + j * (m + 1)
    double* cs;        // [m]
    double* sn;        // [m]
    double* sv;        // [m+1] least-squares right-hand side
    double* y;         // [m]
    double* h1;        // [m+1] first Gram-Schmidt coefficients
    double* h2;        // [m+1] second (re-orthogonalisation) coefficients
    double* partials;  // [(m+1) * G]
    double* history;   // [hist_cap] residual estimate per Arnoldi step, may be null
    int hist_cap;
    int G;             // grid of the basis kernels, gm_grid(n)
    unsigned Gs;       // grid of the stream SpMV kernels
};

// Workspace of a GMRES(m) solver: the basis in the precision of the Arnoldi
// steps, the small FP64 arrays, block partials and the polled state.
struct GmWork
{
    int m = 0;
    int64_t n = -1;
    size_t ld = 0;
    unsigned G = 0;
    DevBuf<double> V64, w64;
    DevBuf<float> V32, w32;
    DevBuf<double> small, partials;
    DevBuf<GmState> st;
    GmState* h_st = nullptr; // pinned mirror of st
    GmWork() = default;
    GmWork(const GmWork&) = delete;
    GmWork& operator=(const GmWork&) = delete;
    ~GmWork()
    {
        if (h_st)
            (void)hipHostFree(h_st);
    }
};

// Grid of the basis kernels: at least n / (kGmRowsPerThread * kBlock) blocks,
// a multiple of 8.
unsigned gm_grid(int64_t n);
// FP64 scratch of the small arrays for restart length m.
inline size_t gm_small_size(int m) { return size_t(m + 1) * size_t(m) + 3 * size_t(m) + 3 * size_t(m + 1); }

// Zeroes the state; bnorm_dev holds ||b||.
template <class T> int gm_init(const GmArgs<T>& a, const double* bnorm_dev, double tol,
                                int max_iter, bool mixed, hipStream_t s);
// Cycle start, part 1: w = b - A64 x (FP64, rounded to T); with Jacobi the
// preconditioned residual goes straight to v_0.
template <class T> int gm_cycle_residual(const GmArgs<T>& a, hipStream_t s);
// Cycle start, part 2 (after z = M^-1 w has been written to v_0): beta, the
// convergence and iteration-limit tests, v_0 /= beta.
template <class T> int gm_cycle_begin(const GmArgs<T>& a, hipStream_t s);
// Arnoldi step j, part 1: w = A v_j (with Jacobi: v_{j+1} = D^-1 A v_j).
template <class T> int gm_step_spmv(const GmArgs<T>& a, int j, hipStream_t s);
// Arnoldi step j, part 2 (after z = M^-1 w has been written to v_{j+1}):
// two Gram-Schmidt passes, the Hessenberg column, Givens rotations, the
// residual estimate and v_{j+1} /= h_{j+1,j}.
template <class T> int gm_step_orth(const GmArgs<T>& a, int j, hipStream_t s);
// Cycle end: solve H y = s and x += V y over the steps of the cycle.
template <class T> int gm_cycle_end(const GmArgs<T>& a, hipStream_t s);

} // namespace mi
