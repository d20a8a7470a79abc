// amg.hpp — device-side smoothed-aggregation hierarchy.
#pragma once

#include "amg_setup.hpp"
#include "solver.hpp"
#include "stream.hpp"

#include <vector>

namespace mi
{

// Chebyshev coefficients of a degree-ldeg smoother on [lmin_ratio lmax, lmax]
// of D^-1 A, lmax the setup's estimate (degree 1: damped Jacobi on
// [lmax / 2, lmax]): step k is d = beta[k] d + alpha[k] D^-1 r.
void chebyshev(double lmax_est, int ldeg, double lmin_ratio, std::vector<double>& alpha,
               std::vector<double>& beta);

template <class T> struct DMat
{
    int nr = 0, nc = 0;
    int64_t nnz = 0;
    int lanes = 8;
    DevBuf<int> rp, ci;
    DevBuf<T> v;    // in SELL-64 slot order when sell.built(), else in the column order of rb when rb.sorted()
    DevBuf<int> dg; // diagonal positions (CSR) when v holds the FP32 difference form
    RowBlocks rb;
    SellForm sell;  // short rows only (the finest operator, the prolongators)
    // diff_form: store the FP32 difference form (stream_rows_diff) when every
    // row holds its diagonal; ignored for FP64.
    int upload(const HCsr& h, hipStream_t s, bool diff_form = false);
    // New values on the uploaded pattern, FP64 in CSR order on the device:
    // the stored form is rebuilt from them as upload builds it from host
    // values.  tmp: nnz doubles of scratch (used by the FP32 difference form).
    int refresh(const double* csr_values, double* tmp, hipStream_t s);
#if defined(__HIPCC__)
    StreamMat<T> sm() const
    {
        StreamMat<T> m{nr,     rb.nrb, rb.blk.p, rp.p, ci.p, v.p, dg.n ? dg.p : nullptr,
                       rb.lpr, rb.sorted() ? rb.cw.p : nullptr, rb.hdr.p};
        if (sell.built()) {
            m.cw = nullptr;
            m.soff = sell.soff.p;
            m.eci = sell.eci.p;
            m.nsl = sell.nsl;
            if constexpr (std::is_same<T, float>::value)
                m.ecd = sell.ecd.n ? sell.ecd.p : nullptr;
        }
        return m;
    }
    // Workgroups of a launch over sm().
    unsigned grid() const { return sell.built() ? sell_grid(sell.nsl) : stream_grid(rb.nrb); }
#endif
};

template <class T> struct DLevel
{
    DMat<T> A, P, R;
    DevBuf<T> invD;
    DevBuf<T> x, b, r, d, d2, e; // x and b of level 0 are the caller's vectors
    std::vector<double> alpha, beta; // Chebyshev step coefficients
    int deg = 2;                     // smoother degree on this level
};

// Device values of one level for DHierarchy::refresh (FP64, CSR order of the
// level's host matrices).
struct LevelValues
{
    const double* A = nullptr;
    const double* P = nullptr; // nullptr on the coarsest level
    const double* R = nullptr;
    double lmax = 0.0;         // estimate of lambda_max(D^-1 A), x 1.05
};

template <class T> struct DHierarchy
{
    std::vector<DLevel<T>> L;
    DevBuf<double> cinv; // dense FP64 inverse of the coarsest operator
    int nc = 0;
    int deg = 2;
    double lmin_ratio = 0.3;
    GraphCache graphs;    // captured V-cycles
    GraphCache cg_graphs; // captured V-cycles fused with a CG iteration
    // degree: Chebyshev degree on level 0; coarse_degree on the coarser
    // levels (< 1: the same as level 0).
    // cinv_device: the coarsest inverse already on the device (n_c x n_c,
    // copied), or nullptr to upload H.coarse_inverse.
    int build(const AmgHierarchyHost& H, int degree, int coarse_degree, double lmin_ratio,
              hipStream_t s, const double* cinv_device = nullptr);
    // New values on the same patterns: every level's operators, inverse
    // diagonal and Chebyshev coefficients, and the coarsest inverse (copied
    // from cinv_device).  tmp: scratch of the largest level's nnz doubles.
    int refresh(const std::vector<LevelValues>& v, const double* cinv_device, double* tmp,
                hipStream_t s);
    // x0 = V-cycle(b0); x0 needs no initialisation.  With use_graph the
    // cycle's launches are replayed from a graph captured for (b0, x0, s).
    int vcycle(const T* b0, T* x0, hipStream_t s, bool use_graph = false);
    // x0 = V-cycle(f.r) inside a CG iteration (CgFuse): the first level-0
    // launch applies the CG update to f.r before the cycle reads it and
    // stores r.r partials (*grr workgroups), the last stores r.z partials
    // (*grz workgroups).  Needs two levels or more.
    int vcycle_cg(const CgFuse<T>& f, T* x0, hipStream_t s, bool use_graph, unsigned* grr,
                  unsigned* grz);
    // the launches of one V-cycle, issued through L (at once or into a graph);
    // cf: the fused CG iteration, or nullptr
    int cycle(const T* b0, T* x0, Launcher& L, const CgFuse<T>* cf = nullptr);
};

std::unique_ptr<Precond> make_amg(const mi_solver_cfg& c);

// Numeric re-setup on the device (amg_numeric.hip): the values of a hierarchy
// recomputed for new values of its finest operator, with the aggregates,
// strong-coupling masks and sparsity patterns of the host setup it was built
// by.  record() uploads what the recomputation needs (once per hierarchy);
// run() recomputes every level's values and leaves the coarsest operator
// dense in `dense`.
struct NumLevel;
struct AmgNumeric
{
    std::vector<NumLevel> L;
    DevBuf<double> scal;   // per level: lambda, 1/||y||, lambda_max (x 1.05)
    DevBuf<double> dense;  // coarsest operator, row-major
    std::vector<double> lmax; // lambda_max per level after run()
    AmgNumeric();
    ~AmgNumeric();
    // H: a hierarchy built by amg_build (with aggregates and strong masks);
    // rp0/ci0: the finest operator's pattern on the device (H's level 0).
    int record(const AmgHierarchyHost& H, const int* rp0, const int* ci0, hipStream_t s);
    // a0: the finest operator's new FP64 values (CSR order of H's level 0).
    // truncated: P's pattern came from a truncated full setup, whose rows
    // are scaled to the row sums of the complete smoothed rows
    int run(const double* a0, double omega_scale, bool truncated, hipStream_t s);
    // Values of level l for DHierarchy::refresh.
    LevelValues values(int l) const;
    // Copies the recomputed values into H (same patterns).
    int download(AmgHierarchyHost& H, hipStream_t s) const;
    size_t max_nnz() const;
};

// Inverse of a dense row-major n x n FP64 matrix on the device: blocked
// Gauss-Jordan with partial pivoting on [A | I] that performs, for every
// entry, the host dense_inverse's operations in the same order (identical
// results).  Panels of NB pivot columns: one workgroup factors the panel in
// registers (pivot search, row exchange, scaling, the panel's own
// eliminations, the factors f_r of every row); the rows the panel exchanged
// or scaled replay its steps per column (k_gj_rows); every other row takes
// its NB updates a_rk -= f_r u_k in one pass (k_gj_update).  Three launches
// per NB columns instead of two per column.  n <= kAmgMaxDense.
struct DenseInverse
{
    DevBuf<double> F, U, ipv, amax;
    DevBuf<int> info;
    // a (overwritten) holds A; inv receives A^-1.
    int run(int n, double* a, double* inv, hipStream_t s);
};

} // namespace mi
