// amg.hpp — device-side smoothed-aggregation hierarchy.
#pragma once

#include "amg_setup.hpp"
#include "solver.hpp"
#include "stream.hpp"

#include <vector>

namespace mi
{

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

template <class T> struct DHierarchy
{
    std::vector<DLevel<T>> L;
    DevBuf<double> cinv; // dense FP64 inverse of the coarsest operator
    int nc = 0;
    int deg = 2;
    GraphCache graphs; // captured V-cycles
    // degree: Chebyshev degree on level 0; coarse_degree on the coarser
    // levels (< 1: the same as level 0).
    // cinv_device: the coarsest inverse already on the device (n_c x n_c,
    // copied), or nullptr to upload H.coarse_inverse.
    int build(const AmgHierarchyHost& H, int degree, int coarse_degree, double lmin_ratio,
              hipStream_t s, const double* cinv_device = nullptr);
    // x0 = V-cycle(b0); x0 needs no initialisation.  With use_graph the
    // cycle's launches are replayed from a graph captured for (b0, x0, s).
    int vcycle(const T* b0, T* x0, hipStream_t s, bool use_graph = false);
    // the launches of one V-cycle, issued through L (at once or into a graph)
    int cycle(const T* b0, T* x0, Launcher& L);
};

std::unique_ptr<Precond> make_amg(const mi_solver_cfg& c);

// Inverse of a dense row-major n x n matrix on the device: Gauss-Jordan with
// partial pivoting on [A | I], the operations of the host dense_inverse in
// the same order (identical results), one pivot column per two launches.
// a is overwritten; inv receives the inverse.
int dense_inverse_device(int n, double amax, double* a, double* inv, hipStream_t s);

} // namespace mi
