// krylov.hpp — device state and launchers of the fused Krylov kernels.
#pragma once

#include "blas1.hpp"
#include "spmv.hpp"

namespace mi
{

// Scalars of one Krylov solve, resident on the device.  Kernels read the
// step lengths from here and the last block of each reduction writes them, so
// an iteration needs no host round trip.  `done`: 0 running, 1 converged,
// 2 iteration limit, 3 breakdown.
struct KState
{
    double rz;     // r.z (CG) / rho (BiCGStab)
    double alpha;
    double beta;
    double omega;  // BiCGStab
    double bnorm;  // max(||b||, 1e-30)
    double tol;
    double last;   // last relative residual
    double scratch[4];
    int iter;
    int max_iter;
    int done;
    int pad;
};

// Row range of logical block `lb` when `G` blocks split `n` rows into
// contiguous ranges (rounded to `align` rows).
struct RowRange
{
    int beg, end;
};

#if defined(__HIPCC__)
// Workgroups are dispatched round-robin over the 8 XCDs of the chip; this maps
// hardware block b to a logical block so that each XCD walks one contiguous
// slice of rows and keeps its gathered x in its own L2.
__device__ __forceinline__ int xcd_logical_block(int b, int G)
{
    constexpr int kXcd = 8;
    if (G % kXcd != 0)
        return b;
    return (b % kXcd) * (G / kXcd) + b / kXcd;
}

__device__ __forceinline__ RowRange block_rows(int n, int G, int align)
{
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int64_t per = ((int64_t(n) + G - 1) / G + align - 1) / align * align;
    int64_t beg = int64_t(lb) * per, end = beg + per;
    if (beg > n)
        beg = n;
    if (end > n)
        end = n;
    return {int(beg), int(end)};
}
#endif

// out[0] = sum (or its square root) of G one-value block partials.
int launch_fin_sum(const double* partials, unsigned G, bool take_sqrt, double* out, hipStream_t s);

// Grid size for the row-range kernels: a multiple of 8 (one share per XCD),
// at most kMaxReduceBlocks, at least one block per `min_rows_per_block` rows.
unsigned range_grid(int64_t rows, int min_rows_per_block);

// Workspace of a CG / BiCGStab solve in element type T.
template <class T> struct KrylovVectors
{
    DevBuf<T> r, z, p, q, s, t, y, rh, v;
};

// --------------------------------------------------------------- CG launchers
// Stage tags let one set of kernels serve both precisions: T is the vector
// and matrix-value type (double, or float for the inner iterations of the
// mixed solve); every reduction accumulates in double.
template <class T> struct CgArgs
{
    int n;
    int64_t nnz;
    const int* rp;
    const int* ci;
    const T* val;
    const T* invD; // Jacobi: inverse diagonal; nullptr for an external preconditioner
    const T* b;
    T* x;
    T* r;
    T* z;
    T* p;
    T* Ap;
    KState* st;
    double* partials;
    double* history; // [max_iter], may be null
    int lanes;
    unsigned grid;   // range_grid(n, ...)
    const int* blk;  // CSR-stream row blocks of the matrix
    int nrb;
    const int* diag; // FP32 difference form: diagonal positions, else nullptr
    const unsigned* cw; // column-ordered row blocks (val in that order), else nullptr
    const int* hdr;     // row block boundaries {row, nonzero, column base, 0}, else nullptr
    const int* soff;    // SELL-64 form (val in its slot order), else nullptr
    const int* eci;
    const short* ecd; // FP32 SELL-64: 16-bit column offsets, or nullptr
    int nsl;
    // FP32 inner solves of the mixed form: when set, the iterate is
    // accumulated here in FP64 (xacc += alpha p) instead of in x, which is
    // then unused.  The caller's FP64 solution takes the correction directly
    // and the FP32 rounding of a stored correction never reaches it.
    double* xacc;
};

// r = b (x0 = 0) or r = b - A x (warm); z = M r (Jacobi) ; p = z; rz, ||b||.
// For an external preconditioner call cg_init_r, apply M, then cg_init_p.
template <class T> int cg_init_jacobi(const CgArgs<T>& a, bool warm, double tol, int max_iter,
                                       hipStream_t s);
template <class T> int cg_init_r(const CgArgs<T>& a, bool warm, hipStream_t s);
template <class T> int cg_init_p(const CgArgs<T>& a, double tol, int max_iter, hipStream_t s);
// One CG iteration (three to four launches).  With a Jacobi preconditioner
// the apply is fused; otherwise `apply` is invoked between the update and the
// r.z reduction.
template <class T> int cg_iteration_jacobi(const CgArgs<T>& a, hipStream_t s);
template <class T> int cg_iteration_part1(const CgArgs<T>& a, hipStream_t s);  // Ap, alpha, x, r, ||r||
template <class T> int cg_iteration_part2(const CgArgs<T>& a, hipStream_t s);  // rz, beta, p

// A CG iteration whose update and r.z reduction run inside the
// preconditioner's own launches (the AMG V-cycle): the preconditioner's
// first kernel applies x += alpha p, r -= alpha Ap and stores partials of
// r.r, its last stores partials of r.z (r = the updated residual, z = M r).
template <class T> struct CgFuse
{
    T* r;          // the CG residual (the apply's input), updated in place
    const T* p;
    const T* Ap;
    T* x;          // the iterate, or nullptr when xacc is set
    double* xacc;  // FP64 accumulation of the iterate (CgArgs::xacc)
    KState* st;    // alpha; done: the fused kernels do nothing
    double* prr;   // r.r partials, one per workgroup of the first kernel
    double* prz;   // r.z partials, one per workgroup of the last kernel
};
template <class T> int cg_iteration_spmv(const CgArgs<T>& a, hipStream_t s); // Ap, alpha
// After a fused apply: the stop test on r.r (grr partials), then beta and
// rz from r.z (grz partials), then p = z + beta p.
template <class T> int cg_iteration_fused_tail(const CgArgs<T>& a, const double* prr, unsigned grr,
                                               const double* prz, unsigned grz, hipStream_t s);

// --------------------------------------------------------- BiCGStab
// Right-preconditioned BiCGStab of BiCGSTABSolver::solve
// (src/solver/LinearSolvers.cpp:360-404).  KState: rz = rho_new (rh.r),
// scratch[0] = rho, alpha, omega, beta as in the reference.
template <class T> struct BiArgs
{
    int n;
    int64_t nnz;
    const int* rp;
    const int* ci;
    const T* val;
    const T* b;
    T *x, *r, *rh, *p, *v, *s, *t, *y, *z;
    KState* st;
    double* partials;
    double* history;
    int lanes;
    unsigned grid;
    const int* blk;
    int nrb;
    const int* diag;    // as in CgArgs
    const unsigned* cw; // as in CgArgs
    const int* hdr;     // as in CgArgs
    const int* soff;
    const int* eci;
    const short* ecd; // FP32 SELL-64: 16-bit column offsets, or nullptr
    int nsl;
    double* xacc;       // as in CgArgs
};

template <class T> int bicg_init(const BiArgs<T>& a, bool warm, double tol, int max_iter,
                                  hipStream_t s);
template <class T> int bicg_stage1(const BiArgs<T>& a, hipStream_t s); // p update
template <class T> int bicg_stage2(const BiArgs<T>& a, hipStream_t s); // v = A y, alpha, s, x += alpha y
template <class T> int bicg_stage3(const BiArgs<T>& a, hipStream_t s); // t = A z, omega, x, r, rho

} // namespace mi
