// batch.hip — multi-RHS solve of equations that share one CSR pattern.
//
// SimAll solves every scalar transport equation (energy, k/omega, species)
// with its own lin_.solve on an identical pattern
// (src/solver/ScalarTransport.cpp:84-242, the solve at :230-232), and the
// three momentum components one after another on one matrix
// (src/solver/SimpleAlgorithm.cpp:529-532).  Here the E equations of one
// pattern are solved together, each stage of the iteration one launch for all
// of them.
//
// Layout.  The Krylov working set is slot-interleaved: the equations being
// iterated occupy W slots (1 <= W <= 64), and every per-equation
// quantity of a nonzero k or a row i is a line of W consecutive values:
// v32[k * W + s], r[i * W + s], ...  A row of the batched SpMV is handled by
// W / V threads, each taking V consecutive slots (V = 4, 2 or 1, the widest
// dividing W under the autotuned cap) with V-wide loads: one load of
// colIdx[k] serves every equation, the values of nonzero k are one
// contiguous W-wide line (one value for all when the equations share one
// matrix), and the gather of x at column c is one contiguous line too.  The
// shared pattern is streamed once per launch and no gather is scattered
// across equations.  The SpMV and the ILU(0) sweeps run on an autotuned grid
// and row walk; the vector stages, norms and true residuals on one workgroup
// per 256 rows.  When the pattern has a SELL-64 form (short rows: 7-point
// operators) the nonzeros are stored in its slot order,
// m = soff[slice] + 64 k + row-in-slice, so the loads of colIdx and of the
// value lines of the rows of a workgroup step are contiguous for any W.
//
// As equations converge their slots go idle; at the next poll the live ones
// are repacked into a layout exactly as wide as their number (the other of
// two slot buffer sets), so launches stream the bytes of the equations still
// iterating and little more.  Between refinement sweeps everything per
// equation lives equation-major ("home"): FP64 and FP32 values
// v[e * nnz + k], b, x and the FP64 true residual [e * n + i].
//
// Arithmetic: PCG (CGSolver::solve, src/solver/LinearSolvers.cpp:420-452),
// BiCGStab (BiCGSTABSolver::solve, :360-404) or GMRES(m) (GMRESmSolver::solve,
// :266-340, one cycle per sweep) per equation, with Jacobi, ILU(0) (factorized
// on the device under a two-colour ordering), on a shared matrix the SA-AMG
// V-cycle of its one hierarchy, or no preconditioner; FP32 values (in the
// difference form when every row holds its diagonal) and vectors, the sweep's
// correction in FP32 or, for weakly dominant operators, FP64, and FP64 dot
// products and norms; the FP64 true residual with the FP64 values refines
// each equation until ||b_e - A_e x_e|| <= tol ||b_e|| (x_e accumulates each
// sweep's correction in FP64).  Each equation has its own scalars and
// convergence flag.  One inner
// CG iteration is five launches (eight with ILU(0)), one BiCGStab iteration
// nine (twelve); `chunk` iterations, or a whole GMRES cycle, are replayed
// from a HIP graph between polls.
#include "amg.hpp"
#include "csr.hpp"
#include "krylov.hpp"
#include "reorder.hpp"
#include "stream.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace mi
{

constexpr int kMaxEq = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int kBatchMaxGrid = 4096;
constexpr int kSpmmUnroll = 8; // nonzeros of a row loaded ahead of their gathers
// Operators whose mean diagonal-dominance margin (|a_ii| - sum_j!=i |a_ij|)
// / |a_ii| is below this sum the sweep's correction in FP64: rows that
// mostly sum to zero (a conservative diffusion operator without a time term,
// the pressure equation) make a nearly singular operator, on which FP32
// rounding of the correction costs the refinement sweeps iterations (the
// 128^3 pressure batch: 55 against 66); an operator with a time term keeps
// the FP32 correction, 8 bytes less per row and equation in each update
// (the 100^3 transport batch, mean margins 0.03 and up: 1270 against 1259
// iterations for its hardest equation, 7 % less time).
constexpr double kX64Margin = 1e-3;

struct BEq
{
    double rz, alpha, beta, r0, bnorm, tol_in, rel, last;
    double rho, omega; // BiCGStab: rho of the previous iteration (rz holds the current), omega
    double gscale, gbeta; // GMRES: 1 / norm of the newest basis vector; the cycle's start norm
    int iter, total, outer, done, fin, status, sconv;
    int gj, gend;         // GMRES: Arnoldi steps of the cycle; the cycle has ended
    int pad;
};

// Stages.  B_BI_*: BiCGStab; B_BI_T skips equations whose s already met the
// inner tolerance (sconv).
// With ILU(0) the preconditioner apply is two launches of its own
// (k_bilu), so the vector stages that would fold D^-1 into them come in
// variants without it: B_UPD_X, B_DOT_RZ, B_INIT_Z and B_PZ for CG,
// B_BI_PN and B_BI_SN for BiCGStab.
enum BOp : int {
    B_BNORM = 0, B_TRUE = 1, B_INIT = 2, B_ALPHA = 3, B_UPDATE = 4, B_P = 5,
    B_BI_INIT = 6, B_BI_P = 7, B_BI_ALPHA = 8, B_BI_S = 9, B_BI_T = 10, B_BI_OMEGA = 11,
    B_BI_UPDATE = 12, B_UPD_X = 13, B_DOT_RZ = 14, B_INIT_Z = 15, B_PZ = 16, B_BI_PN = 17,
    B_BI_SN = 18, B_GM = 19
};

// Slot-layout stages.
struct BArgs
{
    int n, W;
    int64_t nnz;
    const int* rp;
    const int* ci;
    const int* soff;   // SELL-64 slice offsets (sell layout)
    const int* eci;    // SELL-64 slot columns (diagonal flag, -1 padding)
    int shared;        // one matrix for every equation: v32[k], invD[i], lv[k]
    int band;          // row walk of the SpMM and ILU sweeps (row_walk)
    int diff;          // v32 is the FP32 difference form (FP64 row sums on the diagonal)
    const float* v32;  // [nnz or padded SELL slots][W] ([..] when shared)
    const float* invD; // [n][W] ([n] when shared): 1/a_ii, or the ILU(0) 1/pivot
    const float* lv;   // ILU(0) lower factor l_ik at the entries of the second colour's rows
    const int* eq;     // [W] equation of each slot, -1: empty
    float* xs;         // [n][W] correction of the current sweep
    double* xs64;      // the same in FP64 when not null (weakly dominant operators)
    float* r;          // BiCGStab: r, then s in place
    float* p;
    float* Ap;         // BiCGStab: v
    float* rh;         // BiCGStab shadow residual, y = D^-1 p, z = D^-1 s, t = A z
    float* y;
    float* z;
    float* t;
    BEq* st;           // nullptr: every occupied slot active (autotuner)
    double* partials;  // [2][W][G]
};

// Equation-major stages (norm of b, FP64 true residual).
struct HArgs
{
    int n, E, nrb;
    int64_t nnz;
    int64_t vstride; // distance between equations' FP64 values: nnz, or 0 when shared
    const int* blk; // pattern row blocks
    const int* rp;
    const int* ci;
    const double* v64; // [E][nnz]
    const double* b;   // [E][n]
    const double* x;   // [E][n]
    double* rt;        // [E][n]
    BEq* st;
    double* partials;  // [E][G]
};

__device__ __forceinline__ bool slot_on(const BArgs& a, int op, int e)
{
    if (e < 0)
        return false;
    if (a.st == nullptr)
        return true;
    const BEq& q = a.st[e];
    if (op == B_BI_T)
        return !q.done && !q.sconv;
    if (op == B_GM)
        return !q.done && !q.gend;
    return !q.done;
}

// Contiguous range [beg, end) of `count` items for this workgroup, the
// ranges of one XCD adjacent (xcd_logical_block), `align` dividing every
// boundary.
__device__ __forceinline__ void block_range(int64_t count, int64_t align, int64_t& beg,
                                            int64_t& end)
{
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int64_t per = ((count + G - 1) / G + align - 1) / align * align;
    beg = min(count, int64_t(lb) * per);
    end = min(count, beg + per);
}

// Rows [lo, lo + count) of a launch whose workgroups take passes of R rows.
// Contiguous: each workgroup one range (block_range).  Banded: the rows are
// cut in eight ranges, one per XCD (a workgroup runs on XCD blockIdx % 8),
// and the workgroups of an XCD take its passes in turn, so that together
// they sweep a band of G/8 passes that moves through the range: the rows a
// band gathers (its own +- the bandwidth) stay few enough to be reused from
// the XCD's L2, where contiguous ranges spread the XCD's workgroups over its
// whole range at once.  Walk: passes start at base, base + stride, ... < lim.
struct RowWalk
{
    int64_t base, lim, stride;
};
__device__ __forceinline__ RowWalk row_walk(int64_t lo, int64_t count, int R, int band)
{
    RowWalk w;
    const int G = gridDim.x;
    if (band && G % 8 == 0) {
        const int Gx = G / 8;
        const int x = int(blockIdx.x % 8), j = int(blockIdx.x / 8);
        const int64_t per = ((count + 7) / 8 + R - 1) / R * R;
        const int64_t xb = min(count, int64_t(x) * per), xe = min(count, xb + per);
        w.base = lo + xb + int64_t(j) * R;
        w.lim = lo + xe;
        w.stride = int64_t(Gx) * R;
    } else {
        int64_t r0, r1;
        block_range(count, R, r0, r1);
        w.base = lo + r0;
        w.lim = lo + r1;
        w.stride = R;
    }
    return w;
}

// ------------------------------------------------------------------ kernels
// V consecutive FP32 slots of one row as one load (V = 1, 2 or 4).
template <int V> struct SlotVec
{
    using type = float __attribute__((ext_vector_type(V)));
    static __device__ __forceinline__ float at(const type& x, int u) { return x[u]; }
    static __device__ __forceinline__ void set(type& x, int u, float v) { x[u] = v; }
};
template <> struct SlotVec<1>
{
    using type = float;
    static __device__ __forceinline__ float at(float x, int) { return x; }
    static __device__ __forceinline__ void set(float& x, int, float v) { x = v; }
};
// The same V slots in FP64 (the sweep's correction).
template <int V> struct SlotVec64
{
    using type = double __attribute__((ext_vector_type(V)));
    static __device__ __forceinline__ void set(type& x, int u, double v) { x[u] = v; }
    static __device__ __forceinline__ double at(const type& x, int u) { return x[u]; }
};
template <> struct SlotVec64<1>
{
    using type = double;
    static __device__ __forceinline__ void set(double& x, int, double v) { x = v; }
    static __device__ __forceinline__ double at(double x, int) { return x; }
};

// FP64 sums of up to two values per slot over a workgroup whose threads
// each hold V consecutive slots of one row group (TW = W / V threads per
// row, R = kBlock / TW row groups), in a fixed order (row groups 0 .. R-1
// of each slot); partials[v][s][block].
template <int V>
__device__ __forceinline__ void slot_partials_v(const BArgs& a, int R, int grp, int s0, bool live,
                                                int nv, const double (&d0)[V],
                                                const double (&d1)[V])
{
    __shared__ double lds[2][kBlock * V];
    if (live)
#pragma unroll
        for (int u = 0; u < V; ++u) {
            lds[0][grp * a.W + s0 + u] = d0[u];
            lds[1][grp * a.W + s0 + u] = d1[u];
        }
    __syncthreads();
    const int G = gridDim.x;
    for (int t = threadIdx.x; t < nv * a.W; t += kBlock) {
        const int v = t / a.W, sl = t % a.W;
        double sum = 0.0;
        for (int q = 0; q < R; ++q)
            sum += lds[v][q * a.W + sl];
        a.partials[(size_t(v) * a.W + sl) * G + blockIdx.x] = sum;
    }
}

// Batched SpMV out = A in over the slot layout, with FP64 partials per slot:
//   B_ALPHA     Ap = A p;  p.Ap                (CG)
//   B_BI_ALPHA  v = A y;   rh.v                (BiCGStab)
//   B_BI_T      t = A z;   t.s, t.t            (BiCGStab, s held in r)
// W / V threads per row, each taking V consecutive slots (equations) with
// V-wide loads of the values and of the gathered rows of `in`: kBlock V / W
// rows per workgroup at a time; every slot's row is summed in CSR order.
// kShared: one matrix for every equation, its value of a nonzero loaded once
// and applied to the thread's V slots.  With a.diff the FP32 values are the
// difference form of the single-equation solver's inner operator: a_ij
// (x_j - x_i) off the diagonal, the FP64 row sum times x_i on it (the same
// product in exact arithmetic, without the cancellation that rounding a_ii
// and a_ij apart leaves in A x for the nearly constant vectors of a pinned
// pressure operator); without it x_i is 0 and the sums are the plain ones.
template <int OP, bool kSell, int V, bool kShared>
__global__ __launch_bounds__(kBlock) void k_bspmm(BArgs a)
{
    using VT = typename SlotVec<V>::type;
    constexpr int U = V == 1 ? kSpmmUnroll : kSpmmUnroll / 2; // row entries per pass
    const int W = a.W;
    const int TW = W / V;          // threads per row
    const int R = kBlock / TW;     // rows per workgroup pass
    const int t = int(threadIdx.x);
    const bool live = t < R * TW;
    const int grp = t / TW;
    const int s0 = (t % TW) * V;   // first slot of this thread
    bool on[V];
    bool any = false, all = true;
#pragma unroll
    for (int u = 0; u < V; ++u) {
        on[u] = live && slot_on(a, OP, a.eq[s0 + u]);
        any = any || on[u];
        all = all && on[u];
    }
    const float* vin = (OP == B_ALPHA || OP == B_GM) ? a.p : OP == B_BI_ALPHA ? a.y : a.z;
    float* vout = OP == B_BI_T ? a.t : a.Ap;
    const float* vw = (OP == B_ALPHA || OP == B_GM) ? a.p : OP == B_BI_ALPHA ? a.rh : a.r;
    const RowWalk rw = row_walk(0, a.n, R, a.band);
    double d0[V], d1[V];
#pragma unroll
    for (int u = 0; u < V; ++u)
        d0[u] = d1[u] = 0.0;
    if (any) {
        for (int64_t pass = rw.base; pass < rw.lim; pass += rw.stride) {
            const int64_t i = pass + grp;
            if (i >= rw.lim)
                break;
            const size_t q = size_t(i) * W + s0;
            // the dot product's operand, loaded ahead of the row's streams
            const VT wq = *reinterpret_cast<const VT*>(vw + q);
            const VT xi = a.diff ? *reinterpret_cast<const VT*>(vin + q) : VT(0.f);
            VT acc = VT(0.f);
            if constexpr (kSell) {
                // slot k of row i: m0 + 64 k, in CSR order
                const int slc = int(i / kSellSlice);
                const int m0 = a.soff[slc] + int(i % kSellSlice);
                const int w = (a.soff[slc + 1] - a.soff[slc]) / kSellSlice;
                for (int k0 = 0; k0 < w; k0 += U) {
                    int c[U];
                    VT v[U];
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        const size_t mk = size_t(m0) + size_t(k0 + j) * kSellSlice;
                        const bool in = k0 + j < w;
                        c[j] = in ? a.eci[mk] : -1;
                        if constexpr (kShared)
                            v[j] = in ? VT(a.v32[mk]) : VT(0.f);
                        else
                            v[j] = in ? __builtin_nontemporal_load(
                                            reinterpret_cast<const VT*>(a.v32 + mk * W + s0))
                                      : VT(0.f);
                    }
                    VT xv[U];
#pragma unroll
                    for (int j = 0; j < U; ++j)
                        xv[j] = c[j] >= 0 ? *reinterpret_cast<const VT*>(
                                                vin + size_t(c[j] & kSellColMask) * W + s0)
                                          : VT(0.f);
#pragma unroll
                    for (int j = 0; j < U; ++j)
                        if (c[j] >= 0)
                            acc += v[j] * ((c[j] & kSellDiag) ? xv[j] : xv[j] - xi);
                }
            } else {
                const int kb = a.rp[i], ke = a.rp[i + 1];
                for (int k0 = kb; k0 < ke; k0 += U) {
                    int c[U];
                    VT v[U];
#pragma unroll
                    for (int j = 0; j < U; ++j) {
                        const int k = k0 + j;
                        c[j] = k < ke ? a.ci[k] : -1;
                        if constexpr (kShared)
                            v[j] = k < ke ? VT(a.v32[k]) : VT(0.f);
                        else
                            v[j] = k < ke ? __builtin_nontemporal_load(reinterpret_cast<const VT*>(
                                                a.v32 + size_t(k) * W + s0))
                                          : VT(0.f);
                    }
                    VT xv[U];
#pragma unroll
                    for (int j = 0; j < U; ++j)
                        xv[j] = c[j] >= 0 ? *reinterpret_cast<const VT*>(vin + size_t(c[j]) * W + s0)
                                          : VT(0.f);
#pragma unroll
                    for (int j = 0; j < U; ++j)
                        if (c[j] >= 0)
                            acc += v[j] * (c[j] == int(i) ? xv[j] : xv[j] - xi);
                }
            }
            if (all) {
                *reinterpret_cast<VT*>(vout + q) = acc;
            } else {
#pragma unroll
                for (int u = 0; u < V; ++u)
                    if (on[u])
                        vout[q + u] = SlotVec<V>::at(acc, u);
            }
#pragma unroll
            for (int u = 0; u < V; ++u) {
                if (!on[u])
                    continue;
                const double au = double(SlotVec<V>::at(acc, u));
                d0[u] += double(SlotVec<V>::at(wq, u)) * au;
                if (OP == B_BI_T)
                    d1[u] += au * au;
            }
        }
    }
    slot_partials_v<V>(a, R, grp, s0, live, OP == B_BI_T ? 2 : 1, d0, d1);
}

// The SpMV kernel for the slot width W of a launch: V = 4, 2 or 1 slots per
// thread (slots_per_thread; the cap is autotuned per batch).
template <int OP> using BSpmmKernel = void (*)(BArgs);
// Slots per thread of a launch of width W: the widest of 4, 2, 1 that divides
// W and does not exceed the autotuned cap vmax.
inline int slots_per_thread(int W, int vmax)
{
    if (W % 4 == 0 && vmax >= 4)
        return 4;
    if (W % 2 == 0 && vmax >= 2)
        return 2;
    return 1;
}
template <int OP, bool kShared> BSpmmKernel<OP> bspmm_kernel_v(bool sell, int V)
{
    if (V == 4)
        return sell ? k_bspmm<OP, true, 4, kShared> : k_bspmm<OP, false, 4, kShared>;
    if (V == 2)
        return sell ? k_bspmm<OP, true, 2, kShared> : k_bspmm<OP, false, 2, kShared>;
    return sell ? k_bspmm<OP, true, 1, kShared> : k_bspmm<OP, false, 1, kShared>;
}
template <int OP> BSpmmKernel<OP> bspmm_kernel(bool sell, bool shared, int W, int vmax)
{
    const int V = slots_per_thread(W, vmax);
    return shared ? bspmm_kernel_v<OP, true>(sell, V) : bspmm_kernel_v<OP, false>(sell, V);
}

// Vector stages over the slot layout (element q = i W + s; thread t of a
// workgroup takes slot t % W of rows t / W, t / W + P, ...).  D^-1 is
// invD[q], or invD[i] for a shared matrix:
//   B_INIT      p = D^-1 r; r.z, r.r
//   B_UPDATE    xs += alpha p, r -= alpha Ap; r.r, r.z
//   B_P         p = D^-1 r + beta p
//   B_BI_INIT   rh = r; p = v = 0; r.r, rh.r
//   B_BI_P      p = r + beta (p - omega v); y = D^-1 p
//   B_BI_S      s = r - alpha v (in r); z = D^-1 s; s.s
//   B_BI_UPDATE xs += alpha y + omega z; r = s - omega t; r.r, rh.r
// and with the preconditioner applied by its own launches (z = M^-1 r in z):
//   B_UPD_X     xs += alpha p, r -= alpha Ap
//   B_DOT_RZ    r.r, r.z
//   B_INIT_Z    p = z; r.z, r.r
//   B_PZ        p = z + beta p
//   B_BI_PN     p = r + beta (p - omega v)
//   B_BI_SN     s = r - alpha v (in r); s.s
// Threads take V consecutive slots of a row (V-wide loads and stores, like
// the SpMM): W / V threads per row, R = kBlock V / W rows per pass; a slot of
// an equation not iterating keeps its values.  With a.xs64 the sweep's
// correction is summed in FP64 with the step lengths in FP64 (the products
// of FP32 directions exact, as the single-equation solver's FP64 iterate):
// rounding it to FP32 leaves a residual of order kappa * 6e-8, which on a
// weakly dominant (ill-conditioned) operator costs the refinement sweeps
// iterations; a strongly dominant one keeps the FP32 correction (a.xs).
template <int V>
__global__ __launch_bounds__(kBlock) void k_bvec(BArgs a, int op)
{
    using VT = typename SlotVec<V>::type;
    using DT = typename SlotVec64<V>::type;
    const int W = a.W;
    const int TW = W / V;
    const int R = kBlock / TW;
    const int t = int(threadIdx.x);
    const bool live = t < R * TW;
    const int grp = t / TW;
    const int s0 = (t % TW) * V;
    bool on[V], keep[V];
    bool any = false, all = true, allk = true;
    VT al = VT(0.f), be = VT(0.f), om = VT(0.f);
    double al64[V], om64[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const int e = live ? a.eq[s0 + u] : -1;
        on[u] = slot_on(a, op, e);
        keep[u] = on[u];
        al64[u] = om64[u] = 0.0;
        if (on[u]) {
            const BEq& st = a.st[e];
            SlotVec<V>::set(al, u, float(st.alpha));
            SlotVec<V>::set(be, u, float(st.beta));
            SlotVec<V>::set(om, u, float(st.omega));
            al64[u] = st.alpha;
            om64[u] = st.omega;
            if (op == B_BI_UPDATE && st.sconv)
                keep[u] = false; // s met the tolerance: only x moves (:387-390)
        }
        any = any || on[u];
        all = all && on[u];
        allk = allk && keep[u];
    }
    int64_t r0, r1;
    block_range(a.n, R, r0, r1);
    double d0[V], d1[V];
#pragma unroll
    for (int u = 0; u < V; ++u)
        d0[u] = d1[u] = 0.0;
    const bool sh = a.shared != 0;
    if (any) {
        for (int64_t i = r0 + grp; i < r1; i += R) {
            const size_t q = size_t(i) * W + s0;
            auto ld = [&](const float* p) { return *reinterpret_cast<const VT*>(p + q); };
            auto put = [&](float* p, const VT& v, const bool (&m)[V], bool every) {
                if (every) {
                    *reinterpret_cast<VT*>(p + q) = v;
                } else {
#pragma unroll
                    for (int u = 0; u < V; ++u)
                        if (m[u])
                            p[q + u] = SlotVec<V>::at(v, u);
                }
            };
            auto dinv = [&]() { return sh ? VT(a.invD[i]) : ld(a.invD); };
            // xs64 += dx for the slots in m
            auto xs_add = [&](const DT& dx, const bool (&m)[V], bool every) {
                double* px = a.xs64 + q;
                if (every) {
                    *reinterpret_cast<DT*>(px) = *reinterpret_cast<const DT*>(px) + dx;
                } else {
#pragma unroll
                    for (int u = 0; u < V; ++u)
                        if (m[u])
                            px[u] += SlotVec64<V>::at(dx, u);
                }
            };
            // alpha p in FP64
            auto step64 = [&](const VT& pv) {
                DT dx;
#pragma unroll
                for (int u = 0; u < V; ++u)
                    SlotVec64<V>::set(dx, u, al64[u] * double(SlotVec<V>::at(pv, u)));
                return dx;
            };
            auto acc = [&](double (&d)[V], const VT& x, const VT& y, const bool (&m)[V]) {
#pragma unroll
                for (int u = 0; u < V; ++u)
                    if (m[u])
                        d[u] += double(SlotVec<V>::at(x, u)) * double(SlotVec<V>::at(y, u));
            };
            if (op == B_INIT) {
                const VT r = ld(a.r), z = dinv() * r;
                put(a.p, z, on, all);
                acc(d0, r, z, on);
                acc(d1, r, r, on);
            } else if (op == B_UPDATE) {
                if (a.xs64)
                    xs_add(step64(ld(a.p)), on, all);
                else
                    put(a.xs, ld(a.xs) + al * ld(a.p), on, all);
                const VT r = ld(a.r) - al * ld(a.Ap);
                put(a.r, r, on, all);
                const VT z = dinv() * r;
                acc(d0, r, r, on);
                acc(d1, r, z, on);
            } else if (op == B_P) {
                put(a.p, dinv() * ld(a.r) + be * ld(a.p), on, all);
            } else if (op == B_UPD_X) {
                if (a.xs64)
                    xs_add(step64(ld(a.p)), on, all);
                else
                    put(a.xs, ld(a.xs) + al * ld(a.p), on, all);
                put(a.r, ld(a.r) - al * ld(a.Ap), on, all);
            } else if (op == B_DOT_RZ) {
                const VT r = ld(a.r), z = ld(a.z);
                acc(d0, r, r, on);
                acc(d1, r, z, on);
            } else if (op == B_INIT_Z) {
                const VT r = ld(a.r), z = ld(a.z);
                put(a.p, z, on, all);
                acc(d0, r, z, on);
                acc(d1, r, r, on);
            } else if (op == B_PZ) {
                put(a.p, ld(a.z) + be * ld(a.p), on, all);
            } else if (op == B_BI_INIT) {
                // rh = r; p = v = 0 (BiCGSTABSolver::solve, :366-372)
                const VT r = ld(a.r);
                put(a.rh, r, on, all);
                put(a.p, VT(0.f), on, all);
                put(a.Ap, VT(0.f), on, all);
                acc(d0, r, r, on);
                acc(d1, r, r, on);
            } else if (op == B_BI_P || op == B_BI_PN) {
                // (:380-382)
                const VT pv = ld(a.r) + be * (ld(a.p) - om * ld(a.Ap));
                put(a.p, pv, on, all);
                if (op == B_BI_P)
                    put(a.y, dinv() * pv, on, all);
            } else if (op == B_BI_S || op == B_BI_SN) {
                // (:384-386)
                const VT sv = ld(a.r) - al * ld(a.Ap);
                put(a.r, sv, on, all);
                if (op == B_BI_S)
                    put(a.z, dinv() * sv, on, all);
                acc(d0, sv, sv, on);
            } else { // B_BI_UPDATE (:387-397)
                // slots whose s met the tolerance take x += alpha y alone
                const VT y = ld(a.y), z = ld(a.z);
                if (a.xs64) {
                    DT dx = step64(y);
#pragma unroll
                    for (int u = 0; u < V; ++u)
                        if (keep[u])
                            SlotVec64<V>::set(dx, u, SlotVec64<V>::at(dx, u) +
                                                         om64[u] * double(SlotVec<V>::at(z, u)));
                    xs_add(dx, on, all);
                } else {
                    VT xs = ld(a.xs) + al * y;
#pragma unroll
                    for (int u = 0; u < V; ++u)
                        if (keep[u])
                            SlotVec<V>::set(xs, u, SlotVec<V>::at(xs, u) +
                                                       SlotVec<V>::at(om, u) * SlotVec<V>::at(z, u));
                    put(a.xs, xs, on, all);
                }
                bool anyk = false;
#pragma unroll
                for (int u = 0; u < V; ++u)
                    anyk = anyk || keep[u];
                if (anyk) {
                    const VT r = ld(a.r) - om * ld(a.t);
                    put(a.r, r, keep, allk);
                    acc(d0, r, r, keep);
                    acc(d1, ld(a.rh), r, keep);
                }
            }
        }
    }
    if (op == B_P || op == B_BI_P || op == B_UPD_X || op == B_PZ || op == B_BI_PN)
        return;
    slot_partials_v<V>(a, R, grp, s0, live, (op == B_BI_S || op == B_BI_SN) ? 1 : 2, d0, d1);
}

// One slot per thread: on the 32-equation 1M-row batch V = 4 ran the CG
// update in the same time (188.8 vs 187.6 us at full width) and the whole
// solve 3-4 % slower (519-525 vs 506 ms).
using BVecKernel = void (*)(BArgs, int);
BVecKernel bvec_kernel(int) { return k_bvec<1>; }

// ---------------------------------------------------------------- GMRES(m)
// Left-preconditioned GMRES(m) for every equation of the slot layout: one
// refinement sweep is one cycle, the Arnoldi steps of GMRESmSolver::solve
// (src/solver/LinearSolvers.cpp:266-340) with classical Gram-Schmidt applied
// twice (the single-equation solver's orthogonalisation), per-equation
// Hessenberg matrix, Givens rotations and least-squares right-hand side in
// FP64 on the device.  The basis is m+1 slot-layout FP32 vectors.
constexpr int kBGmMax = 32; // longest batched cycle (registers of the dot kernel)

struct BGm
{
    float* V;      // basis: vector k at V + k * vs
    size_t vs;     // n * W
    int m;
    double* small; // per equation, `ss` doubles: H, cs, sn, sv, y, h1, h2
    int ss;
    double* part;  // [W][m + 1][G] partials of the basis kernels
};
__device__ __forceinline__ double* gm_H(const BGm& g, int e) { return g.small + size_t(e) * g.ss; }
__device__ __forceinline__ double* gm_cs(const BGm& g, int e) { return gm_H(g, e) + (g.m + 1) * g.m; }
__device__ __forceinline__ double* gm_sn(const BGm& g, int e) { return gm_cs(g, e) + g.m; }
__device__ __forceinline__ double* gm_sv(const BGm& g, int e) { return gm_sn(g, e) + g.m; }
__device__ __forceinline__ double* gm_y(const BGm& g, int e) { return gm_sv(g, e) + g.m + 1; }
__device__ __forceinline__ double* gm_h(const BGm& g, int e, int which)
{
    return gm_y(g, e) + g.m + (which == 2 ? g.m + 1 : 0);
}
inline int gm_small_stride(int m) { return (m + 1) * m + 3 * m + 3 * (m + 1); }

__device__ __forceinline__ bool gm_on(const BArgs& a, int e)
{
    return e >= 0 && !a.st[e].done && !a.st[e].gend;
}

// The basis kernels hold rows of z in registers and take the basis vectors
// in groups of four (the single-equation GMRES kernels' arrangement): a
// thread owns slot t % W of rows rg, rg + P, ..., rg + (kGmRows - 1) P of a
// pass (rg = t / W, P = kBlock / W row groups), a workgroup walks the passes
// of its contiguous row range, and every group of four dot products is
// reduced per slot through LDS and stored as partials
// g.part[(slot (m+1) + k) G + block].
constexpr int kGmRows = 4;

// Stores the per-slot workgroup sums of four values per thread (k0 .. k0+3,
// those below nv).
__device__ __forceinline__ void gm_store4(const BArgs& a, const BGm& g, int P, int k0, int nv,
                                          const double (&acc)[4])
{
    __shared__ double lds[4][kBlock];
    const int W = a.W;
    const int G = gridDim.x;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        lds[u][threadIdx.x] = acc[u];
    __syncthreads();
    const int nk = min(4, nv - k0);
    for (int t = threadIdx.x; t < nk * W; t += kBlock) {
        const int u = t / W, sl = t % W;
        double sum = 0.0;
        for (int q = 0; q < P; ++q)
            sum += lds[u][q * W + sl];
        g.part[(size_t(sl) * (g.m + 1) + k0 + u) * G + blockIdx.x] = sum;
    }
    __syncthreads();
}

// Partials of v_k . z for k < nv over this workgroup's rows, z read as
// stored (one FP32 vector per call).
__device__ __forceinline__ void gm_dots4(const BArgs& a, const BGm& g, int P, int rg, int sl,
                                         bool on, int64_t i0, int64_t i1, int nv, const float* z)
{
    const int W = a.W;
    for (int k0 = 0; k0 < nv; k0 += 4) {
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        if (on) {
            for (int64_t base = i0; base < i1; base += int64_t(kGmRows) * P) {
                double zr[kGmRows];
                size_t q[kGmRows];
#pragma unroll
                for (int r = 0; r < kGmRows; ++r) {
                    const int64_t i = base + rg + int64_t(r) * P;
                    const bool in = i < i1;
                    q[r] = size_t(in ? i : i0) * W + sl; // clamped, dropped below
                    zr[r] = in ? double(z[q[r]]) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = min(k0 + u, nv - 1);
                    const float* v = g.V + size_t(k) * g.vs;
                    float vv[kGmRows];
#pragma unroll
                    for (int r = 0; r < kGmRows; ++r)
                        vv[r] = v[q[r]];
#pragma unroll
                    for (int r = 0; r < kGmRows; ++r)
                        acc[u] += double(vv[r]) * zr[r];
                }
            }
        }
        gm_store4(a, g, P, k0, nv, acc);
    }
}

// Partials of v_k . z for k < nv, z = v_zslot.
__global__ __launch_bounds__(kBlock) void k_bgm_dots(BArgs a, BGm g, int nv, int zslot)
{
    const int W = a.W, P = kBlock / W, t = int(threadIdx.x);
    const bool live = t < P * W;
    const int sl = t % W, rg = t / W;
    const bool on = live && gm_on(a, a.eq[sl]);
    int64_t i0, i1;
    block_range(a.n, int64_t(kGmRows) * P, i0, i1);
    gm_dots4(a, g, P, rg, sl, on, i0, i1, nv, g.V + size_t(zslot) * g.vs);
}

// z = v_zslot; z -= sum_k h_k v_k (k < nv, h = h1 or h2 of the slot's
// equation, accumulated in FP64 in order k), stored back; then the partials
// of v_k . z (norm == 0) or of z . z (norm == 1, one value).
__global__ __launch_bounds__(kBlock) void k_bgm_orth(BArgs a, BGm g, int nv, int zslot, int which,
                                                     int norm)
{
    const int W = a.W, P = kBlock / W, t = int(threadIdx.x);
    const bool live = t < P * W;
    const int sl = t % W, rg = t / W;
    const int e = live ? a.eq[sl] : -1;
    const bool on = live && gm_on(a, e);
    int64_t i0, i1;
    block_range(a.n, int64_t(kGmRows) * P, i0, i1);
    float* z = g.V + size_t(zslot) * g.vs;
    double nrm = 0.0;
    if (on) {
        const double* h = gm_h(g, e, which);
        for (int64_t base = i0; base < i1; base += int64_t(kGmRows) * P) {
            double zr[kGmRows];
            size_t q[kGmRows];
            bool in[kGmRows];
#pragma unroll
            for (int r = 0; r < kGmRows; ++r) {
                const int64_t i = base + rg + int64_t(r) * P;
                in[r] = i < i1;
                q[r] = size_t(in[r] ? i : i0) * W + sl;
                zr[r] = double(z[q[r]]);
            }
            for (int k0 = 0; k0 < nv; k0 += 4) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int k = k0 + u;
                    const int kc = min(k, nv - 1);
                    const double hk = k < nv ? h[kc] : 0.0;
                    const float* v = g.V + size_t(kc) * g.vs;
                    float vv[kGmRows];
#pragma unroll
                    for (int r = 0; r < kGmRows; ++r)
                        vv[r] = v[q[r]];
                    if (k < nv)
#pragma unroll
                        for (int r = 0; r < kGmRows; ++r)
                            zr[r] -= hk * double(vv[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < kGmRows; ++r) {
                if (!in[r])
                    continue; // the clamped row's update is dropped
                const float zt = float(zr[r]);
                z[q[r]] = zt;
                nrm += double(zt) * double(zt);
            }
        }
    }
    if (norm) {
        double acc[4] = {nrm, 0.0, 0.0, 0.0};
        gm_store4(a, g, P, 0, 1, acc);
        return;
    }
    // each thread reads back only the rows it wrote
    gm_dots4(a, g, P, rg, sl, on, i0, i1, nv, z);
}

// h_which[k] of each slot's equation = the fold of its partials; one block
// per (slot, k < nv).
__global__ __launch_bounds__(kBlock) void k_bgm_fold(BArgs a, BGm g, int nv, int which, int G)
{
    const int sl = int(blockIdx.x) / nv, k = int(blockIdx.x) % nv;
    const int e = a.eq[sl];
    if (!gm_on(a, e))
        return;
    double tot[1] = {strided_sum(g.part + (size_t(sl) * (g.m + 1) + k) * G, G)};
    __shared__ double lds[kBlock / kWave];
    block_sum<1>(tot, lds);
    if (threadIdx.x == 0)
        gm_h(g, e, which)[k] = tot[0];
}

// Cycle start, one block per slot: beta = ||v_0|| (v_0 = M^-1 r of the
// sweep), s = (beta, 0, ...), scale 1 / beta; a zero or non-finite beta ends
// the slot's cycle at once (no correction).
__global__ __launch_bounds__(kBlock) void k_bgm_begin(BArgs a, BGm g, int G)
{
    const int sl = int(blockIdx.x);
    const int e = a.eq[sl];
    if (e < 0 || a.st[e].done)
        return;
    double tot[1] = {strided_sum(g.part + size_t(sl) * (g.m + 1) * G, G)};
    __shared__ double lds[kBlock / kWave];
    block_sum<1>(tot, lds);
    if (threadIdx.x != 0)
        return;
    BEq& q = a.st[e];
    const double beta = sqrt(tot[0]);
    q.gj = 0;
    q.gbeta = beta;
    double* sv = gm_sv(g, e);
    for (int i = 0; i <= g.m; ++i)
        sv[i] = 0.0;
    sv[0] = beta;
    q.gend = (beta > 0.0 && isfinite(beta)) ? 0 : 1;
    q.gscale = beta > 0.0 ? 1.0 / beta : 1.0;
}

// v_slot *= the slot's scale
__global__ __launch_bounds__(kBlock) void k_bgm_scale(BArgs a, BGm g, int vslot)
{
    const int W = a.W, P = kBlock / W, t = int(threadIdx.x);
    if (t >= P * W)
        return;
    const int sl = t % W, rr = t / W;
    const int e = a.eq[sl];
    if (!gm_on(a, e))
        return;
    const double sc = a.st[e].gscale;
    float* v = g.V + size_t(vslot) * g.vs;
    int64_t i0, i1;
    block_range(a.n, P, i0, i1);
    for (int64_t i = i0 + rr; i < i1; i += P) {
        const size_t q = size_t(i) * W + sl;
        v[q] = float(double(v[q]) * sc);
    }
}

// Arnoldi step j, one block per slot: Hessenberg column j = h1 + h2 and
// ||z||, the Givens rotations and the residual estimate of GMRESmSolver::solve
// (LinearSolvers.cpp:303-326); the cycle ends when the estimate relative to
// the cycle's start norm meets the sweep's inner tolerance, at m steps, at
// the iteration limit or on a singular column.
__global__ __launch_bounds__(kBlock) void k_bgm_step(BArgs a, BGm g, int j, int G, int max_iter)
{
    const int sl = int(blockIdx.x);
    const int e = a.eq[sl];
    if (!gm_on(a, e))
        return;
    double tot[1] = {strided_sum(g.part + size_t(sl) * (g.m + 1) * G, G)};
    __shared__ double lds[kBlock / kWave];
    block_sum<1>(tot, lds);
    if (threadIdx.x != 0)
        return;
    BEq& q = a.st[e];
    double* Hj = gm_H(g, e) + size_t(j) * size_t(g.m + 1);
    double* cs = gm_cs(g, e);
    double* sn = gm_sn(g, e);
    double* sv = gm_sv(g, e);
    const double* h1 = gm_h(g, e, 1);
    const double* h2 = gm_h(g, e, 2);
    for (int i = 0; i <= j; ++i)
        Hj[i] = h1[i] + h2[i];
    const double hn = sqrt(tot[0]);
    Hj[j + 1] = hn;
    q.gscale = hn != 0.0 ? 1.0 / hn : 1.0;
    for (int i = 0; i < j; ++i) {
        const double tt = cs[i] * Hj[i] + sn[i] * Hj[i + 1];
        Hj[i + 1] = -sn[i] * Hj[i] + cs[i] * Hj[i + 1];
        Hj[i] = tt;
    }
    const double dd = hypot(Hj[j], Hj[j + 1]);
    if (!(dd > 0.0)) {
        q.gend = 1; // singular column: the cycle keeps the steps before it
        return;
    }
    cs[j] = Hj[j] / dd;
    sn[j] = Hj[j + 1] / dd;
    Hj[j] = dd;
    Hj[j + 1] = 0.0;
    const double tt = cs[j] * sv[j];
    sv[j + 1] = -sn[j] * sv[j];
    sv[j] = tt;
    q.gj = j + 1;
    ++q.iter;
    ++q.total;
    q.last = fabs(sv[j + 1]) / q.gbeta;
    if (q.last < q.tol_in || j + 1 == g.m || q.total >= max_iter)
        q.gend = 1;
}

// Cycle end, one thread per slot: y = H^-1 s over the cycle's steps.
__global__ void k_bgm_backsolve(BArgs a, BGm g)
{
    const int sl = int(threadIdx.x);
    if (sl >= a.W)
        return;
    const int e = a.eq[sl];
    if (e < 0 || a.st[e].done)
        return;
    const int jc = a.st[e].gj;
    const double* H = gm_H(g, e);
    const double* sv = gm_sv(g, e);
    double* y = gm_y(g, e);
    const size_t ldh = size_t(g.m + 1);
    for (int i = jc - 1; i >= 0; --i) {
        double sum = sv[i];
        for (int k = i + 1; k < jc; ++k)
            sum -= H[size_t(k) * ldh + i] * y[k];
        y[i] = sum / H[size_t(i) * ldh + i];
    }
}

// xs += sum_{k < gj} y_k v_k for every slot of the sweep
__global__ __launch_bounds__(kBlock) void k_bgm_xupdate(BArgs a, BGm g)
{
    const int W = a.W, P = kBlock / W, t = int(threadIdx.x);
    if (t >= P * W)
        return;
    const int sl = t % W, rr = t / W;
    const int e = a.eq[sl];
    if (e < 0 || a.st[e].done)
        return;
    const int jc = a.st[e].gj;
    if (jc == 0)
        return;
    double y[kBGmMax];
    const double* yg = gm_y(g, e);
#pragma unroll
    for (int k = 0; k < kBGmMax; ++k)
        y[k] = k < jc ? yg[k] : 0.0;
    int64_t i0, i1;
    block_range(a.n, P, i0, i1);
    for (int64_t i = i0 + rr; i < i1; i += P) {
        const size_t q = size_t(i) * W + sl;
        double xv = a.xs64 ? a.xs64[q] : double(a.xs[q]);
#pragma unroll
        for (int k = 0; k < kBGmMax; ++k)
            if (k < jc)
                xv += y[k] * double(g.V[size_t(k) * g.vs + q]);
        if (a.xs64)
            a.xs64[q] = xv;
        else
            a.xs[q] = float(xv);
    }
}

// out = D^-1 in (Jacobi, or the identity with invD = 1) for the slots of a
// running GMRES cycle.
__global__ __launch_bounds__(kBlock) void k_bgm_jacobi(BArgs a, const float* __restrict__ in,
                                                       float* __restrict__ out)
{
    const int W = a.W, P = kBlock / W, t = int(threadIdx.x);
    if (t >= P * W)
        return;
    const int sl = t % W, rr = t / W;
    if (!gm_on(a, a.eq[sl]))
        return;
    int64_t i0, i1;
    block_range(a.n, P, i0, i1);
    for (int64_t i = i0 + rr; i < i1; i += P) {
        const size_t q = size_t(i) * W + sl;
        out[q] = (a.shared ? a.invD[i] : a.invD[q]) * in[q];
    }
}

// ------------------------------------------------------------ batched AMG
// The smoothed-aggregation V-cycle of one shared matrix applied to every
// slot of the layout at once: the sequence of the single-vector cycle
// (DHierarchy::cycle: Chebyshev pre-smoothing from x = 0, restriction of the
// residual, the dense FP64 coarsest solve, prolongation and post-smoothing)
// on FP32 SELL-64 operators (square ones in the difference form), each
// streamed once per launch for all W right-hand sides.  Vectors of level l
// are [n_l][W] lines.
struct BSell
{
    int nr;
    int diff; // square operator in the FP32 difference form (kSellDiag marks a_ii's slot)
    int band; // row walk (row_walk)
    const int* soff;  // slice offsets, 64 rows a slice
    const int* eci;   // column per slot, kSellDiag on a diagonal, -1 padding
    const float* v;   // value per slot
};

// Smoother start from x = 0 over n W elements: r = b, d = a0 D^-1 b, x = d
// (the levels whose rows are long: there the separate start is cheaper than
// gathering a0 D^-1 b in the first product, 14 + 78 against 116 us on the
// 451k-row, 24-entry level 1 of the 128^3 pressure hierarchy at W = 8).
__global__ __launch_bounds__(kBlock) void k_bamg_init(int64_t n, int W, const float* __restrict__ b,
                                                      const float* __restrict__ invD, float a0,
                                                      float* __restrict__ x, float* __restrict__ r,
                                                      float* __restrict__ d)
{
    const int64_t m = n * W;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x; q < m; q += stride) {
        const float bi = b[q];
        const float di = a0 * invD[q / W] * bi;
        r[q] = bi;
        d[q] = di;
        x[q] = di;
    }
}

// One launch of the batched V-cycle over the rows of A and lines of W
// slots, in the SpMM's thread layout: TW = W / V threads per row, each
// taking V consecutive slots (V-wide loads of the gathered rows of din, one
// load of a nonzero for all of them); every operator in SELL-64 slot order
// (slot k of row i at soff[i / 64] + 64 k + i % 64, CSR order within the
// row), so that the threads of a wave load consecutive slots:
//   mode 0 (Chebyshev step): r -= A din; dout = beta din + alpha D^-1 r; x += dout
//   mode 1 (residual):       r -= A din
//   mode 2 (restriction):    y = A din
//   mode 3 (prolongation):   y = A din; x += y
//   mode 4 (last step):      mode 0 without storing r and dout
//   mode 5 (first step):     din = b, d0 = a0 D^-1 b: r = b - A d0;
//                            dout = beta d0 + alpha D^-1 r; x = d0 + dout
//   mode 6 (start, degree 1): din = b, d0 = a0 D^-1 b: r = b - A d0; x = d0
// (modes 5 and 6 start the smoother from x = 0 in the same launch as its
// first product, gathering a0 D^-1 b in place of d0).
template <int V>
__global__ __launch_bounds__(kBlock) void k_bamg_rows(int W, BSell A, int mode,
                                                      const float* __restrict__ din,
                                                      float* __restrict__ dout,
                                                      float* __restrict__ r,
                                                      const float* __restrict__ invD,
                                                      float* __restrict__ x,
                                                      float* __restrict__ y, float beta, float alpha,
                                                      float a0)
{
    using VT = typename SlotVec<V>::type;
    constexpr int U = V == 1 ? kSpmmUnroll : kSpmmUnroll / 2; // row entries per pass
    const int TW = W / V;
    const int R = kBlock / TW;
    const int t = int(threadIdx.x);
    if (t >= R * TW) // no barrier below
        return;
    const int grp = t / TW;
    const int s0 = (t % TW) * V;
    const RowWalk rw = row_walk(0, A.nr, R, A.band);
    auto ld = [&](const float* p, size_t q) { return *reinterpret_cast<const VT*>(p + q); };
    auto st = [&](float* p, size_t q, VT v) { *reinterpret_cast<VT*>(p + q) = v; };
    const bool from_b = mode >= 5;
    // the vector the operator multiplies, at row j (entries q .. q + V - 1)
    auto xin = [&](size_t qq, int64_t j) {
        const VT v = ld(din, qq);
        return from_b ? VT(a0 * invD[j]) * v : v;
    };
    for (int64_t pass = rw.base; pass < rw.lim; pass += rw.stride) {
        const int64_t i = pass + grp;
        if (i >= rw.lim)
            break;
        const size_t q = size_t(i) * W + s0;
        const VT xi = A.diff ? xin(q, i) : VT(0.f);
        VT acc = VT(0.f);
        const int slc = int(i / kSellSlice);
        const int m0 = A.soff[slc] + int(i % kSellSlice);
        const int w = (A.soff[slc + 1] - A.soff[slc]) / kSellSlice;
        for (int k0 = 0; k0 < w; k0 += U) {
            int c[U];
            float v[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const size_t mk = size_t(m0) + size_t(k0 + j) * kSellSlice;
                const bool in = k0 + j < w;
                c[j] = in ? A.eci[mk] : -1;
                v[j] = in ? A.v[mk] : 0.f;
            }
            VT xv[U];
#pragma unroll
            for (int j = 0; j < U; ++j) {
                const int cj = c[j] & kSellColMask;
                xv[j] = c[j] >= 0 ? xin(size_t(cj) * W + s0, cj) : VT(0.f);
            }
#pragma unroll
            for (int j = 0; j < U; ++j)
                if (c[j] >= 0)
                    acc += VT(v[j]) * ((c[j] & kSellDiag) ? xv[j] : xv[j] - xi);
        }
        if (from_b) {
            const VT b = ld(din, q);
            const VT d0 = VT(a0 * invD[i]) * b;
            const VT ri = b - acc;
            st(r, q, ri);
            if (mode == 5) {
                const VT dn = VT(beta) * d0 + VT(alpha * invD[i]) * ri;
                st(dout, q, dn);
                st(x, q, d0 + dn);
            } else {
                st(x, q, d0);
            }
        } else if (mode == 4) {
            const VT ri = ld(r, q) - acc;
            const VT dn = VT(beta) * ld(din, q) + VT(alpha * invD[i]) * ri;
            st(x, q, ld(x, q) + dn);
        } else if (mode <= 1) {
            const VT ri = ld(r, q) - acc;
            st(r, q, ri);
            if (mode == 0) {
                const VT dn = VT(beta) * ld(din, q) + VT(alpha * invD[i]) * ri;
                st(dout, q, dn);
                st(x, q, ld(x, q) + dn);
            }
        } else {
            st(y, q, acc);
            if (mode == 3)
                st(x, q, ld(x, q) + acc);
        }
    }
}
using BAmgRows = void (*)(int, BSell, int, const float*, float*, float*, const float*, float*,
                          float*, float, float, float);
inline BAmgRows bamg_rows_kernel(int V)
{
    return V == 4 ? k_bamg_rows<4> : V == 2 ? k_bamg_rows<2> : k_bamg_rows<1>;
}

// x = Ainv b on the coarsest level (FP64 dense), one workgroup per row:
// thread t sums the columns j = t / W (mod kBlock / W) of slot t % W.
__global__ __launch_bounds__(kBlock) void k_bamg_coarse(int nc, int W,
                                                        const double* __restrict__ inv,
                                                        const float* __restrict__ b,
                                                        float* __restrict__ x)
{
    __shared__ double lds[kBlock];
    const int row = int(blockIdx.x);
    const int P = kBlock / W, t = int(threadIdx.x);
    const int s = t % W, jg = t / W;
    double acc = 0.0;
    if (t < P * W) {
        const double* a = inv + size_t(row) * size_t(nc);
        for (int j = jg; j < nc; j += P)
            acc += a[j] * double(b[size_t(j) * W + s]);
    }
    lds[t] = acc;
    __syncthreads();
    if (t < W) {
        double sum = 0.0;
        for (int q = 0; q < P; ++q)
            sum += lds[q * W + t];
        x[size_t(row) * W + t] = float(sum);
    }
}

// ILU(0) apply out = (LU)^-1 in under the two-colour ordering: rows [0, n0)
// have no lower entries (their neighbours all come later), rows [n0, n) no
// upper entries, and the factorization changes only the second colour's
// pivots (ILU0Preconditioner::setup, src/solver/LinearSolvers.cpp:81-124,
// on this ordering).  The forward and backward sweeps of apply (:126-150)
// are then one launch each:
//   phase 0, rows [n0, n):  out_i = (in_i - sum_k l_ik in_k) / d_i
//            (the forward sweep leaves the first colour's y_k = in_k)
//   phase 1, rows [0, n0):  out_i = (in_i - sum_j a_ij out_j) / a_ii
// with 1/d_i in invD and l_ik in lv at the positions of a_ik.  Slots of
// equations not iterating are left alone.
template <bool kSell, int V, bool kShared>
__global__ __launch_bounds__(kBlock) void k_bilu(BArgs a, const float* __restrict__ in,
                                                 float* __restrict__ out, int phase, int n0)
{
    using VT = typename SlotVec<V>::type;
    constexpr int U = V == 1 ? kSpmmUnroll : kSpmmUnroll / 2;
    const int W = a.W;
    const int TW = W / V;
    const int R = kBlock / TW;
    const int t = int(threadIdx.x);
    const bool live = t < R * TW;
    const int grp = t / TW;
    const int s0 = (t % TW) * V;
    bool on[V];
    bool any = false, all = true;
#pragma unroll
    for (int u = 0; u < V; ++u) {
        on[u] = live && slot_on(a, B_ALPHA, a.eq[s0 + u]);
        any = any || on[u];
        all = all && on[u];
    }
    if (!any)
        return;
    const int64_t lo = phase == 0 ? n0 : 0, hi = phase == 0 ? a.n : n0;
    const float* fac = phase == 0 ? a.lv : a.v32;
    const float* src = phase == 0 ? in : out;
    const RowWalk rw = row_walk(lo, hi - lo, R, a.band);
    for (int64_t pass = rw.base; pass < rw.lim; pass += rw.stride) {
        const int64_t i = pass + grp;
        if (i >= rw.lim)
            break;
        const size_t q = size_t(i) * W + s0;
        VT acc = VT(0.f);
        auto fetch = [&](int64_t mk, int c, VT& f, VT& xv) {
            if constexpr (kShared)
                f = VT(fac[mk]);
            else
                f = __builtin_nontemporal_load(reinterpret_cast<const VT*>(fac + mk * W + s0));
            xv = *reinterpret_cast<const VT*>(src + size_t(c) * W + s0);
        };
        if constexpr (kSell) {
            const int slc = int(i / kSellSlice);
            const int m0 = a.soff[slc] + int(i % kSellSlice);
            const int w = (a.soff[slc + 1] - a.soff[slc]) / kSellSlice;
            for (int k0 = 0; k0 < w; k0 += U) {
                int c[U];
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    const size_t mk = size_t(m0) + size_t(k0 + j) * kSellSlice;
                    c[j] = k0 + j < w ? a.eci[mk] : -1;
                }
                VT f[U], xv[U];
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    const bool off = c[j] >= 0 && !(c[j] & kSellDiag);
                    if (off)
                        fetch(int64_t(m0) + int64_t(k0 + j) * kSellSlice, c[j] & kSellColMask, f[j],
                              xv[j]);
                    else
                        f[j] = xv[j] = VT(0.f);
                }
#pragma unroll
                for (int j = 0; j < U; ++j)
                    acc += f[j] * xv[j];
            }
        } else {
            const int kb = a.rp[i], ke = a.rp[i + 1];
            for (int k0 = kb; k0 < ke; k0 += U) {
                int c[U];
#pragma unroll
                for (int j = 0; j < U; ++j)
                    c[j] = k0 + j < ke ? a.ci[k0 + j] : -1;
                VT f[U], xv[U];
#pragma unroll
                for (int j = 0; j < U; ++j) {
                    if (c[j] >= 0 && c[j] != int(i))
                        fetch(k0 + j, c[j], f[j], xv[j]);
                    else
                        f[j] = xv[j] = VT(0.f);
                }
#pragma unroll
                for (int j = 0; j < U; ++j)
                    acc += f[j] * xv[j];
            }
        }
        const VT xi = *reinterpret_cast<const VT*>(in + q);
        VT dv;
        if constexpr (kShared)
            dv = VT(a.invD[i]);
        else
            dv = *reinterpret_cast<const VT*>(a.invD + q);
        const VT res = (xi - acc) * dv;
        if (all) {
            *reinterpret_cast<VT*>(out + q) = res;
        } else {
#pragma unroll
            for (int u = 0; u < V; ++u)
                if (on[u])
                    out[q + u] = SlotVec<V>::at(res, u);
        }
    }
}

using BIluKernel = void (*)(BArgs, const float*, float*, int, int);
template <bool kShared> BIluKernel bilu_kernel_v(bool sell, int V)
{
    if (V == 4)
        return sell ? k_bilu<true, 4, kShared> : k_bilu<false, 4, kShared>;
    if (V == 2)
        return sell ? k_bilu<true, 2, kShared> : k_bilu<false, 2, kShared>;
    return sell ? k_bilu<true, 1, kShared> : k_bilu<false, 1, kShared>;
}
BIluKernel bilu_kernel(bool sell, bool shared, int W, int vmax)
{
    const int V = slots_per_thread(W, vmax);
    return shared ? bilu_kernel_v<true>(sell, V) : bilu_kernel_v<false>(sell, V);
}

// ILU(0) factors of the equations e0 .. e0 + ne - 1 (FP64 from the FP64
// values, stored in FP32), one thread per (equation, row), under the
// two-colour ordering: a row i >= n0 takes, in its CSR order (every entry
// but the diagonal precedes it), l_ik = a_ik / a_kk and d_i -= l_ik a_ki, the
// IKJ loop of ILU0Preconditioner::setup (LinearSolvers.cpp:98-123) on this
// ordering, where row k's only upper entry found in row i is a_ki; a zero
// pivot a_kk leaves l_ik = a_ik and d_i alone (:104-106).  A row i < n0 keeps
// d_i = a_ii.  ipiv = 1/d_i (1 for d_i = 0: the apply then leaves s, :148).
// tpos[k]: position of the transposed entry; vpos[k]: the value line of
// entry k in the working layout (the SELL-64 slot, or k).
__global__ __launch_bounds__(kBlock) void k_bilu_factor(
    int n, int n0, int ne, int64_t vstride, int64_t lstride, const int* __restrict__ rp,
    const int* __restrict__ ci, const int* __restrict__ diag, const int* __restrict__ tpos,
    const int* __restrict__ vpos, const double* __restrict__ v64, float* __restrict__ lv,
    float* __restrict__ ipiv)
{
    const int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= int64_t(n) * ne)
        return;
    const int e = int(t / n), i = int(t - int64_t(e) * n);
    const double* va = v64 + size_t(e) * vstride;
    float* le = lv + size_t(e) * lstride;
    const int kd = diag[i];
    double d = va[kd];
    if (i >= n0) {
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int c = ci[k];
            if (c >= i)
                break;
            const double piv = va[diag[c]];
            double l = va[k];
            if (piv != 0.0) {
                l = l / piv;
                d -= l * va[tpos[k]];
            }
            le[vpos[k]] = float(l);
        }
    }
    ipiv[size_t(e) * n + i] = float(d != 0.0 ? 1.0 / d : 1.0);
}

// Equation-major stages.  Per-workgroup sums of one value per equation,
// stored as partials[e][block].
__device__ __forceinline__ void eq_partial(int e, double v, double* lds)
{
    v = wave_sum(v);
    if ((threadIdx.x & (kWave - 1)) == 0)
        lds[e * kWaves + threadIdx.x / kWave] += v;
}

__device__ __forceinline__ void eq_store(const HArgs& a, const double* lds)
{
    __syncthreads();
    const int G = gridDim.x;
    for (int e = threadIdx.x; e < a.E; e += kBlock) {
        double s = 0.0;
        for (int q = 0; q < kWaves; ++q)
            s += lds[e * kWaves + q];
        a.partials[size_t(e) * G + blockIdx.x] = s;
    }
}

// ||b_e||^2 of every equation, over units (equation, chunk of rows).
constexpr int kChunkRows = kBlock * 8;
__global__ __launch_bounds__(kBlock) void k_bnorm(HArgs a)
{
    __shared__ double lds[kMaxEq * kWaves];
    for (int t = threadIdx.x; t < a.E * kWaves; t += kBlock)
        lds[t] = 0.0;
    __syncthreads();
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int nch = (a.n + kChunkRows - 1) / kChunkRows;
    const int64_t units = int64_t(a.E) * nch;
    const int64_t per = (units + G - 1) / G;
    const int64_t u0 = lb * per, u1 = min(units, u0 + per);
    for (int64_t u = u0; u < u1; ++u) {
        const int e = int(u / nch);
        const int c0 = int(u % nch) * kChunkRows, c1 = min(a.n, c0 + kChunkRows);
        const double* be = a.b + size_t(e) * a.n;
        double s = 0.0;
        for (int i = c0 + threadIdx.x; i < c1; i += kBlock)
            s += be[i] * be[i];
        eq_partial(e, s, lds);
    }
    eq_store(a, lds);
}

// rt_e = b_e - A_e x_e in FP64 for the unfinished equations, over units
// (equation, row block) of the pattern's CSR-stream row blocks; partials of
// ||rt_e||^2.
__global__ __launch_bounds__(kBlock) void k_btrue(HArgs a)
{
    __shared__ double prod[kStreamNnz];
    __shared__ double lds[kMaxEq * kWaves];
    __shared__ int act[kMaxEq];
    __shared__ int nact;
    for (int t = threadIdx.x; t < a.E * kWaves; t += kBlock)
        lds[t] = 0.0;
    if (threadIdx.x == 0) {
        int c = 0;
        for (int e = 0; e < a.E; ++e)
            if (!a.st[e].fin)
                act[c++] = e;
        nact = c;
    }
    __syncthreads();
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int64_t units = int64_t(nact) * a.nrb;
    const int64_t per = (units + G - 1) / G;
    const int64_t u0 = lb * per, u1 = min(units, u0 + per);
    const int tid = threadIdx.x;
    for (int64_t u = u0; u < u1; ++u) {
        const int e = act[u / a.nrb];
        const int rb = int(u % a.nrb);
        const double* __restrict__ ve = a.v64 + size_t(e) * a.vstride;
        const double* __restrict__ xe = a.x + size_t(e) * a.n;
        const size_t off = size_t(e) * a.n;
        const int r0 = a.blk[rb], r1 = a.blk[rb + 1];
        const int k0 = a.rp[r0], k1 = a.rp[r1];
        if (k1 - k0 > kStreamNnz) {
            double s = 0.0;
            for (int k = k0 + tid; k < k1; k += kBlock)
                s += ve[k] * xe[a.ci[k]];
            prod[tid] = s;
            __syncthreads();
            double rr = 0.0;
            if (tid == 0) {
                double t = 0.0;
                for (int j = 0; j < kBlock; ++j)
                    t += prod[j];
                const double ri = a.b[off + r0] - t;
                a.rt[off + r0] = ri;
                rr = ri * ri;
            }
            eq_partial(e, rr, lds);
            __syncthreads();
            continue;
        }
        for (int k = k0 + tid; k < k1; k += kBlock)
            prod[k - k0] = ve[k] * xe[a.ci[k]];
        __syncthreads();
        const int i = r0 + tid;
        double rr = 0.0;
        if (i < r1) {
            double s = 0.0;
            const int end = a.rp[i + 1] - k0;
            for (int k = a.rp[i] - k0; k < end; ++k)
                s += prod[k];
            const double ri = a.b[off + i] - s;
            a.rt[off + i] = ri;
            rr = ri * ri;
        }
        eq_partial(e, rr, lds);
        __syncthreads();
    }
    eq_store(a, lds);
}

struct BFinArgs
{
    int E, W, G, max_iter, max_outer;
    double tol, inner_tol;
};

// Updates one equation's scalars from the totals of a stage (the
// finalizations of CGSolver::solve and BiCGSTABSolver::solve).
__device__ void b_finalize(BEq& s, int op, const double (&tot)[2], const BFinArgs& f)
{
    switch (op) {
    case B_BNORM:
        s.bnorm = fmax(sqrt(tot[0]), 1e-30);
        s.fin = 0;
        s.done = 1;
        s.total = 0;
        s.outer = 0;
        s.status = 0;
        s.rel = 1.0;
        if (tot[0] == 0.0) { // zero rhs: x = 0 (x is zeroed before the solve)
            s.fin = 1;
            s.rel = 0.0;
        }
        break;
    case B_TRUE: {
        if (s.fin)
            break;
        const double rn = sqrt(tot[0]);
        s.rel = rn / s.bnorm;
        if (s.rel < f.tol) {
            s.fin = 1;
            s.done = 1;
        } else if (s.outer >= f.max_outer || s.total >= f.max_iter || !isfinite(rn)) {
            s.fin = 1; // not converged: status marks it
            s.done = 1;
            s.status = 1;
        } else {
            s.done = 0;
            s.iter = 0;
            s.gj = 0;
            s.gend = 0; // a GMRES cycle starts with the sweep
            ++s.outer;
            s.tol_in = fmax(f.inner_tol, 0.5 * f.tol * s.bnorm / rn);
        }
        break;
    }
    case B_INIT:
        if (s.done)
            break;
        s.rz = tot[0];
        s.r0 = sqrt(tot[1]);
        if (!(s.r0 > 0.0))
            s.done = 1;
        break;
    case B_ALPHA:
        if (s.done)
            break;
        if (tot[0] > 0.0 && isfinite(tot[0])) {
            s.alpha = s.rz / tot[0];
        } else {
            s.alpha = 0.0;
            s.done = 1; // breakdown: the outer true residual decides
        }
        break;
    case B_UPDATE:
        if (s.done)
            break;
        ++s.iter;
        ++s.total;
        s.last = sqrt(tot[0]) / s.r0;
        if (s.last < s.tol_in || s.total >= f.max_iter) {
            s.done = 1;
        } else {
            s.beta = tot[1] / s.rz;
            s.rz = tot[1];
        }
        break;
    case B_BI_INIT:
        if (s.done)
            break;
        s.r0 = sqrt(tot[0]);
        s.sconv = 0;
        s.rho = 1.0;
        s.alpha = 1.0;
        s.omega = 1.0;
        s.rz = tot[1]; // rho_new = rh.r
        if (!(s.r0 > 0.0) || s.rz == 0.0)
            s.done = 1;
        else
            s.beta = (s.rz / s.rho) * (s.alpha / s.omega);
        break;
    case B_BI_ALPHA:
        if (s.done)
            break;
        if (tot[0] != 0.0 && isfinite(tot[0])) {
            s.alpha = s.rz / tot[0];
        } else {
            s.alpha = 0.0;
            s.done = 1; // breakdown: the outer true residual decides
        }
        break;
    case B_BI_S:
        if (s.done)
            break;
        if (sqrt(tot[0]) / s.r0 < s.tol_in)
            s.sconv = 1;
        break;
    case B_BI_OMEGA:
        if (s.done || s.sconv)
            break;
        s.omega = tot[0] / fmax(tot[1], 1e-30);
        break;
    case B_BI_UPDATE:
        if (s.done)
            break;
        ++s.iter;
        ++s.total;
        if (s.sconv) {
            s.done = 1;
            break;
        }
        s.last = sqrt(tot[0]) / s.r0;
        if (s.last < s.tol_in || s.total >= f.max_iter || tot[1] == 0.0 || s.omega == 0.0) {
            s.done = 1;
        } else {
            s.rho = s.rz;
            s.rz = tot[1];
            s.beta = (s.rz / s.rho) * (s.alpha / s.omega);
        }
        break;
    default:
        break;
    }
}

// One workgroup per slot: fold the slot's G partials (fixed order) and
// update its equation's scalars.
__global__ __launch_bounds__(kBlock) void k_bfin(const double* __restrict__ partials,
                                                 const int* __restrict__ eq, BEq* st, int op,
                                                 BFinArgs f)
{
    const int sl = blockIdx.x;
    const int e = eq[sl];
    if (e < 0)
        return;
    const int nv = (op == B_INIT || op == B_UPDATE || op == B_BI_INIT || op == B_BI_OMEGA ||
                    op == B_BI_UPDATE)
                       ? 2
                       : 1;
    double tot[2] = {0.0, 0.0};
    for (int v = 0; v < nv; ++v)
        tot[v] = strided_sum(partials + (size_t(v) * f.W + sl) * f.G, f.G);
    __shared__ double lds[2 * kWaves];
    block_sum<2>(tot, lds);
    if (threadIdx.x == 0)
        b_finalize(st[e], op, tot, f);
}

// One workgroup per equation for the equation-major stages.
__global__ __launch_bounds__(kBlock) void k_bfin_eq(const double* __restrict__ partials, BEq* st,
                                                    int op, BFinArgs f)
{
    const int e = blockIdx.x;
    double tot[2] = {0.0, 0.0};
    tot[0] = strided_sum(partials + size_t(e) * f.G, f.G);
    __shared__ double lds[2 * kWaves];
    block_sum<2>(tot, lds);
    if (threadIdx.x == 0)
        b_finalize(st[e], op, tot, f);
}

// Home (equation-major, rows of length `rows`) <-> slot layout, through a
// tile of 64 rows x W slots in LDS so both sides are accessed in lines:
//   mode 0: slot[i W + s] = float(home[eq[s] * rows + i]), or `fill` for an
//           empty slot (for every slot when use_fill)
//   mode 1: home[eq[s] * rows + i] += H(slot[i W + s]) for the slots set in
//           `mask`
constexpr int kTileRows = 64;
template <class H, class T>
__global__ __launch_bounds__(kBlock) void k_btile(int64_t rows, int W, const int* __restrict__ eq,
                                                  uint64_t mask, int mode, int use_fill,
                                                  float fill, H* home, T* slot)
{
    __shared__ T tile[kTileRows * kMaxEq];
    const int64_t ntile = (rows + kTileRows - 1) / kTileRows;
    for (int64_t tb = blockIdx.x; tb < ntile; tb += gridDim.x) {
        const int64_t i0 = tb * kTileRows;
        const int nr = int(min(int64_t(kTileRows), rows - i0));
        if (mode == 0) {
            // read each slot's run of rows, write whole lines
            for (int t = threadIdx.x; t < kTileRows * W; t += kBlock) {
                const int sl = t / kTileRows, r = t % kTileRows;
                const int e = eq[sl];
                T v = T(fill);
                if (r < nr && e >= 0 && !use_fill)
                    v = T(home[size_t(e) * rows + i0 + r]);
                tile[r * W + sl] = v;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < nr * W; t += kBlock)
                slot[size_t(i0) * W + t] = tile[t];
        } else {
            for (int t = threadIdx.x; t < nr * W; t += kBlock)
                tile[t] = slot[size_t(i0) * W + t];
            __syncthreads();
            for (int t = threadIdx.x; t < kTileRows * W; t += kBlock) {
                const int sl = t / kTileRows, r = t % kTileRows;
                const int e = eq[sl];
                if (r < nr && e >= 0 && ((mask >> sl) & 1ull))
                    home[size_t(e) * rows + i0 + r] += H(tile[r * W + sl]);
            }
        }
        __syncthreads();
    }
}

// Narrower slot layout from a wider one: dst[i W2 + s] = src[i W1 + map[s]].
template <class T>
__global__ __launch_bounds__(kBlock) void k_brepack(int64_t rows, int W1, int W2,
                                                    const int* __restrict__ map,
                                                    const T* __restrict__ src,
                                                    T* __restrict__ dst)
{
    const int64_t m = rows * W2;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x; q < m; q += stride) {
        const int64_t i = q / W2;
        const int sl = int(q - i * W2);
        dst[q] = src[size_t(i) * W1 + map[sl]];
    }
}

// d[k] = src[map[k]] (map null: identity) in FP64 and, unless d32 is null, FP32
__global__ __launch_bounds__(kBlock) void k_bconvert(int64_t m, const int* __restrict__ map,
                                                     const double* __restrict__ src,
                                                     double* __restrict__ d64,
                                                     float* __restrict__ d32)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
        const double v = src[map ? map[i] : i];
        d64[i] = v;
        if (d32)
            d32[i] = float(v);
    }
}

// Equation-major vectors between the caller's numbering and the batch's:
// dir 0: dst[e n + i] = src[e n + perm[i]]; dir 1: dst[e n + perm[i]] = src[e n + i].
__global__ __launch_bounds__(kBlock) void k_bpermute(int64_t n, int E, const int* __restrict__ perm,
                                                     const double* __restrict__ src,
                                                     double* __restrict__ dst, int dir)
{
    const int64_t m = n * E;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < m; t += stride) {
        const int64_t e = t / n, i = t - e * n;
        const int64_t j = e * n + perm[i];
        if (dir == 0)
            dst[t] = src[j];
        else
            dst[j] = src[t];
    }
}

// FP32 values of one equation in SELL-64 slot order (0 in padding slots).
__global__ __launch_bounds__(kBlock) void k_bsell32(int64_t m, const int* __restrict__ eperm,
                                                    const double* __restrict__ v64,
                                                    float* __restrict__ v32)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
        const int k = eperm[i];
        v32[i] = k >= 0 ? float(v64[k]) : 0.f;
    }
}

// The FP32 difference form's diagonal: v32 at the slot of a_ii (vpos: SELL
// slot of each CSR entry, null for CSR order) = the FP64 sum of row i.
__global__ __launch_bounds__(kBlock) void k_bdiffdiag(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ diag,
                                                      const double* __restrict__ v64,
                                                      const int* __restrict__ vpos,
                                                      float* __restrict__ v32)
{
#pragma clang fp contract(off)
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double sum = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        sum += v64[k];
    const int k = diag[i];
    v32[vpos ? vpos[k] : k] = float(sum);
}

// Sum of the row margins (|a_ii| - sum_j!=i |a_ij|) / |a_ii| (-1 for a zero
// or missing diagonal) of each workgroup's rows: part[blockIdx].
__global__ __launch_bounds__(kBlock) void k_bmargin(int n, const int* __restrict__ rp,
                                                    const int* __restrict__ ci,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ part)
{
    __shared__ double lds[kBlock];
    double m = 0.0;
    for (int i = int(blockIdx.x) * kBlock + threadIdx.x; i < n; i += int(gridDim.x) * kBlock) {
        double d = 0.0, off = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            if (ci[k] == i)
                d = fabs(v[k]);
            else
                off += fabs(v[k]);
        }
        m += d > 0.0 ? (d - off) / d : -1.0;
    }
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h /= 2) {
        if (int(threadIdx.x) < h)
            lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        part[blockIdx.x] = lds[0];
}

// invD_e[i] = 1 / a_ii (1 when missing or zero: LinearSolvers.cpp:54-65)
__global__ __launch_bounds__(kBlock) void k_binvdiag(int n, const int* __restrict__ diag,
                                                     const double* __restrict__ val,
                                                     float* __restrict__ invD)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int k = diag[i];
    const double v = k >= 0 ? val[k] : 0.0;
    invD[i] = float(v != 0.0 ? 1.0 / v : 1.0);
}

// out = D^-1 in over the whole slot layout (every slot).
__global__ __launch_bounds__(kBlock) void k_bscale(BArgs a, const float* __restrict__ in,
                                                   float* __restrict__ out)
{
    const int64_t m = int64_t(a.n) * a.W;
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t q = int64_t(blockIdx.x) * kBlock + threadIdx.x; q < m; q += stride)
        out[q] = (a.shared ? a.invD[q / a.W] : a.invD[q]) * in[q];
}

__global__ __launch_bounds__(kBlock) void k_bfill(int64_t m, double* __restrict__ x)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride)
        x[i] = 0.0;
}

__global__ __launch_bounds__(kBlock) void k_bzero32(int64_t m, float* __restrict__ x)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride)
        x[i] = 0.f;
}

// Grid-barrier timing (mi_time_grid_barrier): one element per thread updated
// between barriers, cooperative grid.sync() against kernel boundaries.
__global__ __launch_bounds__(kBlock) void k_gbar_sync(int iterations, float* __restrict__ buf)
{
    cooperative_groups::grid_group g = cooperative_groups::this_grid();
    const size_t i = size_t(blockIdx.x) * kBlock + threadIdx.x;
    for (int it = 0; it < iterations; ++it) {
        buf[i] += 1.f;
        g.sync();
    }
}

__global__ __launch_bounds__(kBlock) void k_gbar_one(float* __restrict__ buf)
{
    buf[size_t(blockIdx.x) * kBlock + threadIdx.x] += 1.f;
}

} // namespace mi

using namespace mi;

// One slot layout: its width, slot table and the arrays of the working set.
// v32, invD (and lv with ILU(0)) are per-slot copies of the equations'
// values; a batch with one shared matrix reads its home arrays instead.
struct BSlots
{
    int W = 0;
    std::vector<int> eqs; // equation of each slot, -1 empty (size W)
    DevBuf<int> d_eq;
    DevBuf<float> v32, invD, lv, xs, r, p, Ap, rh, y, z, t;
    DevBuf<double> xs64; // the sweep's correction in FP64 (weakly dominant operators)
};

struct mi_batch
{
    std::shared_ptr<Pattern> src; // the caller's pattern
    std::shared_ptr<Pattern> pat; // working pattern: src, or src reordered (RCM / two colours)
    DevBuf<int> d_perm, d_vmap;   // perm[new] = old; vmap[k] = source entry of entry k
    bool reordered = false;
    bool diff = false; // FP32 values in the difference form (every row holds its diagonal)
    // mean diagonal-dominance margin (|a_ii| - sum_j!=i |a_ij|) / |a_ii| of
    // each matrix slot; below kX64Margin the sweeps sum their correction in FP64
    std::vector<double> margin;
    bool x64 = false;
    int E = 0, Wmax = 1;
    int G = 8;    // workgroups of the batched SpMM and ILU sweeps (autotuned)
    int Vmax = 4; // cap on slots per thread of those launches (autotuned)
    int band = 0; // banded row walk for those launches (autotuned)
    int Gv = 8;   // workgroups of the vector, norm and true-residual stages (from n)
    int64_t n = 0, nnz = 0;
    bool sell = false;   // values in the pattern's SELL-64 slot order
    int64_t vrows = 0;   // value lines: nnz, or the SELL-64 slots
    bool shared = false; // one matrix for every equation (multi-RHS)
    int Em = 1;          // matrices held: E, or 1 when shared
    // two-colour ordering for ILU(0): rows [0, n0) and [n0, n)
    bool ilu_order = false;
    int n0 = 0;
    DevBuf<int> d_tpos, d_vpos; // transposed entry; value line of each entry
    bool ilu_valid = false;     // lvh / ipivh hold the factors of the current values
    // equation-major ("home") arrays
    DevBuf<double> v64, b, x, rt;
    DevBuf<float> v32h, invDh, lvh, ipivh, ones;
    DevBuf<double> partials;
    DevBuf<BEq> st;
    BEq* h_st = nullptr;
    // two slot layouts (packing and repacking write the other one)
    BSlots set[2];
    int cur = 0;
    bool slots_valid = false; // set[cur] holds the home values of its equations
    int pc = MI_PC_JACOBI;    // preconditioner the slots were packed for
    bool ext_alloc = false;   // rh, y, z, t allocated
    // batched AMG (shared matrix): the hierarchy of the last setup, FP32 CSR
    // operators and [n][Wmax] work vectors per level, the dense FP64 inverse
    // of the coarsest operator
    struct ALevel
    {
        int n = 0, deg = 1, diff = 0;
        DevBuf<int> asoff, aeci, psoff, peci, rsoff, reci; // SELL-64 forms
        DevBuf<float> av, pv, rv, invD;
        DevBuf<float> r, d, d2, x, b, e;
        std::vector<double> alpha, beta;
    };
    std::vector<ALevel> amg;
    DevBuf<double> amg_cinv;
    int amg_nc = 0;
    bool amg_valid = false; // the hierarchy belongs to the current values
    // batched GMRES(m): basis ((m+1) slot-layout vectors of the widest
    // layout), per-equation small FP64 arrays, partials of the basis kernels
    int gm_m = 0;
    DevBuf<float> gmV;
    DevBuf<double> gmSmall, gmPart;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0, graph_kind = -1, graph_pc = -1, graph_set = -1, graph_W = 0;
    BFinArgs graph_f{}; // finalize arguments baked into the graph
    double tune_us = 0.0, tune_vec_us = 0.0;

    ~mi_batch()
    {
        if (graph)
            (void)hipGraphExecDestroy(graph);
        if (h_st)
            (void)hipHostFree(h_st);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
    BArgs args(int which) const
    {
        const BSlots& S = set[which];
        BArgs a{};
        a.n = int(n);
        a.W = S.W;
        a.nnz = nnz;
        a.rp = pat->rp.p;
        a.ci = pat->ci.p;
        a.soff = sell ? pat->sell.soff.p : nullptr;
        a.eci = sell ? pat->sell.eci.p : nullptr;
        a.shared = shared ? 1 : 0;
        a.band = band;
        a.diff = diff ? 1 : 0;
        if (shared) {
            a.v32 = v32h.p;
            a.invD = pc == MI_PC_ILU ? ipivh.p : pc == MI_PC_NONE ? ones.p : invDh.p;
            a.lv = lvh.p;
        } else {
            a.v32 = S.v32.p;
            a.invD = S.invD.p;
            a.lv = S.lv.p;
        }
        a.eq = S.d_eq.p;
        a.xs = S.xs.p;
        a.xs64 = x64 ? S.xs64.p : nullptr;
        a.r = S.r.p;
        a.p = S.p.p;
        a.Ap = S.Ap.p;
        a.rh = S.rh.p;
        a.y = S.y.p;
        a.z = S.z.p;
        a.t = S.t.p;
        a.st = st.p;
        a.partials = partials.p;
        return a;
    }
    HArgs hargs() const
    {
        HArgs a{};
        a.n = int(n);
        a.E = E;
        a.nrb = pat->rb.nrb;
        a.nnz = nnz;
        a.vstride = shared ? 0 : nnz;
        a.blk = pat->rb.blk.p;
        a.rp = pat->rp.p;
        a.ci = pat->ci.p;
        a.v64 = v64.p;
        a.b = b.p;
        a.x = x.p;
        a.rt = rt.p;
        a.st = st.p;
        a.partials = partials.p;
        return a;
    }
};

namespace
{

// Working-set arrays of a slot layout (sized for the widest layout; the
// vectors of BiCGStab and of a separate preconditioner apply when first
// needed).
int alloc_slots(mi_batch* B, BSlots& S, bool ext)
{
    const size_t nW = size_t(std::max<int64_t>(B->n, 1)) * size_t(B->Wmax);
    const size_t zW = size_t(std::max<int64_t>(B->vrows, 1)) * size_t(B->Wmax);
    if (!S.xs.p) {
        if (!B->shared) {
            MI_TRY(S.v32.alloc(zW));
            MI_TRY(S.invD.alloc(nW));
            if (B->ilu_order)
                MI_TRY(S.lv.alloc(zW));
        }
        for (DevBuf<float>* v : {&S.xs, &S.r, &S.p, &S.Ap})
            MI_TRY(v->alloc(nW));
        MI_TRY(S.d_eq.alloc(size_t(kMaxEq)));
    }
    if (B->x64 && !S.xs64.p) {
        MI_TRY(S.xs64.alloc(nW));
        MI_HIP(hipMemsetAsync(S.xs64.p, 0, nW * sizeof(double), B->stream));
    }
    if (ext && !S.rh.p)
        for (DevBuf<float>* v : {&S.rh, &S.y, &S.z, &S.t})
            MI_TRY(v->alloc(nW));
    return MI_OK;
}

int upload_eqs(mi_batch* B, BSlots& S)
{
    MI_HIP(hipMemcpyAsync(S.d_eq.p, S.eqs.data(), sizeof(int) * size_t(S.W),
                          hipMemcpyHostToDevice, B->stream));
    MI_HIP(hipStreamSynchronize(B->stream)); // S.eqs is host memory that may change
    return MI_OK;
}

unsigned tile_grid(int64_t rows)
{
    const int64_t t = (rows + kTileRows - 1) / kTileRows;
    return unsigned(std::max<int64_t>(1, std::min<int64_t>(t, 8192)));
}

// Slots for `eqs` in the other layout, values and preconditioner data from
// the home arrays (per-equation matrices); it becomes the current layout.
int pack_from_home(mi_batch* B, const std::vector<int>& eqs, bool ext)
{
    const int nxt = B->cur ^ 1;
    BSlots& S = B->set[nxt];
    MI_TRY(alloc_slots(B, S, ext));
    S.W = int(eqs.size());
    S.eqs = eqs;
    MI_TRY(upload_eqs(B, S));
    hipStream_t s = B->stream;
    if (!B->shared) {
        if (B->vrows)
            k_btile<float><<<tile_grid(B->vrows), kBlock, 0, s>>>(
                B->vrows, S.W, S.d_eq.p, 0ull, 0, 0, 0.f, B->v32h.p, S.v32.p);
        const bool ilu = B->pc == MI_PC_ILU;
        if (B->n)
            k_btile<float><<<tile_grid(B->n), kBlock, 0, s>>>(
                B->n, S.W, S.d_eq.p, 0ull, 0, B->pc == MI_PC_NONE ? 1 : 0, 1.f,
                ilu ? B->ipivh.p : B->invDh.p, S.invD.p);
        if (ilu && B->vrows)
            k_btile<float><<<tile_grid(B->vrows), kBlock, 0, s>>>(
                B->vrows, S.W, S.d_eq.p, 0ull, 0, 0, 0.f, B->lvh.p, S.lv.p);
        MI_HIP(hipGetLastError());
    }
    B->cur = nxt;
    B->slots_valid = true;
    return MI_OK;
}

// x_e += the current sweep's correction for the slots in `mask`.
int retire(mi_batch* B, uint64_t mask)
{
    const BSlots& S = B->set[B->cur];
    if (!mask || B->n == 0)
        return MI_OK;
    if (B->x64)
        k_btile<double><<<tile_grid(B->n), kBlock, 0, B->stream>>>(B->n, S.W, S.d_eq.p, mask, 1,
                                                                   0, 0.f, B->x.p, S.xs64.p);
    else
        k_btile<double><<<tile_grid(B->n), kBlock, 0, B->stream>>>(B->n, S.W, S.d_eq.p, mask, 1,
                                                                   0, 0.f, B->x.p, S.xs.p);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// The slots `keep` (positions in the current layout) of every working array
// into the other layout, keep.size() wide.
int repack(mi_batch* B, const std::vector<int>& keep, bool ext)
{
    const BSlots& A = B->set[B->cur];
    BSlots& S = B->set[B->cur ^ 1];
    MI_TRY(alloc_slots(B, S, ext));
    S.W = int(keep.size());
    S.eqs.assign(size_t(S.W), -1);
    std::vector<int> map(size_t(S.W), 0);
    for (size_t q = 0; q < keep.size(); ++q) {
        S.eqs[q] = A.eqs[size_t(keep[q])];
        map[q] = keep[q];
    }
    MI_TRY(upload_eqs(B, S));
    DevBuf<int> d_map;
    MI_TRY(d_map.upload(map.data(), map.size(), B->stream));
    hipStream_t s = B->stream;
    auto cp = [&](int64_t rows, const auto& from, auto& to) {
        if (rows > 0)
            k_brepack<<<reduce_grid(rows * S.W, 4), kBlock, 0, s>>>(rows, A.W, S.W, d_map.p,
                                                                    from.p, to.p);
    };
    if (!B->shared) {
        cp(B->vrows, A.v32, S.v32);
        cp(B->n, A.invD, S.invD);
        if (B->pc == MI_PC_ILU)
            cp(B->vrows, A.lv, S.lv);
    }
    if (B->x64)
        cp(B->n, A.xs64, S.xs64);
    else
        cp(B->n, A.xs, S.xs);
    cp(B->n, A.r, S.r);
    cp(B->n, A.p, S.p);
    cp(B->n, A.Ap, S.Ap);
    if (ext) {
        cp(B->n, A.rh, S.rh);
        cp(B->n, A.y, S.y);
        cp(B->n, A.z, S.z);
        cp(B->n, A.t, S.t);
    }
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s)); // d_map is released on return
    B->cur ^= 1;
    return MI_OK;
}

// Folds the partials of a producing launch of G workgroups.
int fin(mi_batch* B, int op, const BFinArgs& f, int G, Launcher& L)
{
    const BSlots& S = B->set[B->cur];
    BFinArgs g = f;
    g.G = G;
    return L(k_bfin, unsigned(S.W), kBlock, static_cast<const double*>(B->partials.p),
             static_cast<const int*>(S.d_eq.p), B->st.p, op, g);
}

// The SA-AMG hierarchy of the shared matrix (host setup, the single-equation
// solver's parameters), uploaded as FP32 CSR operators.
int amg_setup_batch(mi_batch* B, const mi_solver_cfg& c)
{
    const Pattern& P = *B->pat;
    const int n = int(B->n);
    HCsr A;
    A.nr = A.nc = n;
    A.rp = P.h_rp;
    A.ci = P.h_ci;
    A.v.resize(size_t(B->nnz));
    if (B->nnz)
        MI_HIP(hipMemcpy(A.v.data(), B->v64.p, size_t(B->nnz) * sizeof(double),
                         hipMemcpyDeviceToHost));
    AmgParams p;
    if (c.amg_max_levels > 0)
        p.max_levels = c.amg_max_levels;
    if (c.amg_coarse_size > 0)
        p.coarse_size = c.amg_coarse_size;
    if (c.amg_strength >= 0)
        p.theta = c.amg_strength;
    if (c.amg_jacobi_omega > 0)
        p.omega_scale = c.amg_jacobi_omega;
    if (c.amg_truncation >= 0)
        p.truncation = c.amg_truncation;
    AmgHierarchyHost H;
    if (amg_build(A, p, H) != 0)
        return set_error(MI_ERR_UNSUPPORTED, "batched AMG: hierarchy construction failed");
    const int nl = int(H.levels.size());
    const int deg0 = c.amg_sweeps > 0 ? c.amg_sweeps : 2;
    const int degc = c.amg_coarse_sweeps > 0 ? c.amg_coarse_sweeps : deg0;
    hipStream_t s = B->stream;
    B->amg.clear();
    B->amg.resize(size_t(nl));
    // SELL-64 form of an operator (slot k of row i at soff[i / 64] + 64 k +
    // i % 64, CSR order within the row); square: kSellDiag on the diagonal
    // slots; diff: the FP32 difference form (the FP64 row sum in the
    // diagonal's place), for square operators whose rows all hold a diagonal
    auto up = [&](const HCsr& h, DevBuf<int>& soff, DevBuf<int>& eci, DevBuf<float>& v,
                  bool square, bool diff) -> int {
        const int nr = h.nr;
        const int nsl = (nr + kSellSlice - 1) / kSellSlice;
        std::vector<int> off(size_t(nsl) + 1, 0);
        for (int sl = 0; sl < nsl; ++sl) {
            int w = 0;
            for (int i = sl * kSellSlice; i < std::min(nr, (sl + 1) * kSellSlice); ++i)
                w = std::max(w, h.rp[size_t(i) + 1] - h.rp[size_t(i)]);
            off[size_t(sl) + 1] = off[size_t(sl)] + w * kSellSlice;
        }
        const size_t tot = std::max<size_t>(size_t(off[size_t(nsl)]), 1);
        std::vector<int> c(tot, -1);
        std::vector<float> f(tot, 0.f);
        for (int i = 0; i < nr; ++i) {
            const size_t base = size_t(off[size_t(i / kSellSlice)]) + size_t(i % kSellSlice);
            double sum = 0.0;
            size_t dpos = 0;
            bool has = false;
            for (int k = h.rp[size_t(i)], j = 0; k < h.rp[size_t(i) + 1]; ++k, ++j) {
                const size_t m = base + size_t(j) * kSellSlice;
                const int col = h.ci[size_t(k)];
                const bool dg = square && col == i;
                c[m] = dg ? (col | kSellDiag) : col;
                f[m] = float(h.v[size_t(k)]);
                sum += h.v[size_t(k)];
                if (dg) {
                    dpos = m;
                    has = true;
                }
            }
            if (diff && has)
                f[dpos] = float(sum);
        }
        MI_TRY(soff.upload(off.data(), off.size(), s));
        MI_TRY(eci.upload(c.data(), c.size(), s));
        MI_TRY(v.upload(f.data(), f.size(), s));
        MI_HIP(hipStreamSynchronize(s)); // the host arrays are released on return
        return MI_OK;
    };
    for (int l = 0; l < nl; ++l) {
        const AmgLevelHost& h = H.levels[size_t(l)];
        mi_batch::ALevel& d = B->amg[size_t(l)];
        d.n = h.A.nr;
        d.deg = l == 0 ? deg0 : degc;
        bool all_diag = true;
        for (int i = 0; i < h.A.nr && all_diag; ++i) {
            bool has = false;
            for (int k = h.A.rp[size_t(i)]; k < h.A.rp[size_t(i) + 1]; ++k)
                has = has || h.A.ci[size_t(k)] == i;
            all_diag = has;
        }
        d.diff = all_diag ? 1 : 0;
        // level 0 streams the batch's own SELL-64 operator when it has one;
        // the coarsest level is solved with its dense inverse
        if (l + 1 < nl && (l > 0 || !B->sell))
            MI_TRY(up(h.A, d.asoff, d.aeci, d.av, true, all_diag));
        if (l + 1 < nl) {
            MI_TRY(up(h.P, d.psoff, d.peci, d.pv, false, false));
            MI_TRY(up(h.R, d.rsoff, d.reci, d.rv, false, false));
        }
        std::vector<float> inv(size_t(std::max(d.n, 1)), 1.f);
        for (int i = 0; i < d.n; ++i)
            for (int k = h.A.rp[size_t(i)]; k < h.A.rp[size_t(i) + 1]; ++k)
                if (h.A.ci[size_t(k)] == i && h.A.v[size_t(k)] != 0.0)
                    inv[size_t(i)] = float(1.0 / h.A.v[size_t(k)]);
        MI_TRY(d.invD.upload(inv.data(), inv.size(), s));
        MI_HIP(hipStreamSynchronize(s));
        const size_t nw = size_t(std::max(d.n, 1)) * size_t(B->Wmax);
        for (DevBuf<float>* v : {&d.r, &d.d, &d.d2, &d.x, &d.b, &d.e})
            MI_TRY(v->alloc(nw));
        chebyshev(h.lmax, d.deg, 0.3, d.alpha, d.beta);
    }
    B->amg_nc = H.coarse_n;
    MI_TRY(B->amg_cinv.upload(H.coarse_inverse.data(),
                              std::max<size_t>(H.coarse_inverse.size(), 1), s));
    MI_HIP(hipStreamSynchronize(s));
    B->amg_valid = true;
    return MI_OK;
}

// out = V-cycle(in) for every slot of the layout of a: the launch sequence
// of DHierarchy::cycle.  Level 0 streams the batch's own SELL-64 operator
// (the hierarchy's level-0 matrix, in the same FP32 form) when it has one.
int amg_apply_batch(mi_batch* B, const BArgs& a, const float* in, float* out, Launcher& L)
{
    const int nl = int(B->amg.size());
    const int W = a.W;
    const int V = slots_per_thread(W, B->Vmax);
    // operators of few rows take one slot per thread (W threads a row): more
    // workgroups for their latency-bound launches (8 pressure right-hand
    // sides at 128^3: 68.9 -> 65.5 ms with levels 2 and 3 so; level 1's 451k
    // rows as well: 69.4 ms)
    constexpr int64_t kFewRows = 100000;
    auto vfor = [&](int64_t rows) { return rows < kFewRows ? 1 : V; };
    auto kern = [&](int64_t rows) { return bamg_rows_kernel(vfor(rows)); };
    auto rows_grid = [&](int64_t rows) {
        const int R = kBlock / (W / vfor(rows));
        return unsigned(std::max<int64_t>(1, std::min<int64_t>((rows + R - 1) / R, 16384)));
    };
    auto vec_grid = [&](int64_t rows) {
        return unsigned(std::max<int64_t>(1, std::min<int64_t>((rows * W + kBlock - 1) / kBlock,
                                                               16384)));
    };
    auto sell = [&](const DevBuf<int>& soff, const DevBuf<int>& eci, const DevBuf<float>& v,
                    int nr, int diff) { return BSell{nr, diff, 0, soff.p, eci.p, v.p}; };
    auto bptr = [&](int l) -> const float* { return l == 0 ? in : B->amg[size_t(l)].b.p; };
    auto xptr = [&](int l) -> float* { return l == 0 ? out : B->amg[size_t(l)].x.p; };
    float* fnull = nullptr;
    const float* cnull = nullptr;
    // r -= A din, and for modes 0, 4, 5 the Chebyshev update, on level l's
    // operator (level 0: the batch's own, with the SpMM's grid and row walk)
    auto smooth = [&](int l, int mode, const float* din, float* dout, float* x, double beta,
                      double alpha, double a0 = 0.0) -> int {
        mi_batch::ALevel& d = B->amg[size_t(l)];
        const bool own = l == 0 && B->sell;
        const BSell A = own ? BSell{d.n, a.diff, a.band, a.soff, a.eci, a.v32}
                            : sell(d.asoff, d.aeci, d.av, d.n, d.diff);
        return L(own ? bamg_rows_kernel(V) : kern(d.n), own ? unsigned(B->G) : rows_grid(d.n),
                 kBlock, W, A, mode, din, dout, d.r.p,
                 static_cast<const float*>(d.invD.p), x, fnull, float(beta), float(alpha),
                 float(a0));
    };
    for (int l = 0; l + 1 < nl; ++l) {
        mi_batch::ALevel& d = B->amg[size_t(l)];
        float* x = xptr(l);
        // pre-smoothing from x = 0; on level 0 its start is fused with the
        // first product
        if (l > 0) {
            MI_TRY(L(k_bamg_init, vec_grid(d.n), kBlock, int64_t(d.n), W, bptr(l),
                     static_cast<const float*>(d.invD.p), float(d.alpha[0]), x, d.r.p, d.d.p));
            float* din = d.d.p;
            float* dout = d.d2.p;
            for (int k = 1; k < d.deg; ++k) {
                MI_TRY(smooth(l, 0, din, dout, x, d.beta[size_t(k)], d.alpha[size_t(k)]));
                std::swap(din, dout);
            }
            MI_TRY(smooth(l, 1, din, nullptr, x, 0.0, 0.0));
        } else if (d.deg == 1) {
            MI_TRY(smooth(l, 6, bptr(l), fnull, x, 0.0, 0.0, d.alpha[0]));
        } else {
            MI_TRY(smooth(l, 5, bptr(l), d.d.p, x, d.beta[1], d.alpha[1], d.alpha[0]));
            float* din = d.d.p;
            float* dout = d.d2.p;
            for (int k = 2; k < d.deg; ++k) {
                MI_TRY(smooth(l, 0, din, dout, x, d.beta[size_t(k)], d.alpha[size_t(k)]));
                std::swap(din, dout);
            }
            MI_TRY(smooth(l, 1, din, nullptr, x, 0.0, 0.0));
        }
        const int nc = B->amg[size_t(l) + 1].n;
        MI_TRY(L(kern(nc), rows_grid(nc), kBlock, W, sell(d.rsoff, d.reci, d.rv, nc, 0), 2,
                 static_cast<const float*>(d.r.p), fnull, fnull, cnull, fnull,
                 B->amg[size_t(l) + 1].b.p, 0.f, 0.f, 0.f));
    }
    if (B->amg_nc > 0)
        MI_TRY(L(k_bamg_coarse, unsigned(B->amg_nc), kBlock, B->amg_nc, W,
                 static_cast<const double*>(B->amg_cinv.p), bptr(nl - 1), xptr(nl - 1)));
    for (int l = nl - 2; l >= 0; --l) {
        mi_batch::ALevel& d = B->amg[size_t(l)];
        float* x = xptr(l);
        MI_TRY(L(kern(d.n), rows_grid(d.n), kBlock, W, sell(d.psoff, d.peci, d.pv, d.n, 0), 3,
                 static_cast<const float*>(xptr(l + 1)), fnull, fnull, cnull, x, d.e.p, 0.f,
                 0.f, 0.f));
        // post-smoothing; its last step stores only x
        MI_TRY(smooth(l, d.deg == 1 ? 4 : 0, d.e.p, d.d.p, x, 0.0, d.alpha[0]));
        float* din = d.d.p;
        float* dout = d.d2.p;
        for (int k = 1; k < d.deg; ++k) {
            MI_TRY(smooth(l, k + 1 == d.deg ? 4 : 0, din, dout, x, d.beta[size_t(k)],
                          d.alpha[size_t(k)]));
            std::swap(din, dout);
        }
    }
    return MI_OK;
}

// z = M^-1 in for the iterating slots: the two sweeps of the ILU(0) apply.
int ilu_apply(mi_batch* B, const BArgs& a, const float* in, float* out, Launcher& L)
{
    const BIluKernel k = bilu_kernel(B->sell, B->shared, a.W, B->Vmax);
    MI_TRY(L(k, B->G, kBlock, a, in, out, 0, B->n0));
    return L(k, B->G, kBlock, a, in, out, 1, B->n0);
}

// A preconditioner applied by launches of its own (ILU(0), the AMG V-cycle).
bool separate_pc(int pc) { return pc == MI_PC_ILU || pc == MI_PC_AMG; }
int separate_apply(mi_batch* B, const BArgs& a, const float* in, float* out, Launcher& L)
{
    if (B->pc == MI_PC_AMG)
        return amg_apply_batch(B, a, in, out, L);
    return ilu_apply(B, a, in, out, L);
}

// One CG iteration for every iterating equation: five launches with Jacobi,
// eight with ILU(0), the V-cycle's and seven with AMG.
int one_iteration(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    MI_TRY(L(bspmm_kernel<B_ALPHA>(B->sell, B->shared, a.W, B->Vmax), B->G, kBlock, a));
    MI_TRY(fin(B, B_ALPHA, f, B->G, L));
    if (separate_pc(B->pc)) {
        MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_UPD_X)));
        MI_TRY(separate_apply(B, a, a.r, a.z, L));
        MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_DOT_RZ)));
        MI_TRY(fin(B, B_UPDATE, f, B->Gv, L));
        return L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_PZ));
    }
    MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_UPDATE)));
    MI_TRY(fin(B, B_UPDATE, f, B->Gv, L));
    return L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_P));
}

// One BiCGStab iteration for every iterating equation: nine launches with
// Jacobi, twelve with ILU(0).
int one_iteration_bicg(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    const bool ilu = separate_pc(B->pc);
    MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(ilu ? B_BI_PN : B_BI_P)));
    if (ilu)
        MI_TRY(separate_apply(B, a, a.p, a.y, L));
    MI_TRY(L(bspmm_kernel<B_BI_ALPHA>(B->sell, B->shared, a.W, B->Vmax), B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_ALPHA, f, B->G, L));
    MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(ilu ? B_BI_SN : B_BI_S)));
    MI_TRY(fin(B, B_BI_S, f, B->Gv, L));
    if (ilu)
        MI_TRY(separate_apply(B, a, a.r, a.z, L));
    MI_TRY(L(bspmm_kernel<B_BI_T>(B->sell, B->shared, a.W, B->Vmax), B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_OMEGA, f, B->G, L));
    MI_TRY(L(bvec_kernel(a.W), B->Gv, kBlock, a, int(B_BI_UPDATE)));
    return fin(B, B_BI_UPDATE, f, B->Gv, L);
}

// Autotuner, timed on this pattern in the widest layout with every
// equation's slot active: the SpMM's tile shape (slots per thread V, which
// sets the rows of a workgroup pass, 256 V / W), its row walk (banded or
// contiguous) and its workgroup count (rows per workgroup); the ILU(0)
// sweeps use the same shape.
int autotune(mi_batch* B)
{
    B->G = 8;
    B->Vmax = 4;
    // vector stages: one workgroup per 256 rows, up to kBatchMaxGrid, before tuning
    B->Gv = int(std::min<int64_t>(kBatchMaxGrid, std::max<int64_t>(8, (B->n + 2047) / 2048 * 8)));
    if (B->n == 0)
        return MI_OK;
    std::vector<int> all(size_t(B->E));
    for (int e = 0; e < B->E; ++e)
        all[size_t(e)] = e;
    MI_TRY(pack_from_home(B, all, false));
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    auto time_us = [&](auto&& launch) -> double {
        if (launch() != MI_OK)
            return -1.0;
        (void)hipEventRecord(e0, B->stream);
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
            if (launch() != MI_OK)
                return -1.0;
        (void)hipEventRecord(e1, B->stream);
        if (hipEventSynchronize(e1) != hipSuccess)
            return -1.0;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return 1000.0 * ms / reps;
    };
    double best = 1e300;
    BArgs a = B->args(B->cur);
    a.st = nullptr;
    const int W = a.W;
    int last_v = 0;
    for (int vmax : {4, 2, 1}) {
        const int v = slots_per_thread(W, vmax);
        if (v == last_v)
            continue; // the same kernel as the previous cap
        last_v = v;
        auto k = bspmm_kernel<B_ALPHA>(B->sell, B->shared, W, vmax);
        for (int band : {0, 1})
            for (int g : {256, 512, 1024, 2048, 4096}) {
                if (g > kBatchMaxGrid)
                    break;
                a.band = band;
                Launcher L{B->stream};
                const double us = time_us([&] { return L(k, unsigned(g), kBlock, a); });
                if (us < 0.0)
                    return set_error(MI_ERR_DEVICE, "batch autotuner: SpMM launch failed");
                if (us < best) {
                    best = us;
                    B->G = g;
                    B->Vmax = vmax;
                    B->band = band;
                }
            }
    }
    B->tune_us = best;
    // vector stages: one workgroup per 256 rows up to 2048 (whole 32-equation
    // solves on 1M rows: 2048 workgroups 506 ms, 3912 518 ms; single launches
    // of the CG update time alike on either grid, so they are not tuned)
    B->Gv = std::min(B->Gv, 2048);
    {
        BArgs av = B->args(B->cur);
        Launcher L{B->stream};
        B->tune_vec_us = time_us([&] { return L(bvec_kernel(av.W), unsigned(B->Gv), kBlock, av,
                                                int(B_UPDATE)); });
        if (B->tune_vec_us < 0.0)
            return set_error(MI_ERR_DEVICE, "batch autotuner: vector launch failed");
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->slots_valid = false; // packed from values not yet set
    return MI_OK;
}

// The workspace of batched GMRES(m).
int gm_alloc(mi_batch* B, int m)
{
    if (B->gm_m == m)
        return MI_OK;
    B->gmV = DevBuf<float>();
    B->gmSmall = DevBuf<double>();
    B->gmPart = DevBuf<double>();
    const size_t nW = size_t(std::max<int64_t>(B->n, 1)) * size_t(B->Wmax);
    MI_TRY(B->gmV.alloc(size_t(m + 1) * nW));
    MI_TRY(B->gmSmall.alloc(size_t(B->E) * size_t(gm_small_stride(m))));
    MI_TRY(B->gmPart.alloc(size_t(B->Wmax) * size_t(m + 1) * size_t(kBatchMaxGrid)));
    B->gm_m = m;
    return MI_OK;
}

BGm gm_args(const mi_batch* B, int W)
{
    BGm g{};
    g.V = B->gmV.p;
    g.vs = size_t(B->n) * size_t(W);
    g.m = B->gm_m;
    g.small = B->gmSmall.p;
    g.ss = gm_small_stride(B->gm_m);
    g.part = B->gmPart.p;
    return g;
}

// z = M^-1 in for the slots of a running GMRES cycle.
int gm_precond(mi_batch* B, const BArgs& a, const float* in, float* out, Launcher& L)
{
    if (separate_pc(B->pc))
        return separate_apply(B, a, in, out, L);
    return L(k_bgm_jacobi, unsigned(B->Gv), kBlock, a, in, out);
}

// One GMRES(m) cycle for every equation of the sweep, from r (FP32 residual
// of the sweep) to xs += V y.
int gm_sweep(mi_batch* B, const BArgs& a0, int max_iter, Launcher& L)
{
    const BGm g = gm_args(B, a0.W);
    const unsigned Gv = unsigned(B->Gv);
    const int m = g.m;
    MI_TRY(gm_precond(B, a0, a0.r, g.V, L));
    MI_TRY(L(k_bgm_orth, Gv, kBlock, a0, g, 0, 0, 1, 1));
    MI_TRY(L(k_bgm_begin, unsigned(a0.W), kBlock, a0, g, int(Gv)));
    MI_TRY(L(k_bgm_scale, Gv, kBlock, a0, g, 0));
    for (int j = 0; j < m; ++j) {
        BArgs aj = a0;
        aj.p = g.V + size_t(j) * g.vs; // w = A v_j into Ap
        float* vn = g.V + size_t(j + 1) * g.vs;
        MI_TRY(L(bspmm_kernel<B_GM>(B->sell, B->shared, a0.W, B->Vmax), B->G, kBlock, aj));
        MI_TRY(gm_precond(B, aj, aj.Ap, vn, L));
        const unsigned fb = unsigned(a0.W * (j + 1));
        MI_TRY(L(k_bgm_dots, Gv, kBlock, a0, g, j + 1, j + 1));
        MI_TRY(L(k_bgm_fold, fb, kBlock, a0, g, j + 1, 1, int(Gv)));
        MI_TRY(L(k_bgm_orth, Gv, kBlock, a0, g, j + 1, j + 1, 1, 0));
        MI_TRY(L(k_bgm_fold, fb, kBlock, a0, g, j + 1, 2, int(Gv)));
        MI_TRY(L(k_bgm_orth, Gv, kBlock, a0, g, j + 1, j + 1, 2, 1));
        MI_TRY(L(k_bgm_step, unsigned(a0.W), kBlock, a0, g, j, int(Gv), max_iter));
        MI_TRY(L(k_bgm_scale, Gv, kBlock, a0, g, j + 1));
    }
    MI_TRY(L(k_bgm_backsolve, 1u, unsigned(kWave), a0, g));
    return L(k_bgm_xupdate, Gv, kBlock, a0, g);
}

int check_cfg(const mi_batch* B, const mi_solver_cfg* c)
{
    if (c->kind == MI_SOLVER_GMRES) {
        if (c->restart > kBGmMax)
            return set_error(MI_ERR_UNSUPPORTED, "batched GMRES: restart %d above %d", c->restart,
                             kBGmMax);
    } else if (c->kind != MI_SOLVER_CG && c->kind != MI_SOLVER_BICGSTAB) {
        return set_error(MI_ERR_UNSUPPORTED,
                         "batched solve: CG, BiCGStab or GMRES (kind %d given)", c->kind);
    }
    if (c->preconditioner == MI_PC_ILU) {
        if (!B->ilu_order)
            return set_error(MI_ERR_UNSUPPORTED,
                             "batched solve: ILU(0) needs a batch created with MI_BATCH_ILU0");
    } else if (c->preconditioner == MI_PC_AMG) {
        if (!B->shared)
            return set_error(MI_ERR_UNSUPPORTED,
                             "batched solve: AMG needs one shared matrix (MI_BATCH_SHARED_VALUES)");
    } else if (c->preconditioner != MI_PC_JACOBI && c->preconditioner != MI_PC_NONE) {
        return set_error(MI_ERR_UNSUPPORTED,
                         "batched solve: Jacobi, ILU(0), AMG or no preconditioner (%d given)",
                         c->preconditioner);
    }
    if (!(c->tolerance > 0.0) || c->max_iterations < 0)
        return set_error(MI_ERR_BAD_ARG, "bad tolerance or iteration limit");
    return MI_OK;
}

int read_state(mi_batch* B)
{
    MI_HIP(hipMemcpyAsync(B->h_st, B->st.p, sizeof(BEq) * size_t(B->E), hipMemcpyDeviceToHost,
                          B->stream));
    MI_HIP(hipStreamSynchronize(B->stream));
    return MI_OK;
}

// ILU(0) factors of every held matrix from its FP64 values.
int factor_ilu(mi_batch* B)
{
    if (B->n == 0) {
        B->ilu_valid = true;
        return MI_OK;
    }
    const Pattern& P = *B->pat;
    const int64_t work = B->n * int64_t(B->Em);
    k_bilu_factor<<<grid_for(work, kBlock), kBlock, 0, B->stream>>>(
        int(B->n), B->n0, B->Em, B->shared ? 0 : B->nnz, B->shared ? 0 : B->vrows, P.rp.p, P.ci.p,
        P.diag.p, B->d_tpos.p, B->d_vpos.p, B->v64.p, B->lvh.p, B->ipivh.p);
    MI_HIP(hipGetLastError());
    B->ilu_valid = true;
    return MI_OK;
}

int solve(mi_batch* B, const mi_solver_cfg& c, const double* b_dev, double* x_dev,
          int* iterations, double* rel)
{
    const int E = B->E;
    const int64_t n = B->n;
    hipStream_t s = B->stream;
    if (n == 0) {
        for (int e = 0; e < E; ++e) {
            if (iterations)
                iterations[e] = 0;
            if (rel)
                rel[e] = 0.0;
        }
        return MI_OK;
    }
    const bool bicg = c.kind == MI_SOLVER_BICGSTAB;
    const bool gmres = c.kind == MI_SOLVER_GMRES;
    const int gm_m = std::max(1, c.restart);
    const int pc = c.preconditioner;
    if (pc != B->pc) {
        B->pc = pc;
        B->slots_valid = false;
    }
    if (pc == MI_PC_ILU && !B->ilu_valid) {
        MI_TRY(factor_ilu(B));
        B->slots_valid = false;
    }
    if (pc == MI_PC_AMG && !B->amg_valid) {
        if (B->graph) { // its launches hold the previous hierarchy's buffers
            (void)hipGraphExecDestroy(B->graph);
            B->graph = nullptr;
        }
        MI_TRY(amg_setup_batch(B, c));
    }
    const bool ext = bicg || separate_pc(pc);
    {
        double m = 1.0;
        for (double v : B->margin)
            m = std::min(m, v);
        const bool x64 = m < kX64Margin;
        if (x64 != B->x64) {
            B->x64 = x64;
            if (B->graph) { // its launches hold the other correction
                (void)hipGraphExecDestroy(B->graph);
                B->graph = nullptr;
            }
            for (BSlots& S : B->set)
                MI_TRY(alloc_slots(B, S, false));
        }
    }
    if (gmres)
        MI_TRY(gm_alloc(B, gm_m));
    if (ext && !B->ext_alloc) {
        MI_TRY(alloc_slots(B, B->set[0], true));
        MI_TRY(alloc_slots(B, B->set[1], true));
        B->ext_alloc = true;
    }
    BFinArgs f{};
    f.E = E;
    f.G = B->Gv;
    f.max_iter = c.max_iterations;
    // a GMRES sweep is one cycle of at most m steps: by default as many
    // sweeps as the iteration limit allows
    f.max_outer = c.max_outer > 0 ? c.max_outer
                  : gmres        ? std::max(20, c.max_iterations / gm_m + 2)
                                 : 20;
    f.tol = c.tolerance;
    f.inner_tol = c.inner_tolerance > 0 ? c.inner_tolerance : 1e-5;
    const size_t m = size_t(n) * size_t(E);
    if (B->reordered) {
        k_bpermute<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(n, E, B->d_perm.p, b_dev, B->b.p, 0);
        MI_HIP(hipGetLastError());
    } else {
        MI_HIP(hipMemcpyAsync(B->b.p, b_dev, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    k_bfill<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(int64_t(m), B->x.p);
    MI_HIP(hipGetLastError());
    const HArgs h = B->hargs();
    k_bnorm<<<B->Gv, kBlock, 0, s>>>(h);
    k_bfin_eq<<<E, kBlock, 0, s>>>(B->partials.p, B->st.p, B_BNORM, f);
    MI_HIP(hipGetLastError());
    const int chunk = c.check_interval > 0 ? c.check_interval : 8;
    const bool use_graph = c.use_graph != 0 && graphs_enabled();
    for (;;) {
        // FP64 true residual of the unfinished equations; which need a sweep
        k_btrue<<<B->Gv, kBlock, 0, s>>>(h);
        k_bfin_eq<<<E, kBlock, 0, s>>>(B->partials.p, B->st.p, B_TRUE, f);
        MI_HIP(hipGetLastError());
        MI_TRY(read_state(B));
        std::vector<int> want;
        for (int e = 0; e < E; ++e)
            if (!B->h_st[e].fin)
                want.push_back(e);
        if (want.empty())
            break;
        {
            std::vector<int> have(B->set[B->cur].eqs);
            have.erase(std::remove(have.begin(), have.end(), -1), have.end());
            if (!B->slots_valid || have != want)
                MI_TRY(pack_from_home(B, want, ext));
        }
        // r = float(rt) and xs = 0 for this sweep, then the first stage
        {
            BSlots& S = B->set[B->cur];
            k_btile<double><<<tile_grid(n), kBlock, 0, s>>>(n, S.W, S.d_eq.p, 0ull, 0, 0, 0.f,
                                                            B->rt.p, S.r.p);
            if (B->x64)
                k_bfill<<<reduce_grid(n * S.W, 4), kBlock, 0, s>>>(n * S.W, S.xs64.p);
            else
                k_bzero32<<<reduce_grid(n * S.W, 4), kBlock, 0, s>>>(n * S.W, S.xs.p);
            MI_HIP(hipGetLastError());
            Launcher L{s};
            f.W = S.W;
            const BArgs a0 = B->args(B->cur);
            if (gmres) {
                // the sweep is one GMRES(m) cycle, one graph
                if (use_graph) {
                    if (!B->graph || B->graph_chunk != gm_m || B->graph_kind != c.kind ||
                        B->graph_pc != pc || B->graph_set != B->cur || B->graph_W != a0.W ||
                        std::memcmp(&B->graph_f, &f, sizeof(BFinArgs)) != 0) {
                        if (B->graph) {
                            (void)hipGraphExecDestroy(B->graph);
                            B->graph = nullptr;
                        }
                        MI_TRY(build_graph(s, [&](Launcher& Lg) -> int {
                            return gm_sweep(B, a0, c.max_iterations, Lg);
                        }, &B->graph));
                        B->graph_chunk = gm_m;
                        B->graph_kind = c.kind;
                        B->graph_pc = pc;
                        B->graph_set = B->cur;
                        B->graph_W = a0.W;
                        B->graph_f = f;
                    }
                    MI_HIP(hipGraphLaunch(B->graph, s));
                } else {
                    MI_TRY(gm_sweep(B, a0, c.max_iterations, L));
                }
                MI_TRY(retire(B, S.W >= 64 ? ~0ull : ((1ull << S.W) - 1)));
                continue;
            }
            if (bicg) {
                MI_TRY(L(bvec_kernel(a0.W), B->Gv, kBlock, a0, int(B_BI_INIT)));
                MI_TRY(fin(B, B_BI_INIT, f, B->Gv, L));
            } else if (separate_pc(pc)) {
                MI_TRY(separate_apply(B, a0, a0.r, a0.z, L));
                MI_TRY(L(bvec_kernel(a0.W), B->Gv, kBlock, a0, int(B_INIT_Z)));
                MI_TRY(fin(B, B_INIT, f, B->Gv, L));
            } else {
                MI_TRY(L(bvec_kernel(a0.W), B->Gv, kBlock, a0, int(B_INIT)));
                MI_TRY(fin(B, B_INIT, f, B->Gv, L));
            }
        }
        for (;;) {
            const BArgs a = B->args(B->cur);
            f.W = a.W;
            auto iteration = [&](Launcher& L) {
                return bicg ? one_iteration_bicg(B, a, f, L) : one_iteration(B, a, f, L);
            };
            if (use_graph) {
                if (!B->graph || B->graph_chunk != chunk || B->graph_kind != c.kind ||
                    B->graph_pc != pc || B->graph_set != B->cur || B->graph_W != a.W ||
                    std::memcmp(&B->graph_f, &f, sizeof(BFinArgs)) != 0) {
                    if (B->graph) {
                        (void)hipGraphExecDestroy(B->graph);
                        B->graph = nullptr;
                    }
                    MI_TRY(build_graph(s, [&](Launcher& L) -> int {
                        for (int it = 0; it < chunk; ++it)
                            MI_TRY(iteration(L));
                        return MI_OK;
                    }, &B->graph));
                    B->graph_chunk = chunk;
                    B->graph_kind = c.kind;
                    B->graph_pc = pc;
                    B->graph_set = B->cur;
                    B->graph_W = a.W;
                    B->graph_f = f;
                }
                MI_HIP(hipGraphLaunch(B->graph, s));
            } else {
                Launcher L{s};
                for (int it = 0; it < chunk; ++it)
                    MI_TRY(iteration(L));
            }
            MI_TRY(read_state(B));
            const BSlots& S = B->set[B->cur];
            std::vector<int> live;
            uint64_t idle = 0;
            for (int q = 0; q < S.W; ++q) {
                const int e = S.eqs[size_t(q)];
                if (e < 0)
                    continue;
                if (B->h_st[e].done)
                    idle |= 1ull << q;
                else
                    live.push_back(q);
            }
            if (live.empty()) {
                MI_TRY(retire(B, idle));
                break;
            }
            if (int(live.size()) < S.W) {
                // idle slots: their corrections go home and the live slots
                // move to a layout as wide as their number
                MI_TRY(retire(B, idle));
                MI_TRY(repack(B, live, ext));
                B->slots_valid = false; // no longer the home values of `want`
            }
        }
    }
    int status = MI_OK;
    for (int e = 0; e < E; ++e) {
        if (iterations)
            iterations[e] = B->h_st[e].total;
        if (rel)
            rel[e] = B->h_st[e].rel;
        if (B->h_st[e].status)
            status = MI_ERR_NOT_CONVERGED;
    }
    if (B->reordered) {
        k_bpermute<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(n, E, B->d_perm.p, B->x.p, x_dev, 1);
        MI_HIP(hipGetLastError());
    } else {
        MI_HIP(hipMemcpyAsync(x_dev, B->x.p, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    MI_HIP(hipStreamSynchronize(s));
    if (status != MI_OK)
        set_error(MI_ERR_NOT_CONVERGED, "batched solve: some equations did not converge");
    return status;
}

// The two-colour ordering of a pattern (rows of the first colour first, RCM
// order within each colour), the first row of the second colour, and for
// every entry of the reordered pattern the position of its transposed entry.
// Refused unless the pattern is structurally symmetric, holds every
// diagonal and has two colours.
int two_colour_order(const Pattern& src, std::vector<int>& perm, std::vector<int>& brp,
                     std::vector<int>& bci, std::vector<int>& vmap, std::vector<int>& tpos,
                     int& n0)
{
    const int n = int(src.n);
    int colors = 0;
    color_order(n, src.h_rp.data(), src.h_ci.data(), perm, &colors);
    if (colors > 2)
        return set_error(MI_ERR_UNSUPPORTED,
                         "batched ILU(0): the pattern needs %d colours (two supported)", colors);
    permute_pattern(n, src.h_rp.data(), src.h_ci.data(), perm, brp, bci, vmap);
    n0 = n;
    for (int i = 0; i < n && n0 == n; ++i)
        for (int k = brp[size_t(i)]; k < brp[size_t(i) + 1]; ++k)
            if (bci[size_t(k)] < i) {
                n0 = i;
                break;
            }
    tpos.assign(bci.size(), -1);
    for (int i = 0; i < n; ++i) {
        bool has_diag = false;
        for (int k = brp[size_t(i)]; k < brp[size_t(i) + 1]; ++k) {
            const int c = bci[size_t(k)];
            if (c == i) {
                has_diag = true;
                tpos[size_t(k)] = k;
                continue;
            }
            if ((i < n0) == (c < n0))
                return set_error(MI_ERR_UNSUPPORTED,
                                 "batched ILU(0): rows %d and %d of one colour are coupled", i, c);
            const int* b0 = bci.data() + brp[size_t(c)];
            const int* b1 = bci.data() + brp[size_t(c) + 1];
            const int* f = std::lower_bound(b0, b1, i);
            if (f == b1 || *f != i)
                return set_error(MI_ERR_UNSUPPORTED,
                                 "batched ILU(0): the pattern is not symmetric (%d, %d)", i, c);
            tpos[size_t(k)] = int(f - bci.data());
        }
        if (!has_diag)
            return set_error(MI_ERR_UNSUPPORTED, "batched ILU(0): row %d has no diagonal", i);
    }
    return MI_OK;
}

// Value line of each CSR entry in the working layout: its SELL-64 slot.
__global__ __launch_bounds__(kBlock) void k_bslot_of(int64_t m, const int* __restrict__ eperm,
                                                     int* __restrict__ vpos)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
        const int k = eperm[i];
        if (k >= 0)
            vpos[k] = int(i);
    }
}

// Values of matrix slot e (0 when shared) from the caller's FP64 values.
int write_matrix(mi_batch* B, int e, const mi_csr* A)
{
    hipStream_t s = B->stream;
    double* v64 = B->v64.p + size_t(e) * size_t(B->nnz);
    float* v32 = B->v32h.p + size_t(e) * size_t(B->vrows);
    if (B->nnz)
        k_bconvert<<<reduce_grid(B->nnz, 4), kBlock, 0, s>>>(
            B->nnz, B->reordered ? B->d_vmap.p : nullptr, A->val.p, v64, B->sell ? nullptr : v32);
    if (B->sell && B->vrows)
        k_bsell32<<<reduce_grid(B->vrows, 4), kBlock, 0, s>>>(B->vrows, B->pat->sell.eperm.p, v64,
                                                              v32);
    if (B->n && B->diff)
        k_bdiffdiag<<<grid_for(B->n, kBlock), kBlock, 0, s>>>(
            int(B->n), B->pat->rp.p, B->pat->diag.p, v64, B->sell ? B->d_vpos.p : nullptr, v32);
    if (B->n)
        k_binvdiag<<<grid_for(B->n, kBlock), kBlock, 0, s>>>(
            int(B->n), B->pat->diag.p, v64, B->invDh.p + size_t(e) * size_t(B->n));
    constexpr int kMarginBlocks = 256;
    double part[kMarginBlocks];
    if (B->n)
        k_bmargin<<<kMarginBlocks, kBlock, 0, s>>>(int(B->n), B->pat->rp.p, B->pat->ci.p, v64,
                                                   B->partials.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipMemcpyAsync(part, B->partials.p, sizeof(part), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    double m = 0.0;
    for (int k = 0; B->n && k < kMarginBlocks; ++k)
        m += part[k];
    B->margin[size_t(e)] = B->n ? m / double(B->n) : 1.0;
    B->slots_valid = false;
    B->ilu_valid = false;
    B->amg_valid = false;
    return MI_OK;
}

bool same_pattern(const mi_batch* B, const mi_csr* A)
{
    return A->pat == B->src || (A->n() == B->n && A->nnz() == B->nnz &&
                                A->pat->h_rp == B->src->h_rp && A->pat->h_ci == B->src->h_ci);
}

} // namespace

extern "C" {

MI_API int mi_batch_create_ex(const mi_csr* A, int n_eq, int flags, mi_batch** out)
{
    if (!A || !out || n_eq < 1 || n_eq > kMaxEq)
        return set_error(MI_ERR_BAD_ARG, "need a pattern matrix and 1..%d equations", kMaxEq);
    if (flags & ~(MI_BATCH_SHARED_VALUES | MI_BATCH_ILU0))
        return set_error(MI_ERR_BAD_ARG, "unknown batch flags 0x%x", flags);
    MI_TRY(ensure_device());
    auto B = std::make_unique<mi_batch>();
    B->src = A->pat;
    B->pat = A->pat;
    B->E = n_eq;
    B->Wmax = n_eq;
    B->shared = (flags & MI_BATCH_SHARED_VALUES) != 0;
    B->Em = B->shared ? 1 : n_eq;
    B->margin.assign(size_t(B->Em), -1.0); // no values yet
    B->ilu_order = (flags & MI_BATCH_ILU0) != 0;
    B->n = A->n();
    B->nnz = A->nnz();
    MI_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    const Pattern& src = *A->pat;
    std::vector<int> tpos;
    if (B->ilu_order || B->n >= 100000) {
        // two colours for ILU(0), else the single-equation solver's bandwidth reduction
        std::vector<int> perm, brp, bci, vmap;
        if (B->ilu_order) {
            MI_TRY(two_colour_order(src, perm, brp, bci, vmap, tpos, B->n0));
        } else {
            rcm_order(int(src.n), src.h_rp.data(), src.h_ci.data(), perm);
            permute_pattern(int(src.n), src.h_rp.data(), src.h_ci.data(), perm, brp, bci, vmap);
        }
        MI_TRY(make_pattern(src.n, src.nnz, brp.data(), bci.data(), B->pat, B->stream));
        MI_TRY(B->d_perm.upload(perm.data(), perm.size(), B->stream));
        MI_TRY(B->d_vmap.upload(vmap.data(), vmap.size(), B->stream));
        MI_HIP(hipStreamSynchronize(B->stream));
        B->reordered = true;
    }
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&B->h_st), sizeof(BEq) * kMaxEq, 0));
    B->sell = B->pat->sell.built() && B->pat->sell.eperm.p != nullptr;
    B->vrows = B->sell ? B->pat->sell.padded : B->nnz;
    B->diff = B->pat->all_diag;
    if (B->sell && (B->diff || B->ilu_order)) {
        // the SELL-64 slot of every CSR entry
        MI_TRY(B->d_vpos.alloc(size_t(std::max<int64_t>(B->nnz, 1))));
        k_bslot_of<<<reduce_grid(B->vrows, 4), kBlock, 0, B->stream>>>(
            B->vrows, B->pat->sell.eperm.p, B->d_vpos.p);
        MI_HIP(hipGetLastError());
    }
    const size_t nE = size_t(std::max<int64_t>(B->n, 1)) * size_t(n_eq);
    const size_t nM = size_t(std::max<int64_t>(B->n, 1)) * size_t(B->Em);
    const size_t zM = size_t(std::max<int64_t>(B->nnz, 1)) * size_t(B->Em);
    const size_t vM = size_t(std::max<int64_t>(B->vrows, 1)) * size_t(B->Em);
    MI_TRY(B->v64.alloc(zM));
    MI_TRY(B->v32h.alloc(vM));
    MI_TRY(B->invDh.alloc(nM));
    for (DevBuf<double>* v : {&B->b, &B->x, &B->rt})
        MI_TRY(v->alloc(nE));
    MI_TRY(B->partials.alloc(size_t(2) * kBatchMaxGrid * size_t(kMaxEq)));
    MI_TRY(B->st.alloc(size_t(n_eq)));
    MI_HIP(hipMemsetAsync(B->v32h.p, 0, vM * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->v64.p, 0, zM * sizeof(double), B->stream));
    MI_HIP(hipMemsetAsync(B->invDh.p, 0, nM * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->st.p, 0, sizeof(BEq) * size_t(n_eq), B->stream));
    if (B->shared) {
        std::vector<float> one(nM / size_t(B->Em), 1.f);
        MI_TRY(B->ones.upload(one.data(), one.size(), B->stream));
    }
    if (B->ilu_order) {
        MI_TRY(B->lvh.alloc(vM));
        MI_TRY(B->ipivh.alloc(nM));
        MI_HIP(hipMemsetAsync(B->lvh.p, 0, vM * sizeof(float), B->stream));
        MI_HIP(hipMemsetAsync(B->ipivh.p, 0, nM * sizeof(float), B->stream));
        MI_TRY(B->d_tpos.upload(tpos.data(), std::max<size_t>(tpos.size(), 1), B->stream));
        if (!B->sell) {
            std::vector<int> id(size_t(std::max<int64_t>(B->nnz, 1)));
            for (size_t k = 0; k < id.size(); ++k)
                id[k] = int(k);
            MI_TRY(B->d_vpos.upload(id.data(), id.size(), B->stream));
        }
    }
    MI_HIP(hipStreamSynchronize(B->stream));
    MI_TRY(alloc_slots(B.get(), B->set[0], false));
    MI_TRY(alloc_slots(B.get(), B->set[1], false));
    for (BSlots& S : B->set)
        for (DevBuf<float>* v : {&S.p, &S.Ap, &S.xs, &S.r})
            MI_HIP(hipMemsetAsync(v->p, 0, v->n * sizeof(float), B->stream));
    MI_HIP(hipStreamSynchronize(B->stream));
    MI_TRY(autotune(B.get()));
    *out = B.release();
    return MI_OK;
}

MI_API int mi_batch_create(const mi_csr* A, int n_eq, mi_batch** out)
{
    return mi_batch_create_ex(A, n_eq, 0, out);
}

MI_API int mi_batch_set_equation(mi_batch* B, int eq, const mi_csr* A)
{
    if (!B || !A || eq < 0 || eq >= B->E)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (B->shared)
        return set_error(MI_ERR_BAD_ARG,
                         "a batch with shared values holds one matrix: use mi_batch_set_matrix");
    if (!same_pattern(B, A))
        return set_error(MI_ERR_BAD_ARG, "matrix does not have the batch's pattern");
    MI_TRY(ensure_device());
    return write_matrix(B, eq, A);
}

MI_API int mi_batch_set_matrix(mi_batch* B, const mi_csr* A)
{
    if (!B || !A)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (!same_pattern(B, A))
        return set_error(MI_ERR_BAD_ARG, "matrix does not have the batch's pattern");
    MI_TRY(ensure_device());
    for (int e = 0; e < B->Em; ++e)
        MI_TRY(write_matrix(B, e, A));
    return MI_OK;
}

MI_API int mi_batch_set_values(mi_batch* B, const double* values_host)
{
    if (!B || (B->nnz && !values_host))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (B->shared)
        return set_error(MI_ERR_BAD_ARG,
                         "a batch with shared values holds one matrix: use mi_batch_set_matrix");
    MI_TRY(ensure_device());
    mi_csr tmp;
    tmp.pat = B->src;
    MI_TRY(tmp.val.alloc(size_t(std::max<int64_t>(B->nnz, 1))));
    for (int e = 0; e < B->E; ++e) {
        MI_TRY(tmp.val.upload(values_host + size_t(e) * size_t(B->nnz), size_t(B->nnz), B->stream));
        MI_TRY(mi_batch_set_equation(B, e, &tmp));
    }
    return MI_OK;
}

MI_API int mi_batch_solve_device(mi_batch* B, const mi_solver_cfg* cfg, const double* b_device,
                                 double* x_device, int* iterations, double* rel_residual)
{
    clear_error();
    if (!B || (B->n && (!b_device || !x_device)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    mi_solver_cfg c;
    mi_solver_cfg_default(&c);
    if (cfg)
        c = *cfg;
    MI_TRY(check_cfg(B, &c));
    MI_TRY(ensure_device());
    return solve(B, c, b_device, x_device, iterations, rel_residual);
}

MI_API int mi_batch_apply_preconditioner(mi_batch* B, int preconditioner, const double* r_device,
                                         double* z_device)
{
    clear_error();
    if (!B || (B->n && (!r_device || !z_device)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    mi_solver_cfg c;
    mi_solver_cfg_default(&c);
    c.kind = MI_SOLVER_BICGSTAB;
    c.preconditioner = preconditioner;
    MI_TRY(check_cfg(B, &c));
    MI_TRY(ensure_device());
    const int64_t n = B->n;
    if (n == 0)
        return MI_OK;
    hipStream_t s = B->stream;
    if (preconditioner != B->pc) {
        B->pc = preconditioner;
        B->slots_valid = false;
    }
    if (preconditioner == MI_PC_ILU && !B->ilu_valid)
        MI_TRY(factor_ilu(B));
    if (preconditioner == MI_PC_AMG && !B->amg_valid) {
        if (B->graph) {
            (void)hipGraphExecDestroy(B->graph);
            B->graph = nullptr;
        }
        MI_TRY(amg_setup_batch(B, c));
    }
    if (!B->ext_alloc) {
        MI_TRY(alloc_slots(B, B->set[0], true));
        MI_TRY(alloc_slots(B, B->set[1], true));
        B->ext_alloc = true;
    }
    std::vector<int> all(size_t(B->E));
    for (int e = 0; e < B->E; ++e)
        all[size_t(e)] = e;
    MI_TRY(pack_from_home(B, all, true));
    BSlots& S = B->set[B->cur];
    const size_t m = size_t(n) * size_t(B->E);
    if (B->reordered)
        k_bpermute<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(n, B->E, B->d_perm.p, r_device,
                                                                 B->rt.p, 0);
    else
        MI_HIP(hipMemcpyAsync(B->rt.p, r_device, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    k_btile<double><<<tile_grid(n), kBlock, 0, s>>>(n, S.W, S.d_eq.p, 0ull, 0, 0, 0.f, B->rt.p,
                                                    S.r.p);
    MI_HIP(hipGetLastError());
    BArgs a = B->args(B->cur);
    a.st = nullptr; // every slot
    Launcher L{s};
    if (separate_pc(preconditioner)) {
        MI_TRY(separate_apply(B, a, a.r, a.z, L));
    } else {
        k_bscale<<<reduce_grid(n * S.W, 4), kBlock, 0, s>>>(a, a.r, a.z);
        MI_HIP(hipGetLastError());
    }
    k_bfill<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(int64_t(m), B->x.p);
    MI_HIP(hipGetLastError());
    const uint64_t every = S.W >= 64 ? ~0ull : ((1ull << S.W) - 1);
    k_btile<double><<<tile_grid(n), kBlock, 0, s>>>(n, S.W, S.d_eq.p, every, 1, 0, 0.f, B->x.p,
                                                    a.z);
    MI_HIP(hipGetLastError());
    if (B->reordered)
        k_bpermute<<<reduce_grid(int64_t(m), 4), kBlock, 0, s>>>(n, B->E, B->d_perm.p, B->x.p,
                                                                 z_device, 1);
    else
        MI_HIP(hipMemcpyAsync(z_device, B->x.p, m * sizeof(double), hipMemcpyDeviceToDevice, s));
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_batch_solve(mi_batch* B, const mi_solver_cfg* cfg, const double* b_host,
                          double* x_host, int* iterations, double* rel_residual)
{
    clear_error();
    if (!B || (B->n && (!b_host || !x_host)))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    const size_t m = size_t(B->n) * size_t(B->E);
    DevBuf<double> db, dx;
    MI_TRY(db.upload(b_host, m, B->stream));
    MI_TRY(dx.alloc(std::max<size_t>(m, 1)));
    MI_HIP(hipStreamSynchronize(B->stream));
    const int rc = mi_batch_solve_device(B, cfg, db.p, dx.p, iterations, rel_residual);
    if (rc != MI_OK && rc != MI_ERR_NOT_CONVERGED)
        return rc;
    MI_HIP(hipMemcpy(x_host, dx.p, m * sizeof(double), hipMemcpyDeviceToHost));
    return rc;
}

MI_API int mi_batch_tuning(const mi_batch* B, int* rows_per_block, double* spmm_us,
                           double* spmm_bytes)
{
    if (!B)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (rows_per_block) // rows of one workgroup
        *rows_per_block = int((B->n + B->G - 1) / B->G);
    if (spmm_us)
        *spmm_us = B->tune_us;
    if (spmm_bytes) // FP32 values per matrix, shared columns and row pointers, p once, Ap once
        *spmm_bytes = double(B->nnz) * (4.0 * B->Em + 4.0) + 4.0 * double(B->n + 1) +
                      2.0 * 4.0 * double(B->n) * B->E;
    return MI_OK;
}

MI_API int mi_batch_tuning_shape(const mi_batch* B, int* slots_per_thread_max,
                                 int* spmm_workgroups, int* banded, int* vector_workgroups,
                                 double* vector_us)
{
    if (!B)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (banded)
        *banded = B->band;
    if (slots_per_thread_max)
        *slots_per_thread_max = B->Vmax;
    if (spmm_workgroups)
        *spmm_workgroups = B->G;
    if (vector_workgroups)
        *vector_workgroups = B->Gv;
    if (vector_us)
        *vector_us = B->tune_vec_us;
    return MI_OK;
}

MI_API int mi_batch_set_shape(mi_batch* B, int slots_per_thread_max, int spmm_workgroups,
                              int banded)
{
    clear_error();
    if (!B)
        return set_error(MI_ERR_BAD_ARG, "null handle");
    if (slots_per_thread_max != 1 && slots_per_thread_max != 2 && slots_per_thread_max != 4)
        return set_error(MI_ERR_BAD_ARG, "slots per thread cap must be 1, 2 or 4 (%d given)",
                         slots_per_thread_max);
    if (spmm_workgroups < 8 || spmm_workgroups > kBatchMaxGrid || spmm_workgroups % 8 != 0)
        return set_error(MI_ERR_BAD_ARG, "SpMM workgroups must be a multiple of 8 in [8, %d] (%d given)",
                         kBatchMaxGrid, spmm_workgroups);
    if (banded != 0 && banded != 1)
        return set_error(MI_ERR_BAD_ARG, "banded must be 0 or 1");
    B->Vmax = slots_per_thread_max;
    B->G = spmm_workgroups;
    B->band = banded;
    if (B->graph) { // captured with the previous shape
        (void)hipGraphExecDestroy(B->graph);
        B->graph = nullptr;
    }
    return MI_OK;
}

MI_API int mi_time_grid_barrier(int workgroups, int iterations, double* sync_us,
                                double* launch_us)
{
    clear_error();
    if (workgroups < 1 || iterations < 1 || iterations > 100000)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    int dev = 0, ncu = 0, per = 0, coop = 0;
    MI_HIP(hipGetDevice(&dev));
    MI_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    MI_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    MI_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_gbar_sync, kBlock, 0));
    if (!coop)
        return set_error(MI_ERR_UNSUPPORTED, "device has no cooperative launch");
    if (int64_t(workgroups) > int64_t(per) * ncu)
        return set_error(MI_ERR_BAD_ARG, "%d workgroups exceed the %d x %d that can be resident",
                         workgroups, per, ncu);
    hipStream_t s = nullptr;
    MI_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    DevBuf<float> buf;
    int rc = buf.alloc(size_t(workgroups) * kBlock);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float ms_sync = 0.f, ms_launch = 0.f;
    auto run = [&]() -> int {
        MI_TRY(rc);
        MI_HIP(hipMemsetAsync(buf.p, 0, buf.n * sizeof(float), s));
        MI_HIP(hipEventCreate(&e0));
        MI_HIP(hipEventCreate(&e1));
        int it = iterations;
        float* p = buf.p;
        void* args[] = {&it, &p};
        MI_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_gbar_sync),
                                          dim3(workgroups), dim3(kBlock), args, 0, s));
        MI_HIP(hipEventRecord(e0, s));
        MI_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_gbar_sync),
                                          dim3(workgroups), dim3(kBlock), args, 0, s));
        MI_HIP(hipEventRecord(e1, s));
        MI_HIP(hipEventSynchronize(e1));
        MI_HIP(hipEventElapsedTime(&ms_sync, e0, e1));
        k_gbar_one<<<workgroups, kBlock, 0, s>>>(p);
        MI_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < iterations; ++i)
            k_gbar_one<<<workgroups, kBlock, 0, s>>>(p);
        MI_HIP(hipEventRecord(e1, s));
        MI_HIP(hipEventSynchronize(e1));
        MI_HIP(hipGetLastError());
        MI_HIP(hipEventElapsedTime(&ms_launch, e0, e1));
        return MI_OK;
    };
    rc = run();
    if (e0)
        (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    MI_TRY(rc);
    if (sync_us)
        *sync_us = 1000.0 * ms_sync / iterations;
    if (launch_us)
        *launch_us = 1000.0 * ms_launch / iterations;
    return MI_OK;
}

MI_API int mi_batch_destroy(mi_batch* B)
{
    if (B && B->stream)
        (void)hipStreamSynchronize(B->stream);
    delete B;
    return MI_OK;
}

} // extern "C"
