// batch.hip — multi-RHS solve of equations that share one CSR pattern.
//
// SimAll solves every scalar transport equation (energy, k/omega, species)
// with its own lin_.solve on an identical pattern
// (src/solver/ScalarTransport.cpp:84-242, the solve at :230-232).  Here the
// E equations of one pattern are solved together, each stage of the
// iteration one launch for all of them.
//
// Layout.  The Krylov working set is slot-interleaved: the equations being
// iterated occupy W slots (1 <= W <= 64), and every per-equation
// quantity of a nonzero k or a row i is a line of W consecutive values:
// v32[k * W + s], r[i * W + s], ...  A row of the batched SpMV is handled by
// W / V threads, each taking V consecutive slots (V = 4, 2 or 1, the widest
// dividing W) with V-wide loads: one load of colIdx[k] serves every
// equation, the values of nonzero k are one contiguous W-wide line, and the
// gather of x at column c is one contiguous line too.  The shared pattern is
// streamed once per launch and no gather is scattered across equations.  The
// SpMV runs on an autotuned grid; the vector stages, norms and true
// residuals on one workgroup per 256 rows.  When the
// pattern has a SELL-64 form (short rows: 7-point operators) the nonzeros
// are stored in its slot order, m = soff[slice] + 64 k + row-in-slice, so the
// loads of colIdx and of the value lines of the rows of a workgroup step are
// contiguous for any W.
//
// As equations converge their slots go idle; at the next poll the live ones
// are repacked into a layout exactly as wide as their number (the other of
// two slot buffer sets), so launches stream the bytes of the equations still
// iterating and little more.  Between refinement sweeps everything per
// equation lives equation-major ("home"): FP64 and FP32 values
// v[e * nnz + k], b, x and the FP64 true residual [e * n + i].
//
// Arithmetic: Jacobi-PCG per equation (CGSolver::solve,
// src/solver/LinearSolvers.cpp:420-452), or for nonsymmetric equations
// (convection) Jacobi-preconditioned BiCGStab (BiCGSTABSolver::solve,
// :360-404), with FP32 values, vectors and correction and FP64 dot products
// and norms; the FP64 true residual with the FP64 values refines each
// equation until ||b_e - A_e x_e|| <= tol ||b_e|| (x_e accumulates each
// sweep's correction in FP64).  Each equation has its own scalars and
// convergence flag.  One inner CG iteration is five launches (batched SpMV
// + p.Ap, finalize, update, finalize, direction), one BiCGStab iteration
// nine; `chunk` iterations are replayed from a HIP graph between polls.
#include "csr.hpp"
#include "krylov.hpp"
#include "reorder.hpp"
#include "stream.hpp"

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

struct BEq
{
    double rz, alpha, beta, r0, bnorm, tol_in, rel, last;
    double rho, omega; // BiCGStab: rho of the previous iteration (rz holds the current), omega
    int iter, total, outer, done, fin, status, sconv, pad;
};

// Stages.  B_BI_*: BiCGStab; B_BI_T skips equations whose s already met the
// inner tolerance (sconv).
enum BOp : int {
    B_BNORM = 0, B_TRUE = 1, B_INIT = 2, B_ALPHA = 3, B_UPDATE = 4, B_P = 5,
    B_BI_INIT = 6, B_BI_P = 7, B_BI_ALPHA = 8, B_BI_S = 9, B_BI_T = 10, B_BI_OMEGA = 11,
    B_BI_UPDATE = 12
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
    const float* v32;  // [nnz or padded SELL slots][W]
    const float* invD; // [n][W]
    const int* eq;     // [W] equation of each slot, -1: empty
    float* xs;         // [n][W] correction of the current sweep
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
    return !q.done;
}

// Threads of a workgroup as P = kBlock / W rows of W slots: thread t takes
// slot t % W of row t / W (threads from P W on are idle).
struct SlotMap
{
    int W, P, sl, rr;
    bool live;
    __device__ __forceinline__ explicit SlotMap(int w)
        : W(w), P(kBlock / w), sl(int(threadIdx.x) % w), rr(int(threadIdx.x) / w),
          live(int(threadIdx.x) < (kBlock / w) * w)
    {
    }
};

// FP64 sums of up to two values per slot over a workgroup, in a fixed
// order (rows 0 .. P-1 of each slot); partials[v][s][block].
__device__ __forceinline__ void slot_partials(const BArgs& a, const SlotMap& m, int nv, double d0,
                                              double d1)
{
    __shared__ double lds[2][kBlock];
    lds[0][threadIdx.x] = d0;
    lds[1][threadIdx.x] = d1;
    __syncthreads();
    const int G = gridDim.x;
    if (int(threadIdx.x) < nv * m.W) {
        const int v = threadIdx.x / m.W, sl = threadIdx.x % m.W;
        double s = 0.0;
        for (int q = 0; q < m.P; ++q)
            s += lds[v][q * m.W + sl];
        a.partials[(size_t(v) * m.W + sl) * G + blockIdx.x] = s;
    }
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

// ------------------------------------------------------------------ kernels
// V consecutive FP32 slots of one row as one load (V = 1, 2 or 4).
template <int V> struct SlotVec
{
    using type = float __attribute__((ext_vector_type(V)));
    static __device__ __forceinline__ float at(const type& x, int u) { return x[u]; }
};
template <> struct SlotVec<1>
{
    using type = float;
    static __device__ __forceinline__ float at(float x, int) { return x; }
};

// FP64 sums of up to two values per slot over a workgroup whose threads
// each hold V consecutive slots of one row group (TW = W / V threads per
// row, R = kBlock / TW row groups), in a fixed order (row groups 0 .. R-1
// of each slot); partials[v][s][block].  V = 1 is the SlotMap arrangement.
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
template <int OP, bool kSell, int V>
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
    const float* vin = OP == B_ALPHA ? a.p : OP == B_BI_ALPHA ? a.y : a.z;
    float* vout = OP == B_BI_T ? a.t : a.Ap;
    const float* vw = OP == B_ALPHA ? a.p : OP == B_BI_ALPHA ? a.rh : a.r;
    int64_t r0, r1;
    block_range(a.n, R, r0, r1);
    double d0[V], d1[V];
#pragma unroll
    for (int u = 0; u < V; ++u)
        d0[u] = d1[u] = 0.0;
    if (any) {
        for (int64_t i = r0 + grp; i < r1; i += R) {
            const size_t q = size_t(i) * W + s0;
            // the dot product's operand, loaded ahead of the row's streams
            const VT wq = *reinterpret_cast<const VT*>(vw + q);
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
                            acc += v[j] * xv[j];
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
                        v[j] = k < ke ? __builtin_nontemporal_load(
                                            reinterpret_cast<const VT*>(a.v32 + size_t(k) * W + s0))
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
                            acc += v[j] * xv[j];
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
// thread, the widest that divides W (measured on a 1M-row 7-point batch:
// V = 4 is the fastest at every width from 4 to 32 equations).
template <int OP> using BSpmmKernel = void (*)(BArgs);
template <int OP> BSpmmKernel<OP> bspmm_kernel(bool sell, int W)
{
    if (W % 4 == 0)
        return sell ? k_bspmm<OP, true, 4> : k_bspmm<OP, false, 4>;
    if (W % 2 == 0)
        return sell ? k_bspmm<OP, true, 2> : k_bspmm<OP, false, 2>;
    return sell ? k_bspmm<OP, true, 1> : k_bspmm<OP, false, 1>;
}

// Vector stages over the slot layout (element q = i W + s; thread t of a
// workgroup takes slot t % W of rows t / W, t / W + P, ...):
//   B_INIT      p = D^-1 r; r.z, r.r
//   B_UPDATE    xs += alpha p, r -= alpha Ap; r.r, r.z
//   B_P         p = D^-1 r + beta p
//   B_BI_INIT   rh = r; p = v = 0; r.r, rh.r
//   B_BI_P      p = r + beta (p - omega v); y = D^-1 p
//   B_BI_S      s = r - alpha v (in r); z = D^-1 s; s.s
//   B_BI_UPDATE xs += alpha y + omega z; r = s - omega t; r.r, rh.r
__global__ __launch_bounds__(kBlock) void k_bvec(BArgs a, int op)
{
    const SlotMap m(a.W);
    const int W = m.W;
    const int e = m.live ? a.eq[m.sl] : -1;
    const bool on = slot_on(a, op, e);
    int64_t i0, i1;
    block_range(a.n, m.P, i0, i1);
    // this thread's elements: rows i0 + rr, i0 + rr + P, ... of its slot
    const int64_t q0 = (i0 + m.rr) * W + m.sl, q1 = i1 * W;
    const int64_t step = int64_t(m.P) * W;
    double s0 = 0.0, s1 = 0.0;
    if (on) {
        const BEq& st = a.st[e];
        if (op == B_INIT) {
            for (int64_t q = q0; q < q1; q += step) {
                const float ri = a.r[q], zi = a.invD[q] * ri;
                a.p[q] = zi;
                s0 += double(ri) * double(zi);
                s1 += double(ri) * double(ri);
            }
        } else if (op == B_UPDATE) {
            const float al = float(st.alpha);
            for (int64_t q = q0; q < q1; q += step) {
                a.xs[q] += al * a.p[q];
                const float ri = a.r[q] - al * a.Ap[q];
                a.r[q] = ri;
                const float zi = a.invD[q] * ri;
                s0 += double(ri) * double(ri);
                s1 += double(ri) * double(zi);
            }
        } else if (op == B_P) {
            const float be = float(st.beta);
            for (int64_t q = q0; q < q1; q += step)
                a.p[q] = a.invD[q] * a.r[q] + be * a.p[q];
        } else if (op == B_BI_INIT) {
            // rh = r; p = v = 0 (BiCGSTABSolver::solve, :366-372)
            for (int64_t q = q0; q < q1; q += step) {
                const float ri = a.r[q];
                a.rh[q] = ri;
                a.p[q] = 0.f;
                a.Ap[q] = 0.f;
                s0 += double(ri) * double(ri);
            }
            s1 = s0;
        } else if (op == B_BI_P) {
            // (:380-382)
            const float be = float(st.beta), om = float(st.omega);
            for (int64_t q = q0; q < q1; q += step) {
                const float pi = a.r[q] + be * (a.p[q] - om * a.Ap[q]);
                a.p[q] = pi;
                a.y[q] = a.invD[q] * pi;
            }
        } else if (op == B_BI_S) {
            // (:384-386)
            const float al = float(st.alpha);
            for (int64_t q = q0; q < q1; q += step) {
                const float si = a.r[q] - al * a.Ap[q];
                a.r[q] = si;
                a.z[q] = a.invD[q] * si;
                s0 += double(si) * double(si);
            }
        } else { // B_BI_UPDATE (:387-397)
            const float al = float(st.alpha);
            if (st.sconv) {
                // s met the tolerance: x += alpha y and stop (:387-390)
                for (int64_t q = q0; q < q1; q += step)
                    a.xs[q] += al * a.y[q];
            } else {
                const float om = float(st.omega);
                for (int64_t q = q0; q < q1; q += step) {
                    a.xs[q] += al * a.y[q] + om * a.z[q];
                    const float ri = a.r[q] - om * a.t[q];
                    a.r[q] = ri;
                    s0 += double(ri) * double(ri);
                    s1 += double(a.rh[q]) * double(ri);
                }
            }
        }
    }
    if (op == B_P || op == B_BI_P)
        return;
    slot_partials(a, m, (op == B_BI_S) ? 1 : 2, s0, s1);
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
        const double* __restrict__ ve = a.v64 + size_t(e) * a.nnz;
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
template <class H>
__global__ __launch_bounds__(kBlock) void k_btile(int64_t rows, int W, const int* __restrict__ eq,
                                                  uint64_t mask, int mode, int use_fill,
                                                  float fill, H* home, float* slot)
{
    __shared__ float tile[kTileRows * kMaxEq];
    const int64_t ntile = (rows + kTileRows - 1) / kTileRows;
    for (int64_t tb = blockIdx.x; tb < ntile; tb += gridDim.x) {
        const int64_t i0 = tb * kTileRows;
        const int nr = int(min(int64_t(kTileRows), rows - i0));
        if (mode == 0) {
            // read each slot's run of rows, write whole lines
            for (int t = threadIdx.x; t < kTileRows * W; t += kBlock) {
                const int sl = t / kTileRows, r = t % kTileRows;
                const int e = eq[sl];
                float v = fill;
                if (r < nr && e >= 0 && !use_fill)
                    v = float(home[size_t(e) * rows + i0 + r]);
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
__global__ __launch_bounds__(kBlock) void k_brepack(int64_t rows, int W1, int W2,
                                                    const int* __restrict__ map,
                                                    const float* __restrict__ src,
                                                    float* __restrict__ dst)
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

} // namespace mi

using namespace mi;

// One slot layout: its width, slot table and the arrays of the working set.
struct BSlots
{
    int W = 0;
    std::vector<int> eqs; // equation of each slot, -1 empty (size W)
    DevBuf<int> d_eq;
    DevBuf<float> v32, invD, xs, r, p, Ap, rh, y, z, t;
};

struct mi_batch
{
    std::shared_ptr<Pattern> src; // the caller's pattern
    std::shared_ptr<Pattern> pat; // working pattern: src, or src reordered (RCM)
    DevBuf<int> d_perm, d_vmap;   // perm[new] = old; vmap[k] = source entry of entry k
    bool reordered = false;
    int E = 0, Wmax = 1;
    int G = 8;  // workgroups of the batched SpMM (autotuned)
    int Gv = 8; // workgroups of the vector, norm and true-residual stages (from n)
    int64_t n = 0, nnz = 0;
    bool sell = false;   // values in the pattern's SELL-64 slot order
    int64_t vrows = 0;   // value lines: nnz, or the SELL-64 slots
    // equation-major ("home") arrays
    DevBuf<double> v64, b, x, rt;
    DevBuf<float> v32h, invDh;
    DevBuf<double> partials;
    DevBuf<BEq> st;
    BEq* h_st = nullptr;
    // two slot layouts (packing and repacking write the other one)
    BSlots set[2];
    int cur = 0;
    bool slots_valid = false; // set[cur] holds the home values of its equations
    bool pc_none = false;     // invD of the packed slots is 1
    bool bicg_alloc = false;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0, graph_kind = -1, graph_set = -1, graph_W = 0;
    BFinArgs graph_f{}; // finalize arguments baked into the graph
    double tune_us = 0.0;

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
        a.v32 = S.v32.p;
        a.invD = S.invD.p;
        a.eq = S.d_eq.p;
        a.xs = S.xs.p;
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

// Working-set arrays of a slot layout (sized for the widest layout;
// BiCGStab's extra vectors when first needed).
int alloc_slots(mi_batch* B, BSlots& S, bool bicg)
{
    const size_t nW = size_t(std::max<int64_t>(B->n, 1)) * size_t(B->Wmax);
    const size_t zW = size_t(std::max<int64_t>(B->vrows, 1)) * size_t(B->Wmax);
    if (!S.v32.p) {
        MI_TRY(S.v32.alloc(zW));
        for (DevBuf<float>* v : {&S.invD, &S.xs, &S.r, &S.p, &S.Ap})
            MI_TRY(v->alloc(nW));
        MI_TRY(S.d_eq.alloc(size_t(kMaxEq)));
    }
    if (bicg && !S.rh.p)
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

// Slots for `eqs` in the other layout, values and inverse diagonals from the
// home arrays; it becomes the current layout.
int pack_from_home(mi_batch* B, const std::vector<int>& eqs, bool bicg)
{
    const int nxt = B->cur ^ 1;
    BSlots& S = B->set[nxt];
    MI_TRY(alloc_slots(B, S, bicg));
    S.W = int(eqs.size());
    S.eqs = eqs;
    MI_TRY(upload_eqs(B, S));
    hipStream_t s = B->stream;
    if (B->vrows)
        k_btile<float><<<tile_grid(B->vrows), kBlock, 0, s>>>(B->vrows, S.W, S.d_eq.p, 0ull, 0, 0,
                                                              0.f, B->v32h.p, S.v32.p);
    if (B->n)
        k_btile<float><<<tile_grid(B->n), kBlock, 0, s>>>(B->n, S.W, S.d_eq.p, 0ull, 0,
                                                          B->pc_none ? 1 : 0, 1.f, B->invDh.p,
                                                          S.invD.p);
    MI_HIP(hipGetLastError());
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
    k_btile<double><<<tile_grid(B->n), kBlock, 0, B->stream>>>(B->n, S.W, S.d_eq.p, mask, 1, 0,
                                                               0.f, B->x.p, S.xs.p);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

// The slots `keep` (positions in the current layout) of every working array
// into the other layout, keep.size() wide.
int repack(mi_batch* B, const std::vector<int>& keep, bool bicg)
{
    const BSlots& A = B->set[B->cur];
    BSlots& S = B->set[B->cur ^ 1];
    MI_TRY(alloc_slots(B, S, bicg));
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
    auto cp = [&](int64_t rows, const DevBuf<float>& from, DevBuf<float>& to) {
        if (rows > 0)
            k_brepack<<<reduce_grid(rows * S.W, 4), kBlock, 0, s>>>(rows, A.W, S.W, d_map.p,
                                                                    from.p, to.p);
    };
    cp(B->vrows, A.v32, S.v32);
    cp(B->n, A.invD, S.invD);
    cp(B->n, A.xs, S.xs);
    cp(B->n, A.r, S.r);
    cp(B->n, A.p, S.p);
    cp(B->n, A.Ap, S.Ap);
    if (bicg) {
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

// One CG iteration for every iterating equation (five launches).
int one_iteration(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    MI_TRY(L(bspmm_kernel<B_ALPHA>(B->sell, a.W), B->G, kBlock, a));
    MI_TRY(fin(B, B_ALPHA, f, B->G, L));
    MI_TRY(L(k_bvec, B->Gv, kBlock, a, int(B_UPDATE)));
    MI_TRY(fin(B, B_UPDATE, f, B->Gv, L));
    return L(k_bvec, B->Gv, kBlock, a, int(B_P));
}

// One BiCGStab iteration for every iterating equation (nine launches).
int one_iteration_bicg(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    MI_TRY(L(k_bvec, B->Gv, kBlock, a, int(B_BI_P)));
    MI_TRY(L(bspmm_kernel<B_BI_ALPHA>(B->sell, a.W), B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_ALPHA, f, B->G, L));
    MI_TRY(L(k_bvec, B->Gv, kBlock, a, int(B_BI_S)));
    MI_TRY(fin(B, B_BI_S, f, B->Gv, L));
    MI_TRY(L(bspmm_kernel<B_BI_T>(B->sell, a.W), B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_OMEGA, f, B->G, L));
    MI_TRY(L(k_bvec, B->Gv, kBlock, a, int(B_BI_UPDATE)));
    return fin(B, B_BI_UPDATE, f, B->Gv, L);
}

// Autotuner: workgroups of the batched launches, timed on this pattern in
// the widest layout with every equation's slot active.
int autotune(mi_batch* B)
{
    B->G = 8;
    // vector stages: one workgroup per 256 rows, up to kBatchMaxGrid (the
    // SpMM's best grid can be far smaller and would leave them short of waves)
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
    double best = 1e300;
    BArgs a = B->args(B->cur);
    a.st = nullptr;
    for (int g : {256, 512, 1024, 2048, 4096}) {
        if (g > kBatchMaxGrid)
            break;
        Launcher L{B->stream};
        auto k = bspmm_kernel<B_ALPHA>(B->sell, a.W);
        MI_TRY(L(k, unsigned(g), kBlock, a));
        MI_HIP(hipEventRecord(e0, B->stream));
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
            MI_TRY(L(k, unsigned(g), kBlock, a));
        MI_HIP(hipEventRecord(e1, B->stream));
        MI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MI_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1000.0 * ms / reps;
        if (us < best) {
            best = us;
            B->G = g;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->tune_us = best;
    B->slots_valid = false; // packed from values not yet set
    return MI_OK;
}

int check_cfg(const mi_solver_cfg* c)
{
    if (c->kind != MI_SOLVER_CG && c->kind != MI_SOLVER_BICGSTAB)
        return set_error(MI_ERR_UNSUPPORTED, "batched solve: CG or BiCGStab (kind %d given)",
                         c->kind);
    if (c->preconditioner != MI_PC_JACOBI && c->preconditioner != MI_PC_NONE)
        return set_error(MI_ERR_UNSUPPORTED, "batched solve: Jacobi preconditioner only");
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
    const bool pc_none = c.preconditioner == MI_PC_NONE;
    if (pc_none != B->pc_none) {
        B->pc_none = pc_none;
        B->slots_valid = false;
    }
    if (bicg && !B->bicg_alloc) {
        MI_TRY(alloc_slots(B, B->set[0], true));
        MI_TRY(alloc_slots(B, B->set[1], true));
        B->bicg_alloc = true;
    }
    BFinArgs f{};
    f.E = E;
    f.G = B->Gv;
    f.max_iter = c.max_iterations;
    f.max_outer = c.max_outer > 0 ? c.max_outer : 20;
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
                MI_TRY(pack_from_home(B, want, bicg));
        }
        // r = float(rt) and xs = 0 for this sweep, then the first stage
        {
            BSlots& S = B->set[B->cur];
            k_btile<double><<<tile_grid(n), kBlock, 0, s>>>(n, S.W, S.d_eq.p, 0ull, 0, 0, 0.f,
                                                            B->rt.p, S.r.p);
            k_bzero32<<<reduce_grid(n * S.W, 4), kBlock, 0, s>>>(n * S.W, S.xs.p);
            MI_HIP(hipGetLastError());
            Launcher L{s};
            f.W = S.W;
            MI_TRY(L(k_bvec, B->Gv, kBlock, B->args(B->cur), int(bicg ? B_BI_INIT : B_INIT)));
            MI_TRY(fin(B, bicg ? B_BI_INIT : B_INIT, f, B->Gv, L));
        }
        for (;;) {
            const BArgs a = B->args(B->cur);
            f.W = a.W;
            auto iteration = [&](Launcher& L) {
                return bicg ? one_iteration_bicg(B, a, f, L) : one_iteration(B, a, f, L);
            };
            if (use_graph) {
                if (!B->graph || B->graph_chunk != chunk || B->graph_kind != c.kind ||
                    B->graph_set != B->cur || B->graph_W != a.W ||
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
                MI_TRY(repack(B, live, bicg));
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

} // namespace

extern "C" {

MI_API int mi_batch_create(const mi_csr* A, int n_eq, mi_batch** out)
{
    if (!A || !out || n_eq < 1 || n_eq > kMaxEq)
        return set_error(MI_ERR_BAD_ARG, "need a pattern matrix and 1..%d equations", kMaxEq);
    MI_TRY(ensure_device());
    auto B = std::make_unique<mi_batch>();
    B->src = A->pat;
    B->pat = A->pat;
    B->E = n_eq;
    B->Wmax = n_eq;
    B->n = A->n();
    B->nnz = A->nnz();
    MI_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    if (B->n >= 100000) {
        // the same bandwidth reduction as the single-equation solver's setup
        const Pattern& src = *A->pat;
        std::vector<int> perm, brp, bci, vmap;
        rcm_order(int(src.n), src.h_rp.data(), src.h_ci.data(), perm);
        permute_pattern(int(src.n), src.h_rp.data(), src.h_ci.data(), perm, brp, bci, vmap);
        MI_TRY(make_pattern(src.n, src.nnz, brp.data(), bci.data(), B->pat, B->stream));
        MI_TRY(B->d_perm.upload(perm.data(), perm.size(), B->stream));
        MI_TRY(B->d_vmap.upload(vmap.data(), vmap.size(), B->stream));
        MI_HIP(hipStreamSynchronize(B->stream));
        B->reordered = true;
    }
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&B->h_st), sizeof(BEq) * kMaxEq, 0));
    B->sell = B->pat->sell.built() && B->pat->sell.eperm.p != nullptr;
    B->vrows = B->sell ? B->pat->sell.padded : B->nnz;
    const size_t nE = size_t(std::max<int64_t>(B->n, 1)) * size_t(n_eq);
    const size_t zE = size_t(std::max<int64_t>(B->nnz, 1)) * size_t(n_eq);
    const size_t vE = size_t(std::max<int64_t>(B->vrows, 1)) * size_t(n_eq);
    MI_TRY(B->v64.alloc(zE));
    MI_TRY(B->v32h.alloc(vE));
    MI_TRY(B->invDh.alloc(nE));
    for (DevBuf<double>* v : {&B->b, &B->x, &B->rt})
        MI_TRY(v->alloc(nE));
    MI_TRY(B->partials.alloc(size_t(2) * kBatchMaxGrid * size_t(kMaxEq)));
    MI_TRY(B->st.alloc(size_t(n_eq)));
    MI_HIP(hipMemsetAsync(B->v32h.p, 0, vE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->v64.p, 0, zE * sizeof(double), B->stream));
    MI_HIP(hipMemsetAsync(B->invDh.p, 0, nE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->st.p, 0, sizeof(BEq) * size_t(n_eq), B->stream));
    MI_HIP(hipStreamSynchronize(B->stream));
    MI_TRY(alloc_slots(B.get(), B->set[0], false));
    MI_TRY(alloc_slots(B.get(), B->set[1], false));
    for (BSlots& S : B->set) {
        MI_HIP(hipMemsetAsync(S.p.p, 0, S.p.n * sizeof(float), B->stream));
        MI_HIP(hipMemsetAsync(S.Ap.p, 0, S.Ap.n * sizeof(float), B->stream));
    }
    MI_HIP(hipStreamSynchronize(B->stream));
    MI_TRY(autotune(B.get()));
    *out = B.release();
    return MI_OK;
}

MI_API int mi_batch_set_equation(mi_batch* B, int eq, const mi_csr* A)
{
    if (!B || !A || eq < 0 || eq >= B->E)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (A->pat != B->src && (A->n() != B->n || A->nnz() != B->nnz ||
                             A->pat->h_rp != B->src->h_rp || A->pat->h_ci != B->src->h_ci))
        return set_error(MI_ERR_BAD_ARG, "matrix does not have the batch's pattern");
    MI_TRY(ensure_device());
    hipStream_t s = B->stream;
    double* v64 = B->v64.p + size_t(eq) * size_t(B->nnz);
    float* v32 = B->v32h.p + size_t(eq) * size_t(B->vrows);
    if (B->nnz)
        k_bconvert<<<reduce_grid(B->nnz, 4), kBlock, 0, s>>>(
            B->nnz, B->reordered ? B->d_vmap.p : nullptr, A->val.p, v64, B->sell ? nullptr : v32);
    if (B->sell && B->vrows)
        k_bsell32<<<reduce_grid(B->vrows, 4), kBlock, 0, s>>>(B->vrows, B->pat->sell.eperm.p, v64,
                                                              v32);
    if (B->n)
        k_binvdiag<<<grid_for(B->n, kBlock), kBlock, 0, s>>>(
            int(B->n), B->pat->diag.p, v64, B->invDh.p + size_t(eq) * size_t(B->n));
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    B->slots_valid = false;
    return MI_OK;
}

MI_API int mi_batch_set_values(mi_batch* B, const double* values_host)
{
    if (!B || (B->nnz && !values_host))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
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
    MI_TRY(check_cfg(&c));
    MI_TRY(ensure_device());
    return solve(B, c, b_device, x_device, iterations, rel_residual);
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
    if (spmm_bytes) // FP32 values per equation, shared columns and row pointers, p once, Ap once
        *spmm_bytes = double(B->nnz) * (4.0 * B->E + 4.0) + 4.0 * double(B->n + 1) +
                      2.0 * 4.0 * double(B->n) * B->E;
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
