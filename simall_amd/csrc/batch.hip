// batch.hip — multi-RHS solve of equations that share one CSR pattern.
//
// SimAll solves every scalar transport equation (energy, k/omega, species)
// with its own lin_.solve on an identical pattern
// (src/solver/ScalarTransport.cpp:84-242, the solve at :230-232).  Here the
// E equations of one pattern are solved together: each stage of the
// iteration is one launch for all equations.
//
// Layout: values, vectors and inverse diagonals are interleaved by equation
// (element (k, e) at k * Ep + e, Ep = E rounded up to a power of two).  A
// workgroup of 256 threads covers 256 / Ep rows x Ep equations, thread
// (row slot, e) owns row i of equation e and walks the row in column order.
// For each nonzero the Ep lanes of a row read Ep consecutive values and
// gather Ep consecutive x entries: one 4*Ep-byte segment each, and the
// shared column index once for all equations.
//
// Arithmetic: Jacobi-PCG per equation (CGSolver::solve,
// src/solver/LinearSolvers.cpp:420-452) with FP32 values and vectors and FP64
// dot products, norms and iterate (the iterate is accumulated in FP64 as
// in the mixed single-equation solver); an FP64 true residual with the FP64
// values refines each equation until ||b_e - A_e x_e|| <= tol ||b_e||.
// Each equation has its own scalars and convergence flag: converged
// equations drop out of every launch.  One inner iteration is five launches
// (batched SpMV + p.Ap, finalize, update, finalize, direction) captured once
// into a HIP graph of `chunk` iterations that the host replays between polls.
#include "csr.hpp"
#include "krylov.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

namespace mi
{

struct BEq
{
    double rz, alpha, beta, r0, bnorm, tol_in, rel, last;
    int iter, total, outer, done, fin, status, pad[2];
};

enum BOp : int { B_BNORM = 0, B_TRUE = 1, B_INIT = 2, B_ALPHA = 3, B_UPDATE = 4 };

struct BArgs
{
    int n, E, Ep, RB;
    const int* rp;
    const int* ci;
    const float* v32;
    const double* v64;
    const float* invD;
    const double* b;
    double* x;
    double* rt;
    float* r;
    float* p;
    float* Ap;
    BEq* st;
    double* partials; // [2][G][Ep]
};

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ bool b_active(const BArgs& a, int e, bool inner)
{
    return e < a.E && (a.st == nullptr || (inner ? !a.st[e].done : !a.st[e].fin));
}

// Sum of v over the row slots of each equation into partials[j][block][e].
template <int NV>
__device__ __forceinline__ void b_partials(const BArgs& a, double (&v)[NV])
{
    __shared__ double red[NV][kBlock];
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < NV; ++j)
        red[j][t] = v[j];
    __syncthreads();
    const int Ep = a.Ep, RG = kBlock / Ep;
    if (t < Ep) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
            double s = 0.0;
            for (int q = 0; q < RG; ++q)
                s += red[j][q * Ep + t];
            a.partials[(size_t(j) * gridDim.x + blockIdx.x) * Ep + t] = s;
        }
    }
}

// ------------------------------------------------------------------ kernels
// Ap = A p for the active equations; partials of p.Ap.
__global__ __launch_bounds__(kBlock) void k_bspmv_dot(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    double v[1] = {0.0};
    if (b_active(a, e, true)) {
        const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
        for (int i = r0 + slot; i < r1; i += RG) {
            float s = 0.f;
            for (int k = a.rp[i]; k < a.rp[i + 1]; ++k) {
                const int c = a.ci[k];
                s += __builtin_nontemporal_load(&a.v32[size_t(k) * Ep + e]) *
                     a.p[size_t(c) * Ep + e];
            }
            const size_t q = size_t(i) * Ep + e;
            a.Ap[q] = s;
            v[0] += double(a.p[q]) * double(s);
        }
    }
    b_partials<1>(a, v);
}

// rt = b - A x in FP64 (FP64 values), r = float(rt); partials of ||rt||^2.
__global__ __launch_bounds__(kBlock) void k_btrue(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    double v[1] = {0.0};
    if (b_active(a, e, false)) {
        const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
        for (int i = r0 + slot; i < r1; i += RG) {
            double s = 0.0;
            for (int k = a.rp[i]; k < a.rp[i + 1]; ++k)
                s += a.v64[size_t(k) * Ep + e] * a.x[size_t(a.ci[k]) * Ep + e];
            const size_t q = size_t(i) * Ep + e;
            const double ri = a.b[q] - s;
            a.rt[q] = ri;
            a.r[q] = float(ri);
            v[0] += ri * ri;
        }
    }
    b_partials<1>(a, v);
}

// partials of ||b||^2 (all equations)
__global__ __launch_bounds__(kBlock) void k_bnorm(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    double v[1] = {0.0};
    if (e < a.E) {
        const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
        for (int i = r0 + slot; i < r1; i += RG) {
            const double bi = a.b[size_t(i) * Ep + e];
            v[0] += bi * bi;
        }
    }
    b_partials<1>(a, v);
}

// inner start from d = 0: p = D^-1 r; partials of r.z, r.r
__global__ __launch_bounds__(kBlock) void k_binit(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    double v[2] = {0.0, 0.0};
    if (b_active(a, e, true)) {
        const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
        for (int i = r0 + slot; i < r1; i += RG) {
            const size_t q = size_t(i) * Ep + e;
            const float ri = a.r[q], zi = a.invD[q] * ri;
            a.p[q] = zi;
            v[0] += double(ri) * double(zi);
            v[1] += double(ri) * double(ri);
        }
    }
    b_partials<2>(a, v);
}

// x += alpha p (FP64), r -= alpha Ap; partials of r.r, r.z
__global__ __launch_bounds__(kBlock) void k_bupdate(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    double v[2] = {0.0, 0.0};
    if (b_active(a, e, true)) {
        const double al = a.st[e].alpha;
        const float alf = float(al);
        const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
        for (int i = r0 + slot; i < r1; i += RG) {
            const size_t q = size_t(i) * Ep + e;
            a.x[q] += al * double(a.p[q]);
            const float ri = a.r[q] - alf * a.Ap[q];
            a.r[q] = ri;
            const float zi = a.invD[q] * ri;
            v[0] += double(ri) * double(ri);
            v[1] += double(ri) * double(zi);
        }
    }
    b_partials<2>(a, v);
}

// p = D^-1 r + beta p
__global__ __launch_bounds__(kBlock) void k_bp(BArgs a)
{
    const int lb = xcd_logical_block(blockIdx.x, gridDim.x);
    const int Ep = a.Ep, RG = kBlock / Ep;
    const int e = threadIdx.x & (Ep - 1), slot = threadIdx.x / Ep;
    if (!b_active(a, e, true))
        return;
    const float be = float(a.st[e].beta);
    const int r0 = lb * a.RB, r1 = min(a.n, r0 + a.RB);
    for (int i = r0 + slot; i < r1; i += RG) {
        const size_t q = size_t(i) * Ep + e;
        a.p[q] = a.invD[q] * a.r[q] + be * a.p[q];
    }
}

struct BFinArgs
{
    int E, Ep, G, max_iter, max_outer;
    double tol, inner_tol;
};

// One block: fold the G partials of each equation (fixed order) and update
// the equation's scalars.
__global__ __launch_bounds__(kBlock) void k_bfin(const double* __restrict__ partials, BEq* st,
                                                 int op, BFinArgs f)
{
    __shared__ double red[2][kBlock];
    const int Ep = f.Ep, RG = kBlock / Ep;
    const int t = threadIdx.x, e = t & (Ep - 1), j = t / Ep;
    const int nv = (op == B_INIT || op == B_UPDATE) ? 2 : 1;
    for (int v = 0; v < 2; ++v) {
        double s = 0.0;
        if (v < nv)
            for (int g = j; g < f.G; g += RG)
                s += partials[(size_t(v) * f.G + g) * Ep + e];
        red[v][t] = s;
    }
    __syncthreads();
    if (t >= Ep || e >= f.E)
        return;
    double tot[2] = {0.0, 0.0};
    for (int v = 0; v < nv; ++v)
        for (int q = 0; q < RG; ++q)
            tot[v] += red[v][q * Ep + e];
    BEq& s = st[e];
    switch (op) {
    case B_BNORM:
        s.bnorm = fmax(sqrt(tot[0]), 1e-30);
        s.fin = 0;
        s.done = 1;
        s.total = 0;
        s.outer = 0;
        s.status = 0;
        s.rel = 1.0;
        if (tot[0] == 0.0) { // zero rhs: x = 0 (the caller zeroed x)
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
        } else if (s.outer >= f.max_outer || s.total >= f.max_iter) {
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
    default:
        break;
    }
}

// interleave: dst[i * Ep + e] = src[i] (FP64 and optional FP32 copy)
__global__ __launch_bounds__(kBlock) void k_bscatter(int64_t m, const double* __restrict__ src,
                                                     int Ep, int e, double* __restrict__ d64,
                                                     float* __restrict__ d32)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
        const double v = src[i];
        if (d64)
            d64[i * Ep + e] = v;
        if (d32)
            d32[i * Ep + e] = float(v);
    }
}

__global__ __launch_bounds__(kBlock) void k_bgather(int64_t m, const double* __restrict__ src,
                                                    int Ep, int e, double* __restrict__ dst)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride)
        dst[i] = src[i * Ep + e];
}

// invD[i * Ep + e] = 1 / a_ii (1 when missing or zero: LinearSolvers.cpp:54-65)
__global__ __launch_bounds__(kBlock) void k_binvdiag(int n, const int* __restrict__ diag,
                                                     const double* __restrict__ val, int Ep, int e,
                                                     float* __restrict__ invD)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int k = diag[i];
    const double a = k >= 0 ? val[k] : 0.0;
    invD[size_t(i) * Ep + e] = float(a != 0.0 ? 1.0 / a : 1.0);
}

__global__ __launch_bounds__(kBlock) void k_bfill(int64_t m, double* __restrict__ x)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride)
        x[i] = 0.0;
}

} // namespace mi

using namespace mi;

struct mi_batch
{
    std::shared_ptr<Pattern> pat;
    int E = 0, Ep = 1, RB = 256, G = 8;
    int64_t n = 0, nnz = 0;
    DevBuf<float> v32, invD, r, p, Ap;
    DevBuf<double> v64, b, x, rt, partials, io;
    DevBuf<BEq> st;
    BEq* h_st = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0;
    BFinArgs graph_f{}; // finalize arguments baked into the captured graph
    double tune_us = 0.0;
    std::vector<int> tried_rb;
    std::vector<double> tried_us;

    ~mi_batch()
    {
        if (graph)
            (void)hipGraphExecDestroy(graph);
        if (h_st)
            (void)hipHostFree(h_st);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
    BArgs args() const
    {
        BArgs a{};
        a.n = int(n);
        a.E = E;
        a.Ep = Ep;
        a.RB = RB;
        a.rp = pat->rp.p;
        a.ci = pat->ci.p;
        a.v32 = v32.p;
        a.v64 = v64.p;
        a.invD = invD.p;
        a.b = b.p;
        a.x = x.p;
        a.rt = rt.p;
        a.r = r.p;
        a.p = p.p;
        a.Ap = Ap.p;
        a.st = st.p;
        a.partials = partials.p;
        return a;
    }
    unsigned grid_for_rb(int rb) const
    {
        const int64_t g = (n + rb - 1) / rb;
        return unsigned((g + 7) / 8 * 8);
    }
};

namespace
{

int launch_all(mi_batch* B, void (*k)(BArgs), const BArgs& a)
{
    k<<<B->G, kBlock, 0, B->stream>>>(a);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int fin(mi_batch* B, int op, const BFinArgs& f)
{
    k_bfin<<<1, kBlock, 0, B->stream>>>(B->partials.p, B->st.p, op, f);
    MI_HIP(hipGetLastError());
    return MI_OK;
}

int one_iteration(mi_batch* B, const BArgs& a, const BFinArgs& f)
{
    MI_TRY(launch_all(B, k_bspmv_dot, a));
    MI_TRY(fin(B, B_ALPHA, f));
    MI_TRY(launch_all(B, k_bupdate, a));
    MI_TRY(fin(B, B_UPDATE, f));
    return launch_all(B, k_bp, a);
}

// Autotuner: rows per workgroup for the batched SpMV, timed on this pattern
// with all equations active.
int autotune(mi_batch* B)
{
    if (B->n == 0)
        return MI_OK;
    const int RG = kBlock / B->Ep;
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int best_rb = RG;
    B->tried_rb.clear();
    B->tried_us.clear();
    for (int mult : {1, 2, 4, 8, 16, 32, 64}) {
        const int rb = RG * mult;
        if (mult > 1 && int64_t(rb) > B->n)
            break;
        BArgs a = B->args();
        a.RB = rb;
        a.st = nullptr; // all equations active
        const unsigned g = B->grid_for_rb(rb);
        if (size_t(g) * size_t(B->Ep) * 2 > B->partials.n)
            continue;
        k_bspmv_dot<<<g, kBlock, 0, B->stream>>>(a);
        MI_HIP(hipGetLastError());
        MI_HIP(hipEventRecord(e0, B->stream));
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
            k_bspmv_dot<<<g, kBlock, 0, B->stream>>>(a);
        MI_HIP(hipEventRecord(e1, B->stream));
        MI_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MI_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double us = 1000.0 * ms / reps;
        B->tried_rb.push_back(rb);
        B->tried_us.push_back(us);
        if (us < best) {
            best = us;
            best_rb = rb;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    B->RB = best_rb;
    B->G = int(B->grid_for_rb(best_rb));
    B->tune_us = best;
    return MI_OK;
}

int check_cfg(const mi_solver_cfg* c)
{
    if (c->kind != MI_SOLVER_CG)
        return set_error(MI_ERR_UNSUPPORTED, "batched solve: only CG (kind %d given)", c->kind);
    if (c->preconditioner != MI_PC_JACOBI && c->preconditioner != MI_PC_NONE)
        return set_error(MI_ERR_UNSUPPORTED, "batched solve: Jacobi preconditioner only");
    if (!(c->tolerance > 0.0) || c->max_iterations < 0)
        return set_error(MI_ERR_BAD_ARG, "bad tolerance or iteration limit");
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
    BFinArgs f{};
    f.E = E;
    f.Ep = B->Ep;
    f.G = B->G;
    f.max_iter = c.max_iterations;
    f.max_outer = c.max_outer > 0 ? c.max_outer : 20;
    f.tol = c.tolerance;
    f.inner_tol = c.inner_tolerance > 0 ? c.inner_tolerance : 1e-5;
    const unsigned vg = reduce_grid(n, 4);
    // b, x = 0 into the interleaved layout
    for (int e = 0; e < E; ++e)
        if (n)
            k_bscatter<<<vg, kBlock, 0, s>>>(n, b_dev + size_t(e) * n, B->Ep, e, B->b.p, nullptr);
    k_bfill<<<reduce_grid(n * B->Ep, 4), kBlock, 0, s>>>(n * B->Ep, B->x.p);
    MI_HIP(hipGetLastError());
    const BArgs a = B->args();
    MI_TRY(launch_all(B, k_bnorm, a));
    MI_TRY(fin(B, B_BNORM, f));
    const int chunk = c.check_interval > 0 ? c.check_interval : 8;
    const bool use_graph = c.use_graph != 0;
    if (use_graph && (!B->graph || B->graph_chunk != chunk ||
        std::memcmp(&B->graph_f, &f, sizeof(BFinArgs)) != 0)) {
        if (B->graph) {
            (void)hipGraphExecDestroy(B->graph);
            B->graph = nullptr;
        }
        MI_HIP(hipStreamSynchronize(s));
        hipGraph_t g = nullptr;
        MI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = MI_OK;
        for (int it = 0; it < chunk && rc == MI_OK; ++it)
            rc = one_iteration(B, a, f);
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != MI_OK || ec != hipSuccess || !g) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc != MI_OK ? rc : set_error(MI_ERR_DEVICE, "graph capture failed");
        }
        const hipError_t ei = hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        MI_HIP(ei);
        B->graph_chunk = chunk;
        B->graph_f = f;
    }
    for (;;) {
        MI_TRY(launch_all(B, k_btrue, a));
        MI_TRY(fin(B, B_TRUE, f));
        MI_HIP(hipMemcpyAsync(B->h_st, B->st.p, sizeof(BEq) * size_t(E), hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        bool all_fin = true;
        for (int e = 0; e < E; ++e)
            all_fin = all_fin && B->h_st[e].fin;
        if (all_fin)
            break;
        MI_TRY(launch_all(B, k_binit, a));
        MI_TRY(fin(B, B_INIT, f));
        for (;;) {
            if (use_graph) {
                MI_HIP(hipGraphLaunch(B->graph, s));
            } else {
                for (int it = 0; it < chunk; ++it)
                    MI_TRY(one_iteration(B, a, f));
            }
            MI_HIP(hipMemcpyAsync(B->h_st, B->st.p, sizeof(BEq) * size_t(E), hipMemcpyDeviceToHost,
                                  s));
            MI_HIP(hipStreamSynchronize(s));
            bool all_done = true;
            for (int e = 0; e < E; ++e)
                all_done = all_done && B->h_st[e].done;
            if (all_done)
                break;
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
    for (int e = 0; e < E; ++e)
        if (n)
            k_bgather<<<vg, kBlock, 0, s>>>(n, B->x.p, B->Ep, e, x_dev + size_t(e) * n);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    if (status != MI_OK)
        set_error(MI_ERR_NOT_CONVERGED, "batched solve: some equations did not converge");
    return status;
}

} // namespace

extern "C" {

MI_API int mi_batch_create(const mi_csr* A, int n_eq, mi_batch** out)
{
    if (!A || !out || n_eq < 1 || n_eq > 64)
        return set_error(MI_ERR_BAD_ARG, "need a pattern matrix and 1..64 equations");
    MI_TRY(ensure_device());
    auto B = std::make_unique<mi_batch>();
    B->pat = A->pat;
    B->E = n_eq;
    while (B->Ep < n_eq)
        B->Ep *= 2;
    B->n = A->n();
    B->nnz = A->nnz();
    MI_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    MI_HIP(hipHostMalloc(reinterpret_cast<void**>(&B->h_st), sizeof(BEq) * 64, 0));
    const size_t nE = size_t(std::max<int64_t>(B->n, 1)) * size_t(B->Ep);
    const size_t zE = size_t(std::max<int64_t>(B->nnz, 1)) * size_t(B->Ep);
    MI_TRY(B->v32.alloc(zE));
    MI_TRY(B->v64.alloc(zE));
    for (DevBuf<float>* v : {&B->invD, &B->r, &B->p, &B->Ap})
        MI_TRY(v->alloc(nE));
    for (DevBuf<double>* v : {&B->b, &B->x, &B->rt})
        MI_TRY(v->alloc(nE));
    // partials for the smallest candidate row block (the autotuner's first)
    const int RG = kBlock / B->Ep;
    MI_TRY(B->partials.alloc(size_t(2) * B->grid_for_rb(RG) * size_t(B->Ep)));
    MI_TRY(B->st.alloc(size_t(B->Ep)));
    MI_HIP(hipMemsetAsync(B->v32.p, 0, zE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->p.p, 0, nE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->st.p, 0, sizeof(BEq) * size_t(B->Ep), B->stream));
    MI_HIP(hipStreamSynchronize(B->stream));
    MI_TRY(autotune(B.get()));
    *out = B.release();
    return MI_OK;
}

MI_API int mi_batch_set_equation(mi_batch* B, int eq, const mi_csr* A)
{
    if (!B || !A || eq < 0 || eq >= B->E)
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    if (A->pat != B->pat && (A->n() != B->n || A->nnz() != B->nnz ||
                             A->pat->h_rp != B->pat->h_rp || A->pat->h_ci != B->pat->h_ci))
        return set_error(MI_ERR_BAD_ARG, "matrix does not have the batch's pattern");
    MI_TRY(ensure_device());
    hipStream_t s = B->stream;
    if (B->nnz)
        k_bscatter<<<reduce_grid(B->nnz, 4), kBlock, 0, s>>>(B->nnz, A->val.p, B->Ep, eq, B->v64.p,
                                                            B->v32.p);
    if (B->n)
        k_binvdiag<<<grid_for(B->n, kBlock), kBlock, 0, s>>>(int(B->n), B->pat->diag.p, A->val.p,
                                                            B->Ep, eq, B->invD.p);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

MI_API int mi_batch_set_values(mi_batch* B, const double* values_host)
{
    if (!B || (B->nnz && !values_host))
        return set_error(MI_ERR_BAD_ARG, "bad argument");
    MI_TRY(ensure_device());
    mi_csr tmp;
    tmp.pat = B->pat;
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
    if (rows_per_block)
        *rows_per_block = B->RB;
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
