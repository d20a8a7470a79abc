// batch.hip — multi-RHS solve of equations that share one CSR pattern.
//
// SimAll solves every scalar transport equation (energy, k/omega, species)
// with its own lin_.solve on an identical pattern
// (src/solver/ScalarTransport.cpp:84-242, the solve at :230-232).  Here the
// E equations of one pattern are solved together: each stage of the
// iteration is one launch for all equations.
//
// Layout: equation-major.  Values v[e * nnz + k] (FP32 for the iteration,
// FP64 for the true residual), vectors x[e * n + i].  A persistent workgroup
// walks its range of the pattern's row blocks (<= 256 rows, <= 2048
// nonzeros); for each row block it stages the shared column indices in LDS
// once, then for every equation still iterating streams that equation's
// values, gathers its x, and sums each row in column order (thread per row).
// An equation that has converged is skipped for the whole launch, so the
// traffic of a launch is that of the active equations plus one pass over the
// shared pattern.
//
// Arithmetic: Jacobi-PCG per equation (CGSolver::solve,
// src/solver/LinearSolvers.cpp:420-452), or for nonsymmetric equations
// (convection) Jacobi-preconditioned BiCGStab (BiCGSTABSolver::solve,
// :360-404), with FP32 values and vectors and FP64 dot products, norms and
// iterate (accumulated in FP64 as in the mixed single-equation solver); an
// FP64 true residual with the FP64 values refines each equation until
// ||b_e - A_e x_e|| <= tol ||b_e||.  Each equation has its own scalars and
// convergence flag.  One inner CG iteration is five launches (batched SpMV +
// p.Ap, finalize, update, finalize, direction), one BiCGStab iteration nine;
// they are captured once into a HIP graph of `chunk` iterations that the
// host replays between polls.
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

struct BArgs
{
    int n, E, nrb;
    int64_t nnz;
    const int* blk; // pattern row blocks
    const int* rp;
    const int* ci;
    const float* v32;  // [E][nnz]
    const double* v64; // [E][nnz]
    const float* invD; // [E][n]
    const double* b;   // [E][n]
    double* x;
    double* rt;
    float* r;          // BiCGStab: r, then s in place
    float* p;
    float* Ap;         // BiCGStab: v
    float* rh;         // BiCGStab shadow residual, y = D^-1 p, z = D^-1 s, t = A z
    float* y;
    float* z;
    float* t;
    BEq* st;           // nullptr: every equation active (autotuner)
    double* partials;  // [2][E][G]
};

__device__ __forceinline__ bool b_iterating(const BArgs& a, int e)
{
    return a.st == nullptr || !a.st[e].done;
}

// Per-workgroup FP64 sums of up to two values per equation: each wavefront
// adds its share into its own slot (fixed order), the slots are folded at
// the end and stored as partials[v][e][block].
struct BSums
{
    double (*w)[kMaxEq][kWaves];
    __device__ __forceinline__ void clear(int E)
    {
        for (int t = threadIdx.x; t < 2 * kMaxEq * kWaves; t += kBlock)
            (&w[0][0][0])[t] = 0.0;
        __syncthreads();
        (void)E;
    }
    __device__ __forceinline__ void add(int v, int e, double s)
    {
        s = wave_sum(s);
        if ((threadIdx.x & (kWave - 1)) == 0)
            w[v][e][threadIdx.x / kWave] += s;
    }
    __device__ __forceinline__ void store(const BArgs& a, int nv)
    {
        __syncthreads();
        const int G = gridDim.x;
        for (int t = threadIdx.x; t < nv * a.E; t += kBlock) {
            const int v = t / a.E, e = t % a.E;
            double s = 0.0;
            for (int q = 0; q < kWaves; ++q)
                s += w[v][e][q];
            a.partials[(size_t(v) * a.E + e) * G + blockIdx.x] = s;
        }
    }
};

// ------------------------------------------------------------------ kernels
// Work of a launch: units (equation, row block) for SpMV-shaped stages and
// (equation, row chunk) for vector stages, over the equations the stage
// applies to.  Every workgroup builds the same list of those equations, then
// walks its contiguous range of units, so units of one equation are
// consecutive and each is processed exactly like one block of the
// single-equation CSR-stream engine.
constexpr int kChunkRows = kBlock * 8;

__device__ __forceinline__ int b_list(const BArgs& a, int op, int* act)
{
    __shared__ int cnt;
    if (threadIdx.x == 0) {
        int c = 0;
        for (int e = 0; e < a.E; ++e) {
            bool on = true;
            if (op == B_TRUE)
                on = !a.st[e].fin;
            else if (op == B_BI_T)
                on = b_iterating(a, e) && !a.st[e].sconv;
            else if (op != B_BNORM)
                on = b_iterating(a, e);
            if (on)
                act[c++] = e;
        }
        cnt = c;
    }
    __syncthreads();
    return cnt;
}

// Batched SpMV out_e = A_e in_e (FP32) for the equations of the stage, with
// FP64 partials of per-row products:
//   B_ALPHA     Ap = A p;  p.Ap                (CG)
//   B_BI_ALPHA  v = A y;   rh.v                (BiCGStab)
//   B_BI_T      t = A z;   t.s, t.t            (BiCGStab, s held in r)
template <int OP>
__global__ __launch_bounds__(kBlock) void k_bspmv(BArgs a)
{
    __shared__ float prod[kStreamNnz];
    __shared__ int act[kMaxEq];
    __shared__ double wsum[2][kMaxEq][kWaves];
    BSums S{wsum};
    S.clear(a.E);
    const int nact = b_list(a, OP, act);
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int64_t units = int64_t(nact) * a.nrb;
    const int64_t per = (units + G - 1) / G;
    const int64_t u0 = lb * per, u1 = min(units, u0 + per);
    const int tid = threadIdx.x;
    const float* vin = OP == B_ALPHA ? a.p : OP == B_BI_ALPHA ? a.y : a.z;
    float* vout = OP == B_BI_T ? a.t : a.Ap;
    const float* vw = OP == B_ALPHA ? a.p : OP == B_BI_ALPHA ? a.rh : a.r;
    for (int64_t u = u0; u < u1; ++u) {
        const int e = act[u / a.nrb];
        const int rb = int(u % a.nrb);
        const float* __restrict__ ve = a.v32 + size_t(e) * a.nnz;
        const float* __restrict__ pe = vin + size_t(e) * a.n;
        const float* __restrict__ we = vw + size_t(e) * a.n;
        float* __restrict__ Ae = vout + size_t(e) * a.n;
        const int r0 = a.blk[rb], r1 = a.blk[rb + 1];
        const int k0 = a.rp[r0], k1 = a.rp[r1];
        if (k1 - k0 > kStreamNnz) {
            // one long row: strided sums, fixed-order fold over the block
            float s = 0.f;
            for (int k = k0 + tid; k < k1; k += kBlock)
                s += ve[k] * pe[a.ci[k]];
            prod[tid] = s;
            __syncthreads();
            double d0 = 0.0, d1 = 0.0;
            if (tid == 0) {
                float t = 0.f;
                for (int j = 0; j < kBlock; ++j)
                    t += prod[j];
                Ae[r0] = t;
                d0 = double(we[r0]) * double(t);
                d1 = double(t) * double(t);
            }
            S.add(0, e, d0);
            if (OP == B_BI_T)
                S.add(1, e, d1);
            __syncthreads();
            continue;
        }
        const int i = r0 + tid;
        float wi = 0.f;
        if (i < r1)
            wi = we[i];
        int c[kLoadsPerThread];
        float v[kLoadsPerThread];
#pragma unroll
        for (int j = 0; j < kLoadsPerThread; ++j) {
            const int k = k0 + tid + j * kBlock;
            c[j] = k < k1 ? __builtin_nontemporal_load(&a.ci[k]) : -1;
            v[j] = k < k1 ? __builtin_nontemporal_load(&ve[k]) : 0.f;
        }
#pragma unroll
        for (int j = 0; j < kLoadsPerThread; ++j)
            if (c[j] >= 0)
                prod[tid + j * kBlock] = v[j] * pe[c[j]];
        __syncthreads();
        double d0 = 0.0, d1 = 0.0;
        if (i < r1) {
            float s = 0.f;
            const int end = a.rp[i + 1] - k0;
            for (int k = a.rp[i] - k0; k < end; ++k)
                s += prod[k];
            Ae[i] = s;
            d0 = double(wi) * double(s);
            d1 = double(s) * double(s);
        }
        S.add(0, e, d0);
        if (OP == B_BI_T)
            S.add(1, e, d1);
        __syncthreads();
    }
    S.store(a, OP == B_BI_T ? 2 : 1);
}

// rt_e = b_e - A_e x_e in FP64 for the unfinished equations, r_e = float(rt_e);
// partials of ||rt_e||^2.
__global__ __launch_bounds__(kBlock) void k_btrue(BArgs a)
{
    __shared__ double prod[kStreamNnz];
    __shared__ int act[kMaxEq];
    __shared__ double wsum[2][kMaxEq][kWaves];
    BSums S{wsum};
    S.clear(a.E);
    const int nact = b_list(a, B_TRUE, act);
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
                a.r[off + r0] = float(ri);
                rr = ri * ri;
            }
            S.add(0, e, rr);
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
            a.r[off + i] = float(ri);
            rr = ri * ri;
        }
        S.add(0, e, rr);
        __syncthreads();
    }
    S.store(a, 1);
}

// Vector stages over units (equation, chunk of kChunkRows rows):
//   B_BNORM  partials ||b_e||^2 (all equations)
//   B_INIT   p = D^-1 r; partials r.z, r.r
//   B_UPDATE x += alpha p (FP64), r -= alpha Ap; partials r.r, r.z
//   B_P      p = D^-1 r + beta p
__global__ __launch_bounds__(kBlock) void k_bvec(BArgs a, int op)
{
    __shared__ int act[kMaxEq];
    __shared__ double wsum[2][kMaxEq][kWaves];
    BSums S{wsum};
    S.clear(a.E);
    const int nact = b_list(a, op, act);
    const int G = gridDim.x;
    const int lb = xcd_logical_block(blockIdx.x, G);
    const int nch = (a.n + kChunkRows - 1) / kChunkRows;
    const int64_t units = int64_t(nact) * nch;
    const int64_t per = (units + G - 1) / G;
    const int64_t u0 = lb * per, u1 = min(units, u0 + per);
    for (int64_t u = u0; u < u1; ++u) {
        const int e = act[u / nch];
        const int c0 = int(u % nch) * kChunkRows;
        const int c1 = min(a.n, c0 + kChunkRows);
        const size_t off = size_t(e) * a.n;
        double s0 = 0.0, s1 = 0.0;
        if (op == B_BNORM) {
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock)
                s0 += a.b[off + i] * a.b[off + i];
        } else if (op == B_INIT) {
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                const float ri = a.r[off + i], zi = a.invD[off + i] * ri;
                a.p[off + i] = zi;
                s0 += double(ri) * double(zi);
                s1 += double(ri) * double(ri);
            }
        } else if (op == B_UPDATE) {
            const double al = a.st[e].alpha;
            const float alf = float(al);
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                a.x[off + i] += al * double(a.p[off + i]);
                const float ri = a.r[off + i] - alf * a.Ap[off + i];
                a.r[off + i] = ri;
                const float zi = a.invD[off + i] * ri;
                s0 += double(ri) * double(ri);
                s1 += double(ri) * double(zi);
            }
        } else if (op == B_P) {
            const float be = float(a.st[e].beta);
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock)
                a.p[off + i] = a.invD[off + i] * a.r[off + i] + be * a.p[off + i];
            continue;
        } else if (op == B_BI_INIT) {
            // rh = r; p = v = 0; partials r.r, rh.r (BiCGSTABSolver::solve, :366-372)
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                const float ri = a.r[off + i];
                a.rh[off + i] = ri;
                a.p[off + i] = 0.f;
                a.Ap[off + i] = 0.f;
                s0 += double(ri) * double(ri);
            }
            s1 = s0;
        } else if (op == B_BI_P) {
            // p = r + beta (p - omega v); y = D^-1 p (:380-382)
            const float be = float(a.st[e].beta), om = float(a.st[e].omega);
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                const float pi = a.r[off + i] + be * (a.p[off + i] - om * a.Ap[off + i]);
                a.p[off + i] = pi;
                a.y[off + i] = a.invD[off + i] * pi;
            }
            continue;
        } else if (op == B_BI_S) {
            // s = r - alpha v (in r); z = D^-1 s; partials s.s (:384-386)
            const float al = float(a.st[e].alpha);
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                const float si = a.r[off + i] - al * a.Ap[off + i];
                a.r[off + i] = si;
                a.z[off + i] = a.invD[off + i] * si;
                s0 += double(si) * double(si);
            }
        } else { // B_BI_UPDATE: x += alpha y + omega z; r = s - omega t (:393-397)
            const BEq& q = a.st[e];
            const double al = q.alpha;
            if (q.sconv) {
                // s met the tolerance: x += alpha y and stop (:387-390)
                for (int i = c0 + threadIdx.x; i < c1; i += kBlock)
                    a.x[off + i] += al * double(a.y[off + i]);
                continue;
            }
            const double om = q.omega;
            const float omf = float(om);
            for (int i = c0 + threadIdx.x; i < c1; i += kBlock) {
                a.x[off + i] += al * double(a.y[off + i]) + om * double(a.z[off + i]);
                const float ri = a.r[off + i] - omf * a.t[off + i];
                a.r[off + i] = ri;
                s0 += double(ri) * double(ri);
                s1 += double(a.rh[off + i]) * double(ri);
            }
        }
        S.add(0, e, s0);
        if (op != B_BNORM && op != B_BI_S)
            S.add(1, e, s1);
    }
    if (op != B_P && op != B_BI_P)
        S.store(a, (op == B_BNORM || op == B_BI_S) ? 1 : 2);
}

struct BFinArgs
{
    int E, G, max_iter, max_outer;
    double tol, inner_tol;
};

// One workgroup per equation: fold the G partials (fixed order) and update
// the equation's scalars.
__global__ __launch_bounds__(kBlock) void k_bfin(const double* __restrict__ partials, BEq* st,
                                                 int op, BFinArgs f)
{
    const int e = blockIdx.x;
    const int nv = (op == B_INIT || op == B_UPDATE || op == B_BI_INIT || op == B_BI_OMEGA ||
                    op == B_BI_UPDATE)
                       ? 2
                       : 1;
    double tot[2] = {0.0, 0.0};
    for (int v = 0; v < nv; ++v)
        for (int g = threadIdx.x; g < f.G; g += kBlock)
            tot[v] += partials[(size_t(v) * f.E + e) * f.G + g];
    __shared__ double lds[2 * kWaves];
    block_sum<2>(tot, lds);
    if (threadIdx.x != 0)
        return;
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

// d[k] = src[map[k]] (map null: identity) in FP64 and FP32
__global__ __launch_bounds__(kBlock) void k_bconvert(int64_t m, const int* __restrict__ map,
                                                     const double* __restrict__ src,
                                                     double* __restrict__ d64,
                                                     float* __restrict__ d32)
{
    const int64_t stride = int64_t(gridDim.x) * kBlock;
    for (int64_t i = int64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
        const double v = src[map ? map[i] : i];
        d64[i] = v;
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

} // namespace mi

using namespace mi;

struct mi_batch
{
    std::shared_ptr<Pattern> src; // the caller's pattern
    std::shared_ptr<Pattern> pat; // working pattern: src, or src reordered (RCM)
    DevBuf<int> d_perm, d_vmap;   // perm[new] = old; vmap[k] = source entry of entry k
    bool reordered = false;
    int E = 0, G = 8;
    int64_t n = 0, nnz = 0;
    DevBuf<float> v32, invD, r, p, Ap;
    DevBuf<float> rh, y, z, t; // BiCGStab vectors, allocated on its first solve
    int graph_kind = -1;
    DevBuf<double> v64, b, x, rt, partials;
    DevBuf<BEq> st;
    BEq* h_st = nullptr;
    hipStream_t stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_chunk = 0;
    BFinArgs graph_f{}; // finalize arguments baked into the captured graph
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
    BArgs args() const
    {
        BArgs a{};
        a.n = int(n);
        a.E = E;
        a.nrb = pat->rb.nrb;
        a.nnz = nnz;
        a.blk = pat->rb.blk.p;
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
        a.rh = rh.p;
        a.y = y.p;
        a.z = z.p;
        a.t = t.p;
        a.st = st.p;
        a.partials = partials.p;
        return a;
    }
};

namespace
{

constexpr int kBatchMaxGrid = 4096;

int fin(mi_batch* B, int op, const BFinArgs& f, Launcher& L)
{
    return L(k_bfin, unsigned(B->E), kBlock, static_cast<const double*>(B->partials.p), B->st.p,
             op, f);
}

int fin(mi_batch* B, int op, const BFinArgs& f)
{
    Launcher L{B->stream};
    return fin(B, op, f, L);
}

// One CG iteration for every iterating equation (five launches).
int one_iteration(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    MI_TRY(L(k_bspmv<B_ALPHA>, B->G, kBlock, a));
    MI_TRY(fin(B, B_ALPHA, f, L));
    MI_TRY(L(k_bvec, B->G, kBlock, a, int(B_UPDATE)));
    MI_TRY(fin(B, B_UPDATE, f, L));
    return L(k_bvec, B->G, kBlock, a, int(B_P));
}

// One BiCGStab iteration for every iterating equation (nine launches).
int one_iteration_bicg(mi_batch* B, const BArgs& a, const BFinArgs& f, Launcher& L)
{
    MI_TRY(L(k_bvec, B->G, kBlock, a, int(B_BI_P)));
    MI_TRY(L(k_bspmv<B_BI_ALPHA>, B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_ALPHA, f, L));
    MI_TRY(L(k_bvec, B->G, kBlock, a, int(B_BI_S)));
    MI_TRY(fin(B, B_BI_S, f, L));
    MI_TRY(L(k_bspmv<B_BI_T>, B->G, kBlock, a));
    MI_TRY(fin(B, B_BI_OMEGA, f, L));
    MI_TRY(L(k_bvec, B->G, kBlock, a, int(B_BI_UPDATE)));
    return fin(B, B_BI_UPDATE, f, L);
}

// Autotuner: workgroups of the persistent batched launches, timed on this
// pattern with every equation active.
int autotune(mi_batch* B)
{
    B->G = 8;
    if (B->n == 0)
        return MI_OK;
    hipEvent_t e0, e1;
    MI_HIP(hipEventCreate(&e0));
    MI_HIP(hipEventCreate(&e1));
    double best = 1e300;
    BArgs a = B->args();
    a.st = nullptr; // all equations active
    for (int g : {512, 1024, 2048, 4096}) {
        if (g > kBatchMaxGrid)
            break;
        k_bspmv<B_ALPHA><<<g, kBlock, 0, B->stream>>>(a);
        MI_HIP(hipGetLastError());
        MI_HIP(hipEventRecord(e0, B->stream));
        const int reps = 3;
        for (int r = 0; r < reps; ++r)
            k_bspmv<B_ALPHA><<<g, kBlock, 0, B->stream>>>(a);
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
    f.G = B->G;
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
    const bool bicg = c.kind == MI_SOLVER_BICGSTAB;
    if (bicg && !B->rh.p) {
        const size_t nE = size_t(n) * size_t(E);
        for (DevBuf<float>* v : {&B->rh, &B->y, &B->z, &B->t})
            MI_TRY(v->alloc(nE));
    }
    const BArgs a = B->args();
    k_bvec<<<B->G, kBlock, 0, s>>>(a, B_BNORM);
    MI_HIP(hipGetLastError());
    MI_TRY(fin(B, B_BNORM, f));
    const int chunk = c.check_interval > 0 ? c.check_interval : 8;
    const bool use_graph = c.use_graph != 0 && graphs_enabled();
    auto iteration = [&](Launcher& L) {
        return bicg ? one_iteration_bicg(B, a, f, L) : one_iteration(B, a, f, L);
    };
    if (use_graph && (!B->graph || B->graph_chunk != chunk || B->graph_kind != c.kind ||
                      std::memcmp(&B->graph_f, &f, sizeof(BFinArgs)) != 0)) {
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
        B->graph_f = f;
    }
    for (;;) {
        k_btrue<<<B->G, kBlock, 0, s>>>(a);
        MI_HIP(hipGetLastError());
        MI_TRY(fin(B, B_TRUE, f));
        MI_HIP(hipMemcpyAsync(B->h_st, B->st.p, sizeof(BEq) * size_t(E), hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        bool all_fin = true;
        for (int e = 0; e < E; ++e)
            all_fin = all_fin && B->h_st[e].fin;
        if (all_fin)
            break;
        k_bvec<<<B->G, kBlock, 0, s>>>(a, bicg ? B_BI_INIT : B_INIT);
        MI_HIP(hipGetLastError());
        MI_TRY(fin(B, bicg ? B_BI_INIT : B_INIT, f));
        for (;;) {
            if (use_graph) {
                MI_HIP(hipGraphLaunch(B->graph, s));
            } else {
                Launcher L{s};
                for (int it = 0; it < chunk; ++it)
                    MI_TRY(iteration(L));
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
    const size_t nE = size_t(std::max<int64_t>(B->n, 1)) * size_t(n_eq);
    const size_t zE = size_t(std::max<int64_t>(B->nnz, 1)) * size_t(n_eq);
    MI_TRY(B->v32.alloc(zE));
    MI_TRY(B->v64.alloc(zE));
    for (DevBuf<float>* v : {&B->invD, &B->r, &B->p, &B->Ap})
        MI_TRY(v->alloc(nE));
    for (DevBuf<double>* v : {&B->b, &B->x, &B->rt})
        MI_TRY(v->alloc(nE));
    MI_TRY(B->partials.alloc(size_t(2) * kBatchMaxGrid * size_t(n_eq)));
    MI_TRY(B->st.alloc(size_t(n_eq)));
    MI_HIP(hipMemsetAsync(B->v32.p, 0, zE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->p.p, 0, nE * sizeof(float), B->stream));
    MI_HIP(hipMemsetAsync(B->st.p, 0, sizeof(BEq) * size_t(n_eq), B->stream));
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
    if (B->nnz)
        k_bconvert<<<reduce_grid(B->nnz, 4), kBlock, 0, s>>>(
            B->nnz, B->reordered ? B->d_vmap.p : nullptr, A->val.p, v64,
            B->v32.p + size_t(eq) * size_t(B->nnz));
    if (B->n)
        k_binvdiag<<<grid_for(B->n, kBlock), kBlock, 0, s>>>(
            int(B->n), B->pat->diag.p, v64, B->invD.p + size_t(eq) * size_t(B->n));
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
    if (rows_per_block) // rows of one persistent workgroup
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
