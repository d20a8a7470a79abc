// amg_numeric.hip — numeric re-setup of the SA-AMG hierarchy on the device.
//
// AMGPreconditioner::setup (src/solver/AMGPreconditioner.cpp:248-326) builds
// the whole hierarchy again at every solve (src/solver/LinearSolvers.cpp:422).
// In a SIMPLE/PISO loop the pressure matrix keeps its sparsity pattern and
// only its values change, so after one full host setup (amg_setup.cpp) this
// file recomputes the hierarchy's values on the device with that setup's
// aggregates, strong-coupling masks and sparsity patterns:
//   lambda_max  20 power iterations on D^-1/2 A D^-1/2 from the host's start
//               vector, Rayleigh quotient x 1.05 (amg_lambda_max)
//   P           (I - omega D_F^-1 A_F) T on P's pattern, omega = omega_scale /
//               lambda_max, D_F = diagonal + weak couplings of the row
//               (amg_smoothed_prolongator)
//   R           P^T, a gather through a fixed transpose map (csr_transpose)
//   A_c         R (A P): two row-wise products into the fixed patterns of A P
//               and A_c (csr_multiply), each as a list of term pairs per
//               target entry recorded with the patterns (k_num_pairs)
// Every value is formed with the host setup's operations in the host's order,
// so the hierarchy equals a host rebuild with the same aggregates up to the
// rounding of the power iteration's dot products (the device sums them in
// another order).  Entries the host dropped as exact zeros stay outside the
// patterns.  All work is row-parallel (one thread per row, no atomics), so
// the results do not depend on the launch.
#include "amg.hpp"
#include "host_threads.hpp"

#include <algorithm>
#include <climits>
#include <cmath>

namespace mi
{

struct NumLevel
{
    int n = 0, nc = 0;
    int64_t nnz = 0;
    const int* arp = nullptr; // A_l's pattern on the device
    const int* aci = nullptr;
    DevBuf<int> arp_own, aci_own; // levels >= 1
    DevBuf<double> av;            // A_l's values, CSR order
    DevBuf<unsigned char> strong; // per entry of A_l (all levels but the coarsest)
    DevBuf<int> agg;              // per row of A_l
    DevBuf<double> dF, s, x0, y0, y1, part;
    unsigned pgrid = 1;           // blocks of the power-iteration launches
    DevBuf<int> prp, pci;         // P_l
    DevBuf<double> pv;
    DevBuf<int> rrp, rci, rmap;   // R_l = P_l^T: rv[t] = pv[rmap[t]]
    DevBuf<double> rv;
    DevBuf<int> aprp, apci;       // A_l P_l
    DevBuf<double> apv;
    // Term lists of the products A P and R (A P) (product_pairs): offsets
    // per target entry and (A position, B position) pairs; empty: the
    // row-wise kernels instead
    DevBuf<int> ap_off, rap_off;
    DevBuf<int2> ap_pab, rap_pab;
    int64_t p_nnz = 0, r_nnz = 0;
    int lanes = 1;                // lanes per row of the power iteration
    bool wave_rows = false;       // P and A P a wavefront per row (few rows), else a thread
    bool ap_wave = false;         // A P and R (A P) by k_num_product_wave
    bool rap_wave = false;
};

namespace
{
int max_row(const std::vector<int>& rp)
{
    int m = 0;
    for (size_t i = 0; i + 1 < rp.size(); ++i)
        m = std::max(m, rp[i + 1] - rp[i]);
    return m;
}

// Term lists of C = A B on C's fixed pattern (C's rows sorted): for every
// entry of C in CSR order, the A and B positions of its terms a_ik b_kj in
// the order the row-wise product adds them (k in A's row order); terms whose
// column is not in C's pattern are dropped, as the row-wise product drops
// them.  Returns false when the lists would hold 2^31 pairs or more.
bool product_pairs(const HCsr& A, const HCsr& B, const HCsr& C, std::vector<int>& off,
                   std::vector<int2>& pab)
{
    const int n = A.nr;
    const int64_t m = C.nnz();
    std::vector<int64_t> cnt(size_t(m) + 1, 0);
    const int T = host_threads();
    // pass 1: terms per entry of C
    host_parallel_for(n, T, [&](int b, int e, int) {
        std::vector<int> mark(size_t(std::max(C.nc, 1)), -1);
        for (int i = b; i < e; ++i) {
            for (int t = C.rp[size_t(i)]; t < C.rp[size_t(i) + 1]; ++t)
                mark[size_t(C.ci[size_t(t)])] = t;
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                const int r = A.ci[size_t(k)];
                for (int q = B.rp[size_t(r)]; q < B.rp[size_t(r) + 1]; ++q) {
                    const int t = mark[size_t(B.ci[size_t(q)])];
                    if (t >= 0)
                        ++cnt[size_t(t) + 1];
                }
            }
            for (int t = C.rp[size_t(i)]; t < C.rp[size_t(i) + 1]; ++t)
                mark[size_t(C.ci[size_t(t)])] = -1;
        }
    });
    for (int64_t t = 0; t < m; ++t)
        cnt[size_t(t) + 1] += cnt[size_t(t)];
    if (cnt[size_t(m)] >= INT_MAX)
        return false;
    off.assign(size_t(m) + 1, 0);
    for (int64_t t = 0; t <= m; ++t)
        off[size_t(t)] = int(cnt[size_t(t)]);
    pab.assign(size_t(std::max<int64_t>(cnt[size_t(m)], 1)), int2{0, 0});
    // pass 2: the pairs, each entry's in term order
    host_parallel_for(n, T, [&](int b, int e, int) {
        std::vector<int> mark(size_t(std::max(C.nc, 1)), -1);
        std::vector<int> cur;
        for (int i = b; i < e; ++i) {
            const int t0 = C.rp[size_t(i)], t1 = C.rp[size_t(i) + 1];
            cur.assign(off.begin() + t0, off.begin() + t1);
            for (int t = t0; t < t1; ++t)
                mark[size_t(C.ci[size_t(t)])] = t;
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                const int r = A.ci[size_t(k)];
                for (int q = B.rp[size_t(r)]; q < B.rp[size_t(r) + 1]; ++q) {
                    const int t = mark[size_t(B.ci[size_t(q)])];
                    if (t >= 0) {
                        const int w = cur[size_t(t - t0)]++;
                        pab[size_t(w)] = int2{k, q};
                    }
                }
            }
            for (int t = t0; t < t1; ++t)
                mark[size_t(C.ci[size_t(t)])] = -1;
        }
    });
    return true;
}
} // namespace

AmgNumeric::AmgNumeric() = default;
AmgNumeric::~AmgNumeric() = default;

// ---------------------------------------------------------------- kernels
// Row i: the diagonal d (sum of the row's diagonal entries), the filtered
// diagonal dF = d + the row's weak off-diagonal entries (strong == nullptr:
// dF = d), and the power iteration's scaling s = 1 / sqrt|d| (1 when d = 0).
__global__ __launch_bounds__(kBlock) void k_num_diag(int n, const int* __restrict__ rp,
                                                     const int* __restrict__ ci,
                                                     const double* __restrict__ v,
                                                     const unsigned char* __restrict__ strong,
                                                     double* __restrict__ dF,
                                                     double* __restrict__ s)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    double d = 0.0;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        if (ci[k] == i)
            d += v[k];
    double f = d;
    if (strong) {
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            if (ci[k] != i && !strong[k])
                f += v[k];
        if (fabs(f) <= kAmgFilterCancel * fabs(d))
            f = 0.0; // the couplings, all weak, cancel the diagonal
    }
    dF[i] = f;
    s[i] = d != 0.0 ? 1.0 / sqrt(fabs(d)) : 1.0;
}

// One power-iteration step y = S A S x, x = x0 (first step) or the previous
// y times 1/||y|| (scal[1]); block partials of x.y, x.x and y.y.  G lanes
// per row (a power of two) sum the row's entries and fold them by shuffles.
template <int G>
__global__ __launch_bounds__(kBlock) void k_num_power(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ ci,
                                                      const double* __restrict__ v,
                                                      const double* __restrict__ s,
                                                      const double* __restrict__ xin,
                                                      const double* __restrict__ scal,
                                                      double* __restrict__ y,
                                                      double* __restrict__ part)
{
#pragma clang fp contract(off)
    const int sub = int(threadIdx.x) % G;
    const double inv = scal ? scal[1] : 1.0;
    double acc3[3] = {0.0, 0.0, 0.0};
    // rows in a grid-stride loop (a bounded grid keeps the partials few)
    const int64_t stride = int64_t(gridDim.x) * (kBlock / G);
    for (int64_t i0 = int64_t(blockIdx.x) * (kBlock / G); i0 < n; i0 += stride) {
        const int i = int(i0) + int(threadIdx.x) / G;
        double acc = 0.0;
        if (i < n)
            for (int k = rp[i] + sub; k < rp[i + 1]; k += G) {
                const int j = ci[k];
                const double xj = scal ? xin[j] * inv : xin[j];
                acc += v[k] * s[j] * xj;
            }
#pragma unroll
        for (int off = G / 2; off >= 1; off >>= 1)
            acc += __shfl_xor(acc, off, G);
        if (i < n && sub == 0) {
            const double yi = s[i] * acc;
            const double xi = scal ? xin[i] * inv : xin[i];
            y[i] = yi;
            acc3[0] += xi * yi;
            acc3[1] += xi * xi;
            acc3[2] += yi * yi;
        }
    }
    block_partials<3>(acc3, part);
}

// lambda = |x.y / x.x|, 1/||y||; on the last step also lambda_max = 1.05 lambda.
__global__ __launch_bounds__(kBlock) void k_num_power_fin(const double* __restrict__ part, int G,
                                                          int last, double* __restrict__ scal)
{
    double t[3];
    fold_partials<3>(part, G, t);
    if (threadIdx.x == 0) {
        const double lam = t[1] > 0 ? fabs(t[0] / t[1]) : 1.0;
        scal[0] = lam;
        scal[1] = t[2] > 0 ? 1.0 / sqrt(t[2]) : 1.0;
        if (last)
            scal[2] = lam * 1.05;
    }
}

// Row i of P = (I - omega D_F^-1 A_F) T on P's fixed pattern, W lanes per row
// (1 or a wavefront), one lane per entry of P's row (column c) at a time:
// the entry starts at 1 when c is the row's own aggregate, then takes -w dF
// (diagonal) or -w a_ij (strong coupling) for every entry of A's row whose
// column lies in aggregate c, in A's row order, w = omega / dF.  truncated:
// the pattern holds the entries a truncated setup kept, and the row is
// scaled by its complete sum (the same additions, in the same order, as the
// host setup's) over the sum of the kept entries (in column order).
template <int W>
__global__ __launch_bounds__(kBlock) void k_num_prolongator(
    int n, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ v,
    const unsigned char* __restrict__ strong, const int* __restrict__ agg,
    const double* __restrict__ dF, const double* __restrict__ lmax, double omega_scale,
    const int* __restrict__ prp, const int* __restrict__ pci, double* __restrict__ pv,
    int truncated)
{
#pragma clang fp contract(off)
    const int lane = int(threadIdx.x) % W;
    const int i = int(blockIdx.x) * (kBlock / W) + int(threadIdx.x) / W;
    if (i >= n)
        return;
    const double omega = omega_scale / lmax[0];
    const double df = dF[i];
    const double w = df != 0.0 ? omega / df : 0.0;
    const int own = agg[i];
    for (int q = prp[i] + lane; q < prp[i + 1]; q += W) {
        const int c = pci[q];
        double acc = 0.0;
        if (own == c)
            acc += 1.0;
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (agg[j] != c)
                continue;
            if (j == i)
                acc += -w * df;
            else if (strong[k])
                acc += -w * v[k];
        }
        pv[q] = acc;
    }
    if (truncated) {
        // one lane sums the row (W > 1: the row's lanes are one wavefront,
        // whose stores the fence makes visible to that lane) and every lane
        // scales the entries it computed
        if (W > 1)
            __threadfence_block();
        double scale = 1.0;
        if (lane == 0) {
            double full = own >= 0 ? 1.0 : 0.0;
            for (int k = rp[i]; k < rp[i + 1]; ++k) {
                const int j = ci[k];
                if (agg[j] < 0)
                    continue;
                if (j == i)
                    full += -w * df;
                else if (strong[k])
                    full += -w * v[k];
            }
            double kept = 0.0;
            for (int q = prp[i]; q < prp[i + 1]; ++q)
                kept += pv[q];
            scale = kept != 0.0 ? full / kept : 1.0;
        }
        if (W > 1)
            scale = __shfl(scale, 0, W);
        for (int q = prp[i] + lane; q < prp[i + 1]; q += W)
            pv[q] = pv[q] * scale;
    }
}

__global__ __launch_bounds__(kBlock) void k_num_gather(int64_t m, const int* __restrict__ map,
                                                       const double* __restrict__ src,
                                                       double* __restrict__ dst)
{
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < m;
         t += int64_t(gridDim.x) * kBlock)
        dst[t] = src[map[t]];
}

// First position in the sorted columns ci[b, e) whose column is >= c.
__device__ __forceinline__ int find_col(const int* __restrict__ ci, int b, int e, int c)
{
    while (b < e) {
        const int m = (b + e) >> 1;
        if (ci[m] < c)
            b = m + 1;
        else
            e = m;
    }
    return b;
}

// Row i of C = A B into C's fixed pattern: for every entry a_ik of A's row in
// order, for every entry b_kj of B's row k, c_ij += a_ik b_kj -- for each
// c_ij the terms in A's row order, as the host's row-wise product adds them.
// One thread per row.
__global__ __launch_bounds__(kBlock) void k_num_product(
    int n, const int* __restrict__ arp, const int* __restrict__ aci, const double* __restrict__ av,
    const int* __restrict__ brp, const int* __restrict__ bci, const double* __restrict__ bv,
    const int* __restrict__ crp, const int* __restrict__ cci, double* __restrict__ cv)
{
#pragma clang fp contract(off)
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    const int c0 = crp[i], c1 = crp[i + 1];
    for (int q = c0; q < c1; ++q)
        cv[q] = 0.0;
    for (int k = arp[i]; k < arp[i + 1]; ++k) {
        const int r = aci[k];
        const double a = av[k];
        for (int q = brp[r]; q < brp[r + 1]; ++q) {
            const int c = bci[q];
            const int pos = find_col(cci, c0, c1, c);
            if (pos < c1 && cci[pos] == c)
                cv[pos] += a * bv[q];
        }
    }
}

// Orders the LDS accesses of one wavefront's lanes (no workgroup barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The same product with one wavefront per row of C (rows of at most
// kProductRow entries): the row's columns and sums sit in LDS; for each
// entry a_ik in A's row order the lanes take the entries of B's row k (all
// in different columns of C) at once, so every c_ij receives its terms in
// the same order as above.
constexpr int kProductRow = 512;

__global__ __launch_bounds__(kBlock) void k_num_product_wave(
    int n, const int* __restrict__ arp, const int* __restrict__ aci, const double* __restrict__ av,
    const int* __restrict__ brp, const int* __restrict__ bci, const double* __restrict__ bv,
    const int* __restrict__ crp, const int* __restrict__ cci, double* __restrict__ cv)
{
#pragma clang fp contract(off)
    __shared__ int cols[kBlock / kWave][kProductRow];
    __shared__ double sums[kBlock / kWave][kProductRow];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = int(threadIdx.x) / kWave;
    const int i = int(blockIdx.x) * (kBlock / kWave) + w;
    if (i >= n)
        return;
    const int c0 = crp[i];
    const int L = crp[i + 1] - c0;
    int* cw = cols[w];
    double* sw = sums[w];
    for (int t = lane; t < L; t += kWave) {
        cw[t] = cci[c0 + t];
        sw[t] = 0.0;
    }
    wave_sync();
    for (int k = arp[i]; k < arp[i + 1]; ++k) {
        const int r = aci[k];
        const double a = av[k];
        for (int q = brp[r] + lane; q < brp[r + 1]; q += kWave) {
            const int c = bci[q];
            int lo = 0, hi = L;
            while (lo < hi) {
                const int m = (lo + hi) >> 1;
                if (cw[m] < c)
                    lo = m + 1;
                else
                    hi = m;
            }
            if (lo < L && cw[lo] == c)
                sw[lo] += a * bv[q];
        }
        wave_sync();
    }
    for (int t = lane; t < L; t += kWave)
        cv[c0 + t] = sw[t];
}

// C = A B from the term lists of product_pairs: entry t of C sums its terms
// a[pab.x] b[pab.y] in list order from 0 -- the additions of k_num_product
// in the same order, one thread per entry of C instead of a search per term.
// Four terms' loads are issued before their four additions.
__global__ __launch_bounds__(kBlock) void k_num_pairs(int64_t m, const int* __restrict__ off,
                                                      const int2* __restrict__ pab,
                                                      const double* __restrict__ av,
                                                      const double* __restrict__ bv,
                                                      double* __restrict__ cv)
{
#pragma clang fp contract(off)
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < m;
         t += int64_t(gridDim.x) * kBlock) {
        double sum = 0.0;
        int u = off[t];
        const int ue = off[t + 1];
        for (; u + 4 <= ue; u += 4) {
            const int2 p0 = pab[u], p1 = pab[u + 1], p2 = pab[u + 2], p3 = pab[u + 3];
            const double a0 = av[p0.x], a1 = av[p1.x], a2 = av[p2.x], a3 = av[p3.x];
            const double b0 = bv[p0.y], b1 = bv[p1.y], b2 = bv[p2.y], b3 = bv[p3.y];
            sum += a0 * b0;
            sum += a1 * b1;
            sum += a2 * b2;
            sum += a3 * b3;
        }
        for (; u < ue; ++u) {
            const int2 p = pab[u];
            sum += av[p.x] * bv[p.y];
        }
        cv[t] = sum;
    }
}

__global__ __launch_bounds__(kBlock) void k_num_zero(int64_t m, double* __restrict__ x)
{
    for (int64_t t = int64_t(blockIdx.x) * kBlock + threadIdx.x; t < m;
         t += int64_t(gridDim.x) * kBlock)
        x[t] = 0.0;
}

// dense[i][j] += a_ij over the CSR entries (dense zeroed first).
__global__ __launch_bounds__(kBlock) void k_num_dense(int n, const int* __restrict__ rp,
                                                      const int* __restrict__ ci,
                                                      const double* __restrict__ v,
                                                      double* __restrict__ dense)
{
    const int i = int(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n)
        return;
    for (int k = rp[i]; k < rp[i + 1]; ++k)
        dense[size_t(i) * size_t(n) + size_t(ci[k])] += v[k];
}

// --------------------------------------------------------------- record
namespace
{

bool rows_sorted(const HCsr& h)
{
    for (int i = 0; i < h.nr; ++i)
        for (int k = h.rp[size_t(i)] + 1; k < h.rp[size_t(i) + 1]; ++k)
            if (h.ci[size_t(k)] <= h.ci[size_t(k) - 1])
                return false;
    return true;
}

} // namespace

int AmgNumeric::record(const AmgHierarchyHost& H, const int* rp0, const int* ci0, hipStream_t s)
{
    const int nl = int(H.levels.size());
    L.clear();
    if (nl == 0)
        return set_error(MI_ERR_BAD_ARG, "empty hierarchy");
    L.resize(size_t(nl));
    for (int l = 0; l < nl; ++l) {
        const AmgLevelHost& h = H.levels[size_t(l)];
        NumLevel& d = L[size_t(l)];
        d.n = h.A.nr;
        d.nnz = h.A.nnz();
        if (l == 0) {
            d.arp = rp0;
            d.aci = ci0;
        } else {
            if (!rows_sorted(h.A))
                return set_error(MI_ERR_UNSUPPORTED, "level %d has unsorted rows", l);
            MI_TRY(d.arp_own.upload(h.A.rp.data(), h.A.rp.size(), s));
            MI_TRY(d.aci_own.upload(h.A.ci.data(), h.A.ci.size(), s));
            d.arp = d.arp_own.p;
            d.aci = d.aci_own.p;
        }
        MI_TRY(d.av.alloc(size_t(std::max<int64_t>(d.nnz, 1))));
        const size_t n1 = size_t(std::max(d.n, 1));
        std::vector<double> x0;
        amg_lambda_start(d.n, x0);
        MI_TRY(d.x0.upload(x0.data(), x0.size(), s));
        MI_HIP(hipStreamSynchronize(s)); // x0 goes out of scope
        for (DevBuf<double>* b : {&d.dF, &d.s, &d.y0, &d.y1})
            MI_TRY(b->alloc(n1));
        {
            const int64_t avg = d.n > 0 ? (d.nnz + d.n - 1) / d.n : 1;
            d.lanes = 1;
            while (d.lanes < 64 && d.lanes < avg)
                d.lanes *= 2;
        }
        d.pgrid = std::min<unsigned>(grid_for(int64_t(d.n) * d.lanes, kBlock), kMaxReduceBlocks);
        MI_TRY(d.part.alloc(size_t(d.pgrid) * 3));
        if (l + 1 == nl)
            break;
        if (h.agg.size() != size_t(d.n) || h.strong.size() != size_t(d.nnz))
            return set_error(MI_ERR_UNSUPPORTED, "level %d records no aggregates", l);
        const HCsr& next = H.levels[size_t(l) + 1].A;
        d.nc = next.nr;
        MI_TRY(d.strong.upload(h.strong.data(), h.strong.size(), s));
        MI_TRY(d.agg.upload(h.agg.data(), h.agg.size(), s));
        // P_l as built
        if (!rows_sorted(h.P))
            return set_error(MI_ERR_UNSUPPORTED, "prolongator %d has unsorted rows", l);
        d.p_nnz = h.P.nnz();
        MI_TRY(d.prp.upload(h.P.rp.data(), h.P.rp.size(), s));
        MI_TRY(d.pci.upload(h.P.ci.data(), h.P.ci.size(), s));
        MI_TRY(d.pv.alloc(size_t(std::max<int64_t>(d.p_nnz, 1))));
        // R_l = P_l^T in csr_transpose's order, and the map into P's values
        {
            const int nc = h.P.nc;
            std::vector<int> rrp(size_t(nc) + 1, 0), rci(size_t(d.p_nnz)), rmap(size_t(d.p_nnz));
            for (int c : h.P.ci)
                ++rrp[size_t(c) + 1];
            for (int c = 0; c < nc; ++c)
                rrp[size_t(c) + 1] += rrp[size_t(c)];
            std::vector<int> pos(rrp.begin(), rrp.end() - 1);
            for (int i = 0; i < h.P.nr; ++i)
                for (int k = h.P.rp[size_t(i)]; k < h.P.rp[size_t(i) + 1]; ++k) {
                    const int w = pos[size_t(h.P.ci[size_t(k)])]++;
                    rci[size_t(w)] = i;
                    rmap[size_t(w)] = k;
                }
            if (rrp != h.R.rp || rci != h.R.ci)
                return set_error(MI_ERR_UNSUPPORTED, "level %d: R is not the transpose of P", l);
            d.r_nnz = d.p_nnz;
            MI_TRY(d.rrp.upload(rrp.data(), rrp.size(), s));
            MI_TRY(d.rci.upload(rci.data(), rci.size(), s));
            MI_TRY(d.rmap.upload(rmap.data(), rmap.size(), s));
            MI_TRY(d.rv.alloc(size_t(std::max<int64_t>(d.p_nnz, 1))));
            MI_HIP(hipStreamSynchronize(s)); // rrp, rci and rmap go out of scope
        }
        // the pattern of A_l P_l (rows sorted)
        {
            HCsr AP;
            csr_multiply(h.A, h.P, AP, 0);
            // Below 64K rows a thread per row leaves the device mostly idle
            // (measured at 96^3: level 2, 3584 rows, P 128 -> 33 us and A P
            // 722 -> 70 us a wavefront per row; level 0, 884,736 rows, P
            // 156 vs 911 us and A P 858 vs 1531 us a thread per row).  R (A P)
            // has long rows at every level: a wavefront per row (level 0
            // 2115 -> 962 us, level 1 7529 -> 307 us).
            d.wave_rows = d.n < 65536;
            d.ap_wave = d.wave_rows && max_row(AP.rp) <= kProductRow;
            d.rap_wave = max_row(next.rp) <= kProductRow;
            MI_TRY(d.aprp.upload(AP.rp.data(), AP.rp.size(), s));
            MI_TRY(d.apci.upload(AP.ci.data(), AP.ci.size(), s));
            MI_TRY(d.apv.alloc(size_t(std::max<int64_t>(AP.nnz(), 1))));
            // the two products as term lists
            std::vector<int> off;
            std::vector<int2> pab;
            if (product_pairs(h.A, h.P, AP, off, pab)) {
                MI_TRY(d.ap_off.upload(off.data(), off.size(), s));
                MI_TRY(d.ap_pab.upload(pab.data(), pab.size(), s));
                MI_HIP(hipStreamSynchronize(s));
            }
            if (product_pairs(h.R, AP, next, off, pab)) {
                MI_TRY(d.rap_off.upload(off.data(), off.size(), s));
                MI_TRY(d.rap_pab.upload(pab.data(), pab.size(), s));
            }
            MI_HIP(hipStreamSynchronize(s)); // AP's host arrays go out of scope
        }
    }
    const int nc = L.back().n;
    MI_TRY(dense.alloc(size_t(std::max(nc, 1)) * size_t(std::max(nc, 1))));
    MI_TRY(scal.alloc(size_t(nl) * 3));
    lmax.assign(size_t(nl), 0.0);
    MI_HIP(hipStreamSynchronize(s));
    return MI_OK;
}

// ------------------------------------------------------------------ run
int AmgNumeric::run(const double* a0, double omega_scale, bool truncated, hipStream_t s)
{
    const int nl = int(L.size());
    if (nl == 0)
        return set_error(MI_ERR_BAD_ARG, "numeric re-setup before record");
    if (L[0].nnz > 0)
        MI_HIP(hipMemcpyAsync(L[0].av.p, a0, size_t(L[0].nnz) * sizeof(double),
                              hipMemcpyDeviceToDevice, s));
    for (int l = 0; l < nl; ++l) {
        NumLevel& d = L[size_t(l)];
        const bool last = l + 1 == nl;
        double* sc = scal.p + 3 * l;
        const unsigned g = grid_for(d.n, kBlock);
        if (d.n > 0) {
            k_num_diag<<<g, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p,
                                            last ? nullptr : d.strong.p, d.dF.p, d.s.p);
            // lambda_max(D^-1 A): 20 steps, as amg_lambda_max
            constexpr int kIters = 20;
            const double* xin = d.x0.p;
            for (int it = 0; it < kIters; ++it) {
                double* y = (it & 1) ? d.y1.p : d.y0.p;
                auto kern = d.lanes >= 64 ? k_num_power<64>
                          : d.lanes >= 32 ? k_num_power<32>
                          : d.lanes >= 16 ? k_num_power<16>
                          : d.lanes >= 8  ? k_num_power<8>
                          : d.lanes >= 4  ? k_num_power<4>
                          : d.lanes >= 2  ? k_num_power<2>
                                          : k_num_power<1>;
                kern<<<d.pgrid, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p, d.s.p, xin,
                                                it == 0 ? nullptr : sc, y, d.part.p);
                k_num_power_fin<<<1, kBlock, 0, s>>>(d.part.p, int(d.pgrid), it + 1 == kIters, sc);
                xin = y;
            }
        }
        MI_HIP(hipGetLastError());
        if (last)
            break;
        NumLevel& nx = L[size_t(l) + 1];
        const unsigned gw = grid_for(d.n, kBlock / kWave); // a wavefront per row
        if (d.wave_rows)
            k_num_prolongator<kWave><<<gw, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p, d.strong.p,
                                                           d.agg.p, d.dF.p, sc + 2, omega_scale,
                                                           d.prp.p, d.pci.p, d.pv.p,
                                                           truncated ? 1 : 0);
        else
            k_num_prolongator<1><<<g, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p, d.strong.p,
                                                      d.agg.p, d.dF.p, sc + 2, omega_scale,
                                                      d.prp.p, d.pci.p, d.pv.p, truncated ? 1 : 0);
        if (d.r_nnz > 0)
            k_num_gather<<<reduce_grid(d.r_nnz, 4), kBlock, 0, s>>>(d.r_nnz, d.rmap.p, d.pv.p,
                                                                   d.rv.p);
        const int64_t ap_nnz = int64_t(d.ap_off.n) - 1;
        const int64_t c_nnz = int64_t(d.rap_off.n) - 1;
        if (d.ap_off.n > 0 && ap_nnz > 0)
            k_num_pairs<<<reduce_grid(ap_nnz, 2), kBlock, 0, s>>>(ap_nnz, d.ap_off.p, d.ap_pab.p,
                                                                 d.av.p, d.pv.p, d.apv.p);
        else if (d.ap_wave)
            k_num_product_wave<<<gw, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p, d.prp.p, d.pci.p,
                                                     d.pv.p, d.aprp.p, d.apci.p, d.apv.p);
        else
            k_num_product<<<g, kBlock, 0, s>>>(d.n, d.arp, d.aci, d.av.p, d.prp.p, d.pci.p,
                                               d.pv.p, d.aprp.p, d.apci.p, d.apv.p);
        if (d.rap_off.n > 0 && c_nnz > 0)
            k_num_pairs<<<reduce_grid(c_nnz, 2), kBlock, 0, s>>>(c_nnz, d.rap_off.p, d.rap_pab.p,
                                                                d.rv.p, d.apv.p, nx.av.p);
        else if (d.rap_wave)
            k_num_product_wave<<<grid_for(nx.n, kBlock / kWave), kBlock, 0, s>>>(
                nx.n, d.rrp.p, d.rci.p, d.rv.p, d.aprp.p, d.apci.p, d.apv.p, nx.arp, nx.aci,
                nx.av.p);
        else
            k_num_product<<<grid_for(nx.n, kBlock), kBlock, 0, s>>>(
                nx.n, d.rrp.p, d.rci.p, d.rv.p, d.aprp.p, d.apci.p, d.apv.p, nx.arp, nx.aci,
                nx.av.p);
        MI_HIP(hipGetLastError());
    }
    // the coarsest operator, dense
    const NumLevel& c = L.back();
    const int64_t nn = int64_t(c.n) * c.n;
    if (nn > 0) {
        k_num_zero<<<reduce_grid(nn, 4), kBlock, 0, s>>>(nn, dense.p);
        k_num_dense<<<grid_for(c.n, kBlock), kBlock, 0, s>>>(c.n, c.arp, c.aci, c.av.p, dense.p);
        MI_HIP(hipGetLastError());
    }
    std::vector<double> h(size_t(nl) * 3);
    MI_HIP(hipMemcpyAsync(h.data(), scal.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    for (int l = 0; l < nl; ++l)
        lmax[size_t(l)] = L[size_t(l)].n > 0 ? h[size_t(l) * 3 + 2] : 1.05;
    return MI_OK;
}

LevelValues AmgNumeric::values(int l) const
{
    const NumLevel& d = L[size_t(l)];
    LevelValues v;
    v.A = d.av.p;
    if (size_t(l) + 1 < L.size()) {
        v.P = d.pv.p;
        v.R = d.rv.p;
    }
    v.lmax = lmax[size_t(l)];
    return v;
}

size_t AmgNumeric::max_nnz() const
{
    size_t m = 1;
    for (const NumLevel& d : L)
        m = std::max({m, size_t(d.nnz), size_t(d.p_nnz)});
    return m;
}

int AmgNumeric::download(AmgHierarchyHost& H, hipStream_t s) const
{
    if (H.levels.size() != L.size())
        return set_error(MI_ERR_BAD_ARG, "hierarchy does not match the recorded one");
    for (size_t l = 0; l < L.size(); ++l) {
        AmgLevelHost& h = H.levels[l];
        const NumLevel& d = L[l];
        if (d.nnz > 0)
            MI_HIP(hipMemcpyAsync(h.A.v.data(), d.av.p, size_t(d.nnz) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        if (l + 1 < L.size() && d.p_nnz > 0) {
            MI_HIP(hipMemcpyAsync(h.P.v.data(), d.pv.p, size_t(d.p_nnz) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
            MI_HIP(hipMemcpyAsync(h.R.v.data(), d.rv.p, size_t(d.r_nnz) * sizeof(double),
                                  hipMemcpyDeviceToHost, s));
        }
        h.lmax = lmax[l];
    }
    MI_HIP(hipStreamSynchronize(s));
    const AmgLevelHost& C = H.levels.back();
    const int nc = C.A.nr;
    H.coarse_dense.assign(size_t(nc) * size_t(nc), 0.0);
    for (int i = 0; i < nc; ++i)
        for (int k = C.A.rp[size_t(i)]; k < C.A.rp[size_t(i) + 1]; ++k)
            H.coarse_dense[size_t(i) * size_t(nc) + size_t(C.A.ci[size_t(k)])] += C.A.v[size_t(k)];
    return MI_OK;
}

} // namespace mi
