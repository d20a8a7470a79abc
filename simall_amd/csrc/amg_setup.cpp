// amg_setup.cpp — smoothed-aggregation AMG setup (Vanek, Mandel, Brezina 1996)
// on the host in FP64.
//
// The reference preconditioner is classical Ruge-Stueben AMG
// (src/solver/AMGPreconditioner.cpp:43-437) whose C/F splitting rescans all
// rows for every coarse point; it cannot build a 2M-row hierarchy.  This
// setup replaces it with smoothed aggregation, whose pieces map onto
// data-parallel GPU kernels in the solve phase:
//   strength    |a_ij| >= theta sqrt(|a_ii a_jj|)
//   aggregates  greedy root + strong neighbourhood, leftovers joined to the
//               most strongly connected aggregate, isolated rows as singletons
//   T           piecewise-constant tentative prolongator (near null space 1)
//   P           (I - omega D^-1 A) T, omega = 4/3 / lambda_max(D^-1 A)
//   A_c         R A P with R = P^T (symmetric V-cycle for CG)
// and the coarsest operator is inverted densely in FP64.
#include "amg_setup.hpp"
#include "host_threads.hpp"
#include "reorder.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <barrier>
#include <thread>

namespace mi
{

namespace
{

int thread_count(int requested)
{
    return host_threads(requested);
}

template <class F> void parallel_for(int n, int threads, F&& f)
{
    host_parallel_for(n, threads, f);
}

std::vector<double> diagonal(const HCsr& A)
{
    std::vector<double> d(size_t(A.nr), 0.0);
    parallel_for(A.nr, thread_count(0), [&](int b, int e, int) {
        for (int i = b; i < e; ++i)
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k)
                if (A.ci[size_t(k)] == i)
                    d[size_t(i)] += A.v[size_t(k)];
    });
    return d;
}

// Concatenates per-range CSR pieces (rows in range order) into one matrix.
struct RowPart
{
    std::vector<int> len, ci;
    std::vector<double> v;
};

void join_parts(std::vector<RowPart>& parts, int n, std::vector<int>& rp, std::vector<int>& ci,
                std::vector<double>* v)
{
    rp.assign(size_t(n) + 1, 0);
    size_t total = 0;
    int row = 0;
    for (const RowPart& p : parts)
        for (int l : p.len) {
            total += size_t(l);
            rp[size_t(++row)] = int(total);
        }
    ci.resize(total);
    if (v)
        v->resize(total);
    size_t off = 0;
    for (RowPart& p : parts) {
        std::copy(p.ci.begin(), p.ci.end(), ci.begin() + long(off));
        if (v)
            std::copy(p.v.begin(), p.v.end(), v->begin() + long(off));
        off += p.ci.size();
        p = RowPart{};
    }
}

} // namespace

void amg_strength(const HCsr& A, double theta, std::vector<int>& srp, std::vector<int>& sci)
{
    const std::vector<double> d = diagonal(A);
    const int threads = thread_count(0);
    std::vector<RowPart> parts(static_cast<size_t>(threads));
    parallel_for(A.nr, threads, [&](int b, int e, int t) {
        RowPart& P = parts[size_t(t)];
        for (int i = b; i < e; ++i) {
            const size_t before = P.ci.size();
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                const int j = A.ci[size_t(k)];
                if (j == i)
                    continue;
                const double a = std::abs(A.v[size_t(k)]);
                if (a > 0.0 && a >= theta * std::sqrt(std::abs(d[size_t(i)] * d[size_t(j)])))
                    P.ci.push_back(j);
            }
            P.len.push_back(int(P.ci.size() - before));
        }
    });
    join_parts(parts, A.nr, srp, sci, nullptr);
}

int amg_aggregate(int n, const std::vector<int>& srp, const std::vector<int>& sci, const HCsr& A,
                  std::vector<int>& agg)
{
    agg.assign(size_t(n), -1);
    int nagg = 0;
    // pass 1: a node whose strong neighbours are all free roots an aggregate
    for (int i = 0; i < n; ++i) {
        if (agg[size_t(i)] >= 0)
            continue;
        bool free = true;
        for (int k = srp[size_t(i)]; k < srp[size_t(i) + 1]; ++k)
            if (agg[size_t(sci[size_t(k)])] >= 0) {
                free = false;
                break;
            }
        if (!free)
            continue;
        agg[size_t(i)] = nagg;
        for (int k = srp[size_t(i)]; k < srp[size_t(i) + 1]; ++k)
            agg[size_t(sci[size_t(k)])] = nagg;
        ++nagg;
    }
    // pass 2: join the aggregate of the most strongly connected neighbour
    std::vector<int> joined(size_t(n), -1);
    for (int i = 0; i < n; ++i) {
        if (agg[size_t(i)] >= 0)
            continue;
        double best = -1.0;
        int target = -1;
        for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
            const int j = A.ci[size_t(k)];
            if (j == i || agg[size_t(j)] < 0)
                continue;
            bool strong = false;
            for (int s = srp[size_t(i)]; s < srp[size_t(i) + 1]; ++s)
                if (sci[size_t(s)] == j) {
                    strong = true;
                    break;
                }
            const double a = std::abs(A.v[size_t(k)]);
            if (strong && a > best) {
                best = a;
                target = agg[size_t(j)];
            }
        }
        joined[size_t(i)] = target;
    }
    for (int i = 0; i < n; ++i)
        if (agg[size_t(i)] < 0 && joined[size_t(i)] >= 0)
            agg[size_t(i)] = joined[size_t(i)];
    // pass 3: what is left forms new aggregates with its free strong neighbours
    for (int i = 0; i < n; ++i) {
        if (agg[size_t(i)] >= 0)
            continue;
        agg[size_t(i)] = nagg;
        for (int k = srp[size_t(i)]; k < srp[size_t(i) + 1]; ++k)
            if (agg[size_t(sci[size_t(k)])] < 0)
                agg[size_t(sci[size_t(k)])] = nagg;
        ++nagg;
    }
    return nagg;
}

void amg_lambda_start(int n, std::vector<double>& x)
{
    x.resize(size_t(n));
    uint64_t st = 0x9E3779B97F4A7C15ULL;
    for (int i = 0; i < n; ++i) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        x[size_t(i)] = double(st >> 11) * (1.0 / 9007199254740992.0) - 0.5;
    }
}

double amg_lambda_max(const HCsr& A, int iters)
{
    // Power iteration on the symmetric scaling D^-1/2 A D^-1/2 (same spectrum
    // as D^-1 A) with a Rayleigh quotient estimate.
    const int n = A.nr;
    if (n == 0)
        return 1.0;
    const std::vector<double> d = diagonal(A);
    std::vector<double> s(static_cast<size_t>(n)), x(static_cast<size_t>(n)), y(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        s[size_t(i)] = d[size_t(i)] != 0.0 ? 1.0 / std::sqrt(std::abs(d[size_t(i)])) : 1.0;
    amg_lambda_start(n, x);
    const int threads = thread_count(0);
    double lam = 1.0;
    for (int it = 0; it < iters; ++it) {
        parallel_for(n, threads, [&](int b, int e, int) {
            for (int i = b; i < e; ++i) {
                double acc = 0.0;
                for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                    const int j = A.ci[size_t(k)];
                    acc += A.v[size_t(k)] * s[size_t(j)] * x[size_t(j)];
                }
                y[size_t(i)] = s[size_t(i)] * acc;
            }
        });
        // dot products over fixed 64K-row chunks added in chunk order: the
        // estimate does not depend on the number of threads
        constexpr int kChunk = 1 << 16;
        const int nch = (n + kChunk - 1) / kChunk;
        std::vector<double> part(size_t(nch) * 3, 0.0);
        host_parallel_for(nch, std::min(threads, nch), [&](int b, int e, int) {
            for (int c = b; c < e; ++c) {
                double xy = 0.0, xx = 0.0, yy = 0.0;
                for (int i = c * kChunk; i < std::min(n, (c + 1) * kChunk); ++i) {
                    xy += x[size_t(i)] * y[size_t(i)];
                    xx += x[size_t(i)] * x[size_t(i)];
                    yy += y[size_t(i)] * y[size_t(i)];
                }
                part[size_t(c) * 3] = xy;
                part[size_t(c) * 3 + 1] = xx;
                part[size_t(c) * 3 + 2] = yy;
            }
        }, 2);
        double xy = 0.0, xx = 0.0, yy = 0.0;
        for (int c = 0; c < nch; ++c) {
            xy += part[size_t(c) * 3];
            xx += part[size_t(c) * 3 + 1];
            yy += part[size_t(c) * 3 + 2];
        }
        lam = xx > 0 ? std::abs(xy / xx) : 1.0;
        const double inv = yy > 0 ? 1.0 / std::sqrt(yy) : 1.0;
        parallel_for(n, threads, [&](int b, int e, int) {
            for (int i = b; i < e; ++i)
                x[size_t(i)] = y[size_t(i)] * inv;
        });
    }
    return lam;
}

void amg_smoothed_prolongator(const HCsr& A, const std::vector<int>& agg, int nagg, double omega,
                              HCsr& P, const std::vector<int>* srp, const std::vector<int>* sci,
                              double truncation)
{
    // With a strength pattern, the smoothing uses the filtered operator A_F:
    // weak off-diagonal entries are dropped and added to the diagonal (row
    // sums, hence A_F 1 = A 1, are preserved), so P spreads only along strong
    // couplings and the Galerkin operators stay sparse.
    const int n = A.nr;
    const std::vector<double> d = diagonal(A);
    P.nr = n;
    P.nc = nagg;
    const int threads = thread_count(0);
    std::vector<RowPart> parts(static_cast<size_t>(threads));
    parallel_for(n, threads, [&](int rb, int re, int t) {
    RowPart& out = parts[size_t(t)];
    std::vector<int> cols;
    std::vector<double> acc(static_cast<size_t>(nagg), 0.0);
    std::vector<char> used(static_cast<size_t>(nagg), 0);
    std::vector<char> strong(static_cast<size_t>(n), 0);
    for (int i = rb; i < re; ++i) {
        cols.clear();
        double full = 0.0; // the complete row's sum: its additions in order
        auto add = [&](int c, double v) {
            if (!used[size_t(c)]) {
                used[size_t(c)] = 1;
                acc[size_t(c)] = 0.0;
                cols.push_back(c);
            }
            acc[size_t(c)] += v;
            full += v;
        };
        double dF = d[size_t(i)];
        if (srp) {
            for (int s = (*srp)[size_t(i)]; s < (*srp)[size_t(i) + 1]; ++s)
                strong[size_t((*sci)[size_t(s)])] = 1;
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                const int j = A.ci[size_t(k)];
                if (j != i && !strong[size_t(j)])
                    dF += A.v[size_t(k)];
            }
            if (std::fabs(dF) <= kAmgFilterCancel * std::fabs(d[size_t(i)]))
                dF = 0.0;
        }
        if (agg[size_t(i)] >= 0)
            add(agg[size_t(i)], 1.0);
        const double w = dF != 0.0 ? omega / dF : 0.0;
        for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
            const int j = A.ci[size_t(k)];
            const int a = agg[size_t(j)];
            if (a < 0)
                continue;
            if (j == i)
                add(a, -w * dF);
            else if (!srp || strong[size_t(j)])
                add(a, -w * A.v[size_t(k)]);
        }
        if (srp)
            for (int s = (*srp)[size_t(i)]; s < (*srp)[size_t(i) + 1]; ++s)
                strong[size_t((*sci)[size_t(s)])] = 0;
        std::sort(cols.begin(), cols.end());
        // truncation: entries below truncation x the row's largest magnitude
        // are dropped and the kept ones scaled to the complete row's sum
        double cut = 0.0, scale = 1.0;
        if (truncation > 0.0) {
            double rmax = 0.0;
            for (int c : cols)
                rmax = std::max(rmax, std::fabs(acc[size_t(c)]));
            cut = truncation * rmax;
            double kept = 0.0;
            for (int c : cols)
                if (acc[size_t(c)] != 0.0 && std::fabs(acc[size_t(c)]) >= cut)
                    kept += acc[size_t(c)];
            scale = kept != 0.0 ? full / kept : 1.0;
        }
        const size_t before = out.ci.size();
        for (int c : cols) {
            used[size_t(c)] = 0;
            const double v = acc[size_t(c)];
            if (v != 0.0 && (truncation <= 0.0 || std::fabs(v) >= cut)) {
                out.ci.push_back(c);
                out.v.push_back(truncation > 0.0 ? v * scale : v);
            }
        }
        out.len.push_back(int(out.ci.size() - before));
    }
    });
    join_parts(parts, n, P.rp, P.ci, &P.v);
}

void csr_transpose(const HCsr& A, HCsr& T)
{
    T.nr = A.nc;
    T.nc = A.nr;
    T.rp.assign(size_t(A.nc) + 1, 0);
    for (int c : A.ci)
        ++T.rp[size_t(c) + 1];
    for (int i = 0; i < A.nc; ++i)
        T.rp[size_t(i) + 1] += T.rp[size_t(i)];
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    std::vector<int> pos(T.rp.begin(), T.rp.end() - 1);
    for (int i = 0; i < A.nr; ++i)
        for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
            const int c = A.ci[size_t(k)];
            const int w = pos[size_t(c)]++;
            T.ci[size_t(w)] = i;
            T.v[size_t(w)] = A.v[size_t(k)];
        }
}

void csr_multiply(const HCsr& A, const HCsr& B, HCsr& C, int threads)
{
    // Row-wise Gustavson product with a sparse accumulator per thread; rows
    // keep their columns sorted.
    threads = thread_count(threads);
    const int n = A.nr;
    struct Part
    {
        std::vector<int> rp, ci;
        std::vector<double> v;
    };
    const int parts = (threads > 1 && n >= 4096) ? threads : 1;
    std::vector<Part> part(static_cast<size_t>(parts));
    const int chunk = (n + parts - 1) / parts;
    auto work = [&](int b, int e, int t) {
        Part& P = part[size_t(t)];
        P.rp.assign(1, 0);
        std::vector<int> mark(size_t(B.nc), -1);
        std::vector<double> acc(size_t(B.nc), 0.0);
        std::vector<int> cols;
        for (int i = b; i < e; ++i) {
            cols.clear();
            for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                const int r = A.ci[size_t(k)];
                const double a = A.v[size_t(k)];
                for (int q = B.rp[size_t(r)]; q < B.rp[size_t(r) + 1]; ++q) {
                    const int c = B.ci[size_t(q)];
                    if (mark[size_t(c)] != i) {
                        mark[size_t(c)] = i;
                        acc[size_t(c)] = 0.0;
                        cols.push_back(c);
                    }
                    acc[size_t(c)] += a * B.v[size_t(q)];
                }
            }
            std::sort(cols.begin(), cols.end());
            for (int c : cols) {
                P.ci.push_back(c);
                P.v.push_back(acc[size_t(c)]);
            }
            P.rp.push_back(int(P.ci.size()));
        }
    };
    if (parts == 1) {
        work(0, n, 0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < parts; ++t) {
            const int b = t * chunk, e = std::min(n, b + chunk);
            pool.emplace_back([&, b, e, t] {
                if (b < e)
                    work(b, e, t);
                else
                    part[size_t(t)].rp.assign(1, 0);
            });
        }
        for (auto& th : pool)
            th.join();
    }
    C.nr = n;
    C.nc = B.nc;
    size_t total = 0;
    for (auto& p : part)
        total += p.ci.size();
    C.rp.assign(size_t(n) + 1, 0);
    C.ci.resize(total);
    C.v.resize(total);
    size_t off = 0;
    int row = 0;
    for (auto& p : part) {
        const int rows = int(p.rp.size()) - 1;
        for (int r = 0; r < rows; ++r)
            C.rp[size_t(row + r) + 1] = int(off) + p.rp[size_t(r) + 1];
        std::memcpy(C.ci.data() + off, p.ci.data(), p.ci.size() * sizeof(int));
        std::memcpy(C.v.data() + off, p.v.data(), p.v.size() * sizeof(double));
        off += p.ci.size();
        row += rows;
    }
}

int dense_inverse(int n, std::vector<double>& a, std::vector<double>& inv, int threads)
{
    // Gauss-Jordan elimination with partial pivoting on [A | I].  A pivot that
    // vanishes (a singular coarse operator, e.g. an unpinned Neumann problem)
    // is replaced by a tiny multiple of the largest diagonal so the coarse
    // correction stays finite; the Krylov outer iteration absorbs it.
    // One team of threads splits the row updates of every pivot step; thread
    // 0 picks, swaps and scales the pivot row between two barriers.
    inv.assign(size_t(n) * size_t(n), 0.0);
    const size_t N = size_t(n);
    for (size_t i = 0; i < N; ++i)
        inv[i * N + i] = 1.0;
    double amax = 0.0;
    for (size_t i = 0; i < N; ++i)
        amax = std::max(amax, std::abs(a[i * N + i]));
    int regularised = 0;
    const int team = (n >= 256) ? std::max(1, std::min(threads, 16)) : 1;
    std::barrier sync(team);
    auto pivot_step = [&](size_t c) {
        size_t piv = c;
        for (size_t r = c + 1; r < N; ++r)
            if (std::abs(a[r * N + c]) > std::abs(a[piv * N + c]))
                piv = r;
        if (piv != c) {
            for (size_t k = 0; k < N; ++k) {
                std::swap(a[c * N + k], a[piv * N + k]);
                std::swap(inv[c * N + k], inv[piv * N + k]);
            }
        }
        double p = a[c * N + c];
        if (std::abs(p) <= 1e-13 * amax) {
            p = (p >= 0 ? 1.0 : -1.0) * 1e-10 * (amax > 0 ? amax : 1.0);
            a[c * N + c] = p;
            ++regularised;
        }
        const double ip = 1.0 / p;
        for (size_t k = 0; k < N; ++k) {
            a[c * N + k] *= ip;
            inv[c * N + k] *= ip;
        }
    };
    auto worker = [&](int t) {
        const size_t chunk = (N + size_t(team) - 1) / size_t(team);
        const size_t r0 = size_t(t) * chunk, r1 = std::min(N, r0 + chunk);
        for (size_t c = 0; c < N; ++c) {
            if (t == 0)
                pivot_step(c);
            sync.arrive_and_wait();
            const double* ac = &a[c * N];
            const double* ic = &inv[c * N];
            for (size_t r = r0; r < r1; ++r) {
                if (r == c)
                    continue;
                double* ar = &a[r * N];
                const double f = ar[c];
                if (f == 0.0)
                    continue;
                // columns < c of the pivot row are already zero
                for (size_t k = c; k < N; ++k)
                    ar[k] -= f * ac[k];
                double* ir = &inv[r * N];
                for (size_t k = 0; k < N; ++k)
                    ir[k] -= f * ic[k];
            }
            sync.arrive_and_wait();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < team; ++t)
        pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool)
        th.join();
    return regularised;
}

// Renumbers the coarse unknowns of a new level with reverse Cuthill-McKee:
// A_c <- Q A_c Q^T and P <- P Q^T (R is rebuilt as P^T).  An equivalent
// hierarchy whose coarse operators have a narrow band, so the device gathers
// of x on those levels touch few cache lines.
void amg_renumber_coarse(HCsr& Ac, HCsr& P, std::vector<int>& agg)
{
    const int nc = Ac.nr;
    if (nc < 2)
        return;
    std::vector<int> perm, brp, bci, vmap;
    rcm_order(nc, Ac.rp.data(), Ac.ci.data(), perm);
    permute_pattern(nc, Ac.rp.data(), Ac.ci.data(), perm, brp, bci, vmap);
    std::vector<double> bv(vmap.size());
    parallel_for(int(vmap.size()), thread_count(0), [&](int b, int e, int) {
        for (int k = b; k < e; ++k)
            bv[size_t(k)] = Ac.v[size_t(vmap[size_t(k)])];
    });
    Ac.rp = std::move(brp);
    Ac.ci = std::move(bci);
    Ac.v = std::move(bv);
    std::vector<int> inv(static_cast<size_t>(nc));
    for (int i = 0; i < nc; ++i)
        inv[size_t(perm[size_t(i)])] = i;
    for (int& a : agg)
        if (a >= 0)
            a = inv[size_t(a)];
    parallel_for(P.nr, thread_count(0), [&](int b, int e, int) {
        std::vector<std::pair<int, double>> row;
        for (int i = b; i < e; ++i) {
            row.clear();
            for (int k = P.rp[size_t(i)]; k < P.rp[size_t(i) + 1]; ++k)
                row.emplace_back(inv[size_t(P.ci[size_t(k)])], P.v[size_t(k)]);
            std::sort(row.begin(), row.end(),
                      [](const std::pair<int, double>& x, const std::pair<int, double>& y) {
                          return x.first < y.first;
                      });
            int k = P.rp[size_t(i)];
            for (const auto& en : row) {
                P.ci[size_t(k)] = en.first;
                P.v[size_t(k)] = en.second;
                ++k;
            }
        }
    });
}

int amg_build(const HCsr& A0, const AmgParams& p, AmgHierarchyHost& out)
{
    out.levels.clear();
    out.levels.emplace_back();
    out.levels.back().A = A0;
    const int threads = thread_count(p.threads);
    int64_t nnz_total = A0.nnz(), rows_total = A0.nr;
    for (int lev = 0; lev + 1 < p.max_levels; ++lev) {
        AmgLevelHost& L = out.levels.back();
        const HCsr& A = L.A;
        if (A.nr <= std::min(p.coarse_size, kAmgMaxDense))
            break;
        std::vector<int> srp, sci, agg;
        amg_strength(A, p.theta, srp, sci);
        int nagg = amg_aggregate(A.nr, srp, sci, A, agg);
        if (nagg >= A.nr * 9 / 10 && p.theta > 0.0) {
            // too few strong couplings to coarsen: treat every coupling as strong
            amg_strength(A, 0.0, srp, sci);
            nagg = amg_aggregate(A.nr, srp, sci, A, agg);
        }
        // the strong entries of A that the filtered smoothing keeps (every
        // off-diagonal entry when unfiltered), for numeric re-setups
        L.strong.assign(size_t(A.nnz()), 0);
        parallel_for(A.nr, thread_count(0), [&](int b, int e, int) {
            std::vector<char> mark(size_t(A.nc), 0);
            for (int i = b; i < e; ++i) {
                for (int s = srp[size_t(i)]; s < srp[size_t(i) + 1]; ++s)
                    mark[size_t(sci[size_t(s)])] = 1;
                for (int k = A.rp[size_t(i)]; k < A.rp[size_t(i) + 1]; ++k) {
                    const int j = A.ci[size_t(k)];
                    L.strong[size_t(k)] = j != i && (!p.filter || mark[size_t(j)]);
                }
                for (int s = srp[size_t(i)]; s < srp[size_t(i) + 1]; ++s)
                    mark[size_t(sci[size_t(s)])] = 0;
            }
        });
        if (nagg <= 0 || nagg >= A.nr * 9 / 10)
            break; // no useful coarsening
        L.lmax = amg_lambda_max(A, 20) * 1.05;
        L.n_aggregates = nagg;
        const double omega = p.omega_scale / L.lmax;
        if (p.filter)
            amg_smoothed_prolongator(A, agg, nagg, omega, L.P, &srp, &sci, p.truncation);
        else
            amg_smoothed_prolongator(A, agg, nagg, omega, L.P, nullptr, nullptr, p.truncation);
        HCsr AP, Ac;
        {
            HCsr R0;
            csr_transpose(L.P, R0);
            csr_multiply(A, L.P, AP, threads);
            csr_multiply(R0, AP, Ac, threads);
        }
        if (p.renumber)
            amg_renumber_coarse(Ac, L.P, agg);
        L.agg = std::move(agg);
        csr_transpose(L.P, L.R);
        out.levels.emplace_back();
        out.levels.back().A = std::move(Ac);
        nnz_total += out.levels.back().A.nnz();
        rows_total += out.levels.back().A.nr;
    }
    AmgLevelHost& C = out.levels.back();
    C.lmax = amg_lambda_max(C.A, 20) * 1.05;
    const int nc = C.A.nr;
    if (nc > kAmgMaxDense)
        return 1; // coarsening stalled: the coarsest level is too large for a dense solve
    std::vector<double>& dense = out.coarse_dense;
    dense.assign(size_t(nc) * size_t(nc), 0.0);
    for (int i = 0; i < nc; ++i)
        for (int k = C.A.rp[size_t(i)]; k < C.A.rp[size_t(i) + 1]; ++k)
            dense[size_t(i) * size_t(nc) + size_t(C.A.ci[size_t(k)])] += C.A.v[size_t(k)];
    out.coarse_inverse.clear();
    if (p.host_coarse_inverse) {
        std::vector<double> work(dense);
        dense_inverse(nc, work, out.coarse_inverse, threads);
    }
    out.coarse_n = nc;
    out.operator_complexity = double(nnz_total) / double(std::max<int64_t>(A0.nnz(), 1));
    out.grid_complexity = double(rows_total) / double(std::max(A0.nr, 1));
    return 0;
}

} // namespace mi
