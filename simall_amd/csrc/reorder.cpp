// reorder.cpp — reverse Cuthill-McKee (host, O(nnz log d)).
//
// The randomly numbered meshes of the pressure workloads scatter each row's
// x-gather over the whole vector; after RCM a row touches a window of about
// one level set (N^2 cells for an N^3 grid), which stays in each XCD's L2.
#include "reorder.hpp"

#include "host_threads.hpp"

#include <algorithm>
#include <numeric>

namespace mi
{

namespace
{

// Symmetrised adjacency without self loops.
void symmetrise(int n, const int* rp, const int* ci, std::vector<int>& arp, std::vector<int>& aci)
{
    std::vector<int> deg(size_t(n) + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (j != i) {
                ++deg[size_t(i)];
                ++deg[size_t(j)];
            }
        }
    arp.assign(size_t(n) + 1, 0);
    for (int i = 0; i < n; ++i)
        arp[size_t(i) + 1] = arp[size_t(i)] + deg[size_t(i)];
    aci.assign(size_t(arp[size_t(n)]), 0);
    std::vector<int> pos(arp.begin(), arp.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k) {
            const int j = ci[k];
            if (j != i) {
                aci[size_t(pos[size_t(i)]++)] = j;
                aci[size_t(pos[size_t(j)]++)] = i;
            }
        }
    // sort + unique each row (structurally symmetric inputs give duplicates),
    // rows in parallel, then compaction
    std::vector<int> ulen(static_cast<size_t>(n), 0);
    host_parallel_for(n, host_threads(), [&](int b0, int e0, int) {
        for (int i = b0; i < e0; ++i) {
            auto b = aci.begin() + arp[size_t(i)], e = aci.begin() + arp[size_t(i) + 1];
            std::sort(b, e);
            ulen[size_t(i)] = int(std::unique(b, e) - b);
        }
    });
    std::vector<int> nrp(size_t(n) + 1, 0);
    int w = 0;
    for (int i = 0; i < n; ++i) {
        const int b = arp[size_t(i)];
        for (int q = 0; q < ulen[size_t(i)]; ++q)
            aci[size_t(w++)] = aci[size_t(b + q)];
        nrp[size_t(i) + 1] = w;
    }
    aci.resize(size_t(w));
    arp.swap(nrp);
}

// BFS levels from `root` within unvisited nodes; returns the last level's
// nodes in `last` and the number of levels.
int bfs_levels(int root, const std::vector<int>& arp, const std::vector<int>& aci,
               const std::vector<char>& done, std::vector<int>& mark, int stamp,
               std::vector<int>& queue, std::vector<int>& last)
{
    queue.clear();
    queue.push_back(root);
    mark[size_t(root)] = stamp;
    size_t head = 0;
    int levels = 0;
    while (head < queue.size()) {
        const size_t end = queue.size();
        last.assign(queue.begin() + long(head), queue.begin() + long(end));
        for (; head < end; ++head) {
            const int u = queue[head];
            for (int k = arp[size_t(u)]; k < arp[size_t(u) + 1]; ++k) {
                const int v = aci[size_t(k)];
                if (!done[size_t(v)] && mark[size_t(v)] != stamp) {
                    mark[size_t(v)] = stamp;
                    queue.push_back(v);
                }
            }
        }
        ++levels;
    }
    return levels;
}

} // namespace

void rcm_order(int n, const int* rp, const int* ci, std::vector<int>& perm)
{
    std::vector<int> arp, aci;
    symmetrise(n, rp, ci, arp, aci);
    auto degree = [&](int u) { return arp[size_t(u) + 1] - arp[size_t(u)]; };
    std::vector<char> done(size_t(n), 0);
    std::vector<int> mark(size_t(n), -1), queue, last, order;
    order.reserve(size_t(n));
    int stamp = 0;
    std::vector<int> nbrs;
    for (int seed = 0; seed < n; ++seed) {
        if (done[size_t(seed)])
            continue;
        // pseudo-peripheral root: repeat BFS from a minimum-degree node of the
        // deepest level while the depth grows
        int root = seed;
        int depth = bfs_levels(root, arp, aci, done, mark, stamp++, queue, last);
        for (int rep = 0; rep < 8; ++rep) {
            int cand = last[0];
            for (int u : last)
                if (degree(u) < degree(cand))
                    cand = u;
            const int d = bfs_levels(cand, arp, aci, done, mark, stamp++, queue, last);
            if (d <= depth)
                break;
            depth = d;
            root = cand;
        }
        // Cuthill-McKee from root
        const size_t start = order.size();
        order.push_back(root);
        done[size_t(root)] = 1;
        for (size_t h = start; h < order.size(); ++h) {
            const int u = order[h];
            nbrs.clear();
            for (int k = arp[size_t(u)]; k < arp[size_t(u) + 1]; ++k) {
                const int v = aci[size_t(k)];
                if (!done[size_t(v)]) {
                    done[size_t(v)] = 1;
                    nbrs.push_back(v);
                }
            }
            std::stable_sort(nbrs.begin(), nbrs.end(),
                             [&](int a, int b) { return degree(a) < degree(b); });
            order.insert(order.end(), nbrs.begin(), nbrs.end());
        }
    }
    perm.assign(order.rbegin(), order.rend());
}

void permute_pattern(int n, const int* rp, const int* ci, const std::vector<int>& perm,
                     std::vector<int>& brp, std::vector<int>& bci, std::vector<int>& vmap)
{
    std::vector<int> inv(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        inv[size_t(perm[size_t(i)])] = i;
    brp.assign(size_t(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        const int o = perm[size_t(i)];
        brp[size_t(i) + 1] = brp[size_t(i)] + (rp[o + 1] - rp[o]);
    }
    bci.resize(size_t(brp[size_t(n)]));
    vmap.resize(size_t(brp[size_t(n)]));
    host_parallel_for(n, host_threads(), [&](int b0, int e0, int) {
        std::vector<std::pair<int, int>> row;
        for (int i = b0; i < e0; ++i) {
            const int o = perm[size_t(i)];
            row.clear();
            for (int k = rp[o]; k < rp[o + 1]; ++k)
                row.emplace_back(inv[size_t(ci[k])], k);
            std::sort(row.begin(), row.end());
            int w = brp[size_t(i)];
            for (auto& [c, k] : row) {
                bci[size_t(w)] = c;
                vmap[size_t(w)] = k;
                ++w;
            }
        }
    });
}

int bandwidth(int n, const int* rp, const int* ci)
{
    int bw = 0;
    for (int i = 0; i < n; ++i)
        for (int k = rp[i]; k < rp[i + 1]; ++k)
            bw = std::max(bw, std::abs(ci[k] - i));
    return bw;
}

void color_order(int n, const int* rp, const int* ci, std::vector<int>& perm, int* colors)
{
    std::vector<int> rcm;
    rcm_order(n, rp, ci, rcm);
    std::vector<int> arp, aci;
    symmetrise(n, rp, ci, arp, aci);
    // greedy colouring in RCM order: the smallest colour unused by coloured neighbours
    std::vector<int> col(static_cast<size_t>(n), -1), mark;
    int nc = 0;
    for (int t = 0; t < n; ++t) {
        const int u = rcm[size_t(t)];
        for (int k = arp[size_t(u)]; k < arp[size_t(u) + 1]; ++k) {
            const int c = col[size_t(aci[size_t(k)])];
            if (c >= 0) {
                if (c >= int(mark.size()))
                    mark.resize(size_t(c) + 1, -1);
                mark[size_t(c)] = u;
            }
        }
        int c = 0;
        while (c < int(mark.size()) && mark[size_t(c)] == u)
            ++c;
        col[size_t(u)] = c;
        nc = std::max(nc, c + 1);
    }
    // rows grouped by colour, RCM order within each colour
    std::vector<int> start(static_cast<size_t>(nc) + 1, 0);
    for (int u = 0; u < n; ++u)
        ++start[size_t(col[size_t(u)]) + 1];
    for (int c = 0; c < nc; ++c)
        start[size_t(c) + 1] += start[size_t(c)];
    perm.assign(static_cast<size_t>(n), 0);
    for (int t = 0; t < n; ++t) {
        const int u = rcm[size_t(t)];
        perm[size_t(start[size_t(col[size_t(u)])]++)] = u;
    }
    if (colors)
        *colors = nc;
}

} // namespace mi
