// host_threads.hpp — row-range parallelism for the host setup code (AMG
// construction, reorderings).  Work is split into contiguous row ranges whose
// results do not depend on the number of threads.
#pragma once

#include <algorithm>
#include <thread>
#include <vector>

namespace mi
{

// Threads for host setup: `requested`, or the hardware concurrency; at most 16.
inline int host_threads(int requested = 0)
{
    int t = requested > 0 ? requested : int(std::thread::hardware_concurrency());
    return std::clamp(t, 1, 16);
}

// f(begin, end, part) over `parts` contiguous ranges of [0, n), run on
// `parts` threads (inline when parts == 1 or n < min_n).
template <class F> void host_parallel_for(int n, int parts, F&& f, int min_n = 4096)
{
    if (parts <= 1 || n < min_n) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (n + parts - 1) / parts;
    for (int t = 0; t < parts; ++t) {
        const int b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back([&f, b, e, t] { f(b, e, t); });
    }
    for (auto& th : pool)
        th.join();
}

} // namespace mi
