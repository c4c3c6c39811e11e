// host_parallel.h -- a parallel loop over host threads for the upload's
// per-triangle passes (memory-bound reads of the caller's entities).
#pragma once

#include <algorithm>
#include <cstddef>
#include <thread>
#include <vector>

namespace az {

// Runs f(i) for i in [0, n), split into contiguous ranges over up to 16
// host threads (serially below minParallel items).  f must only write state
// of its own index.
template <class F>
void parallel_for(size_t n, F f, size_t minParallel = 65536) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t threads = std::min<size_t>(std::min<unsigned>(hw, 16u), n / std::max<size_t>(1, minParallel / 4));
    if (n < minParallel || threads < 2) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve(threads);
    const size_t per = (n + threads - 1) / threads;
    for (size_t t = 0; t < threads; ++t) {
        const size_t b = t * per, e = std::min(n, b + per);
        if (b >= e) break;
        pool.emplace_back([b, e, &f]() {
            for (size_t i = b; i < e; ++i) f(i);
        });
    }
    for (std::thread &th : pool) th.join();
}

}  // namespace az
