// Host worker threads for the compiler passes that work per instance or per tile.
// SYMLIB_THREADS caps the count (default: hardware threads, at most 16).
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

namespace Sym {

inline unsigned workerThreads() {
    unsigned n = std::max(1u, std::thread::hardware_concurrency());
    n = std::min(n, 16u);
    if (const char* env = std::getenv("SYMLIB_THREADS")) {
        const int v = std::atoi(env);
        if (v > 0) n = (unsigned)v;
    }
    return n;
}

// f(i) for i in [0, n), dynamically distributed; the first exception is rethrown.
// f must only touch data owned by index i (or read shared data).
template <class F>
void parallelFor(size_t n, F&& f) {
    const unsigned T = (unsigned)std::min<size_t>(workerThreads(), n);
    if (T <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::exception_ptr err;
    std::mutex m;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&] {
            try {
                for (size_t i; (i = next.fetch_add(1)) < n;) f(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!err) err = std::current_exception();
                next = n;
            }
        });
    for (auto& th : pool) th.join();
    if (err) std::rethrow_exception(err);
}

}  // namespace Sym
