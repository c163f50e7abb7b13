// Host-side parallel loops of the class templates: contiguous index ranges, one thread each, at most 16 threads
// (a GPU node's CPU share per device), so that host phases that are per key (string packing, string forms of
// bumped keys, their placement) do not run on one core while the device waits.
#pragma once
#include <algorithm>
#include <cstddef>
#include <exception>
#include <thread>
#include <vector>

namespace shockhash {

// Threads for n items of work, at least `grain` items each.
inline size_t hostThreads(size_t n, size_t grain = 4096) {
    return std::max<size_t>(1, std::min<size_t>({16, (size_t)std::thread::hardware_concurrency(), n / grain}));
}

// body(t, lo, hi) for the nt ranges [n * t / nt, n * (t + 1) / nt), t < nt, each on its own thread (the calling
// thread when nt is 1); rethrows the first exception a range threw, after every thread has ended.
template <class F>
void parallelRanges(size_t n, size_t nt, F body) {
    if (nt <= 1) {
        body(size_t(0), size_t(0), n);
        return;
    }
    std::vector<std::thread> threads;
    std::vector<std::exception_ptr> errors(nt);
    threads.reserve(nt);
    for (size_t t = 0; t < nt; t++)
        threads.emplace_back([&, t] {
            try {
                body(t, n * t / nt, n * (t + 1) / nt);
            } catch (...) {
                errors[t] = std::current_exception();
            }
        });
    for (auto &th : threads) th.join();
    for (auto &e : errors)
        if (e) std::rethrow_exception(e);
}

// body(i) for every i in [0, n).
template <class F>
void parallelFor(size_t n, F body) {
    parallelRanges(n, hostThreads(n), [&](size_t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) body(i);
    });
}

}  // namespace shockhash
