// Host-side data parallelism for the mesh builder and the upload: a loop over
// independent indices split into contiguous chunks on a few host threads.
// Each index is processed exactly once by the same code, so results do not
// depend on the thread count. Helper threads are counted process-wide and
// capped (at most 15 besides the callers), so loops started from several
// threads at once (blocks built concurrently) do not multiply the threads.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <thread>
#include <vector>

namespace e3d {

inline int host_threads() {
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, h));
}

inline std::atomic<int>& helper_threads_alive() {
    static std::atomic<int> n{0};
    return n;
}

// body(i0, i1) over [0, n) in contiguous chunks of at least `grain` indices;
// the first exception thrown by a chunk is rethrown after all chunks ended.
template <class Body>
void parallel_for(size_t n, Body body, size_t grain = (size_t)1 << 15) {
    std::atomic<int>& alive = helper_threads_alive();
    const int cap = host_threads() - 1;
    const int want = (int)std::min<size_t>((size_t)cap, n / std::max<size_t>(grain, 1));
    int got = 0;
    while (got < want) {
        int cur = alive.load();
        if (cur >= cap) break;
        if (alive.compare_exchange_weak(cur, cur + 1)) ++got;
    }
    if (got == 0) {
        body((size_t)0, n);
        return;
    }
    const int parts = got + 1;
    std::vector<std::exception_ptr> err(parts);
    auto run = [&](int p) {
        try {
            body(n * p / parts, n * (p + 1) / parts);
        } catch (...) {
            err[p] = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    int p = 1;
    try {
        for (; p < parts; ++p) pool.emplace_back(run, p);
    } catch (...) {  // no more threads: the caller runs the remaining chunks
    }
    for (int q = p; q < parts; ++q) run(q);
    run(0);
    for (auto& t : pool) t.join();
    alive -= got;
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

}  // namespace e3d
