// parallel.hpp -- minimal std::thread fork/join helpers for the host-side code
// (generators, CSR assembly, the pGOrder CPU phase).  No OpenMP runtime is
// linked into libmr, so it can share a process with any other OpenMP user.
#pragma once
#include <sched.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace mr {

// per-thread cap on the worker count of the helpers below (0 = no cap), so a
// background task (a reorder beside a serving loop) can leave cores free
inline int& host_thread_cap()
{
    static thread_local int cap = 0;
    return cap;
}

inline int host_threads()
{
    static const int n = [] {
        // the CPUs this process may run on (a container's share), not every
        // CPU of the machine
        int h = 0;
        cpu_set_t set;
        CPU_ZERO(&set);
        if (sched_getaffinity(0, sizeof(set), &set) == 0) h = CPU_COUNT(&set);
        if (h <= 0) h = (int)std::thread::hardware_concurrency();
        if (h <= 0) h = 4;
        if (const char* e = std::getenv("MR_HOST_THREADS")) {
            const int v = std::atoi(e);
            if (v > 0) h = v;
        }
        return std::min(h, 64);
    }();
    const int cap = host_thread_cap();
    return cap > 0 ? std::min(n, cap) : n;
}

// fn(t, begin, end) on contiguous chunks of [0, n), one chunk per thread
template <typename F> void parallel_chunks(int64_t n, F&& fn, int threads = 0)
{
    if (threads <= 0) threads = host_threads();
    if (n < 4096 || threads == 1) {
        fn(0, (int64_t)0, n);
        return;
    }
    threads = (int)std::min<int64_t>(threads, n);
    std::vector<std::thread> th;
    th.reserve((size_t)threads);
    for (int t = 0; t < threads; ++t) {
        const int64_t b = n * t / threads, e = n * (t + 1) / threads;
        th.emplace_back([&fn, t, b, e] { fn(t, b, e); });
    }
    for (auto& x : th) x.join();
}

inline uint64_t splitmix64_next(uint64_t& s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// independent stream for (seed, index): finalised index mixed into the seed
inline uint64_t stream_state(uint64_t seed, uint64_t index)
{
    uint64_t s = index * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull;
    uint64_t h = splitmix64_next(s);
    return seed ^ h;
}

inline double u01(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace mr
