// host_par.h -- host-side thread helpers for the ordering and symbolic analysis.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

namespace sehip {

// worker threads for host analysis: SPEIGS_HIP_THREADS if set, else the
// hardware concurrency, at most 16
inline int host_threads()
{
    if (const char *e = getenv("SPEIGS_HIP_THREADS")) {
        const int t = atoi(e);
        if (t >= 1) return std::min(t, 64);
    }
    const int hw = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(hw > 0 ? hw : 1, 16));
}

// fn(i, tid) for i in [0, n), dynamically scheduled in chunks over nt threads
template <class Fn>
void parallel_for(int64_t n, int nt, int64_t chunk, Fn fn)
{
    if (n <= 0) return;
    nt = (int)std::max<int64_t>(1, std::min<int64_t>(nt, (n + chunk - 1) / chunk));
    if (nt == 1) {
        for (int64_t i = 0; i < n; i++) fn(i, 0);
        return;
    }
    std::atomic<int64_t> next{0};
    auto work = [&](int tid) {
        for (;;) {
            const int64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const int64_t e = std::min(n, b + chunk);
            for (int64_t i = b; i < e; i++) fn(i, tid);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &t : th) t.join();
}

}  // namespace sehip
