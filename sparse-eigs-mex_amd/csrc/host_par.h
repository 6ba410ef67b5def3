// host_par.h -- host-side thread helpers for the ordering and symbolic analysis.
#pragma once
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <mutex>
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

// Worker threads kept for the process's lifetime, so that a parallel loop
// costs a wake-up instead of thread creations (the analysis runs a few dozen
// loops per call).  One loop at a time: run() returns false when the pool is
// busy (a loop started from inside a loop, or from two host threads at once),
// and the caller then starts threads of its own.
class HostPool {
public:
    static HostPool &get()
    {
        static HostPool *p = new HostPool();   // never destroyed: its threads stay blocked until exit
        return *p;
    }
    // work(tid) for tid in [0, nt): tid 0 on the caller, the others on workers
    bool run(int nt, const std::function<void(int)> &work)
    {
        std::unique_lock<std::mutex> busy(busy_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        {
            std::lock_guard<std::mutex> lk(m_);
            while ((int)workers_.size() < nt - 1) {
                const int id = (int)workers_.size() + 1;
                workers_.emplace_back([this, id] { loop(id); });
            }
            job_ = &work;
            want_ = nt;
            remaining_ = nt - 1;
            gen_++;
        }
        cv_.notify_all();
        work(0);
        std::unique_lock<std::mutex> lk(m_);
        done_.wait(lk, [&] { return remaining_ == 0; });
        job_ = nullptr;
        return true;
    }

private:
    void loop(int id)
    {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            if (id >= want_) continue;
            const std::function<void(int)> *job = job_;
            lk.unlock();
            (*job)(id);
            lk.lock();
            if (--remaining_ == 0) done_.notify_one();
        }
    }
    std::mutex busy_, m_;
    std::condition_variable cv_, done_;
    std::vector<std::thread> workers_;
    const std::function<void(int)> *job_ = nullptr;
    uint64_t gen_ = 0;
    int want_ = 0, remaining_ = 0;
};

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
    std::function<void(int)> work = [&](int tid) {
        for (;;) {
            const int64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const int64_t e = std::min(n, b + chunk);
            for (int64_t i = b; i < e; i++) fn(i, tid);
        }
    };
    if (HostPool::get().run(nt, work)) return;
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &t : th) t.join();
}

}  // namespace sehip
