// devcache.cpp -- device memory for DevMem, with freed blocks kept for reuse.
//
// A call at n = 10^6 allocates ~20 device buffers (the 13 GB front pool, the
// solve state, the eigensolver's blocks) and returning them to the runtime
// took ~9 ms of a 0.37 s call.  Freed blocks are instead kept per device and
// handed to a later request of the same size class (best fit, at most 1/8
// larger), so repeated calls on one device allocate and free nothing.  The
// blocks kept on a device are capped (SPEIGS_HIP_CACHE_MB, default a quarter
// of the device's memory; 0 turns the cache off); a failed allocation first
// returns every kept block of its device to the runtime and tries again, and
// speigs_hip_release_cache() returns all of them.  A block is handed back only
// after its device has finished all queued work (hipDeviceSynchronize: the
// same wait hipFree makes), so no kernel of the releasing call still uses it.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "dev.h"

namespace sehip {

namespace {

constexpr int kMaxDev = 64;

struct Cache {
    std::mutex mu;
    std::multimap<size_t, void *> free_blocks[kMaxDev];   // bytes -> block
    size_t held[kMaxDev] = {};
    size_t cap[kMaxDev] = {};
    bool cap_known[kMaxDev] = {};
};

Cache &cache()
{
    static Cache *c = new Cache();   // never destroyed: the runtime may be gone at exit
    return *c;
}

size_t round_size(size_t bytes)
{
    const size_t g = bytes >= ((size_t)64 << 20) ? ((size_t)2 << 20) : ((size_t)4 << 10);
    return (bytes + g - 1) / g * g;
}

// the cap of device dev (the caller holds the lock and has dev current)
size_t cap_of(Cache &c, int dev)
{
    const char *e = std::getenv("SPEIGS_HIP_CACHE_MB");
    if (e && *e) return (size_t)std::max(0LL, std::atoll(e)) << 20;
    if (!c.cap_known[dev]) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            tot = 0;
        }
        c.cap[dev] = tot / 4;
        c.cap_known[dev] = true;
    }
    return c.cap[dev];
}

// returns every kept block of device dev to the runtime (lock held, dev current)
void drain(Cache &c, int dev)
{
    for (auto &kv : c.free_blocks[dev]) (void)hipFree(kv.second);
    c.free_blocks[dev].clear();
    c.held[dev] = 0;
}

}  // namespace

int dev_alloc(void **p, size_t bytes, int *dev_out, size_t *bytes_out)
{
    *p = nullptr;
    int dev = 0;
    SEHIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return SPEIGS_ERR_ARG;
    const size_t want = round_size(bytes ? bytes : 1);
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    auto &fl = c.free_blocks[dev];
    auto it = fl.lower_bound(want);
    if (it != fl.end() && it->first <= want + want / 8) {
        *p = it->second;
        *bytes_out = it->first;
        c.held[dev] -= it->first;
        fl.erase(it);
        *dev_out = dev;
        return 0;
    }
    hipError_t e = hipMalloc(p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        drain(c, dev);
        e = hipMalloc(p, want);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return SPEIGS_ERR_MEM;
    }
    *bytes_out = want;
    *dev_out = dev;
    return 0;
}

void dev_release(std::vector<DevBlock> &blocks)
{
    if (blocks.empty()) return;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) (void)hipGetLastError();
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    std::vector<int> devs;
    for (const DevBlock &b : blocks)
        if (std::find(devs.begin(), devs.end(), b.dev) == devs.end()) devs.push_back(b.dev);
    for (int dev : devs) {
        if (hipSetDevice(dev) != hipSuccess) (void)hipGetLastError();
        const size_t cap = cap_of(c, dev);
        // the releasing call's kernels may still be queued (error returns)
        const bool idle = hipDeviceSynchronize() == hipSuccess;
        if (!idle) (void)hipGetLastError();
        for (const DevBlock &b : blocks) {
            if (b.dev != dev) continue;
            if (idle && cap > 0 && b.bytes <= cap) {
                c.free_blocks[dev].emplace(b.bytes, b.p);
                c.held[dev] += b.bytes;
            } else {
                (void)hipFree(b.p);
            }
        }
        // over the cap: the largest kept blocks go back first
        while (c.held[dev] > cap && !c.free_blocks[dev].empty()) {
            auto last = std::prev(c.free_blocks[dev].end());
            (void)hipFree(last->second);
            c.held[dev] -= last->first;
            c.free_blocks[dev].erase(last);
        }
    }
    blocks.clear();
    if (hipSetDevice(cur) != hipSuccess) (void)hipGetLastError();
}

}  // namespace sehip

void speigs_hip_release_cache(void)
{
    using namespace sehip;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) (void)hipGetLastError();
    Cache &c = cache();
    std::lock_guard<std::mutex> lk(c.mu);
    for (int dev = 0; dev < kMaxDev; dev++) {
        if (c.free_blocks[dev].empty()) continue;
        if (hipSetDevice(dev) != hipSuccess) {
            (void)hipGetLastError();
            continue;
        }
        drain(c, dev);
    }
    if (hipSetDevice(cur) != hipSuccess) (void)hipGetLastError();
}
