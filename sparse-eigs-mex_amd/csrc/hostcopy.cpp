// hostcopy.cpp -- copies between pageable host memory and the device through
// pinned staging buffers, and pre-faulting of host output buffers.
//
// A hipMemcpy to or from pageable memory is staged by the runtime one small
// piece at a time, with the host-side copy of each piece on the calling
// thread, and a result buffer that the caller has just allocated also takes a
// page fault per page on its first write: ~4 GB/s device-to-host into a fresh
// n x k eigenvector array.  Here the copy is cut into chunks of kChunk bytes
// dealt round-robin to up to kThreads host threads; each thread moves its
// chunks through two pinned buffers of its own (the DMA into one overlapping
// the host copy out of the other), so the host copies run in parallel while
// the DMA engine stays busy.  The pinned buffers are allocated on first use
// and kept, per device, until the process ends (2 kThreads kChunk bytes per
// concurrent copy).  Within a copy a buffer whose transfer may still be in
// flight carries the event that ends it; every transfer has ended before the
// copy returns its buffers to the pool.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dev.h"

namespace sehip {

namespace {

constexpr size_t kChunk = (size_t)4 << 20;
constexpr int kThreads = 4;
constexpr size_t kDirect = (size_t)1 << 20;   // smaller copies: one hipMemcpyAsync
constexpr int kMaxDev = 64;

struct Slot {
    void *buf = nullptr;
    hipEvent_t ev = nullptr;
    bool pending = false;   // ev marks a transfer that may still use buf
};

struct Pool {
    std::mutex mu;
    std::vector<Slot *> free_slots[kMaxDev];
};

Pool &pool()
{
    static Pool *p = new Pool();   // never destroyed: the runtime may be gone at exit
    return *p;
}

int acquire(int dev, int count, std::vector<Slot *> &out)
{
    Pool &p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    std::vector<Slot *> &fl = p.free_slots[dev];
    while ((int)out.size() < count) {
        if (!fl.empty()) {
            out.push_back(fl.back());
            fl.pop_back();
            continue;
        }
        Slot *s = new Slot();
        if (hipHostMalloc(&s->buf, kChunk, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            delete s;
            return SPEIGS_ERR_MEM;
        }
        const hipError_t e = hipEventCreateWithFlags(&s->ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipHostFree(s->buf);
            delete s;
            return hip_fail(e, __FILE__, __LINE__, "hipEventCreateWithFlags(staging slot)");
        }
        out.push_back(s);
    }
    return 0;
}

void release(int dev, std::vector<Slot *> &slots)
{
    Pool &p = pool();
    std::lock_guard<std::mutex> lk(p.mu);
    for (Slot *s : slots) p.free_slots[dev].push_back(s);
    slots.clear();
}

int wait_slot(Slot *s)
{
    if (!s->pending) return 0;
    SEHIP_CHECK(hipEventSynchronize(s->ev));
    s->pending = false;
    return 0;
}

// rows x width bytes; row r at dst + r dpitch and src + r spitch
struct Layout {
    char *dst;
    const char *src;
    size_t dpitch, spitch, width, rows, per_row, nchunks;
    size_t len(size_t c) const
    {
        const size_t part = c % per_row;
        return std::min(kChunk, width - part * kChunk);
    }
    size_t doff(size_t c) const { return (c / per_row) * dpitch + (c % per_row) * kChunk; }
    size_t soff(size_t c) const { return (c / per_row) * spitch + (c % per_row) * kChunk; }
};

// thread t's chunks t, t + T, ...: device -> slot -> host, two slots in flight
int d2h_worker(const Layout &L, int t, int T, Slot *s0, Slot *s1, hipStream_t st, int dev)
{
    SEHIP_CHECK(hipSetDevice(dev));
    Slot *sl[2] = {s0, s1};
    int rc = 0;
    for (Slot *s : sl)
        if ((rc = wait_slot(s))) return rc;
    auto issue = [&](size_t c, Slot *s) -> int {
        SEHIP_CHECK(hipMemcpyAsync(s->buf, L.src + L.soff(c), L.len(c), hipMemcpyDeviceToHost, st));
        SEHIP_CHECK(hipEventRecord(s->ev, st));
        s->pending = true;
        return 0;
    };
    size_t j = 0;
    for (size_t c = t; c < L.nchunks && j < 2; c += T, j++)
        if ((rc = issue(c, sl[j]))) return rc;
    j = 0;
    for (size_t c = t; c < L.nchunks; c += T, j++) {
        Slot *s = sl[j & 1];
        if ((rc = wait_slot(s))) return rc;
        std::memcpy(L.dst + L.doff(c), s->buf, L.len(c));
        const size_t c2 = c + 2 * (size_t)T;
        if (c2 < L.nchunks && (rc = issue(c2, s))) return rc;
    }
    return 0;
}

// thread t's chunks: host -> slot -> device; a slot is refilled once its
// previous DMA has ended, the last two stay pending
int h2d_worker(const Layout &L, int t, int T, Slot *s0, Slot *s1, hipStream_t st, int dev)
{
    SEHIP_CHECK(hipSetDevice(dev));
    Slot *sl[2] = {s0, s1};
    int rc = 0;
    size_t j = 0;
    for (size_t c = t; c < L.nchunks; c += T, j++) {
        Slot *s = sl[j & 1];
        if ((rc = wait_slot(s))) return rc;
        std::memcpy(s->buf, L.src + L.soff(c), L.len(c));
        SEHIP_CHECK(hipMemcpyAsync(L.dst + L.doff(c), s->buf, L.len(c), hipMemcpyHostToDevice, st));
        SEHIP_CHECK(hipEventRecord(s->ev, st));
        s->pending = true;
    }
    return 0;
}

int staged(bool to_host, void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
           hipStream_t st)
{
    if (width == 0 || rows == 0) return 0;
    const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
    if (width * rows < kDirect) {
        if (rows == 1)
            SEHIP_CHECK(hipMemcpyAsync(dst, src, width, kind, st));
        else
            SEHIP_CHECK(hipMemcpy2DAsync(dst, dpitch, src, spitch, width, rows, kind, st));
        if (to_host) SEHIP_CHECK(hipStreamSynchronize(st));
        return 0;
    }
    int dev = 0;
    SEHIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDev) return SPEIGS_ERR_ARG;
    Layout L{static_cast<char *>(dst), static_cast<const char *>(src), dpitch, spitch, width, rows,
             (width + kChunk - 1) / kChunk, 0};
    L.nchunks = L.per_row * rows;
    const int T = (int)std::min<size_t>(kThreads, L.nchunks);
    std::vector<Slot *> slots;
    int rc = acquire(dev, 2 * T, slots);
    if (rc) {
        release(dev, slots);
        return rc;
    }
    std::vector<int> rcs(T, 0);
    std::vector<std::string> why(T);
    auto run = [&](int t) {
        rcs[t] = to_host ? d2h_worker(L, t, T, slots[2 * t], slots[2 * t + 1], st, dev)
                         : h2d_worker(L, t, T, slots[2 * t], slots[2 * t + 1], st, dev);
        if (rcs[t]) why[t] = last_error_text();
    };
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
    // the last transfers of a host-to-device copy may still be in flight:
    // wait for them here, so that no slot goes back to the pool holding an
    // event of this call's stream (which the caller may destroy)
    for (Slot *s : slots) {
        const int r = wait_slot(s);
        if (r && !rc) {
            rc = r;
            why[0] = last_error_text();
            rcs[0] = r;
        }
    }
    release(dev, slots);
    for (int t = 0; t < T; t++)
        if (rcs[t]) {
            note_error(why[t].c_str());
            return rcs[t];
        }
    return 0;
}

}  // namespace

int copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    return staged(true, dst, bytes, src, bytes, bytes, 1, st);
}

int copy_to_host_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                    hipStream_t st)
{
    return staged(true, dst, dpitch, src, spitch, width, rows, st);
}

int copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t st)
{
    return staged(false, dst, bytes, src, bytes, bytes, 1, st);
}

void HostTouch::start(void *p, size_t bytes)
{
    join();
    if (!p || bytes < ((size_t)1 << 22)) return;
    const int T = 4;
    const size_t page = 4096, per = (bytes / T + page - 1) / page * page;
    for (int t = 0; t < T; t++) {
        char *b = static_cast<char *>(p) + std::min(bytes, (size_t)t * per);
        char *e = static_cast<char *>(p) + std::min(bytes, (size_t)(t + 1) * per);
        if (b >= e) break;
        th_.emplace_back([b, e, page] {
            // one write per page, the last byte too (b need not be page-aligned)
            for (volatile char *c = b; c < e; c += page) *c = *c;
            volatile char *last = e - 1;
            *last = *last;
        });
    }
}

void HostTouch::join()
{
    for (auto &t : th_) t.join();
    th_.clear();
}

}  // namespace sehip
