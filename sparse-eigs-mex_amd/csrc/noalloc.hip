// noalloc.hip -- speigs_hip_op_apply_noalloc: speigs_hip_op_apply for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  No counterpart in the
// reference (its operator, src/speigs_core.c:57-116, is host code).
//
// speigs_hip_op_apply stages the caller's host vectors through device buffers
// of the handle that grow on demand (an allocation whenever nvec grows).  This
// variant stages them through `scratch`, a file-scope device array of fixed
// size that every device holds from the library's load until the process
// ends; the handle's own buffers (the factor, the solve workspace) were all
// allocated when the handle was created.  In host code `scratch` names a
// host-side stand-in, not device memory: its device address is taken once per
// device (hipGetSymbolAddress, with that device current, at the first call on
// it) and only that address is ever handed to kernels and copies.
//
// All calls on one device share its array: calls that overlap on one device
// (from several host threads) race on it.  Nothing in it has to start at zero:
// each call writes the input half before any kernel reads it.
#include <mutex>

#include "factor_impl.h"
#include "op_impl.h"

namespace sehip {
namespace {

constexpr int64_t kScratchDoubles = (int64_t)8 << 20;   // 64 MiB per device
constexpr int64_t kHalf = kScratchDoubles / 2;           // input half, output half
constexpr int kMaxDev = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_mu;
double *g_scratch[kMaxDev] = {};

// scratch's device address on device dev (taken once, with dev current)
int scratch_on(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDev) return SEHIP_FAIL(SPEIGS_ERR_ARG, "op_apply_noalloc: device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        int cur = 0;
        SEHIP_CHECK(hipGetDevice(&cur));
        if (cur != dev) SEHIP_CHECK(hipSetDevice(dev));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) SEHIP_CHECK(hipSetDevice(cur));
        if (e != hipSuccess) return hip_fail(e, __FILE__, __LINE__, "hipGetSymbolAddress(scratch)");
        if (!p) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "op_apply_noalloc: scratch has no device address");
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return 0;
}

struct DeviceGuard {
    int prev = -1;
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

}  // namespace
}  // namespace sehip

extern "C" int speigs_hip_op_apply_noalloc(speigs_hip_op *op, const double *x, double *y, speigs_int nvec)
{
    using namespace sehip;
    clear_error();
    if (!op || !x || !y || nvec <= 0) return SPEIGS_ERR_ARG;
    if (op->mode == SPEIGS_HIP_OP_SOLVE && (!op->F || !factor_apply_preallocated(op->F)))
        return SEHIP_FAIL(SPEIGS_ERR_ARG, "op_apply_noalloc: this operator's solve allocates at first use "
                                          "(precision SPEIGS_HIP_FP32_BLOCK); use speigs_hip_op_apply");
    if (op->n > kHalf / nvec) {
        char what[200];
        snprintf(what, sizeof what, "op_apply_noalloc: n * nvec = %lld * %lld exceeds the %lld values of scratch",
                 (long long)op->n, (long long)nvec, (long long)kHalf);
        return SEHIP_FAIL(SPEIGS_ERR_ARG, what);
    }
    double *base = nullptr;
    int rc = scratch_on(op->dev, &base);
    if (rc) return rc;
    DeviceGuard g;
    int cur = 0;
    SEHIP_CHECK(hipGetDevice(&cur));
    if (cur != op->dev) {
        SEHIP_CHECK(hipSetDevice(op->dev));
        g.prev = cur;
    }
    const int64_t n = op->n;
    double *xs = base, *ys = base + kHalf;
    const size_t bytes = (size_t)n * (size_t)nvec * sizeof(double);
    SEHIP_CHECK(hipMemcpyAsync(xs, x, bytes, hipMemcpyHostToDevice, op->stream));
    if (op->mode == SPEIGS_HIP_OP_MULT) {
        spmv_multi(op->A, xs, n, ys, n, (int)nvec, op->stream);
    } else {
        for (int64_t v = 0; v < nvec; v++)
            if ((rc = factor_apply(op->F, xs + v * n, ys + v * n))) return rc;
    }
    SEHIP_CHECK(hipMemcpyAsync(y, ys, bytes, hipMemcpyDeviceToHost, op->stream));
    SEHIP_CHECK(hipStreamSynchronize(op->stream));
    return op->mode == SPEIGS_HIP_OP_MULT ? 0 : factor_check(op->F);
}
