// op_impl.h -- state behind the opaque speigs_hip_op handle (include/speigs_hip.h):
// the device SELL matrix, the device factor for OP_SOLVE, the handle's stream
// and the device staging of the host vectors that speigs_hip_op_apply moves.
#pragma once
#include "factor.h"

struct speigs_hip_op {
    int mode = 0;
    int64_t n = 0;
    int dev = 0;                     // the device the handle was created on
    hipStream_t stream = nullptr;
    sehip::DevMem mem;
    sehip::SellA A;
    sehip::Factor *F = nullptr;
    double *x = nullptr, *y = nullptr;   // staging of op_apply (grown on demand)
    int64_t cap = 0;
    int precision = 0;
    ~speigs_hip_op()
    {
        sehip::factor_destroy(F);
        mem.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};
