// mpnn.h -- internal interface of the learned start (mpnn.hip).
#pragma once
#include <cstdint>

#include "../../include/speigs_hip.h"

namespace sehip {

constexpr int MP_WIDTH = 32;   // channels of every layer and of the output block
constexpr int MP_NOISE = 16;   // noise channels among the node features

int64_t mp_nparams(int layers);
// output block X (n x 32, row-major FP32, device memory) for the graph of A
int mp_infer_device(speigs_hip_model *m, int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                    uint64_t seed, float *Xdev);
// the same from the solver's device SELL matrix of A (no host graph build)
struct SellA;
int mp_infer_sell(speigs_hip_model *m, const SellA &A, uint64_t seed, float *Xdev);

}  // namespace sehip
