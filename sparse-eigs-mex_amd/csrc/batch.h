// batch.h -- the batched component solver: many independent components (the
// diagonal blocks the driver splits a matrix into) share one device factor of
// their block-diagonal matrix and advance their thick-restart Lanczos
// recurrences in lockstep, one operator apply and one set of segmented vector
// kernels per step for all of them.
#pragma once
#include <vector>

#include "dev.h"

namespace sehip {

// rows [r0, r1) of component comp, r1 - r0 <= 1024
struct SegChunk {
    int32_t comp, pad;
    int64_t r0, r1;
};

struct SegLayout {
    int64_t ncomp = 0, nchunks = 0;
    std::vector<int64_t> cptr;        // chunks of component c: [cptr[c], cptr[c+1])
    SegChunk *chunks = nullptr;       // device
    int64_t *cptr_d = nullptr;        // device
};

// component c owns rows [off[c], off[c+1])
int seg_layout_build(const std::vector<int64_t> &off, SegLayout &L, DevMem &mem, hipStream_t st);

// Per-component arrays (device, indexed by component): mcomp / pcomp / cnt =
// active column counts (0 skips the component), act = participation flag,
// jcol / jdst = column index or -1, h / lam / s / out = ps-strided rows.
void seg_dots(const SegLayout &L, const double *V, int64_t ld, const int32_t *mcomp, const double *w, double *partial,
              int ps, hipStream_t st);
void seg_reduce(const SegLayout &L, const double *partial, int ps, const int32_t *cnt, double *out, hipStream_t st);
void seg_update(const SegLayout &L, const double *V, int64_t ld, const int32_t *act, const int32_t *mcomp,
                const double *h, double *w, double *partial, int ps, int mode, hipStream_t st);
void seg_gather_col(const SegLayout &L, const double *V, int64_t ld, const int32_t *jcol, double *x, hipStream_t st);
void seg_store_col(const SegLayout &L, const double *x, const double *s, const int32_t *jdst, double *V, int64_t ld,
                   hipStream_t st);
void seg_gemm(const SegLayout &L, const double *V, int64_t ld, const int32_t *mcomp, const int32_t *pcomp,
              const double *S, int64_t ss, double *X, int64_t ldx, hipStream_t st);
void seg_pairs(const SegLayout &L, const double *X, int64_t ldx, const double *Y, int64_t ldy, const int32_t *pcomp,
               const double *lam, double *partial, int ps, int mode, hipStream_t st);
void seg_scale_cols(const SegLayout &L, double *X, int64_t ldx, const int32_t *pcomp, const double *s, int ps,
                    hipStream_t st);
void seg_copy_cols(const SegLayout &L, const double *Sv, double *D, int64_t ld, const int32_t *cnt, hipStream_t st);

// one component of the batch: rows [off, off + n) of the batch vector
struct BatchComp {
    int64_t off = 0, n = 0, k = 0;
    double anorm = 1.0;
    // results
    std::vector<double> lambda, resid, V;   // V: n x k column-major (if requested)
    int64_t nops = 0, restarts = 0, nconv = 0;
    int flag = 0;
};

// Thick-restart Lanczos (reference src/speigs_core.c:146-315) for every
// component at once over the block-diagonal device operator op (its vectors
// are the batch vector).
int lanczos_batched(DevOperator *op, int64_t n, std::vector<BatchComp> &comps, const speigs_opts *o,
                    bool want_vectors);

}  // namespace sehip
