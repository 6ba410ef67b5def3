/*
 * speigs_hip.h -- MI355X extension of the speigs C API (libspeigs_hip.so).
 *
 * speigs.h stays ABI-identical to the reference; everything specific to the
 * GPU build lives here, in separate structs, so stack-allocated speigs_opts /
 * speigs_info of existing callers keep their layout.
 *
 * Two levels are exported, both plain C (no C++ or HIP types cross):
 *   1. speigs_ex(): speigs() plus precision / start / telemetry knobs.
 *   2. the device operator (speigs_hip_op_*): the reference's pluggable
 *      operator seam -- speigs_op, op_apply, op_multA at
 *      src/speigs_core.c:57-116 -- as an opaque handle owning the device CSR
 *      matrix and the device supernodal factor.
 * Error convention as in speigs.h: 0 or a negative SPEIGS_ERR_* code; nothing
 * is printed and no process state is changed.  Every call is re-entrant: all
 * state is per call or per handle, and each handle owns its HIP stream.
 */
#ifndef SPEIGS_HIP_H
#define SPEIGS_HIP_H

#include <stdint.h>
#include "speigs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* arithmetic of the shift-invert factor */
#define SPEIGS_HIP_FP64     0   /* FP64 factor and solves: the drop-in path            */
#define SPEIGS_HIP_FP32_IR  1   /* FP32 factor on MFMA, each apply refined to FP64 by
                                 * iterative refinement on an FP64 SpMV residual      */
#define SPEIGS_HIP_FP32_BLOCK 2 /* FP32 factor (L11 blocks inverted) as the preconditioner
                                 * of a block eigensolver (LOBPCG, block size <= 32): one
                                 * multi-right-hand-side FP32 solve per iteration, Ritz
                                 * pairs and residuals against A in FP64 ('sm', A+dI SPD) */

/* start vector of the Lanczos recurrence */
#define SPEIGS_HIP_START_UNIFORM  0  /* xorshift64* uniform [-0.5,0.5): the reference's
                                      * start (src/speigs_core.c:185-191)               */
#define SPEIGS_HIP_START_VECTOR   1  /* caller-supplied start_vec[n]                    */
#define SPEIGS_HIP_START_LEARNED  2  /* learned start block (falls back to uniform)     */

typedef struct {
    int           precision;    /* SPEIGS_HIP_FP64 (default), _FP32_IR or _FP32_BLOCK        */
    double        ir_tol;       /* refined apply: ||b - (A+dI)x|| <= ir_tol ||b||; 0 -> 1e-13 */
    int           ir_maxit;     /* refinement sweeps per apply; 0 -> 12                      */
    int           start;        /* SPEIGS_HIP_START_*                                        */
    int           start_block;  /* learned start: number of proposed vectors; 0 -> k        */
    int           use_graph;    /* 1 -> capture the triangular solves in a hipGraph          */
    int           device;       /* HIP device ordinal; default 0                             */
    int           verbose;      /* 0 silent                                                  */
    const double *start_vec;    /* n values for SPEIGS_HIP_START_VECTOR                      */
    const struct speigs_hip_model *model;   /* trained start network for
                                             * SPEIGS_HIP_START_LEARNED (NULL -> uniform) */
    int64_t       reserved[7];
} speigs_hip_opts;

typedef struct {
    int64_t nnz_L;          /* stored entries of the supernodal factor (dense fronts)  */
    double  factor_flops;   /* flops of the numeric factorization                      */
    int64_t nsuper;         /* supernodes (fronts)                                     */
    int64_t nlevels;        /* levels of the supernodal elimination tree               */
    int64_t ir_steps;       /* refinement corrections over all applies                 */
    int64_t solves;         /* triangular solve pairs executed (block path: multi-
                             * right-hand-side solves, each over <= 32 columns)       */
    int     precision_used; /* factor precision actually used                          */
    int     fallback;       /* 1 if FP32 factor/refinement failed and FP64 took over    */
    double  t_order, t_symbolic, t_upload, t_numeric, t_solve, t_spmv;
    double  bytes_solve;    /* algorithmic bytes of one solve pair (factor + vectors)  */
    double  bytes_spmv;     /* algorithmic bytes of one SpMV                           */
    double  t_invert;       /* in-place inversion of the diagonal blocks (block path)  */
    double  perturbed_pivots; /* indefinite fallback (path 5): pivots of the signed
                             * factorization A = L S L^T replaced by +-sqrt(eps)||A||_1
                             * (the applies are refined against A)                  */
    double  reserved[6];
} speigs_hip_info;

void speigs_hip_default_opts(speigs_hip_opts *o);

/* speigs() with the extension knobs; hopts/hinfo may be NULL. */
int speigs_ex(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, const double *Ax,
              speigs_int k, speigs_which which, const speigs_opts *opts,
              const speigs_hip_opts *hopts, double *lambda, double *V, double *resid,
              speigs_info *info, speigs_hip_info *hinfo);

/* number of visible HIP devices (0 when none); never fails */
int speigs_hip_device_count(void);

/* Where the calling thread's most recent SPEIGS_ERR_INTERNAL / _MEM came from:
 * "file:line: failing step (code c)", with the HIP runtime's error name for a
 * failed HIP call, or "" if the last speigs_ex / speigs_hip_op_* call on this
 * thread recorded none.  No counterpart in the reference (whose only internal
 * failures are CHOLMOD's and LAPACK's status codes); for diagnosis only. */
const char *speigs_hip_last_error(void);

/* Device memory freed at the end of a call is kept per device and reused by
 * later calls (up to SPEIGS_HIP_CACHE_MB megabytes per device, default a
 * quarter of its memory; 0 disables keeping).  This returns every kept block
 * to the HIP runtime.  No counterpart in the reference (its solver runs on the
 * host); for callers that share the device with other allocators. */
void speigs_hip_release_cache(void);

/* Ordering + supernodal symbolic analysis only: host code, no device needed.
 * Replaces cholmod_l_analyze (src/speigs_core.c:933-936).  perm[n] (may be
 * NULL) receives the fill-reducing order, perm[new] = old; snode[n] (may be
 * NULL) the supernode of each new column; hinfo receives nnz_L, factor_flops,
 * nsuper, nlevels, t_order and t_symbolic. */
int speigs_hip_analyze(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                       speigs_int *perm, speigs_int *snode, speigs_hip_info *hinfo);

/* The reverse Cuthill-McKee order the banded gate uses (src/speigs_core.c:471-503),
 * perm[new] = old.  Host code, no device needed. */
int speigs_hip_rcm(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, speigs_int *perm);

/* ---- device operator: the pluggable seam (src/speigs_core.c:57-116) ------ */
typedef struct speigs_hip_op speigs_hip_op;

#define SPEIGS_HIP_OP_MULT  0   /* y = A x                        (op_multA, mode 0) */
#define SPEIGS_HIP_OP_SOLVE 1   /* y = (A + delta I)^-1 x         (op_apply, mode 1) */

/* Build the operator for A (CSC, upper triangle read).  For OP_SOLVE this runs
 * ordering + symbolic analysis on the host and the numeric factorization on the
 * device; returns SPEIGS_ERR_FACTOR if A + delta I is not positive definite. */
int  speigs_hip_op_create(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                          const double *Ax, int mode, int precision, double delta,
                          speigs_hip_op **op, speigs_hip_info *hinfo);
/* refactor with a new shift (same pattern) */
int  speigs_hip_op_refactor(speigs_hip_op *op, double delta);
/* y[:,j] = M x[:,j] for nvec host vectors (column-major, leading dim n).
 * precision SPEIGS_HIP_FP32_BLOCK: one unrefined FP32 multi-right-hand-side
 * solve per 32 columns (accuracy ~kappa * 6e-8, the block eigensolver's
 * preconditioner); FP64 and FP32_IR solve each column to FP64 accuracy. */
int  speigs_hip_op_apply(speigs_hip_op *op, const double *x, double *y, speigs_int nvec);
/* y = A x (always A, never the shifted matrix) */
int  speigs_hip_op_multA(speigs_hip_op *op, const double *x, double *y, speigs_int nvec);
/* speigs_hip_op_apply without any device allocation.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * host vectors are staged through a fixed 64 MiB array that every device holds
 * until the process ends, so n * nvec <= 4,194,304 (else SPEIGS_ERR_ARG, with
 * the reason in speigs_hip_last_error()); operators of precision
 * SPEIGS_HIP_FP32_BLOCK, whose solve plan is allocated at first use, are
 * refused with SPEIGS_ERR_ARG.  All calls on one device share that array:
 * calls that overlap on one device race. */
int  speigs_hip_op_apply_noalloc(speigs_hip_op *op, const double *x, double *y, speigs_int nvec);
void speigs_hip_op_destroy(speigs_hip_op *op);

/* ---- learned start: a message-passing network over the graph of A --------
 * No counterpart in the reference (its plan, docs/REQUIREMENTS.md, names a
 * learned starting block as future work).  Three layers of width 32 over the
 * graph with edge weights |a_ij| / a_ii; node features from the diagonal, the
 * off-diagonal row mass and 16 seeded noise channels; the 32 output channels
 * are the proposed start block (BF16 weights, FP32 arithmetic on the matrix
 * cores).  Trained on the GPU against eigenvectors of the plain solver; the
 * eigensolver's residual test still decides convergence, so the network can
 * change the iteration count but not the answer. */
typedef struct speigs_hip_model speigs_hip_model;

/* new network with Glorot-uniform weights from `seed` (layers 1..8, usually 3) */
int     speigs_hip_model_create(int layers, uint64_t seed, speigs_hip_model **model);
int64_t speigs_hip_model_nparams(const speigs_hip_model *model);
/* parameters: per layer W (32 x 64 row-major, [output][input], input = [H | A_hat H])
 * then b (32), FP32 */
int     speigs_hip_model_set_params(speigs_hip_model *model, const float *params, int64_t nparams);
int     speigs_hip_model_get_params(const speigs_hip_model *model, float *params, int64_t nparams);
/* X = the n x 32 output block (row-major FP32) for the graph of A */
int     speigs_hip_model_infer(speigs_hip_model *model, speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                               const double *Ax, uint64_t seed, float *X);
/* loss 1 - tr(S^-1 X^T U U^T X)/k (S = X^T X) against the k orthonormal columns
 * of U (n x k column-major) and its gradient (nparams floats; NULL skips it) */
int     speigs_hip_model_grad(speigs_hip_model *model, speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                              const double *Ax, const double *U, speigs_int k, uint64_t seed, double *loss,
                              float *grad);
/* one Adam step on that loss (learning rate lr) */
int     speigs_hip_model_train_step(speigs_hip_model *model, speigs_int n, const speigs_int *Ap,
                                    const speigs_int *Ai, const double *Ax, const double *U, speigs_int k,
                                    uint64_t seed, float lr, double *loss);
void    speigs_hip_model_destroy(speigs_hip_model *model);

#ifdef __cplusplus
}
#endif
#endif /* SPEIGS_HIP_H */
