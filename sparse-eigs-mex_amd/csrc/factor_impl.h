// factor_impl.h -- internal state of the device supernodal factor, shared by
// factor.hip (factorization, blocked solves), blocksolve.hip (in-place inverse
// of the diagonal blocks, multi-right-hand-side solves) and lobpcg.cpp.
#pragma once
#include <vector>

#include "factor.h"

namespace sehip {

constexpr int TB = 64;   // tile edge
constexpr int KB = 4;    // panels per outer block of the Schur updates (K = 256)

struct EaTask {   // extend-add: child front columns [cb*64, cb*64+64) into the parent
    int32_t child, parent, cb, pad;
};

struct PcbTask {   // a block of a front: (front, block index)
    int32_t parent, pcb;
};
struct EaBlk {     // extend-add target: parent front, 64-column block, 256-row block
    int32_t parent, cb, rb, pad;
};

struct TileTask {   // panel k of front s; tile row i, tile column j
    int32_t s, k, i, j;
};

struct SolveArgs {
    const int64_t *foff, *wsoff, *col0, *rptr, *relptr, *chptr;
    const int32_t *Fs, *ncol, *rows, *rel, *chl;
    double *ws;
};

struct BlkTask {
    int32_t s, b;
};

struct UpdTask {
    int32_t s, p, j, pad;
};

constexpr int64_t kXinvSlots = 2048;   // diagonal-tile inverses held at once (64 x 64 each)

struct LevelPlan {
    std::vector<int64_t> ea_off;      // per child ordinal: offset in the extend-add task array
    std::vector<int64_t> ea_cnt;
    std::vector<int64_t> k_potrf_off, k_potrf_cnt, k_trsm_off, k_trsm_cnt, k_upd_off, k_upd_cnt, k_upref_off;
    // the outer (block-end) updates: the next block's columns (out) on the
    // factor stream, the columns right of it (out2) on the look-ahead stream
    std::vector<int64_t> k_out_off, k_out_cnt, k_outref_off;
    std::vector<int64_t> k_out2_off, k_out2_cnt, k_out2ref_off;
    // per panel step: start of each group of kXinvSlots fronts in the TRSM tasks
    // (plus the end); a group's inverses fit the scratch, slot = ordinal % kXinvSlots
    std::vector<std::vector<int64_t>> k_trsm_cut;
    int64_t fr_off = 0, fr_cnt = 0;   // fronts of the level in the fronts array
    std::vector<int64_t> sea_off, sea_cnt;   // (unused) per-ordinal solve extend-add
    int64_t pcb_off = 0, pcb_cnt = 0;        // factor extend-add: parent (column, row) blocks
    int64_t fin_off = 0, fin_cnt = 0;        // solve init: (front, 256-row chunk)
    int64_t fb_off = 0, fb_cnt = 0, bb_off = 0, bb_cnt = 0;   // blocked solve work items
    int64_t ub_off = 0, ub_cnt = 0;                            // backward update-row chunks
};

struct Factor {
    Symbolic S;
    FactorOptions opt;
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;              // look-ahead: the far part of block-end updates
    hipEvent_t ev_ready = nullptr, ev_far = nullptr;
    hipEvent_t ev_zero = nullptr;    // the pool's zero fill queued at set-up (on st2)
    bool pool_zeroed = false;        // the next numeric phase waits for ev_zero instead of zeroing
    DevMem mem;
    int64_t n = 0, nnzA = 0;
    bool fp32 = false;
    void *pool = nullptr;
    double *Ax = nullptr, *diagA = nullptr;
    int64_t *Apd = nullptr, *Aid = nullptr, *iperm = nullptr;   // A's CSC arrays, iperm[old] = new
    int32_t *snode = nullptr;                                     // supernode of each permuted column
    int64_t *bndptr = nullptr;                                    // extend-add block boundaries (k_ea_bounds)
    int32_t *bnd = nullptr;
    int64_t *foff = nullptr, *wsoff = nullptr, *col0 = nullptr, *rptr = nullptr, *relptr = nullptr, *chptr = nullptr;
    int32_t *Fs = nullptr, *ncol = nullptr, *rows = nullptr, *rel = nullptr, *chl = nullptr, *fronts = nullptr;
    int64_t *perm = nullptr;
    EaTask *ea = nullptr;
    TileTask *tiles = nullptr;
    int32_t *ufid = nullptr;          // update tiles: front ordinal in the step's POTRF list
    int64_t *upref = nullptr;         // per step and front: first update tile
    std::vector<LevelPlan> plan;
    int *flag = nullptr;
    double *ws = nullptr, *xp = nullptr, *yp = nullptr, *r = nullptr, *d = nullptr, *ax = nullptr, *partial = nullptr;
    double *hpart = nullptr;
    int64_t wssize = 0;
    double delta = 0.0;
    const SellA *A = nullptr;
    int64_t solves = 0, ir_steps = 0;
    bool factored = false;
    bool use_mfma = true;
    bool use_reg_panel = true;       // register POTRF + TRSM as a product with L_kk^-1
    int potrf_waves = 4;             // register POTRF on 4 waves (columns split) or 1
    bool potrf_blocked = true;       // blocked register POTRF (16-column blocks, MFMA updates)
    int solve_mode = 3;               // 0: one workgroup per front, 1: blocked multi-workgroup,
                                      // 2: generic inverse-based, 3: inverse-based on the block plan
    bool wait_test = false;           // test knob: the blocked solve's waits all give up
    bool plan_ready = false;          // the single-vector solves' plan is on the device
    EaTask *sea = nullptr;
    PcbTask *fin = nullptr;
    EaBlk *eab = nullptr;             // factor extend-add blocks
    UpdTask *upd = nullptr;
    double *bpart = nullptr;
    int64_t *partoff = nullptr;
    unsigned *tick_f = nullptr, *tick_b = nullptr;
    BlkTask *fblk = nullptr, *bblk = nullptr;
    int64_t *panbase = nullptr;
    int *fflag = nullptr, *bflag = nullptr, *err = nullptr, *herr = nullptr;
    unsigned *counters = nullptr;
    int epoch = 0;
    // HIP events around each solve pair, harvested at the next apply (the
    // Krylov loop synchronizes between applies) or at factor_stats
    std::vector<hipEvent_t> ev;
    std::vector<int> ev_pending;      // indices of recorded pairs not yet read
    int ev_next = 0;
    double t_solve = 0.0, t_invert = 0.0;
    int64_t block_solves = 0, block_rhs = 0;   // multi-right-hand-side solves (FP32 vectors)
    bool inverted = false;            // L11 of every front holds its inverse
    // signed factorization A = L S L^T of an indefinite matrix (the fallback
    // where no shift makes A + delta I positive definite): the column signs,
    // the pivot floor and how many pivots it replaced
    bool signed_fac = false;
    void *sgn = nullptr;
    double tau = 0.0;
    int perturbed = 0;
    void *xinv = nullptr;             // per panel launch: inverses of the diagonal tiles (TRSM)
    void *rs = nullptr, *bs_d = nullptr;     // block-solve state (row-major FP32 / generic FP64 vectors)
    ~Factor()
    {
        if (st2) (void)hipStreamSynchronize(st2);
        if (ev_ready) (void)hipEventDestroy(ev_ready);
        if (ev_far) (void)hipEventDestroy(ev_far);
        if (ev_zero) (void)hipEventDestroy(ev_zero);
        if (st2) (void)hipStreamDestroy(st2);
        if (hpart) (void)hipHostFree(hpart);
        if (herr) (void)hipHostFree(herr);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
};

// HIP event pair around one solve on F->st, accumulated into t_solve
int ev_begin(Factor *F);
void ev_end(Factor *F, int i);

// in-place inverse of every front's L11 (chain-free solves); 0 ok
int factor_invert_l11(Factor *F);
// X = (L L^T)^-1 B for nb <= 32 right-hand sides, FP32 arithmetic, after
// factor_invert_l11; element (i, v) of B at B[i * brs + v * bcs] (likewise X,
// which may alias B with the same strides)
int factor_block_solve_rows(Factor *F, const double *B, int64_t brs, int64_t bcs, double *X, int64_t xrs,
                            int64_t xcs, int nb);
// starts building the row solve's plan on a helper thread (it reads only the
// symbolic structure); the first factor_block_solve_rows joins it
void rowsolve_prepare_async(Factor *F);
// the single-right-hand-side solve over the same plan and the inverted L11
// (solve mode 3): prepare joins the plan and inverts (allocates nothing after
// the plan's upload); solve takes device FP64 vectors in the original order
int factor_vec_prepare(Factor *F);
bool factor_vec_prepared(const Factor *F);
int factor_vec_solve(Factor *F, const double *b, double *x);
void rowsolve_free(Factor *F);
int factor_block_solve_d(Factor *F, const double *B, double *X, int nb);
void factor_block_free(Factor *F);

// block eigensolver preconditioned by the FP32 factor (lobpcg.cpp): one
// block of m = k + 8 vectors (fewer when n is small), m <= 32, the columns of
// one multi-right-hand-side solve; lobpcg_block_size is m, or 0 when k does not
// fit (the driver then runs Lanczos over the same FP32 factor with refinement)
constexpr int kLobpcgMaxBlock = 32;
int lobpcg_block_size(int64_t n, int64_t k);
// Xstart: optional proposed start block (device, n x 32 row-major FP32)
int lobpcg_dev(const SellA &A, Factor *F, hipStream_t st, int64_t n, int64_t k, const speigs_opts *o,
               const speigs_hip_opts *ho, double anorm, double *lambda, double *Vout, double *resid,
               speigs_info *info, speigs_hip_info *hinfo, const float *Xstart);

}  // namespace sehip
