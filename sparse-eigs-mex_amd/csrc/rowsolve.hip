// rowsolve.hip -- the multi-right-hand-side solve of the block eigensolver:
// X = (L L^T)^-1 B for up to 32 right-hand sides, FP32 vectors, with the
// diagonal blocks L11 of every front already inverted in place (blocksolve.hip).
//
// Vectors are row-major with a row of 32 floats (128 bytes, right-hand sides
// beyond nb are zero), so gathering a row by index -- the permutation, the
// children's update rows, X at a front's update rows -- is one contiguous
// line.  One buffer holds, by rows,
//     [0, n)              P  = the permuted right-hand sides
//     [n, n + sum r)      U  = each front's update block (r rows per front)
//     [n + sum r, ...)    T  = per-column scratch: W1 (forward), Z (backward)
// plus Y and X (n rows each).  Front s (c columns, r update rows), forward:
//     W  = [P_c; 0] + sum of the children's U rows at their places in s
//          (a precomputed pull list per front row, children in order, so the
//          sums are deterministic and need no atomics).  Each W row is summed
//          where it is read: W2 rows in the epilogue that writes U_s, W1 rows
//          in the loads of the fronts whose columns fit one tile; only the
//          W1 of wider fronts (read by several output tiles) is assembled
//          into T by one launch per level
//     Y  = L11^-1 W1,   U_s = W2 - L21 Y
// backward:
//     Z  = Y - L21^T X_R,   X_c = L11^-T Z      (X_R = X at the update rows,
//          read by index inside the product)
// All four products run on v_mfma_f32_16x16x4f32 with the vector rows as the
// second operand: lane l of a product holds column l%16 of the right-hand
// sides, so every vector access is 16 consecutive floats.
//
// Fronts with c <= 64 and r <= 64 (most of the tree) are solved whole by one
// wave each, Y / Z staying in LDS.  The others use 64-row output tiles: a
// task is one tile and up to RKCHUNK K tiles, each wave owns 16 rows and
// keeps the next K tile's loads in flight while it multiplies the current
// one; tasks of a split K range leave partials that a finish kernel sums in
// a fixed order.  The transposed products (Z, X_c) read L along columns
// coalesced and transpose each wave's 16 x 64 slice through LDS.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "factor_impl.h"
#include "host_par.h"

namespace sehip {

namespace {

constexpr int RW = 32;            // floats per vector row
constexpr int RTILE = 64;         // K tile depth; output rows per tile of the X11 products
// The two L21 products (modes 1, 2) run with 64- or 128-row output tiles,
// chosen per level: with 128 (each wave 32 rows) a K tile's vector rows,
// staged once per workgroup, are read half as often, but there are half the
// workgroups and the transposed product's LDS allows two per CU -- all 128 was
// 3.64 -> 4.20 ms per solve at 100^3.  Both give the same result to the bit
// (the same K ranges and MFMA order per output element), so the choice can be
// made by timing (the autotuner below) without changing any answer; at 100^3
// it kept 64 rows at every level, and its two timed solves cost ~1.5 ms, so
// 64 is the default and the autotuner runs on request.
constexpr int RT_WIDE = 128;
constexpr int RKCHUNK = 4;        // K tiles per part task
constexpr int RBP = 36;           // LDS pitch of a staged vector row (rows k and k + 4 on other banks)
constexpr int TINY = 64;          // wave-per-front limit on c and r
constexpr int RP = 36;            // LDS row pitch (floats)

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct RsArgs {
    const int64_t *foff, *col0, *rptr, *chptr, *frow, *uoff, *lptr;
    const int32_t *Fs, *ncol, *rows, *lst;
    int64_t n, t0;               // buffer row of U's first row (= n), of T's first row
};

struct RsPart {                  // output tile I of front s, K tiles [k0, k1)
    int32_t s, I, k0, k1, direct, pad;
};
struct RsFin {
    int32_t s, I;
    int64_t p0;                  // first part (relative to the launch)
    int32_t np, pad;
};
struct RsRow {
    int32_t s, I;                // front, 32-row chunk
};

__device__ __forceinline__ f32x4 mfma4(float a, float b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// exchange through LDS between the lanes of one wave: the LDS executes one
// wave's instructions in order, so a read issued after a write sees it (and a
// later write cannot overtake an earlier read); the wave barrier keeps the
// compiler from moving the accesses across this point.  No memory fence: it
// would also wait for the global loads kept in flight across it.
__device__ __forceinline__ void wave_lds_sync() { __builtin_amdgcn_wave_barrier(); }

// Row rid of the fronts' right-hand side W (rows of front s from frow[s]):
// the sum, in pull-list order, of its P row (rows < c) and of the children's
// update rows that land on it, for the W1 rows of the fronts whose columns
// fit one tile (read by one product task only, so they are summed in that
// task's loads rather than assembled).  The list bounds, the first kPull
// indices and those rows are in flight together; longer lists continue one
// entry at a time.  8 consecutive elements from col (two 16-byte loads per
// list entry).
constexpr int kPull = 4;

__device__ __forceinline__ void pull_sum8(const RsArgs &a, const float *__restrict__ buf, int64_t rid, int col,
                                          float4 (&o)[2])
{
    const int64_t q0 = a.lptr[rid], q1 = a.lptr[rid + 1];
    int32_t id[kPull];
#pragma unroll
    for (int u = 0; u < kPull; u++) id[u] = q0 + u < q1 ? a.lst[q0 + u] : -1;
    float4 u0[kPull], u1[kPull];
#pragma unroll
    for (int u = 0; u < kPull; u++) {
        const float *src = buf + (int64_t)(id[u] < 0 ? 0 : id[u]) * RW + col;
        u0[u] = id[u] < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : *reinterpret_cast<const float4 *>(src);
        u1[u] = id[u] < 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : *reinterpret_cast<const float4 *>(src + 4);
    }
    float4 x0 = make_float4(0.f, 0.f, 0.f, 0.f), x1 = x0;
    auto add = [&](const float4 &p, const float4 &q) {
        x0.x += p.x;
        x0.y += p.y;
        x0.z += p.z;
        x0.w += p.w;
        x1.x += q.x;
        x1.y += q.y;
        x1.z += q.z;
        x1.w += q.w;
    };
#pragma unroll
    for (int u = 0; u < kPull; u++)
        if (id[u] >= 0) add(u0[u], u1[u]);
    for (int64_t q = q0 + kPull; q < q1; q++) {
        const float *src = buf + (int64_t)a.lst[q] * RW + col;
        add(*reinterpret_cast<const float4 *>(src), *reinterpret_cast<const float4 *>(src + 4));
    }
    o[0] = x0;
    o[1] = x1;
}

// ---- tiny fronts (c <= 64, r <= 64): one wave per front ----------------------
// Operand layout (MFMA step (kq, t), lane l = 16 g + lc): the product index is
// k = 16 kq + 4 g + t, the first operand holds A[16 rb + lc][k], the second
// V[k][16 h + lc]; the result register v of lane l is row 16 rb + 4 g + v,
// right-hand side 16 h + lc.  Every load of a step is issued before the first
// product (no stores in between), so a front costs a few memory latencies.
template <class T, int NQ, int NR>
__global__ void __launch_bounds__(256) k_rs_tiny_fwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                                     int nf, RsArgs a, float *__restrict__ buf,
                                                     float *__restrict__ Yg)
{
    __shared__ float Ys[4][16 * NQ][RP];   // the wave's Y rows (LDS sized to the class)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, lc = lane & 15;
    const int fi = blockIdx.x * 4 + wv;
    if (fi >= nf) return;       // per wave; no workgroup barrier below
    const int s = fronts[fi];
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s], ub = a.n + a.uoff[s];
    const bool leaf = a.chptr[s + 1] == a.chptr[s];   // not assembled: W1 = P rows, W2 = 0
    const int64_t tb = leaf ? c0 : a.t0 + c0;
    float (*Yw)[RP] = Ys[wv];

    // NQ, NR >= the front's 16-row blocks of c and r (the launch's class);
    // blocks past c or r hold zeros, so their products add exact zeros
    float wb[NQ][4][2], w2[NR][4][2];
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * kq + 4 * g + t;
            const bool in = k < c;
            wb[kq][t][0] = in ? buf[(tb + k) * RW + lc] : 0.f;
            wb[kq][t][1] = in ? buf[(tb + k) * RW + 16 + lc] : 0.f;
        }
#pragma unroll
    for (int kq = 0; kq < NR; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int ur = 16 * kq + 4 * g + t;       // (rb = kq, v = t) of the U rows
            const bool iu = !leaf && ur < r;
            w2[kq][t][0] = iu ? buf[(ub + ur) * RW + lc] : 0.f;
            w2[kq][t][1] = iu ? buf[(ub + ur) * RW + 16 + lc] : 0.f;
        }
    // Y = X11 W1 (X11 lower triangular: A[row][k] = L(row, k), k <= row)
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
        const int row = 16 * rb + lc;
#pragma unroll
        for (int kq = 0; kq <= rb; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                const float av = (row < c && k <= row) ? (float)L[(int64_t)k * F + row] : 0.f;
                acc0 = mfma4(av, wb[kq][t][0], acc0);
                acc1 = mfma4(av, wb[kq][t][1], acc1);
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int yr = 16 * rb + 4 * g + v;
            Yw[yr][lc] = acc0[v];
            Yw[yr][16 + lc] = acc1[v];
            if (yr < c) {
                Yg[(c0 + yr) * RW + lc] = acc0[v];
                Yg[(c0 + yr) * RW + 16 + lc] = acc1[v];
            }
        }
    }
    if (r <= 0) return;
    wave_lds_sync();
    float yb[NQ][4][2];
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * kq + 4 * g + t;
            yb[kq][t][0] = Yw[k][lc];
            yb[kq][t][1] = Yw[k][16 + lc];
        }
    // U = W2 - L21 Y
#pragma unroll
    for (int rb = 0; rb < NR; rb++) {
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
        const int lrow = 16 * rb + lc;
#pragma unroll
        for (int kq = 0; kq < NQ; kq++) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                const float av = (lrow < r && k < c) ? (float)L[(int64_t)k * F + c + lrow] : 0.f;
                acc0 = mfma4(av, yb[kq][t][0], acc0);
                acc1 = mfma4(av, yb[kq][t][1], acc1);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int ur = 16 * rb + 4 * g + v;
            if (ur < r) {
                buf[(ub + ur) * RW + lc] = w2[rb][v][0] - acc0[v];
                buf[(ub + ur) * RW + 16 + lc] = w2[rb][v][1] - acc1[v];
            }
        }
    }
}

template <class T, int NQ, int NR>
__global__ void __launch_bounds__(256) k_rs_tiny_bwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                                     int nf, RsArgs a, const float *__restrict__ Yg,
                                                     float *__restrict__ Xg)
{
    __shared__ float Zs[4][16 * NQ][RP];   // the wave's Z rows (LDS sized to the class)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, lc = lane & 15;
    const int fi = blockIdx.x * 4 + wv;
    if (fi >= nf) return;
    const int s = fronts[fi];
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s];
    const int32_t *R = a.rows + a.rptr[s];
    float (*Zw)[RP] = Zs[wv];

    // NQ, NR >= the front's 16-row blocks of c and r (the launch's class);
    // blocks past c or r hold zeros, so their products add exact zeros
    float xb[NR][4][2], yv[NQ][4][2];
#pragma unroll
    for (int kq = 0; kq < NR; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * kq + 4 * g + t;
            const bool in = k < r;
            const int64_t gr = in ? (int64_t)R[k] : 0;
            xb[kq][t][0] = in ? Xg[gr * RW + lc] : 0.f;
            xb[kq][t][1] = in ? Xg[gr * RW + 16 + lc] : 0.f;
        }
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * kq + 4 * g + t;
            const bool iy = k < c;          // (rb = kq, v = t) of the Y rows
            yv[kq][t][0] = iy ? Yg[(c0 + k) * RW + lc] : 0.f;
            yv[kq][t][1] = iy ? Yg[(c0 + k) * RW + 16 + lc] : 0.f;
        }
    // Z = Y - L21^T X_R (A[j][k] = L(c + k, j), k over the r update rows)
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
        const int j = 16 * rb + lc;
#pragma unroll
        for (int kq = 0; kq < NR; kq++) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                const float av = (j < c && k < r) ? (float)L[(int64_t)j * F + c + k] : 0.f;
                acc0 = mfma4(av, xb[kq][t][0], acc0);
                acc1 = mfma4(av, xb[kq][t][1], acc1);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int zr = 16 * rb + 4 * g + v;
            Zw[zr][lc] = yv[rb][v][0] - acc0[v];
            Zw[zr][16 + lc] = yv[rb][v][1] - acc1[v];
        }
    }
    wave_lds_sync();
    float zb[NQ][4][2];
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int k = 16 * kq + 4 * g + t;
            zb[kq][t][0] = Zw[k][lc];
            zb[kq][t][1] = Zw[k][16 + lc];
        }
    // X_c = X11^T Z (A[j][k] = L(k, j), k >= j)
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
        const int j = 16 * rb + lc;
#pragma unroll
        for (int kq = rb; kq < NQ; kq++) {
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                const float av = (j < c && k < c && k >= j) ? (float)L[(int64_t)j * F + k] : 0.f;
                acc0 = mfma4(av, zb[kq][t][0], acc0);
                acc1 = mfma4(av, zb[kq][t][1], acc1);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int xr = 16 * rb + 4 * g + v;
            if (xr < c) {
                Xg[(c0 + xr) * RW + lc] = acc0[v];
                Xg[(c0 + xr) * RW + 16 + lc] = acc1[v];
            }
        }
    }
}

// ---- narrow fronts of the lowest levels (c <= 64 columns, r <= 512) --------
// A front whose columns fit one 64-row tile is solved by waves that read its
// right-hand side W where it is needed: the W rows of a front with children
// are summed from their pull lists inside the loads (no assembly pass), a
// leaf's W1 is its P rows and its W2 zero.  Vector rows go through the wave's
// own LDS in row layout (lane l: row l / 4, floats 8 (l % 4) .. +7, two
// 16-byte loads per source row) and are read back in MFMA operand layout.
// NQ = 16-row blocks of c (1, 2 or 4: the launch's class); blocks past c hold
// zeros, so their products add exact zeros.  Operand layout as for every
// product here (MFMA step (kq, t), lane l = 16 g + lc): the product index is
// k = 16 kq + 4 g + t, the first operand A[16 rb + lc][k], the second
// V[k][16 h + lc]; result register v of lane l is row 16 rb + 4 g + v,
// right-hand side 16 h + lc.
// A narrow front's task carries everything its first loads need, read by
// scalar loads (the task index is wave-uniform): the chain before the first
// operand load is one load instead of an index, then the front's fields.
struct RsNTask {                 // forward: update rows [u0, u0 + NTR) of a front
    int64_t foff, c0, fr, ub;    // factor panel, first column, first row id, first U row of the task
    int32_t F, c, u0, leaf;
};
struct RsNFront {                // backward: a front
    int64_t foff, c0, rptr;      // factor panel, first column, first update-row index
    int32_t F, c;
};

constexpr int NTR = 64;          // update rows per forward task

// row i of front s's W (zero when i >= lim): a leaf's P row, else the sum of
// the row's pull list; 8 floats from col
__device__ __forceinline__ void w_row8(const RsArgs &a, const float *__restrict__ buf, int64_t fr, int64_t c0,
                                       bool leaf, int i, int lim, int col, float4 (&o)[2])
{
    o[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    o[1] = o[0];
    if (i >= lim) return;
    if (leaf) {
        const float *src = buf + (c0 + i) * RW + col;
        o[0] = *reinterpret_cast<const float4 *>(src);
        o[1] = *reinterpret_cast<const float4 *>(src + 4);
        return;
    }
    pull_sum8(a, buf, fr + i, col, o);
}

__device__ __forceinline__ void stage8(float (*S)[RP], int row, int col, const float4 (&v)[2])
{
    *reinterpret_cast<float4 *>(&S[row][col]) = v[0];
    *reinterpret_cast<float4 *>(&S[row][col + 4]) = v[1];
}

// forward, one wave per task (front s, update rows [u0, u0 + NTR)):
//   Y = X11 W1 (every task; the task with u0 = 0 stores it), U = W2 - L21 Y
//   for the task's update rows, 16 at a time with the next 16 rows' L21 and
//   W2 loads in flight during the current products
template <class T, int NQ>
__global__ void __launch_bounds__(256) k_rs_nfwd(const T *__restrict__ pool, const RsNTask *__restrict__ tasks, int nt,
                                                 RsArgs a, float *__restrict__ buf, float *__restrict__ Yg)
{
    __shared__ __align__(16) float Ws[4][16 * NQ][RP];   // the wave's W1, then Y
    __shared__ __align__(16) float Us[4][16][RP];        // W2 of the current 16 update rows
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, lc = lane & 15;
    const int ti = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wv);
    if (ti >= nt) return;       // per wave; no workgroup barrier below
    const RsNTask tk = tasks[ti];
    const int u0 = tk.u0;
    const int32_t F = tk.F, c = tk.c, r = F - c;
    const int nq = (c + 15) >> 4;                     // 16-row blocks of c (<= NQ)
    const int un = min(NTR, r - u0);                  // the task's update rows (<= 0: none)
    const int nbk = un > 0 ? (un + 15) >> 4 : 0;
    const T *L = pool + tk.foff;
    const int64_t c0 = tk.c0, fr = tk.fr, ub = tk.ub;
    const bool leaf = tk.leaf != 0;
    float (*Wq)[RP] = Ws[wv];
    float (*Uq)[RP] = Us[wv];
    const int srow = lane >> 2, scol = 8 * (lane & 3);

    // every load that does not need Y first: W1, X11, the first 16 rows' L21 and W2
    float4 w1[NQ][2];
#pragma unroll
    for (int q = 0; q < NQ; q++) w_row8(a, buf, fr, c0, leaf, 16 * q + srow, c, scol, w1[q]);
    float a11[NQ * (NQ + 1) / 2][4];
#pragma unroll
    for (int rb = 0; rb < NQ; rb++)
#pragma unroll
        for (int kq = 0; kq <= rb; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int row = 16 * rb + lc, k = 16 * kq + 4 * g + t;
                a11[rb * (rb + 1) / 2 + kq][t] = (row < c && k <= row) ? (float)L[(int64_t)k * F + row] : 0.f;
            }
    float a21[NQ][4];
    float4 w2[2];
    auto load_blk = [&](int b) {
        const int lrow = u0 + 16 * b + lc;            // update row of this lane's L21 operand
#pragma unroll
        for (int kq = 0; kq < NQ; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                a21[kq][t] = (lrow < r && k < c) ? (float)L[(int64_t)k * F + c + lrow] : 0.f;
            }
        w_row8(a, buf, fr, c0, leaf, c + u0 + 16 * b + srow, leaf ? 0 : F, scol, w2);
    };
    if (nbk > 0) load_blk(0);

    // Y = X11 W1
#pragma unroll
    for (int q = 0; q < NQ; q++) stage8(Wq, 16 * q + srow, scol, w1[q]);
    wave_lds_sync();
    float wb[NQ][4][2];
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            wb[kq][t][0] = Wq[16 * kq + 4 * g + t][lc];
            wb[kq][t][1] = Wq[16 * kq + 4 * g + t][16 + lc];
        }
    wave_lds_sync();            // W1 read before Y overwrites it
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        if (rb >= nq) break;          // blocks past c: zero rows, never read below
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
#pragma unroll
        for (int kq = 0; kq <= rb; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc0 = mfma4(a11[rb * (rb + 1) / 2 + kq][t], wb[kq][t][0], acc0);
                acc1 = mfma4(a11[rb * (rb + 1) / 2 + kq][t], wb[kq][t][1], acc1);
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int yr = 16 * rb + 4 * g + v;
            Wq[yr][lc] = acc0[v];
            Wq[yr][16 + lc] = acc1[v];
            if (u0 == 0 && yr < c) {
                Yg[(c0 + yr) * RW + lc] = acc0[v];
                Yg[(c0 + yr) * RW + 16 + lc] = acc1[v];
            }
        }
    }
    if (nbk == 0) return;
    wave_lds_sync();
    float yb[NQ][4][2];
#pragma unroll
    for (int kq = 0; kq < NQ; kq++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            yb[kq][t][0] = kq < nq ? Wq[16 * kq + 4 * g + t][lc] : 0.f;
            yb[kq][t][1] = kq < nq ? Wq[16 * kq + 4 * g + t][16 + lc] : 0.f;
        }

    // U = W2 - L21 Y, 16 update rows at a time
#pragma unroll
    for (int b = 0; b < NTR / 16; b++) {
        if (b >= nbk) break;
        // this block's operands, then the next block's loads (into the same
        // names: the compiler keeps both sets live, so no wait comes between)
        float ca[NQ][4];
#pragma unroll
        for (int kq = 0; kq < NQ; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) ca[kq][t] = a21[kq][t];
        const float4 cw[2] = {w2[0], w2[1]};
        if (b + 1 < nbk) load_blk(b + 1);
        stage8(Uq, srow, scol, cw);
        wave_lds_sync();
        float o0[4], o1[4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            o0[v] = Uq[4 * g + v][lc];
            o1[v] = Uq[4 * g + v][16 + lc];
        }
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
#pragma unroll
        for (int kq = 0; kq < NQ; kq++) {
            if (kq >= nq) break;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                acc0 = mfma4(ca[kq][t], yb[kq][t][0], acc0);
                acc1 = mfma4(ca[kq][t], yb[kq][t][1], acc1);
            }
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int ur = 16 * b + 4 * g + v;
            if (ur < un) {
                buf[(ub + ur) * RW + lc] = o0[v] - acc0[v];
                buf[(ub + ur) * RW + 16 + lc] = o1[v] - acc1[v];
            }
        }
        wave_lds_sync();        // this block's W2 read before the next block's is staged
    }
}

// backward: Z = Y - L21^T X_R over the update rows in 64-row chunks, then
// X_c = X11^T Z.  WPF waves per front: 1 (the wave takes every chunk, four
// fronts per workgroup) or 4 (wave w takes chunks w, w + 4, ...; the partial
// sums are added in wave order through LDS, and wave rb forms output block rb).
template <class T, int NQ, int WPF>
__global__ void __launch_bounds__(256) k_rs_nbwd(const T *__restrict__ pool, const RsNFront *__restrict__ fronts,
                                                 int nf, RsArgs a, const float *__restrict__ Yg,
                                                 float *__restrict__ Xg)
{
    __shared__ __align__(16) float Xs[4][64][RP];   // the wave's X_R chunk, then its (partial) Z
    __shared__ __align__(16) float Zs[WPF == 4 ? 16 * NQ : 1][RP];   // WPF 4: the front's Z
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, lc = lane & 15;
    const int fi = __builtin_amdgcn_readfirstlane(WPF == 1 ? blockIdx.x * 4 + wv : (int)blockIdx.x);
    if (fi >= nf) return;       // WPF 4: uniform over the workgroup (one front each)
    const int sub = WPF == 1 ? 0 : wv;
    const RsNFront fd = fronts[fi];
    const int32_t F = fd.F, c = fd.c, r = F - c;
    const T *L = pool + fd.foff;
    const int64_t c0 = fd.c0;
    const int32_t *R = a.rows + fd.rptr;
    float (*Xq)[RP] = Xs[wv];
    const int srow = lane >> 2, scol = 8 * (lane & 3);

    f32x4 acc[NQ][2];
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        acc[rb][0] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc[rb][1] = acc[rb][0];
    }
    const int nch = (r + 63) >> 6;
    for (int ch = sub; ch < nch; ch += WPF) {
        const int k0 = ch * 64;
        int32_t gi[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + 16 * q + srow;
            gi[q] = k < r ? R[k] : -1;
        }
        // L21^T operand: A[j][k] = L(c + k, j)
        float at[NQ][4][4];
#pragma unroll
        for (int rb = 0; rb < NQ; rb++)
#pragma unroll
            for (int kq = 0; kq < 4; kq++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int j = 16 * rb + lc, k = k0 + 16 * kq + 4 * g + t;
                    at[rb][kq][t] = (j < c && k < r) ? (float)L[(int64_t)j * F + c + k] : 0.f;
                }
        float4 xr[4][2];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            xr[q][0] = make_float4(0.f, 0.f, 0.f, 0.f);
            xr[q][1] = xr[q][0];
            if (gi[q] >= 0) {
                const float *src = Xg + (int64_t)gi[q] * RW + scol;
                xr[q][0] = *reinterpret_cast<const float4 *>(src);
                xr[q][1] = *reinterpret_cast<const float4 *>(src + 4);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) stage8(Xq, 16 * q + srow, scol, xr[q]);
        wave_lds_sync();
#pragma unroll
        for (int kq = 0; kq < 4; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const float b0 = Xq[16 * kq + 4 * g + t][lc], b1 = Xq[16 * kq + 4 * g + t][16 + lc];
#pragma unroll
                for (int rb = 0; rb < NQ; rb++) {
                    acc[rb][0] = mfma4(at[rb][kq][t], b0, acc[rb][0]);
                    acc[rb][1] = mfma4(at[rb][kq][t], b1, acc[rb][1]);
                }
            }
        wave_lds_sync();        // the chunk read before the next one is staged
    }

    float (*Zq)[RP] = WPF == 1 ? Xq : Zs;
    if (WPF == 1) {
        // Z = Y - L21^T X_R, into the wave's LDS
#pragma unroll
        for (int rb = 0; rb < NQ; rb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int zr = 16 * rb + 4 * g + v;
                const float y0 = zr < c ? Yg[(c0 + zr) * RW + lc] : 0.f;
                const float y1 = zr < c ? Yg[(c0 + zr) * RW + 16 + lc] : 0.f;
                Xq[zr][lc] = y0 - acc[rb][0][v];
                Xq[zr][16 + lc] = y1 - acc[rb][1][v];
            }
        wave_lds_sync();
    } else {
        // the partial sums to LDS, then 16 rows of Z per wave: Y minus the four
        // partials added in wave order
#pragma unroll
        for (int rb = 0; rb < NQ; rb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int zr = 16 * rb + 4 * g + v;
                Xq[zr][lc] = acc[rb][0][v];
                Xq[zr][16 + lc] = acc[rb][1][v];
            }
        __syncthreads();
        if (wv < NQ) {
            const int zr = 16 * wv + srow;
            float4 y[2];
            y[0] = make_float4(0.f, 0.f, 0.f, 0.f);
            y[1] = y[0];
            if (zr < c) {
                const float *src = Yg + (c0 + zr) * RW + scol;
                y[0] = *reinterpret_cast<const float4 *>(src);
                y[1] = *reinterpret_cast<const float4 *>(src + 4);
            }
            float4 z[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                float4 p = *reinterpret_cast<const float4 *>(&Xs[0][zr][scol + 4 * h]);
#pragma unroll
                for (int w = 1; w < 4; w++) {
                    const float4 q = *reinterpret_cast<const float4 *>(&Xs[w][zr][scol + 4 * h]);
                    p.x += q.x;
                    p.y += q.y;
                    p.z += q.z;
                    p.w += q.w;
                }
                z[h] = make_float4(y[h].x - p.x, y[h].y - p.y, y[h].z - p.z, y[h].w - p.w);
            }
            stage8(Zs, zr, scol, z);
        }
        __syncthreads();
    }

    // X_c = X11^T Z (A[j][k] = L(k, j), k >= j); WPF 4: wave rb forms block rb
#pragma unroll
    for (int rb = 0; rb < NQ; rb++) {
        if (WPF == 4 && rb != wv) continue;
        f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0;
        const int j = 16 * rb + lc;
#pragma unroll
        for (int kq = rb; kq < NQ; kq++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int k = 16 * kq + 4 * g + t;
                const float av = (j < c && k < c && k >= j) ? (float)L[(int64_t)j * F + k] : 0.f;
                acc0 = mfma4(av, Zq[k][lc], acc0);
                acc1 = mfma4(av, Zq[k][16 + lc], acc1);
            }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int xr = 16 * rb + 4 * g + v;
            if (xr < c) {
                Xg[(c0 + xr) * RW + lc] = acc0[v];
                Xg[(c0 + xr) * RW + 16 + lc] = acc1[v];
            }
        }
    }
}

// ---- tile products of the other fronts ---------------------------------------
// W of the level's fronts with children: W1 rows to T (row t0 + c0 + i), W2
// rows to U (the plan leaves out what is read in place: leaves, and the W1 of
// fronts whose product task sums it from the pull lists)
__global__ void __launch_bounds__(256) k_rs_assemble(const RsRow *__restrict__ tasks, RsArgs a, float *__restrict__ buf)
{
    const RsRow t = tasks[blockIdx.x];
    const int s = t.s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int i = t.I * 32 + (threadIdx.x >> 3), q4 = threadIdx.x & 7;
    if (i >= F) return;
    const int64_t rid = a.frow[s] + i;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const int64_t q1 = a.lptr[rid + 1];
    // four list entries at a time: their indices, then their rows, in flight
    // together; the sums stay in list order
    for (int64_t q = a.lptr[rid]; q < q1; q += 4) {
        int32_t idx[4];
#pragma unroll
        for (int u = 0; u < 4; u++) idx[u] = q + u < q1 ? a.lst[q + u] : 0;
        f32x4 x[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            x[u] = q + u < q1 ? *(const f32x4 *)(buf + (int64_t)idx[u] * RW + 4 * q4) : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < 4; u++)
            if (q + u < q1) {
                acc.x += x[u].x;
                acc.y += x[u].y;
                acc.z += x[u].z;
                acc.w += x[u].w;
            }
    }
    const int64_t dst = i < c ? a.t0 + a.col0[s] + i : a.n + a.uoff[s] + (i - c);
    *(f32x4 *)(buf + dst * RW + 4 * q4) = acc;
}

// the four tile products (output rows of one 64-row tile, all right-hand sides):
//   MODE 0  Y      = X11 W1             out: Y rows of the front's columns
//   MODE 1  U      = W2 - L21 Y         out: U rows (read-modify-write)
//   MODE 2  Z      = Y - L21^T X_R      out: T rows (X_R from the U rows)
//   MODE 3  X_c    = X11^T Z            out: X rows

// operands of one 64-deep K tile for one wave (16 output rows): A in a[] (modes
// 0, 1: directly in operand layout; modes 2, 3: column j = 16 wv + jj of the
// wave's rows at tile row `lane`, to be transposed through LDS), the vector
// rows in b0[] / b1[]; index 4 kq + t
// the wave's A operands of K tile K: modes 0, 1 lane (g, lc) holds
// A(row 16 wv + lc, k = 16 kq + 4 g + t) at x = 4 kq + t; modes 2, 3 the
// wave's 16 output columns of L along k (lane = k), transposed through LDS later
// NB = 16-row blocks of output rows per wave (tile rows / 64)
template <class T, int MODE, int NB>
__device__ __forceinline__ void rs_load_a(const T *__restrict__ L, int F, int c, int I, int K, int kn, int orow0,
                                          int on, int wv, int g, int lc, int lane, float (&av)[16 * NB])
{
    if (MODE == 0 || MODE == 1) {
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int kq = 0; kq < 4; kq++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int k = 16 * kq + 4 * g + t, kk = K * RTILE + k, x = 16 * nb + 4 * kq + t;
                    const int row = 16 * NB * wv + 16 * nb + lc;
                    float v = 0.f;
                    if (row < on && k < kn) {
                        if (MODE == 0) {
                            if (K < I || k <= row) v = (float)L[(int64_t)kk * F + orow0 + row];
                        } else {
                            v = (float)L[(int64_t)kk * F + c + orow0 + row];
                        }
                    }
                    av[x] = v;
                }
    } else {
#pragma unroll
        for (int jj = 0; jj < 16 * NB; jj++) {
            const int j = 16 * NB * wv + jj;   // tile-local output row = front column orow0 + j
            float v = 0.f;
            if (j < on && lane < kn) {
                const int64_t col = (int64_t)(orow0 + j) * F;
                if (MODE == 2) v = (float)L[col + c + K * RTILE + lane];
                else if (K > I || lane >= j) v = (float)L[col + K * RTILE + lane];
            }
            av[jj] = v;
        }
    }
}

// the K tile's vector rows (64 x 32 floats), loaded by the whole workgroup:
// thread e takes row e / 4, floats 8 (e % 4) .. +7.  Mode 0: W1 from T, or
// summed from the pull lists when the front's columns are one tile (pull; its
// W1 rows are read by this task alone); mode 1: Y; mode 2: X at the update
// rows, read by index; mode 3: Z from T.
template <int MODE>
__device__ __forceinline__ void rs_load_b(const RsArgs &a, int64_t c0, int64_t fr, const int32_t *__restrict__ R,
                                          bool pull, bool leaf, int K, int kn, const float *__restrict__ buf,
                                          const float *__restrict__ Yg, const float *__restrict__ Xg, float4 (&bv)[2])
{
    const int row = threadIdx.x >> 2, col = 8 * (threadIdx.x & 3);
    const int64_t kk = (int64_t)K * RTILE + row;
    bv[0] = make_float4(0.f, 0.f, 0.f, 0.f);
    bv[1] = bv[0];
    if (row < kn) {
        if (MODE == 0 && pull) {
            pull_sum8(a, buf, fr + kk, col, bv);
            return;
        }
        const float *src = (MODE == 0 && leaf) ? buf + (c0 + kk) * RW
                         : (MODE == 0 || MODE == 3) ? buf + (a.t0 + c0 + kk) * RW
                         : MODE == 1                ? Yg + (c0 + kk) * RW
                                                    : Xg + (int64_t)R[kk] * RW;
        bv[0] = *reinterpret_cast<const float4 *>(src + col);
        bv[1] = *reinterpret_cast<const float4 *>(src + col + 4);
    }
}

// The finish of a split output tile: the partial sums of its tasks added in
// a fixed order, then the tile written as a direct task would.
template <int MODE, int RT>
__device__ __forceinline__ void fin_tile(const RsFin &t, const RsArgs &a, float *__restrict__ buf,
                                         float *__restrict__ Yg, float *__restrict__ Xg,
                                         const float *__restrict__ part)
{
    constexpr int PER = RT * RW / 256;                // elements per thread
    const int s = t.s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int orow0 = t.I * RT;
    const int on = min(RT, (MODE == 1 ? F - c : c) - orow0);
    const int64_t c0 = a.col0[s];
    const bool leaf = a.chptr[s + 1] == a.chptr[s];   // W2 = 0
    // partial sums in a fixed order, each step's PER loads independent
    float sum[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) sum[u] = 0.f;
    const float *p = part + t.p0 * (RT * RW) + threadIdx.x;
    for (int q = 0; q < t.np; q++)
#pragma unroll
        for (int u = 0; u < PER; u++) {
            sum[u] += p[(int64_t)q * (RT * RW) + 256 * u];
        }
    const float *src = MODE == 1 ? buf + (a.n + a.uoff[s] + orow0) * RW : Yg + (c0 + orow0) * RW;
    float *dst = MODE == 0 ? Yg + (c0 + orow0) * RW
               : MODE == 1 ? buf + (a.n + a.uoff[s] + orow0) * RW
               : MODE == 2 ? buf + (a.t0 + c0 + orow0) * RW : Xg + (c0 + orow0) * RW;
    float o[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u;
        o[u] = (((MODE == 1 && !leaf) || MODE == 2) && (e >> 5) < on) ? src[e] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        const int e = threadIdx.x + 256 * u;
        if ((e >> 5) < on) dst[e] = (MODE == 1 || MODE == 2) ? o[u] - sum[u] : sum[u];
    }
}

template <class T, int MODE, int RT>
__global__ void __launch_bounds__(256) k_rs_part(const T *__restrict__ pool, const RsPart *__restrict__ tasks, RsArgs a,
                                                 float *__restrict__ buf, float *__restrict__ Yg,
                                                 float *__restrict__ Xg, float *__restrict__ part)
{
    constexpr int NB = RT / 64;                       // 16-row blocks of output rows per wave
    constexpr bool TR = MODE == 2 || MODE == 3;       // A read along columns, transposed through LDS
    __shared__ float At[TR ? 4 : 1][TR ? RTILE : 1][TR ? 16 * NB + 4 : 1];   // [k][j] of the wave's A tile
    __shared__ __align__(16) float Bs[2][RTILE][RBP];   // the K tile's vector rows, double-buffered
    const RsPart tk = tasks[blockIdx.x];
    const int s = tk.s, I = tk.I;
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const int orow0 = I * RT;
    const int on = min(RT, (MODE == 1 ? r : c) - orow0);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane >> 4, lc = lane & 15;
    // a wave whose rows are past the tile still loads its share of the vector
    // tiles and takes part in the barriers
    const bool active = 16 * NB * wv < on;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s], fr = a.frow[s];
    const int32_t *R = a.rows + a.rptr[s];   // mode 2: the update rows' indices
    const bool leaf = a.chptr[s + 1] == a.chptr[s];   // W1 = P rows, W2 = 0
    const bool pull = !leaf && c <= RTILE;            // mode 0: W1 from the pull lists
    const int kdim = MODE == 2 ? r : c;
    f32x4 acc0[NB], acc1[NB];
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        acc0[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
        acc1[nb] = acc0[nb];
    }
    float av[16 * NB];
    float4 bv[2];
    auto kn_of = [&](int K) { return min(RTILE, kdim - K * RTILE); };
    if (active) rs_load_a<T, MODE, NB>(L, F, c, I, tk.k0, kn_of(tk.k0), orow0, on, wv, g, lc, lane, av);
    rs_load_b<MODE>(a, c0, fr, R, pull, leaf, tk.k0, kn_of(tk.k0), buf, Yg, Xg, bv);
    {
        float *dst = &Bs[0][threadIdx.x >> 2][8 * (threadIdx.x & 3)];
        *reinterpret_cast<float4 *>(dst) = bv[0];
        *reinterpret_cast<float4 *>(dst + 4) = bv[1];
    }
    __syncthreads();
    for (int K = tk.k0; K < tk.k1; K++) {
        const int bb = (K - tk.k0) & 1;
        float ca[16 * NB];
#pragma unroll
        for (int x = 0; x < 16 * NB; x++) ca[x] = av[x];
        const bool more = K + 1 < tk.k1;
        if (more) {     // next tile's loads in flight during this tile's products
            if (active) rs_load_a<T, MODE, NB>(L, F, c, I, K + 1, kn_of(K + 1), orow0, on, wv, g, lc, lane, av);
            rs_load_b<MODE>(a, c0, fr, R, pull, leaf, K + 1, kn_of(K + 1), buf, Yg, Xg, bv);
        }
        if (active) {
            if constexpr (TR) {
#pragma unroll
                for (int jj = 0; jj < 16 * NB; jj++) At[wv][lane][jj] = ca[jj];
                wave_lds_sync();
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
#pragma unroll
                    for (int kq = 0; kq < 4; kq++)
#pragma unroll
                        for (int t = 0; t < 4; t++) ca[16 * nb + 4 * kq + t] = At[wv][16 * kq + 4 * g + t][16 * nb + lc];
                wave_lds_sync();
            }
#pragma unroll
            for (int kq = 0; kq < 4; kq++)
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int k = 16 * kq + 4 * g + t;
                    const float b0 = Bs[bb][k][lc], b1 = Bs[bb][k][16 + lc];
#pragma unroll
                    for (int nb = 0; nb < NB; nb++) {
                        const int x = 16 * nb + 4 * kq + t;
                        acc0[nb] = mfma4(ca[x], b0, acc0[nb]);
                        acc1[nb] = mfma4(ca[x], b1, acc1[nb]);
                    }
                }
        }
        if (more) {
            // the other buffer was last read in the previous tile, before the
            // barrier that ended it
            float *dst = &Bs[bb ^ 1][threadIdx.x >> 2][8 * (threadIdx.x & 3)];
            *reinterpret_cast<float4 *>(dst) = bv[0];
            *reinterpret_cast<float4 *>(dst + 4) = bv[1];
        }
        __syncthreads();
    }
    if (tk.direct) {
        if (!active) return;
        // modes 1, 2 read before they write: all reads first (a store between
        // them would serialize the loads)
        float o0[NB][4], o1[NB][4];
        const float *src = MODE == 1 ? buf + (a.n + a.uoff[s] + orow0) * RW : Yg + (c0 + orow0) * RW;
        float *dst = MODE == 0 ? Yg + (c0 + orow0) * RW
                   : MODE == 1 ? buf + (a.n + a.uoff[s] + orow0) * RW
                   : MODE == 2 ? buf + (a.t0 + c0 + orow0) * RW : Xg + (c0 + orow0) * RW;
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * NB * wv + 16 * nb + 4 * g + v;
                o0[nb][v] = 0.f;
                o1[nb][v] = 0.f;
                if (((MODE == 1 && !leaf) || MODE == 2) && row < on) {
                    o0[nb][v] = src[row * RW + lc];
                    o1[nb][v] = src[row * RW + 16 + lc];
                }
            }
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * NB * wv + 16 * nb + 4 * g + v;
                if (row < on) {
                    dst[row * RW + lc] = (MODE == 1 || MODE == 2) ? o0[nb][v] - acc0[nb][v] : acc0[nb][v];
                    dst[row * RW + 16 + lc] = (MODE == 1 || MODE == 2) ? o1[nb][v] - acc1[nb][v] : acc1[nb][v];
                }
            }
    } else {
        if (!active) return;
        float *out = part + (int64_t)blockIdx.x * (RT * RW);
#pragma unroll
        for (int nb = 0; nb < NB; nb++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 16 * NB * wv + 16 * nb + 4 * g + v;
                out[row * RW + lc] = acc0[nb][v];
                out[row * RW + 16 + lc] = acc1[nb][v];
            }
    }
}

template <int MODE, int RT>
__global__ void __launch_bounds__(256) k_rs_fin(const RsFin *__restrict__ tasks, RsArgs a, float *__restrict__ buf,
                                                float *__restrict__ Yg, float *__restrict__ Xg,
                                                const float *__restrict__ part)
{
    fin_tile<MODE, RT>(tasks[blockIdx.x], a, buf, Yg, Xg, part);
}

// ---- permutation + precision change at the boundary -------------------------
// buffer row i = B(perm[i], :) (FP64 -> FP32, columns >= nb zero); element
// (i, v) of B at B[i * rs + v * cs]
// A block covers 32 rows, each 32-thread group 4 of them (rows slot, slot + 8,
// ...): the 4 rows' indices, then their 4 gathers, are in flight together
// (129 -> 107 us at 10^6 rows, 28 columns).
constexpr int RS_PERM_ROWS = 32;

__global__ void __launch_bounds__(256) k_rs_in(const double *__restrict__ B, int64_t rs, int64_t cs,
                                               const int64_t *__restrict__ perm, int64_t n, int nb,
                                               float *__restrict__ buf)
{
    const int64_t i0 = (int64_t)blockIdx.x * RS_PERM_ROWS + (threadIdx.x >> 5);
    const int v = threadIdx.x & 31;
    int64_t pr[4];
#pragma unroll
    for (int u = 0; u < 4; u++) pr[u] = i0 + 8 * u < n ? perm[i0 + 8 * u] : 0;
    float val[4];
#pragma unroll
    for (int u = 0; u < 4; u++) val[u] = (i0 + 8 * u < n && v < nb) ? (float)B[pr[u] * rs + (int64_t)v * cs] : 0.f;
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (i0 + 8 * u < n) buf[(i0 + 8 * u) * RW + v] = val[u];
}

// one row per 32-thread group (grouping rows as in k_rs_in made this
// scatter slower: 92 -> 97 us at 10^6 rows)
__global__ void __launch_bounds__(256) k_rs_out(const float *__restrict__ Xg, const int64_t *__restrict__ perm,
                                                int64_t n, int nb, double *__restrict__ X, int64_t rs, int64_t cs)
{
    const int64_t i = (int64_t)blockIdx.x * 8 + (threadIdx.x >> 5);
    const int v = threadIdx.x & 31;
    if (i >= n || v >= nb) return;
    X[perm[i] * rs + (int64_t)v * cs] = (double)Xg[i * RW + v];
}

// ---- one right-hand side: the Lanczos paths' solves ----------------------
// The same plan and the same inverted diagonal blocks as the block solve, one
// FP64 value per buffer row: Y = X11 W1, U = W2 - L21 Y forward, Z = Y -
// L21^T X_R, X_c = X11^T Z backward -- matrix-vector products with no chain
// inside a front, only the tree levels.  Every sum runs in a fixed order.

constexpr int VS_ASM = 8;        // assembly chunks (32 rows) per workgroup

__global__ void __launch_bounds__(256) k_vs_in(const double *__restrict__ B, const int64_t *__restrict__ perm, int64_t n,
                                               double *__restrict__ buf)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) buf[i] = B[perm[i]];
}

__global__ void __launch_bounds__(256) k_vs_out(const double *__restrict__ X, const int64_t *__restrict__ perm, int64_t n,
                                                double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[perm[i]] = X[i];
}

template <class T>
__global__ void __launch_bounds__(256) k_vs_sign(double *__restrict__ y, const T *__restrict__ sgn, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] *= (double)sgn[i];
}

__device__ __forceinline__ double vs_pull(const RsArgs &a, const double *__restrict__ buf, int64_t rid)
{
    double x = 0.0;
    for (int64_t q = a.lptr[rid]; q < a.lptr[rid + 1]; q++) x += buf[a.lst[q]];
    return x;
}

// the assembled rows of a level (the plan's 32-row chunks, eight per workgroup)
__global__ void __launch_bounds__(256) k_vs_assemble(const RsRow *__restrict__ tasks, int64_t cnt, RsArgs a,
                                                     double *__restrict__ buf)
{
    const int64_t tix = (int64_t)blockIdx.x * VS_ASM + (threadIdx.x >> 5);
    if (tix >= cnt) return;
    const RsRow t = tasks[tix];
    const int s = t.s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int i = t.I * 32 + (threadIdx.x & 31);
    if (i >= F) return;
    const double x = vs_pull(a, buf, a.frow[s] + i);
    buf[i < c ? a.t0 + a.col0[s] + i : a.n + a.uoff[s] + (i - c)] = x;
}

// tiny fronts (c, r <= 64), one wave per front, lane = row
template <class T>
__global__ void __launch_bounds__(256) k_vs_tiny_fwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                                     int nf, RsArgs a, double *__restrict__ buf, double *__restrict__ Y)
{
    __shared__ double sw[4][64], sy[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int fi = blockIdx.x * 4 + wv;
    if (fi >= nf) return;
    const int s = fronts[fi];
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s], ub = a.n + a.uoff[s];
    const bool leaf = a.chptr[s + 1] == a.chptr[s];
    const int64_t tb = leaf ? c0 : a.t0 + c0;
    sw[wv][lane] = lane < c ? buf[tb + lane] : 0.0;
    const double w2 = (!leaf && lane < r) ? buf[ub + lane] : 0.0;
    wave_lds_sync();
    double y = 0.0;
    for (int k = 0; k < c; k++)
        if (k <= lane && lane < c) y += (double)L[(int64_t)k * F + lane] * sw[wv][k];
    if (lane < c) Y[c0 + lane] = y;
    sy[wv][lane] = lane < c ? y : 0.0;
    wave_lds_sync();
    if (lane < r) {
        double u = 0.0;
        for (int k = 0; k < c; k++) u += (double)L[(int64_t)k * F + c + lane] * sy[wv][k];
        buf[ub + lane] = w2 - u;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_vs_tiny_bwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                                     int nf, RsArgs a, const double *__restrict__ Y,
                                                     double *__restrict__ X)
{
    __shared__ double sx[4][64], sz[4][64];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int fi = blockIdx.x * 4 + wv;
    if (fi >= nf) return;
    const int s = fronts[fi];
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s];
    const int32_t *R = a.rows + a.rptr[s];
    sx[wv][lane] = lane < r ? X[R[lane]] : 0.0;
    wave_lds_sync();
    // lane j: its column of L21 and of X11, read down the column
    double z = 0.0;
    if (lane < c) {
        const T *col = L + (int64_t)lane * F;
        for (int k = 0; k < r; k++) z += (double)col[c + k] * sx[wv][k];
        z = Y[c0 + lane] - z;
    }
    sz[wv][lane] = lane < c ? z : 0.0;
    wave_lds_sync();
    if (lane < c) {
        const T *col = L + (int64_t)lane * F;
        double x = 0.0;
        for (int k = lane; k < c; k++) x += (double)col[k] * sz[wv][k];
        X[c0 + lane] = x;
    }
}

// The product tasks of the other fronts (the plan's 64-row output tiles and
// K ranges of at most RKCHUNK 64-row tiles).  The task's vector segment (<=
// 256 values: W1, Y, X at the update rows, or Z) is staged in LDS first.
// Modes 0, 1: lane = output row, wave w takes K tile k0 + w, the four waves'
// sums added in wave order.  Modes 2, 3: the output rows are columns of L;
// wave w takes columns 16 w .. 16 w + 15, lane = row within each K tile, and
// each column's sum is reduced over the lanes at the end.
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_vs_part(const T *__restrict__ pool, const RsPart *__restrict__ tasks, RsArgs a,
                                                 double *__restrict__ buf, double *__restrict__ Y,
                                                 double *__restrict__ X, double *__restrict__ part)
{
    __shared__ double sv[RKCHUNK * RTILE];
    __shared__ double sacc[4][RTILE];
    const RsPart tk = tasks[blockIdx.x];
    const int s = tk.s, I = tk.I;
    const int32_t F = a.Fs[s], c = a.ncol[s], r = F - c;
    const int orow0 = I * RTILE;
    const int on = min(RTILE, (MODE == 1 ? r : c) - orow0);
    const int kdim = MODE == 2 ? r : c;
    const int kbeg = tk.k0 * RTILE, kend = min(kdim, tk.k1 * RTILE);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s], fr = a.frow[s];
    const bool leaf = a.chptr[s + 1] == a.chptr[s];
    const bool pull = !leaf && c <= RTILE;
    {
        const int kk = kbeg + (int)threadIdx.x;
        double v = 0.0;
        if (kk < kend) {
            if (MODE == 0) v = pull ? vs_pull(a, buf, fr + kk) : buf[(leaf ? c0 : a.t0 + c0) + kk];
            else if (MODE == 1) v = Y[c0 + kk];
            else if (MODE == 2) v = X[a.rows[a.rptr[s] + kk]];
            else v = buf[a.t0 + c0 + kk];
        }
        sv[threadIdx.x] = v;
    }
    __syncthreads();
    double res = 0.0;              // modes 0, 1: this thread's row (lane), wave 0 after the sum
    if (MODE == 0 || MODE == 1) {
        const int K = tk.k0 + wv;
        double acc = 0.0;
        if (K < tk.k1 && lane < on) {
            const int row = orow0 + lane;
            const int kn = min(RTILE, kdim - K * RTILE);
            const T *base = L + (MODE == 1 ? c + row : row);
            for (int k = 0; k < kn; k++) {
                const int kk = K * RTILE + k;
                if (MODE == 0 && K == I && k > lane) break;     // X11 lower triangular
                acc += (double)base[(int64_t)kk * F] * sv[kk - kbeg];
            }
        }
        sacc[wv][lane] = acc;
        __syncthreads();
        if (wv != 0) return;
        res = sacc[0][lane] + sacc[1][lane] + sacc[2][lane] + sacc[3][lane];
    } else {
        double acc[16];
#pragma unroll
        for (int jj = 0; jj < 16; jj++) acc[jj] = 0.0;
        for (int K = tk.k0; K < tk.k1; K++) {
            const int kk = K * RTILE + lane;
            if (kk >= kend) continue;
            const double xv = sv[kk - kbeg];
#pragma unroll
            for (int jj = 0; jj < 16; jj++) {
                const int j = 16 * wv + jj;
                if (j >= on) continue;
                const T *col = L + (int64_t)(orow0 + j) * F;
                if (MODE == 2) acc[jj] += (double)col[c + kk] * xv;
                else if (kk >= orow0 + j) acc[jj] += (double)col[kk] * xv;   // X11^T: k >= j
            }
        }
#pragma unroll
        for (int jj = 0; jj < 16; jj++) {
            double v = acc[jj];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            acc[jj] = v;
        }
        // lane jj of each wave now holds column 16 wv + jj's sum
        double mine = 0.0;
#pragma unroll
        for (int jj = 0; jj < 16; jj++)
            if (lane == jj) mine = acc[jj];
        if (lane < 16) sacc[0][16 * wv + lane] = mine;
        __syncthreads();
        if (wv != 0) return;
        res = sacc[0][lane];
    }
    // wave 0, lane = output row
    if (lane >= on) return;
    if (tk.direct) {
        if (MODE == 0) Y[c0 + orow0 + lane] = res;
        else if (MODE == 1) {
            const int64_t u = a.n + a.uoff[s] + orow0 + lane;
            buf[u] = (leaf ? 0.0 : buf[u]) - res;
        } else if (MODE == 2) buf[a.t0 + c0 + orow0 + lane] = Y[c0 + orow0 + lane] - res;
        else X[c0 + orow0 + lane] = res;
    } else {
        part[(int64_t)blockIdx.x * RTILE + lane] = res;
    }
}

template <int MODE>
__global__ void __launch_bounds__(64) k_vs_fin(const RsFin *__restrict__ tasks, RsArgs a, double *__restrict__ buf,
                                               double *__restrict__ Y, double *__restrict__ X,
                                               const double *__restrict__ part)
{
    const RsFin t = tasks[blockIdx.x];
    const int s = t.s, lane = threadIdx.x;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int orow0 = t.I * RTILE;
    const int on = min(RTILE, (MODE == 1 ? F - c : c) - orow0);
    if (lane >= on) return;
    double res = 0.0;
    for (int q = 0; q < t.np; q++) res += part[(t.p0 + q) * RTILE + lane];
    const int64_t c0 = a.col0[s];
    const bool leaf = a.chptr[s + 1] == a.chptr[s];
    if (MODE == 0) Y[c0 + orow0 + lane] = res;
    else if (MODE == 1) {
        const int64_t u = a.n + a.uoff[s] + orow0 + lane;
        buf[u] = (leaf ? 0.0 : buf[u]) - res;
    } else if (MODE == 2) buf[a.t0 + c0 + orow0 + lane] = Y[c0 + orow0 + lane] - res;
    else X[c0 + orow0 + lane] = res;
}

// ---- plan -----------------------------------------------------------------
struct RsLevel {
    int64_t sm_off = 0, sm_cnt = 0;          // tiny fronts, by class (tiny_class)
    int64_t tc_off[9] = {}, tc_cnt[9] = {};
    int64_t as_off = 0, as_cnt = 0;          // assembly chunks (every front of the level)
    // product task lists by variant: modes 0-3 with 64-row tiles, then modes
    // 1 and 2 with 128-row tiles (variants 4, 5); wide[m - 1]: mode m runs wide
    int64_t pt_off[6] = {}, pt_cnt[6] = {};
    int64_t ft_off[6] = {}, ft_cnt[6] = {};
    bool wide[2] = {false, false};
    // block solve at the lowest levels (L < kNarrowLevels): the narrow fronts
    // by class of c (16, 32, 64 columns) -- forward tasks, backward fronts
    // with one wave or four -- and the lists of the level's other fronts (tiny
    // classes, assembly chunks, product / finish tasks) without them
    bool narrow = false;
    int64_t nf_off[3] = {}, nf_cnt[3] = {};
    int64_t n1_off[3] = {}, n1_cnt[3] = {}, n4_off[3] = {}, n4_cnt[3] = {};
    int64_t rtc_off[9] = {}, rtc_cnt[9] = {};
    int64_t ras_off = 0, ras_cnt = 0;
    int64_t rpt_off[6] = {}, rpt_cnt[6] = {};
    int64_t rft_off[6] = {}, rft_cnt[6] = {};
};

// host-side plan of the solve (task lists, pull lists), uploaded at first use
struct RsHost {
    std::vector<int32_t> small, lst;
    std::vector<RsNTask> ntask;
    std::vector<RsNFront> nfront;
    std::vector<RsRow> asmt;
    std::vector<RsPart> pt;
    std::vector<RsFin> ft;
    std::vector<int64_t> frow, uoff, lptr;
    int64_t maxpart = 1, sumr = 0;
};

struct RsState {
    // the plan is built and uploaded (on a stream of its own) by a helper
    // thread started while the device factorizes; the first solve joins it and
    // waits for that stream
    std::thread th;
    int plan_rc = 0;
    std::string plan_msg;     // the helper thread's error record
    bool ready = false;
    hipStream_t up_st = nullptr;
    RsHost host;
    DevMem mem;
    std::vector<RsLevel> lev;
    // the two halves below the root (split): lev holds the first half's lists
    // at levels < ljoin (all lists from ljoin on), levR the second half's
    bool split = false, one_stream = false;
    int32_t ljoin = 0;
    std::vector<RsLevel> levR;
    hipEvent_t sev[4] = {};        // fork / join of the halves' streams (up_st runs the second)
    int32_t *small = nullptr;
    RsNTask *ntask = nullptr;        // narrow fronts' forward tasks and backward records
    RsNFront *nfront = nullptr;
    RsRow *asmt = nullptr;
    RsPart *pt = nullptr;
    RsFin *ft = nullptr;
    float *part = nullptr, *part2 = nullptr, *buf = nullptr, *Yg = nullptr, *Xg = nullptr;
    int64_t *frow = nullptr, *uoff = nullptr, *lptr = nullptr;
    int32_t *lst = nullptr;
    int64_t n = 0, t0 = 0;
    // the single-right-hand-side solve's buffers (one FP64 value per row)
    bool rows = false, vec = false;            // which buffers the plan's upload allocates
    double *vbuf = nullptr, *vY = nullptr, *vX = nullptr, *vpart = nullptr, *vpart2 = nullptr;
    // the level launches of one solve captured as a graph (opt.use_graph):
    // they read and write only the plan's own buffers and the factor's pool,
    // so one capture serves every later solve of this factor
    hipGraphExec_t gexec = nullptr;
    const void *gpool = nullptr;
    // tile-height autotuner: 0 not started, 1 / 2 timing the narrow / wide
    // run, 3 decided; tune_run = the run being timed (-1 none)
    int tune_state = 0, tune_run = -1;
    std::vector<hipEvent_t> tune_ev[2];
    ~RsState()
    {
        if (th.joinable()) th.join();
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto &v : tune_ev)
            for (hipEvent_t x : v)
                if (x) (void)hipEventDestroy(x);
        for (hipEvent_t x : sev)
            if (x) (void)hipEventDestroy(x);
        if (up_st) (void)hipStreamDestroy(up_st);
    }
};

bool is_tiny(int64_t c, int64_t r) { return c <= TINY && r <= TINY; }

// launch class of a tiny front: 3 (16-row blocks of c: 1, 2 or 4) + (of r)
int tiny_class(int64_t c, int64_t r)
{
    auto b = [](int64_t x) { return x <= 16 ? 0 : x <= 32 ? 1 : 2; };
    return 3 * b(c) + b(r);
}

// A launch of a class with a handful of fronts costs one front's latency chain
// (8-25 us at 100^3) however little it does, while the kernel of a larger
// class (NQ, NR at least the front's) solves the same front exactly: its
// extra blocks hold zeros.  So a class of fewer than kTinyMin fronts joins the
// cheapest larger class of its level that has kTinyMin or more, and the small
// classes left over join together into the smallest class covering them all.
constexpr size_t kTinyMin = 1024;

void promote_tiny_classes(std::vector<int32_t> (&cls)[9])
{
    auto covers = [](int a, int b) { return a / 3 >= b / 3 && a % 3 >= b % 3; };
    auto cost = [](int q) { return (1 << (q / 3)) * (1 << (q % 3)); };
    auto move = [&](int from, int to) {
        cls[to].insert(cls[to].end(), cls[from].begin(), cls[from].end());
        cls[from].clear();
    };
    for (int q = 0; q < 9; q++) {
        if (cls[q].empty() || cls[q].size() >= kTinyMin) continue;
        int best = -1;
        for (int p = 0; p < 9; p++)
            if (p != q && cls[p].size() >= kTinyMin && covers(p, q) && (best < 0 || cost(p) < cost(best))) best = p;
        if (best >= 0) move(q, best);
    }
    int jc = 0, jr = 0, nsmall = 0;
    for (int q = 0; q < 9; q++)
        if (!cls[q].empty() && cls[q].size() < kTinyMin) {
            jc = std::max(jc, q / 3);
            jr = std::max(jr, q % 3);
            nsmall++;
        }
    if (nsmall > 1)
        for (int q = 0; q < 9; q++)
            if (q != 3 * jc + jr && !cls[q].empty() && cls[q].size() < kTinyMin) move(q, 3 * jc + jr);
}

// Narrow fronts: the block solve's wave-per-front path at the lowest levels,
// where fronts are many and small and a level's separate assembly, product
// and finish launches each cost more than their work (measured at 100^3:
// levels 0-2 went from 1.05 to 0.81 ms per solve; from level 3 on the tile
// products are as fast or faster)
constexpr int kNarrowLevels = 3;
constexpr int64_t kNarrowRmax = 512;

bool is_narrow(int64_t c, int64_t r) { return c <= RTILE && r <= kNarrowRmax; }

int narrow_class(int64_t c) { return c <= 16 ? 0 : c <= 32 ? 1 : 2; }   // 16, 32, 64 columns

// a class of fewer than kTinyMin entries joins the smallest larger class
// present (its kernel solves the same front exactly)
template <class E>
void promote_narrow(std::vector<E> (&cls)[3])
{
    for (int q = 0; q < 2; q++) {
        if (cls[q].empty() || cls[q].size() >= kTinyMin) continue;
        for (int p = q + 1; p < 3; p++)
            if (!cls[p].empty()) {
                cls[p].insert(cls[p].end(), cls[q].begin(), cls[q].end());
                cls[q].clear();
                break;
            }
    }
}

// product tasks of one launch (first task pbase): a tile whose K range is one
// chunk is a direct task; a longer one is split into chunks that leave
// partial sums for a finish record; a mode-2 tile of a front without update
// rows is a direct task with an empty K range (Z = Y)
void plan_parts(const Symbolic &S, const std::vector<int64_t> &fronts, int mode, int64_t RT, std::vector<RsPart> &pt,
                std::vector<RsFin> &ft, int64_t pbase)
{
    for (int64_t s : fronts) {
        const int64_t c = S.c1[s] - S.c0[s], r = S.F[s] - c;
        const int64_t nt11 = (c + RTILE - 1) / RTILE, nupd = (r + RTILE - 1) / RTILE;
        const int64_t nout = mode == 1 ? (r + RT - 1) / RT : (c + RT - 1) / RT;
        for (int64_t I = 0; I < nout; I++) {
            int64_t kb, ke;
            if (mode == 0) { kb = 0; ke = I + 1; }
            else if (mode == 1) { kb = 0; ke = nt11; }
            else if (mode == 2) { kb = 0; ke = nupd; }
            else { kb = I; ke = nt11; }
            const int64_t np = (ke - kb + RKCHUNK - 1) / RKCHUNK;
            if (np <= 1) {
                pt.push_back({(int32_t)s, (int32_t)I, (int32_t)kb, (int32_t)ke, 1, 0});
                continue;
            }
            RsFin f{(int32_t)s, (int32_t)I, (int64_t)pt.size() - pbase, (int32_t)np, 0};
            for (int64_t k0 = kb; k0 < ke; k0 += RKCHUNK)
                pt.push_back({(int32_t)s, (int32_t)I, (int32_t)k0, (int32_t)std::min<int64_t>(ke, k0 + RKCHUNK), 0, 0});
            ft.push_back(f);
        }
    }
}

// the plan: reads only the symbolic structure, so it can run beside the
// numeric factorization
int rs_plan(const Factor *F, RsState &B)
{
    const Symbolic &S = F->S;
    const int64_t ns = S.ns, n = F->n;
    RsHost &H = B.host;
    std::vector<int64_t> &frow = H.frow, &uoff = H.uoff, &lptr = H.lptr;
    frow.assign(ns + 1, 0);
    uoff.assign(ns + 1, 0);
    for (int64_t s = 0; s < ns; s++) {
        frow[s + 1] = frow[s] + S.F[s];
        uoff[s + 1] = uoff[s] + (S.F[s] - (S.c1[s] - S.c0[s]));
    }
    const int64_t nrid = frow[ns], sumr = uoff[ns];
    H.sumr = sumr;
    if (2 * n + sumr >= ((int64_t)1 << 31)) return SPEIGS_ERR_MEM;   // buffer rows are int32 in the pull lists
    B.n = n;
    B.t0 = n + sumr;
    // pull lists: row (s, i) <- P row c0 + i (i < c), then each child's U rows
    // that land on i, children in order (a quarter of the analysis threads:
    // this runs beside the SELL build and the factorization's launches)
    const int nt = std::max(1, host_threads() / 4);
    std::vector<int64_t> tot(ns + 1, 0);
    parallel_for(ns, nt, 256, [&](int64_t s, int) {
        int64_t t = S.c1[s] - S.c0[s];
        for (int64_t q = S.chptr[s]; q < S.chptr[s + 1]; q++) {
            const int64_t ch = S.chl[q];
            t += S.F[ch] - (S.c1[ch] - S.c0[ch]);
        }
        tot[s + 1] = t;
    });
    for (int64_t s = 0; s < ns; s++) tot[s + 1] += tot[s];
    lptr.assign(nrid + 1, 0);
    std::vector<int32_t> &lst = H.lst;
    lst.assign(std::max<int64_t>(tot[ns], 1), 0);
    lptr[nrid] = tot[ns];
    parallel_for(ns, nt, 64, [&](int64_t s, int) {
        const int64_t c = S.c1[s] - S.c0[s], Fv = S.F[s];
        std::vector<int64_t> cnt(Fv + 1, 0);
        for (int64_t i = 0; i < c; i++) cnt[i + 1]++;
        for (int64_t q = S.chptr[s]; q < S.chptr[s + 1]; q++) {
            const int64_t ch = S.chl[q], rc = S.F[ch] - (S.c1[ch] - S.c0[ch]);
            const int32_t *rl = S.rel.data() + S.relptr[ch];
            for (int64_t j = 0; j < rc; j++) cnt[rl[j] + 1]++;
        }
        for (int64_t i = 0; i < Fv; i++) cnt[i + 1] += cnt[i];
        const int64_t base = tot[s];
        for (int64_t i = 0; i < Fv; i++) lptr[frow[s] + i] = base + cnt[i];
        for (int64_t i = 0; i < c; i++) lst[base + cnt[i]++] = (int32_t)(S.c0[s] + i);
        for (int64_t q = S.chptr[s]; q < S.chptr[s + 1]; q++) {
            const int64_t ch = S.chl[q], rc = S.F[ch] - (S.c1[ch] - S.c0[ch]);
            const int32_t *rl = S.rel.data() + S.relptr[ch];
            for (int64_t j = 0; j < rc; j++) lst[base + cnt[rl[j]]++] = (int32_t)(n + uoff[ch] + j);
        }
    });
    // per level: tiny fronts, assembly chunks of every front, product tasks of the others
    std::vector<std::vector<int64_t>> bylevel(S.nlevels);
    for (int64_t s = 0; s < ns; s++) bylevel[S.level[s]].push_back(s);
    std::vector<int32_t> &small = H.small;
    std::vector<RsRow> &asmt = H.asmt;
    std::vector<RsPart> &pt = H.pt;
    std::vector<RsFin> &ft = H.ft;
    int64_t &maxpart = H.maxpart;
    // The tree below a single root splits into two halves that share no front
    // until the root: its first children's subtrees (about half of the fronts'
    // F^2) and the rest.  Every level below the root is planned once per half,
    // and a solve runs the halves' level chains on two streams (their launches
    // are latency-bound, so the chains overlap), joining at the root.  Each
    // front's work is the same either way.
    B.split = false;
    std::vector<char> left(ns, 1);
    // test knob: 0 plans no split, 1 plans it but runs both halves on one stream
    const char *ks = getenv("SPEIGS_HIP_RS_SPLIT");
    B.one_stream = ks && atoi(ks) == 1;
    if (S.nlevels >= 3 && bylevel[S.nlevels - 1].size() == 1 && !(ks && atoi(ks) == 0)) {
        const int64_t R = bylevel[S.nlevels - 1][0], nk = S.chptr[R + 1] - S.chptr[R];
        if (nk >= 2) {
            std::vector<double> wsum(ns + 1, 0.0);
            for (int64_t s = 0; s < ns; s++) wsum[s + 1] = wsum[s] + (double)S.F[s] * (double)S.F[s];
            const double half = 0.5 * (wsum[ns] - (double)S.F[R] * (double)S.F[R]);
            int64_t kj = S.chl[S.chptr[R]];
            for (int64_t q = S.chptr[R]; q + 1 < S.chptr[R + 1]; q++) {   // the last child stays right
                kj = S.chl[q];
                if (wsum[kj + 1] >= half) break;
            }
            for (int64_t s = 0; s < ns; s++) left[s] = s <= kj;
            B.split = true;
            B.ljoin = S.nlevels - 1;
        }
    }
    auto build_level = [&](int32_t L, const std::vector<int64_t> &fronts, RsLevel &P) {
        std::vector<int64_t> big;
        P.sm_off = (int64_t)small.size();
        std::vector<int32_t> cls[9];
        for (int64_t s : fronts) {
            const int64_t c = S.c1[s] - S.c0[s];
            if (is_tiny(c, S.F[s] - c)) cls[tiny_class(c, S.F[s] - c)].push_back((int32_t)s);
            else big.push_back(s);
        }
        promote_tiny_classes(cls);
        for (int q = 0; q < 9; q++) {
            P.tc_off[q] = (int64_t)small.size();
            P.tc_cnt[q] = (int64_t)cls[q].size();
            small.insert(small.end(), cls[q].begin(), cls[q].end());
        }
        P.sm_cnt = (int64_t)small.size() - P.sm_off;
        // assembled: every row of a front with children, except the W1 rows of
        // the non-tiny fronts whose columns fit one tile (their product task
        // sums them from the pull lists).  A leaf is read in place: its W1 is
        // its P rows, its W2 zero.
        P.as_off = (int64_t)asmt.size();
        for (int64_t s : fronts) {
            const int64_t c = S.c1[s] - S.c0[s], r = S.F[s] - c;
            const bool tiny = is_tiny(c, r), leaf = S.chptr[s + 1] == S.chptr[s];
            if (leaf) continue;
            const int64_t i0 = (!tiny && c <= RTILE) ? c : 0;      // first assembled row
            for (int64_t I = i0 / 32; I * 32 < S.F[s]; I++) asmt.push_back({(int32_t)s, (int32_t)I});
        }
        P.as_cnt = (int64_t)asmt.size() - P.as_off;
        for (int v = 0; v < 6; v++) {
            const int mode = v < 4 ? v : v - 3;
            const int64_t p0 = (int64_t)pt.size(), f0 = (int64_t)ft.size();
            plan_parts(S, big, mode, v < 4 ? RTILE : RT_WIDE, pt, ft, p0);
            P.pt_off[v] = p0;
            P.pt_cnt[v] = (int64_t)pt.size() - p0;
            P.ft_off[v] = f0;
            P.ft_cnt[v] = (int64_t)ft.size() - f0;
            maxpart = std::max<int64_t>(maxpart, P.pt_cnt[v]);
        }
        // the block solve's lists at the lowest levels: narrow fronts by class,
        // the level's other fronts as above
        if (!B.rows || L >= kNarrowLevels) return;
        P.narrow = true;
        big.clear();
        std::vector<RsNTask> fw[3];
        std::vector<RsNFront> b1[3], b4[3];
        for (int q = 0; q < 9; q++) cls[q].clear();
        for (int64_t s : fronts) {
            const int64_t c = S.c1[s] - S.c0[s], r = S.F[s] - c;
            if (is_narrow(c, r)) {
                const int q = narrow_class(c);
                const int32_t leaf = S.chptr[s + 1] == S.chptr[s];
                for (int64_t I = 0; I == 0 || I * NTR < r; I++)
                    fw[q].push_back({S.foff[s], S.c0[s], frow[s], n + uoff[s] + I * NTR, (int32_t)S.F[s], (int32_t)c,
                                     (int32_t)(I * NTR), leaf});
                (r <= 2 * 64 ? b1 : b4)[q].push_back({S.foff[s], S.c0[s], S.rptr[s], (int32_t)S.F[s], (int32_t)c});
            } else if (is_tiny(c, r)) {
                cls[tiny_class(c, r)].push_back((int32_t)s);
            } else {
                big.push_back(s);
            }
        }
        promote_narrow(fw);
        promote_narrow(b1);
        promote_narrow(b4);
        promote_tiny_classes(cls);
        for (int q = 0; q < 3; q++) {
            P.nf_off[q] = (int64_t)H.ntask.size();
            P.nf_cnt[q] = (int64_t)fw[q].size();
            H.ntask.insert(H.ntask.end(), fw[q].begin(), fw[q].end());
            P.n1_off[q] = (int64_t)H.nfront.size();
            P.n1_cnt[q] = (int64_t)b1[q].size();
            H.nfront.insert(H.nfront.end(), b1[q].begin(), b1[q].end());
            P.n4_off[q] = (int64_t)H.nfront.size();
            P.n4_cnt[q] = (int64_t)b4[q].size();
            H.nfront.insert(H.nfront.end(), b4[q].begin(), b4[q].end());
        }
        for (int q = 0; q < 9; q++) {
            P.rtc_off[q] = (int64_t)small.size();
            P.rtc_cnt[q] = (int64_t)cls[q].size();
            small.insert(small.end(), cls[q].begin(), cls[q].end());
        }
        P.ras_off = (int64_t)asmt.size();
        for (int32_t q = 0; q < 9; q++)
            for (int32_t s : cls[q])
                if (S.chptr[s + 1] != S.chptr[s])
                    for (int64_t I = 0; I * 32 < S.F[s]; I++) asmt.push_back({s, (int32_t)I});
        for (int64_t s : big) {
            const int64_t c = S.c1[s] - S.c0[s];
            if (S.chptr[s + 1] == S.chptr[s]) continue;
            const int64_t i0 = c <= RTILE ? c : 0;      // first assembled row (as above)
            for (int64_t I = i0 / 32; I * 32 < S.F[s]; I++) asmt.push_back({(int32_t)s, (int32_t)I});
        }
        P.ras_cnt = (int64_t)asmt.size() - P.ras_off;
        for (int v = 0; v < 6; v++) {
            const int mode = v < 4 ? v : v - 3;
            const int64_t p0 = (int64_t)pt.size(), f0 = (int64_t)ft.size();
            plan_parts(S, big, mode, v < 4 ? RTILE : RT_WIDE, pt, ft, p0);
            P.rpt_off[v] = p0;
            P.rpt_cnt[v] = (int64_t)pt.size() - p0;
            P.rft_off[v] = f0;
            P.rft_cnt[v] = (int64_t)ft.size() - f0;
            maxpart = std::max<int64_t>(maxpart, P.rpt_cnt[v]);
        }
    };
    B.lev.assign(S.nlevels, RsLevel());
    B.levR.assign(S.nlevels, RsLevel());
    for (int32_t L = 0; L < S.nlevels; L++) {
        if (B.split && L < B.ljoin) {
            std::vector<int64_t> fl, fr;
            for (int64_t s : bylevel[L]) (left[s] ? fl : fr).push_back(s);
            build_level(L, fl, B.lev[L]);
            build_level(L, fr, B.levR[L]);
        } else {
            build_level(L, bylevel[L], B.lev[L]);
        }
    }
    if (H.ntask.empty()) H.ntask.push_back(RsNTask{});
    if (H.nfront.empty()) H.nfront.push_back(RsNFront{});
    if (small.empty()) small.push_back(0);
    if (asmt.empty()) asmt.push_back({0, 0});
    if (pt.empty()) pt.push_back({0, 0, 0, 0, 0, 0});
    if (ft.empty()) ft.push_back({0, 0, 0, 0, 0});
    return 0;
}

// device buffers and the plan's upload; the host plan is released after
int rs_upload(Factor *F, RsState &B, hipStream_t st)
{
    RsHost &H = B.host;
    const int64_t n = F->n;
    int rc;
    auto up = [&](auto **dst, const auto &v) -> int {
        int e = B.mem.alloc(dst, v.size());
        if (e) return e;
        return copy_to_device(*dst, v.data(), v.size() * sizeof(v[0]), st);
    };
    if ((rc = up(&B.small, H.small)) || (rc = up(&B.asmt, H.asmt)) ||
        (rc = up(&B.pt, H.pt)) || (rc = up(&B.ft, H.ft)) || (rc = up(&B.frow, H.frow)) || (rc = up(&B.uoff, H.uoff)) ||
        (rc = up(&B.lptr, H.lptr)) || (rc = up(&B.lst, H.lst)) || (rc = up(&B.ntask, H.ntask)) ||
        (rc = up(&B.nfront, H.nfront)))
        return rc;
    if (B.split) {
        for (hipEvent_t &e : B.sev)
            SEHIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (B.rows && (rc = B.mem.alloc(&B.part2, (size_t)H.maxpart * RT_WIDE * RW))) return rc;
        if (B.vec && (rc = B.mem.alloc(&B.vpart2, (size_t)H.maxpart * RTILE))) return rc;
    }
    if (B.rows) {
        if ((rc = B.mem.alloc(&B.part, (size_t)H.maxpart * RT_WIDE * RW))) return rc;
        if ((rc = B.mem.alloc(&B.buf, (size_t)(2 * n + H.sumr) * RW))) return rc;
        if ((rc = B.mem.alloc(&B.Yg, (size_t)n * RW))) return rc;
        if ((rc = B.mem.alloc(&B.Xg, (size_t)n * RW))) return rc;
    }
    if (B.vec) {
        if ((rc = B.mem.alloc(&B.vpart, (size_t)H.maxpart * RTILE))) return rc;
        if ((rc = B.mem.alloc(&B.vbuf, (size_t)(2 * n + H.sumr)))) return rc;
        if ((rc = B.mem.alloc(&B.vY, (size_t)n))) return rc;
        if ((rc = B.mem.alloc(&B.vX, (size_t)n))) return rc;
    }
    B.host = RsHost();
    return 0;
}

// the tiny fronts of a level, one launch per class present
template <class T, bool FWD, int NQ, int NR>
void tiny_one(const RsLevel &P, int q, RsState &B, const T *pool, const RsArgs &a, hipStream_t st)
{
    const int64_t cnt = P.narrow ? P.rtc_cnt[q] : P.tc_cnt[q];
    if (cnt <= 0) return;
    const int64_t off = P.narrow ? P.rtc_off[q] : P.tc_off[q];
    const dim3 g((unsigned)((cnt + 3) / 4)), blk(256);
    if (FWD)
        k_rs_tiny_fwd<T, NQ, NR><<<g, blk, 0, st>>>(pool, B.small + off, (int)cnt, a, B.buf, B.Yg);
    else
        k_rs_tiny_bwd<T, NQ, NR><<<g, blk, 0, st>>>(pool, B.small + off, (int)cnt, a, B.Yg, B.Xg);
}

template <class T, bool FWD>
void tiny_launch(const RsLevel &P, RsState &B, const T *pool, const RsArgs &a, hipStream_t st)
{
    tiny_one<T, FWD, 1, 1>(P, 0, B, pool, a, st);
    tiny_one<T, FWD, 1, 2>(P, 1, B, pool, a, st);
    tiny_one<T, FWD, 1, 4>(P, 2, B, pool, a, st);
    tiny_one<T, FWD, 2, 1>(P, 3, B, pool, a, st);
    tiny_one<T, FWD, 2, 2>(P, 4, B, pool, a, st);
    tiny_one<T, FWD, 2, 4>(P, 5, B, pool, a, st);
    tiny_one<T, FWD, 4, 1>(P, 6, B, pool, a, st);
    tiny_one<T, FWD, 4, 2>(P, 7, B, pool, a, st);
    tiny_one<T, FWD, 4, 4>(P, 8, B, pool, a, st);
}

// the narrow fronts of a level: one launch per class present
template <class T>
void narrow_fwd(const RsLevel &P, RsState &B, const T *pool, const RsArgs &a, hipStream_t st)
{
    const dim3 blk(256);
    for (int q = 0; q < 3; q++) {
        const int64_t cnt = P.nf_cnt[q];
        if (cnt <= 0) continue;
        const dim3 g((unsigned)((cnt + 3) / 4));
        const RsNTask *t = B.ntask + P.nf_off[q];
        if (q == 0) k_rs_nfwd<T, 1><<<g, blk, 0, st>>>(pool, t, (int)cnt, a, B.buf, B.Yg);
        else if (q == 1) k_rs_nfwd<T, 2><<<g, blk, 0, st>>>(pool, t, (int)cnt, a, B.buf, B.Yg);
        else k_rs_nfwd<T, 4><<<g, blk, 0, st>>>(pool, t, (int)cnt, a, B.buf, B.Yg);
    }
}

template <class T, int NQ>
void narrow_bwd_class(int64_t c1, const RsNFront *f1, int64_t c4, const RsNFront *f4, const T *pool, const RsArgs &a,
                      RsState &B, hipStream_t st)
{
    const dim3 blk(256);
    if (c1 > 0) k_rs_nbwd<T, NQ, 1><<<dim3((unsigned)((c1 + 3) / 4)), blk, 0, st>>>(pool, f1, (int)c1, a, B.Yg, B.Xg);
    if (c4 > 0) k_rs_nbwd<T, NQ, 4><<<dim3((unsigned)c4), blk, 0, st>>>(pool, f4, (int)c4, a, B.Yg, B.Xg);
}

template <class T>
void narrow_bwd(const RsLevel &P, RsState &B, const T *pool, const RsArgs &a, hipStream_t st)
{
    for (int q = 0; q < 3; q++) {
        const int64_t c1 = P.n1_cnt[q], c4 = P.n4_cnt[q];
        const RsNFront *f1 = B.nfront + P.n1_off[q], *f4 = B.nfront + P.n4_off[q];
        if (q == 0) narrow_bwd_class<T, 1>(c1, f1, c4, f4, pool, a, B, st);
        else if (q == 1) narrow_bwd_class<T, 2>(c1, f1, c4, f4, pool, a, B, st);
        else narrow_bwd_class<T, 4>(c1, f1, c4, f4, pool, a, B, st);
    }
}

// forward and backward over the levels, between the permutations
template <class T>
int rs_levels(Factor *F, RsState &B, hipStream_t st)
{
    const Symbolic &S = F->S;
    const T *pool = (const T *)F->pool;
    const RsArgs a{F->foff, F->col0, F->rptr, F->chptr, B.frow, B.uoff, B.lptr, F->Fs, F->ncol, F->rows, B.lst, B.n,
                   B.t0};
    const dim3 blk(256);
    // wide: -1 the level's choice, else 0 / 1 for every level (autotuner runs);
    // tune: events around each level's L21 products (mode 1 at 2 L, mode 2 at 2 L + 1)
    auto parts = [&](int mode, const RsLevel &P, int wide, hipEvent_t e0, hipEvent_t e1, hipStream_t st,
                     float *part) {
        const bool w = (mode == 1 || mode == 2) && (wide < 0 ? P.wide[mode - 1] : wide == 1);
        const int v = w ? mode + 3 : mode;
        const int64_t pc = P.narrow ? P.rpt_cnt[v] : P.pt_cnt[v], fc = P.narrow ? P.rft_cnt[v] : P.ft_cnt[v];
        const RsPart *pt = B.pt + (P.narrow ? P.rpt_off[v] : P.pt_off[v]);
        const RsFin *ft = B.ft + (P.narrow ? P.rft_off[v] : P.ft_off[v]);
        const dim3 gp((unsigned)pc), gf((unsigned)fc);
        if (e0) (void)hipEventRecord(e0, st);
        switch (v) {
        case 0:
            if (pc) k_rs_part<T, 0, RTILE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<0, RTILE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
            break;
        case 1:
            if (pc) k_rs_part<T, 1, RTILE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<1, RTILE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
            break;
        case 2:
            if (pc) k_rs_part<T, 2, RTILE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<2, RTILE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
            break;
        case 3:
            if (pc) k_rs_part<T, 3, RTILE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<3, RTILE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
            break;
        case 4:
            if (pc) k_rs_part<T, 1, RT_WIDE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<1, RT_WIDE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
            break;
        default:
            if (pc) k_rs_part<T, 2, RT_WIDE><<<gp, blk, 0, st>>>(pool, pt, a, B.buf, B.Yg, B.Xg, part);
            if (fc) k_rs_fin<2, RT_WIDE><<<gf, blk, 0, st>>>(ft, a, B.buf, B.Yg, B.Xg, part);
        }
        if (e1) (void)hipEventRecord(e1, st);
    };
    const int wide = B.tune_run;             // -1 outside the autotuner's runs
    hipEvent_t *ev = wide >= 0 ? B.tune_ev[wide].data() : nullptr;
    auto e_at = [&](int i) -> hipEvent_t { return ev ? ev[i] : nullptr; };
    // ea, eb: the autotuner's events before and after the level's L21 product
    auto fwd = [&](const RsLevel &P, hipStream_t s, float *part, hipEvent_t ea, hipEvent_t eb) {
        if (P.narrow) narrow_fwd<T>(P, B, pool, a, s);
        const int64_t asc = P.narrow ? P.ras_cnt : P.as_cnt;
        if (asc) k_rs_assemble<<<dim3((unsigned)asc), blk, 0, s>>>(B.asmt + (P.narrow ? P.ras_off : P.as_off), a, B.buf);
        if (P.sm_cnt) tiny_launch<T, true>(P, B, pool, a, s);
        parts(0, P, wide, nullptr, nullptr, s, part);
        parts(1, P, wide, ea, eb, s, part);
    };
    auto bwd = [&](const RsLevel &P, hipStream_t s, float *part, hipEvent_t ea, hipEvent_t eb) {
        parts(2, P, wide, ea, eb, s, part);
        parts(3, P, wide, nullptr, nullptr, s, part);
        if (P.sm_cnt) tiny_launch<T, false>(P, B, pool, a, s);
        if (P.narrow) narrow_bwd<T>(P, B, pool, a, s);
    };
    // The halves below the root on two streams.  The autotuner's timed runs
    // use one, the first half's product opening the level's timing and the
    // second's closing it.
    const bool two = B.split && wide < 0 && !B.one_stream;
    hipStream_t s2 = two ? B.up_st : st;
    float *part2 = two ? B.part2 : B.part;
    const int32_t lj = B.split ? B.ljoin : S.nlevels;
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[0], st));
        SEHIP_CHECK(hipStreamWaitEvent(s2, B.sev[0], 0));
    }
    for (int32_t L = 0; L < lj; L++) {
        if (B.split) {
            fwd(B.lev[L], st, B.part, e_at(4 * L), nullptr);
            fwd(B.levR[L], s2, part2, nullptr, e_at(4 * L + 1));
        } else {
            fwd(B.lev[L], st, B.part, e_at(4 * L), e_at(4 * L + 1));
        }
    }
    // the root's level pulls both halves' update rows
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[1], s2));
        SEHIP_CHECK(hipStreamWaitEvent(st, B.sev[1], 0));
    }
    for (int32_t L = lj; L < S.nlevels; L++) fwd(B.lev[L], st, B.part, e_at(4 * L), e_at(4 * L + 1));
    for (int32_t L = S.nlevels - 1; L >= lj; L--) bwd(B.lev[L], st, B.part, e_at(4 * L + 2), e_at(4 * L + 3));
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[2], st));
        SEHIP_CHECK(hipStreamWaitEvent(s2, B.sev[2], 0));
    }
    for (int32_t L = std::min(lj, S.nlevels) - 1; L >= 0; L--) {
        if (B.split) {
            bwd(B.lev[L], st, B.part, e_at(4 * L + 2), nullptr);
            bwd(B.levR[L], s2, part2, nullptr, e_at(4 * L + 3));
        } else {
            bwd(B.lev[L], st, B.part, e_at(4 * L + 2), e_at(4 * L + 3));
        }
    }
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[3], s2));
        SEHIP_CHECK(hipStreamWaitEvent(st, B.sev[3], 0));
    }
    return 0;
}

// The autotuner of the L21 products' tile height (SPEIGS_HIP_RS_WIDE=auto):
// the first solve of a factor runs every level with 64-row tiles, the second
// with 128, each level's mode-1 and mode-2 launches bracketed by events; from
// the third solve on each level and mode takes the faster (128 only if at
// least 3 % faster).  The results are the same to the bit either way.
// SPEIGS_HIP_RS_WIDE=1 runs every level with 128 rows; otherwise 64.
int rs_tune_begin(RsState &B, int nlevels)
{
    if (B.tune_state == 0) {
        const char *e = getenv("SPEIGS_HIP_RS_WIDE");
        if (!e || strcmp(e, "auto") != 0) {
            const bool w = e && e[0] == '1';
            for (RsLevel &P : B.lev) P.wide[0] = P.wide[1] = w;
            B.tune_state = 3;
        } else {
            for (int w = 0; w < 2; w++) {
                B.tune_ev[w].assign((size_t)4 * nlevels, nullptr);
                for (hipEvent_t &x : B.tune_ev[w]) SEHIP_CHECK(hipEventCreate(&x));
            }
            B.tune_state = 1;
        }
    }
    B.tune_run = (B.tune_state == 1 || B.tune_state == 2) ? B.tune_state - 1 : -1;
    return 0;
}

int rs_tune_end(RsState &B, int nlevels)
{
    if (B.tune_state == 1) {
        B.tune_state = 2;
    } else if (B.tune_state == 2) {
        for (int w = 0; w < 2; w++)
            for (hipEvent_t x : B.tune_ev[w]) SEHIP_CHECK(hipEventSynchronize(x));
        for (int L = 0; L < nlevels; L++)
            for (int m = 0; m < 2; m++) {
                float t[2] = {0.f, 0.f};
                for (int w = 0; w < 2; w++)
                    SEHIP_CHECK(hipEventElapsedTime(&t[w], B.tune_ev[w][4 * L + 2 * m], B.tune_ev[w][4 * L + 2 * m + 1]));
                B.lev[L].wide[m] = t[1] < 0.97f * t[0];
            }
        for (int w = 0; w < 2; w++) {
            for (hipEvent_t x : B.tune_ev[w]) (void)hipEventDestroy(x);
            B.tune_ev[w].clear();
        }
        B.tune_state = 3;
    }
    B.tune_run = -1;
    return 0;
}

template <class T>
int rs_run(Factor *F, RsState &B, const double *Bin, int64_t brs, int64_t bcs, double *Xout, int64_t xrs,
           int64_t xcs, int nb)
{
    hipStream_t st = F->st;
    const unsigned nrb = (unsigned)((F->n + RS_PERM_ROWS - 1) / RS_PERM_ROWS);
    const dim3 blk(256);
    k_rs_in<<<dim3(nrb), blk, 0, st>>>(Bin, brs, bcs, F->perm, F->n, nb, B.buf);
    int rc = rs_tune_begin(B, F->S.nlevels);
    if (rc) return rc;
    if (B.tune_state != 3) {
        if ((rc = rs_levels<T>(F, B, st))) return rc;   // the autotuner's timed runs (not captured)
        if ((rc = rs_tune_end(B, F->S.nlevels))) return rc;
    } else if (F->opt.use_graph) {
        if (B.gexec && B.gpool != F->pool) {
            SEHIP_CHECK(hipGraphExecDestroy(B.gexec));
            B.gexec = nullptr;
        }
        if (!B.gexec) {
            // thread-local capture: launches of other threads are not drawn in
            SEHIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            const int lr = rs_levels<T>(F, B, st);
            const hipError_t le = hipGetLastError();
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (lr) {
                if (g) (void)hipGraphDestroy(g);
                return lr;
            }
            SEHIP_CHECK(le);
            SEHIP_CHECK(ce);
            const hipError_t ie = hipGraphInstantiate(&B.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            SEHIP_CHECK(ie);
            B.gpool = F->pool;
        }
        SEHIP_CHECK(hipGraphLaunch(B.gexec, st));
    } else if ((rc = rs_levels<T>(F, B, st))) {
        return rc;
    }
    k_rs_out<<<dim3((unsigned)((F->n + 7) / 8)), blk, 0, st>>>(B.Xg, F->perm, F->n, nb, Xout, xrs, xcs);
    return 0;
}

// one right-hand side: b, x device FP64 vectors of n (original order)
template <class T>
int vs_run(Factor *F, RsState &B, const double *b, double *x)
{
    const Symbolic &S = F->S;
    hipStream_t st = F->st;
    const T *pool = (const T *)F->pool;
    const RsArgs a{F->foff, F->col0, F->rptr, F->chptr, B.frow, B.uoff, B.lptr, F->Fs, F->ncol, F->rows, B.lst, B.n,
                   B.t0};
    const dim3 blk(256);
    const unsigned nb = (unsigned)((F->n + 255) / 256);
    k_vs_in<<<dim3(nb), blk, 0, st>>>(b, F->perm, F->n, B.vbuf);
    auto parts = [&](int mode, const RsLevel &P, hipStream_t st, double *vpart) {
        const int64_t pc = P.pt_cnt[mode], fc = P.ft_cnt[mode];
        const RsPart *pt = B.pt + P.pt_off[mode];
        const RsFin *ft = B.ft + P.ft_off[mode];
        const dim3 gp((unsigned)pc), gf((unsigned)fc), b64(64);
        switch (mode) {
        case 0:
            if (pc) k_vs_part<T, 0><<<gp, blk, 0, st>>>(pool, pt, a, B.vbuf, B.vY, B.vX, vpart);
            if (fc) k_vs_fin<0><<<gf, b64, 0, st>>>(ft, a, B.vbuf, B.vY, B.vX, vpart);
            break;
        case 1:
            if (pc) k_vs_part<T, 1><<<gp, blk, 0, st>>>(pool, pt, a, B.vbuf, B.vY, B.vX, vpart);
            if (fc) k_vs_fin<1><<<gf, b64, 0, st>>>(ft, a, B.vbuf, B.vY, B.vX, vpart);
            break;
        case 2:
            if (pc) k_vs_part<T, 2><<<gp, blk, 0, st>>>(pool, pt, a, B.vbuf, B.vY, B.vX, vpart);
            if (fc) k_vs_fin<2><<<gf, b64, 0, st>>>(ft, a, B.vbuf, B.vY, B.vX, vpart);
            break;
        default:
            if (pc) k_vs_part<T, 3><<<gp, blk, 0, st>>>(pool, pt, a, B.vbuf, B.vY, B.vX, vpart);
            if (fc) k_vs_fin<3><<<gf, b64, 0, st>>>(ft, a, B.vbuf, B.vY, B.vX, vpart);
        }
    };
    auto fwd = [&](const RsLevel &P, hipStream_t st, double *vpart) {
        if (P.as_cnt)
            k_vs_assemble<<<dim3((unsigned)((P.as_cnt + VS_ASM - 1) / VS_ASM)), blk, 0, st>>>(B.asmt + P.as_off,
                                                                                            P.as_cnt, a, B.vbuf);
        if (P.sm_cnt)
            k_vs_tiny_fwd<T><<<dim3((unsigned)((P.sm_cnt + 3) / 4)), blk, 0, st>>>(pool, B.small + P.sm_off,
                                                                                  (int)P.sm_cnt, a, B.vbuf, B.vY);
        parts(0, P, st, vpart);
        parts(1, P, st, vpart);
    };
    auto bwd = [&](const RsLevel &P, hipStream_t st, double *vpart) {
        parts(2, P, st, vpart);
        parts(3, P, st, vpart);
        if (P.sm_cnt)
            k_vs_tiny_bwd<T><<<dim3((unsigned)((P.sm_cnt + 3) / 4)), blk, 0, st>>>(pool, B.small + P.sm_off,
                                                                                  (int)P.sm_cnt, a, B.vY, B.vX);
    };
    // the halves below the root on two streams, joined at the root and at the end
    const bool two = B.split && !B.one_stream;
    hipStream_t s2 = two ? B.up_st : st;
    const int32_t lj = B.split ? B.ljoin : S.nlevels;
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[0], st));
        SEHIP_CHECK(hipStreamWaitEvent(s2, B.sev[0], 0));
    }
    for (int32_t L = 0; L < lj; L++) {
        fwd(B.lev[L], st, B.vpart);
        fwd(B.levR[L], s2, two ? B.vpart2 : B.vpart);
    }
    // the root's level pulls both halves' update rows
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[1], s2));
        SEHIP_CHECK(hipStreamWaitEvent(st, B.sev[1], 0));
    }
    for (int32_t L = lj; L < S.nlevels; L++) fwd(B.lev[L], st, B.vpart);
    if (F->signed_fac) k_vs_sign<T><<<dim3(nb), blk, 0, st>>>(B.vY, (const T *)F->sgn, F->n);
    for (int32_t L = S.nlevels - 1; L >= lj; L--) bwd(B.lev[L], st, B.vpart);
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[2], st));
        SEHIP_CHECK(hipStreamWaitEvent(s2, B.sev[2], 0));
    }
    for (int32_t L = std::min(lj, S.nlevels) - 1; L >= 0; L--) {
        bwd(B.lev[L], st, B.vpart);
        bwd(B.levR[L], s2, two ? B.vpart2 : B.vpart);
    }
    if (two) {
        SEHIP_CHECK(hipEventRecord(B.sev[3], s2));
        SEHIP_CHECK(hipStreamWaitEvent(st, B.sev[3], 0));
    }
    k_vs_out<<<dim3(nb), blk, 0, st>>>(B.vX, F->perm, F->n, x);
    return 0;
}

// join the plan's helper thread and its upload stream
int rs_ready(RsState &B)
{
    if (B.ready) return 0;
    if (B.th.joinable()) B.th.join();
    if (B.plan_rc) {
        note_error(B.plan_msg.c_str());
        return B.plan_rc;
    }
    SEHIP_CHECK(hipStreamSynchronize(B.up_st));
    B.ready = true;
    return 0;
}

}  // namespace

int factor_vec_prepare(Factor *F)
{
    rowsolve_prepare_async(F);
    RsState &B = *reinterpret_cast<RsState *>(F->rs);
    if (!B.vec) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "vector solve: the plan was made without its buffers");
    int rc = rs_ready(B);
    if (rc) return rc;
    return factor_invert_l11(F);
}

bool factor_vec_prepared(const Factor *F)
{
    const RsState *B = reinterpret_cast<const RsState *>(F->rs);
    return B && B->ready && B->vec && F->inverted;
}

int factor_vec_solve(Factor *F, const double *b, double *x)
{
    int rc = factor_vec_prepare(F);
    if (rc) return rc;
    RsState &B = *reinterpret_cast<RsState *>(F->rs);
    return F->fp32 ? vs_run<float>(F, B, b, x) : vs_run<double>(F, B, b, x);
}

int factor_block_solve_rows(Factor *F, const double *Bin, int64_t brs, int64_t bcs, double *Xout, int64_t xrs,
                            int64_t xcs, int nb)
{
    if (nb < 1 || nb > RW) return SPEIGS_ERR_ARG;
    if (!F->inverted) return SPEIGS_ERR_FACTOR;
    rowsolve_prepare_async(F);
    RsState &B = *reinterpret_cast<RsState *>(F->rs);
    if (!B.rows) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "block solve: the plan was made without its buffers");
    int rc0 = rs_ready(B);
    if (rc0) return rc0;
    const int e = ev_begin(F);
    const int rc = F->fp32 ? rs_run<float>(F, B, Bin, brs, bcs, Xout, xrs, xcs, nb)
                           : rs_run<double>(F, B, Bin, brs, bcs, Xout, xrs, xcs, nb);
    if (rc) return rc;
    ev_end(F, e);
    F->block_solves++;
    F->block_rhs += nb;
    return 0;
}

void rowsolve_prepare_async(Factor *F)
{
    if (F->rs) return;
    RsState *B = new RsState();
    F->rs = B;
    // the block path's row buffers, the single-vector path's value buffers
    B->rows = F->opt.precision == SPEIGS_HIP_FP32_BLOCK;
    B->vec = F->solve_mode == 3;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) (void)hipGetLastError();
    B->th = std::thread([F, B, dev] {
        auto run = [&]() -> int {
            int rc = rs_plan(F, *B);
            if (rc) return rc;
            SEHIP_CHECK(hipSetDevice(dev));
            SEHIP_CHECK(hipStreamCreate(&B->up_st));
            return rs_upload(F, *B, B->up_st);
        };
        B->plan_rc = run();
        if (B->plan_rc) B->plan_msg = last_error_text();
    });
}

void rowsolve_free(Factor *F)
{
    delete reinterpret_cast<RsState *>(F->rs);
    F->rs = nullptr;
}

}  // namespace sehip
