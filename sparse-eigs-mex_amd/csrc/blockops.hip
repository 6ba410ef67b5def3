// blockops.hip -- kernels on row-major multivectors (n x w, element (r, c) at
// X[r * ld + c]) for the block eigensolver.  A block of m <= 96 vectors is
// stored row by row so that every kernel streams contiguous memory: a chunk of
// rows of all columns is one contiguous range, the SpMM gathers one contiguous
// row of X per nonzero, and the page footprint of a chunk is one or two pages
// instead of one page per column.
#include <algorithm>
#include <cstdint>

#include "dev.h"

namespace sehip {

// ------------------------------------------------------------------ Gram
// G = A^T B (p x q, p, q <= 96, column-major ld p) on the FP64 matrix cores.
// A workgroup owns a contiguous range of rows; its four waves split the (<= 6
// x 6) output tiles of 16 x 16 into 2 x 2 groups of <= 3 x 3, halving the tile
// rows and the tile columns (each wave loads only its column blocks, the
// shared rows come through L1), which keeps the accumulators at 72 registers
// and several waves per SIMD in flight.
// v_mfma_f64_16x16x4 takes its operands straight from memory, 4 rows per step:
// lane l loads A[r + l/16][16 ti + l%16] and B[r + l/16][16 tj + l%16] (128-byte
// row pieces) and holds G[16 ti + l/16 + 4v][16 tj + l%16].  The operands of
// the next RG_AHEAD steps are in flight during the current step's products (one
// step ahead left the small Grams waiting on memory: 278 -> 215 us for a 28 x 28
// Gram over 10^6 rows, 466 -> 343 us for 84 x 56).  One partial per workgroup,
// then a fixed-order sum.
constexpr int RG_WAVES = 1024;     // workgroups over the rows
constexpr int RG_AHEAD = 4;        // steps of 4 rows loaded ahead

typedef double rd4 __attribute__((ext_vector_type(4)));

template <int GA, int GB>
__global__ void __launch_bounds__(256) k_rgram_partial(const double *__restrict__ A, int64_t lda, int p,
                                                       const double *__restrict__ B, int64_t ldb, int q, int64_t n,
                                                       int64_t rows_per_wave, double *__restrict__ partial)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int pt = (p + 15) >> 4, qt = (q + 15) >> 4;
    // halves of the tile rows and of the tile columns (GA = ceil(pt / 2), GB =
    // ceil(qt / 2) each): the waves' shares differ by at most one row or
    // column of tiles, and a wave's group is always GA x GB products (a tile
    // past the matrix multiplies clamped columns into outputs never stored),
    // so no product waits on a branch
    const int a0 = GA * (wv & 1), b0 = GB * (wv >> 1);   // this wave's tile group
    if (a0 >= pt || b0 >= qt) return;                     // no workgroup barrier below
    rd4 acc[GA][GB];
#pragma unroll
    for (int a = 0; a < GA; a++)
#pragma unroll
        for (int b = 0; b < GB; b++) acc[a][b] = rd4{0.0, 0.0, 0.0, 0.0};
    // (the last workgroups may start past n: no rows)
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_wave, r1 = max(r0, min(n, r0 + rows_per_wave));
    const int64_t rfull = r0 + ((r1 - r0) & ~(int64_t)3);  // whole steps of 4 rows
    // columns clamped into the matrices: a column past p (q) feeds only rows
    // (columns) of the product past p (q), which are not stored
    int64_t ofa[GA], ofb[GB];
#pragma unroll
    for (int a = 0; a < GA; a++) ofa[a] = min(16 * (a0 + a) + lr, p - 1);
#pragma unroll
    for (int b = 0; b < GB; b++) ofb[b] = min(16 * (b0 + b) + lr, q - 1);
    // ring of RG_AHEAD steps: slot s holds the rows r + 4 s of the current
    // round.  While the next round is whole its loads are unconditional (a
    // load under a branch makes the step wait for every load in flight); the
    // last round in the ring is then multiplied out, and the steps left (fewer
    // than a round) go one at a time
    double xa[RG_AHEAD][GA], xb[RG_AHEAD][GB];
    auto load = [&](int s, int64_t r) {
        const int64_t row = r + lk;
#pragma unroll
        for (int a = 0; a < GA; a++) xa[s][a] = A[row * lda + ofa[a]];
#pragma unroll
        for (int b = 0; b < GB; b++) xb[s][b] = B[row * ldb + ofb[b]];
    };
    auto mul = [&](int s) {
#pragma unroll
        for (int a = 0; a < GA; a++)
#pragma unroll
            for (int b = 0; b < GB; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[s][a], xb[s][b], acc[a][b], 0, 0, 0);
    };
    constexpr int64_t RND = 4 * RG_AHEAD;
    int64_t r = r0;
    if (r + RND <= rfull) {
#pragma unroll
        for (int s = 0; s < RG_AHEAD; s++) load(s, r + 4 * s);
        for (; r + 2 * RND <= rfull; r += RND) {
#pragma unroll
            for (int s = 0; s < RG_AHEAD; s++) {
                double ca[GA], cb[GB];
#pragma unroll
                for (int a = 0; a < GA; a++) ca[a] = xa[s][a];
#pragma unroll
                for (int b = 0; b < GB; b++) cb[b] = xb[s][b];
                load(s, r + RND + 4 * s);
#pragma unroll
                for (int a = 0; a < GA; a++)
#pragma unroll
                    for (int b = 0; b < GB; b++)
                        acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[a], cb[b], acc[a][b], 0, 0, 0);
            }
        }
#pragma unroll
        for (int s = 0; s < RG_AHEAD; s++) mul(s);
        r += RND;
    }
    for (; r < rfull; r += 4) {
        load(0, r);
        mul(0);
    }
    if (rfull < r1) {            // the last 1-3 rows: the step's other rows are zero
        const int64_t row = rfull + lk;
        const bool in = row < r1;
        const int64_t rc = in ? row : r1 - 1;
        double ca[GA], cb[GB];
#pragma unroll
        for (int a = 0; a < GA; a++) ca[a] = A[rc * lda + ofa[a]];
#pragma unroll
        for (int b = 0; b < GB; b++) cb[b] = B[rc * ldb + ofb[b]];
#pragma unroll
        for (int a = 0; a < GA; a++) ca[a] = in ? ca[a] : 0.0;
#pragma unroll
        for (int b = 0; b < GB; b++) cb[b] = in ? cb[b] : 0.0;
#pragma unroll
        for (int a = 0; a < GA; a++)
#pragma unroll
            for (int b = 0; b < GB; b++)
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[a], cb[b], acc[a][b], 0, 0, 0);
    }
    double *out = partial + (int64_t)blockIdx.x * p * q;
#pragma unroll
    for (int a = 0; a < GA; a++)
#pragma unroll
        for (int b = 0; b < GB; b++) {
            if (a0 + a >= pt || b0 + b >= qt) continue;
            const int j = 16 * (b0 + b) + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * (a0 + a) + lk + 4 * v;
                if (i < p && j < q) out[(int64_t)j * p + i] = acc[a][b][v];
            }
        }
}

// G[idx] = sum over waves of partial[w][idx]: 4 threads per output over the
// residues of w mod 4, 8 independent accumulators each, fixed combination order
__global__ void __launch_bounds__(256) k_rgram_reduce(const double *__restrict__ partial, int64_t nw, int pq,
                                                      double *__restrict__ G)
{
    __shared__ double red[4][64];
    const int o = threadIdx.x & 63, sg = threadIdx.x >> 6;
    const int idx = blockIdx.x * 64 + o;
    double a[8];
#pragma unroll
    for (int u = 0; u < 8; u++) a[u] = 0.0;
    if (idx < pq) {
        int64_t b = sg;
        for (; b + 28 < nw; b += 32) {
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] += partial[(b + 4 * u) * pq + idx];
        }
        for (int u = 0; b < nw; b += 4, u++) a[u & 7] += partial[b * pq + idx];
    }
    red[sg][o] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
    __syncthreads();
    if (sg == 0 && idx < pq) G[idx] = (red[0][o] + red[1][o]) + (red[2][o] + red[3][o]);
}

static int64_t rgram_rpw(int64_t n) { return ((n + RG_WAVES - 1) / RG_WAVES + 3) / 4 * 4; }

// partial slots (workgroups) of one Gram: a multiple of 4 covering n rows
int64_t rgram_blocks(int64_t n)
{
    const int64_t w = (n + rgram_rpw(n) - 1) / rgram_rpw(n);
    return (w + 3) / 4 * 4;
}

void rgram(const double *A, int64_t lda, int p, const double *B, int64_t ldb, int q, int64_t n, double *partial,
           double *G, hipStream_t st)
{
    const int64_t rpw = rgram_rpw(n), nw = rgram_blocks(n);
    const int ga = ((p + 15) / 16 + 1) / 2, gb = ((q + 15) / 16 + 1) / 2;    // <= 3 each (p, q <= 96)
    const dim3 g((unsigned)nw), t(256);
    switch (3 * (ga - 1) + (gb - 1)) {
    case 0: k_rgram_partial<1, 1><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 1: k_rgram_partial<1, 2><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 2: k_rgram_partial<1, 3><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 3: k_rgram_partial<2, 1><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 4: k_rgram_partial<2, 2><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 5: k_rgram_partial<2, 3><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 6: k_rgram_partial<3, 1><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    case 7: k_rgram_partial<3, 2><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    default: k_rgram_partial<3, 3><<<g, t, 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial); break;
    }
    k_rgram_reduce<<<dim3((unsigned)((p * q + 63) / 64)), dim3(256), 0, st>>>(partial, nw, p * q, G);
}

// ------------------------------------------------------------------ tall GEMM
// X = V C + beta X (V n x m, C m x p column-major, m, p <= 96; X must not alias
// V) on the FP64 matrix cores.  C is staged in LDS once per block; each wave
// owns 16-row groups and all (<= 6) 16-column output tiles, loads a group's
// whole 16 x m slice of V at once (all loads in flight), then runs the steps of 4 inner
// indices: lane l supplies V[r + l%16][k + l/16] and C[k + l/16][16 tj + l%16]
// and holds X[r + l/16 + 4v][16 tj + l%16] (stores of four 128-byte row pieces).
constexpr int RM_CP = 100;   // LDS pitch of C rows

constexpr int RM_GROUPS = 4;   // 16-row groups per wave

// X may be V itself (same rows, same stride): a wave reads all of its 16 rows'
// inputs before it writes any output row, and no other wave touches them.
// PT = ceil(p / 16) output tiles and NS >= ceil(m / 4) steps are fixed per
// instance, so no product or load waits on a branch: rows past n are read at
// row n - 1 (they feed only output rows that are not stored) and inner indices
// past m at column m - 1 (C's rows past m are zero in LDS)
template <int PT, int NS>
__global__ void __launch_bounds__(256) k_rgemm(const double *V, int64_t ldv, int m,
                                               const double *__restrict__ C, int p, double *X,
                                               int64_t ldx, int64_t n, double beta)
{
    __shared__ double Cs[96][RM_CP];   // Cs[k][c] = C[k + c m]
    const int mp = m * p;
    for (int e = threadIdx.x; e < mp; e += 256) {
        const int c = e / m, k = e - c * m;
        Cs[k][c] = C[e];
    }
    // zero the rows past m up to the last step and the columns past p up to
    // the last 16-column tile
    constexpr int K4 = 4 * NS, P16 = 16 * PT;
    for (int e = threadIdx.x; e < K4 * P16; e += 256) {
        const int k = e / P16, c = e - k * P16;
        if (k >= m || c >= p) Cs[k][c] = 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int64_t rbase = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RM_GROUPS * 16;
    // one group's 16 x m slice of V in registers at a time: loading the next
    // group's during the current products (double the registers, occupancy 2
    // instead of 3) was slower -- 375 vs 342 us for 84 -> 56 columns at 10^6 rows
    for (int g = 0; g < RM_GROUPS; g++) {
        const int64_t r0 = rbase + g * 16;
        if (r0 >= n) break;
        const int64_t arow = min(r0 + lr, n - 1);
        double va[NS];
#pragma unroll
        for (int s2 = 0; s2 < NS; s2++) va[s2] = V[arow * ldv + min(4 * s2 + lk, m - 1)];
        rd4 acc[PT];
#pragma unroll
        for (int b = 0; b < PT; b++) acc[b] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < NS; s2++)
#pragma unroll
            for (int b = 0; b < PT; b++)
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[s2], Cs[4 * s2 + lk][16 * b + lr], acc[b], 0, 0, 0);
#pragma unroll
        for (int b = 0; b < PT; b++) {
            const int c = 16 * b + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t row = r0 + lk + 4 * v;
                if (row < n && c < p) {
                    double *xp = X + row * ldx + c;
                    *xp = beta == 0.0 ? acc[b][v] : acc[b][v] + beta * *xp;
                }
            }
        }
    }
}

template <int PT>
static void rgemm_steps(const double *V, int64_t ldv, int m, const double *C, int p, double *X, int64_t ldx,
                        int64_t n, double beta, hipStream_t st)
{
    const int64_t rows = 64 * RM_GROUPS;
    const dim3 g((unsigned)((n + rows - 1) / rows)), t(256);
    if (m <= 28) k_rgemm<PT, 7><<<g, t, 0, st>>>(V, ldv, m, C, p, X, ldx, n, beta);
    else if (m <= 56) k_rgemm<PT, 14><<<g, t, 0, st>>>(V, ldv, m, C, p, X, ldx, n, beta);
    else if (m <= 84) k_rgemm<PT, 21><<<g, t, 0, st>>>(V, ldv, m, C, p, X, ldx, n, beta);
    else k_rgemm<PT, 24><<<g, t, 0, st>>>(V, ldv, m, C, p, X, ldx, n, beta);
}

void rgemm(const double *V, int64_t ldv, int m, const double *C, int p, double *X, int64_t ldx, int64_t n,
           double beta, hipStream_t st)
{
    if (n <= 0 || m <= 0 || p <= 0) return;     // (every caller has 1 <= m, p <= 96)
    switch ((p + 15) / 16) {
    case 1: rgemm_steps<1>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    case 2: rgemm_steps<2>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    case 3: rgemm_steps<3>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    case 4: rgemm_steps<4>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    case 5: rgemm_steps<5>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    default: rgemm_steps<6>(V, ldv, m, C, p, X, ldx, n, beta, st); break;
    }
}

// ------------------------------------------------------------------ Ritz step
// The block eigensolver's Rayleigh-Ritz update and the next residuals in one
// pass over S and AS (row-major, row stride ld, d <= 96 columns): with
// C2 = [C | C_P] (d x 2m, column-major, m <= 32) on the device,
//   Sn[:, 0:2m) = S C2            (new X, new P)
//   ASn[:, 0:m) = AS C            (new A X)
//   Sn[:, 2m:3m) = R = ASn[:, 0:m) - Sn[:, 0:m) diag(lam)
// and, for the first k columns, per-workgroup sums of R's squares (part,
// 32 per workgroup), summed in a fixed order by rritz_norms_reduce.  The
// products are k_rgemm's (the same 16-row groups, inner steps and tiles), so
// Sn and ASn are bitwise those of two rgemm calls, and R is formed from the
// same doubles as a separate residual pass would read back.  Replaces two
// tall products and the residual pass (three reads of n x m blocks).
constexpr int RR_GROUPS = 4;   // 16-row groups per wave

// PT = ceil(2m / 16) tiles of the S product, MT = ceil(m / 16) of the AS product
template <int PT, int MT>
__global__ void __launch_bounds__(256) k_rritz(const double *__restrict__ S, const double *__restrict__ AS, int64_t ld,
                                               int d, const double *__restrict__ C2, int m,
                                               const double *__restrict__ lam, int k, double *__restrict__ Sn,
                                               double *__restrict__ ASn, int64_t n, double *__restrict__ part)
{
    __shared__ double Cs[96][RM_CP];   // Cs[kk][c] = C2[kk + c d]
    __shared__ double red[4][32];
    const int p = 2 * m;
    for (int e = threadIdx.x; e < d * p; e += 256) {
        const int c = e / d, kk = e - c * d;
        Cs[kk][c] = C2[e];
    }
    const int d4 = (d + 3) & ~3, p16 = (p + 15) & ~15;
    for (int e = threadIdx.x; e < d4 * p16; e += 256) {
        const int kk = e / p16, c = e - kk * p16;
        if (kk >= d || c >= p) Cs[kk][c] = 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const double l0 = lr < m ? lam[lr] : 0.0, l1 = 16 + lr < m ? lam[16 + lr] : 0.0;
    double ns0 = 0.0, ns1 = 0.0;      // this lane's sums of squares, columns lr and 16 + lr
    const int64_t rbase = ((int64_t)blockIdx.x * 4 + wv) * RR_GROUPS * 16;
    for (int g = 0; g < RR_GROUPS; g++) {
        const int64_t r0 = rbase + g * 16;
        if (r0 >= n) break;
        const int64_t arow = r0 + lr;
        const bool rin = arow < n;
        double va[24];
        rd4 acc[PT], acc2[MT];
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            const int kk = 4 * s2 + lk;
            va[s2] = (rin && kk < d) ? S[arow * ld + kk] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < PT; b++) acc[b] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            if (4 * s2 >= d) break;
#pragma unroll
            for (int b = 0; b < PT; b++)
                acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[s2], Cs[4 * s2 + lk][16 * b + lr], acc[b], 0, 0, 0);
        }
        // the new P (tiles 2, 3) leaves now
#pragma unroll
        for (int b = MT; b < PT; b++) {
            const int c = 16 * b + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t row = r0 + lk + 4 * v;
                if (row < n && c < p) Sn[row * ld + c] = acc[b][v];
            }
        }
        // (C's first two tiles are read from LDS again below: keeping the
        // values read above would hold 48 more doubles per lane)
        asm volatile("" ::: "memory");
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            const int kk = 4 * s2 + lk;
            va[s2] = (rin && kk < d) ? AS[arow * ld + kk] : 0.0;
        }
#pragma unroll
        for (int b = 0; b < MT; b++) acc2[b] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            if (4 * s2 >= d) break;
#pragma unroll
            for (int b = 0; b < MT; b++)
                acc2[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[s2], Cs[4 * s2 + lk][16 * b + lr], acc2[b], 0, 0, 0);
        }
        // lane (lr, lk), register v: row r0 + lk + 4v, column 16 b + lr
#pragma unroll
        for (int b = 0; b < MT; b++) {
            const int c = 16 * b + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t row = r0 + lk + 4 * v;
                if (row < n && c < p) Sn[row * ld + c] = acc[b][v];
            }
        }
#pragma unroll
        for (int b = 0; b < MT; b++) {
            const int c = 16 * b + lr;
            const double lc = b == 0 ? l0 : l1;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t row = r0 + lk + 4 * v;
                if (row < n && c < m) {
                    const double ax = acc2[b][v], x = acc[b][v];
                    ASn[row * ld + c] = ax;
                    const double rv = ax - lc * x;
                    Sn[row * ld + p + c] = rv;
                    if (c < k) {
                        if (b == 0) ns0 = fma(rv, rv, ns0);
                        else ns1 = fma(rv, rv, ns1);
                    }
                }
            }
        }
    }
    // column sums: the four lane groups of a wave, then the four waves, in order
    ns0 += __shfl_xor(ns0, 16);
    ns0 += __shfl_xor(ns0, 32);
    ns1 += __shfl_xor(ns1, 16);
    ns1 += __shfl_xor(ns1, 32);
    if (lk == 0) {
        red[wv][lr] = ns0;
        red[wv][16 + lr] = ns1;
    }
    __syncthreads();
    if (threadIdx.x < 32)
        part[(int64_t)blockIdx.x * 32 + threadIdx.x] =
            (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// sums of per-workgroup column partials (32 per workgroup) in a fixed order:
// workgroup j of the launch sums partials [64 j, 64 j + 64) -- eight
// interleaved partial sums per column, then their sum -- into out[j][c]
__global__ void __launch_bounds__(256) k_part_reduce(const double *__restrict__ part, int64_t nb,
                                                     double *__restrict__ out)
{
    __shared__ double red[8][32];
    const int c = threadIdx.x & 31, sl = threadIdx.x >> 5;
    const int64_t b0 = (int64_t)blockIdx.x * 64, b1 = min(nb, b0 + 64);
    double t = 0.0;
    for (int64_t b = b0 + sl; b < b1; b += 8) t += part[b * 32 + c];
    red[sl][c] = t;
    __syncthreads();
    if (threadIdx.x < 32) {
        double u = 0.0;
        for (int i = 0; i < 8; i++) u += red[i][c];
        out[(int64_t)blockIdx.x * 32 + c] = u;
    }
}

int64_t rritz_blocks(int64_t n) { return (n + 64 * RR_GROUPS - 1) / (64 * RR_GROUPS); }

// per-workgroup partials and the levels of their sums
int64_t rritz_part(int64_t n)
{
    int64_t cnt = rritz_blocks(n), tot = cnt;
    while (cnt > 1) {
        cnt = (cnt + 63) / 64;
        tot += cnt;
    }
    return 32 * tot;
}

void rritz(const double *S, const double *AS, int64_t ld, int d, const double *C2, int m, const double *lam, int k,
           double *Sn, double *ASn, int64_t n, double *part, double *out, hipStream_t st)
{
    const int64_t nb = rritz_blocks(n);
    const dim3 g((unsigned)nb), t(256);
    if (m <= 8) k_rritz<1, 1><<<g, t, 0, st>>>(S, AS, ld, d, C2, m, lam, k, Sn, ASn, n, part);
    else if (m <= 16) k_rritz<2, 1><<<g, t, 0, st>>>(S, AS, ld, d, C2, m, lam, k, Sn, ASn, n, part);
    else if (m <= 24) k_rritz<3, 2><<<g, t, 0, st>>>(S, AS, ld, d, C2, m, lam, k, Sn, ASn, n, part);
    else k_rritz<4, 2><<<g, t, 0, st>>>(S, AS, ld, d, C2, m, lam, k, Sn, ASn, n, part);
    // groups of 64 partials summed level by level (each level after the last
    // in part), the last level into out
    const double *src = part;
    double *nxt_at = part + 32 * nb;
    int64_t cnt = nb;
    for (;;) {
        const int64_t nxt = (cnt + 63) / 64;
        double *dst = nxt == 1 ? out : nxt_at;
        k_part_reduce<<<dim3((unsigned)nxt), t, 0, st>>>(src, cnt, dst);
        if (nxt == 1) break;
        src = dst;
        nxt_at = dst + 32 * nxt;
        cnt = nxt;
    }
}

// ------------------------------------------------------------------ SpMM
// Y = A X for row-major X, Y (nv <= 96 columns): SELL-64 slices; a 16-lane
// group per row, lane j of the group on columns j, 16 + j, ..., so each
// nonzero's value and column are read once for all columns and each gather of
// row col of X is a contiguous 128-byte piece.  A block = 16 rows.
template <int NC>
__global__ void __launch_bounds__(256) k_sell_rspmm(int64_t nslices, const int64_t *__restrict__ sptr,
                                                    const int32_t *__restrict__ swid, const int32_t *__restrict__ col,
                                                    const double *__restrict__ val, const double *__restrict__ X,
                                                    int64_t ldx, double *__restrict__ Y, int64_t ldy, int64_t n,
                                                    int nv)
{
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int64_t sl = row >> 6;
    if (sl >= nslices || row >= n) return;
    const int t = threadIdx.x & 15;
    int64_t pp = sptr[sl] + (row & 63);
    const int w = swid[sl];
    double s[NC];
#pragma unroll
    for (int j = 0; j < NC; j++) s[j] = 0.0;
    // eight nonzeros at a time: their columns and values, then all of their X
    // loads, are in flight together; the sums run in the nonzeros' order
    for (int k0 = 0; k0 < w; k0 += 8, pp += 8 * 64) {
        const int kc = min(8, w - k0);
        double a[8];
        int64_t cc[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a[u] = u < kc ? val[pp + 64 * u] : 0.0;
            cc[u] = u < kc ? (int64_t)col[pp + 64 * u] : row;
        }
        double xv[8][NC];
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int j = 0; j < NC; j++)
                xv[u][j] = (u < kc && 16 * j + t < nv) ? X[cc[u] * ldx + 16 * j + t] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (u < kc)
#pragma unroll
                for (int j = 0; j < NC; j++)
                    if (16 * j + t < nv) s[j] = fma(a[u], xv[u][j], s[j]);
    }
#pragma unroll
    for (int j = 0; j < NC; j++)
        if (16 * j + t < nv) Y[row * ldy + 16 * j + t] = s[j];
}

// 33..64 columns: lane t of a row's 16-lane group owns columns 4t .. 4t + 3
// and gathers them as two 16-byte loads, one 512-byte load instruction per
// nonzero for four rows instead of four 128-byte ones (the sums per column run
// in the same nonzero order as above; 554 -> 494 us for 56 columns at 10^6
// rows).  X and Y rows must be 16-byte aligned.  FULL (nv a multiple of 4: a
// lane has four columns or none) issues every load with no branch: the slots
// past a row's width in its last chunk read the row's last slot again with a
// zero coefficient (adding exact zeros), a lane with no columns reads columns
// 0..3 and stores nothing.
template <bool FULL>
__global__ void __launch_bounds__(256) k_sell_rspmm_q(int64_t nslices, const int64_t *__restrict__ sptr,
                                                      const int32_t *__restrict__ swid,
                                                      const int32_t *__restrict__ col,
                                                      const double *__restrict__ val, const double *__restrict__ X,
                                                      int64_t ldx, double *__restrict__ Y, int64_t ldy, int64_t n,
                                                      int nv)
{
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int64_t sl = row >> 6;
    if (sl >= nslices || row >= n) return;
    const int c0 = 4 * (threadIdx.x & 15);
    const int nc = min(4, nv - c0);   // columns of this lane (<= 0: none)
    int64_t pp = sptr[sl] + (row & 63);
    const int w = swid[sl];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (FULL) {
        const int cx = nc > 0 ? c0 : 0;
        for (int k0 = 0; k0 < w; k0 += 8, pp += 8 * 64) {
            const int kc = min(8, w - k0);
            double a[8];
            int64_t cc[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int64_t q = pp + 64 * min(u, kc - 1);
                a[u] = val[q];
                cc[u] = col[q];
            }
            double2 xv[8][2];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double2 *xr = reinterpret_cast<const double2 *>(X + cc[u] * ldx + cx);
                xv[u][0] = xr[0];
                xv[u][1] = xr[1];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const double au = u < kc ? a[u] : 0.0;
                s[0] = fma(au, xv[u][0].x, s[0]);
                s[1] = fma(au, xv[u][0].y, s[1]);
                s[2] = fma(au, xv[u][1].x, s[2]);
                s[3] = fma(au, xv[u][1].y, s[3]);
            }
        }
        if (nc > 0) {
            double2 *yr = reinterpret_cast<double2 *>(Y + row * ldy + c0);
            yr[0] = double2{s[0], s[1]};
            yr[1] = double2{s[2], s[3]};
        }
        return;
    }
    for (int k0 = 0; k0 < w; k0 += 8, pp += 8 * 64) {
        const int kc = min(8, w - k0);
        double a[8];
        int64_t cc[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            a[u] = u < kc ? val[pp + 64 * u] : 0.0;
            cc[u] = u < kc ? (int64_t)col[pp + 64 * u] : row;
        }
        double2 xv[8][2];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const double *xr = X + cc[u] * ldx + c0;
            if (u < kc && nc == 4) {
                xv[u][0] = reinterpret_cast<const double2 *>(xr)[0];
                xv[u][1] = reinterpret_cast<const double2 *>(xr)[1];
            } else {
                const bool ok = u < kc;
                xv[u][0] = double2{ok && nc > 0 ? xr[0] : 0.0, ok && nc > 1 ? xr[1] : 0.0};
                xv[u][1] = double2{ok && nc > 2 ? xr[2] : 0.0, 0.0};
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (u < kc) {
                s[0] = fma(a[u], xv[u][0].x, s[0]);
                s[1] = fma(a[u], xv[u][0].y, s[1]);
                s[2] = fma(a[u], xv[u][1].x, s[2]);
                s[3] = fma(a[u], xv[u][1].y, s[3]);
            }
    }
    double *yr = Y + row * ldy + c0;
    if (nc == 4) {
        reinterpret_cast<double2 *>(yr)[0] = double2{s[0], s[1]};
        reinterpret_cast<double2 *>(yr)[1] = double2{s[2], s[3]};
    } else {
        for (int j = 0; j < nc; j++) yr[j] = s[j];
    }
}

void rspmm(const SellA &S, const double *X, int64_t ldx, double *Y, int64_t ldy, int nv, hipStream_t st)
{
    const dim3 g((unsigned)((S.n + 15) / 16)), b(256);
    const bool al16 = ((uintptr_t)X | (uintptr_t)Y) % 16 == 0 && ldx % 2 == 0 && ldy % 2 == 0;
    if (nv > 32 && nv <= 64 && al16) {
        if (nv % 4 == 0)
            k_sell_rspmm_q<true><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv);
        else
            k_sell_rspmm_q<false><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv);
        return;
    }
    switch ((nv + 15) / 16) {
    case 1: k_sell_rspmm<1><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv); break;
    case 2: k_sell_rspmm<2><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv); break;
    case 3: k_sell_rspmm<3><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv); break;
    case 4: k_sell_rspmm<4><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv); break;
    case 5: k_sell_rspmm<5><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv); break;
    default: k_sell_rspmm<6><<<g, b, 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv);
    }
}

__global__ void k_rresid(const double *__restrict__ AX, int64_t ldax, const double *__restrict__ X, int64_t ldx,
                         const double *__restrict__ lam, int m, int64_t n, double *__restrict__ R, int64_t ldr)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * m) return;
    const int64_t r = e / m;
    const int c = (int)(e - r * m);
    R[r * ldr + c] = AX[r * ldax + c] - lam[c] * X[r * ldx + c];
}

void rresid(const double *AX, int64_t ldax, const double *X, int64_t ldx, const double *lam, int m, int64_t n,
            double *R, int64_t ldr, hipStream_t st)
{
    const int64_t tot = n * m;
    k_rresid<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st>>>(AX, ldax, X, ldx, lam, m, n, R, ldr);
}

__global__ void k_rcopy(const double *__restrict__ X, int64_t ldx, int m, int64_t n, double *__restrict__ Y,
                        int64_t ldy)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * m) return;
    const int64_t r = e / m;
    const int c = (int)(e - r * m);
    Y[r * ldy + c] = X[r * ldx + c];
}

// R = AX - X diag(lam) (m <= 32 columns) and, for the first k columns, the sums
// of squares: per block in a fixed grid-stride order, then one fixed-order sum
constexpr int RN_BLOCKS = 1024;
__global__ void __launch_bounds__(256) k_rresid_norm(const double *__restrict__ AX, int64_t ldax,
                                                     const double *__restrict__ X, int64_t ldx,
                                                     const double *__restrict__ lam, int m, int k, int64_t n,
                                                     double *__restrict__ R, int64_t ldr, double *__restrict__ part)
{
    __shared__ double red[8][33];
    const int c = threadIdx.x & 31, rl = threadIdx.x >> 5;
    const double lc = c < m ? lam[c] : 0.0;
    double acc = 0.0;
    for (int64_t r = (int64_t)blockIdx.x * 8 + rl; r < n; r += (int64_t)RN_BLOCKS * 8) {
        if (c < m) {
            const double v = AX[r * ldax + c] - lc * X[r * ldx + c];
            R[r * ldr + c] = v;
            acc = fma(v, v, acc);
        }
    }
    red[rl][c] = c < k ? acc : 0.0;
    __syncthreads();
    if (threadIdx.x < 32) {
        double t = 0.0;
        for (int i = 0; i < 8; i++) t += red[i][threadIdx.x];
        part[(int64_t)blockIdx.x * 32 + threadIdx.x] = t;
    }
}

__global__ void k_norm_reduce(const double *__restrict__ part, int k, double *__restrict__ out)
{
    const int c = threadIdx.x;
    if (c >= k) return;
    double t = 0.0;
    for (int b = 0; b < RN_BLOCKS; b++) t += part[(int64_t)b * 32 + c];
    out[c] = t;
}

void rresid_norms(const double *AX, int64_t ldax, const double *X, int64_t ldx, const double *lam, int m, int k,
                  int64_t n, double *R, int64_t ldr, double *part, double *out, hipStream_t st)
{
    k_rresid_norm<<<dim3(RN_BLOCKS), dim3(256), 0, st>>>(AX, ldax, X, ldx, lam, m, k, n, R, ldr, part);
    k_norm_reduce<<<dim3(1), dim3(32), 0, st>>>(part, k, out);
}

int64_t rresid_norms_part() { return (int64_t)RN_BLOCKS * 32; }

void rcopy(const double *X, int64_t ldx, int m, int64_t n, double *Y, int64_t ldy, hipStream_t st)
{
    const int64_t tot = n * m;
    if (tot > 0) k_rcopy<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st>>>(X, ldx, m, n, Y, ldy);
}

// transposes through a 64-row LDS tile: row-major (ld) <-> column-major (ld n)
template <class S, class D, bool TO_COLS>
__global__ void __launch_bounds__(256) k_rtrans(const S *__restrict__ in, int64_t ldr, int m, int64_t n,
                                                D *__restrict__ out)
{
    __shared__ double T[64][97];   // odd pitch: the rows of a wave hit distinct bank pairs
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    const int nr = (int)min((int64_t)64, n - r0);
    if (TO_COLS) {
        for (int e = threadIdx.x; e < 64 * m; e += 256) {
            const int r = e / m, c = e - r * m;
            if (r < nr) T[r][c] = (double)in[(r0 + r) * ldr + c];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * m; e += 256) {
            const int r = e & 63, c = e >> 6;
            if (r < nr) out[(int64_t)c * n + r0 + r] = (D)T[r][c];
        }
    } else {
        for (int e = threadIdx.x; e < 64 * m; e += 256) {
            const int r = e & 63, c = e >> 6;
            if (r < nr) T[r][c] = (double)in[(int64_t)c * n + r0 + r];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * m; e += 256) {
            const int r = e / m, c = e - r * m;
            if (r < nr) out[(r0 + r) * ldr + c] = (D)T[r][c];
        }
    }
}

void rto_cols(const double *X, int64_t ldx, int m, int64_t n, double *Y, hipStream_t st)
{
    k_rtrans<double, double, true><<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st>>>(X, ldx, m, n, Y);
}

void rfrom_cols(const double *Y, int m, int64_t n, double *X, int64_t ldx, hipStream_t st)
{
    k_rtrans<double, double, false><<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st>>>(Y, ldx, m, n, X);
}

__global__ void k_rfrom_f32rows(const float *__restrict__ X, int64_t ldx, int m, int64_t n, double *__restrict__ Y,
                                int64_t ldy)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * m) return;
    const int64_t r = e / m;
    const int c = (int)(e - r * m);
    Y[r * ldy + c] = (double)X[r * ldx + c];
}

void rfrom_f32rows(const float *X, int64_t ldx, int m, int64_t n, double *Y, int64_t ldy, hipStream_t st)
{
    const int64_t tot = n * m;
    if (tot > 0) k_rfrom_f32rows<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st>>>(X, ldx, m, n, Y, ldy);
}

// columns [c0, c1) of X: uniform [-0.5, 0.5) from splitmix64 of
// (seed, column, row) -- a counter-based stream, the same on every run
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void k_rrandom(double *__restrict__ X, int64_t ldx, int c0, int c1, int64_t n, uint64_t seed)
{
    const int w = c1 - c0;
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * w) return;
    const int64_t r = e / w;
    const int c = c0 + (int)(e - r * w);
    const uint64_t z = splitmix64(seed ^ splitmix64(((uint64_t)c << 40) ^ (uint64_t)r));
    X[r * ldx + c] = (double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

void rrandom(double *X, int64_t ldx, int c0, int c1, int64_t n, uint64_t seed, hipStream_t st)
{
    const int64_t tot = n * (c1 - c0);
    if (tot > 0) k_rrandom<<<dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st>>>(X, ldx, c0, c1, n, seed);
}

}  // namespace sehip
