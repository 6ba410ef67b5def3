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
// Each wave owns a contiguous range of rows and all (<= 6 x 6) output tiles of
// 16 x 16; v_mfma_f64_16x16x4 takes its operands straight from memory, 4 rows
// per step: lane l loads A[r + l/16][16 ti + l%16] and B[r + l/16][16 tj + l%16]
// (four 128-byte row pieces per load) and holds G[16 ti + l/16 + 4v][16 tj +
// l%16].  The next step's operands are loaded before the current step's
// products.  One partial per wave, then a fixed-order sum.
constexpr int RG_WAVES = 1024;     // waves over the rows (4 per block)

typedef double rd4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_rgram_partial(const double *__restrict__ A, int64_t lda, int p,
                                                       const double *__restrict__ B, int64_t ldb, int q, int64_t n,
                                                       int64_t rows_per_wave, double *__restrict__ partial)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lr = lane & 15, lk = lane >> 4;
    const int pt = (p + 15) >> 4, qt = (q + 15) >> 4;
    rd4 acc[6][6];
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) acc[a][b] = rd4{0.0, 0.0, 0.0, 0.0};
    const int64_t r0 = wid * rows_per_wave, r1 = min(n, r0 + rows_per_wave);
    double xa[6], xb[6];
    auto load = [&](int64_t r) {
        const int64_t row = r + lk;
        const bool in = row < r1;
#pragma unroll
        for (int a = 0; a < 6; a++) {
            const int c = 16 * a + lr;
            xa[a] = (a < pt && in && c < p) ? A[row * lda + c] : 0.0;
            xb[a] = (a < qt && in && c < q) ? B[row * ldb + c] : 0.0;
        }
    };
    if (r0 < r1) load(r0);
    for (int64_t r = r0; r < r1; r += 4) {
        double ca[6], cb[6];
#pragma unroll
        for (int a = 0; a < 6; a++) { ca[a] = xa[a]; cb[a] = xb[a]; }
        if (r + 4 < r1) load(r + 4);
#pragma unroll
        for (int a = 0; a < 6; a++)
#pragma unroll
            for (int b = 0; b < 6; b++)
                if (a < pt && b < qt) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ca[a], cb[b], acc[a][b], 0, 0, 0);
    }
    double *out = partial + wid * p * q;
#pragma unroll
    for (int a = 0; a < 6; a++)
#pragma unroll
        for (int b = 0; b < 6; b++) {
            if (a >= pt || b >= qt) continue;
            const int j = 16 * b + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int i = 16 * a + lk + 4 * v;
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

// partial slots (waves) of one Gram: a multiple of 4 waves covering n rows
int64_t rgram_blocks(int64_t n)
{
    const int64_t w = (n + rgram_rpw(n) - 1) / rgram_rpw(n);
    return (w + 3) / 4 * 4;
}

void rgram(const double *A, int64_t lda, int p, const double *B, int64_t ldb, int q, int64_t n, double *partial,
           double *G, hipStream_t st)
{
    const int64_t rpw = rgram_rpw(n), nw = rgram_blocks(n);
    k_rgram_partial<<<dim3((unsigned)(nw / 4)), dim3(256), 0, st>>>(A, lda, p, B, ldb, q, n, rpw, partial);
    k_rgram_reduce<<<dim3((unsigned)((p * q + 63) / 64)), dim3(256), 0, st>>>(partial, nw, p * q, G);
}

// ------------------------------------------------------------------ tall GEMM
// X = V C + beta X (V n x m, C m x p column-major, m, p <= 96; X must not alias
// V) on the FP64 matrix cores.  C is staged in LDS once per block; each wave
// owns 16 rows and all (<= 6) 16-column output tiles, loads its whole 16 x m
// slice of V up front (all loads in flight), then runs the steps of 4 inner
// indices: lane l supplies V[r + l%16][k + l/16] and C[k + l/16][16 tj + l%16]
// and holds X[r + l/16 + 4v][16 tj + l%16] (stores of four 128-byte row pieces).
constexpr int RM_CP = 100;   // LDS pitch of C rows

constexpr int RM_GROUPS = 4;   // 16-row groups per wave

__global__ void __launch_bounds__(256) k_rgemm(const double *__restrict__ V, int64_t ldv, int m,
                                               const double *__restrict__ C, int p, double *__restrict__ X,
                                               int64_t ldx, int64_t n, double beta)
{
    __shared__ double Cs[96][RM_CP];   // Cs[k][c] = C[k + c m]
    const int mp = m * p;
    for (int e = threadIdx.x; e < mp; e += 256) {
        const int c = e / m, k = e - c * m;
        Cs[k][c] = C[e];
    }
    // zero the pad of the last step of 4 and of the last 16-column tile
    const int m4 = (m + 3) & ~3, p16 = (p + 15) & ~15;
    for (int e = threadIdx.x; e < m4 * p16; e += 256) {
        const int k = e / p16, c = e - k * p16;
        if (k >= m || c >= p) Cs[k][c] = 0.0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int pt = (p + 15) >> 4;
    for (int g = 0; g < RM_GROUPS; g++) {
        const int64_t r0 = (((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * RM_GROUPS + g) * 16;
        if (r0 >= n) break;
        const int64_t arow = r0 + lr;
        const bool rin = arow < n;
        double va[24];
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            const int kk = 4 * s2 + lk;
            va[s2] = (rin && kk < m) ? V[arow * ldv + kk] : 0.0;
        }
        rd4 acc[6];
#pragma unroll
        for (int b = 0; b < 6; b++) acc[b] = rd4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s2 = 0; s2 < 24; s2++) {
            if (4 * s2 >= m) break;
#pragma unroll
            for (int b = 0; b < 6; b++)
                if (b < pt)
                    acc[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(va[s2], Cs[4 * s2 + lk][16 * b + lr], acc[b], 0, 0, 0);
        }
#pragma unroll
        for (int b = 0; b < 6; b++) {
            if (b >= pt) continue;
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

void rgemm(const double *V, int64_t ldv, int m, const double *C, int p, double *X, int64_t ldx, int64_t n,
           double beta, hipStream_t st)
{
    const int64_t rows = 64 * RM_GROUPS;
    k_rgemm<<<dim3((unsigned)((n + rows - 1) / rows)), dim3(256), 0, st>>>(V, ldv, m, C, p, X, ldx, n, beta);
}

// ------------------------------------------------------------------ SpMM
// Y = A X for row-major X, Y (nv <= 96 columns): SELL-64 slices; a 16-lane
// group per row, lane j of the group on column t0 + j, so each nonzero gathers
// one contiguous 128-byte piece of row col of X.  A block = 16 rows, grid.y
// over 16-column chunks.
__global__ void __launch_bounds__(256) k_sell_rspmm(int64_t nslices, const int64_t *__restrict__ sptr,
                                                    const int32_t *__restrict__ swid, const int32_t *__restrict__ col,
                                                    const double *__restrict__ val, const double *__restrict__ X,
                                                    int64_t ldx, double *__restrict__ Y, int64_t ldy, int64_t n,
                                                    int nv)
{
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int64_t sl = row >> 6;
    if (sl >= nslices || row >= n) return;
    const int t = blockIdx.y * 16 + (threadIdx.x & 15);
    const bool on = t < nv;
    int64_t pp = sptr[sl] + (row & 63);
    const int w = swid[sl];
    double s = 0.0;
    for (int k = 0; k < w; k++, pp += 64) {
        const double a = val[pp];
        const int64_t c = col[pp];
        if (on) s = fma(a, X[c * ldx + t], s);
    }
    if (on) Y[row * ldy + t] = s;
}

void rspmm(const SellA &S, const double *X, int64_t ldx, double *Y, int64_t ldy, int nv, hipStream_t st)
{
    k_sell_rspmm<<<dim3((unsigned)((S.n + 15) / 16), (unsigned)((nv + 15) / 16)), dim3(256), 0, st>>>(
        S.nslices, S.sptr, S.swid, S.col, S.val, X, ldx, Y, ldy, S.n, nv);
}

// ------------------------------------------------------------------ elementwise
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
