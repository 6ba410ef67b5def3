// factor.hip -- numeric multifrontal Cholesky and triangular solves on gfx950.
//
// Replaces cholmod_l_factorize_p and cholmod_l_solve2 (reference
// src/speigs_core.c:938-970 and :86-100).  Data layout (HBM):
//   pool   : every front s is a dense F_s x F_s column-major block (FP64 or
//            FP32) at foff[s]; after factorization its first c_s columns hold
//            the factor panel [L11; L21] and its trailing r_s x r_s block the
//            update (Schur complement) passed to the parent by extend-add.
//   ws     : per front a vector of F_s doubles used by the solves (the
//            forward solve leaves the child's update vector in its tail).
// Factorization runs level by level over the supernodal tree (leaves first).
// Per level: extend-add of the children's update blocks (one launch per child
// ordinal, so no two blocks of a launch write the same front), then per
// 64-column panel k: POTRF of the diagonal tile, TRSM of the tiles below it,
// and the Schur update of the trailing tiles -- each a batched launch over a
// task list that spans every front of the level.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>

#include "factor_impl.h"
#include "mfma.h"
#include "host_par.h"

namespace sehip {


// ------------------------------------------------------------------ kernels
// p[0 .. n) = 0 with 16-byte stores over a grid stride (p 16-byte aligned);
// one element per thread ran at ~0.7 TB/s on the 2.3 GB FP32 pool
template <class T>
__global__ void __launch_bounds__(256) k_zero(T *p, int64_t n)
{
    const int64_t n16 = n * (int64_t)sizeof(T) / 16;
    float4 *q = reinterpret_cast<float4 *>(p);
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += stride)
        q[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    const int64_t i = n16 * 16 / (int64_t)sizeof(T) + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = T(0);
}

// ---- assembly of A into the fronts, from A's own CSC arrays on the device.
// An entry (i, j), i <= j, of the stored pattern lands in front s = snode[col]
// of its permuted column col = min(iperm i, iperm j), at the front row of
// row = max(...): row - c0[s] inside the supernode, else c + the position of
// row in the supernode's sorted row list (binary search).

// snode[j] = the supernode of permuted column j (a wave per supernode: the
// root's thousands of columns are not left to one thread)
__global__ void __launch_bounds__(256) k_snode_fill(const int64_t *__restrict__ col0,
                                                    const int32_t *__restrict__ ncol, int64_t ns,
                                                    int32_t *__restrict__ snode)
{
    const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (s >= ns) return;
    const int64_t c0 = col0[s];
    for (int j = threadIdx.x & 63; j < ncol[s]; j += 64) snode[c0 + j] = (int32_t)s;
}

// diagA[iperm j] = A(j, j) for the stored diagonal entries (diagA zeroed before)
__global__ void __launch_bounds__(256) k_diag_values(const int64_t *__restrict__ Ap, const int64_t *__restrict__ Ai,
                                                     const double *__restrict__ Ax,
                                                     const int64_t *__restrict__ iperm, int64_t n,
                                                     double *__restrict__ diagA)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
        if (Ai[q] == j) diagA[iperm[j]] = Ax[q];
}

// every diagonal of the pool = delta: the value of a diagonal A does not store
template <class T>
__global__ void __launch_bounds__(256) k_diag_delta(T *pool, const int32_t *__restrict__ snode,
                                                    const int64_t *__restrict__ col0, const int32_t *__restrict__ Fs,
                                                    const int64_t *__restrict__ foff, int64_t n, double delta)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int32_t s = snode[j];
    const int64_t l = j - col0[s];
    pool[foff[s] + l + l * Fs[s]] = (T)(0.0 + delta);
}

// pool entry of (i, j) = A value (+ delta on diagonals), one thread per stored
// column; an entry outside the symbolic structure sets bit 2 of flag[0] and
// the first one found leaves its (i, j) in flag[2], flag[3]
template <class T>
__global__ void __launch_bounds__(256) k_assemble(T *pool, const int64_t *__restrict__ Ap,
                                                  const int64_t *__restrict__ Ai, const double *__restrict__ Ax,
                                                  const int64_t *__restrict__ iperm, const int32_t *__restrict__ snode,
                                                  const int64_t *__restrict__ col0, const int32_t *__restrict__ ncol,
                                                  const int32_t *__restrict__ Fs, const int64_t *__restrict__ rptr,
                                                  const int32_t *__restrict__ rows, const int64_t *__restrict__ foff,
                                                  int64_t n, double delta, int *__restrict__ flag)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const int64_t nj = iperm[j];
    for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
        const int64_t i = Ai[q];
        if (i > j) continue;
        const int64_t ni = iperm[i];
        const int64_t col = min(ni, nj), row = max(ni, nj);
        const int32_t s = snode[col];
        const int64_t c0 = col0[s];
        const int c = ncol[s];
        int64_t lr;
        if (row < c0 + c) {
            lr = row - c0;
        } else {
            int64_t lo = rptr[s], hi = rptr[s + 1];
            const int64_t end = hi;
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (rows[mid] < row) lo = mid + 1;
                else hi = mid;
            }
            if (lo == end || rows[lo] != row) {
                // the first such entry is kept for the error report
                if ((atomicOr(flag, 2) & 2) == 0) {
                    flag[2] = (int)i;
                    flag[3] = (int)j;
                }
                continue;
            }
            lr = c + (lo - rptr[s]);
        }
        const double v = i == j ? Ax[q] + delta : Ax[q];
        pool[foff[s] + lr + (col - c0) * Fs[s]] = (T)v;
    }
}



template <class T>
__global__ void __launch_bounds__(256) k_extend_add(T *pool, const EaTask *__restrict__ tasks,
                                                    const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                    const int32_t *__restrict__ ncol,
                                                    const int64_t *__restrict__ relptr,
                                                    const int32_t *__restrict__ rel)
{
    const EaTask t = tasks[blockIdx.x];
    const int32_t Fc = Fs[t.child], cc = ncol[t.child], rc = Fc - cc, Fp = Fs[t.parent];
    const T *U = pool + foff[t.child] + (int64_t)cc * Fc + cc;     // r x r, ld Fc
    T *P = pool + foff[t.parent];
    const int32_t *rl = rel + relptr[t.child];
    const int j0 = t.cb * 64, j1 = min(rc, j0 + 64);
    for (int j = j0; j < j1; j++) {
        const int64_t pc = (int64_t)rl[j] * Fp;
        for (int i = j + threadIdx.x; i < rc; i += 256) P[pc + rl[i]] += U[(int64_t)j * Fc + i];
    }
}



__device__ __forceinline__ int lower_bound_i32(const int32_t *a, int n, int v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Block boundaries of the children's update rows in their parents: for child
// ch (rc = its update rows, rl = rel + relptr[ch] increasing parent rows) and k
// = 0 .. ceil(F_parent / 64), bnd[bndptr[ch] + k] = the first j with rl[j] >=
// 64 k (rc if none) -- the lower bounds the extend-add needs for its 64-column
// and 256-row blocks, written once per factor by one thread per rel entry.
__global__ void __launch_bounds__(256) k_ea_bounds(const int64_t *__restrict__ relptr, const int32_t *__restrict__ rel,
                                                   const int64_t *__restrict__ bndptr, int64_t ns, int32_t *__restrict__ bnd)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= relptr[ns]) return;
    int64_t lo = 0, hi = ns;                  // the child whose list holds e: relptr[ch] <= e < relptr[ch + 1]
    while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (relptr[mid] <= e) lo = mid; else hi = mid;
    }
    const int64_t ch = lo, j = e - relptr[ch], rc = relptr[ch + 1] - relptr[ch];
    const int64_t nb = bndptr[ch + 1] - bndptr[ch];
    int32_t *bd = bnd + bndptr[ch];
    const int kcur = rel[e] >> 6, kprev = j == 0 ? -1 : rel[e - 1] >> 6;
    for (int k = kprev + 1; k <= kcur && k < nb; k++) bd[k] = (int32_t)j;
    if (j == rc - 1)
        for (int64_t k = kcur + 1; k < nb; k++) bd[k] = (int32_t)rc;
}

__device__ __forceinline__ int32_t lane_i32(int32_t v, int q) { return __builtin_amdgcn_readlane(v, q); }
__device__ __forceinline__ int64_t lane_i64(int64_t v, int q)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, q);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(v >> 32), q);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// extend-add into the parent block (columns [64 cb, +64), rows [256 rb, +256)):
// each wave takes one 64-row block of it and applies the children in order,
// so every element's sum is deterministic.  The lanes first read the
// children's sizes, offsets and block boundaries together (lane q for the
// q-th child of a run of 64; the rel maps are increasing, so a child
// contributes a contiguous range of its columns and of its rows, from the
// boundaries of k_ea_bounds), then child by child lane t takes child row ilo
// + t, 16 columns at a time with all 16 loads issued before the stores.  Two
// children may add into the same parent element from different lanes: the
// wave's stores complete before the next child's loads.
template <class T>
__global__ void __launch_bounds__(256) k_extend_add_blk(T *pool, const EaBlk *__restrict__ tasks,
                                                        const int64_t *__restrict__ foff,
                                                        const int32_t *__restrict__ Fs, const int32_t *__restrict__ ncol,
                                                        const int64_t *__restrict__ relptr,
                                                        const int32_t *__restrict__ rel,
                                                        const int64_t *__restrict__ chptr,
                                                        const int32_t *__restrict__ chl,
                                                        const int64_t *__restrict__ bndptr,
                                                        const int32_t *__restrict__ bnd)
{
    constexpr int NB = 16;
    const EaBlk t = tasks[blockIdx.x];
    const int lane = threadIdx.x & 63;
    const int r64 = 4 * t.rb + (int)(threadIdx.x >> 6);      // the wave's 64-row block of the parent
    const int32_t Fp = Fs[t.parent];
    // rows above the column block get nothing (a child's U is lower triangular)
    if (r64 < t.cb || 64 * r64 >= Fp) return;
    T *P = pool + foff[t.parent];
    const int64_t q0 = chptr[t.parent], q1 = chptr[t.parent + 1];
    for (int64_t qb = q0; qb < q1; qb += 64) {
        const int nq = (int)min((int64_t)64, q1 - qb);
        const int ch = chl[qb + min(lane, nq - 1)];
        const int32_t Fc = Fs[ch], cc = ncol[ch];
        const int64_t fo = foff[ch], ro = relptr[ch];
        // every child has ceil(Fp / 64) + 1 boundaries and cb, r64 < ceil(Fp / 64);
        // none is written for a child with no update rows
        const int32_t *bd = bnd + bndptr[ch];
        const int32_t b0 = bd[t.cb], b1 = bd[t.cb + 1], b2 = bd[r64], b3 = bd[r64 + 1];
        const bool any = Fc > cc;
        const int32_t jlo_l = any ? b0 : 0, jhi_l = any ? b1 : 0, ilo_l = any ? b2 : 0, ihi_l = any ? b3 : 0;
        for (int q = 0; q < nq; q++) {
            // (read out of lane q here, where every lane is active)
            const int jlo = lane_i32(jlo_l, q), jhi = lane_i32(jhi_l, q);
            const int ilo = lane_i32(ilo_l, q), ihi = lane_i32(ihi_l, q);
            const int32_t Fcq = lane_i32(Fc, q), ccq = lane_i32(cc, q);
            const int64_t roq = lane_i64(ro, q), foq = lane_i64(fo, q);
            const int i = ilo + lane;
            if (jlo < jhi && i < ihi) {
                const int32_t *rl = rel + roq;
                const T *U = pool + foq + (int64_t)ccq * Fcq + ccq;     // r x r, ld Fc
                const int64_t prow = rl[i];
                const int jend = min(jhi, i + 1);
                for (int j = jlo; j < jend; j += NB) {
                    // every load unconditional at an index clamped into [j, jend)
                    // (a select after each load, or a use between loads, makes
                    // each wait for the one before); only the stores are predicated
                    T u[NB], pv[NB];
                    int64_t pos[NB];
                    int32_t rj[NB];
#pragma unroll
                    for (int b = 0; b < NB; b++) rj[b] = rl[min(j + b, jend - 1)];
#pragma unroll
                    for (int b = 0; b < NB; b++) {
                        const int jj = min(j + b, jend - 1);
                        pos[b] = (int64_t)rj[b] * Fp + prow;
                        u[b] = U[(int64_t)jj * Fcq + i];
                        pv[b] = P[pos[b]];
                    }
#pragma unroll
                    for (int b = 0; b < NB; b++)
                        if (j + b < jend) P[pos[b]] = pv[b] + u[b];
                }
            }
            // the stores done before the next child's loads (other lanes may
            // read these elements); the fences also keep the compiler from
            // moving memory operations across
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
    }
}



// diagonal tile of panel k: in-LDS unblocked Cholesky with the pivot test
template <class T>
__global__ void __launch_bounds__(256) k_potrf(T *pool, const TileTask *__restrict__ tasks,
                                               const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                               const int32_t *__restrict__ ncol, const int64_t *__restrict__ col0,
                                               const double *__restrict__ diagA, double delta, double piv_rel,
                                               int *flag)
{
    __shared__ T a[TB][TB + 1];
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0);
    T *base = pool + foff[t.s] + (int64_t)k0 * F + k0;
    for (int e = threadIdx.x; e < w * w; e += 256) {
        const int r = e % w, q = e / w;
        a[r][q] = (r >= q) ? base[(int64_t)q * F + r] : T(0);
    }
    __syncthreads();
    for (int p = 0; p < w; p++) {
        if (threadIdx.x == 0) {
            T d = a[p][p];
            const double ref = fabs(diagA[col0[t.s] + k0 + p] + delta);
            if (!((double)d > piv_rel * ref) || !(d == d)) {
                atomicOr(flag, 1);
                d = T(1);
            }
            a[p][p] = sqrt(d);
        }
        __syncthreads();
        const T dp = a[p][p];
        for (int r = p + 1 + threadIdx.x; r < w; r += 256) a[r][p] /= dp;
        __syncthreads();
        const int m = w - p - 1;
        for (int e = threadIdx.x; e < m * m; e += 256) {
            const int r = p + 1 + e % m, q = p + 1 + e / m;
            if (r >= q) a[r][q] -= a[r][p] * a[q][p];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < w * w; e += 256) {
        const int r = e % w, q = e / w;
        if (r >= q) base[(int64_t)q * F + r] = a[r][q];
    }
}

// rows of tile row i below the diagonal block: B <- B L_kk^-T
template <class T>
__global__ void __launch_bounds__(256) k_trsm(T *pool, const TileTask *__restrict__ tasks,
                                              const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                              const int32_t *__restrict__ ncol)
{
    __shared__ T L[TB][TB + 1];
    __shared__ T b[TB][TB + 1];
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0);
    const int r0 = max(t.i * TB, k0 + w), r1 = min(F, (t.i + 1) * TB), nr = r1 - r0;
    if (nr <= 0) return;
    T *fr = pool + foff[t.s];
    for (int e = threadIdx.x; e < w * w; e += 256) {
        const int r = e % w, q = e / w;
        L[r][q] = fr[(int64_t)(k0 + q) * F + k0 + r];
    }
    for (int e = threadIdx.x; e < nr * w; e += 256) {
        const int r = e % nr, q = e / nr;
        b[r][q] = fr[(int64_t)(k0 + q) * F + r0 + r];
    }
    __syncthreads();
    const int row = threadIdx.x & 63, grp = threadIdx.x >> 6;
    for (int p = 0; p < w; p++) {
        if (grp == (p & 3) && row < nr) b[row][p] /= L[p][p];
        __syncthreads();
        if (row < nr) {
            const T xp = b[row][p];
            for (int q = p + 1 + ((grp - (p + 1)) & 3); q < w; q += 4) b[row][q] -= xp * L[q][p];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nr * w; e += 256) {
        const int r = e % nr, q = e / nr;
        fr[(int64_t)(k0 + q) * F + r0 + r] = b[r][q];
    }
}

// Schur updates are outer-blocked: blocks of KB panels (256 columns).  After
// panel k the columns left in its block, [kend, bend), are updated at once with
// panel k alone (the "inner" update, so the next panels of the block see their
// final values); everything right of the block, [bend, F), is updated once per
// block with all of its panels, K = bend - 256 kb <= 256 (the "outer" update,
// after the block's last panel).  The trailing matrix's C tiles are then read
// and written once per 256 columns instead of once per panel.
//
// Update tile gi of a panel step, decoded instead of listed: ufid[gi] is the
// front's ordinal f in the step's POTRF list pt, upref[f] its first tile; a
// front's tiles run over column tiles j >= j0 and, within a column, rows
// i >= j.  The result is the tile's K range and its clipped rows / columns.
// outer: 0 inner update; 1 the outer update's first min(KB, m) tile columns
// (the next block's panels); 2 the rest of the outer update.
struct UpdTile {
    int32_t s;
    int k0, kw, r0, r1, q0, q1;
};

__device__ __forceinline__ UpdTile upd_decode(const TileTask *__restrict__ pt, const int32_t *__restrict__ ufid,
                                              const int64_t *__restrict__ upref, const int32_t *__restrict__ Fs,
                                              const int32_t *__restrict__ ncol, int64_t gi, int outer)
{
    const int f = ufid[gi];
    const TileTask t = pt[f];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int kend = t.k * TB + min(TB, c - t.k * TB);
    const int kb = t.k / KB, bend = min((kb + 1) * KB * TB, c);
    const int nt = (F + TB - 1) / TB;
    int lo, hi, j0;
    UpdTile u;
    u.s = t.s;
    if (outer) {
        lo = bend;
        hi = F;
        j0 = bend / TB;
        u.k0 = kb * KB * TB;
        u.kw = bend - u.k0;
    } else {
        lo = kend;
        hi = bend;
        j0 = kend < min(F, (t.k + 1) * TB) ? t.k : t.k + 1;
        u.k0 = t.k * TB;
        u.kw = kend - u.k0;
    }
    const int m = nt - j0;
    int64_t rem = gi - upref[f];
    if (outer == 2) {
        // the far part: after the first min(KB, m) tile columns of the outer range
        const int64_t kk = min(KB, m);
        rem += kk * m - kk * (kk - 1) / 2;
    }
    int tt = 0;
    while (rem >= m - tt) {
        rem -= m - tt;
        tt++;
    }
    const int j = j0 + tt, i = j + (int)rem;
    u.r0 = max(i * TB, lo);
    u.r1 = min(F, (i + 1) * TB);
    u.q0 = max(j * TB, lo);
    u.q1 = min(hi, (j + 1) * TB);
    return u;
}

// Schur update of a C tile: C -= L_i L_j^T over K = [k0, k0 + kw) (vector FMA,
// 4 x 4 per thread, 64-column slabs through LDS)
template <class T>
__global__ void __launch_bounds__(256) k_update(T *pool, const TileTask *__restrict__ pt,
                                                const int32_t *__restrict__ ufid, const int64_t *__restrict__ upref,
                                                int outer, const int64_t *__restrict__ foff,
                                                const int32_t *__restrict__ Fs, const int32_t *__restrict__ ncol)
{
    __shared__ T Ai[TB][TB + 4];   // [kk][row]
    __shared__ T Aj[TB][TB + 4];
    const UpdTile u = upd_decode(pt, ufid, upref, Fs, ncol, blockIdx.x, outer);
    const int32_t F = Fs[u.s];
    const int r0 = u.r0, q0 = u.q0, nr = u.r1 - u.r0, nq = u.q1 - u.q0;
    if (nr <= 0 || nq <= 0) return;
    T *fr = pool + foff[u.s];
    const int ri = 4 * (threadIdx.x & 15), qj = 4 * (threadIdx.x >> 4);
    T acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = T(0);
    for (int sb = 0; sb < u.kw; sb += TB) {
        const int w = min(TB, u.kw - sb), kc = u.k0 + sb;
        __syncthreads();
        for (int e = threadIdx.x; e < TB * w; e += 256) {
            const int r = e % TB, kk = e / TB;
            Ai[kk][r] = r < nr ? fr[(int64_t)(kc + kk) * F + r0 + r] : T(0);
            Aj[kk][r] = r < nq ? fr[(int64_t)(kc + kk) * F + q0 + r] : T(0);
        }
        __syncthreads();
        for (int kk = 0; kk < w; kk++) {
            T x[4], y[4];
#pragma unroll
            for (int a = 0; a < 4; a++) { x[a] = Ai[kk][ri + a]; y[a] = Aj[kk][qj + a]; }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++) acc[a][b] = fma(x[a], y[b], acc[a][b]);
        }
    }
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int q = qj + b;
        if (q >= nq) continue;
        T *col = fr + (int64_t)(q0 + q) * F + r0;
#pragma unroll
        for (int a = 0; a < 4; a++)
            if (ri + a < nr) col[ri + a] -= acc[a][b];
    }
}

// ---- panel factorization on one wave -----------------------------------------
// The diagonal tile of panel k lives in registers: lane r holds row r (a[q] =
// L(r, q) for q <= r; entries above the diagonal are scratch that never feeds
// the lower part).  Column p: the pivot comes from lane p by readlane, the
// column is scaled, and each later column q is updated with L(q, p) broadcast
// from lane q.  The same broadcasts drive the inverse: lane r carries column r
// of X = L_kk^-1 as a right-looking forward substitution (t <- e_r;
// x_p = t_p / L(p, p); t_q -= L(q, p) x_p), so the inverse costs FMAs only.
// X is stored column-major to the panel's scratch slot for the TRSM product.
__device__ __forceinline__ float bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double bcast(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <class T>
__global__ void __launch_bounds__(64) k_potrf_reg(T *pool, const TileTask *__restrict__ tasks,
                                                  const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                  const int32_t *__restrict__ ncol, const int64_t *__restrict__ col0,
                                                  const double *__restrict__ diagA, double delta, double piv_rel,
                                                  int *flag, T *__restrict__ xinv)
{
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0);
    const int r = threadIdx.x;
    T *base = pool + foff[t.s] + (int64_t)k0 * F + k0;
    // rows beyond w are the identity (their columns stay e_r, their X rows too)
    T a[TB], x[TB];
#pragma unroll
    for (int q = 0; q < TB; q++) {
        a[q] = (r < w && q < w && q <= r) ? base[(int64_t)q * F + r] : T(q == r ? 1 : 0);
        x[q] = T(q == r ? 1 : 0);
    }
    bool bad = false;
    // lane r holds A(k0 + r, k0 + r); the pivot test of step p broadcasts it
    // (a load inside the step chain would put a memory latency on every step)
    const double dAr = r < w ? diagA[col0[t.s] + k0 + r] : 0.0;
#pragma unroll
    for (int p = 0; p < TB; p++) {
        if (p < w) {
            T d = bcast(a[p], p);
            const double ref = fabs(bcast(dAr, p) + delta);
            if (!((double)d > piv_rel * ref) || !(d == d)) {
                bad = true;
                d = T(1);
            }
            // true divisions (not a reciprocal): the same rounding as a
            // left-looking substitution, which the multiplicity tests rely on
            const T sd = sqrt(d);
            a[p] = (r == p) ? sd : a[p] / sd;
            const T xp = x[p] / sd;
            x[p] = xp;
#pragma unroll
            for (int q = p + 1; q < TB; q++) {
                const T lq = bcast(a[p], q);          // L(q, p)
                a[q] -= a[p] * lq;
                x[q] -= lq * xp;
            }
        }
    }
    if (bad && r == 0) atomicOr(flag, 1);
#pragma unroll
    for (int q = 0; q < TB; q++)
        if (r < w && q < w && q <= r) base[(int64_t)q * F + r] = a[q];
    // column r of X (column-major 64 x 64) is lane r's x; transposed through
    // LDS so that each store instruction writes one contiguous column
    __shared__ T xs[TB][TB + 1];
#pragma unroll
    for (int i = 0; i < TB; i++) xs[i][r] = x[i];
    __syncthreads();
    T *X = xinv + (int64_t)blockIdx.x * (TB * TB);
#pragma unroll
    for (int j = 0; j < TB; j++) X[(int64_t)j * TB + r] = xs[r][j];
}

// The same factorization and inverse on four waves: wave w holds columns
// 16w .. 16w+15 of every row (lane r = row r), so the rank-1 updates of a step
// are spread over the four SIMDs instead of one wave issuing all of them.
// Step p: the wave owning column p forms the pivot and column p of L and of X
// and leaves them in LDS (two buffers, alternating by step, so one barrier per
// step suffices); every wave then reads the step's column with one wait and
// updates its own columns q > p.  Each element sees the same operations in the
// same order as in k_potrf_reg, so the factor and the inverse are the same to
// the bit.
//
// SG: the signed factorization of a symmetric indefinite block, A = L S L^T
// with S = diag(+-1) and L's diagonal positive (the LDL^T factorization with
// L scaled by |D|^1/2): step p's pivot d gives s_p = sign(d), L(p, p) =
// sqrt|d|, L(r, p) = A(r, p) / (L(p, p) s_p), and the rank-1 update carries
// s_p.  A pivot with |d| < tau is replaced by s_p tau (static pivoting: the
// apply refines the solution against A, and flag[1] counts the
// replacements).  X is S L^-1 (rows scaled by the signs), so that the TRSM's
// B X^T is A21 L^-T S, the signed L21; the signs go to sgn[column].
template <class T, bool SG>
__global__ void __launch_bounds__(256) k_potrf_reg4(T *pool, const TileTask *__restrict__ tasks,
                                                    const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                    const int32_t *__restrict__ ncol, const int64_t *__restrict__ col0,
                                                    const double *__restrict__ diagA, double delta, double piv_rel,
                                                    int *flag, T *__restrict__ xinv, T *__restrict__ sgn, double tau)
{
    __shared__ __align__(16) T Lb[2][TB];
    __shared__ T Xb[2][TB];
    __shared__ T xs[TB][TB + 1];
    __shared__ T Sb[TB];                      // SG: the pivots' signs
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0);
    const int r = threadIdx.x & 63, wv = threadIdx.x >> 6, q0 = 16 * wv;
    T *base = pool + foff[t.s] + (int64_t)k0 * F + k0;
    // rows beyond w are the identity (their columns stay e_r, their X rows too)
    T a[16], x[16];
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int q = q0 + j;
        a[j] = (r < w && q < w && q <= r) ? base[(int64_t)q * F + r] : T(q == r ? 1 : 0);
        x[j] = T(q == r ? 1 : 0);
    }
    bool bad = false;
    int npert = 0;                            // SG: replaced pivots (owner waves' lane 0)
    const double dAr = r < w ? diagA[col0[t.s] + k0 + r] : 0.0;
#pragma unroll
    for (int p = 0; p < TB; p++) {
        if (p >= w) continue;                     // w is the same for the whole block
        const int jp = p & 15, b = p & 1;
        if (wv == (p >> 4)) {
            T d = bcast(a[jp], p);
            if (SG) {
                if (!(d == d)) {
                    bad = true;
                    d = T(1);
                }
                const T sp = d < T(0) ? T(-1) : T(1);
                if (!(fabs((double)d) >= tau)) {
                    d = sp * T(tau);
                    npert++;
                }
                const T sd = sqrt(d * sp);
                a[jp] = (r == p) ? sd : a[jp] / (sd * sp);
                const T xp = x[jp] / sd;
                x[jp] = xp;
                Lb[b][r] = a[jp];
                Xb[b][r] = xp;
                if (r == 0) Sb[p] = sp;
            } else {
                const double ref = fabs(bcast(dAr, p) + delta);
                if (!((double)d > piv_rel * ref) || !(d == d)) {
                    bad = true;
                    d = T(1);
                }
                // true divisions (not a reciprocal), as in k_potrf_reg
                const T sd = sqrt(d);
                a[jp] = (r == p) ? sd : a[jp] / sd;
                const T xp = x[jp] / sd;
                x[jp] = xp;
                Lb[b][r] = a[jp];
                Xb[b][r] = xp;
            }
        }
        __syncthreads();
        if (q0 + 15 > p) {
            const T lr = Lb[b][r], xr = Xb[b][r];
            const T lrs = SG ? lr * Sb[p] : lr;
            T lq[16];                                   // L(q, p), the same for every lane
#pragma unroll
            for (int j = 0; j < 16; j++) lq[j] = Lb[b][q0 + j];
#pragma unroll
            for (int j = 0; j < 16; j++) {
                const T an = a[j] - lrs * lq[j], xn = x[j] - lq[j] * xr;
                const bool upd = q0 + j > p;
                a[j] = upd ? an : a[j];
                x[j] = upd ? xn : x[j];
            }
        }
    }
    if (bad && r == 0) atomicOr(flag, 1);
    if (SG) {
        if (npert && r == 0) atomicAdd(flag + 1, npert);
        // X = S L^-1: the stored X's row q (register column q of every lane)
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (q0 + j < w) x[j] *= Sb[q0 + j];
        if (wv == 0 && r < w) sgn[col0[t.s] + k0 + r] = Sb[r];
    }
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const int q = q0 + j;
        if (r < w && q < w && q <= r) base[(int64_t)q * F + r] = a[j];
    }
    // column r of X (column-major 64 x 64) is row r's x, transposed through LDS
#pragma unroll
    for (int j = 0; j < 16; j++) xs[q0 + j][r] = x[j];
    __syncthreads();
    T *X = xinv + (int64_t)blockIdx.x * (TB * TB);
#pragma unroll
    for (int j = 0; j < 16; j++) X[(int64_t)(q0 + j) * TB + r] = xs[r][q0 + j];
}

// 1 / sqrt(d) for a pivot: the hardware reciprocal square root in FP32 (about
// 1 ulp), a rounded square root and division in FP64
__device__ __forceinline__ float rsqrt_piv(float d) { return __builtin_amdgcn_rsqf(d); }
__device__ __forceinline__ double rsqrt_piv(double d) { return 1.0 / sqrt(d); }

// Cholesky of the 16 x 16 diagonal block at (o, o) of the LDS tile, with its
// inverse, on one wave: lane 16g + r holds row r, columns 4g .. 4g+3 (a) and
// X(4g+j, r) (x, the right-looking forward substitution t <- e_r of the
// unblocked kernels).  Group gp factors its four columns alone (pivots and
// multipliers by readlane inside the group), leaves them with its x in LDS,
// and the groups to its right apply the rank-4 update: four LDS round trips
// per block instead of one per column, each element updated column by column
// in the same order.  thr (lane l) is the pivot threshold of tile row l; a
// failed pivot is replaced by 1 and sets bad.
template <class T>
__device__ __forceinline__ void potrf_diag16(T (*S)[TB + 1], T (*XS)[TB + 1], T (*Cb)[16][8], int o, int w,
                                             double thr, bool &bad)
{
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    T a[4], x[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        a[j] = S[o + 4 * g + j][o + r];
        x[j] = T(4 * g + j == r ? 1 : 0);
    }
#pragma unroll
    for (int gp = 0; gp < 4; gp++) {
        const int b = gp & 1;
        if (g == gp) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int p = 4 * gp + i;
                T d = bcast(a[i], 16 * gp + p);
                if (o + p < w) {
                    const double th = bcast(thr, o + p);
                    if (!((double)d > th) || !(d == d)) {
                        bad = true;
                        d = T(1);
                    }
                }
                const T ri = rsqrt_piv(d);
                a[i] = (r == p) ? d * ri : a[i] * ri;
                x[i] = x[i] * ri;
#pragma unroll
                for (int i2 = i + 1; i2 < 4; i2++) {
                    const T lq = bcast(a[i], 16 * gp + 4 * gp + i2);     // L(4 gp + i2, p)
                    a[i2] -= a[i] * lq;
                    x[i2] -= lq * x[i];
                }
            }
#pragma unroll
            for (int i = 0; i < 4; i++) {
                Cb[b][r][i] = a[i];
                Cb[b][r][4 + i] = x[i];
            }
        }
        if (gp == 3) break;
        // the group's columns are read by the other lanes of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (g > gp) {
            T lrow[4], xrow[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                lrow[i] = Cb[b][r][i];
                xrow[i] = Cb[b][r][4 + i];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                T lq[4];
#pragma unroll
                for (int i = 0; i < 4; i++) lq[i] = Cb[b][4 * g + j][i];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    a[j] -= lrow[i] * lq[i];
                    x[j] -= lq[i] * xrow[i];
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int q = 4 * g + j;
        if (q <= r) S[o + q][o + r] = a[j];
        XS[o + r][o + q] = x[j];
    }
}

// Blocked panel factorization and inverse: the 64 x 64 tile as 4 x 4 blocks of
// 16, held in LDS (S[c][r] = A(r, c), overwritten by L; XS[c][r] = X(r, c)).
// Block column kb: wave 0 factors the 16 x 16 diagonal block with its inverse
// (potrf_diag16); the blocks below it are solved as products with that inverse
// (L_ib = A_ib X_bb^T, one wave per block) and the trailing lower blocks
// updated (S_ij -= L_ib L_jb^T, one wave per block pair), both on the matrix
// cores.  Wave 0 updates the next diagonal block first and factors it while
// waves 1-3 update the rest.  The off-diagonal blocks of X = L^-1 then follow
// one block diagonal at a time by X_ij = -X_ii sum_{m=j}^{i-1} L_im X_mj.  The
// factor agrees with the unblocked kernels' to rounding (a reciprocal square
// root per pivot, products summed by the matrix cores).  Rows and columns
// beyond w are the identity, so their pivots are 1 and their couplings 0; the
// pivot test covers columns < w only, and the blocks wholly beyond w are
// skipped (a leaf front of 16 columns factors one diagonal block).
template <class T>
__global__ void __launch_bounds__(256) k_potrf_blk(T *pool, const TileTask *__restrict__ tasks,
                                                   const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                   const int32_t *__restrict__ ncol, const int64_t *__restrict__ col0,
                                                   const double *__restrict__ diagA, double delta, double piv_rel,
                                                   int *flag, T *__restrict__ xinv)
{
    typedef typename MfmaAcc<T>::V V;
    __shared__ T S[TB][TB + 1];
    __shared__ T XS[TB][TB + 1];
    __shared__ T TS[4][16][17];     // per wave: TS[w][c][r] = T(r, c) of the X recurrence
    __shared__ __align__(16) T Cb[2][16][8];
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0), wl = w - 1;
    // the 16-column blocks holding columns < w; those after them are the
    // identity in L and in X and are not factored
    const int nbk = (w + 15) >> 4;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, lr = lane & 15, lk = lane >> 4;
    T *base = pool + foff[t.s] + (int64_t)k0 * F + k0;
    {
        // every load in flight before the LDS stores (a select after each load
        // would wait for it); entries outside the tile's lower part read a
        // lower entry inside it and are replaced by the identity's
        // (column groups 4e .. 4e+3 at or beyond w are not loaded: e and w are
        // uniform, so that test is a scalar branch)
        T v[TB * TB / 256];
#pragma unroll
        for (int e = 0; e < TB * TB / 256; e++) {
            const int i = tid + 256 * e, cc = min(i >> 6, wl), rr = max(min(i & 63, wl), cc);
            v[e] = 4 * e < w ? base[(int64_t)cc * F + rr] : T(0);
        }
#pragma unroll
        for (int e = 0; e < TB * TB / 256; e++) {
            const int i = tid + 256 * e, cc = i >> 6, rr = i & 63;
            S[cc][rr] = (rr < w && cc < w && cc <= rr) ? v[e] : T(cc == rr ? 1 : 0);
            XS[cc][rr] = (cc == rr && cc >= 16 * nbk) ? T(1) : T(0);
        }
    }
    // lane l of wave 0: the pivot threshold of tile row l (A(k0 + l, k0 + l))
    const double thr = piv_rel * fabs(diagA[col0[t.s] + k0 + min(lane, wl)] + delta);
    bool bad = false;
    __syncthreads();
    if (wv == 0) potrf_diag16<T>(S, XS, Cb, 0, w, thr, bad);
    __syncthreads();
#pragma unroll 1
    for (int kb = 0; kb < nbk - 1; kb++) {
        const int o = 16 * kb, nb = nbk - 1 - kb;
        if (wv < nb) {
            const int ri = 16 * (kb + 1 + wv);
            V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                acc = MfmaAcc<T>::blk(S[o + kk + lk][ri + lr], XS[o + kk + lk][o + lr], acc);
#pragma unroll
            for (int v = 0; v < 4; v++) S[o + MfmaAcc<T>::colof(lk, v)][ri + lr] = acc[v];
        }
        __syncthreads();
        // block pairs (kb + ib, kb + 1 + jb), jb < ib: pair 0 (the next diagonal
        // block) on wave 0, which then factors it; the others on waves 1-3
        const int np = nb * (nb + 1) / 2;
        for (int pi = wv == 0 ? 0 : wv; pi < (wv == 0 ? 1 : np); pi += 3) {
            int ib = 1, jb = pi;
            while (jb >= ib) {
                jb -= ib;
                ib++;
            }
            const int ri = 16 * (kb + ib), rj = 16 * (kb + 1 + jb);
            V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                acc = MfmaAcc<T>::blk(S[o + kk + lk][ri + lr], S[o + kk + lk][rj + lr], acc);
#pragma unroll
            for (int v = 0; v < 4; v++) S[rj + MfmaAcc<T>::colof(lk, v)][ri + lr] -= acc[v];
        }
        if (wv == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            potrf_diag16<T>(S, XS, Cb, o + 16, w, thr, bad);
        }
        __syncthreads();
    }
#pragma unroll
    for (int dd = 1; dd < 4; dd++) {
        if (dd >= nbk) break;
        const bool act = wv < nbk - dd;
        const int ri = 16 * (dd + wv), rj = 16 * wv;     // block (i, j = i - dd)
        if (act) {
            V acc = {T(0), T(0), T(0), T(0)};
            for (int rm = rj; rm < ri; rm += 16)
#pragma unroll
                for (int kk = 0; kk < 16; kk += 4)
                    acc = MfmaAcc<T>::blk(S[rm + kk + lk][ri + lr], XS[rj + lr][rm + kk + lk], acc);
#pragma unroll
            for (int v = 0; v < 4; v++) TS[wv][MfmaAcc<T>::colof(lk, v)][lr] = acc[v];
        }
        __syncthreads();
        if (act) {
            V acc = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < 16; kk += 4)
                acc = MfmaAcc<T>::blk(XS[ri + kk + lk][ri + lr], TS[wv][lr][kk + lk], acc);
#pragma unroll
            for (int v = 0; v < 4; v++) XS[rj + MfmaAcc<T>::colof(lk, v)][ri + lr] = -acc[v];
        }
        __syncthreads();
    }
    // a failed pivot is seen by the lanes of its column's group
    if (__syncthreads_or(bad) && tid == 0) atomicOr(flag, 1);
    T *X = xinv + (int64_t)blockIdx.x * (TB * TB);
#pragma unroll
    for (int e = 0; e < TB * TB / 256; e++) {
        const int i = tid + 256 * e, cc = i >> 6, rr = i & 63;
        if (rr < w && cc < w && cc <= rr) base[(int64_t)cc * F + rr] = S[cc][rr];
        X[i] = XS[cc][rr];
    }
}

// tile row i below the diagonal block: B <- B X^T with X = L_kk^-1 from the
// panel's scratch slot t.j (MFMA, the product formed as for the Schur update)
template <class T>
__global__ void __launch_bounds__(256) k_trsm_mfma(T *pool, const TileTask *__restrict__ tasks,
                                                   const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                   const int32_t *__restrict__ ncol, const T *__restrict__ xinv)
{
    __shared__ T Bt[TB][TB + 4];   // Bt[k][row] = B(row, k)
    __shared__ T Xt[TB][TB + 4];   // Xt[k][col] = X(col, k)
    typedef typename MfmaAcc<T>::V V;
    const TileTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.k * TB, w = min(TB, c - k0);
    const int r0 = max(t.i * TB, k0 + w), r1 = min(F, (t.i + 1) * TB), nr = r1 - r0;
    if (nr <= 0) return;
    T *fr = pool + foff[t.s];
    const T *X = xinv + (int64_t)t.j * (TB * TB);
    const int w4 = (w + 3) & ~3;
    // all 32 loads of a thread in flight before its LDS stores
    T vb[TB * TB / 256], vx[TB * TB / 256];
#pragma unroll
    for (int u = 0; u < TB * TB / 256; u++) {
        const int e = threadIdx.x + 256 * u, rr = e % TB, kk = e / TB;
        vb[u] = (rr < nr && kk < w) ? fr[(int64_t)(k0 + kk) * F + r0 + rr] : T(0);
        vx[u] = (kk < w && rr < w) ? X[(int64_t)kk * TB + rr] : T(0);   // X(rr, kk): column kk of X
    }
#pragma unroll
    for (int u = 0; u < TB * TB / 256; u++) {
        const int e = threadIdx.x + 256 * u, rr = e % TB, kk = e / TB;
        Bt[kk][rr] = vb[u];
        Xt[kk][rr] = vx[u];
    }
    __syncthreads();
    V acc[2][2];
    zero_acc<T>(acc);
    tile_mma<T, TB + 4>(Bt, Xt, w4, acc);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rb = (wv & 1) * 32, cb = (wv >> 1) * 32, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a2 = 0; a2 < 2; a2++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb + 16 * a2 + lr, col = cb + 16 * b + MfmaAcc<T>::colof(lk, v);
                if (row < nr && col < w) fr[(int64_t)(k0 + col) * F + r0 + row] = acc[a2][b][v];
            }
}

// Schur update of a C tile on the matrix cores: each wave owns a 32 x 32
// quarter of the 64 x 64 tile as 2 x 2 blocks of 16 x 16, K = 4 per MFMA, the K
// range in 64-column slabs staged through LDS while the next slab's loads are
// in flight.  Lanes 0-15 of each result value hold 16 consecutive rows of one
// column, so the read-modify-write of the column-major front is coalesced.
// Every global load is issued before the first store (a store into the front
// between loads would order them).  SG: the signed factorization's update
// C -= L_i S L_j^T, the column signs applied to the j operand as it is loaded.
template <class T, bool SG>
__global__ void __launch_bounds__(256) k_update_mfma(T *pool, const TileTask *__restrict__ pt,
                                                     const int32_t *__restrict__ ufid,
                                                     const int64_t *__restrict__ upref, int outer,
                                                     const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                     const int32_t *__restrict__ ncol, const int64_t *__restrict__ col0,
                                                     const T *__restrict__ sgn)
{
    __shared__ T Ai[TB][TB + 4];   // [kk][row]
    __shared__ T Aj[TB][TB + 4];
    typedef typename MfmaAcc<T>::V V;
    const UpdTile u = upd_decode(pt, ufid, upref, Fs, ncol, blockIdx.x, outer);
    const int32_t F = Fs[u.s];
    const int r0 = u.r0, q0 = u.q0, nr = u.r1 - u.r0, nq = u.q1 - u.q0;
    if (nr <= 0 || nq <= 0) return;
    T *fr = pool + foff[u.s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rb = (wv & 1) * 32, cb = (wv >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    T va[16], vb[16], cv[2][2][4];
    const T *sg = SG ? sgn + col0[u.s] : nullptr;
    auto load = [&](int kc, int w) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int kk = wv + 4 * q;
            const bool in = kk < w;
            va[q] = (in && lane < nr) ? fr[(int64_t)(kc + kk) * F + r0 + lane] : T(0);
            vb[q] = (in && lane < nq) ? fr[(int64_t)(kc + kk) * F + q0 + lane] : T(0);
            if (SG) vb[q] = in ? vb[q] * sg[kc + kk] : vb[q];
        }
    };
    load(u.k0, min(TB, u.kw));
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb + 16 * a + lr, col = cb + 16 * b + MfmaAcc<T>::colof(lk, v);
                cv[a][b][v] = (row < nr && col < nq) ? fr[(int64_t)(q0 + col) * F + r0 + row] : T(0);
            }
    V acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = V{T(0), T(0), T(0), T(0)};
    for (int sb = 0; sb < u.kw; sb += TB) {
        const int w = min(TB, u.kw - sb), w4 = (w + 3) & ~3;
        if (sb > 0) __syncthreads();          // the previous slab's products are done
#pragma unroll
        for (int q = 0; q < 16; q++) {
            Ai[wv + 4 * q][lane] = va[q];
            Aj[wv + 4 * q][lane] = vb[q];
        }
        __syncthreads();
        if (sb + TB < u.kw) load(u.k0 + sb + TB, min(TB, u.kw - sb - TB));
        for (int kk = 0; kk < w4; kk += 4) {
            T xi[2], xj[2];
#pragma unroll
            for (int a = 0; a < 2; a++) {
                xi[a] = Ai[kk + lk][rb + 16 * a + lr];
                xj[a] = Aj[kk + lk][cb + 16 * a + lr];
            }
#pragma unroll
            for (int a = 0; a < 2; a++)
#pragma unroll
                for (int b = 0; b < 2; b++) acc[a][b] = MfmaAcc<T>::blk(xi[a], xj[b], acc[a][b]);
        }
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) {
            const int row = rb + 16 * a + lr;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int col = cb + 16 * b + MfmaAcc<T>::colof(lk, v);
                if (row < nr && col < nq) fr[(int64_t)(q0 + col) * F + r0 + row] = cv[a][b][v] - acc[a][b][v];
            }
        }
}

// ------------------------------------------------------------------ solves
__global__ void k_gather_perm(const double *__restrict__ x, const int64_t *__restrict__ perm, double *__restrict__ y,
                              int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = x[perm[i]];
}

__global__ void k_scatter_perm(const double *__restrict__ x, const int64_t *__restrict__ perm, double *__restrict__ y,
                               int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[perm[i]] = x[i];
}

// between the two sweeps of a signed factor's solve: y <- S y
template <class T>
__global__ void k_sign_scale(double *__restrict__ y, const T *__restrict__ sgn, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] *= (double)sgn[i];
}



// forward: L y = b for every front of a level (one block per front)
template <class T>
__global__ void __launch_bounds__(256) k_fwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                             SolveArgs a, const double *__restrict__ b, double *__restrict__ y)
{
    __shared__ double yp[TB];
    const int s = fronts[blockIdx.x];
    const int32_t F = a.Fs[s], c = a.ncol[s];
    double *w = a.ws + a.wsoff[s];
    const int64_t c0 = a.col0[s];
    for (int i = threadIdx.x; i < F; i += 256) w[i] = i < c ? b[c0 + i] : 0.0;
    __syncthreads();
    for (int64_t p = a.chptr[s]; p < a.chptr[s + 1]; p++) {
        const int ch = a.chl[p];
        const int32_t cc = a.ncol[ch], rc = a.Fs[ch] - cc;
        const double *u = a.ws + a.wsoff[ch] + cc;
        const int32_t *rl = a.rel + a.relptr[ch];
        for (int t = threadIdx.x; t < rc; t += 256) w[rl[t]] += u[t];
        __syncthreads();
    }
    const T *L = pool + a.foff[s];
    for (int k0 = 0; k0 < c; k0 += TB) {
        const int wd = min(TB, c - k0);
        if (threadIdx.x < 64) {
            const int q = threadIdx.x;
            double v = q < wd ? w[k0 + q] : 0.0;
            for (int p = 0; p < wd; p++) {
                if (q == p) v = v / (double)L[(int64_t)(k0 + p) * F + k0 + p];
                const double yv = __shfl(v, p, 64);
                if (q > p && q < wd) v -= (double)L[(int64_t)(k0 + p) * F + k0 + q] * yv;
            }
            if (q < wd) { w[k0 + q] = v; yp[q] = v; }
        }
        __syncthreads();
        for (int r = k0 + wd + threadIdx.x; r < F; r += 256) {
            double s2 = 0.0;
            for (int p = 0; p < wd; p++) s2 = fma((double)L[(int64_t)(k0 + p) * F + r], yp[p], s2);
            w[r] -= s2;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < c; i += 256) y[c0 + i] = w[i];
}

// backward: L^T x = y for every front of a level
template <class T>
__global__ void __launch_bounds__(256) k_bwd(const T *__restrict__ pool, const int32_t *__restrict__ fronts,
                                             SolveArgs a, const double *__restrict__ y, double *__restrict__ x)
{
    __shared__ double z[TB];
    const int s = fronts[blockIdx.x];
    const int32_t F = a.Fs[s], c = a.ncol[s];
    double *w = a.ws + a.wsoff[s];
    const int64_t c0 = a.col0[s];
    const int32_t *R = a.rows + a.rptr[s];
    for (int i = threadIdx.x; i < F; i += 256) w[i] = i < c ? y[c0 + i] : x[R[i - c]];
    __syncthreads();
    const T *L = pool + a.foff[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int npan = (c + TB - 1) / TB;
    for (int kp = npan - 1; kp >= 0; kp--) {
        const int k0 = kp * TB, wd = min(TB, c - k0);
        for (int p = wv; p < wd; p += 4) {
            const T *colp = L + (int64_t)(k0 + p) * F;
            double s2 = 0.0;
            for (int r = k0 + wd + lane; r < F; r += 64) s2 = fma((double)colp[r], w[r], s2);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) s2 += __shfl_xor(s2, off, 64);
            if (lane == 0) z[p] = w[k0 + p] - s2;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int q = threadIdx.x;
            double v = q < wd ? z[q] : 0.0;
            for (int p = wd - 1; p >= 0; p--) {
                if (q == p) v = v / (double)L[(int64_t)(k0 + p) * F + k0 + p];
                const double xv = __shfl(v, p, 64);
                if (q < p) v -= (double)L[(int64_t)(k0 + q) * F + k0 + p] * xv;
            }
            if (q < wd) w[k0 + q] = v;
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < c; i += 256) x[c0 + i] = w[i];
}

// ---- blocked multi-workgroup solves ------------------------------------------
// A front of order F is cut into 64-row blocks (forward) or 64-column panels
// (backward).  Work items take tickets from a per-launch counter in dependency
// order, so an item only ever waits on items holding smaller tickets -- items
// that are already running -- and the launch cannot deadlock whatever the
// residency.  Completion of a panel is published as flag[panel] = epoch (the
// solve's sequence number, so flags never need clearing); the panel values
// themselves cross workgroups through agent-scope atomic loads and stores.
// Every wait is bounded: after ~2^22 polls the item records an error and
// continues, so a logic error shows up as SPEIGS_ERR_INTERNAL, never a hang.


// The wait's bound and the value it waits for come from the host (WaitSpec):
// normally 2^22 polls for the solve's own epoch; a host test sets a value no
// item ever publishes and a short bound, so that every wait gives up and the
// error path runs with every kernel still exiting normally.
struct WaitSpec {
    int mask;    // awaited value = epoch ^ mask (0 in normal operation)
    int spin;    // polls before giving up
};

__device__ __forceinline__ void wait_epoch(const int *f, int epoch, WaitSpec ws, int *err)
{
    epoch ^= ws.mask;
    // relaxed polling: an agent-scope acquire here would invalidate the L2 of
    // the XCD on every poll; the published data is read with coherent loads
    for (int it = 0; it < ws.spin; it++) {
        if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch) return;
        if ((it & 63) == 63 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
        __builtin_amdgcn_s_sleep(1);
    }
    atomicOr(err, 2);
}

__device__ __forceinline__ double ld_coherent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_coherent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// after coherent (write-through) stores of a panel: wait until every memory
// access the wave issued is acknowledged (s_waitcnt 0: vmcnt, expcnt and
// lgkmcnt all zero), then raise the flag
__device__ __forceinline__ void publish(int *f, int epoch)
{
    __builtin_amdgcn_s_waitcnt(0);
    __hip_atomic_store(f, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double bcast_lane(double v, int src)
{
    const int64_t bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffff), src);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), src);
    return __longlong_as_double(((int64_t)hi << 32) | (uint32_t)lo);
}

__global__ void k_zero_counters(unsigned *c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

// ws[s] rows [ck*256, ck*256+256) = [b(c0 .. c0+c) ; 0] plus the children's
// update vectors (children in order: deterministic)
__global__ void __launch_bounds__(256) k_fwd_init(const PcbTask *__restrict__ tasks, SolveArgs a,
                                                  const double *__restrict__ b)
{
    const PcbTask t = tasks[blockIdx.x];
    const int s = t.parent;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    double *w = a.ws + a.wsoff[s];
    const int64_t c0 = a.col0[s];
    const int r0 = t.pcb * 256, r1 = min(F, r0 + 256);
    const int i = r0 + threadIdx.x;
    if (i < r1) w[i] = i < c ? b[c0 + i] : 0.0;
    for (int64_t q = a.chptr[s]; q < a.chptr[s + 1]; q++) {
        __syncthreads();
        const int ch = a.chl[q];
        const int32_t cc = a.ncol[ch], rc = a.Fs[ch] - cc;
        const int32_t *rl = a.rel + a.relptr[ch];
        const int lo = lower_bound_i32(rl, rc, r0), hi = lower_bound_i32(rl, rc, r1);
        const double *u = a.ws + a.wsoff[ch] + cc;
        for (int e = lo + threadIdx.x; e < hi; e += 256) w[rl[e]] += u[e];
    }
}

// ws[parent][rel[t]] += u_child[t] for one child ordinal (EaTask: child, parent)
__global__ void __launch_bounds__(256) k_fwd_ea(const EaTask *__restrict__ tasks, SolveArgs a)
{
    const EaTask t = tasks[blockIdx.x];
    const int32_t cc = a.ncol[t.child], rc = a.Fs[t.child] - cc;
    const double *u = a.ws + a.wsoff[t.child] + cc;
    double *w = a.ws + a.wsoff[t.parent];
    const int32_t *rl = a.rel + a.relptr[t.child];
    for (int i = threadIdx.x; i < rc; i += 256) w[rl[i]] += u[i];
}

template <class T>
__global__ void __launch_bounds__(256) k_fwd_blk(const T *__restrict__ pool, const BlkTask *__restrict__ tasks,
                                                 SolveArgs a, const int64_t *__restrict__ panbase, int *flags,
                                                 int epoch, WaitSpec wsp, unsigned *ticket, int *err, double *y)
{
    __shared__ double wl[64];
    __shared__ double yp[64];
    __shared__ double part[4][64];
    __shared__ double Ld[64][65];        // diagonal block, Ld[p][q] = L[r0+q][r0+p]
    __shared__ unsigned tk;
    const int s = tasks[blockIdx.x].s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int npan = (c + 63) / 64;
    const unsigned nblk = (unsigned)((F + 63) / 64);
    if (threadIdx.x == 0) tk = nblk == 1 ? 0u : atomicAdd(&ticket[s], 1u) % nblk;
    __syncthreads();
    const int b = (int)tk;
    const int r0 = b * 64, nr = min(64, F - r0);
    double *ws = a.ws + a.wsoff[s];
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s];
    int *fl = flags + panbase[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const bool has_diag = b < npan;
    const int wd = has_diag ? min(64, c - r0) : 0;
    if (threadIdx.x < 64) wl[lane] = lane < nr ? ws[r0 + lane] : 0.0;
    if (has_diag)
        for (int e = threadIdx.x; e < 64 * 64; e += 256) {
            const int q = e & 63, p = e >> 6;
            Ld[p][q] = (p < wd && q < nr) ? (double)L[(int64_t)(r0 + p) * F + r0 + q] : 0.0;
        }
    const int pend = min(b, npan);
    // rows r0+lane, columns p*64 + wv + 4j: prefetched one panel ahead of its flag
    T pre[16];
    auto fetch = [&](int p) {
        const int pw = min(64, c - p * 64);
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const int col = wv + 4 * j;
            pre[j] = (lane < nr && col < pw) ? L[(int64_t)(p * 64 + col) * F + r0 + lane] : T(0);
        }
    };
    if (pend > 0) fetch(0);
    __syncthreads();
    for (int p = 0; p < pend; p++) {
        T cur[16];
#pragma unroll
        for (int j = 0; j < 16; j++) cur[j] = pre[j];
        if (p + 1 < pend) fetch(p + 1);
        if (threadIdx.x == 0) wait_epoch(&fl[p], epoch, wsp, err);
        __syncthreads();
        const int pw = min(64, c - p * 64);
        if (threadIdx.x < 64) yp[lane] = lane < pw ? ld_coherent(&y[c0 + p * 64 + lane]) : 0.0;
        __syncthreads();
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 16; j++) acc = fma((double)cur[j], yp[wv + 4 * j], acc);
        part[wv][lane] = acc;
        __syncthreads();
        if (threadIdx.x < 64) wl[lane] -= (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
        __syncthreads();
    }
    if (has_diag) {
        if (wv == 0) {
            double v = wl[lane];
            const double dinv = lane < wd ? 1.0 / Ld[lane][lane] : 0.0;
#pragma unroll 8
            for (int p = 0; p < wd; p++) {
                const double lq = Ld[p][lane];
                const double tq = v * dinv;
                const double yv = bcast_lane(tq, p);
                if (lane == p) v = tq;
                else if (lane > p) v = fma(-lq, yv, v);
            }
            if (lane < wd) st_coherent(&y[c0 + r0 + lane], v);
            else if (lane < nr) ws[r0 + lane] = v;
            if (lane == 0) publish(&fl[b], epoch);
        }
    } else if (threadIdx.x < 64 && lane < nr) {
        ws[r0 + lane] = wl[lane];
    }
}

// backward, update rows: part[s][p][j][q] = sum over rows r of chunk j (512 rows
// of the front's update part) of L[r][p*64+q] * x[R[r-c]] -- fully parallel


template <class T>
__global__ void __launch_bounds__(256) k_bwd_upd(const T *__restrict__ pool, const UpdTask *__restrict__ tasks,
                                                 SolveArgs a, const double *__restrict__ x, double *__restrict__ part,
                                                 const int64_t *__restrict__ partoff)
{
    __shared__ double xs[512];
    __shared__ double red[4][64];
    const UpdTask t = tasks[blockIdx.x];
    const int s = t.s, p = t.p;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int k0 = p * 64, wd = min(64, c - k0);
    const int rb = c + t.j * 512, nrr = min(512, F - rb);
    const T *L = pool + a.foff[s];
    const int32_t *R = a.rows + a.rptr[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < nrr; i += 256) xs[i] = x[R[rb - c + i]];
    __syncthreads();
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0;
    for (int r = lane; r < nrr; r += 64) {
        const double xv = xs[r];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int q = wv + 4 * i;
            if (q < wd) acc[i] = fma((double)L[(int64_t)(k0 + q) * F + rb + r], xv, acc[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == i) red[wv][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int nch = (F - c + 511) / 512;
        const int q = lane;
        part[partoff[s] + ((int64_t)p * nch + t.j) * 64 + q] = (q < wd) ? red[q & 3][q >> 2] : 0.0;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_bwd_blk(const T *__restrict__ pool, const BlkTask *__restrict__ tasks,
                                                 SolveArgs a, const int64_t *__restrict__ panbase, int *flags,
                                                 int epoch, WaitSpec wsp, unsigned *ticket, int *err,
                                                 const double *__restrict__ y,
                                                 double *x, const double *__restrict__ part,
                                                 const int64_t *__restrict__ partoff)
{
    __shared__ double xs[256];
    __shared__ double z[64];
    __shared__ double Ld[64][65];        // Ld[q][r] = L[k0+r][k0+q] (column q of the diagonal block)
    __shared__ double red[4][64];
    __shared__ unsigned tk;
    const int s = tasks[blockIdx.x].s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int npan = (c + 63) / 64;
    if (threadIdx.x == 0) tk = npan == 1 ? 0u : atomicAdd(&ticket[s], 1u) % (unsigned)npan;
    __syncthreads();
    const int p = npan - 1 - (int)tk;
    const int k0 = p * 64, wd = min(64, c - k0);
    const T *L = pool + a.foff[s];
    const int64_t c0 = a.col0[s];
    const int32_t *R = a.rows + a.rptr[s];
    int *fl = flags + panbase[s];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int r = e & 63, q = e >> 6;
        Ld[q][r] = (r < wd && q < wd) ? (double)L[(int64_t)(k0 + q) * F + k0 + r] : 0.0;
    }
    if (threadIdx.x < 64) z[lane] = lane < wd ? y[c0 + k0 + lane] : 0.0;
    // thread (lane, wv) accumulates columns q = wv + 4i over rows lane, lane+64, ...
    double acc[16];
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.0;
    // update rows (>= c): partial sums from k_bwd_upd, added in fixed chunk order
    {
        const int nch = (F - c + 511) / 512;
        const double *pt = part + partoff[s] + (int64_t)p * nch * 64;
        __syncthreads();
        if (threadIdx.x < 64) {
            double sum = 0.0;
            for (int j = 0; j < nch; j++) sum += pt[(int64_t)j * 64 + lane];
            if (lane < wd) z[lane] -= sum;
        }
    }
    // later panels of this front, prefetching each L block before its flag
    T pre[16];
    auto fetch = [&](int pp) {
        const int pw = min(64, c - pp * 64);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int q = wv + 4 * i;
            pre[i] = (lane < pw && q < wd) ? L[(int64_t)(k0 + q) * F + pp * 64 + lane] : T(0);
        }
    };
    if (npan - 1 > p) fetch(npan - 1);
    for (int pp = npan - 1; pp > p; pp--) {
        T cur[16];
#pragma unroll
        for (int i = 0; i < 16; i++) cur[i] = pre[i];
        if (pp - 1 > p) fetch(pp - 1);
        __syncthreads();
        if (threadIdx.x == 0) wait_epoch(&fl[pp], epoch, wsp, err);
        __syncthreads();
        if (threadIdx.x < 64) xs[lane] = (pp * 64 + lane < c) ? ld_coherent(&x[c0 + pp * 64 + lane]) : 0.0;
        __syncthreads();
        const double xv = xs[lane];
#pragma unroll
        for (int i = 0; i < 16; i++) acc[i] = fma((double)cur[i], xv, acc[i]);
    }
    // reduce the 64 lanes of each column: transpose through LDS, 16 columns per wave
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double v = acc[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == i) red[wv][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int q = lane, w2 = q & 3, i = q >> 2;
        if (q < wd) z[q] -= red[w2][i];
    }
    __syncthreads();
    if (wv == 0) {
        const double dinv = lane < wd ? 1.0 / Ld[lane][lane] : 0.0;
        double v = z[lane];
#pragma unroll 8
        for (int r = wd - 1; r >= 0; r--) {
            const double lq = Ld[lane][r];
            const double tq = v * dinv;
            const double xv = bcast_lane(tq, r);
            if (lane == r) v = tq;
            else if (lane < r) v = fma(-lq, xv, v);
        }
        if (lane < wd) st_coherent(&x[c0 + k0 + lane], v);
        if (lane == 0) publish(&fl[p], epoch);
    }
}

__global__ void k_norm2_partial(const double *x, int64_t n, double *partial)
{
    __shared__ double lds[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) s = fma(x[i], x[i], s);
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = lds[0];
}

// ------------------------------------------------------------------ host side




constexpr int kEvPairs = 32;

void ev_harvest(Factor *F, bool wait)
{
    std::vector<int> keep;
    for (int i : F->ev_pending) {
        if (!wait && hipEventQuery(F->ev[2 * i + 1]) != hipSuccess) { keep.push_back(i); continue; }
        if (wait) (void)hipEventSynchronize(F->ev[2 * i + 1]);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, F->ev[2 * i], F->ev[2 * i + 1]) == hipSuccess) F->t_solve += 1e-3 * ms;
    }
    F->ev_pending.swap(keep);
}

int ev_begin(Factor *F)
{
    if (F->ev.empty()) return -1;
    for (int tries = 0; tries < kEvPairs; tries++) {
        const int i = F->ev_next;
        F->ev_next = (F->ev_next + 1) % kEvPairs;
        if (std::find(F->ev_pending.begin(), F->ev_pending.end(), i) == F->ev_pending.end()) {
            (void)hipEventRecord(F->ev[2 * i], F->st);
            return i;
        }
    }
    return -1;
}

void ev_end(Factor *F, int i)
{
    if (i < 0) return;
    (void)hipEventRecord(F->ev[2 * i + 1], F->st);
    F->ev_pending.push_back(i);
}

namespace {

template <class T>
int run_numeric(Factor *F, double delta)
{
    const Symbolic &S = F->S;
    hipStream_t st = F->st;
    T *pool = (T *)F->pool;
    const double piv_rel = F->fp32 ? 1e-5 : 1e-11;
    // the signed factorization runs on the register panel and the MFMA update
    const bool sg = F->signed_fac;
    if (sg && !F->use_reg_panel) return SEHIP_FAIL(SPEIGS_ERR_ARG, "signed factorization needs the register panel");
    SEHIP_CHECK(hipMemsetAsync(F->flag, 0, 4 * sizeof(int), st));
    if ((uintptr_t)pool % 16) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "front pool not 16-byte aligned");
    if (F->pool_zeroed) {
        // the first factorization: the fill queued at set-up
        SEHIP_CHECK(hipStreamWaitEvent(st, F->ev_zero, 0));
        F->pool_zeroed = false;
    } else {
        k_zero<T><<<dim3((unsigned)std::min<int64_t>(4096, (S.pool * (int64_t)sizeof(T) / 16 + 255) / 256 + 1)),
                    dim3(256), 0, st>>>(pool, S.pool);
    }
    const dim3 gn((unsigned)((F->n + 255) / 256));
    k_diag_delta<T><<<gn, dim3(256), 0, st>>>(pool, F->snode, F->col0, F->Fs, F->foff, F->n, delta);
    k_assemble<T><<<gn, dim3(256), 0, st>>>(pool, F->Apd, F->Aid, F->Ax, F->iperm, F->snode, F->col0, F->ncol, F->Fs,
                                            F->rptr, F->rows, F->foff, F->n, delta, F->flag);
    bool far_pending = false;     // a far part of a block-end update is queued on st2
    for (int32_t L = 0; L < S.nlevels; L++) {
        const LevelPlan &P = F->plan[L];
        if (P.pcb_cnt > 0)
            k_extend_add_blk<T><<<dim3((unsigned)P.pcb_cnt), dim3(256), 0, st>>>(
                pool, F->eab + P.pcb_off, F->foff, F->Fs, F->ncol, F->relptr, F->rel, F->chptr, F->chl, F->bndptr,
                F->bnd);
        for (size_t k = 0; k < P.k_potrf_cnt.size(); k++) {
            if (F->use_reg_panel) {
                // panel in registers (+ its inverse), then TRSM as a product with the inverse
                // (in groups of kXinvSlots fronts, the inverses' scratch capacity)
                T *xinv = (T *)F->xinv;
                const std::vector<int64_t> &cut = P.k_trsm_cut[k];
                for (size_t g = 0; g + 1 < cut.size(); g++) {
                    const int64_t p0 = (int64_t)g * kXinvSlots;
                    const int64_t pc = std::min(kXinvSlots, P.k_potrf_cnt[k] - p0);
                    if (sg)
                        k_potrf_reg4<T, true><<<dim3((unsigned)pc), dim3(256), 0, st>>>(
                            pool, F->tiles + P.k_potrf_off[k] + p0, F->foff, F->Fs, F->ncol, F->col0, F->diagA,
                            delta, piv_rel, F->flag, xinv, (T *)F->sgn, F->tau);
                    else if (F->potrf_blocked)
                        k_potrf_blk<T><<<dim3((unsigned)pc), dim3(256), 0, st>>>(
                            pool, F->tiles + P.k_potrf_off[k] + p0, F->foff, F->Fs, F->ncol, F->col0, F->diagA,
                            delta, piv_rel, F->flag, xinv);
                    else if (F->potrf_waves == 4)
                        k_potrf_reg4<T, false><<<dim3((unsigned)pc), dim3(256), 0, st>>>(
                            pool, F->tiles + P.k_potrf_off[k] + p0, F->foff, F->Fs, F->ncol, F->col0, F->diagA,
                            delta, piv_rel, F->flag, xinv, nullptr, 0.0);
                    else
                        k_potrf_reg<T><<<dim3((unsigned)pc), dim3(64), 0, st>>>(
                            pool, F->tiles + P.k_potrf_off[k] + p0, F->foff, F->Fs, F->ncol, F->col0, F->diagA,
                            delta, piv_rel, F->flag, xinv);
                    if (cut[g + 1] > cut[g])
                        k_trsm_mfma<T><<<dim3((unsigned)(cut[g + 1] - cut[g])), dim3(256), 0, st>>>(
                            pool, F->tiles + cut[g], F->foff, F->Fs, F->ncol, xinv);
                }
            } else {
                if (P.k_potrf_cnt[k] > 0)
                    k_potrf<T><<<dim3((unsigned)P.k_potrf_cnt[k]), dim3(256), 0, st>>>(
                        pool, F->tiles + P.k_potrf_off[k], F->foff, F->Fs, F->ncol, F->col0, F->diagA, delta, piv_rel,
                        F->flag);
                if (P.k_trsm_cnt[k] > 0)
                    k_trsm<T><<<dim3((unsigned)P.k_trsm_cnt[k]), dim3(256), 0, st>>>(pool, F->tiles + P.k_trsm_off[k],
                                                                                     F->foff, F->Fs, F->ncol);
            }
            // inner then outer updates: disjoint columns of the trailing matrix
            const TileTask *pt = F->tiles + P.k_potrf_off[k];
            auto launch = [&](int outer, int64_t cnt, int64_t off, int64_t refoff, hipStream_t s2) {
                if (cnt <= 0) return;
                const int32_t *uf = F->ufid + off;
                const int64_t *up = F->upref + refoff;
                if (sg)
                    k_update_mfma<T, true><<<dim3((unsigned)cnt), dim3(256), 0, s2>>>(
                        pool, pt, uf, up, outer, F->foff, F->Fs, F->ncol, F->col0, (const T *)F->sgn);
                else if (F->use_mfma)
                    k_update_mfma<T, false><<<dim3((unsigned)cnt), dim3(256), 0, s2>>>(
                        pool, pt, uf, up, outer, F->foff, F->Fs, F->ncol, nullptr, nullptr);
                else
                    k_update<T><<<dim3((unsigned)cnt), dim3(256), 0, s2>>>(pool, pt, uf, up, outer, F->foff, F->Fs,
                                                                           F->ncol);
            };
            launch(0, P.k_upd_cnt[k], P.k_upd_off[k], P.k_upref_off[k], st);
            // Look-ahead: the block-end update's near part (the next block's
            // columns, which the next panels need) stays on this stream, its far
            // part runs on st2 beside the next block's POTRF / TRSM chain.  A
            // far part and the next block-end update write the same columns, so
            // that one waits for it; tiles see their updates in the same order,
            // and the factor is bitwise the same.
            const int64_t nnear = P.k_out_cnt[k], nfar = P.k_out2_cnt[k];
            if (nnear > 0 || nfar > 0) {
                if (far_pending) {
                    SEHIP_CHECK(hipStreamWaitEvent(st, F->ev_far, 0));
                    far_pending = false;
                }
                if (nfar > 0) {
                    SEHIP_CHECK(hipEventRecord(F->ev_ready, st));
                    SEHIP_CHECK(hipStreamWaitEvent(F->st2, F->ev_ready, 0));
                }
                launch(1, nnear, P.k_out_off[k], P.k_outref_off[k], st);
                if (nfar > 0) {
                    launch(2, nfar, P.k_out2_off[k], P.k_out2ref_off[k], F->st2);
                    SEHIP_CHECK(hipEventRecord(F->ev_far, F->st2));
                    far_pending = true;
                }
            }
        }
        // the next level's extend-add reads every update block of this one
        if (far_pending) {
            SEHIP_CHECK(hipStreamWaitEvent(st, F->ev_far, 0));
            far_pending = false;
        }
    }
    int hflag[4] = {0, 0, 0, 0};
    SEHIP_CHECK(hipMemcpyAsync(hflag, F->flag, sizeof hflag, hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    if (hflag[0] & 2) {   // an entry of A outside the symbolic structure
        char what[160];
        snprintf(what, sizeof what, "assembly: A(%d, %d) has no place in the symbolic structure", hflag[2], hflag[3]);
        return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, what);
    }
    F->perturbed = sg ? hflag[1] : 0;
    return hflag[0] ? 1 : 0;
}

template <class T>
void run_solve(Factor *F, const double *b, double *x)
{
    const Symbolic &S = F->S;
    hipStream_t st = F->st;
    const T *pool = (const T *)F->pool;
    SolveArgs a{F->foff, F->wsoff, F->col0, F->rptr, F->relptr, F->chptr, F->Fs, F->ncol, F->rows, F->rel, F->chl, F->ws};
    const unsigned nb = (unsigned)((F->n + 255) / 256);
    if (F->solve_mode == 2) {           // inverse-based, chain-free (see blocksolve.hip)
        if (factor_block_solve_d(F, b, x, 1) == 0) F->solves++;
        return;
    }
    k_gather_perm<<<dim3(nb), dim3(256), 0, st>>>(b, F->perm, F->xp, F->n);
    if (F->solve_mode == 1) {
        const int ep = ++F->epoch;
        const WaitSpec wsp = F->wait_test ? WaitSpec{1 << 30, 64} : WaitSpec{0, 1 << 22};
        for (int32_t L = 0; L < S.nlevels; L++) {
            const LevelPlan &P = F->plan[L];
            if (!P.fr_cnt) continue;
            k_fwd_init<<<dim3((unsigned)P.fin_cnt), dim3(256), 0, st>>>(F->fin + P.fin_off, a, F->xp);
            k_fwd_blk<T><<<dim3((unsigned)P.fb_cnt), dim3(256), 0, st>>>(pool, F->fblk + P.fb_off, a, F->panbase,
                                                                        F->fflag, ep, wsp, F->tick_f, F->err, F->yp);
        }
        if (F->signed_fac) k_sign_scale<T><<<dim3(nb), dim3(256), 0, st>>>(F->yp, (const T *)F->sgn, F->n);
        for (int32_t L = S.nlevels - 1; L >= 0; L--) {
            const LevelPlan &P = F->plan[L];
            if (!P.fr_cnt) continue;
            if (P.ub_cnt)
                k_bwd_upd<T><<<dim3((unsigned)P.ub_cnt), dim3(256), 0, st>>>(pool, F->upd + P.ub_off, a, F->xp,
                                                                            F->bpart, F->partoff);
            k_bwd_blk<T><<<dim3((unsigned)P.bb_cnt), dim3(256), 0, st>>>(pool, F->bblk + P.bb_off, a, F->panbase,
                                                                        F->bflag, ep, wsp, F->tick_b, F->err, F->yp, F->xp,
                                                                        F->bpart, F->partoff);
        }
        k_scatter_perm<<<dim3(nb), dim3(256), 0, st>>>(F->xp, F->perm, x, F->n);
        (void)hipMemcpyAsync(F->herr, F->err, sizeof(int), hipMemcpyDeviceToHost, st);
        F->solves++;
        return;
    }
    for (int32_t L = 0; L < S.nlevels; L++) {
        const LevelPlan &P = F->plan[L];
        if (P.fr_cnt) k_fwd<T><<<dim3((unsigned)P.fr_cnt), dim3(256), 0, st>>>(pool, F->fronts + P.fr_off, a, F->xp, F->yp);
    }
    if (F->signed_fac) k_sign_scale<T><<<dim3(nb), dim3(256), 0, st>>>(F->yp, (const T *)F->sgn, F->n);
    for (int32_t L = S.nlevels - 1; L >= 0; L--) {
        const LevelPlan &P = F->plan[L];
        if (P.fr_cnt) k_bwd<T><<<dim3((unsigned)P.fr_cnt), dim3(256), 0, st>>>(pool, F->fronts + P.fr_off, a, F->yp, F->xp);
    }
    k_scatter_perm<<<dim3(nb), dim3(256), 0, st>>>(F->xp, F->perm, x, F->n);
    F->solves++;
}

double device_norm(Factor *F, const double *v)
{
    const int nb = 256;
    k_norm2_partial<<<dim3(nb), dim3(256), 0, F->st>>>(v, F->n, F->partial);
    if (hipMemcpyAsync(F->hpart, F->partial, nb * sizeof(double), hipMemcpyDeviceToHost, F->st) != hipSuccess) return -1;
    if (hipStreamSynchronize(F->st) != hipSuccess) return -1;
    double s = 0.0;
    for (int i = 0; i < nb; i++) s += F->hpart[i];
    return std::sqrt(s);
}

template <class D, class Al>
int upload(Factor *F, D **dst, const std::vector<D, Al> &src)
{
    int rc = F->mem.alloc(dst, src.size());
    if (rc) return rc;
    if (!src.empty()) return copy_to_device(*dst, src.data(), src.size() * sizeof(D), F->st);
    return 0;
}

// The plan of the blocked single-vector solves (the Lanczos paths): built
// from the symbolic structure and uploaded once per factor.
int single_solve_plan(Factor *F)
{
    if (F->plan_ready) return 0;
    const Symbolic &S = F->S;
    const int64_t ns = S.ns;
    hipStream_t st = F->st;
    int rc = 0;
    double tp = wall();
    std::vector<std::vector<int64_t>> bylevel(S.nlevels);
    for (int64_t s = 0; s < ns; s++) bylevel[S.level[s]].push_back(s);
    // blocked solve plans: forward row blocks in increasing order per front,
    // backward panels in decreasing order per front (ticket = dependency order)
    std::vector<EaTask> sea;
    std::vector<BlkTask> fblk, bblk;
    std::vector<int64_t> panbase(ns + 1, 0);
    for (int64_t s2 = 0; s2 < ns; s2++) panbase[s2 + 1] = panbase[s2] + (S.c1[s2] - S.c0[s2] + 63) / 64;
    for (int32_t L = 0; L < S.nlevels; L++) {
        LevelPlan &P = F->plan[L];
        int64_t maxch = 0;
        for (int64_t s2 : bylevel[L]) maxch = std::max(maxch, S.chptr[s2 + 1] - S.chptr[s2]);
        for (int64_t o2 = 0; o2 < maxch; o2++) {
            P.sea_off.push_back((int64_t)sea.size());
            for (int64_t s2 : bylevel[L])
                if (S.chptr[s2] + o2 < S.chptr[s2 + 1])
                    sea.push_back({(int32_t)S.chl[S.chptr[s2] + o2], (int32_t)s2, 0, 0});
            P.sea_cnt.push_back((int64_t)sea.size() - P.sea_off.back());
        }
        P.fb_off = (int64_t)fblk.size();
        for (int64_t s2 : bylevel[L])
            for (int64_t b = 0; b * 64 < S.F[s2]; b++) fblk.push_back({(int32_t)s2, (int32_t)b});
        P.fb_cnt = (int64_t)fblk.size() - P.fb_off;
        P.bb_off = (int64_t)bblk.size();
        for (int64_t s2 : bylevel[L])
            for (int64_t pp = (S.c1[s2] - S.c0[s2] + 63) / 64 - 1; pp >= 0; pp--) bblk.push_back({(int32_t)s2, (int32_t)pp});
        P.bb_cnt = (int64_t)bblk.size() - P.bb_off;
    }
    std::vector<PcbTask> fin;
    for (int32_t L = 0; L < S.nlevels; L++) {
        LevelPlan &P = F->plan[L];
        P.fin_off = (int64_t)fin.size();
        for (int64_t s2 : bylevel[L])
            for (int64_t cb = 0; cb * 256 < S.F[s2]; cb++) fin.push_back({(int32_t)s2, (int32_t)cb});
        P.fin_cnt = (int64_t)fin.size() - P.fin_off;
    }
    // backward update-row chunks and their partial-sum slots (reused across levels)
    std::vector<UpdTask> updv;
    std::vector<int64_t> partoff(ns + 1, 0);
    int64_t partmax = 1;
    for (int32_t L = 0; L < S.nlevels; L++) {
        LevelPlan &P = F->plan[L];
        P.ub_off = (int64_t)updv.size();
        int64_t off = 0;
        for (int64_t s2 : bylevel[L]) {
            const int64_t c = S.c1[s2] - S.c0[s2], r = S.F[s2] - c;
            const int64_t npan = (c + 63) / 64, nch = (r + 511) / 512;
            partoff[s2] = off;
            off += npan * nch * 64;
            for (int64_t pp = 0; pp < npan; pp++)
                for (int64_t j = 0; j < nch; j++) updv.push_back({(int32_t)s2, (int32_t)pp, (int32_t)j, 0});
        }
        partmax = std::max(partmax, off);
        P.ub_cnt = (int64_t)updv.size() - P.ub_off;
    }
    if (updv.empty()) updv.push_back({0, 0, 0, 0});
    if ((rc = upload(F, &F->upd, updv))) return rc;
    if ((rc = upload(F, &F->partoff, partoff))) return rc;
    if ((rc = F->mem.alloc(&F->bpart, (size_t)partmax))) return rc;
    if ((rc = F->mem.alloc(&F->tick_f, (size_t)ns + 1))) return rc;
    if ((rc = F->mem.alloc(&F->tick_b, (size_t)ns + 1))) return rc;
    SEHIP_CHECK(hipMemsetAsync(F->tick_f, 0, ((size_t)ns + 1) * sizeof(unsigned), st));
    SEHIP_CHECK(hipMemsetAsync(F->tick_b, 0, ((size_t)ns + 1) * sizeof(unsigned), st));
    if (fin.empty()) fin.push_back({0, 0});
    if ((rc = upload(F, &F->fin, fin))) return rc;
    if (sea.empty()) sea.push_back({0, 0, 0, 0});
    if ((rc = upload(F, &F->sea, sea))) return rc;
    if ((rc = upload(F, &F->fblk, fblk))) return rc;
    if ((rc = upload(F, &F->bblk, bblk))) return rc;
    if ((rc = upload(F, &F->panbase, panbase))) return rc;
    if ((rc = F->mem.alloc(&F->fflag, (size_t)panbase[ns] + 1))) return rc;
    if ((rc = F->mem.alloc(&F->bflag, (size_t)panbase[ns] + 1))) return rc;
    SEHIP_CHECK(hipMemsetAsync(F->fflag, 0, ((size_t)panbase[ns] + 1) * sizeof(int), st));
    SEHIP_CHECK(hipMemsetAsync(F->bflag, 0, ((size_t)panbase[ns] + 1) * sizeof(int), st));
    trace_phase("single-vector solve plan", tp);
    F->plan_ready = true;
    return 0;
}

}  // namespace

int factor_create(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, const FactorOptions &o,
                  hipStream_t st, Factor **out, speigs_hip_info *hinfo)
{
    std::unique_ptr<Factor> F(new Factor());
    F->opt = o;
    F->st = st;
    SEHIP_CHECK(hipStreamCreate(&F->st2));
    SEHIP_CHECK(hipEventCreateWithFlags(&F->ev_ready, hipEventDisableTiming));
    SEHIP_CHECK(hipEventCreateWithFlags(&F->ev_far, hipEventDisableTiming));
    SEHIP_CHECK(hipEventCreateWithFlags(&F->ev_zero, hipEventDisableTiming));
    F->n = n;
    F->fp32 = o.precision == SPEIGS_HIP_FP32_IR || o.precision == SPEIGS_HIP_FP32_BLOCK;
    // Single-vector solves: by the inverted diagonal blocks (mode 3) when the
    // caller's tolerance leaves room for one such solve per apply (apply_tol
    // >= 1e-14, i.e. an eigensolver tolerance >= 1e-12: ~3x faster per apply);
    // for tighter ones each apply would need a refinement step, and the
    // chained blocked solve (mode 1), exact to rounding in one pass, is kept
    F->solve_mode = o.apply_tol >= 1e-14 ? 3 : 1;
    // debug knobs for A/B checks of kernel variants
    if (const char *e = getenv("SPEIGS_HIP_SOLVE_MODE")) F->solve_mode = atoi(e);
    if (const char *e = getenv("SPEIGS_HIP_MFMA")) F->use_mfma = atoi(e) != 0;
    if (const char *e = getenv("SPEIGS_HIP_REG_PANEL")) F->use_reg_panel = atoi(e) != 0;
    // 1 or 4: the unblocked panel kernel on that many waves; anything else: blocked
    if (const char *e = getenv("SPEIGS_HIP_POTRF_WAVES")) {
        F->potrf_waves = atoi(e) == 1 ? 1 : 4;
        F->potrf_blocked = atoi(e) != 1 && atoi(e) != 4;
    }
    // test knob: every bounded wait of the blocked single-vector solve gives up
    if (const char *e = getenv("SPEIGS_HIP_WAIT_TEST")) F->wait_test = atoi(e) == 1;
    // A's CSC arrays do not depend on the ordering: a helper thread uploads
    // them while this one orders (the top of the dissection leaves cores free)
    int dev = 0;
    SEHIP_CHECK(hipGetDevice(&dev));
    {
        int r = 0;
        if ((r = F->mem.alloc(&F->Ax, std::max<size_t>((size_t)Ap[n], 1))) ||
            (r = F->mem.alloc(&F->Apd, (size_t)n + 1)) || (r = F->mem.alloc(&F->Aid, std::max<size_t>((size_t)Ap[n], 1))))
            return r;
    }
    int up_rc = 0;
    std::string up_msg;
    std::thread up_th([&] {
        auto run = [&]() -> int {
            SEHIP_CHECK(hipSetDevice(dev));
            int r = 0;
            if (Ap[n] > 0 && ((r = copy_to_device(F->Ax, Ax, (size_t)Ap[n] * sizeof(double), st)) ||
                              (r = copy_to_device(F->Aid, Ai, (size_t)Ap[n] * sizeof(int64_t), st))))
                return r;
            return copy_to_device(F->Apd, Ap, ((size_t)n + 1) * sizeof(int64_t), st);
        };
        up_rc = run();
        if (up_rc) up_msg = last_error_text();
    });
    struct Join {
        std::thread &t;
        ~Join()
        {
            if (t.joinable()) t.join();
        }
    } join_up{up_th};
    SymbolicOptions so;
    int rc = symbolic_analyze(n, Ap, Ai, so, &F->S);
    if (rc) return rc;
    const Symbolic &S = F->S;
    double tu = wall(), tp = tu;
    // the front pool, zeroed on the second stream while this thread builds the
    // assembly data and task lists and the arrays go up on the first (the
    // first numeric phase waits for it)
    if (F->fp32) rc = F->mem.alloc((float **)&F->pool, (size_t)S.pool);
    else rc = F->mem.alloc((double **)&F->pool, (size_t)S.pool);
    if (rc) return rc;
    if ((uintptr_t)F->pool % 16) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "front pool not 16-byte aligned");
    {
        const unsigned zg = (unsigned)std::min<int64_t>(4096, (S.pool * (F->fp32 ? 4 : 8) / 16 + 255) / 256 + 1);
        if (F->fp32) k_zero<float><<<dim3(zg), dim3(256), 0, F->st2>>>((float *)F->pool, S.pool);
        else k_zero<double><<<dim3(zg), dim3(256), 0, F->st2>>>((double *)F->pool, S.pool);
        SEHIP_CHECK(hipGetLastError());
        SEHIP_CHECK(hipEventRecord(F->ev_zero, F->st2));
        F->pool_zeroed = true;
    }
    trace_phase("pool allocation, zero fill queued", tp);
    // per-supernode arrays
    const int64_t ns = S.ns;
    std::vector<int32_t> Fs(ns), nc(ns), rows32(std::max<size_t>(S.rows.size(), 1)), chl32(std::max<size_t>(S.chl.size(), 1));
    std::vector<int64_t> wsoff(ns + 1, 0);
    for (int64_t s = 0; s < ns; s++) {
        if (S.F[s] >= ((int64_t)1 << 31) / 2) return SPEIGS_ERR_MEM;
        Fs[s] = (int32_t)S.F[s];
        nc[s] = (int32_t)(S.c1[s] - S.c0[s]);
        wsoff[s + 1] = wsoff[s] + S.F[s];
    }
    const int nt = host_threads();
    parallel_for((int64_t)S.rows.size(), nt, 65536, [&](int64_t i, int) { rows32[i] = (int32_t)S.rows[i]; });
    for (size_t i = 0; i < S.chl.size(); i++) chl32[i] = (int32_t)S.chl[i];
    F->nnzA = Ap[n];
    // The arrays that are final by now go to the device (A's CSC arrays
    // straight from the caller's).  Copying them on a helper thread while the task
    // lists below are built was measured: the serial list building ran twice as
    // slowly beside the staging copies, for no net gain.
    struct Early {
        void *d;
        const void *h;
        size_t bytes;
    };
    std::vector<Early> early;
    auto stage = [&](auto **d, const auto *h, size_t count) -> int {
        static_assert(sizeof(**d) == sizeof(*h), "device and host element types differ");
        const int r = F->mem.alloc(d, std::max<size_t>(count, 1));
        if (r == 0 && count > 0) early.push_back({(void *)*d, (const void *)h, count * sizeof(**d)});
        return r;
    };
    if ((rc = stage(&F->iperm, S.iperm.data(), S.iperm.size()))) return rc;
    if ((rc = stage(&F->foff, S.foff.data(), S.foff.size()))) return rc;
    if ((rc = stage(&F->wsoff, wsoff.data(), wsoff.size()))) return rc;
    if ((rc = stage(&F->col0, S.c0.data(), S.c0.size()))) return rc;
    if ((rc = stage(&F->rptr, S.rptr.data(), S.rptr.size()))) return rc;
    if ((rc = stage(&F->relptr, S.relptr.data(), S.relptr.size()))) return rc;
    if ((rc = stage(&F->chptr, S.chptr.data(), S.chptr.size()))) return rc;
    if ((rc = stage(&F->Fs, Fs.data(), Fs.size()))) return rc;
    if ((rc = stage(&F->ncol, nc.data(), nc.size()))) return rc;
    if ((rc = stage(&F->rows, rows32.data(), rows32.size()))) return rc;
    if ((rc = stage(&F->rel, S.rel.data(), S.rel.size()))) return rc;
    if ((rc = stage(&F->chl, chl32.data(), chl32.size()))) return rc;
    if ((rc = stage(&F->perm, S.perm.data(), S.perm.size()))) return rc;
    // the extend-add's block boundaries: ceil(F_parent / 64) + 1 per child
    std::vector<int64_t> bndptr(ns + 1, 0);
    for (int64_t s2 = 0; s2 < ns; s2++)
        bndptr[s2 + 1] = bndptr[s2] + (S.parent[s2] >= 0 ? (S.F[S.parent[s2]] + 63) / 64 + 1 : 0);
    if ((rc = stage(&F->bndptr, bndptr.data(), bndptr.size()))) return rc;
    if ((rc = F->mem.alloc(&F->bnd, (size_t)std::max<int64_t>(bndptr[ns], 1)))) return rc;
    up_th.join();
    if (up_rc) {
        note_error(up_msg.c_str());
        return up_rc;
    }
    for (const Early &e : early)
        if ((rc = copy_to_device(e.d, e.h, e.bytes, st))) return rc;
    // the supernode of each column and A's diagonal, once
    if ((rc = F->mem.alloc(&F->snode, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->diagA, (size_t)n))) return rc;
    k_snode_fill<<<dim3((unsigned)((ns + 3) / 4)), dim3(256), 0, st>>>(F->col0, F->ncol, ns, F->snode);
    if (S.relptr[ns] > 0)
        k_ea_bounds<<<dim3((unsigned)((S.relptr[ns] + 255) / 256)), dim3(256), 0, st>>>(F->relptr, F->rel, F->bndptr,
                                                                                       ns, F->bnd);
    SEHIP_CHECK(hipMemsetAsync(F->diagA, 0, (size_t)n * sizeof(double), st));
    k_diag_values<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(F->Apd, F->Aid, F->Ax, F->iperm, n,
                                                                          F->diagA);
    trace_phase("assembly data", tp);
    // plans: extend-add tasks, tile tasks, fronts per level.  The levels'
    // lists are built in parallel, each with offsets into lists of its own,
    // then placed one after another (the offsets moved by the lengths of the
    // levels before; front ordinals and tile prefixes are per step already)
    std::vector<int32_t> fronts;
    std::vector<std::vector<int64_t>> bylevel(S.nlevels);
    for (int64_t s = 0; s < ns; s++) bylevel[S.level[s]].push_back(s);
    F->plan.resize(S.nlevels);
    for (int32_t L = 0; L < S.nlevels; L++) {
        F->plan[L].fr_off = (int64_t)fronts.size();
        for (int64_t s : bylevel[L]) fronts.push_back((int32_t)s);
        F->plan[L].fr_cnt = (int64_t)bylevel[L].size();
    }
    struct LevelLists {
        std::vector<EaTask> ea;
        std::vector<TileTask> tiles;
        std::vector<int32_t> ufid;
        std::vector<int64_t> upref;
        int64_t xinv_max = 1;
    };
    std::vector<LevelLists> LL(S.nlevels);
    parallel_for(S.nlevels, nt, 1, [&](int64_t L, int) {
        LevelPlan &P = F->plan[L];
        std::vector<EaTask> &ea = LL[L].ea;
        std::vector<TileTask> &tiles = LL[L].tiles;
        std::vector<int32_t> &ufid = LL[L].ufid;
        std::vector<int64_t> &upref = LL[L].upref;
        int64_t &xinv_max = LL[L].xinv_max;
        int64_t maxch = 0, maxpan = 0;
        for (int64_t s : bylevel[L]) {
            maxch = std::max(maxch, S.chptr[s + 1] - S.chptr[s]);
            maxpan = std::max<int64_t>(maxpan, (S.c1[s] - S.c0[s] + TB - 1) / TB);
        }
        for (int64_t o2 = 0; o2 < maxch; o2++) {
            P.ea_off.push_back((int64_t)ea.size());
            for (int64_t s : bylevel[L]) {
                if (S.chptr[s] + o2 >= S.chptr[s + 1]) continue;
                const int64_t ch = S.chl[S.chptr[s] + o2];
                const int64_t rc2 = S.F[ch] - (S.c1[ch] - S.c0[ch]);
                for (int64_t cb = 0; cb * 64 < rc2; cb++) ea.push_back({(int32_t)ch, (int32_t)s, (int32_t)cb, 0});
            }
            P.ea_cnt.push_back((int64_t)ea.size() - P.ea_off.back());
        }
        for (int64_t k = 0; k < maxpan; k++) {
            P.k_potrf_off.push_back((int64_t)tiles.size());
            for (int64_t s : bylevel[L])
                if ((S.c1[s] - S.c0[s]) > k * TB) tiles.push_back({(int32_t)s, (int32_t)k, (int32_t)k, (int32_t)k});
            P.k_potrf_cnt.push_back((int64_t)tiles.size() - P.k_potrf_off.back());
            xinv_max = std::max(xinv_max, std::min(kXinvSlots, P.k_potrf_cnt.back()));
            P.k_trsm_off.push_back((int64_t)tiles.size());
            P.k_trsm_cut.emplace_back();
            int64_t ord = 0;     // the front's potrf task ordinal in this step; its inverse's slot = ord % kXinvSlots
            for (int64_t s : bylevel[L]) {
                const int64_t c = S.c1[s] - S.c0[s], Fv = S.F[s];
                if (c <= k * TB) continue;
                if (ord % kXinvSlots == 0) P.k_trsm_cut.back().push_back((int64_t)tiles.size());
                const int64_t kend = k * TB + std::min<int64_t>(TB, c - k * TB);
                const int64_t nt = (Fv + TB - 1) / TB;
                for (int64_t i = k; i < nt; i++)
                    if (std::max(i * TB, kend) < std::min(Fv, (i + 1) * TB))
                        tiles.push_back({(int32_t)s, (int32_t)k, (int32_t)i, (int32_t)(ord % kXinvSlots)});
                ord++;
            }
            P.k_trsm_cut.back().push_back((int64_t)tiles.size());
            P.k_trsm_cnt.push_back((int64_t)tiles.size() - P.k_trsm_off.back());
            // update tiles, per front (in POTRF order) its tile count and ordinal:
            // inner = columns [kend, bend) of the panel's block, outer = columns
            // [bend, F) after the block's last panel (see upd_decode)
            P.k_upd_off.push_back((int64_t)ufid.size());
            P.k_upref_off.push_back((int64_t)upref.size());
            P.k_out_off.push_back((int64_t)ufid.size());
            int64_t acc = 0, acco = 0;
            int32_t f = 0;
            std::vector<int64_t> ocnt, omv;
            for (int64_t s : bylevel[L]) {
                const int64_t c = S.c1[s] - S.c0[s], Fv = S.F[s];
                if (c <= k * TB) continue;
                const int64_t kend = k * TB + std::min<int64_t>(TB, c - k * TB);
                const int64_t nt = (Fv + TB - 1) / TB;
                const int64_t bend = std::min<int64_t>((k / KB + 1) * KB * TB, c);
                const int64_t j0 = kend < std::min(Fv, (k + 1) * TB) ? k : k + 1;
                const int64_t j1 = (bend + TB - 1) / TB;
                int64_t cnt = 0;
                for (int64_t j = j0; j < j1; j++) cnt += nt - j;
                upref.push_back(acc);
                ufid.insert(ufid.end(), (size_t)cnt, f);
                acc += cnt;
                int64_t oc = 0, om = 0;
                if (k == j1 - 1 && bend < Fv) {
                    om = nt - bend / TB;          // tile columns of the outer range
                    oc = om * (om + 1) / 2;
                }
                ocnt.push_back(oc);
                omv.push_back(om);
                f++;
            }
            upref.push_back(acc);
            P.k_upd_cnt.push_back(acc);
            // the outer tiles of this step follow in ufid, their prefix in upref:
            // first the next block's tile columns, then (a list of their own)
            // the columns right of it
            P.k_out_off.back() = (int64_t)ufid.size();
            P.k_outref_off.push_back((int64_t)upref.size());
            std::vector<int64_t> ofar(ocnt.size(), 0);
            for (size_t q = 0; q < ocnt.size(); q++) {
                if (ocnt[q] == 0) continue;
                const int64_t m = omv[q], kk = std::min<int64_t>(KB, m);
                ofar[q] = ocnt[q] - (kk * m - kk * (kk - 1) / 2);
                ocnt[q] -= ofar[q];
            }
            for (size_t q = 0; q < ocnt.size(); q++) {
                upref.push_back(acco);
                ufid.insert(ufid.end(), (size_t)ocnt[q], (int32_t)q);
                acco += ocnt[q];
            }
            upref.push_back(acco);
            P.k_out_cnt.push_back(acco);
            P.k_out2_off.push_back((int64_t)ufid.size());
            P.k_out2ref_off.push_back((int64_t)upref.size());
            int64_t accf = 0;
            for (size_t q = 0; q < ofar.size(); q++) {
                upref.push_back(accf);
                ufid.insert(ufid.end(), (size_t)ofar[q], (int32_t)q);
                accf += ofar[q];
            }
            upref.push_back(accf);
            P.k_out2_cnt.push_back(accf);
        }
    });
    int64_t xinv_max = 1;
    std::vector<int64_t> b_ea(S.nlevels + 1, 0), b_tl(S.nlevels + 1, 0), b_uf(S.nlevels + 1, 0), b_ur(S.nlevels + 1, 0);
    for (int32_t L = 0; L < S.nlevels; L++) {
        b_ea[L + 1] = b_ea[L] + (int64_t)LL[L].ea.size();
        b_tl[L + 1] = b_tl[L] + (int64_t)LL[L].tiles.size();
        b_uf[L + 1] = b_uf[L] + (int64_t)LL[L].ufid.size();
        b_ur[L + 1] = b_ur[L] + (int64_t)LL[L].upref.size();
        xinv_max = std::max(xinv_max, LL[L].xinv_max);
        LevelPlan &P = F->plan[L];
        for (int64_t &x : P.ea_off) x += b_ea[L];
        for (int64_t &x : P.k_potrf_off) x += b_tl[L];
        for (int64_t &x : P.k_trsm_off) x += b_tl[L];
        for (auto &cut : P.k_trsm_cut)
            for (int64_t &x : cut) x += b_tl[L];
        for (int64_t &x : P.k_upd_off) x += b_uf[L];
        for (int64_t &x : P.k_out_off) x += b_uf[L];
        for (int64_t &x : P.k_out2_off) x += b_uf[L];
        for (int64_t &x : P.k_upref_off) x += b_ur[L];
        for (int64_t &x : P.k_outref_off) x += b_ur[L];
        for (int64_t &x : P.k_out2ref_off) x += b_ur[L];
    }
    std::vector<EaTask, NoInitAlloc<EaTask>> ea(b_ea[S.nlevels]);
    std::vector<TileTask, NoInitAlloc<TileTask>> tiles(b_tl[S.nlevels]);
    std::vector<int32_t, NoInitAlloc<int32_t>> ufid(b_uf[S.nlevels]);
    std::vector<int64_t, NoInitAlloc<int64_t>> upref(b_ur[S.nlevels]);
    parallel_for(S.nlevels, nt, 1, [&](int64_t L, int) {
        std::copy(LL[L].ea.begin(), LL[L].ea.end(), ea.begin() + b_ea[L]);
        std::copy(LL[L].tiles.begin(), LL[L].tiles.end(), tiles.begin() + b_tl[L]);
        std::copy(LL[L].ufid.begin(), LL[L].ufid.end(), ufid.begin() + b_uf[L]);
        std::copy(LL[L].upref.begin(), LL[L].upref.end(), upref.begin() + b_ur[L]);
        LL[L] = LevelLists();
    });
    trace_phase("factor task lists", tp);
    if (trace_on())
        for (int32_t L = 0; L < S.nlevels; L++) {
            int64_t sc = 0, sr = 0, mf = 0;
            for (int64_t s : bylevel[L]) {
                sc += S.c1[s] - S.c0[s];
                sr += S.F[s] - (S.c1[s] - S.c0[s]);
                mf = std::max<int64_t>(mf, S.F[s]);
            }
            fprintf(stderr, "[speigs_hip] level %2d: %6zu fronts, sum c %8lld, sum r %9lld, max F %6lld\n", (int)L,
                    bylevel[L].size(), (long long)sc, (long long)sr, (long long)mf);
        }
    std::vector<EaBlk> eab;
    for (int32_t L = 0; L < S.nlevels; L++) {
        LevelPlan &P = F->plan[L];
        P.pcb_off = (int64_t)eab.size();
        for (int64_t s2 : bylevel[L])
            if (S.chptr[s2 + 1] > S.chptr[s2])
                for (int64_t cb = 0; cb * 64 < S.F[s2]; cb++)
                    for (int64_t rb = cb / 4; rb * 256 < S.F[s2]; rb++)   // the lower part: rows >= 64 cb
                        eab.push_back({(int32_t)s2, (int32_t)cb, (int32_t)rb, 0});
        P.pcb_cnt = (int64_t)eab.size() - P.pcb_off;
    }
    if (eab.empty()) eab.push_back({0, 0, 0, 0});
    if ((rc = upload(F.get(), &F->eab, eab))) return rc;
    if (ea.empty()) ea.push_back({0, 0, 0, 0});
    if (tiles.empty()) tiles.push_back({0, 0, 0, 0});
    // the chained single-vector solves' plan (modes 0, 1): now, unless the
    // factor is the block eigensolver's or solves by the inverted L11 (mode 3),
    // whose solves have plans of their own; then at the first factor_apply, if any
    if (o.precision != SPEIGS_HIP_FP32_BLOCK && F->solve_mode != 3 && (rc = single_solve_plan(F.get()))) return rc;
    if ((rc = F->mem.alloc(&F->counters, (size_t)2 * S.nlevels + 2))) return rc;
    if ((rc = F->mem.alloc(&F->err, 1))) return rc;
    SEHIP_CHECK(hipMemsetAsync(F->err, 0, sizeof(int), st));
    if (hipHostMalloc((void **)&F->herr, sizeof(int)) != hipSuccess) return SPEIGS_ERR_MEM;
    *F->herr = 0;
    F->ev.resize(2 * kEvPairs);
    for (auto &e : F->ev)
        SEHIP_CHECK(hipEventCreate(&e));
    trace_phase("small uploads, events", tp);
    // uploads
    if (F->fp32) {
        if ((rc = F->mem.alloc((float **)&F->xinv, (size_t)xinv_max * TB * TB))) return rc;
    } else {
        if ((rc = F->mem.alloc((double **)&F->xinv, (size_t)xinv_max * TB * TB))) return rc;
    }
    trace_phase("inverse slots", tp);
    if ((rc = upload(F.get(), &F->fronts, fronts))) return rc;
    if ((rc = upload(F.get(), &F->ea, ea))) return rc;
    if ((rc = upload(F.get(), &F->tiles, tiles))) return rc;
    if (ufid.empty()) ufid.push_back(0);
    if ((rc = upload(F.get(), &F->ufid, ufid))) return rc;
    if ((rc = upload(F.get(), &F->upref, upref))) return rc;
    if ((rc = F->mem.alloc(&F->flag, 4))) return rc;
    F->wssize = wsoff[ns];
    if ((rc = F->mem.alloc(&F->ws, (size_t)std::max<int64_t>(F->wssize, 1)))) return rc;
    if ((rc = F->mem.alloc(&F->xp, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->yp, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->r, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->d, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->ax, (size_t)n))) return rc;
    if ((rc = F->mem.alloc(&F->partial, 256))) return rc;
    if (hipHostMalloc((void **)&F->hpart, 256 * sizeof(double)) != hipSuccess) return SPEIGS_ERR_MEM;
    SEHIP_CHECK(hipStreamSynchronize(st));
    trace_phase("uploads", tp);
    if (hinfo) {
        hinfo->t_order = S.t_order;
        hinfo->t_symbolic = S.t_symbolic;
        hinfo->t_upload = wall() - tu;
    }
    // the single-vector solve's plan is built by a helper thread while the
    // device factorizes (solve mode 3; the block path starts the same plan)
    if (F->solve_mode == 3) rowsolve_prepare_async(F.get());
    *out = F.release();
    return SPEIGS_OK;
}

namespace {
// a solve mode 3 factor of the single-vector paths: plan joined and L11
// inverted right after the factorization, so that no apply allocates
int vec_after_numeric(Factor *F)
{
    if (F->solve_mode != 3 || F->opt.precision == SPEIGS_HIP_FP32_BLOCK) return 0;
    return factor_vec_prepare(F);
}

// one solve with the factor: the inverse-based one of solve mode 3 (also for
// any factor whose L11 is inverted), else the blocked or per-front chains
int one_solve(Factor *F, const double *b, double *x)
{
    if (F->solve_mode == 3 || (F->inverted && F->solve_mode != 2)) {
        const int rc = factor_vec_solve(F, b, x);
        if (rc == 0) F->solves++;
        return rc;
    }
    if (F->fp32) run_solve<float>(F, b, x);
    else run_solve<double>(F, b, x);
    return 0;
}
}  // namespace

int factor_numeric(Factor *F, double delta)
{
    F->delta = delta;
    F->inverted = false;
    F->signed_fac = false;
    int st = F->fp32 ? run_numeric<float>(F, delta) : run_numeric<double>(F, delta);
    F->factored = (st == 0);
    if (st == 0) {
        const int rc = vec_after_numeric(F);
        if (rc) return rc;
    }
    return st;
}

int factor_numeric_signed(Factor *F, double tau)
{
    F->delta = 0.0;
    F->inverted = false;
    F->factored = false;
    if (!(tau > 0.0)) return SPEIGS_ERR_ARG;
    if (!F->sgn) {
        const size_t cnt = (size_t)std::max<int64_t>(F->n, 1);
        int rc;
        if (F->fp32) {
            float *p = nullptr;
            rc = F->mem.alloc(&p, cnt);
            F->sgn = p;
        } else {
            double *p = nullptr;
            rc = F->mem.alloc(&p, cnt);
            F->sgn = p;
        }
        if (rc) return rc;
    }
    // the signed kernels: register panel (4 waves), MFMA update; the solve
    // scales by S between its sweeps (modes 1 and 3)
    F->use_reg_panel = true;
    F->use_mfma = true;
    F->potrf_waves = 4;
    if (F->solve_mode == 2 || F->solve_mode == 0) F->solve_mode = 1;
    F->signed_fac = true;
    F->tau = tau;
    int st = F->fp32 ? run_numeric<float>(F, 0.0) : run_numeric<double>(F, 0.0);
    if (st == 1) return SEHIP_FAIL(SPEIGS_ERR_FACTOR, "signed factorization: a pivot is not a number");
    if (st < 0) return st;
    F->factored = true;
    return vec_after_numeric(F);
}

int factor_perturbed(const Factor *F) { return F->perturbed; }

void factor_attach_matrix(Factor *F, const SellA *A) { F->A = A; }

bool factor_apply_preallocated(const Factor *F)
{
    if (F->opt.precision == SPEIGS_HIP_FP32_BLOCK || F->solve_mode == 2) return false;
    return F->solve_mode != 3 || factor_vec_prepared(F);
}

int factor_check(const Factor *F)
{
    if (F->herr && *F->herr)
        return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "solve: a bounded wait between work items gave up");
    return 0;
}

int factor_apply(Factor *F, const double *x, double *y)
{
    if (!F->factored) return SPEIGS_ERR_FACTOR;
    int rc0 = factor_check(F);   // a bounded wait gave up in an earlier solve
    if (rc0) return rc0;
    const bool vec = F->solve_mode == 3 || (F->inverted && F->solve_mode != 2);
    if (!vec && (rc0 = single_solve_plan(F))) return rc0;
    if (F->solve_mode == 2 && !F->inverted) {
        const int rc = factor_invert_l11(F);
        if (rc) return rc;
    }
    ev_harvest(F, false);
    auto solve = [&](const double *b, double *out) -> int {
        const int e = ev_begin(F);
        const int rc = one_solve(F, b, out);
        ev_end(F, e);
        return rc;
    };
    // FP64 solves by the inverted L11 blocks: refined against A to apply_tol
    const bool vec64 = vec && !F->fp32 && !F->signed_fac;
    if (!F->fp32 && !F->signed_fac && !vec64) return solve(x, y);
    // FP32 factor, a signed factor (pivots may have been replaced) or the
    // inverse-based FP64 solve: iterative refinement on the FP64 residual of
    // (A + delta I)
    if ((rc0 = solve(x, y))) return rc0;
    if (!F->A) return 0;
    const double bn = device_norm(F, x);
    if (bn < 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "refinement: norm of the right-hand side");
    if (bn == 0.0) return 0;
    // The signed factor stands in for the reference's LU with its refinement
    // steps, whose eigenpairs are judged at tol = 1e-14 ||A||: its applies are
    // refined until the residual stops falling (to the rounding level), not
    // only to ir_tol.  A singular A stops after one step.  The FP64 solve by
    // inverted blocks is refined to apply_tol, which the eigensolver sets from
    // its own tolerance, the same way.
    const double tol = vec64 ? F->opt.apply_tol : F->signed_fac ? F->opt.ir_tol * 1e-3 : F->opt.ir_tol;
    const bool stall_stop = F->signed_fac || vec64;
    double prev = -1.0;
    for (int it = 0; it < F->opt.ir_maxit; it++) {
        spmv(*F->A, y, F->ax, F->st);
        axpby_residual(x, F->ax, y, F->delta, F->r, F->n, F->st);
        const double rn = device_norm(F, F->r);
        if (rn < 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "refinement: norm of the residual");
        if (rn <= tol * bn) break;
        if (stall_stop && prev >= 0.0 && rn > 0.5 * prev) break;
        prev = rn;
        if ((rc0 = solve(F->r, F->d))) return rc0;
        vadd(y, F->d, F->n, F->st);
        F->ir_steps++;
    }
    return 0;
}

void factor_stats(const Factor *Fc, speigs_hip_info *h)
{
    if (!h) return;
    Factor *F = const_cast<Factor *>(Fc);
    ev_harvest(F, true);
    h->t_solve = F->t_solve;
    h->nnz_L = F->S.nnzL;
    h->factor_flops = F->S.flops;
    h->nsuper = F->S.ns;
    h->nlevels = F->S.nlevels;
    h->ir_steps = F->ir_steps;
    h->solves = F->solves;
    h->precision_used = F->opt.precision == SPEIGS_HIP_FP32_BLOCK ? SPEIGS_HIP_FP32_BLOCK : F->fp32 ? SPEIGS_HIP_FP32_IR : SPEIGS_HIP_FP64;
    const double es = F->fp32 ? 4.0 : 8.0;
    h->bytes_solve = 2.0 * (double)F->S.nnzL * es + 2.0 * 2.0 * (double)F->n * 8.0;
    h->t_invert = F->t_invert;
    if (F->block_solves > 0) {
        // multi-right-hand-side solves: factor read once per sweep, FP32 vectors in and out
        const double nb = (double)F->block_rhs / (double)F->block_solves;
        h->solves = F->block_solves;
        h->bytes_solve = 2.0 * (double)F->S.nnzL * es + 2.0 * nb * (double)F->n * 4.0;
    }
}

void factor_destroy(Factor *F)
{
    if (F) factor_block_free(F);
    delete F;
}

const Symbolic *factor_symbolic(const Factor *F) { return &F->S; }

}  // namespace sehip
