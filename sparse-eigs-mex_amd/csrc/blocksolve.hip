// blocksolve.hip -- chain-free multi-right-hand-side solves with the supernodal
// factor: each front's diagonal block L11 is inverted in place after the
// factorization, so a front's forward solve is two GEMM-shaped products
//     Y = L11^-1 W1,   U = W2 - L21 Y
// and its backward solve two more
//     Z = Y - L21^T X_R,   X1 = L11^-T Z,
// with no sequential dependence inside a front; only the tree levels remain.
//
// In-place inverse (lower triangular, 64-column tiles, from the last tile up):
//     X_JJ = L_JJ^-1 ;  X_IJ = -(sum_{J<K<=I} X_IK L_KJ) X_JJ   for I > J,
// where X_IK (K > J) are the already-inverted trailing tiles.  Per step J the
// diagonal tiles of all fronts of a level are inverted in LDS, the products
// T_I = sum_K X_IK L_KJ go to scratch, then X_IJ = -T_I X_JJ overwrite L_IJ.
//
// The block solve serves as the preconditioner of the block eigensolver: the
// FP32 factor and FP32 vectors give ~kappa u_32 accuracy, which the outer
// iteration (residuals against A in FP64) absorbs.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "factor_impl.h"
#include "mfma.h"

namespace sehip {

// ------------------------------------------------------------------ inverse
struct InvTask {
    int32_t s, J, I, pad;
};

// tile (I, K) of front s, rows/cols clipped to the L11 order c
__device__ __forceinline__ int tile_w(int c, int K) { return min(TB, c - K * TB); }

// inverse of the lower-triangular diagonal tile, recursively blocked in LDS:
// the four 16 x 16 diagonal blocks by forward substitution (a thread per
// column), then two merges  X21 = -X22 (L21 X11)  (16 -> 32 -> 64), each product
// spread over all threads.  A partial tile (w < 64) is padded with the identity.
template <class T>
__global__ void __launch_bounds__(256) k_inv_diag(T *pool, const InvTask *__restrict__ tasks,
                                                  const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                  const int32_t *__restrict__ ncol)
{
    __shared__ T a[TB][TB + 1];
    __shared__ T x[TB][TB + 1];
    __shared__ T tmp[TB / 2][TB / 2 + 1];
    const InvTask t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    const int k0 = t.J * TB, w = tile_w(c, t.J);
    T *base = pool + foff[t.s] + (int64_t)k0 * F + k0;
    T va[TB * TB / 256];      // all of a thread's loads in flight before its LDS stores
#pragma unroll
    for (int u = 0; u < TB * TB / 256; u++) {
        const int e = threadIdx.x + 256 * u, r = e % TB, q = e / TB;
        T v = T(0);
        if (r < w && q < w) v = r >= q ? base[(int64_t)q * F + r] : T(0);
        else if (r == q) v = T(1);
        va[u] = v;
    }
#pragma unroll
    for (int u = 0; u < TB * TB / 256; u++) {
        const int e = threadIdx.x + 256 * u, r = e % TB, q = e / TB;
        a[r][q] = va[u];
        x[r][q] = T(0);
    }
    __syncthreads();
    if (threadIdx.x < TB) {                 // 16 x 16 diagonal blocks, one column per thread
        const int j = threadIdx.x, b1 = (j / 16) * 16 + 16;
        for (int i = j; i < b1; i++) {
            T s = (i == j) ? T(1) : T(0);
            for (int k = j; k < i; k++) s -= a[i][k] * x[k][j];
            x[i][j] = s / a[i][i];
        }
    }
    __syncthreads();
    for (int h = 16; h < TB; h *= 2) {      // merge blocks of size h into 2h
        const int nb = TB / (2 * h);        // pairs
        // tmp = L21 X11 for every pair p: rows [p2h + h, p2h + 2h), cols [p2h, p2h + h)
        for (int e = threadIdx.x; e < nb * h * h; e += 256) {
            const int p = e / (h * h), rr = (e / h) % h, cc = e % h;
            const int o = p * 2 * h;
            T s = T(0);
            for (int k = 0; k < h; k++) s += a[o + h + rr][o + k] * x[o + k][o + cc];
            tmp[p * h + rr][cc] = s;
        }
        __syncthreads();
        // X21 = -X22 tmp
        for (int e = threadIdx.x; e < nb * h * h; e += 256) {
            const int p = e / (h * h), rr = (e / h) % h, cc = e % h;
            const int o = p * 2 * h;
            T s = T(0);
            for (int k = 0; k < h; k++) s += x[o + h + rr][o + h + k] * tmp[p * h + k][cc];
            x[o + h + rr][o + cc] = -s;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < w * w; e += 256) {
        const int r = e % w, q = e / w;
        if (r >= q) base[(int64_t)q * F + r] = x[r][q];
    }
}

// Recursive doubling over the 64-column tiles of L11: with the inverses X11,
// X22 of two adjacent diagonal blocks of b tiles known,
//     inv([L11 0; L21 L22]) = [X11 0; -X22 L21 X11  X22],
// formed as T = L21 X11 (scratch) then X21 = -X22 T (over L21).  Level b
// merges every pair of b-tile blocks of every front at once (two launches), so
// the whole inverse takes 2 log2(tiles) launches after the diagonal tiles.
// Task: output tile (i, j) of the pair whose second block starts at tile row
// R1 and whose first block starts at tile C1, with b tiles per block; T of the
// pair at scratch tiles tslot + i b + j (column-major 64 x 64).
struct InvG {
    int32_t s, b, R1, C1, i, j;
    int64_t tslot;
};

constexpr int IP = TB + 4;   // LDS row pitch of the inverse kernels

// MODE 0: T(i, j) = sum_{k=j}^{b-1} L(R1+i, C1+k) X(C1+k, C1+j)
// MODE 1: L(R1+i, C1+j) = -sum_{k=0}^{i} X(R1+i, R1+k) T(k, j)
// (X diagonal tiles lower triangular; tiles clipped to the order c of L11)
template <class T, int MODE>
__global__ void __launch_bounds__(256) k_inv_gemm(T *pool, const InvG *__restrict__ tasks,
                                                  const int64_t *__restrict__ foff, const int32_t *__restrict__ Fs,
                                                  const int32_t *__restrict__ ncol, T *__restrict__ scratch)
{
    __shared__ T At[TB][IP];   // At[kk][row] = A(row, kk)
    __shared__ T Bt[TB][IP];   // Bt[kk][col] = B(kk, col)
    typedef typename MfmaAcc<T>::V V;
    const InvG t = tasks[blockIdx.x];
    const int32_t F = Fs[t.s], c = ncol[t.s];
    T *fr = pool + foff[t.s];
    const int ti = t.R1 + t.i;                   // output tile row (absolute)
    const int tj = t.C1 + t.j;                   // output tile column (absolute)
    const int wi = tile_w(c, ti), wj = tile_w(c, tj);
    const int kbeg = MODE == 0 ? t.j : 0, kend = MODE == 0 ? t.b : t.i + 1;
    V acc[2][2];
    zero_acc<T>(acc);
    for (int k = kbeg; k < kend; k++) {
        const int tk = (MODE == 0 ? t.C1 : t.R1) + k;   // the inner tile (absolute)
        const int wk = tile_w(c, tk);
        if (wk <= 0) break;
        // all 32 loads of a thread in flight before its LDS stores
        T va[TB * TB / 256], vb[TB * TB / 256];
#pragma unroll
        for (int u = 0; u < TB * TB / 256; u++) {
            const int e = threadIdx.x + 256 * u, r = e & 63, q = e >> 6;
            T av = T(0), bv = T(0);
            if (MODE == 0) {
                // A = L(ti, tk)(r, q); B = X(tk, tj)(r, q), lower on the diagonal tile
                if (r < wi && q < wk) av = fr[(int64_t)(tk * TB + q) * F + ti * TB + r];
                if (r < wk && q < wj && (k > t.j || r >= q)) bv = fr[(int64_t)(tj * TB + q) * F + tk * TB + r];
            } else {
                // A = X(ti, tk)(r, q), lower on the diagonal tile; B = T(k, j)(r, q)
                if (r < wi && q < wk && (k < t.i || r >= q)) av = fr[(int64_t)(tk * TB + q) * F + ti * TB + r];
                bv = scratch[(t.tslot + (int64_t)k * t.b + t.j) * (TB * TB) + q * TB + r];
            }
            va[u] = av;
            vb[u] = bv;
        }
        __syncthreads();          // the previous tile's products are done
#pragma unroll
        for (int u = 0; u < TB * TB / 256; u++) {
            const int e = threadIdx.x + 256 * u, r = e & 63, q = e >> 6;
            At[q][r] = va[u];
            Bt[r][q] = vb[u];
        }
        __syncthreads();
        tile_mma<T, IP>(At, Bt, (wk + 3) & ~3, acc);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rb = (wv & 1) * 32, cb = (wv >> 1) * 32, lr = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = rb + 16 * a + lr, col = cb + 16 * b + MfmaAcc<T>::colof(lk, v);
                if (MODE == 0)
                    scratch[(t.tslot + (int64_t)t.i * t.b + t.j) * (TB * TB) + col * TB + row] =
                        (row < wi && col < wj) ? acc[a][b][v] : T(0);
                else if (row < wi && col < wj)
                    fr[(int64_t)(tj * TB + col) * F + ti * TB + row] = -acc[a][b][v];
            }
}

template <class T>
static int invert_impl(Factor *F)
{
    const Symbolic &S = F->S;
    hipStream_t st = F->st;
    T *pool = (T *)F->pool;
    // every diagonal tile of every front, then the merge levels
    std::vector<InvTask> diag;
    int64_t nbmax = 0;
    for (int64_t s = 0; s < S.ns; s++) {
        const int64_t nb = (S.c1[s] - S.c0[s] + TB - 1) / TB;
        nbmax = std::max(nbmax, nb);
        for (int32_t J = 0; J < nb; J++) diag.push_back({(int32_t)s, J, J, 0});
    }
    if (diag.empty()) return 0;
    std::vector<InvG> g;
    struct Level { int64_t off, cnt; };
    std::vector<Level> lev;
    int64_t tmax = 1;
    for (int64_t b = 1; b < nbmax; b *= 2) {
        Level L{(int64_t)g.size(), 0};
        int64_t tslot = 0;
        for (int64_t s = 0; s < S.ns; s++) {
            const int64_t nb = (S.c1[s] - S.c0[s] + TB - 1) / TB;
            for (int64_t c1 = 0; c1 + b < nb; c1 += 2 * b) {
                const int64_t r1 = c1 + b, b2 = std::min(b, nb - r1);
                for (int64_t i = 0; i < b2; i++)
                    for (int64_t j = 0; j < b; j++)
                        g.push_back({(int32_t)s, (int32_t)b, (int32_t)r1, (int32_t)c1, (int32_t)i, (int32_t)j, tslot});
                tslot += b2 * b;
            }
        }
        L.cnt = (int64_t)g.size() - L.off;
        if (L.cnt) lev.push_back(L);
        tmax = std::max(tmax, tslot);
    }
    if (g.empty()) g.push_back({0, 1, 0, 0, 0, 0, 0});
    DevMem tmp;
    InvTask *d_diag = nullptr;
    InvG *d_g = nullptr;
    T *scratch = nullptr;
    int rc;
    if ((rc = tmp.alloc(&d_diag, diag.size()))) return rc;
    if ((rc = tmp.alloc(&d_g, g.size()))) return rc;
    if ((rc = tmp.alloc(&scratch, (size_t)tmax * TB * TB))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(d_diag, diag.data(), diag.size() * sizeof(InvTask), hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipMemcpyAsync(d_g, g.data(), g.size() * sizeof(InvG), hipMemcpyHostToDevice, st));
    k_inv_diag<T><<<dim3((unsigned)diag.size()), dim3(256), 0, st>>>(pool, d_diag, F->foff, F->Fs, F->ncol);
    for (const Level &L : lev) {
        k_inv_gemm<T, 0><<<dim3((unsigned)L.cnt), dim3(256), 0, st>>>(pool, d_g + L.off, F->foff, F->Fs, F->ncol,
                                                                     scratch);
        k_inv_gemm<T, 1><<<dim3((unsigned)L.cnt), dim3(256), 0, st>>>(pool, d_g + L.off, F->foff, F->Fs, F->ncol,
                                                                     scratch);
    }
    SEHIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

int factor_invert_l11(Factor *F)
{
    if (!F->factored) return SPEIGS_ERR_FACTOR;
    if (F->inverted) return 0;
    const double t0 = wall();
    int rc = F->fp32 ? invert_impl<float>(F) : invert_impl<double>(F);
    F->t_invert += wall() - t0;
    if (rc == 0) F->inverted = true;
    return rc;
}

// ------------------------------------------------------------------ block solve
// Vectors: FP32 (or FP64) n x nb column-major with leading dimension n (new
// order), nb <= 32; per front a workspace of F x nb (ld F).
constexpr int NBMAX = 32;

struct BsArgs {
    const int64_t *foff, *wsoff, *col0, *rptr, *relptr, *chptr;
    const int32_t *Fs, *ncol, *rows, *rel, *chl;
};

struct RowTask {   // front s, 64-row tile I (of L11 when I < ceil(c/64), else of the update part)
    int32_t s, I;
};

__device__ __forceinline__ int lb_i32(const int32_t *a, int n, int v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// W_s rows [ck*256, +256) = [B(c0..c0+c) ; 0] + children's U (children in order).
// Each thread batches its nb loads into registers before the adds and stores
// (the workspace regions of parent and children are distinct, so the loads of
// one child may all be in flight at once).
template <class V>
__global__ void __launch_bounds__(256) k_bs_init(const RowTask *__restrict__ tasks, BsArgs a, const V *__restrict__ B,
                                                 int64_t n, int nb, V *__restrict__ ws)
{
    const RowTask t = tasks[blockIdx.x];
    const int s = t.s;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    V *w = ws + a.wsoff[s] * NBMAX;
    const int64_t c0 = a.col0[s];
    const int r0 = t.I * 256, r1 = min(F, r0 + 256);
    const int i = r0 + threadIdx.x;
    if (i < r1) {
        V x[NBMAX];
#pragma unroll
        for (int v = 0; v < NBMAX; v++) x[v] = (v < nb && i < c) ? B[(int64_t)v * n + c0 + i] : V(0);
#pragma unroll
        for (int v = 0; v < NBMAX; v++)
            if (v < nb) w[(int64_t)v * F + i] = x[v];
    }
    for (int64_t q = a.chptr[s]; q < a.chptr[s + 1]; q++) {
        __syncthreads();
        const int ch = a.chl[q];
        const int32_t Fc = a.Fs[ch], cc = a.ncol[ch], rc = Fc - cc;
        const int32_t *rl = a.rel + a.relptr[ch];
        const int lo = lb_i32(rl, rc, r0), hi = lb_i32(rl, rc, r1);
        const V *u = ws + a.wsoff[ch] * NBMAX;
        for (int e = lo + threadIdx.x; e < hi; e += 256) {
            const int dst = rl[e];
            V x[NBMAX], y[NBMAX];
#pragma unroll
            for (int v = 0; v < NBMAX; v++) {
                x[v] = v < nb ? u[(int64_t)v * Fc + cc + e] : V(0);
                y[v] = v < nb ? w[(int64_t)v * F + dst] : V(0);
            }
#pragma unroll
            for (int v = 0; v < NBMAX; v++)
                if (v < nb) w[(int64_t)v * F + dst] = y[v] + x[v];
        }
    }
}

// generic tile-row product on one front (vectors n x nb, column-major, ld n):
//   acc[r][v] (64 x nb) = sum over K-tiles of op(A)[r][k] * Xk[k][v]
// MODE 0 forward L11:    A = X11 tile row I, tiles K <= I (lower on K == I), Xk = W1 (ws) -> Y (global)
// MODE 1 forward L21:    A = L21 tile row I (rows c + 64I ..), all K-tiles of L11, Xk = Y -> W2 -= acc (ws)
// MODE 2 backward L21^T: A = L21 column tile I, K over update tiles, Xk = X at rows R -> Z = Y - acc (ws rows < c)
// MODE 3 backward L11^T: A = X11 column tile I, tiles K >= I (lower on K == I), Xk = Z (ws) -> X (global)
template <class T, class V, int MODE>
__global__ void __launch_bounds__(256) k_bs_tile(const T *__restrict__ pool, const RowTask *__restrict__ tasks,
                                                 BsArgs a, int64_t n, int nb, V *ws, V *Yg, V *Xg)
{
    __shared__ V At[TB][TB + 1];     // At[r][k] = op(A)[r][k]
    __shared__ V Xs[TB][NBMAX + 1];  // Xs[k][v]
    const RowTask t = tasks[blockIdx.x];
    const int s = t.s, I = t.I;
    const int32_t F = a.Fs[s], c = a.ncol[s];
    const int nt11 = (c + TB - 1) / TB;
    const int nupd = (F - c + TB - 1) / TB;
    const T *L = pool + a.foff[s];
    V *w = ws + a.wsoff[s] * NBMAX;
    const int64_t c0 = a.col0[s];
    const int32_t *R = a.rows + a.rptr[s];
    int orow0, on;
    if (MODE == 1) { orow0 = c + I * TB; on = min(TB, F - orow0); }
    else { orow0 = I * TB; on = min(TB, c - orow0); }
    int Kbeg, Kend;
    if (MODE == 0) { Kbeg = 0; Kend = I + 1; }
    else if (MODE == 1) { Kbeg = 0; Kend = nt11; }
    else if (MODE == 2) { Kbeg = 0; Kend = nupd; }
    else { Kbeg = I; Kend = nt11; }
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    V acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = V(0);
    for (int K = Kbeg; K < Kend; K++) {
        int krow0, kn;
        if (MODE == 2) { krow0 = c + K * TB; kn = min(TB, F - krow0); }
        else { krow0 = K * TB; kn = min(TB, c - krow0); }
        __syncthreads();
        for (int e = threadIdx.x; e < TB * TB; e += 256) {
            const int p = e % TB, q = e / TB;   // p fastest: consecutive rows of one column
            if (MODE == 0 || MODE == 1) {       // A[r][k] = L(orow0 + r, krow0 + k): column krow0+q, row orow0+p
                V av = V(0);
                if (p < on && q < kn && (MODE == 1 || K < I || p >= q)) av = (V)L[(int64_t)(krow0 + q) * F + orow0 + p];
                At[p][q] = av;
            } else {                            // op(A)[r][k] = L(krow0 + k, orow0 + r): column orow0+q, row krow0+p
                V av = V(0);
                if (p < kn && q < on && (MODE == 2 || K > I || p >= q)) av = (V)L[(int64_t)(orow0 + q) * F + krow0 + p];
                At[q][p] = av;
            }
        }
        for (int e = threadIdx.x; e < TB * nb; e += 256) {
            const int kk = e % TB, v = e / TB;
            V xv = V(0);
            if (kk < kn) {
                if (MODE == 0 || MODE == 3) xv = w[(int64_t)v * F + krow0 + kk];
                else if (MODE == 1) xv = Yg[(int64_t)v * n + c0 + krow0 + kk];
                else xv = Xg[(int64_t)v * n + R[krow0 - c + kk]];
            }
            Xs[kk][v] = xv;
        }
        __syncthreads();
        for (int k = 0; k < kn; k++) {
            const V av = At[r][k];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int v = g + 4 * j;
                if (v < nb) acc[j] = fma(av, Xs[k][v], acc[j]);
            }
        }
    }
    if (r < on) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int v = g + 4 * j;
            if (v >= nb) continue;
            if (MODE == 0) Yg[(int64_t)v * n + c0 + orow0 + r] = acc[j];
            else if (MODE == 1) w[(int64_t)v * F + orow0 + r] -= acc[j];
            else if (MODE == 2) w[(int64_t)v * F + orow0 + r] = Yg[(int64_t)v * n + c0 + orow0 + r] - acc[j];
            else Xg[(int64_t)v * n + c0 + orow0 + r] = acc[j];
        }
    }
}

template <class V>
__global__ void k_bs_perm(const V *__restrict__ src, const int64_t *__restrict__ perm, V *__restrict__ dst, int64_t n,
                          int nb, int gather)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t p = perm[i];
    for (int v = 0; v < nb; v++) {
        if (gather) dst[(int64_t)v * n + i] = src[(int64_t)v * n + p];
        else dst[(int64_t)v * n + p] = src[(int64_t)v * n + i];
    }
}

namespace {

struct BsPlan {
    std::vector<int64_t> init_off, init_cnt, m0_off, m0_cnt, m1_off, m1_cnt;
};

template <class V>
struct BsState {
    bool ready = false;
    DevMem mem;
    RowTask *tasks = nullptr;
    BsPlan plan;
    V *ws = nullptr, *Pg = nullptr, *Yg = nullptr, *Xg = nullptr;
};

template <class V>
int bs_prepare(Factor *F, BsState<V> &B)
{
    if (B.ready) return 0;
    const Symbolic &S = F->S;
    std::vector<std::vector<int64_t>> bylevel(S.nlevels);
    for (int64_t s = 0; s < S.ns; s++) bylevel[S.level[s]].push_back(s);
    std::vector<RowTask> tk;
    for (int32_t L = 0; L < S.nlevels; L++) {
        B.plan.init_off.push_back((int64_t)tk.size());
        for (int64_t s : bylevel[L])
            for (int64_t ck = 0; ck * 256 < S.F[s]; ck++) tk.push_back({(int32_t)s, (int32_t)ck});
        B.plan.init_cnt.push_back((int64_t)tk.size() - B.plan.init_off.back());
        B.plan.m0_off.push_back((int64_t)tk.size());   // tiles of L11 (modes 0, 2, 3)
        for (int64_t s : bylevel[L])
            for (int64_t I = 0; I * TB < S.c1[s] - S.c0[s]; I++) tk.push_back({(int32_t)s, (int32_t)I});
        B.plan.m0_cnt.push_back((int64_t)tk.size() - B.plan.m0_off.back());
        B.plan.m1_off.push_back((int64_t)tk.size());   // tiles of the update rows (mode 1)
        for (int64_t s : bylevel[L]) {
            const int64_t r = S.F[s] - (S.c1[s] - S.c0[s]);
            for (int64_t I = 0; I * TB < r; I++) tk.push_back({(int32_t)s, (int32_t)I});
        }
        B.plan.m1_cnt.push_back((int64_t)tk.size() - B.plan.m1_off.back());
    }
    if (tk.empty()) tk.push_back({0, 0});
    int rc;
    if ((rc = B.mem.alloc(&B.tasks, tk.size()))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(B.tasks, tk.data(), tk.size() * sizeof(RowTask), hipMemcpyHostToDevice, F->st));
    if ((rc = B.mem.alloc(&B.ws, (size_t)std::max<int64_t>(F->wssize, 1) * NBMAX))) return rc;
    if ((rc = B.mem.alloc(&B.Pg, (size_t)F->n * NBMAX))) return rc;
    if ((rc = B.mem.alloc(&B.Yg, (size_t)F->n * NBMAX))) return rc;
    if ((rc = B.mem.alloc(&B.Xg, (size_t)F->n * NBMAX))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(F->st));
    B.ready = true;
    return 0;
}

template <class T, class V>
int bs_run(Factor *F, BsState<V> &B, const V *Bin, V *Xout, int nb)
{
    const Symbolic &S = F->S;
    hipStream_t st = F->st;
    const T *pool = (const T *)F->pool;
    BsArgs a{F->foff, F->wsoff, F->col0, F->rptr, F->relptr, F->chptr, F->Fs, F->ncol, F->rows, F->rel, F->chl};
    const unsigned nbk = (unsigned)((F->n + 255) / 256);
    k_bs_perm<V><<<dim3(nbk), dim3(256), 0, st>>>(Bin, F->perm, B.Pg, F->n, nb, 1);
    for (int32_t L = 0; L < S.nlevels; L++) {
        const BsPlan &P = B.plan;
        if (P.init_cnt[L])
            k_bs_init<V><<<dim3((unsigned)P.init_cnt[L]), dim3(256), 0, st>>>(B.tasks + P.init_off[L], a, B.Pg, F->n,
                                                                              nb, B.ws);
        if (P.m0_cnt[L])
            k_bs_tile<T, V, 0><<<dim3((unsigned)P.m0_cnt[L]), dim3(256), 0, st>>>(pool, B.tasks + P.m0_off[L], a, F->n,
                                                                                  nb, B.ws, B.Yg, B.Xg);
        if (P.m1_cnt[L])
            k_bs_tile<T, V, 1><<<dim3((unsigned)P.m1_cnt[L]), dim3(256), 0, st>>>(pool, B.tasks + P.m1_off[L], a, F->n,
                                                                                  nb, B.ws, B.Yg, B.Xg);
    }
    for (int32_t L = S.nlevels - 1; L >= 0; L--) {
        const BsPlan &P = B.plan;
        if (P.m0_cnt[L]) {
            k_bs_tile<T, V, 2><<<dim3((unsigned)P.m0_cnt[L]), dim3(256), 0, st>>>(pool, B.tasks + P.m0_off[L], a, F->n,
                                                                                  nb, B.ws, B.Yg, B.Xg);
            k_bs_tile<T, V, 3><<<dim3((unsigned)P.m0_cnt[L]), dim3(256), 0, st>>>(pool, B.tasks + P.m0_off[L], a, F->n,
                                                                                  nb, B.ws, B.Yg, B.Xg);
        }
    }
    k_bs_perm<V><<<dim3(nbk), dim3(256), 0, st>>>(B.Xg, F->perm, Xout, F->n, nb, 0);
    return 0;
}

}  // namespace

int factor_block_solve_d(Factor *F, const double *Bin, double *Xout, int nb)
{
    if (nb < 1 || nb > NBMAX) return SPEIGS_ERR_ARG;
    if (!F->inverted) return SPEIGS_ERR_FACTOR;
    if (!F->bs_d) F->bs_d = new BsState<double>();
    BsState<double> &B = *reinterpret_cast<BsState<double> *>(F->bs_d);
    int rc = bs_prepare(F, B);
    if (rc) return rc;
    return F->fp32 ? bs_run<float, double>(F, B, Bin, Xout, nb) : bs_run<double, double>(F, B, Bin, Xout, nb);
}

void factor_block_free(Factor *F)
{
    delete reinterpret_cast<BsState<double> *>(F->bs_d);
    F->bs_d = nullptr;
    rowsolve_free(F);
}

}  // namespace sehip
