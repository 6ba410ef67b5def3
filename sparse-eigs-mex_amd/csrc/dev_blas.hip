// dev_blas.hip -- bandwidth-bound kernels of the Lanczos loop on gfx950:
// SELL-64 SpMV (reference op_multA -> cholmod_l_sdmult, src/speigs_core.c:75-83),
// the CGS2 block GEMV passes (orth_block, :122-132), the Ritz-vector GEMM
// (dgemv per Ritz vector at :250 and :278) and the fused residual reductions
// (:246-264).
//
// All reductions are deterministic: each 1024-row block writes one partial per
// column, then a fixed-order tree sums the partials.  Blocks are 256 threads
// (4 wave64s); wave reductions use cross-lane shuffles, block reductions LDS.
#include "dev.h"
#include "host_par.h"

#include <algorithm>
#include <atomic>
#include <memory>
#include <cstring>
#include <ctime>

namespace sehip {

int hip_error_code(hipError_t e)
{
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SPEIGS_ERR_MEM;
    return SPEIGS_ERR_INTERNAL;
}

namespace {
thread_local char g_last_error[512];

const char *base_name(const char *path)
{
    const char *b = path;
    for (const char *p = path; *p; p++)
        if (*p == '/') b = p + 1;
    return b;
}
}  // namespace

int fail(int code, const char *file, int line, const char *what)
{
    snprintf(g_last_error, sizeof g_last_error, "%s:%d: %s (code %d)", base_name(file), line, what, code);
    if (trace_on()) fprintf(stderr, "[speigs_hip] error %s\n", g_last_error);
    return code;
}

int hip_fail(hipError_t e, const char *file, int line, const char *call)
{
    (void)hipGetLastError();
    const int code = hip_error_code(e);
    snprintf(g_last_error, sizeof g_last_error, "%s:%d: %s -> %s (code %d)", base_name(file), line, call,
             hipGetErrorName(e), code);
    if (trace_on()) fprintf(stderr, "[speigs_hip] error %s\n", g_last_error);
    return code;
}

const char *last_error_text() { return g_last_error; }

void clear_error() { g_last_error[0] = 0; }

void note_error(const char *text)
{
    if (text && text != g_last_error) snprintf(g_last_error, sizeof g_last_error, "%s", text);
}

double wall()
{
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

bool trace_on()
{
    static const bool on = [] {
        const char *e = getenv("SPEIGS_HIP_TRACE");
        return e && atoi(e) > 0;
    }();
    return on;
}

void trace_phase(const char *what, double &t)
{
    const double now = wall();
    if (trace_on()) fprintf(stderr, "[speigs_hip] %-22s %8.2f ms\n", what, 1e3 * (now - t));
    t = now;
}

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum NV accumulators over the block; thread t < NV gets total t in `out`
template <int NV>
__device__ __forceinline__ void block_sum(double (&acc)[NV], double *lds /*[4*NV]*/, double *out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double s = wave_sum(acc[i]);
        if (lane == 0) lds[wv * NV + i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int i = threadIdx.x;
        *out = ((lds[i] + lds[NV + i]) + (lds[2 * NV + i] + lds[3 * NV + i]));
    }
    __syncthreads();
}

// ------------------------------------------------------------------ SELL SpMV
__global__ void __launch_bounds__(256) k_sell_spmv(int64_t nslices, const int64_t *__restrict__ sptr,
                                                   const int32_t *__restrict__ swid,
                                                   const int32_t *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ x, double *__restrict__ y,
                                                   int64_t n)
{
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t sl = row >> 6;
    if (sl >= nslices) return;
    const int lane = threadIdx.x & 63;
    int64_t p = sptr[sl] + lane;
    const int w = swid[sl];
    double s = 0.0;
    for (int k = 0; k < w; k++, p += 64) s = fma(val[p], x[col[p]], s);
    if (row < n) y[row] = s;
}

template <int NV>
__global__ void __launch_bounds__(256) k_sell_spmm(int64_t nslices, const int64_t *__restrict__ sptr,
                                                   const int32_t *__restrict__ swid,
                                                   const int32_t *__restrict__ col,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ X, int64_t ldx,
                                                   double *__restrict__ Y, int64_t ldy, int64_t n, int nv)
{
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t sl = row >> 6;
    if (sl >= nslices) return;
    const int lane = threadIdx.x & 63;
    int64_t p = sptr[sl] + lane;
    const int w = swid[sl];
    double s[NV];
#pragma unroll
    for (int v = 0; v < NV; v++) s[v] = 0.0;
    for (int k = 0; k < w; k++, p += 64) {
        const double a = val[p];
        const int64_t c = col[p];
#pragma unroll
        for (int v = 0; v < NV; v++)
            if (v < nv) s[v] = fma(a, X[c + v * ldx], s[v]);
    }
    if (row < n) {
#pragma unroll
        for (int v = 0; v < NV; v++)
            if (v < nv) Y[row + v * ldy] = s[v];
    }
}

int sell_build(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, SellA *S,
               DevMem &mem, hipStream_t st)
{
    // Host construction in parallel over the stored columns.  Each row of the
    // full symmetric matrix lists its entries in increasing column order (ties
    // by value), whatever the order of the input, so the SpMV's summation
    // order is a function of the matrix alone.  It runs beside the numeric
    // factorization, whose launching thread must keep a core: half the
    // analysis threads.
    const int nt = std::max(1, host_threads() / 2);
    std::unique_ptr<std::atomic<int32_t>[]> rl(new std::atomic<int32_t>[n]);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { rl[i].store(0, std::memory_order_relaxed); });
    parallel_for(n, nt, 4096, [&](int64_t j, int) {
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i > j) continue;
            rl[i].fetch_add(1, std::memory_order_relaxed);
            if (i != j) rl[j].fetch_add(1, std::memory_order_relaxed);
        }
    });
    const int64_t nsl = (n + 63) / 64;
    std::vector<int64_t> sptr(nsl + 1, 0);
    std::vector<int32_t> swid(nsl, 0);
    parallel_for(nsl, nt, 1024, [&](int64_t s, int) {
        int32_t w = 0;
        for (int64_t r = s * 64; r < std::min(n, s * 64 + 64); r++) w = std::max(w, rl[r].load(std::memory_order_relaxed));
        swid[s] = w;
    });
    for (int64_t s = 0; s < nsl; s++) sptr[s + 1] = sptr[s] + 64 * (int64_t)swid[s];
    const int64_t tot = sptr[nsl];
    std::vector<int32_t> col(tot);
    std::vector<double> val(tot);
    // padding points at the row itself with value 0
    parallel_for(nsl, nt, 256, [&](int64_t s, int) {
        for (int64_t k = 0; k < swid[s]; k++)
            for (int64_t lane = 0; lane < 64; lane++) {
                col[sptr[s] + k * 64 + lane] = (int32_t)std::min(n - 1, s * 64 + lane);
                val[sptr[s] + k * 64 + lane] = 0.0;
            }
    });
    std::unique_ptr<std::atomic<int32_t>[]> cur(new std::atomic<int32_t>[n]);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { cur[i].store(0, std::memory_order_relaxed); });
    auto put = [&](int64_t r, int64_t c, double v) {
        const int64_t k = cur[r].fetch_add(1, std::memory_order_relaxed);
        const int64_t pos = sptr[r >> 6] + k * 64 + (r & 63);
        col[pos] = (int32_t)c;
        val[pos] = v;
    };
    parallel_for(n, nt, 4096, [&](int64_t j, int) {
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i > j) continue;
            put(i, j, Ax[q]);
            if (i != j) put(j, i, Ax[q]);
        }
    });
    std::atomic<int64_t> nnz_sum{0};
    parallel_for(n, nt, 4096, [&](int64_t r, int) {
        const int64_t base = sptr[r >> 6] + (r & 63), len = rl[r].load(std::memory_order_relaxed);
        for (int64_t a = 1; a < len; a++) {       // insertion sort of the row by (column, value)
            const int32_t c = col[base + a * 64];
            const double v = val[base + a * 64];
            int64_t b = a - 1;
            while (b >= 0 && (col[base + b * 64] > c || (col[base + b * 64] == c && val[base + b * 64] > v))) {
                col[base + (b + 1) * 64] = col[base + b * 64];
                val[base + (b + 1) * 64] = val[base + b * 64];
                b--;
            }
            col[base + (b + 1) * 64] = c;
            val[base + (b + 1) * 64] = v;
        }
        nnz_sum.fetch_add(len, std::memory_order_relaxed);
    });
    const int64_t nnz_full = nnz_sum.load();
    S->n = n;
    S->nslices = nsl;
    S->stored = tot;
    S->nnz_full = nnz_full;
    int rc;
    if ((rc = mem.alloc(&S->sptr, nsl + 1))) return rc;
    if ((rc = mem.alloc(&S->swid, nsl))) return rc;
    if ((rc = mem.alloc(&S->col, tot))) return rc;
    if ((rc = mem.alloc(&S->val, tot))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(S->sptr, sptr.data(), (nsl + 1) * 8, hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipMemcpyAsync(S->swid, swid.data(), nsl * 4, hipMemcpyHostToDevice, st));
    if ((rc = copy_to_device(S->col, col.data(), tot * 4, st))) return rc;
    if ((rc = copy_to_device(S->val, val.data(), tot * 8, st))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

void spmv(const SellA &S, const double *x, double *y, hipStream_t st)
{
    int64_t nb = (S.nslices * 64 + 255) / 256;
    k_sell_spmv<<<dim3((unsigned)nb), dim3(256), 0, st>>>(S.nslices, S.sptr, S.swid, S.col, S.val, x, y,
                                                         S.n);
}

void spmv_multi(const SellA &S, const double *X, int64_t ldx, double *Y, int64_t ldy, int nv,
                hipStream_t st)
{
    int64_t nb = (S.nslices * 64 + 255) / 256;
    for (int v0 = 0; v0 < nv; v0 += 4) {
        int c = std::min(4, nv - v0);
        if (c == 1)
            spmv(S, X + v0 * ldx, Y + v0 * ldy, st);
        else
            k_sell_spmm<4><<<dim3((unsigned)nb), dim3(256), 0, st>>>(S.nslices, S.sptr, S.swid, S.col,
                                                                    S.val, X + v0 * ldx, ldx,
                                                                    Y + v0 * ldy, ldy, S.n, c);
    }
}

double spmv_bytes(const SellA &S)
{
    // algorithmic: nnz of the full symmetric matrix (8 B value + 4 B index) + row pointers
    // + x read + y written
    return (double)S.nnz_full * 12.0 + (double)(S.n + 1) * 4.0 + 2.0 * (double)S.n * 8.0;
}

// ------------------------------------------------------------------ CGS2 passes
// Each block owns rows [b*1024, b*1024+1024); thread t holds rows b*1024 + r*256 + t.
__global__ void __launch_bounds__(256) k_coldots(const double *__restrict__ V, int64_t ldv, int m,
                                                 const double *__restrict__ w, int64_t n,
                                                 double *__restrict__ partial)
{
    __shared__ double lds[4 * 8];
    const int64_t base = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
    double wr[kRowsPerThread];
#pragma unroll
    for (int r = 0; r < kRowsPerThread; r++) {
        int64_t row = base + r * 256;
        wr[r] = row < n ? w[row] : 0.0;
    }
    for (int c0 = 0; c0 < m; c0 += 8) {
        double acc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = 0.0;
#pragma unroll
        for (int r = 0; r < kRowsPerThread; r++) {
            int64_t row = base + r * 256;
            if (row < n) {
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (c0 + c < m) acc[c] = fma(V[row + (int64_t)(c0 + c) * ldv], wr[r], acc[c]);
            }
        }
        double tot;
        block_sum<8>(acc, lds, &tot);
        if (threadIdx.x < 8 && c0 + (int)threadIdx.x < m)
            partial[(int64_t)blockIdx.x * m + c0 + threadIdx.x] = tot;
    }
}

__global__ void __launch_bounds__(256) k_reduce_partial(const double *__restrict__ partial, int64_t nb,
                                                        int m, double *__restrict__ out)
{
    __shared__ double lds[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) s += partial[b * m + c];
    lds[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = lds[0];
}

__global__ void __launch_bounds__(256) k_cgs_update(const double *__restrict__ V, int64_t ldv, int m,
                                                    const double *__restrict__ h, double *__restrict__ w,
                                                    int64_t n, double *__restrict__ partial, int mode)
{
    extern __shared__ double sh[];       // h[m] then 32 doubles of reduction scratch
    double *hs = sh;
    double *lds = sh + m;
    for (int c = threadIdx.x; c < m; c += 256) hs[c] = h[c];
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
    double wr[kRowsPerThread];
#pragma unroll
    for (int r = 0; r < kRowsPerThread; r++) {
        int64_t row = base + r * 256;
        if (row < n) {
            double s = 0.0;
            for (int c = 0; c < m; c++) s = fma(V[row + (int64_t)c * ldv], hs[c], s);
            double v = w[row] - s;
            w[row] = v;
            wr[r] = v;
        } else {
            wr[r] = 0.0;
        }
    }
    if (mode == 2) {
        double acc[1] = {0.0};
#pragma unroll
        for (int r = 0; r < kRowsPerThread; r++) acc[0] = fma(wr[r], wr[r], acc[0]);
        double tot;
        block_sum<1>(acc, lds, &tot);
        if (threadIdx.x == 0) partial[blockIdx.x] = tot;
    } else if (mode == 1) {
        for (int c0 = 0; c0 < m; c0 += 8) {
            double acc[8];
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = 0.0;
#pragma unroll
            for (int r = 0; r < kRowsPerThread; r++) {
                int64_t row = base + r * 256;
                if (row < n) {
#pragma unroll
                    for (int c = 0; c < 8; c++)
                        if (c0 + c < m) acc[c] = fma(V[row + (int64_t)(c0 + c) * ldv], wr[r], acc[c]);
                }
            }
            double tot;
            block_sum<8>(acc, lds, &tot);
            if (threadIdx.x < 8 && c0 + (int)threadIdx.x < m)
                partial[(int64_t)blockIdx.x * m + c0 + threadIdx.x] = tot;
        }
    }
}

void coldots_partial(const double *V, int64_t ldv, int m, const double *w, int64_t n, double *partial,
                     hipStream_t st)
{
    k_coldots<<<dim3((unsigned)nblocks_rows(n)), dim3(256), 0, st>>>(V, ldv, m, w, n, partial);
}

void reduce_partial(const double *partial, int64_t nb, int m, double *out, hipStream_t st)
{
    k_reduce_partial<<<dim3((unsigned)m), dim3(256), 0, st>>>(partial, nb, m, out);
}

void cgs_update(const double *V, int64_t ldv, int m, const double *h, double *w, int64_t n, double *partial,
                int mode, hipStream_t st)
{
    size_t shm = (size_t)(m + 32) * sizeof(double);
    k_cgs_update<<<dim3((unsigned)nblocks_rows(n)), dim3(256), shm, st>>>(V, ldv, m, h, w, n, partial, mode);
}

// ------------------------------------------------------------------ Ritz GEMM
// X[:, t] = V S[:, t] (+ beta X[:, t]): one row per thread, up to 32 outputs
// per pass (V re-read per pass), 4 rows of V in flight per step
template <int PT>
__global__ void __launch_bounds__(256) k_tsgemm(const double *__restrict__ V, int64_t ldv, int m,
                                                const double *__restrict__ S, int p, int t0,
                                                double *__restrict__ X, int64_t ldx, int64_t n, double beta)
{
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    double acc[PT];
#pragma unroll
    for (int t = 0; t < PT; t++) acc[t] = 0.0;
    const double *vr = V + row;
    int c = 0;
    for (; c + 4 <= m; c += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) v[u] = vr[(int64_t)(c + u) * ldv];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int t = 0; t < PT; t++)
                if (t0 + t < p) acc[t] = fma(v[u], S[c + u + (int64_t)(t0 + t) * m], acc[t]);
    }
    for (; c < m; c++) {
        const double v = vr[(int64_t)c * ldv];
#pragma unroll
        for (int t = 0; t < PT; t++)
            if (t0 + t < p) acc[t] = fma(v, S[c + (int64_t)(t0 + t) * m], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < PT; t++)
        if (t0 + t < p) {
            double *xp = X + row + (int64_t)(t0 + t) * ldx;
            *xp = beta == 0.0 ? acc[t] : acc[t] + beta * *xp;
        }
}

static void tsgemm_launch(const double *V, int64_t ldv, int m, const double *S, int p, double *X, int64_t ldx,
                          int64_t n, double beta, hipStream_t st)
{
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
    for (int t0 = 0; t0 < p; t0 += 32) {
        if (p - t0 <= 16)
            k_tsgemm<16><<<g, b, 0, st>>>(V, ldv, m, S, p, t0, X, ldx, n, beta);
        else
            k_tsgemm<32><<<g, b, 0, st>>>(V, ldv, m, S, p, t0, X, ldx, n, beta);
    }
}

void tsgemm(const double *V, int64_t ldv, int m, const double *S, int p, double *X, int64_t ldx, int64_t n,
            hipStream_t st)
{
    tsgemm_launch(V, ldv, m, S, p, X, ldx, n, 0.0, st);
}


// partial dots / residual norms over p column pairs
__global__ void __launch_bounds__(256) k_pair_partial(const double *__restrict__ X, int64_t ldx,
                                                      const double *__restrict__ Y, int64_t ldy,
                                                      const double *__restrict__ lam, int p, int64_t n,
                                                      double *__restrict__ partial)
{
    __shared__ double lds[4];
    const int64_t base = (int64_t)blockIdx.x * kRowsPerBlock + threadIdx.x;
    for (int t = 0; t < p; t++) {
        double acc[1] = {0.0};
        const double l = lam ? lam[t] : 0.0;
#pragma unroll
        for (int r = 0; r < kRowsPerThread; r++) {
            int64_t row = base + r * 256;
            if (row < n) {
                const double xv = X[row + (int64_t)t * ldx], yv = Y[row + (int64_t)t * ldy];
                if (lam) {
                    const double d = yv - l * xv;
                    acc[0] = fma(d, d, acc[0]);
                } else {
                    acc[0] = fma(xv, yv, acc[0]);
                }
            }
        }
        double tot;
        block_sum<1>(acc, lds, &tot);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * p + t] = tot;
    }
}

void pairdots_partial(const double *X, int64_t ldx, const double *Y, int64_t ldy, int p, int64_t n,
                      double *partial, hipStream_t st)
{
    k_pair_partial<<<dim3((unsigned)nblocks_rows(n)), dim3(256), 0, st>>>(X, ldx, Y, ldy, nullptr, p, n,
                                                                           partial);
}

void residual_partial(const double *X, int64_t ldx, const double *Y, int64_t ldy, const double *lam, int p,
                      int64_t n, double *partial, hipStream_t st)
{
    k_pair_partial<<<dim3((unsigned)nblocks_rows(n)), dim3(256), 0, st>>>(X, ldx, Y, ldy, lam, p, n,
                                                                           partial);
}

// ------------------------------------------------------------------ elementwise
__global__ void k_scale_cols(double *X, int64_t ldx, const double *s, int p, int64_t n)
{
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    for (int t = 0; t < p; t++) X[row + (int64_t)t * ldx] *= s[t];
}

__global__ void k_scale_copy(const double *x, double a, double *y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ void k_resid_shift(const double *b, const double *Ax, const double *x, double delta, double *r,
                              int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) r[i] = b[i] - Ax[i] - delta * x[i];
}

__global__ void k_vadd(double *x, const double *d, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) x[i] += d[i];
}

void scale_cols(double *X, int64_t ldx, const double *s, int p, int64_t n, hipStream_t st)
{
    k_scale_cols<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(X, ldx, s, p, n);
}

void scale_copy(const double *x, double a, double *y, int64_t n, hipStream_t st)
{
    k_scale_copy<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(x, a, y, n);
}

void axpby_residual(const double *b, const double *Ax, const double *x, double delta, double *r, int64_t n,
                    hipStream_t st)
{
    k_resid_shift<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(b, Ax, x, delta, r, n);
}

void vadd(double *x, const double *d, int64_t n, hipStream_t st)
{
    k_vadd<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(x, d, n);
}

}  // namespace sehip
