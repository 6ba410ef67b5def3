// batch.hip -- segmented vector kernels for the batched component solver: the
// rows of all components are one vector (component c owns rows [off_c,
// off_c + n_c)), cut into chunks of at most 1024 rows that never straddle two
// components.  One block per chunk; per-component parameters (active column
// count, coefficients, scales, target columns) are read from small arrays
// indexed by the chunk's component, so a single launch advances every
// component's Krylov recurrence.  Reductions go through per-chunk partials
// summed in chunk order: the result does not depend on scheduling.
#include <algorithm>

#include "batch.h"

namespace sehip {

namespace {

__device__ __forceinline__ double bwave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// sum of NV accumulators over the block; thread t < NV receives total t
template <int NV>
__device__ __forceinline__ void bblock_sum(double (&acc)[NV], double *lds, double *out)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++) {
        const double s = bwave_sum(acc[i]);
        if (lane == 0) lds[wv * NV + i] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        const int i = threadIdx.x;
        *out = (lds[i] + lds[NV + i]) + (lds[2 * NV + i] + lds[3 * NV + i]);
    }
    __syncthreads();
}

constexpr int RPT = 4;   // rows per thread (1024-row chunks)

// partial[q * ps + i] = sum over the chunk's rows of V[r, i] w[r], i < m_c
__global__ void __launch_bounds__(256) k_seg_dots(const double *__restrict__ V, int64_t ld,
                                                  const int32_t *__restrict__ mcomp, const double *__restrict__ w,
                                                  const SegChunk *__restrict__ chunks, double *__restrict__ partial,
                                                  int ps)
{
    __shared__ double lds[32];
    const SegChunk ch = chunks[blockIdx.x];
    const int m = mcomp[ch.comp];
    if (m <= 0) return;
    double wr[RPT];
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = ch.r0 + threadIdx.x + 256 * u;
        wr[u] = r < ch.r1 ? w[r] : 0.0;
    }
    for (int c0 = 0; c0 < m; c0 += 8) {
        double acc[8];
#pragma unroll
        for (int c = 0; c < 8; c++) acc[c] = 0.0;
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = ch.r0 + threadIdx.x + 256 * u;
            if (r < ch.r1) {
#pragma unroll
                for (int c = 0; c < 8; c++)
                    if (c0 + c < m) acc[c] = fma(V[r + (int64_t)(c0 + c) * ld], wr[u], acc[c]);
            }
        }
        double tot;
        bblock_sum<8>(acc, lds, &tot);
        if (threadIdx.x < 8 && c0 + (int)threadIdx.x < m) partial[(int64_t)blockIdx.x * ps + c0 + threadIdx.x] = tot;
    }
}

// out[c * ps + i] = sum over the chunks of c (in order) of partial[q * ps + i], i < cnt_c
__global__ void __launch_bounds__(64) k_seg_reduce(const double *__restrict__ partial, int ps,
                                                   const int64_t *__restrict__ cptr, const int32_t *__restrict__ cnt,
                                                   double *__restrict__ out)
{
    const int c = blockIdx.x;
    const int m = cnt[c];
    for (int i = threadIdx.x; i < m; i += 64) {
        double s = 0.0;
        for (int64_t q = cptr[c]; q < cptr[c + 1]; q++) s += partial[q * ps + i];
        out[(int64_t)c * ps + i] = s;
    }
}

// w -= V[:, 0:m_c] h_c for active components; then mode 1: partial dots
// V^T w (next pass), mode 2: partial |w|^2 in slot 0
__global__ void __launch_bounds__(256) k_seg_update(const double *__restrict__ V, int64_t ld,
                                                    const int32_t *__restrict__ act, const int32_t *__restrict__ mcomp,
                                                    const double *__restrict__ h, double *__restrict__ w,
                                                    const SegChunk *__restrict__ chunks, double *__restrict__ partial,
                                                    int ps, int mode)
{
    extern __shared__ double sh[];   // h_c[ps] then 32 doubles of reduction scratch
    double *hs = sh, *lds = sh + ps;
    const SegChunk ch = chunks[blockIdx.x];
    if (!act[ch.comp]) return;
    const int m = mcomp[ch.comp];
    for (int i = threadIdx.x; i < m; i += 256) hs[i] = h[(int64_t)ch.comp * ps + i];
    __syncthreads();
    double wr[RPT];
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = ch.r0 + threadIdx.x + 256 * u;
        if (r < ch.r1) {
            double s = 0.0;
            for (int i = 0; i < m; i++) s = fma(V[r + (int64_t)i * ld], hs[i], s);
            const double v = w[r] - s;
            w[r] = v;
            wr[u] = v;
        } else {
            wr[u] = 0.0;
        }
    }
    if (mode == 2) {
        double acc[1] = {0.0};
#pragma unroll
        for (int u = 0; u < RPT; u++) acc[0] = fma(wr[u], wr[u], acc[0]);
        double tot;
        bblock_sum<1>(acc, lds, &tot);
        if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * ps] = tot;
    } else if (mode == 1) {
        for (int c0 = 0; c0 < m; c0 += 8) {
            double acc[8];
#pragma unroll
            for (int c = 0; c < 8; c++) acc[c] = 0.0;
#pragma unroll
            for (int u = 0; u < RPT; u++) {
                const int64_t r = ch.r0 + threadIdx.x + 256 * u;
                if (r < ch.r1) {
#pragma unroll
                    for (int c = 0; c < 8; c++)
                        if (c0 + c < m) acc[c] = fma(V[r + (int64_t)(c0 + c) * ld], wr[u], acc[c]);
                }
            }
            double tot;
            bblock_sum<8>(acc, lds, &tot);
            if (threadIdx.x < 8 && c0 + (int)threadIdx.x < m)
                partial[(int64_t)blockIdx.x * ps + c0 + threadIdx.x] = tot;
        }
    }
}

// x[r] = V[r, j_c] (j_c >= 0), else 0
__global__ void __launch_bounds__(256) k_seg_gather_col(const double *__restrict__ V, int64_t ld,
                                                        const int32_t *__restrict__ jcol,
                                                        const SegChunk *__restrict__ chunks, double *__restrict__ x)
{
    const SegChunk ch = chunks[blockIdx.x];
    const int j = jcol[ch.comp];
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = ch.r0 + threadIdx.x + 256 * u;
        if (r < ch.r1) x[r] = j >= 0 ? V[r + (int64_t)j * ld] : 0.0;
    }
}

// V[r, j_c] = s_c x[r] for j_c >= 0
__global__ void __launch_bounds__(256) k_seg_store_col(const double *__restrict__ x, const double *__restrict__ s,
                                                       const int32_t *__restrict__ jdst,
                                                       const SegChunk *__restrict__ chunks, double *__restrict__ V,
                                                       int64_t ld)
{
    const SegChunk ch = chunks[blockIdx.x];
    const int j = jdst[ch.comp];
    if (j < 0) return;
    const double sc = s[ch.comp];
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = ch.r0 + threadIdx.x + 256 * u;
        if (r < ch.r1) V[r + (int64_t)j * ld] = sc * x[r];
    }
}

// X[r, t] = sum_{i < m_c} V[r, i] S_c[i + t m_c] for t < p_c (S_c at S + c * ss)
__global__ void __launch_bounds__(256) k_seg_gemm(const double *__restrict__ V, int64_t ld,
                                                  const int32_t *__restrict__ mcomp, const int32_t *__restrict__ pcomp,
                                                  const double *__restrict__ S, int64_t ss,
                                                  const SegChunk *__restrict__ chunks, double *__restrict__ X,
                                                  int64_t ldx)
{
    const SegChunk ch = chunks[blockIdx.x];
    const int m = mcomp[ch.comp], p = pcomp[ch.comp];
    if (p <= 0) return;
    const double *Sc = S + (int64_t)ch.comp * ss;
#pragma unroll
    for (int u = 0; u < RPT; u++) {
        const int64_t r = ch.r0 + threadIdx.x + 256 * u;
        if (r >= ch.r1) continue;
        for (int t0 = 0; t0 < p; t0 += 8) {
            double acc[8];
#pragma unroll
            for (int t = 0; t < 8; t++) acc[t] = 0.0;
            for (int i = 0; i < m; i++) {
                const double v = V[r + (int64_t)i * ld];
#pragma unroll
                for (int t = 0; t < 8; t++)
                    if (t0 + t < p) acc[t] = fma(v, Sc[i + (int64_t)(t0 + t) * m], acc[t]);
            }
#pragma unroll
            for (int t = 0; t < 8; t++)
                if (t0 + t < p) X[r + (int64_t)(t0 + t) * ldx] = acc[t];
        }
    }
}

// partial[q * ps + t] = sum X[r, t] Y[r, t]  (mode 0)  or  sum (Y - lam_c[t] X)^2 (mode 1), t < p_c
__global__ void __launch_bounds__(256) k_seg_pairs(const double *__restrict__ X, int64_t ldx,
                                                   const double *__restrict__ Y, int64_t ldy,
                                                   const int32_t *__restrict__ pcomp, const double *__restrict__ lam,
                                                   const SegChunk *__restrict__ chunks, double *__restrict__ partial,
                                                   int ps, int mode)
{
    __shared__ double lds[32];
    const SegChunk ch = chunks[blockIdx.x];
    const int p = pcomp[ch.comp];
    if (p <= 0) return;
    for (int t0 = 0; t0 < p; t0 += 8) {
        double acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = 0.0;
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = ch.r0 + threadIdx.x + 256 * u;
            if (r >= ch.r1) continue;
#pragma unroll
            for (int t = 0; t < 8; t++) {
                if (t0 + t >= p) continue;
                const double x = X[r + (int64_t)(t0 + t) * ldx], y = Y[r + (int64_t)(t0 + t) * ldy];
                if (mode == 0) {
                    acc[t] = fma(x, y, acc[t]);
                } else {
                    const double d = y - lam[(int64_t)ch.comp * ps + t0 + t] * x;
                    acc[t] = fma(d, d, acc[t]);
                }
            }
        }
        double tot;
        bblock_sum<8>(acc, lds, &tot);
        if (threadIdx.x < 8 && t0 + (int)threadIdx.x < p) partial[(int64_t)blockIdx.x * ps + t0 + threadIdx.x] = tot;
    }
}

// X[r, t] *= s_c[t], t < p_c
__global__ void __launch_bounds__(256) k_seg_scale_cols(double *__restrict__ X, int64_t ldx,
                                                        const int32_t *__restrict__ pcomp, const double *__restrict__ s,
                                                        int ps, const SegChunk *__restrict__ chunks)
{
    const SegChunk ch = chunks[blockIdx.x];
    const int p = pcomp[ch.comp];
    for (int t = 0; t < p; t++) {
        const double sc = s[(int64_t)ch.comp * ps + t];
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = ch.r0 + threadIdx.x + 256 * u;
            if (r < ch.r1) X[r + (int64_t)t * ldx] *= sc;
        }
    }
}

// D[r, i] = S[r, i] for i < cnt_c
__global__ void __launch_bounds__(256) k_seg_copy_cols(const double *__restrict__ Sv, double *__restrict__ D,
                                                       int64_t ld, const int32_t *__restrict__ cnt,
                                                       const SegChunk *__restrict__ chunks)
{
    const SegChunk ch = chunks[blockIdx.x];
    const int m = cnt[ch.comp];
    for (int i = 0; i < m; i++) {
#pragma unroll
        for (int u = 0; u < RPT; u++) {
            const int64_t r = ch.r0 + threadIdx.x + 256 * u;
            if (r < ch.r1) D[r + (int64_t)i * ld] = Sv[r + (int64_t)i * ld];
        }
    }
}

}  // namespace

int seg_layout_build(const std::vector<int64_t> &off, SegLayout &L, DevMem &mem, hipStream_t st)
{
    const int64_t nc = (int64_t)off.size() - 1;
    std::vector<SegChunk> ch;
    L.cptr.assign(nc + 1, 0);
    for (int64_t c = 0; c < nc; c++) {
        for (int64_t r = off[c]; r < off[c + 1]; r += 1024) ch.push_back({(int32_t)c, 0, r, std::min(off[c + 1], r + 1024)});
        L.cptr[c + 1] = (int64_t)ch.size();
    }
    if (ch.empty()) ch.push_back({0, 0, 0, 0});
    L.ncomp = nc;
    L.nchunks = (int64_t)ch.size();
    int rc;
    if ((rc = mem.alloc(&L.chunks, ch.size()))) return rc;
    if ((rc = mem.alloc(&L.cptr_d, L.cptr.size()))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(L.chunks, ch.data(), ch.size() * sizeof(SegChunk), hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipMemcpyAsync(L.cptr_d, L.cptr.data(), L.cptr.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    return 0;
}

void seg_dots(const SegLayout &L, const double *V, int64_t ld, const int32_t *mcomp, const double *w, double *partial,
              int ps, hipStream_t st)
{
    k_seg_dots<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(V, ld, mcomp, w, L.chunks, partial, ps);
}

void seg_reduce(const SegLayout &L, const double *partial, int ps, const int32_t *cnt, double *out, hipStream_t st)
{
    k_seg_reduce<<<dim3((unsigned)L.ncomp), dim3(64), 0, st>>>(partial, ps, L.cptr_d, cnt, out);
}

void seg_update(const SegLayout &L, const double *V, int64_t ld, const int32_t *act, const int32_t *mcomp,
                const double *h, double *w, double *partial, int ps, int mode, hipStream_t st)
{
    const size_t shm = (size_t)(ps + 32) * sizeof(double);
    k_seg_update<<<dim3((unsigned)L.nchunks), dim3(256), shm, st>>>(V, ld, act, mcomp, h, w, L.chunks, partial, ps,
                                                                   mode);
}

void seg_gather_col(const SegLayout &L, const double *V, int64_t ld, const int32_t *jcol, double *x, hipStream_t st)
{
    k_seg_gather_col<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(V, ld, jcol, L.chunks, x);
}

void seg_store_col(const SegLayout &L, const double *x, const double *s, const int32_t *jdst, double *V, int64_t ld,
                   hipStream_t st)
{
    k_seg_store_col<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(x, s, jdst, L.chunks, V, ld);
}

void seg_gemm(const SegLayout &L, const double *V, int64_t ld, const int32_t *mcomp, const int32_t *pcomp,
              const double *S, int64_t ss, double *X, int64_t ldx, hipStream_t st)
{
    k_seg_gemm<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(V, ld, mcomp, pcomp, S, ss, L.chunks, X, ldx);
}

void seg_pairs(const SegLayout &L, const double *X, int64_t ldx, const double *Y, int64_t ldy, const int32_t *pcomp,
               const double *lam, double *partial, int ps, int mode, hipStream_t st)
{
    k_seg_pairs<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(X, ldx, Y, ldy, pcomp, lam, L.chunks, partial, ps,
                                                                 mode);
}

void seg_scale_cols(const SegLayout &L, double *X, int64_t ldx, const int32_t *pcomp, const double *s, int ps,
                    hipStream_t st)
{
    k_seg_scale_cols<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(X, ldx, pcomp, s, ps, L.chunks);
}

void seg_copy_cols(const SegLayout &L, const double *Sv, double *D, int64_t ld, const int32_t *cnt, hipStream_t st)
{
    k_seg_copy_cols<<<dim3((unsigned)L.nchunks), dim3(256), 0, st>>>(Sv, D, ld, cnt, L.chunks);
}

}  // namespace sehip
