// mpnn.hip -- the learned start: a small message-passing network over the
// sparsity graph of A that proposes the starting block of the eigensolver,
// with its forward, backward and optimizer kernels.
//
// Graph operator: A_hat = D^-1 |A_off| (row i: |a_ij| / a_ii, j != i), kept as
// a SELL-64 matrix of |a_ij| plus the diagonal D, so A_hat^T = |A_off| D^-1
// (the backward pass) uses the same matrix.  Node features (32 channels,
// row-major n x 32 FP32):
//   0: log a_ii - mean, 1: log s_i - mean (s_i = sum_j |a_ij|),
//   2: (a_ii - s_i) / a_ii, 3..10: N(0,1) noise from a counter-based stream
//   (the splitmix64 / Box-Muller generator of the workload generator),
//   11..14 / 15..18: four more noise channels diffused by A_hat^8 / A_hat^64
//   and scaled to unit RMS (multi-scale inputs: a 3-hop network alone cannot
//   shape modes that vary over the whole graph), 19..31: 0.
// Layer l = 0..L-1 (width 32):  M = A_hat H,  Z = [H | M] W_l^T + b_l,
//   H' = relu(Z) (the last layer is linear and is the output block X).
// W_l is 32 x 64, stored for the products in BF16, master copy FP32; the
// products run on v_mfma_f32_16x16x4f32 with FP32 accumulation.
// Loss (subspace captured, U = the solver's k orthonormal eigenvectors):
//   L = 1 - tr(S^-1 X^T U U^T X) / k,  S = X^T X + ridge,
//   dL/dX = -(2/k) (U B^T S^-1 - X S^-1 C S^-1),  B = X^T U, C = B B^T.
// The network moves the start, never the answer: the eigensolver's own
// residual test decides convergence.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "dev.h"
#include "host_par.h"
#include "mpnn.h"

struct speigs_hip_model {
    int layers = 3;
    int64_t t = 0;                      // optimizer steps taken
    std::vector<float> params;          // host mirror of the master copy
    hipStream_t st = nullptr;
    sehip::DevMem mem;
    float *p = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;   // master, grad, Adam moments
    uint16_t *pw = nullptr;             // BF16 copy of the weights
    ~speigs_hip_model()
    {
        mem.release();
        if (st) (void)hipStreamDestroy(st);
    }
};

namespace sehip {

namespace {

constexpr int CH = MP_WIDTH;          // 32 channels
constexpr int KIN = 2 * MP_WIDTH;     // [H | M]
constexpr int64_t LAYER_P = (int64_t)CH * KIN + CH;

typedef float mf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);    // round to nearest even
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unif(uint64_t seed, uint64_t id)
{
    const uint64_t key = (seed & 0xFFFFFFFFULL) * (1ULL << 32) + id;
    return (double)(mix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

__device__ __forceinline__ float noise(uint64_t seed, int64_t i, int c)
{
    const uint64_t id = (uint64_t)i * MP_NOISE + (uint64_t)c;
    const double u1 = unif(seed, 2 * id), u2 = unif(seed, 2 * id + 1);
    return (float)(sqrt(-2.0 * log1p(-u1)) * cos(2.0 * M_PI * u2));
}

// local features and raw noise of node i (row-major n x 32); Q gets the four
// noise channels 8..11 that are diffused (channels 0..3 of Q, rest 0)
__global__ void k_mp_features(int64_t n, const double *__restrict__ d, const double *__restrict__ s, double mlogd,
                              double mlogs, uint64_t seed, float *__restrict__ H, float *__restrict__ Q)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * CH) return;
    const int64_t i = e / CH;
    const int c = (int)(e - i * CH);
    float v = 0.f;
    if (c == 0) v = (float)(d[i] > 0 ? log(d[i]) - mlogd : 0.0);
    else if (c == 1) v = (float)(s[i] > 0 ? log(s[i]) - mlogs : 0.0);
    else if (c == 2) v = (float)(d[i] > 0 ? (d[i] - s[i]) / d[i] : 0.0);
    else if (c < 11) v = noise(seed, i, c - 3);
    H[e] = v;
    Q[e] = c < 4 ? noise(seed, i, 8 + c) : 0.f;
}

// per-block partial sums of squares of channels 0..3 of Q
__global__ void __launch_bounds__(256) k_mp_sumsq(const float *__restrict__ Q, int64_t n, double *__restrict__ part)
{
    __shared__ double red[256];
    const int c = threadIdx.x & 3;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 64 + (threadIdx.x >> 2); i < min(n, (int64_t)blockIdx.x * 64 + 64); i += 64) {
        const double q = Q[i * CH + c];
        acc += q * q;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int t = threadIdx.x; t < 256; t += 4) s += red[t];
        part[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// H[:, c0 + c] = Q[:, c] * scale[c], c < 4
__global__ void k_mp_place(const float *__restrict__ Q, const double *__restrict__ scale, int c0, int64_t n,
                           float *__restrict__ H)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * 4) return;
    const int64_t i = e >> 2;
    const int c = (int)(e & 3);
    H[i * CH + c0 + c] = (float)(Q[i * CH + c] * scale[c]);
}

// Out_i = so_i sum_j |a_ij| si_j H_j for the first NCH of the 32 channels (32:
// a layer's messages, 4: the diffused input channels): an NCH-lane group per row
template <int NCH>
__global__ void __launch_bounds__(256) k_mp_spmm(int64_t nslices, const int64_t *__restrict__ sptr,
                                                 const int32_t *__restrict__ swid, const int32_t *__restrict__ col,
                                                 const double *__restrict__ val, const float *__restrict__ H,
                                                 const double *__restrict__ si, const double *__restrict__ so,
                                                 float *__restrict__ Out, int64_t n)
{
    const int64_t row = (int64_t)blockIdx.x * (256 / NCH) + (threadIdx.x / NCH);
    if (row >= n) return;
    const int c = threadIdx.x % NCH;
    const int64_t sl = row >> 6;
    int64_t pp = sptr[sl] + (row & 63);
    const int w = swid[sl];
    float acc = 0.f;
    for (int k = 0; k < w; k++, pp += 64) {
        const int64_t j = col[pp];
        const double a = val[pp];
        if (a == 0.0) continue;
        const float coef = (float)(si ? a * si[j] : a);
        acc = fmaf(coef, H[j * CH + c], acc);
    }
    Out[row * CH + c] = so ? (float)(so[row] * (double)acc) : acc;
}

// Z = [A1 | A2] W^T + b (n x 32); H = relu(Z) or Z.  A wave owns 16 rows and
// the two 16-column output tiles; lane l supplies [A1|A2][r + l%16][k + l/16]
// and W[16t + l%16][k + l/16] and holds Z[r + 4(l/16) + v][16t + l%16].
__global__ void __launch_bounds__(256) k_mp_layer(const float *__restrict__ A1, const float *__restrict__ A2,
                                                  const uint16_t *__restrict__ W, const float *__restrict__ b,
                                                  float *__restrict__ H, int64_t n, int relu)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n) return;
    const int lr = lane & 15, lk = lane >> 4;
    const int64_t ar = r0 + lr;
    const bool rin = ar < n;
    mf4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int k = 0; k < KIN; k += 4) {
        const int kk = k + lk;
        const float a = rin ? (kk < CH ? A1[ar * CH + kk] : A2[ar * CH + kk - CH]) : 0.f;
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const float wv = bf16_to_f32(W[(16 * t + lr) * KIN + kk]);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t row = r0 + 4 * lk + v;
            const int c = 16 * t + lr;
            if (row < n) {
                const float z = acc[t][v] + b[c];
                H[row * CH + c] = relu ? fmaxf(z, 0.f) : z;
            }
        }
}

// dA (n x 64) = dZ (n x 32) W (32 x 64): lane l supplies dZ[r + l%16][k + l/16]
// and W[k + l/16][16t + l%16] and holds dA[r + 4(l/16) + v][16t + l%16]
__global__ void __launch_bounds__(256) k_mp_back(const float *__restrict__ dZ, const uint16_t *__restrict__ W,
                                                 float *__restrict__ dA, int64_t n)
{
    const int lane = threadIdx.x & 63;
    const int64_t r0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (r0 >= n) return;
    const int lr = lane & 15, lk = lane >> 4;
    const int64_t ar = r0 + lr;
    const bool rin = ar < n;
    mf4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = mf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < CH; k += 4) {
        const int kk = k + lk;
        const float a = rin ? dZ[ar * CH + kk] : 0.f;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const float wv = bf16_to_f32(W[kk * KIN + 16 * t + lr]);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wv, acc[t], 0, 0, 0);
        }
    }
#pragma unroll
    for (int t = 0; t < 4; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t row = r0 + 4 * lk + v;
            if (row < n) dA[row * KIN + 16 * t + lr] = acc[t][v];
        }
}

// dH = dA[:, 0:32] + dMt (both n x 32), then * (Hprev > 0) when masked
__global__ void k_mp_combine(const float *__restrict__ dA, const float *__restrict__ dMt,
                             const float *__restrict__ Hprev, float *__restrict__ dH, int64_t n, int mask)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * CH) return;
    const int64_t i = e / CH;
    const int c = (int)(e - i * CH);
    float v = dA[i * KIN + c] + dMt[e];
    if (mask && !(Hprev[e] > 0.f)) v = 0.f;
    dH[e] = v;
}

// dM part of dA into its own n x 32 array
__global__ void k_mp_split(const float *__restrict__ dA, float *__restrict__ dM, int64_t n)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n * CH) return;
    const int64_t i = e / CH;
    dM[e] = dA[i * KIN + CH + (e - i * CH)];
}

// weight-gradient partials: block q, rows [q R, q R + R): P_q[o][j] = sum dZ[r][o] [A1|A2][r][j]
// (32 x 64) and P_q[32][o] bias column sums; wave w owns output tile w>>1 and
// input tiles 2(w&1), 2(w&1)+1: lane l supplies dZ[r + l/16][16to + l%16] and
// [A1|A2][r + l/16][16ti + l%16] and holds P[16to + 4(l/16) + v][16ti + l%16]
constexpr int WG_ROWS = 1024;
constexpr int WG_SLOT = CH * KIN + CH;

__global__ void __launch_bounds__(256) k_mp_wgrad(const float *__restrict__ dZ, const float *__restrict__ A1,
                                                  const float *__restrict__ A2, int64_t n, float *__restrict__ part)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int to = wv >> 1, ti0 = 2 * (wv & 1);
    const int64_t r0 = (int64_t)blockIdx.x * WG_ROWS, r1 = min(n, r0 + WG_ROWS);
    mf4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
    float bsum = 0.f;
    for (int64_t r = r0; r < r1; r += 4) {
        const int64_t row = r + lk;
        const bool in = row < r1;
        const float a = in ? dZ[row * CH + 16 * to + lr] : 0.f;
        if ((wv & 1) == 0) bsum += a;                    // waves 0 and 2: bias sums of their output tile
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int j = 16 * (ti0 + u) + lr;
            const float bv = in ? (j < CH ? A1[row * CH + j] : A2[row * CH + j - CH]) : 0.f;
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bv, acc[u], 0, 0, 0);
        }
    }
    float *P = part + (int64_t)blockIdx.x * WG_SLOT;
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) P[(16 * to + 4 * lk + v) * KIN + 16 * (ti0 + u) + lr] = acc[u][v];
    // bias: lanes with equal l%16 hold rows r + l/16; sum the 4 row groups
    __shared__ float bs[2][64];
    if ((wv & 1) == 0) bs[to][lane] = bsum;
    __syncthreads();
    if (threadIdx.x < CH) {
        const int o = threadIdx.x, tt = o >> 4, ll = o & 15;
        P[CH * KIN + o] = (bs[tt][ll] + bs[tt][16 + ll]) + (bs[tt][32 + ll] + bs[tt][48 + ll]);
    }
}

// G[e] = sum over blocks (fixed order)
__global__ void k_mp_wreduce(const float *__restrict__ part, int64_t nb, float *__restrict__ G)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= WG_SLOT) return;
    float s = 0.f;
    for (int64_t q = 0; q < nb; q++) s += part[q * WG_SLOT + e];
    G[e] = s;
}

__global__ void k_mp_to64(const float *__restrict__ X, double *__restrict__ Y, int64_t tot)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < tot) Y[e] = (double)X[e];
}

__global__ void k_mp_to32(const double *__restrict__ Y, float *__restrict__ X, int64_t tot)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < tot) X[e] = (float)Y[e];
}

// Adam (Kingma and Ba 2015) on the FP32 master copy, then the BF16 copy
__global__ void k_mp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                          float *__restrict__ v, uint16_t *__restrict__ pw, int64_t np, float lr, float b1, float b2,
                          float eps, float c1, float c2)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= np) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg, vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    const float pn = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[e] = pn;
    pw[e] = f32_to_bf16(pn);
}

__global__ void k_mp_tobf16(const float *__restrict__ p, uint16_t *__restrict__ pw, int64_t np)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < np) pw[e] = f32_to_bf16(p[e]);
}

inline unsigned nb256(int64_t t) { return (unsigned)((t + 255) / 256); }

// per-graph state of one call
struct Graph {
    int64_t n = 0;
    SellA S;
    double *d = nullptr, *s = nullptr, *dinv = nullptr;
    double mlogd = 0.0, mlogs = 0.0;
};

int graph_build(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, Graph &G, DevMem &mem,
                hipStream_t st)
{
    G.n = n;
    std::vector<double> d(n, 0.0), s(n, 0.0), aabs((size_t)std::max<int64_t>(Ap[n], 1), 0.0);
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i > j) continue;
            if (i == j) {
                d[j] = Ax[q];
            } else {
                aabs[q] = std::fabs(Ax[q]);
                s[i] += aabs[q];
                s[j] += aabs[q];
            }
        }
    double sd = 0.0, ss = 0.0;
    int64_t cd = 0, cs = 0;
    std::vector<double> dinv(n, 0.0);
    for (int64_t i = 0; i < n; i++) {
        if (d[i] > 0) { sd += std::log(d[i]); cd++; dinv[i] = 1.0 / d[i]; }
        if (s[i] > 0) { ss += std::log(s[i]); cs++; }
    }
    G.mlogd = cd ? sd / cd : 0.0;
    G.mlogs = cs ? ss / cs : 0.0;
    int rc;
    if ((rc = sell_build(n, Ap, Ai, aabs.data(), &G.S, mem, st))) return rc;
    if ((rc = mem.alloc(&G.d, (size_t)n))) return rc;
    if ((rc = mem.alloc(&G.s, (size_t)n))) return rc;
    if ((rc = mem.alloc(&G.dinv, (size_t)n))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(G.d, d.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipMemcpyAsync(G.s, s.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipMemcpyAsync(G.dinv, dinv.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    return 0;
}

// The graph from the solver's SELL matrix of A (already on the device): the
// same slices, columns and padding, values |a_ij| off the diagonal and 0 on it;
// per row the diagonal d (first entry in the row's own column: padding entries
// follow the stored ones), s = sum |a_ij| in the row's column order, 1/d, and
// per block the partial sums of log d, log s and their counts.
__global__ void __launch_bounds__(256) k_mp_graph(int64_t n, const int64_t *__restrict__ sptr,
                                                  const int32_t *__restrict__ swid, const int32_t *__restrict__ col,
                                                  const double *__restrict__ val, double *__restrict__ val2,
                                                  double *__restrict__ d, double *__restrict__ s,
                                                  double *__restrict__ dinv, double *__restrict__ part)
{
    __shared__ double red[4][256];
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double ld = 0.0, cd = 0.0, ls = 0.0, cs = 0.0;
    if (row < n) {
        const int64_t sl = row >> 6;
        int64_t pp = sptr[sl] + (row & 63);
        const int w = swid[sl];
        double dv = 0.0, sv = 0.0;
        bool seen = false;
        for (int k = 0; k < w; k++, pp += 64) {
            const double v = val[pp];
            if (col[pp] == row) {
                if (!seen) dv = v;
                seen = true;
                val2[pp] = 0.0;
            } else {
                val2[pp] = fabs(v);
                sv += fabs(v);
            }
        }
        d[row] = dv;
        s[row] = sv;
        dinv[row] = dv > 0 ? 1.0 / dv : 0.0;
        if (dv > 0) { ld = log(dv); cd = 1.0; }
        if (sv > 0) { ls = log(sv); cs = 1.0; }
    }
    red[0][threadIdx.x] = ld;
    red[1][threadIdx.x] = cd;
    red[2][threadIdx.x] = ls;
    red[3][threadIdx.x] = cs;
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int i = 0; i < 256; i++) t += red[threadIdx.x][i];
        part[(int64_t)blockIdx.x * 4 + threadIdx.x] = t;
    }
}

int graph_from_sell(const SellA &A, Graph &G, DevMem &mem, hipStream_t st)
{
    const int64_t n = A.n, nb = (n + 255) / 256;
    G.n = n;
    G.S = A;
    double *val2 = nullptr, *part = nullptr;
    int rc;
    if ((rc = mem.alloc(&val2, (size_t)std::max<int64_t>(A.stored, 1)))) return rc;
    if ((rc = mem.alloc(&G.d, (size_t)n))) return rc;
    if ((rc = mem.alloc(&G.s, (size_t)n))) return rc;
    if ((rc = mem.alloc(&G.dinv, (size_t)n))) return rc;
    if ((rc = mem.alloc(&part, (size_t)nb * 4))) return rc;
    k_mp_graph<<<dim3((unsigned)nb), dim3(256), 0, st>>>(n, A.sptr, A.swid, A.col, A.val, val2, G.d, G.s, G.dinv,
                                                         part);
    G.S.val = val2;
    std::vector<double> hp((size_t)nb * 4);
    SEHIP_CHECK(hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    for (int64_t b = 0; b < nb; b++)
        for (int c = 0; c < 4; c++) t[c] += hp[b * 4 + c];
    G.mlogd = t[1] > 0 ? t[0] / t[1] : 0.0;
    G.mlogs = t[3] > 0 ? t[2] / t[3] : 0.0;
    return 0;
}

void spmm(const Graph &G, const float *H, const double *si, const double *so, float *Out, hipStream_t st)
{
    k_mp_spmm<32><<<dim3((unsigned)((G.n + 7) / 8)), dim3(256), 0, st>>>(G.S.nslices, G.S.sptr, G.S.swid, G.S.col,
                                                                         G.S.val, H, si, so, Out, G.n);
}

// the diffusion of the 4 input noise channels (channels 0..3 of H and Out)
void spmm4(const Graph &G, const float *H, const double *so, float *Out, hipStream_t st)
{
    k_mp_spmm<4><<<dim3((unsigned)((G.n + 63) / 64)), dim3(256), 0, st>>>(G.S.nslices, G.S.sptr, G.S.swid, G.S.col,
                                                                          G.S.val, H, nullptr, so, Out, G.n);
}

constexpr int DIFF1 = 8, DIFF2 = 64;   // diffusion steps of the multi-scale inputs

// H0: local features, raw noise, noise diffused by A_hat^8 and A_hat^64 (unit RMS)
int features(speigs_hip_model *m, const Graph &G, uint64_t seed, float *H, float *Q, float *Q2)
{
    const int64_t n = G.n;
    hipStream_t st = m->st;
    k_mp_features<<<dim3(nb256(n * CH)), dim3(256), 0, st>>>(n, G.d, G.s, G.mlogd, G.mlogs, seed, H, Q);
    const int64_t nb = (n + 63) / 64;
    DevMem tmp;
    double *part = nullptr, *scd = nullptr;
    int rc;
    if ((rc = tmp.alloc(&part, (size_t)nb * 4))) return rc;
    if ((rc = tmp.alloc(&scd, 4))) return rc;
    std::vector<double> hp((size_t)nb * 4), sc(4);
    int done = 0;
    for (int stage = 0; stage < 2; stage++) {
        const int target = stage == 0 ? DIFF1 : DIFF2;
        for (; done < target; done++) {
            spmm4(G, Q, G.dinv, Q2, st);
            std::swap(Q, Q2);
        }
        k_mp_sumsq<<<dim3((unsigned)nb), dim3(256), 0, st>>>(Q, n, part);
        SEHIP_CHECK(hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, st));
        SEHIP_CHECK(hipStreamSynchronize(st));
        for (int c = 0; c < 4; c++) {
            double t = 0.0;
            for (int64_t b = 0; b < nb; b++) t += hp[b * 4 + c];
            sc[c] = t > 0 ? 1.0 / std::sqrt(t / (double)n) : 0.0;
        }
        SEHIP_CHECK(hipMemcpyAsync(scd, sc.data(), 32, hipMemcpyHostToDevice, st));
        k_mp_place<<<dim3(nb256(n * 4)), dim3(256), 0, st>>>(Q, scd, stage == 0 ? 11 : 15, n, H);
        SEHIP_CHECK(hipStreamSynchronize(st));
    }
    return 0;
}

// forward pass; Hs[0..L] (H0 features, H_l outputs), Ms[0..L-1] messages
int forward(speigs_hip_model *m, const Graph &G, uint64_t seed, std::vector<float *> &Hs, std::vector<float *> &Ms)
{
    const int64_t n = G.n;
    hipStream_t st = m->st;
    int rc;
    if ((rc = features(m, G, seed, Hs[0], Ms[0], Hs[1]))) return rc;    // Ms[0], Hs[1]: scratch
    for (int l = 0; l < m->layers; l++) {
        spmm(G, Hs[l], nullptr, G.dinv, Ms[l], st);
        const uint16_t *W = m->pw + l * LAYER_P;
        const float *b = m->p + l * LAYER_P + CH * KIN;
        k_mp_layer<<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st>>>(Hs[l], Ms[l], W, b, Hs[l + 1], n,
                                                                          l + 1 < m->layers ? 1 : 0);
    }
    return 0;
}

struct Buffers {
    DevMem mem;
    std::vector<float *> Hs, Ms;
    float *dH = nullptr, *dA = nullptr, *dM = nullptr, *dMt = nullptr, *part = nullptr;
    double *X64 = nullptr, *U64 = nullptr, *dX64 = nullptr, *gpart = nullptr, *G = nullptr, *C = nullptr;
};

int buffers(speigs_hip_model *m, int64_t n, int64_t k, Buffers &B, bool train)
{
    int rc;
    B.Hs.resize(m->layers + 1);
    B.Ms.resize(m->layers);
    for (auto &h : B.Hs)
        if ((rc = B.mem.alloc(&h, (size_t)n * CH))) return rc;
    for (auto &x : B.Ms)
        if ((rc = B.mem.alloc(&x, (size_t)n * CH))) return rc;
    if (!train) return 0;
    if ((rc = B.mem.alloc(&B.dH, (size_t)n * CH))) return rc;
    if ((rc = B.mem.alloc(&B.dA, (size_t)n * KIN))) return rc;
    if ((rc = B.mem.alloc(&B.dM, (size_t)n * CH))) return rc;
    if ((rc = B.mem.alloc(&B.dMt, (size_t)n * CH))) return rc;
    if ((rc = B.mem.alloc(&B.part, (size_t)((n + WG_ROWS - 1) / WG_ROWS) * WG_SLOT))) return rc;
    if ((rc = B.mem.alloc(&B.X64, (size_t)n * CH))) return rc;
    if ((rc = B.mem.alloc(&B.U64, (size_t)n * std::max<int64_t>(k, 1)))) return rc;
    if ((rc = B.mem.alloc(&B.dX64, (size_t)n * CH))) return rc;
    if ((rc = B.mem.alloc(&B.gpart, (size_t)rgram_blocks(n) * 96 * 96))) return rc;
    if ((rc = B.mem.alloc(&B.G, 96 * 96))) return rc;
    if ((rc = B.mem.alloc(&B.C, 96 * 96))) return rc;
    return 0;
}

// small dense helpers on the host
bool chol_inv(int n, std::vector<double> &S)   // S (n x n sym) -> S^-1
{
    std::vector<double> L((size_t)n * n, 0.0), Inv((size_t)n * n, 0.0);
    for (int j = 0; j < n; j++) {
        double dd = S[j + j * n];
        for (int p = 0; p < j; p++) dd -= L[j + p * n] * L[j + p * n];
        if (!(dd > 0)) return false;
        L[j + j * n] = std::sqrt(dd);
        for (int i = j + 1; i < n; i++) {
            double v = S[i + j * n];
            for (int p = 0; p < j; p++) v -= L[i + p * n] * L[j + p * n];
            L[i + j * n] = v / L[j + j * n];
        }
    }
    for (int c = 0; c < n; c++) {            // solve L L^T x = e_c
        std::vector<double> y(n, 0.0);
        for (int i = 0; i < n; i++) {
            double v = (i == c) ? 1.0 : 0.0;
            for (int p = 0; p < i; p++) v -= L[i + p * n] * y[p];
            y[i] = v / L[i + i * n];
        }
        for (int i = n - 1; i >= 0; i--) {
            double v = y[i];
            for (int p = i + 1; p < n; p++) v -= L[p + i * n] * y[p];
            y[i] = v / L[i + i * n];
        }
        for (int i = 0; i < n; i++) Inv[i + c * n] = y[i];
    }
    S.swap(Inv);
    return true;
}

// loss and dL/dX (into B.dH as FP32) for the output X = Hs[L]
int loss_grad(speigs_hip_model *m, const Graph &G, Buffers &B, const double *Ucol, int64_t k, double *loss)
{
    const int64_t n = G.n;
    hipStream_t st = m->st;
    float *X = B.Hs[m->layers];
    k_mp_to64<<<dim3(nb256(n * CH)), dim3(256), 0, st>>>(X, B.X64, n * CH);
    SEHIP_CHECK(hipMemcpyAsync(B.dX64, Ucol, (size_t)n * k * 8, hipMemcpyHostToDevice, st));   // column-major
    rfrom_cols(B.dX64, (int)k, n, B.U64, k, st);                                              // -> row-major
    std::vector<double> S(CH * CH), Bm((size_t)CH * k);
    rgram(B.X64, CH, CH, B.X64, CH, CH, n, B.gpart, B.G, st);
    SEHIP_CHECK(hipMemcpyAsync(S.data(), B.G, CH * CH * 8, hipMemcpyDeviceToHost, st));
    rgram(B.X64, CH, CH, B.U64, k, (int)k, n, B.gpart, B.G, st);
    SEHIP_CHECK(hipMemcpyAsync(Bm.data(), B.G, (size_t)CH * k * 8, hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    double tr = 0.0;
    for (int i = 0; i < CH; i++) tr += S[i + i * CH];
    const double ridge = 1e-6 * (tr > 0 ? tr / CH : 1.0);
    for (int i = 0; i < CH; i++) S[i + i * CH] += ridge;
    for (int j = 0; j < CH; j++)
        for (int i = 0; i < j; i++) S[i + j * CH] = S[j + i * CH] = 0.5 * (S[i + j * CH] + S[j + i * CH]);
    if (!chol_inv(CH, S)) return SPEIGS_ERR_LAPACK;
    // C = B B^T, f = tr(S^-1 C), M = S^-1 C S^-1, Q = B^T S^-1 (k x 32)
    std::vector<double> C(CH * CH, 0.0), SC(CH * CH, 0.0), M(CH * CH, 0.0), Q((size_t)k * CH, 0.0);
    for (int j = 0; j < CH; j++)
        for (int i = 0; i < CH; i++) {
            double v = 0.0;
            for (int64_t q = 0; q < k; q++) v += Bm[i + q * CH] * Bm[j + q * CH];
            C[i + j * CH] = v;
        }
    double f = 0.0;
    for (int j = 0; j < CH; j++)
        for (int i = 0; i < CH; i++) {
            double v = 0.0;
            for (int q = 0; q < CH; q++) v += S[i + q * CH] * C[q + j * CH];
            SC[i + j * CH] = v;
        }
    for (int i = 0; i < CH; i++) f += SC[i + i * CH];
    for (int j = 0; j < CH; j++)
        for (int i = 0; i < CH; i++) {
            double v = 0.0;
            for (int q = 0; q < CH; q++) v += SC[i + q * CH] * S[q + j * CH];
            M[i + j * CH] = v;
        }
    for (int j = 0; j < CH; j++)
        for (int64_t i = 0; i < k; i++) {
            double v = 0.0;
            for (int q = 0; q < CH; q++) v += Bm[q + i * CH] * S[q + j * CH];
            Q[i + (size_t)j * k] = v;
        }
    *loss = 1.0 - f / (double)k;
    // dX = -(2/k) (U Q - X M)
    const double sc = -2.0 / (double)k;
    for (auto &v : Q) v *= sc;
    for (auto &v : M) v *= -sc;
    SEHIP_CHECK(hipMemcpyAsync(B.C, Q.data(), Q.size() * 8, hipMemcpyHostToDevice, st));
    rgemm(B.U64, k, (int)k, B.C, CH, B.dX64, CH, n, 0.0, st);
    SEHIP_CHECK(hipMemcpyAsync(B.G, M.data(), M.size() * 8, hipMemcpyHostToDevice, st));
    rgemm(B.X64, CH, CH, B.G, CH, B.dX64, CH, n, 1.0, st);
    k_mp_to32<<<dim3(nb256(n * CH)), dim3(256), 0, st>>>(B.dX64, B.dH, n * CH);
    return 0;
}

// backward from dH = dL/dH_L (in B.dH) to the parameter gradient m->g
int backward(speigs_hip_model *m, const Graph &G, Buffers &B)
{
    const int64_t n = G.n;
    hipStream_t st = m->st;
    const int64_t nblk = (n + WG_ROWS - 1) / WG_ROWS;
    for (int l = m->layers - 1; l >= 0; l--) {
        // B.dH holds dZ_l (the relu mask of layer l was applied when it was formed)
        k_mp_wgrad<<<dim3((unsigned)nblk), dim3(256), 0, st>>>(B.dH, B.Hs[l], B.Ms[l], n, B.part);
        k_mp_wreduce<<<dim3(nb256(WG_SLOT)), dim3(256), 0, st>>>(B.part, nblk, m->g + l * LAYER_P);
        if (l == 0) break;
        const uint16_t *W = m->pw + l * LAYER_P;
        k_mp_back<<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st>>>(B.dH, W, B.dA, n);
        k_mp_split<<<dim3(nb256(n * CH)), dim3(256), 0, st>>>(B.dA, B.dM, n);
        spmm(G, B.dM, G.dinv, nullptr, B.dMt, st);     // A_hat^T dM
        k_mp_combine<<<dim3(nb256(n * CH)), dim3(256), 0, st>>>(B.dA, B.dMt, B.Hs[l], B.dH, n, 1);
    }
    return 0;
}

int model_upload(speigs_hip_model *m)
{
    const int64_t np = (int64_t)m->params.size();
    SEHIP_CHECK(hipMemcpyAsync(m->p, m->params.data(), (size_t)np * 4, hipMemcpyHostToDevice, m->st));
    k_mp_tobf16<<<dim3(nb256(np)), dim3(256), 0, m->st>>>(m->p, m->pw, np);
    SEHIP_CHECK(hipStreamSynchronize(m->st));
    return 0;
}

}  // namespace

int64_t mp_nparams(int layers) { return (int64_t)layers * LAYER_P; }

// X (n x 32 row-major FP32, device) for the graph of A
int mp_infer_device(speigs_hip_model *m, int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                    uint64_t seed, float *Xdev)
{
    DevMem gm;
    Graph G;
    int rc;
    if ((rc = graph_build(n, Ap, Ai, Ax, G, gm, m->st))) return rc;
    Buffers B;
    if ((rc = buffers(m, n, 0, B, false))) return rc;
    if ((rc = forward(m, G, seed, B.Hs, B.Ms))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(Xdev, B.Hs[m->layers], (size_t)n * CH * 4, hipMemcpyDeviceToDevice, m->st));
    SEHIP_CHECK(hipStreamSynchronize(m->st));
    return 0;
}

int mp_infer_sell(speigs_hip_model *m, const SellA &A, uint64_t seed, float *Xdev)
{
    double tp = wall();
    DevMem gm;
    Graph G;
    int rc;
    if ((rc = graph_from_sell(A, G, gm, m->st))) return rc;
    trace_phase("start network: graph", tp);
    Buffers B;
    if ((rc = buffers(m, A.n, 0, B, false))) return rc;
    trace_phase("start network: buffers", tp);
    if ((rc = forward(m, G, seed, B.Hs, B.Ms))) return rc;
    SEHIP_CHECK(hipMemcpyAsync(Xdev, B.Hs[m->layers], (size_t)A.n * CH * 4, hipMemcpyDeviceToDevice, m->st));
    SEHIP_CHECK(hipStreamSynchronize(m->st));
    trace_phase("start network: forward", tp);
    return 0;
}

}  // namespace sehip

using namespace sehip;

extern "C" {

int speigs_hip_model_create(int layers, uint64_t seed, speigs_hip_model **out)
{
    if (!out || layers < 1 || layers > 8) return SPEIGS_ERR_ARG;
    *out = nullptr;
    if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
    std::unique_ptr<speigs_hip_model> m(new speigs_hip_model());
    m->layers = layers;
    const int64_t np = mp_nparams(layers);
    m->params.assign(np, 0.f);
    // Glorot-uniform weights from the counter-based stream, zero biases
    uint64_t z = seed ? seed : 0x243F6A8885A308D3ULL;
    const float lim = std::sqrt(6.0f / (float)(KIN + CH));
    for (int l = 0; l < layers; l++)
        for (int64_t e = 0; e < (int64_t)CH * KIN; e++) {
            z ^= z >> 12; z ^= z << 25; z ^= z >> 27;
            const double u = (double)((z * 0x2545F4914F6CDD1DULL) >> 11) * (1.0 / 9007199254740992.0);
            m->params[l * LAYER_P + e] = (float)((2.0 * u - 1.0) * lim);
        }
    SEHIP_CHECK(hipStreamCreate(&m->st));
    int rc;
    if ((rc = m->mem.alloc(&m->p, (size_t)np))) return rc;
    if ((rc = m->mem.alloc(&m->g, (size_t)np))) return rc;
    if ((rc = m->mem.alloc(&m->am, (size_t)np))) return rc;
    if ((rc = m->mem.alloc(&m->av, (size_t)np))) return rc;
    if ((rc = m->mem.alloc(&m->pw, (size_t)np))) return rc;
    SEHIP_CHECK(hipMemsetAsync(m->am, 0, (size_t)np * 4, m->st));
    SEHIP_CHECK(hipMemsetAsync(m->av, 0, (size_t)np * 4, m->st));
    SEHIP_CHECK(hipMemsetAsync(m->g, 0, (size_t)np * 4, m->st));
    if ((rc = model_upload(m.get()))) return rc;
    *out = m.release();
    return SPEIGS_OK;
}

int64_t speigs_hip_model_nparams(const speigs_hip_model *m) { return m ? (int64_t)m->params.size() : 0; }

int speigs_hip_model_set_params(speigs_hip_model *m, const float *p, int64_t np)
{
    if (!m || !p || np != (int64_t)m->params.size()) return SPEIGS_ERR_ARG;
    std::memcpy(m->params.data(), p, (size_t)np * 4);
    m->t = 0;
    SEHIP_CHECK(hipMemsetAsync(m->am, 0, (size_t)np * 4, m->st));
    SEHIP_CHECK(hipMemsetAsync(m->av, 0, (size_t)np * 4, m->st));
    return model_upload(m);
}

int speigs_hip_model_get_params(const speigs_hip_model *m, float *p, int64_t np)
{
    if (!m || !p || np != (int64_t)m->params.size()) return SPEIGS_ERR_ARG;
    SEHIP_CHECK(hipMemcpy(p, m->p, (size_t)np * 4, hipMemcpyDeviceToHost));
    return SPEIGS_OK;
}

int speigs_hip_model_infer(speigs_hip_model *m, speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                           const double *Ax, uint64_t seed, float *X)
{
    if (!m || n <= 0 || !Ap || !Ai || !Ax || !X) return SPEIGS_ERR_ARG;
    DevMem tmp;
    float *Xd;
    int rc;
    if ((rc = tmp.alloc(&Xd, (size_t)n * CH))) return rc;
    if ((rc = mp_infer_device(m, n, Ap, Ai, Ax, seed, Xd))) return rc;
    SEHIP_CHECK(hipMemcpy(X, Xd, (size_t)n * CH * 4, hipMemcpyDeviceToHost));
    return SPEIGS_OK;
}

int speigs_hip_model_grad(speigs_hip_model *m, speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                          const double *Ax, const double *U, speigs_int k, uint64_t seed, double *loss, float *grad)
{
    if (!m || n <= 0 || !Ap || !Ai || !Ax || !U || k < 1 || k > 96 || !loss) return SPEIGS_ERR_ARG;
    DevMem gm;
    Graph G;
    int rc;
    if ((rc = graph_build(n, Ap, Ai, Ax, G, gm, m->st))) return rc;
    Buffers B;
    if ((rc = buffers(m, n, k, B, true))) return rc;
    if ((rc = forward(m, G, seed, B.Hs, B.Ms))) return rc;
    if ((rc = loss_grad(m, G, B, U, k, loss))) return rc;
    if ((rc = backward(m, G, B))) return rc;
    if (grad) SEHIP_CHECK(hipMemcpyAsync(grad, m->g, m->params.size() * 4, hipMemcpyDeviceToHost, m->st));
    SEHIP_CHECK(hipStreamSynchronize(m->st));
    return SPEIGS_OK;
}

int speigs_hip_model_train_step(speigs_hip_model *m, speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                                const double *Ax, const double *U, speigs_int k, uint64_t seed, float lr,
                                double *loss)
{
    int rc = speigs_hip_model_grad(m, n, Ap, Ai, Ax, U, k, seed, loss, nullptr);
    if (rc) return rc;
    const int64_t np = (int64_t)m->params.size();
    m->t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)m->t), c2 = 1.f - std::pow(b2, (float)m->t);
    k_mp_adam<<<dim3(nb256(np)), dim3(256), 0, m->st>>>(m->p, m->g, m->am, m->av, m->pw, np, lr, b1, b2, 1e-8f, c1,
                                                         c2);
    SEHIP_CHECK(hipMemcpyAsync(m->params.data(), m->p, (size_t)np * 4, hipMemcpyDeviceToHost, m->st));
    SEHIP_CHECK(hipStreamSynchronize(m->st));
    return SPEIGS_OK;
}

void speigs_hip_model_destroy(speigs_hip_model *m) { delete m; }

}  // extern "C"
