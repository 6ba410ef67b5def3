// hermitian.hip -- speigs_z: complex Hermitian A = R + iS through the real
// symmetric embedding M = [[R, -S], [S, R]] of order 2n (reference
// src/speigs_core.c:660-768; the embedding is the reference's, the recovery
// below is this library's own).
//
// Every eigenvalue of A is a double eigenvalue of M, with eigenvectors [x; y]
// and [-y; x] for the one complex direction x + iy, so the 2k + 2 pairs asked of
// M hold each wanted complex direction up to a factor of i.  The recovery keeps,
// in the order the real solver returned them, each candidate w_c = a_c + i b_c
// that has at least 0.1 of its (unit) norm outside the complex span of the ones
// kept before it -- the reference's test (a C-linearly dependent candidate is
// dropped, not merely a parallel one) -- and orthonormalises the kept ones.
//
// Here that is one block computation instead of a Gram-Schmidt loop per
// candidate: the complex Gram matrix G = W^H W of all m candidates is formed
// on the device in one pass over W (deterministic partial sums); the selection
// runs on the host on G alone (the part of candidate c outside span(kept) has
// norm^2 = 1 - |R^-H g_c|^2, R the Cholesky factor of the kept block of G);
// the kept directions are then Q = W T with T = D R^-1 (Cholesky QR), formed on
// the device, and a second Cholesky QR on Q's own Gram matrix restores
// orthonormality to rounding (the role of the reference's second Gram-Schmidt
// pass).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "dev.h"
#include "host_par.h"

namespace sehip {
namespace {

using cplx = std::complex<double>;

constexpr int kGT = 16;          // Gram tile: 16 x 16 column pairs per workgroup
constexpr int kGRows = 2048;     // rows per workgroup (one partial per workgroup)
constexpr int kGSub = 64;        // rows staged in LDS at a time

// part[rb][p][q] = sum over the rows of block rb of conj(w_p) w_q, w = a + i b,
// a_c = A[c * ld + i], b_c = B[c * ld + i]; a workgroup per (row block, tile pair)
__global__ void __launch_bounds__(256) k_cgram_partial(const double *__restrict__ A, const double *__restrict__ B,
                                                       int64_t ld, int m, int64_t n, int ntile,
                                                       double *__restrict__ part)
{
    __shared__ double sa[2][kGSub][kGT + 1], sb[2][kGSub][kGT + 1];   // [p or q tile][row][column]
    const int rb = blockIdx.x;
    const int tp = blockIdx.y / ntile, tq = blockIdx.y % ntile;
    const int t = threadIdx.x, tc = t / kGT, tr = t % kGT;
    const int p = tp * kGT + tc, q = tq * kGT + tr;
    const int64_t r0 = (int64_t)rb * kGRows, r1 = min(n, r0 + kGRows);
    double re = 0.0, im = 0.0;
    // loader: thread t stages column t / 16 of a tile, rows 4 (t % 16) .. + 3
    const int lc = t / 16, lr = 4 * (t % 16);
    for (int64_t s0 = r0; s0 < r1; s0 += kGSub) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int col = (h == 0 ? tp : tq) * kGT + lc;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int64_t row = s0 + lr + u;
                const bool in = col < m && row < r1;
                sa[h][lr + u][lc] = in ? A[(int64_t)col * ld + row] : 0.0;
                sb[h][lr + u][lc] = in ? B[(int64_t)col * ld + row] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll 8
        for (int r = 0; r < kGSub; r++) {
            const double ap = sa[0][r][tc], bp = sb[0][r][tc], aq = sa[1][r][tr], bq = sb[1][r][tr];
            re = fma(ap, aq, fma(bp, bq, re));
            im = fma(ap, bq, fma(-bp, aq, im));
        }
        __syncthreads();
    }
    if (p < m && q < m) {
        double *o = part + 2 * ((int64_t)rb * m * m + (int64_t)p * m + q);
        o[0] = re;
        o[1] = im;
    }
}

// G[p][q] = sum over row blocks of part (fixed order)
__global__ void __launch_bounds__(256) k_cgram_reduce(const double *__restrict__ part, int64_t nrb, int mm,
                                                      double *__restrict__ G)
{
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= 2 * mm) return;
    double s = 0.0;
    for (int64_t rb = 0; rb < nrb; rb++) s += part[rb * 2 * mm + e];
    G[e] = s;
}

// Q = W T for p <= 16 * NG output columns: Qr + i Qi, ld n; T complex m x p,
// T[c][t] at Tm[2 (c p + t)]; a thread per row, outputs in groups of 16 so
// that W is read once per group
template <int NG>
__global__ void __launch_bounds__(256) k_cgemm_tall(const double *__restrict__ A, const double *__restrict__ B,
                                                    int64_t ld, int m, const double *__restrict__ Tm, int p,
                                                    int64_t n, double *__restrict__ Qr, double *__restrict__ Qi)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int g0 = 0; g0 < p; g0 += 16 * NG) {
        double accr[16 * NG], acci[16 * NG];
#pragma unroll
        for (int u = 0; u < 16 * NG; u++) accr[u] = acci[u] = 0.0;
        for (int c = 0; c < m; c++) {
            const double a = A[(int64_t)c * ld + i], b = B[(int64_t)c * ld + i];
            const double *tc = Tm + 2 * ((int64_t)c * p + g0);
#pragma unroll
            for (int u = 0; u < 16 * NG; u++) {
                if (g0 + u < p) {
                    const double tr = tc[2 * u], ti = tc[2 * u + 1];
                    accr[u] = fma(a, tr, fma(-b, ti, accr[u]));
                    acci[u] = fma(a, ti, fma(b, tr, acci[u]));
                }
            }
        }
#pragma unroll
        for (int u = 0; u < 16 * NG; u++)
            if (g0 + u < p) {
                Qr[(int64_t)(g0 + u) * n + i] = accr[u];
                Qi[(int64_t)(g0 + u) * n + i] = acci[u];
            }
    }
}

struct Recover {
    hipStream_t st = nullptr;
    DevMem mem;
    double *part = nullptr, *G = nullptr, *Tm = nullptr;
    int64_t nrb = 0;
    ~Recover()
    {
        if (st) (void)hipStreamDestroy(st);
    }
};

// G (m x m complex, row p = conj(w_p) against every w_q) of the columns (A, B)
int cgram(Recover &R, const double *A, const double *B, int64_t ld, int m, int64_t n, std::vector<cplx> &G)
{
    const int ntile = (m + kGT - 1) / kGT;
    k_cgram_partial<<<dim3((unsigned)R.nrb, (unsigned)(ntile * ntile)), dim3(256), 0, R.st>>>(A, B, ld, m, n, ntile,
                                                                                              R.part);
    const int mm = m * m;
    k_cgram_reduce<<<dim3((unsigned)((2 * mm + 255) / 256)), dim3(256), 0, R.st>>>(R.part, R.nrb, mm, R.G);
    G.assign((size_t)mm, cplx(0.0, 0.0));
    SEHIP_CHECK(hipMemcpyAsync(G.data(), R.G, (size_t)mm * sizeof(cplx), hipMemcpyDeviceToHost, R.st));
    SEHIP_CHECK(hipStreamSynchronize(R.st));
    return 0;
}

// (Qr, Qi) = (A + iB) T, T m x p complex
int cgemm(Recover &R, const double *A, const double *B, int64_t ld, int m, const std::vector<cplx> &T, int p,
          int64_t n, double *Qr, double *Qi)
{
    SEHIP_CHECK(hipMemcpyAsync(R.Tm, T.data(), (size_t)m * p * sizeof(cplx), hipMemcpyHostToDevice, R.st));
    const dim3 g((unsigned)((n + 255) / 256));
    if (p <= 16) k_cgemm_tall<1><<<g, dim3(256), 0, R.st>>>(A, B, ld, m, R.Tm, p, n, Qr, Qi);
    else k_cgemm_tall<2><<<g, dim3(256), 0, R.st>>>(A, B, ld, m, R.Tm, p, n, Qr, Qi);
    return 0;
}

// R^-1 of an upper triangular complex r x r matrix (column-major, R[a + b r])
std::vector<cplx> upper_inverse(const std::vector<cplx> &Rm, int r)
{
    std::vector<cplx> X((size_t)r * r, cplx(0.0, 0.0));
    for (int j = 0; j < r; j++) {
        X[j + (size_t)j * r] = 1.0 / Rm[j + (size_t)j * r];
        for (int i = j - 1; i >= 0; i--) {
            cplx s(0.0, 0.0);
            for (int l = i + 1; l <= j; l++) s += Rm[i + (size_t)l * r] * X[l + (size_t)j * r];
            X[i + (size_t)j * r] = -s / Rm[i + (size_t)i * r];
        }
    }
    return X;
}

}  // namespace

// Selects and orthonormalises the complex directions among the m columns of
// V2 (2n x m real, column-major: [a_c; b_c]); writes up to k of them to Vr/Vi
// (n x k column-major, zero columns past the ones found) and their candidate
// indices to sel.  Returns 0 or an error code.
int hermitian_recover(int64_t n, int m, const double *V2, int64_t k, double *Vr, double *Vi,
                      std::vector<int> &sel)
{
    sel.clear();
    if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
    Recover R;
    SEHIP_CHECK(hipStreamCreate(&R.st));
    R.nrb = (n + kGRows - 1) / kGRows;
    const int kq = (int)std::min<int64_t>(k, m);
    double *W = nullptr, *Q1r = nullptr, *Q1i = nullptr, *Q2r = nullptr, *Q2i = nullptr;
    int rc;
    if ((rc = R.mem.alloc(&W, (size_t)2 * n * m))) return rc;
    if ((rc = R.mem.alloc(&Q1r, (size_t)n * kq))) return rc;
    if ((rc = R.mem.alloc(&Q1i, (size_t)n * kq))) return rc;
    if ((rc = R.mem.alloc(&Q2r, (size_t)n * k))) return rc;
    if ((rc = R.mem.alloc(&Q2i, (size_t)n * k))) return rc;
    if ((rc = R.mem.alloc(&R.part, (size_t)R.nrb * 2 * m * m))) return rc;
    if ((rc = R.mem.alloc(&R.G, (size_t)2 * m * m))) return rc;
    if ((rc = R.mem.alloc(&R.Tm, (size_t)2 * m * m))) return rc;
    if ((rc = copy_to_device(W, V2, (size_t)2 * n * m * sizeof(double), R.st))) return rc;
    std::vector<cplx> G;
    if ((rc = cgram(R, W, W + n, 2 * n, m, n, G))) return rc;
    // greedy selection in the solver's order on the Gram matrix alone: Rk is
    // the Cholesky factor of the kept columns' Gram matrix, each scaled to
    // unit norm (sc)
    std::vector<cplx> Rk((size_t)kq * kq, cplx(0.0, 0.0));
    std::vector<double> sc;
    for (int c = 0; c < m && (int)sel.size() < kq; c++) {
        const double g = G[(size_t)c * m + c].real();
        if (!(g > 0.0)) continue;
        const double s = 1.0 / std::sqrt(g);
        const int r = (int)sel.size();
        std::vector<cplx> y(r);
        double y2 = 0.0;
        for (int a = 0; a < r; a++) {      // R^H y = (scaled kept block of G)[:, c]
            cplx v = sc[a] * s * G[(size_t)sel[a] * m + c];
            for (int b = 0; b < a; b++) v -= std::conj(Rk[b + (size_t)a * kq]) * y[b];
            y[a] = v / Rk[a + (size_t)a * kq].real();
            y2 += std::norm(y[a]);
        }
        const double left = std::sqrt(std::max(0.0, 1.0 - y2));
        if (left <= 0.1) continue;            // within the span of the kept ones
        for (int a = 0; a < r; a++) Rk[a + (size_t)r * kq] = y[a];
        Rk[r + (size_t)r * kq] = left;
        sel.push_back(c);
        sc.push_back(s);
    }
    const int r = (int)sel.size();
    if (r > 0) {
        // pass 1: Q1 = W_sel D Rk^-1
        std::vector<cplx> Rr((size_t)r * r);
        for (int a = 0; a < r; a++)
            for (int b = 0; b < r; b++) Rr[a + (size_t)b * r] = Rk[a + (size_t)b * kq];
        const std::vector<cplx> Ri = upper_inverse(Rr, r);
        std::vector<cplx> T((size_t)m * r, cplx(0.0, 0.0));
        for (int a = 0; a < r; a++)
            for (int t = 0; t < r; t++) T[(size_t)sel[a] * r + t] = sc[a] * Ri[a + (size_t)t * r];
        if ((rc = cgemm(R, W, W + n, 2 * n, m, T, r, n, Q1r, Q1i))) return rc;
        // pass 2: Cholesky QR of Q1 on its own Gram matrix
        std::vector<cplx> G2;
        if ((rc = cgram(R, Q1r, Q1i, n, r, n, G2))) return rc;
        std::vector<cplx> R2((size_t)r * r, cplx(0.0, 0.0));
        for (int j = 0; j < r; j++) {
            for (int i = 0; i <= j; i++) {     // R2^H R2 = G2, column j
                cplx v = G2[(size_t)i * r + j];
                for (int l = 0; l < i; l++) v -= std::conj(R2[l + (size_t)i * r]) * R2[l + (size_t)j * r];
                if (i < j) {
                    R2[i + (size_t)j * r] = v / R2[i + (size_t)i * r].real();
                } else {
                    if (!(v.real() > 0.0)) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "speigs_z: second QR pass lost rank");
                    R2[j + (size_t)j * r] = std::sqrt(v.real());
                }
            }
        }
        const std::vector<cplx> T2 = upper_inverse(R2, r);
        if ((rc = cgemm(R, Q1r, Q1i, n, r, T2, r, n, Q2r, Q2i))) return rc;
    }
    if (r < k) {
        SEHIP_CHECK(hipMemsetAsync(Q2r + (size_t)r * n, 0, (size_t)(k - r) * n * sizeof(double), R.st));
        SEHIP_CHECK(hipMemsetAsync(Q2i + (size_t)r * n, 0, (size_t)(k - r) * n * sizeof(double), R.st));
    }
    if (Vr && (rc = copy_to_host(Vr, Q2r, (size_t)n * k * sizeof(double), R.st))) return rc;
    if (Vi && (rc = copy_to_host(Vi, Q2i, (size_t)n * k * sizeof(double), R.st))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(R.st));
    return 0;
}

}  // namespace sehip

extern "C" int speigs_z(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, const double *Axr,
                        const double *Axi, speigs_int k, speigs_which which, const speigs_opts *opts, double *lambda,
                        double *Vr, double *Vi, double *resid, speigs_info *info)
{
    using namespace sehip;
    if (!Axi) return speigs(n, Ap, Ai, Axr, k, which, opts, lambda, Vr, resid, info);
    if (n <= 0 || k <= 0 || k > n || !Ap || !Ai || !Axr || !lambda) return SPEIGS_ERR_ARG;
    speigs_info local;
    if (!info) info = &local;
    // the embedding, column by column in parallel: column j of M holds R(:, j)
    // then S(:, j) (rows n + i), column n + j holds -S(:, j) then R(:, j);
    // explicit zeros of R or S are not stored
    const int64_t n2 = 2 * n;
    const int nt = host_threads();
    std::vector<int64_t> cnt(n2 + 1, 0);
    parallel_for(n, nt, 4096, [&](int64_t j, int) {
        int64_t nr = 0, ni = 0;
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            nr += Axr[q] != 0.0;
            ni += Axi[q] != 0.0;
        }
        cnt[j + 1] = nr + ni;
        cnt[n + j + 1] = nr + ni;
    });
    for (int64_t j = 0; j < n2; j++) cnt[j + 1] += cnt[j];
    const int64_t nzb = cnt[n2];
    std::vector<int64_t> Bi((size_t)std::max<int64_t>(nzb, 1), 0);
    std::vector<double> Bx((size_t)std::max<int64_t>(nzb, 1), 0.0);
    parallel_for(n, nt, 4096, [&](int64_t j, int) {
        int64_t u = cnt[j], v = cnt[n + j];
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
            if (Axr[q] != 0.0) { Bi[u] = Ai[q]; Bx[u++] = Axr[q]; }
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
            if (Axi[q] != 0.0) { Bi[u] = n + Ai[q]; Bx[u++] = Axi[q]; }
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
            if (Axi[q] != 0.0) { Bi[v] = Ai[q]; Bx[v++] = -Axi[q]; }
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
            if (Axr[q] != 0.0) { Bi[v] = n + Ai[q]; Bx[v++] = Axr[q]; }
    });
    // each wanted value is a double one of M: ask for 2k + 2
    const int64_t m = std::min<int64_t>(2 * k + 2, n2);
    std::vector<double> l2(m), r2(m), V2((size_t)n2 * m);
    int rc = speigs(n2, cnt.data(), Bi.data(), Bx.data(), m, which, opts, l2.data(), V2.data(), r2.data(), info);
    if (rc) return rc;
    std::vector<int> sel;
    if ((rc = hermitian_recover(n, (int)m, V2.data(), k, Vr, Vi, sel))) return rc;
    const int64_t nacc = (int64_t)sel.size();
    for (int64_t t = 0; t < nacc; t++) {
        lambda[t] = l2[sel[t]];
        if (resid) resid[t] = r2[sel[t]];
    }
    info->nconv = nacc;
    if (nacc < k) info->flag = 1;
    return SPEIGS_OK;
}
