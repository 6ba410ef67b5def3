// dense_batch.hip -- the dense short-circuit of many small components at once
// (reference src/speigs_core.c:318-353, reached per component through the
// blocks path's recursion, :577-607): the k eigenpairs of smallest (or
// largest) |lambda| of each small dense symmetric matrix.
//
// The O(n^3) parts run on the device, one workgroup per matrix, every matrix
// of the batch in one launch: the Householder tridiagonalisation (reflector k
// a unit vector v, H_k = I - 2 v v^T on rows k+1..n-1, left in column k -- the
// same representation as the host's) and, once the host has the selected
// eigenvectors z of each tridiagonal, their back-transformation Q z.  The
// O(n^2) parts stay on the host: implicit QL for all eigenvalues, the
// selection by |lambda|, inverse iteration for the selected vectors.  Block
// reductions have a fixed order, so a run is reproducible.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dev.h"
#include "host_par.h"

namespace sehip {

namespace {

constexpr int kDT = 1024;        // threads per matrix: thread i owns row i (n <= kDT)

// sum over the workgroup in a fixed order: waves by shuffles, then wave 0
// over the waves' partials; every thread gets the total
__device__ __forceinline__ double block_sum(double x, double *red)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();                       // red's previous use is complete
    if (lane == 0) red[wv] = x;
    __syncthreads();
    if (wv == 0) {
        double t = lane < (int)(blockDim.x >> 6) ? red[lane] : 0.0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
        if (lane == 0) red[32] = t;
    }
    __syncthreads();
    return red[32];
}

// matrix b: n = nn[b], column-major n x n at A + aoff[b] (full symmetric);
// out: d at de + 2 doff[b], e right after it (e[n - 1] = 0)
__global__ void __launch_bounds__(kDT) k_tridiag(double *__restrict__ Aall, const int64_t *__restrict__ aoff,
                                                const int32_t *__restrict__ nn, const int64_t *__restrict__ doff,
                                                double *__restrict__ de)
{
    __shared__ double v[kDT], p[kDT], red[33];
    const int b = blockIdx.x, n = nn[b], tid = threadIdx.x;
    double *A = Aall + aoff[b];
    double *d = de + 2 * doff[b], *e = d + n;
    const int i = tid;                     // this thread's row, in every step
    for (int k = 0; k + 2 < n; k++) {
        const bool mine = i > k && i < n;
        double *ak = A + (int64_t)k * n;
        const double a = mine ? ak[i] : 0.0;
        const double xn = sqrt(block_sum(a * a, red));
        const double x0 = ak[k + 1];
        if (xn == 0.0) {
            if (mine) ak[i] = 0.0;
            if (tid == 0) e[k] = 0.0;
            continue;
        }
        const double alpha = x0 > 0 ? -xn : xn;
        const double vi0 = mine ? (i == k + 1 ? a - alpha : a) : 0.0;
        const double vn = sqrt(block_sum(vi0 * vi0, red));
        if (vn == 0.0) {
            if (mine) ak[i] = 0.0;
            if (tid == 0) e[k] = x0;
            continue;
        }
        const double vi = vi0 / vn;
        if (mine) v[i] = vi;
        __syncthreads();
        // p = A22 v: row i of A22 along the columns (coalesced across rows)
        double pi = 0.0;
        if (mine)
            for (int j = k + 1; j < n; j++) pi += A[i + (int64_t)j * n] * v[j];
        const double K = block_sum(vi * pi, red);
        pi = 2.0 * pi - 2.0 * K * vi;
        if (mine) p[i] = pi;
        __syncthreads();
        // A22 -= v p^T + p v^T
        if (mine)
            for (int j = k + 1; j < n; j++) {
                double *aij = A + i + (int64_t)j * n;
                *aij -= vi * p[j] + pi * v[j];
            }
        if (mine) ak[i] = vi;
        if (tid == 0) e[k] = alpha;
        __syncthreads();                   // v, p reused by the next step
    }
    __syncthreads();
    if (i < n) d[i] = A[i + (int64_t)i * n];
    if (tid == 0) {
        if (n >= 2) e[n - 2] = A[(n - 1) + (int64_t)(n - 2) * n];
        e[n - 1] = 0.0;
    }
}

// back-transformation of the selected vectors: matrix b's k = kk[b] vectors
// (column t at Z + zoff[b] + t n) become H_0 H_1 ... H_{n-3} z; one wave per
// vector (grid y covers the vectors, 16 per workgroup)
__global__ void __launch_bounds__(1024) k_tri_back(const double *__restrict__ Aall, const int64_t *__restrict__ aoff,
                                                   const int32_t *__restrict__ nn, const int32_t *__restrict__ kk,
                                                   const int64_t *__restrict__ zoff, double *__restrict__ Z)
{
    const int b = blockIdx.x, n = nn[b];
    const int t = blockIdx.y * 16 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (t >= kk[b]) return;                // per wave; no workgroup barrier below
    const double *A = Aall + aoff[b];
    double *z = Z + zoff[b] + (int64_t)t * n;
    // lane l owns the entries i = l (mod 64) throughout (no lane reads another's)
    for (int k = n - 3; k >= 0; k--) {
        const double *vk = A + (int64_t)k * n;
        double s = 0.0;
        for (int i = lane; i < n; i += 64)
            if (i > k) s += vk[i] * z[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        s *= 2.0;
        if (s != 0.0)
            for (int i = lane; i < n; i += 64)
                if (i > k) z[i] -= s * vk[i];
    }
}

}  // namespace

int dense_batch(std::vector<DenseJob> &jobs, hipStream_t st)
{
    const int64_t nj = (int64_t)jobs.size();
    if (nj == 0) return 0;
    for (const DenseJob &j : jobs)
        if (j.n < 1 || j.n > kDenseBatchMax || j.k < 1 || j.k > j.n)
            return SEHIP_FAIL(SPEIGS_ERR_ARG, "dense batch: a matrix outside 1 <= k <= n <= 1024");
    std::vector<int64_t> aoff(nj + 1, 0), doff(nj + 1, 0), zoff(nj + 1, 0);
    std::vector<int32_t> nn(nj), kk(nj);
    for (int64_t b = 0; b < nj; b++) {
        const int64_t n = jobs[b].n;
        nn[b] = (int32_t)n;
        kk[b] = (int32_t)jobs[b].k;
        aoff[b + 1] = aoff[b] + n * n;
        doff[b + 1] = doff[b] + n;
        zoff[b + 1] = zoff[b] + n * jobs[b].k;
    }
    DevMem mem;
    double *dA = nullptr, *dde = nullptr, *dZ = nullptr;
    int64_t *daoff = nullptr, *ddoff = nullptr, *dzoff = nullptr;
    int32_t *dnn = nullptr, *dkk = nullptr;
    int rc;
    if ((rc = mem.alloc(&dA, (size_t)aoff[nj])) || (rc = mem.alloc(&dde, (size_t)2 * doff[nj])) ||
        (rc = mem.alloc(&dZ, (size_t)zoff[nj])) || (rc = mem.alloc(&daoff, (size_t)nj + 1)) ||
        (rc = mem.alloc(&ddoff, (size_t)nj + 1)) || (rc = mem.alloc(&dzoff, (size_t)nj + 1)) ||
        (rc = mem.alloc(&dnn, (size_t)nj)) || (rc = mem.alloc(&dkk, (size_t)nj)))
        return rc;
    for (int64_t b = 0; b < nj; b++)
        if ((rc = copy_to_device(dA + aoff[b], jobs[b].A, (size_t)nn[b] * nn[b] * sizeof(double), st))) return rc;
    if ((rc = copy_to_device(daoff, aoff.data(), aoff.size() * sizeof(int64_t), st)) ||
        (rc = copy_to_device(ddoff, doff.data(), doff.size() * sizeof(int64_t), st)) ||
        (rc = copy_to_device(dzoff, zoff.data(), zoff.size() * sizeof(int64_t), st)) ||
        (rc = copy_to_device(dnn, nn.data(), nn.size() * sizeof(int32_t), st)) ||
        (rc = copy_to_device(dkk, kk.data(), kk.size() * sizeof(int32_t), st)))
        return rc;
    k_tridiag<<<dim3((unsigned)nj), dim3(kDT), 0, st>>>(dA, daoff, dnn, ddoff, dde);
    SEHIP_CHECK(hipGetLastError());
    std::vector<double> de((size_t)2 * doff[nj]);
    if ((rc = copy_to_host(de.data(), dde, de.size() * sizeof(double), st))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(st));
    // host: every eigenvalue by implicit QL, the k wanted by |lambda|, their
    // tridiagonal eigenvectors
    std::vector<double> Zh((size_t)zoff[nj]);
    std::vector<int> brc(nj, 0);
    parallel_for(nj, host_threads(), 1, [&](int64_t b, int) {
        DenseJob &J = jobs[b];
        const int64_t n = J.n;
        DenseTri T;
        const double *d = de.data() + 2 * doff[b];
        if (dense_tri_values_from(n, d, d + n, T)) {
            brc[b] = SPEIGS_ERR_LAPACK;
            return;
        }
        std::vector<int64_t> ord(n);
        for (int64_t i = 0; i < n; i++) ord[i] = i;
        std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) {
            const double ax = std::fabs(T.w[x]), ay = std::fabs(T.w[y]);
            return J.largest ? ax > ay : ax < ay;
        });
        J.sel.assign(ord.begin(), ord.begin() + J.k);
        for (int64_t t = 0; t < J.k; t++) J.lambda[t] = T.w[J.sel[t]];
        if (J.V && dense_tri_zvectors(T, J.k, J.sel.data(), Zh.data() + zoff[b], n)) brc[b] = SPEIGS_ERR_LAPACK;
    });
    for (int r : brc)
        if (r) return SEHIP_FAIL(r, "dense batch: eigenvalues or vectors of a tridiagonal");
    bool want_v = false;
    for (const DenseJob &J : jobs) want_v = want_v || J.V;
    if (!want_v) return 0;
    if ((rc = copy_to_device(dZ, Zh.data(), Zh.size() * sizeof(double), st))) return rc;
    int kmax = 1;
    for (int32_t x : kk) kmax = std::max(kmax, (int)x);
    k_tri_back<<<dim3((unsigned)nj, (unsigned)((kmax + 15) / 16)), dim3(1024), 0, st>>>(dA, daoff, dnn, dkk, dzoff, dZ);
    SEHIP_CHECK(hipGetLastError());
    if ((rc = copy_to_host(Zh.data(), dZ, Zh.size() * sizeof(double), st))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(st));
    for (int64_t b = 0; b < nj; b++)
        if (jobs[b].V) std::copy(Zh.begin() + zoff[b], Zh.begin() + zoff[b + 1], jobs[b].V);
    return 0;
}

}  // namespace sehip
