// lobpcg.cpp -- block eigensolver of the FP32-factor path.
//
// With an FP32 factor, refining every shift-invert apply to FP64 costs ~8 FP32
// solves at kappa(A) ~ 1e5 (each refinement sweep gains ~2 digits).  Here the
// FP32 factor is used once per iteration as the preconditioner T ~ (A+dI)^-1
// of a block method, applied to all m residuals in one multi-right-hand-side
// solve (one pass over L for m vectors), and the eigenpairs are refined to FP64
// accuracy by the outer iteration itself: Rayleigh-Ritz and residuals against
// A are FP64.  This is the block form the reference planned
// (docs/REQUIREMENTS.md:172-178: multiple right-hand sides turn BLAS-2 solves
// into BLAS-3), as LOBPCG (Knyazev 2001):
//   S = [X, W = T R, P],  R = A X - X Theta,  Rayleigh-Ritz of A on span(S),
//   X <- smallest m Ritz vectors, P <- their component outside X.
// Robustness: W and P are projected out of X twice and [W P] is made
// orthonormal by SVQB (Stathopoulos and Wu 2002: scale, eigendecompose the Gram
// matrix, drop directions below 1e-12), twice, so S stays orthonormal to
// rounding and the projected problem is a standard symmetric one.
// Convergence is the reference's test: ||A x - lambda x||_2 <= tol*max(||A||_1,1)
// for the k wanted pairs (src/speigs_core.c:243-265).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "factor_impl.h"

namespace sehip {

namespace {

struct LWork {
    int64_t n = 0, ld = 0;
    hipStream_t st = nullptr;
    DevMem mem;
    double *partial = nullptr, *G = nullptr, *C = nullptr, *red = nullptr;
};

int dgram(LWork &W, const double *A, int p, const double *B, int q, std::vector<double> &out)
{
    gram(A, W.ld, p, B, W.ld, q, W.n, W.partial, W.G, W.st);
    out.resize((size_t)p * q);
    SEHIP_CHECK(hipMemcpyAsync(out.data(), W.G, (size_t)p * q * sizeof(double), hipMemcpyDeviceToHost, W.st));
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    return 0;
}

// Y = V Cm + beta Y  (V n x m, Cm m x p on the host, column-major; Y must not alias V)
int dgemm_tall(LWork &W, const double *V, int m, const std::vector<double> &Cm, int p, double *Y, double beta)
{
    SEHIP_CHECK(hipMemcpyAsync(W.C, Cm.data(), (size_t)m * p * sizeof(double), hipMemcpyHostToDevice, W.st));
    tsgemm_beta(V, W.ld, m, W.C, p, Y, W.ld, W.n, beta, W.st);
    return 0;
}

// Z (columns m .. m+nz of S) <- Z - X (X^T Z); returns Z^T Z after the update
int project_out(LWork &W, const double *S, int m, double *Z, int nz, std::vector<double> &zz)
{
    std::vector<double> g;
    int rc = dgram(W, S, m + nz, Z, nz, g);           // rows 0..m-1: X^T Z, rows m..: Z^T Z
    if (rc) return rc;
    const int ld = m + nz;
    std::vector<double> gx((size_t)m * nz);
    zz.assign((size_t)nz * nz, 0.0);
    for (int j = 0; j < nz; j++) {
        for (int i = 0; i < m; i++) gx[(size_t)j * m + i] = -g[(size_t)j * ld + i];
        for (int i = 0; i < nz; i++) zz[(size_t)j * nz + i] = g[(size_t)j * ld + m + i];
    }
    // (Z - X Gx)^T (Z - X Gx) = Z^T Z - Gx^T Gx for orthonormal X
    for (int j = 0; j < nz; j++)
        for (int i = 0; i < nz; i++) {
            double acc = 0.0;
            for (int q = 0; q < m; q++) acc += gx[(size_t)i * m + q] * gx[(size_t)j * m + q];
            zz[(size_t)j * nz + i] -= acc;
        }
    return dgemm_tall(W, S, m, gx, nz, Z, 1.0);
}

// SVQB from the Gram matrix zz of Z: out = Z D U S^-1/2 (kept directions); returns r
int svqb(LWork &W, const double *Z, int nz, const std::vector<double> &zz, double *out, int *r_out)
{
    std::vector<double> d(nz);
    for (int i = 0; i < nz; i++) {
        const double g = zz[(size_t)i * nz + i];
        d[i] = g > 0 ? 1.0 / std::sqrt(g) : 0.0;
    }
    std::vector<double> gs((size_t)nz * nz), ev(nz);
    for (int j = 0; j < nz; j++)
        for (int i = 0; i < nz; i++) gs[(size_t)j * nz + i] = d[i] * 0.5 * (zz[(size_t)j * nz + i] + zz[(size_t)i * nz + j]) * d[j];
    if (dense_syev(nz, gs.data(), ev.data())) return SPEIGS_ERR_LAPACK;
    const double smax = ev[nz - 1];
    int r = 0;
    for (int j = nz - 1; j >= 0 && smax > 0 && ev[j] > 1e-12 * smax; j--) r++;
    std::vector<double> T((size_t)nz * r, 0.0);
    for (int c = 0; c < r; c++) {
        const int j = nz - 1 - c;
        const double sc = 1.0 / std::sqrt(ev[j]);
        for (int i = 0; i < nz; i++) T[(size_t)c * nz + i] = d[i] * gs[(size_t)j * nz + i] * sc;
    }
    if (r > 0) {
        int rc = dgemm_tall(W, Z, nz, T, r, out, 0.0);
        if (rc) return rc;
    }
    *r_out = r;
    return 0;
}

}  // namespace

int lobpcg_dev(const SellA &A, Factor *F, hipStream_t st, int64_t n, int64_t k, const speigs_opts *o,
               const speigs_hip_opts *ho, double anorm, double *lambda, double *Vout, double *resid,
               speigs_info *info, speigs_hip_info *hinfo)
{
    int m = (int)std::min<int64_t>(32, std::max<int64_t>(k + 8, k + k / 2));
    if (m > n / 2) m = (int)std::max<int64_t>(k, n / 2);
    if (k > m || m > 32) return SPEIGS_ERR_ARG;
    const int64_t maxit = o->maxit ? std::max<int64_t>(o->maxit, 50) : 500;
    const double tol = o->tol > 0 ? o->tol : 1e-14;
    const double rtol = tol * (anorm > 1.0 ? anorm : 1.0);
    int rc = factor_invert_l11(F);
    if (rc) return rc;

    LWork W;
    W.n = n;
    W.ld = pad_ld(n);
    W.st = st;
    const size_t col = (size_t)W.ld;
    double *S, *AS, *Sn, *ASn, *Pb, *T2;
    float *Rf, *Wf;
    if ((rc = W.mem.alloc(&S, col * 3 * m))) return rc;     // [X | W | P]
    if ((rc = W.mem.alloc(&AS, col * 3 * m))) return rc;
    if ((rc = W.mem.alloc(&Sn, col * 3 * m))) return rc;
    if ((rc = W.mem.alloc(&ASn, col * 3 * m))) return rc;
    if ((rc = W.mem.alloc(&Pb, col * m))) return rc;
    if ((rc = W.mem.alloc(&T2, col * 2 * m))) return rc;
    if ((rc = W.mem.alloc(&Rf, (size_t)n * m))) return rc;
    if ((rc = W.mem.alloc(&Wf, (size_t)n * m))) return rc;
    if ((rc = W.mem.alloc(&W.partial, (size_t)gram_blocks(n) * 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.G, 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.C, 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.red, 4 * 96))) return rc;

    // start block: the reference's xorshift64* stream (column 0 = its start vector)
    {
        uint64_t seed = o->seed ? o->seed : 88172645463325252ULL;
        std::vector<double> h((size_t)n);
        for (int t = 0; t < m; t++) {
            if (t == 0 && ho && ho->start == SPEIGS_HIP_START_VECTOR && ho->start_vec)
                std::memcpy(h.data(), ho->start_vec, (size_t)n * sizeof(double));
            else
                uniform_vec(&seed, n, h.data());
            SEHIP_CHECK(hipMemcpyAsync(T2 + t * col, h.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, st));
            SEHIP_CHECK(hipStreamSynchronize(st));
        }
        for (int pass = 0; pass < 2; pass++) {
            std::vector<double> g;
            if ((rc = dgram(W, T2, m, T2, m, g))) return rc;
            int r = 0;
            if ((rc = svqb(W, T2, m, g, S, &r))) return rc;
            if (r < m) return SPEIGS_ERR_INTERNAL;
            if (pass == 0) SEHIP_CHECK(hipMemcpyAsync(T2, S, col * m * sizeof(double), hipMemcpyDeviceToDevice, st));
        }
    }
    double *X = S, *AX = AS;
    spmv_multi(A, X, W.ld, AX, W.ld, m, st);
    std::vector<double> theta(m);
    {
        std::vector<double> g;
        if ((rc = dgram(W, X, m, AX, m, g))) return rc;
        for (int j = 0; j < m; j++)
            for (int i = 0; i < j; i++) g[(size_t)j * m + i] = g[(size_t)i * m + j] = 0.5 * (g[(size_t)j * m + i] + g[(size_t)i * m + j]);
        if (dense_syev(m, g.data(), theta.data())) return SPEIGS_ERR_LAPACK;
        if ((rc = dgemm_tall(W, X, m, g, m, Sn, 0.0))) return rc;
        if ((rc = dgemm_tall(W, AX, m, g, m, ASn, 0.0))) return rc;
        std::swap(S, Sn);
        std::swap(AS, ASn);
    }
    int np = 0;
    int64_t it = 0, nconv = 0, solves = 0;
    std::vector<double> res(m);
    double *lam_dev = W.red, *nrm_dev = W.red + 96;
    for (;; it++) {
        X = S;
        AX = AS;
        double *Z = S + col * m, *AZ = AS + col * m;
        SEHIP_CHECK(hipMemcpyAsync(lam_dev, theta.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
        block_resid(AX, X, W.ld, lam_dev, m, n, Z, st);
        pairdots_partial(Z, W.ld, Z, W.ld, m, n, W.partial, st);
        reduce_partial(W.partial, nblocks_rows(n), m, nrm_dev, st);
        SEHIP_CHECK(hipMemcpyAsync(res.data(), nrm_dev, m * sizeof(double), hipMemcpyDeviceToHost, st));
        SEHIP_CHECK(hipStreamSynchronize(st));
        nconv = 0;
        for (int i = 0; i < k; i++) {
            res[i] = std::sqrt(res[i]);
            if (res[i] <= rtol) nconv++;
        }
        if (nconv >= k || it >= maxit) break;
        // W = T R: one FP32 multi-right-hand-side solve
        block_to_f32(Z, W.ld, Rf, m, n, st);
        if ((rc = factor_block_solve(F, Rf, Wf, m))) return rc;
        block_to_f64(Wf, Z, W.ld, m, n, st);
        solves++;
        if (np > 0)
            SEHIP_CHECK(hipMemcpyAsync(Z + col * m, Pb, col * np * sizeof(double), hipMemcpyDeviceToDevice, st));
        int nz = m + np;
        std::vector<double> zz;
        if ((rc = project_out(W, S, m, Z, nz, zz))) return rc;
        if ((rc = project_out(W, S, m, Z, nz, zz))) return rc;
        int r = 0;
        if ((rc = svqb(W, Z, nz, zz, T2, &r))) return rc;
        if (r == 0) break;
        SEHIP_CHECK(hipMemcpyAsync(Z, T2, col * r * sizeof(double), hipMemcpyDeviceToDevice, st));
        nz = r;
        if ((rc = project_out(W, S, m, Z, nz, zz))) return rc;
        if ((rc = svqb(W, Z, nz, zz, T2, &r))) return rc;
        if (r == 0) break;
        SEHIP_CHECK(hipMemcpyAsync(Z, T2, col * r * sizeof(double), hipMemcpyDeviceToDevice, st));
        nz = r;
        spmv_multi(A, Z, W.ld, AZ, W.ld, nz, st);
        // Rayleigh-Ritz on S = [X Z] (orthonormal): H = S^T A S
        const int d = m + nz;
        std::vector<double> H;
        if ((rc = dgram(W, S, d, AS, d, H))) return rc;
        for (int j = 0; j < d; j++)
            for (int i = 0; i < j; i++) {
                const double v = 0.5 * (H[(size_t)j * d + i] + H[(size_t)i * d + j]);
                H[(size_t)j * d + i] = H[(size_t)i * d + j] = v;
            }
        std::vector<double> ev(d);
        if (dense_syev(d, H.data(), ev.data())) return SPEIGS_ERR_LAPACK;
        for (int t = 0; t < m; t++) theta[t] = ev[t];
        std::vector<double> C((size_t)d * m), Cz((size_t)nz * m);
        for (int t = 0; t < m; t++) {
            for (int i = 0; i < d; i++) C[(size_t)t * d + i] = H[(size_t)t * d + i];
            for (int i = 0; i < nz; i++) Cz[(size_t)t * nz + i] = H[(size_t)t * d + m + i];
        }
        if ((rc = dgemm_tall(W, S, d, C, m, Sn, 0.0))) return rc;      // new X
        if ((rc = dgemm_tall(W, AS, d, C, m, ASn, 0.0))) return rc;    // new A X
        if ((rc = dgemm_tall(W, Z, nz, Cz, m, Pb, 0.0))) return rc;    // new P
        std::swap(S, Sn);
        std::swap(AS, ASn);
        np = m;
        if ((it + 1) % 8 == 0) spmv_multi(A, S, W.ld, AS, W.ld, m, st);   // refresh A X against drift
    }
    X = S;
    AX = AS;
    // final pairs: exact A X, Rayleigh quotients, residuals
    spmv_multi(A, X, W.ld, AX, W.ld, m, st);
    pairdots_partial(X, W.ld, AX, W.ld, m, n, W.partial, st);
    reduce_partial(W.partial, nblocks_rows(n), m, lam_dev, st);
    pairdots_partial(X, W.ld, X, W.ld, m, n, W.partial, st);
    reduce_partial(W.partial, nblocks_rows(n), m, nrm_dev, st);
    std::vector<double> lam(m), xx(m);
    SEHIP_CHECK(hipMemcpyAsync(lam.data(), lam_dev, m * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipMemcpyAsync(xx.data(), nrm_dev, m * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    for (int t = 0; t < m; t++) lam[t] /= xx[t];
    SEHIP_CHECK(hipMemcpyAsync(lam_dev, lam.data(), m * sizeof(double), hipMemcpyHostToDevice, st));
    residual_partial(X, W.ld, AX, W.ld, lam_dev, m, n, W.partial, st);
    reduce_partial(W.partial, nblocks_rows(n), m, nrm_dev, st);
    SEHIP_CHECK(hipMemcpyAsync(res.data(), nrm_dev, m * sizeof(double), hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    nconv = 0;
    for (int t = 0; t < k; t++) {
        const double r = std::sqrt(res[t]) / std::sqrt(xx[t]);
        lambda[t] = lam[t];
        if (resid) resid[t] = r;
        if (r <= rtol) nconv++;
    }
    if (Vout)
        SEHIP_CHECK(hipMemcpy2DAsync(Vout, (size_t)n * sizeof(double), X, col * sizeof(double), (size_t)n * sizeof(double),
                                     (size_t)k, hipMemcpyDeviceToHost, st));
    SEHIP_CHECK(hipStreamSynchronize(st));
    info->nconv = nconv;
    info->restarts = it;
    info->nops = solves * m;
    info->flag = nconv >= k ? 0 : 1;
    if (hinfo) hinfo->solves = solves;
    return SPEIGS_OK;
}

}  // namespace sehip
