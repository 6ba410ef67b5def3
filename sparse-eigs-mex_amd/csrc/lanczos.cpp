// lanczos.cpp -- thick-restart Lanczos with full CGS2 reorthogonalisation over
// a device operator (reference src/speigs_core.c:146-315).
//
// Control stays on the host exactly as in the reference loop; every O(n)
// vector lives in HBM.  Per Krylov step the device runs: one operator apply,
// three passes over the basis (V^T w | w -= V h1 and V^T w | w -= V h2 and
// |w|^2), and the host receives j+1 CGS coefficients of each pass plus |w|^2 in
// one small copy.  The m x m projected problem is solved on the host (Jacobi:
// relative accuracy for the positive definite shift-invert projection, whose
// largest Ritz value 1/delta can dwarf the wanted ones).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dev.h"
#include "host_par.h"

namespace sehip {

uint64_t xorshift_next(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}

namespace {

// The xorshift step is linear over GF(2)^64: a map is stored as the images of
// the 64 basis vectors, so that the state J steps ahead is a product with the
// J-th power of the step's map (Brent's jump-ahead for xorshift generators).
struct Gf2Map {
    uint64_t col[64];
};

uint64_t gf2_apply(const Gf2Map &A, uint64_t x)
{
    uint64_t y = 0;
    for (int b = 0; b < 64; b++)
        if ((x >> b) & 1) y ^= A.col[b];
    return y;
}

Gf2Map gf2_compose(const Gf2Map &A, const Gf2Map &B)   // A after B
{
    Gf2Map C;
    for (int b = 0; b < 64; b++) C.col[b] = gf2_apply(A, B.col[b]);
    return C;
}

Gf2Map xorshift_power(int64_t J)
{
    Gf2Map step, P;
    for (int b = 0; b < 64; b++) {
        uint64_t x = (uint64_t)1 << b;
        (void)xorshift_next(&x);
        step.col[b] = x;
        P.col[b] = (uint64_t)1 << b;
    }
    for (; J > 0; J >>= 1) {
        if (J & 1) P = gf2_compose(step, P);
        step = gf2_compose(step, step);
    }
    return P;
}

}  // namespace

// v[i] from the i-th state of the sequence; long vectors in chunks on host
// threads, each chunk started at its state by a jump-ahead (the same values)
void uniform_vec(uint64_t *s, int64_t n, double *v)
{
    auto run = [v](uint64_t st, int64_t i0, int64_t i1) {
        for (int64_t i = i0; i < i1; i++) v[i] = (double)(xorshift_next(&st) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        return st;
    };
    const int T = (int)std::min<int64_t>(host_threads(), n / 65536);
    if (T <= 1) {
        *s = run(*s, 0, n);
        return;
    }
    const int64_t len = (n + T - 1) / T;
    const Gf2Map jump = xorshift_power(len);
    std::vector<uint64_t> st0(T), st1(T);
    st0[0] = *s;
    for (int t = 1; t < T; t++) st0[t] = gf2_apply(jump, st0[t - 1]);
    parallel_for(T, T, 1, [&](int64_t t, int) {
        st1[t] = run(st0[t], t * len, std::min(n, (t + 1) * len));
    });
    *s = st1[T - 1];
}

static double host_sumsq(int64_t n, const double *x)
{
    if (n <= 32) {
        double s = 0.0;
        for (int64_t i = 0; i < n; i++) s += x[i] * x[i];
        return s;
    }
    return host_sumsq(n / 2, x) + host_sumsq(n - n / 2, x + n / 2);
}

namespace {

struct KeyIdx {
    double key;
    int64_t idx;
};
int by_key_desc(const void *a, const void *b)
{
    double x = ((const KeyIdx *)a)->key, y = ((const KeyIdx *)b)->key;
    return (x < y) - (x > y);
}

struct Work {
    int64_t n = 0, ld = 0, nb = 0;
    int ncv0 = 0;
    double *V = nullptr, *Vn = nullptr, *w = nullptr, *X = nullptr, *AX = nullptr, *T = nullptr;
    double *partial = nullptr, *red = nullptr, *S = nullptr;
    double *hred = nullptr;     // pinned host copy of red
    hipStream_t st = nullptr;
    DevMem mem;
    ~Work()
    {
        if (hred) (void)hipHostFree(hred);
    }
};

// CGS2 of w against B[:, 0:m] (device); returns |w|^2 on the host and, when
// hout != nullptr, adds the pass-1 and pass-2 coefficients into hout[0:m].
int cgs2(Work &W, const double *B, int m, double *w, double *hout, double *nrm2)
{
    const int mr = W.ncv0 + 2;   // stride of the red buffer segments
    double *h1 = W.red, *h2 = W.red + mr, *nr = W.red + 2 * mr;
    if (m > 0) {
        coldots_partial(B, W.ld, m, w, W.n, W.partial, W.st);
        reduce_partial(W.partial, W.nb, m, h1, W.st);
        cgs_update(B, W.ld, m, h1, w, W.n, W.partial, 1, W.st);
        reduce_partial(W.partial, W.nb, m, h2, W.st);
        cgs_update(B, W.ld, m, h2, w, W.n, W.partial, 2, W.st);
        reduce_partial(W.partial, W.nb, 1, nr, W.st);
    } else {
        pairdots_partial(w, W.ld, w, W.ld, 1, W.n, W.partial, W.st);
        reduce_partial(W.partial, W.nb, 1, nr, W.st);
    }
    SEHIP_CHECK(hipMemcpyAsync(W.hred, W.red, (size_t)(3 * mr) * sizeof(double), hipMemcpyDeviceToHost, W.st));
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    if (hout)
        for (int i = 0; i < m; i++) hout[i] = (hout[i] + W.hred[i]) + W.hred[mr + i];
    *nrm2 = W.hred[2 * mr];
    return 0;
}

int upload_random(Work &W, uint64_t *seed, double *dst, std::vector<double> &hbuf)
{
    uniform_vec(seed, W.n, hbuf.data());
    SEHIP_CHECK(hipMemcpyAsync(dst, hbuf.data(), (size_t)W.n * sizeof(double), hipMemcpyHostToDevice, W.st));
    return 0;
}

}  // namespace

int lanczos_dev(DevOperator *op, int64_t n, int64_t k, const speigs_opts *o, const speigs_hip_opts *ho,
                double anorm, double *lambda, double *Vout, double *resid, speigs_info *info)
{
    int64_t ncv = o->ncv ? o->ncv : std::max<int64_t>(2 * k + 1, 20);
    if (ncv > n) ncv = n;
    if (ncv <= k) ncv = std::min<int64_t>(n, k + 1);
    if (ncv < 2) ncv = std::min<int64_t>(n, 2);
    if (ncv > 4000) return SPEIGS_ERR_ARG;      // LDS staging of the CGS coefficients
    const int64_t ldh = ncv;
    const int64_t maxit = o->maxit ? o->maxit : 300;
    const double tol = o->tol > 0 ? o->tol : 1e-14;
    const double scale = anorm > 1.0 ? anorm : 1.0;
    const double rtol = tol * scale, small = 1e-13 * scale;
    // the caller's eigenvector array is faulted in on helper threads while the
    // device iterates
    HostTouch touch;
    if (Vout) touch.start(Vout, (size_t)n * std::min<int64_t>(k, ncv) * sizeof(double));

    Work W;
    W.n = n;
    W.ld = pad_ld(n);
    W.nb = nblocks_rows(n);
    W.ncv0 = (int)ncv;
    W.st = op->stream;
    const int64_t kcap = std::max<int64_t>({k, ncv, 1});
    int rc;
    if ((rc = W.mem.alloc(&W.V, (size_t)W.ld * (ncv + 1)))) return rc;
    if ((rc = W.mem.alloc(&W.Vn, (size_t)W.ld * (ncv + 1)))) return rc;
    if ((rc = W.mem.alloc(&W.w, (size_t)W.ld))) return rc;
    if ((rc = W.mem.alloc(&W.X, (size_t)W.ld * kcap))) return rc;
    if ((rc = W.mem.alloc(&W.AX, (size_t)W.ld * kcap))) return rc;
    if ((rc = W.mem.alloc(&W.T, (size_t)W.ld * kcap))) return rc;
    if ((rc = W.mem.alloc(&W.partial, (size_t)W.nb * (kcap + 2)))) return rc;
    if ((rc = W.mem.alloc(&W.red, (size_t)3 * (ncv + 2) + 2 * kcap))) return rc;
    if ((rc = W.mem.alloc(&W.S, (size_t)(ncv + 1) * kcap))) return rc;
    if (hipHostMalloc((void **)&W.hred, (size_t)(3 * (ncv + 2) + 2 * kcap) * sizeof(double)) != hipSuccess)
        return SPEIGS_ERR_MEM;
    SEHIP_CHECK(hipMemsetAsync(W.V, 0, (size_t)W.ld * (ncv + 1) * sizeof(double), W.st));
    SEHIP_CHECK(hipMemsetAsync(W.Vn, 0, (size_t)W.ld * (ncv + 1) * sizeof(double), W.st));

    std::vector<double> H((size_t)ncv * ncv, 0.0), Hs((size_t)ncv * ncv), Z((size_t)ncv * ncv);
    std::vector<double> theta(ncv), keep(ncv), hv(n), sel((size_t)(ncv + 1) * kcap), tmp(kcap);
    std::vector<KeyIdx> ord(ncv);

    uint64_t seed = o->seed ? o->seed : 88172645463325252ULL;
    if (ho && ho->start == SPEIGS_HIP_START_VECTOR && ho->start_vec) {
        std::memcpy(hv.data(), ho->start_vec, (size_t)n * sizeof(double));
    } else {
        uniform_vec(&seed, n, hv.data());
    }
    {
        double nr = std::sqrt(host_sumsq(n, hv.data()));
        if (nr == 0.0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "re-injected start vector has norm 0");
        double s = 1.0 / nr;
        for (int64_t i = 0; i < n; i++) hv[i] *= s;
        if ((rc = copy_to_device(W.V, hv.data(), (size_t)n * sizeof(double), W.st))) return rc;
    }

    int64_t j = 0, restarts = 0, nconv = 0;
    const int mr = (int)ncv + 2;
    double *lam_dev = W.red + 3 * mr;
    for (;;) {
        for (; j < ncv; j++) {
            double ts = wall();
            rc = op->apply(W.V + (size_t)j * W.ld, W.w);
            op->nops++;
            if (rc) return rc;
            SEHIP_CHECK(hipStreamSynchronize(W.st));
            if ((rc = op->check())) return rc;
            info->t_op += wall() - ts;
            ts = wall();
            double *hc = H.data() + (size_t)j * ldh;
            for (int64_t i = 0; i <= j; i++) hc[i] = 0.0;
            double b2 = 0.0;
            if ((rc = cgs2(W, W.V, (int)(j + 1), W.w, hc, &b2))) return rc;
            info->t_ortho += wall() - ts;
            double beta = std::sqrt(b2), beta_rec = beta;
            if (beta <= small) {
                beta_rec = 0.0;
                for (int att = 0; att < 3 && beta <= small; att++) {
                    if ((rc = upload_random(W, &seed, W.w, hv))) return rc;
                    if ((rc = cgs2(W, W.V, (int)(j + 1), W.w, nullptr, &b2))) return rc;
                    beta = std::sqrt(b2);
                }
                if (beta <= small) {
                    ncv = j + 1;
                    break;
                }
            }
            scale_copy(W.w, 1.0 / beta, W.V + (size_t)(j + 1) * W.ld, n, W.st);
            if (j + 1 < ncv) {
                H[(j + 1) + (size_t)j * ldh] = beta_rec;
                H[j + (size_t)(j + 1) * ldh] = beta_rec;
            }
        }
        // projected problem (leading dimension quirk of src/speigs_core.c:233-234 kept:
        // after a breakdown shrank ncv, H is read with the new ncv as its stride)
        const int64_t m = ncv;
        for (int64_t c = 0; c < m; c++)
            for (int64_t r = 0; r < m; r++) Hs[r + c * m] = H[r + c * ncv];
        if (jacobi_eig(m, Hs.data(), m, Z.data(), theta.data())) return SPEIGS_ERR_LAPACK;
        for (int64_t i = 0; i < m; i++) {
            ord[i].key = std::fabs(theta[i]);
            ord[i].idx = i;
        }
        qsort(ord.data(), (size_t)m, sizeof(KeyIdx), by_key_desc);

        // residuals of the wanted Ritz pairs against A
        nconv = 0;
        const int64_t nwant = std::min<int64_t>(k, m);
        double tr = wall();
        for (int64_t t = 0; t < nwant; t++)
            std::memcpy(sel.data() + (size_t)t * m, Z.data() + (size_t)ord[t].idx * m, (size_t)m * sizeof(double));
        SEHIP_CHECK(hipMemcpyAsync(W.S, sel.data(), (size_t)m * nwant * sizeof(double), hipMemcpyHostToDevice, W.st));
        tsgemm(W.V, W.ld, (int)m, W.S, (int)nwant, W.X, W.ld, n, W.st);
        pairdots_partial(W.X, W.ld, W.X, W.ld, (int)nwant, n, W.partial, W.st);
        reduce_partial(W.partial, W.nb, (int)nwant, lam_dev, W.st);
        SEHIP_CHECK(hipMemcpyAsync(W.hred, lam_dev, (size_t)nwant * sizeof(double), hipMemcpyDeviceToHost, W.st));
        SEHIP_CHECK(hipStreamSynchronize(W.st));
        for (int64_t t = 0; t < nwant; t++) {
            double xn = std::sqrt(W.hred[t]);
            tmp[t] = xn > 0 ? 1.0 / xn : 1.0;
        }
        SEHIP_CHECK(hipMemcpyAsync(lam_dev, tmp.data(), (size_t)nwant * sizeof(double), hipMemcpyHostToDevice, W.st));
        scale_cols(W.X, W.ld, lam_dev, (int)nwant, n, W.st);
        if ((rc = op->multA(W.X, W.ld, W.AX, W.ld, (int)nwant))) return rc;
        pairdots_partial(W.X, W.ld, W.AX, W.ld, (int)nwant, n, W.partial, W.st);
        reduce_partial(W.partial, W.nb, (int)nwant, lam_dev, W.st);
        residual_partial(W.X, W.ld, W.AX, W.ld, lam_dev, (int)nwant, n, W.partial, W.st);
        reduce_partial(W.partial, W.nb, (int)nwant, lam_dev + kcap, W.st);
        SEHIP_CHECK(hipMemcpyAsync(W.hred, lam_dev, (size_t)(kcap + nwant) * sizeof(double), hipMemcpyDeviceToHost,
                                   W.st));
        SEHIP_CHECK(hipStreamSynchronize(W.st));
        for (int64_t t = 0; t < nwant; t++) {
            double r = std::sqrt(W.hred[kcap + t]);
            lambda[t] = W.hred[t];
            if (resid) resid[t] = r;
            if (r <= rtol) nconv++;
        }
        info->t_resid += wall() - tr;
        if (nconv >= std::min<int64_t>(k, m) || restarts >= maxit || m < ncv) break;
        restarts++;

        // thick restart: l dominant Ritz vectors, orthonormalised one by one
        int64_t l = std::min<int64_t>(k + 5, ncv - 1);
        if (l < 1) l = 1;
        int64_t nret = 0, computed = 0, base = 0;
        double tk = wall();
        for (int64_t t = 0; t < m && nret < l; t++) {
            if (t >= computed) {    // next batch of candidate Ritz vectors into T
                int64_t p = std::min<int64_t>(l, m - t);
                p = std::min<int64_t>(p, kcap);
                for (int64_t q = 0; q < p; q++)
                    std::memcpy(sel.data() + (size_t)q * m, Z.data() + (size_t)ord[t + q].idx * m,
                                (size_t)m * sizeof(double));
                SEHIP_CHECK(hipMemcpyAsync(W.S, sel.data(), (size_t)m * p * sizeof(double), hipMemcpyHostToDevice, W.st));
                tsgemm(W.V, W.ld, (int)m, W.S, (int)p, W.T, W.ld, n, W.st);
                base = t;
                computed = t + p;
            }
            double *cand = W.T + (size_t)(t - base) * W.ld;
            double nr2 = 0.0;
            if ((rc = cgs2(W, W.Vn, (int)nret, cand, nullptr, &nr2))) return rc;
            double nr = std::sqrt(nr2);
            if (nr < 1e-10) continue;
            scale_copy(cand, 1.0 / nr, W.Vn + (size_t)nret * W.ld, n, W.st);
            keep[nret] = theta[ord[t].idx];
            nret++;
        }
        info->t_ritz += wall() - tk;
        SEHIP_CHECK(hipMemcpyAsync(W.w, W.V + (size_t)m * W.ld, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                   W.st));
        double nr2 = 0.0;
        if ((rc = cgs2(W, W.Vn, (int)nret, W.w, nullptr, &nr2))) return rc;
        double nr = std::sqrt(nr2);
        if (nr < 1e-10) {
            if ((rc = upload_random(W, &seed, W.w, hv))) return rc;
            if ((rc = cgs2(W, W.Vn, (int)nret, W.w, nullptr, &nr2))) return rc;
            nr = std::sqrt(nr2);
        }
        if (nr < 1e-10) break;
        scale_copy(W.w, 1.0 / nr, W.Vn + (size_t)nret * W.ld, n, W.st);
        SEHIP_CHECK(hipMemcpyAsync(W.V, W.Vn, (size_t)W.ld * (nret + 1) * sizeof(double), hipMemcpyDeviceToDevice,
                                   W.st));
        std::fill(H.begin(), H.end(), 0.0);
        for (int64_t c = 0; c < nret; c++) H[c + (size_t)c * ldh] = keep[c];
        j = nret;
    }
    if (Vout) {
        const int64_t nw = std::min<int64_t>(k, ncv);
        touch.join();
        int rc2 = copy_to_host_2d(Vout, (size_t)n * sizeof(double), W.X, (size_t)W.ld * sizeof(double),
                                  (size_t)n * sizeof(double), (size_t)nw, W.st);
        if (rc2) return rc2;
    }
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    if ((rc = op->check())) return rc;
    info->nconv = nconv;
    info->restarts = restarts;
    info->nops = op->nops;
    info->flag = (nconv >= std::min<int64_t>(k, ncv)) ? 0 : 1;
    return SPEIGS_OK;
}

}  // namespace sehip
