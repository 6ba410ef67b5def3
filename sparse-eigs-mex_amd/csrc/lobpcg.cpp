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
// Robustness: [W P] is projected out of X and made orthonormal by SVQB
// (Stathopoulos and Wu 2002: scale, eigendecompose the Gram matrix, drop
// directions below 1e-12) -- both from one Gram matrix [X Z]^T Z and applied as
// one product over [X Z] -- and the pair of steps is repeated when the first can
// have fallen short (cancellation in the projection, a dropped direction, an
// ill-conditioned Gram matrix), so S stays orthonormal to rounding and the
// projected problem is a standard symmetric one.
// Blocks are row-major in HBM (see LWork).
// Convergence is the reference's test: ||A x - lambda x||_2 <= tol*max(||A||_1,1)
// for the k wanted pairs (src/speigs_core.c:243-265).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "factor_impl.h"

namespace sehip {

namespace {

// Blocks live in HBM row-major (element (r, c) at X[r * ld + c]): S = [X | W | P]
// and AS = A S share the row stride ldS, so [X Z] (Z = [W P]) is one block and
// its Gram and products stream contiguous rows.
struct LWork {
    int64_t n = 0;
    hipStream_t st = nullptr;
    DevMem mem;
    double *partial = nullptr, *G = nullptr, *C = nullptr, *red = nullptr, *npart = nullptr, *nrm = nullptr;
    double *rpart = nullptr;    // the Ritz step's per-workgroup residual sums
    // pinned host staging of the small transfers (Gram matrices, coefficient
    // blocks, Ritz values, norms): a pageable copy of a size not seen before
    // can stall the host for milliseconds in the runtime's staging set-up
    double *hbuf = nullptr;
    LWork() = default;
    LWork(const LWork &) = delete;
    LWork &operator=(const LWork &) = delete;
    ~LWork()
    {
        if (hbuf) (void)hipHostFree(hbuf);
    }
};
constexpr size_t kHostIn = 96 * 96, kHostOut = 96 * 96;   // hbuf: [device -> host | host -> device]

// device -> host through the pinned buffer (count <= kHostIn doubles); waits
int to_host(LWork &W, double *h, const double *dev, size_t count)
{
    double tc = 0.0;
    if (trace_on()) {
        SEHIP_CHECK(hipStreamSynchronize(W.st));
        tc = wall();
    }
    SEHIP_CHECK(hipMemcpyAsync(W.hbuf, dev, count * sizeof(double), hipMemcpyDeviceToHost, W.st));
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    if (trace_on() && wall() - tc > 5e-4) {
        char msg[96];
        snprintf(msg, sizeof msg, "  copy back of %zu bytes (slow)", count * sizeof(double));
        trace_phase(msg, tc);
    }
    std::memcpy(h, W.hbuf, count * sizeof(double));
    return 0;
}

// host -> device through the pinned buffer (count <= kHostOut doubles): the
// stream is drained first, so that no copy still reads the buffer (the device
// is idle at every call site: it waits for these coefficients)
int to_dev(LWork &W, double *dev, const double *h, size_t count)
{
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    std::memcpy(W.hbuf + kHostIn, h, count * sizeof(double));
    SEHIP_CHECK(hipMemcpyAsync(dev, W.hbuf + kHostIn, count * sizeof(double), hipMemcpyHostToDevice, W.st));
    return 0;
}

// R = AX - X theta into R (m columns) and the squared norms of its first k columns
int resid_norms(LWork &W, const double *AX, const double *X, int64_t ld, const double *lam_dev, int m, int k,
                double *R, std::vector<double> &out)
{
    rresid_norms(AX, ld, X, ld, lam_dev, m, k, W.n, R, ld, W.npart, W.nrm, W.st);
    out.resize(k);
    return to_host(W, out.data(), W.nrm, (size_t)k);
}

int dgram(LWork &W, const double *A, int64_t lda, int p, const double *B, int64_t ldb, int q,
          std::vector<double> &out)
{
    rgram(A, lda, p, B, ldb, q, W.n, W.partial, W.G, W.st);
    out.resize((size_t)p * q);
    return to_host(W, out.data(), W.G, (size_t)p * q);
}

// Y = V Cm + beta Y  (V n x m, Cm m x p on the host, column-major; Y may be V itself)
int dgemm_tall(LWork &W, const double *V, int64_t ldv, int m, const std::vector<double> &Cm, int p, double *Y,
               int64_t ldy, double beta)
{
    const int rc = to_dev(W, W.C, Cm.data(), (size_t)m * p);
    if (rc) return rc;
    rgemm(V, ldv, m, W.C, p, Y, ldy, W.n, beta, W.st);
    return 0;
}

// SVQB transform from the Gram matrix zz of Z: T = D U S^-1/2 over the kept
// directions (nz x r, column-major), so that Z T is orthonormal; returns r and,
// in *cond, the spread of the kept eigenvalues of the scaled Gram matrix
int svqb_transform(int nz, const std::vector<double> &zz, std::vector<double> &T, int *r_out, double *cond)
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
    if (cond) *cond = r > 0 ? smax / ev[nz - r] : 0.0;
    T.assign((size_t)nz * r, 0.0);
    for (int c = 0; c < r; c++) {
        const int j = nz - 1 - c;
        const double sc = 1.0 / std::sqrt(ev[j]);
        for (int i = 0; i < nz; i++) T[(size_t)c * nz + i] = d[i] * gs[(size_t)j * nz + i] * sc;
    }
    *r_out = r;
    return 0;
}

// Cholesky form of the same orthonormalisation: with the Gram matrix scaled to
// unit diagonal, D zz D = R^T R and T = D R^-1 (upper triangular).  Returns 1
// (use SVQB) when a pivot falls below 1e-8 -- a Gram matrix too close to
// singular, whose near-null directions SVQB drops; *cond = (max R_ii / min
// R_ii)^2 estimates the scaled Gram matrix's condition number.
int chol_transform(int nz, const std::vector<double> &zz, std::vector<double> &T, double *cond)
{
    std::vector<double> d(nz), R((size_t)nz * nz, 0.0);
    for (int i = 0; i < nz; i++) {
        const double g = zz[(size_t)i * nz + i];
        if (!(g > 0)) return 1;
        d[i] = 1.0 / std::sqrt(g);
    }
    // upper R, column-major: R[j * nz + i], i <= j
    for (int j = 0; j < nz; j++) {
        for (int i = 0; i <= j; i++) {
            double acc = d[i] * 0.5 * (zz[(size_t)j * nz + i] + zz[(size_t)i * nz + j]) * d[j];
            for (int q = 0; q < i; q++) acc -= R[(size_t)i * nz + q] * R[(size_t)j * nz + q];
            if (i < j) {
                R[(size_t)j * nz + i] = acc / R[(size_t)i * nz + i];
            } else {
                if (!(acc > 1e-8)) return 1;
                R[(size_t)j * nz + j] = std::sqrt(acc);
            }
        }
    }
    double rmax = 0.0, rmin = 1e300;
    for (int i = 0; i < nz; i++) {
        rmax = std::max(rmax, R[(size_t)i * nz + i]);
        rmin = std::min(rmin, R[(size_t)i * nz + i]);
    }
    *cond = (rmax / rmin) * (rmax / rmin);
    // T = D R^-1: back substitution column by column
    T.assign((size_t)nz * nz, 0.0);
    for (int j = 0; j < nz; j++) {
        T[(size_t)j * nz + j] = 1.0 / R[(size_t)j * nz + j];
        for (int i = j - 1; i >= 0; i--) {
            double acc = 0.0;
            for (int q = i + 1; q <= j; q++) acc += R[(size_t)q * nz + i] * T[(size_t)j * nz + q];
            T[(size_t)j * nz + i] = -acc / R[(size_t)i * nz + i];
        }
    }
    for (int j = 0; j < nz; j++)
        for (int i = 0; i <= j; i++) T[(size_t)j * nz + i] *= d[i];
    return 0;
}

// SVQB of Z: out = Z T (out may be Z itself: the product is row-local)
int svqb(LWork &W, const double *Z, int64_t ldz, int nz, const std::vector<double> &zz, double *out, int64_t ldo,
         int *r_out, double *cond = nullptr)
{
    std::vector<double> T;
    int rc = svqb_transform(nz, zz, T, r_out, cond);
    if (rc) return rc;
    if (*r_out > 0) return dgemm_tall(W, Z, ldz, nz, T, *r_out, out, ldo, 0.0);
    return 0;
}

// Z (columns m .. m+nz of S) <- (Z - X X^T Z) T with T the SVQB transform of
// the projected block, as one product over [X Z]: Z <- [X Z] [-X^T Z T ; T].
// The projected Gram matrix is formed from [X Z]^T Z (X orthonormal); *keep is
// the smallest ratio of a column's norm after the projection to before, *cond
// the SVQB spread; returns the kept count in *r
int project_svqb(LWork &W, double *S, int64_t ldS, int m, int nz, int *r, double *keep, double *cond)
{
    double *Z = S + m;
    std::vector<double> g;
    double tp = wall();
    int rc = dgram(W, S, ldS, m + nz, Z, ldS, nz, g);   // rows 0..m-1: X^T Z, rows m..: Z^T Z
    if (rc) return rc;
    trace_phase("  projection: Gram (device, copy back)", tp);
    const int ld = m + nz;
    std::vector<double> zz((size_t)nz * nz);
    for (int j = 0; j < nz; j++)
        for (int i = 0; i < nz; i++) {
            double acc = g[(size_t)j * ld + m + i];
            for (int q = 0; q < m; q++) acc -= g[(size_t)i * ld + q] * g[(size_t)j * ld + q];
            zz[(size_t)j * nz + i] = acc;
        }
    double kmin = 1.0;
    for (int j = 0; j < nz; j++) {
        const double b = g[(size_t)j * ld + m + j], a = zz[(size_t)j * nz + j];
        if (b > 0) kmin = std::min(kmin, std::sqrt(std::max(a, 0.0) / b));
    }
    *keep = kmin;
    // the Cholesky form when the projected block is well conditioned, SVQB
    // (which can drop directions) otherwise
    std::vector<double> T;
    trace_phase("  projection: projected Gram (host)", tp);
    if (chol_transform(nz, zz, T, cond) == 0) {
        *r = nz;
        trace_phase("  projection: Cholesky form (host)", tp);
    } else {
        if ((rc = svqb_transform(nz, zz, T, r, cond))) return rc;
        trace_phase("  projection: SVQB (host)", tp);
    }
    const int rr = *r;
    if (rr == 0) return 0;
    // C = [-(X^T Z) T ; T], (m + nz) x rr column-major
    std::vector<double> C((size_t)ld * rr);
    for (int c = 0; c < rr; c++) {
        for (int q = 0; q < m; q++) {
            double acc = 0.0;
            for (int j = 0; j < nz; j++) acc += g[(size_t)j * ld + q] * T[(size_t)c * nz + j];
            C[(size_t)c * ld + q] = -acc;
        }
        for (int j = 0; j < nz; j++) C[(size_t)c * ld + m + j] = T[(size_t)c * nz + j];
    }
    trace_phase("  projection: coefficients (host)", tp);
    rc = dgemm_tall(W, S, ldS, ld, C, rr, Z, ldS, 0.0);
    trace_phase("  projection: product queued", tp);
    return rc;
}

void symmetrize(std::vector<double> &H, int d)
{
    for (int j = 0; j < d; j++)
        for (int i = 0; i < j; i++) {
            const double v = 0.5 * (H[(size_t)j * d + i] + H[(size_t)i * d + j]);
            H[(size_t)j * d + i] = H[(size_t)i * d + j] = v;
        }
}

}  // namespace

int lobpcg_block_size(int64_t n, int64_t k)
{
    // k + 8 guard vectors: at 100^3, k = 20, blocks of 28..32 all take 12
    // iterations (26: 13, 24: 14), so the smallest of them is the cheapest
    if (k + 8 > kLobpcgMaxBlock) return 0;
    int64_t m = k + 8;
    if (m > n / 2) m = std::max<int64_t>(k, n / 2);
    return k > m ? 0 : (int)m;
}

int lobpcg_dev(const SellA &A, Factor *F, hipStream_t st, int64_t n, int64_t k, const speigs_opts *o,
               const speigs_hip_opts *ho, double anorm, double *lambda, double *Vout, double *resid,
               speigs_info *info, speigs_hip_info *hinfo, const float *Xstart)
{
    const int m = lobpcg_block_size(n, k);
    if (m == 0) return SEHIP_FAIL(SPEIGS_ERR_ARG, "LOBPCG: k too large for one block");
    // an iteration plays the part of a restart cycle: the reference's cap
    // (maxit restarts, 0 -> 300; src/speigs_core.c:156)
    const int64_t maxit = o->maxit ? o->maxit : 300;
    const double tol = o->tol > 0 ? o->tol : 1e-14;
    const double rtol = tol * (anorm > 1.0 ? anorm : 1.0);
    // the caller's eigenvector array is faulted in on helper threads while the
    // device iterates
    HostTouch touch;
    if (Vout) touch.start(Vout, (size_t)n * k * sizeof(double));
    int rc = 0;

    LWork W;
    W.n = n;
    W.st = st;
    const int64_t ldS = (3 * m + 3) / 4 * 4, ldT = (2 * m + 3) / 4 * 4;
    double *S, *AS, *Sn, *ASn, *T2, *v0;
    if ((rc = W.mem.alloc(&S, (size_t)n * ldS))) return rc;     // [X | W | P]
    if ((rc = W.mem.alloc(&AS, (size_t)n * ldS))) return rc;
    if ((rc = W.mem.alloc(&Sn, (size_t)n * ldS))) return rc;
    if ((rc = W.mem.alloc(&ASn, (size_t)n * ldS))) return rc;
    if ((rc = W.mem.alloc(&T2, (size_t)n * ldT))) return rc;
    if ((rc = W.mem.alloc(&v0, (size_t)n))) return rc;
    if ((rc = W.mem.alloc(&W.partial, (size_t)rgram_blocks(n) * 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.G, 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.C, 96 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.red, 4 * 96))) return rc;
    if ((rc = W.mem.alloc(&W.npart, (size_t)rresid_norms_part()))) return rc;
    if ((rc = W.mem.alloc(&W.nrm, 32))) return rc;
    if (hipHostMalloc((void **)&W.hbuf, (kHostIn + kHostOut) * sizeof(double)) != hipSuccess)
        return SEHIP_FAIL(SPEIGS_ERR_MEM, "LOBPCG: pinned host staging");
    if ((rc = W.mem.alloc(&W.rpart, (size_t)rritz_part(n)))) return rc;

    // The start vector is generated and uploaded (on the factor's second
    // stream) by a helper thread while the device inverts the diagonal blocks.
    uint64_t seed = o->seed ? o->seed : 88172645463325252ULL;
    std::vector<double> h;
    int start_rc = 0;
    std::string start_msg;
    std::thread start_th;
    struct JoinStart {
        std::thread &t;
        ~JoinStart()
        {
            if (t.joinable()) t.join();
        }
    } join_start{start_th};
    if (!Xstart) {
        int dev = 0;
        SEHIP_CHECK(hipGetDevice(&dev));
        start_th = std::thread([&, dev] {
            auto run = [&]() -> int {
                SEHIP_CHECK(hipSetDevice(dev));
                h.resize((size_t)n);
                if (ho && ho->start == SPEIGS_HIP_START_VECTOR && ho->start_vec)
                    std::memcpy(h.data(), ho->start_vec, (size_t)n * sizeof(double));
                else
                    uniform_vec(&seed, n, h.data());
                const int r = copy_to_device(v0, h.data(), (size_t)n * sizeof(double), F->st2);
                if (r) return r;
                SEHIP_CHECK(hipEventRecord(F->ev_ready, F->st2));
                return 0;
            };
            start_rc = run();
            if (start_rc) start_msg = last_error_text();
        });
    }
    rc = factor_invert_l11(F);
    if (start_th.joinable()) start_th.join();
    if (rc) return rc;
    if (start_rc) {
        note_error(start_msg.c_str());
        return start_rc;
    }

    // start block: the proposed block of the learned start (its first m
    // columns), else column 0 = the reference's xorshift64* start vector (or
    // the caller's) and the other columns from a counter-based generator
    {
        if (Xstart) {
            rfrom_f32rows(Xstart, 32, std::min(m, 32), n, T2, ldT, st);
        } else {
            SEHIP_CHECK(hipStreamWaitEvent(st, F->ev_ready, 0));
            rfrom_cols(v0, 1, n, T2, ldT, st);
            rrandom(T2, ldT, 1, m, n, seed, st);
        }
        for (int pass = 0; pass < 2; pass++) {
            std::vector<double> g;
            if ((rc = dgram(W, T2, ldT, m, T2, ldT, m, g))) return rc;
            int r = 0;
            if ((rc = svqb(W, T2, ldT, m, g, S, ldS, &r))) return rc;
            if (r < m) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "LOBPCG: start block lost rank in orthonormalisation");
            if (pass == 0) rcopy(S, ldS, m, n, T2, ldT, st);
        }
    }
    rspmm(A, S, ldS, AS, ldS, m, st);
    std::vector<double> theta(m);
    {
        std::vector<double> g;
        if ((rc = dgram(W, S, ldS, m, AS, ldS, m, g))) return rc;
        symmetrize(g, m);
        if (dense_syev(m, g.data(), theta.data())) return SPEIGS_ERR_LAPACK;
        if ((rc = dgemm_tall(W, S, ldS, m, g, m, Sn, ldS, 0.0))) return rc;
        if ((rc = dgemm_tall(W, AS, ldS, m, g, m, ASn, ldS, 0.0))) return rc;
        std::swap(S, Sn);
        std::swap(AS, ASn);
    }
    int np = 0;
    int64_t it = 0, nconv = 0, solves = 0;
    std::vector<double> res(m);
    double *lam_dev = W.red;
    // the reference's iteration breakdown (include/speigs.h:57-68): t_op the
    // preconditioner solves and A-products, t_ortho the projection and
    // orthonormalisation of [W P], t_resid the residuals, t_ritz Rayleigh-Ritz
    // and the new block; each phase ends at a stream synchronisation
    double tmark = wall();
    auto lap = [&](double &acc) -> int {
        SEHIP_CHECK(hipStreamSynchronize(st));
        const double now = wall();
        acc += now - tmark;
        tmark = now;
        return 0;
    };
    // the residual block W and its norms come from the previous Ritz step's
    // fused pass, except at the start and after a refresh of A X
    bool have_res = false;
    for (;; it++) {
        // S = [X | P | W] (no P in the first iteration): Z = [P W] is contiguous
        double *Z = S + m, *AZ = AS + m, *Wb = S + m + np;
        {
            std::vector<double> g;
            if (have_res) {
                g.resize(k);
                if ((rc = to_host(W, g.data(), W.nrm, (size_t)k))) return rc;
            } else {
                if ((rc = to_dev(W, lam_dev, theta.data(), (size_t)m))) return rc;
                if ((rc = resid_norms(W, AS, S, ldS, lam_dev, m, (int)k, Wb, g))) return rc;
            }
            nconv = 0;
            for (int i = 0; i < k; i++) {
                res[i] = std::sqrt(std::max(g[i], 0.0));
                if (res[i] <= rtol) nconv++;
            }
        }
        if ((rc = lap(info->t_resid))) return rc;
        if (nconv >= k || it >= maxit) break;
        // W = T R: one FP32 multi-right-hand-side solve, in place on the rows of S
        if ((rc = factor_block_solve_rows(F, Wb, ldS, 1, Wb, ldS, 1, m))) return rc;
        solves++;
        if ((rc = lap(info->t_op))) return rc;
        int nz = m + np;
        double keep = 0.0, cond = 0.0;
        int r = 0;
        if ((rc = project_svqb(W, S, ldS, m, nz, &r, &keep, &cond))) return rc;
        if (r == 0) break;
        // the second pass only where the first can have left Z short of
        // orthonormal: a column that lost more than half its norm to the
        // projection (Kahan's "twice is enough" test), a dropped direction, or
        // an ill-conditioned Gram matrix
        const bool again = keep < 0.5 || r < nz || cond > 1e4;
        nz = r;
        if (again) {
            double keep2 = 0.0, cond2 = 0.0;
            if ((rc = project_svqb(W, S, ldS, m, nz, &r, &keep2, &cond2))) return rc;
            if (r == 0) break;
            nz = r;
        }
        if ((rc = lap(info->t_ortho))) return rc;
        rspmm(A, Z, ldS, AZ, ldS, nz, st);
        if ((rc = lap(info->t_op))) return rc;
        // Rayleigh-Ritz on S = [X Z] (orthonormal): H = S^T A S.  X holds the
        // previous Ritz vectors, so X^T A X = diag(theta) (as in Knyazev's and
        // SciPy's LOBPCG) and only S^T A Z is formed: (m + nz) x nz
        const int d = m + nz;
        std::vector<double> G2, H((size_t)d * d, 0.0);
        double tr = wall();
        if ((rc = dgram(W, S, ldS, d, AZ, ldS, nz, G2))) return rc;
        trace_phase("  Rayleigh-Ritz: Gram (device, copy back)", tr);
        for (int t = 0; t < m; t++) H[(size_t)t * d + t] = theta[t];
        for (int j = 0; j < nz; j++)
            for (int t = 0; t < d; t++) {
                H[(size_t)(m + j) * d + t] = G2[(size_t)j * d + t];
                if (t < m) H[(size_t)t * d + m + j] = G2[(size_t)j * d + t];
            }
        symmetrize(H, d);
        // only the m smallest Ritz pairs are kept: all values by QL, those m
        // vectors by inverse iteration on the tridiagonal (O(d^2 m), not the
        // O(d^3) of accumulating every rotation)
        DenseTri tri;
        if (dense_tri_values_rr(d, H.data(), tri)) return SPEIGS_ERR_LAPACK;
        std::vector<int64_t> sel(m);
        for (int t = 0; t < m; t++) sel[t] = t;
        // C2 = [C | C_P]: the Ritz coefficients and those of the new P (the Z
        // rows of C), so [X_new P_new] = S C2 is one pass over S
        std::vector<double> C2((size_t)d * 2 * m, 0.0);
        if (dense_tri_vectors(tri, m, sel.data(), C2.data(), d)) return SPEIGS_ERR_LAPACK;
        trace_phase("  Rayleigh-Ritz: projected eigenproblem (host)", tr);
        for (int t = 0; t < m; t++) {
            theta[t] = tri.w[t];
            for (int i = m; i < d; i++) C2[(size_t)(m + t) * d + i] = C2[(size_t)t * d + i];
        }
        // new X, new P, new A X and the next residual block R = A X - X theta
        // (into the next W slot, columns 2m .. 3m-1) in one pass over S and AS
        // theta rides after C2 in one upload
        C2.resize((size_t)d * 2 * m + m);
        std::copy(theta.begin(), theta.end(), C2.begin() + (size_t)d * 2 * m);
        if ((rc = to_dev(W, W.C, C2.data(), C2.size()))) return rc;
        rritz(S, AS, ldS, d, W.C, m, W.C + (size_t)d * 2 * m, (int)k, Sn, ASn, n, W.rpart, W.nrm, st);
        trace_phase("  Rayleigh-Ritz: step queued", tr);
        std::swap(S, Sn);
        std::swap(AS, ASn);
        np = m;
        have_res = true;
        if ((rc = lap(info->t_ritz))) return rc;
        if ((it + 1) % 8 == 0) {
            rspmm(A, S, ldS, AS, ldS, m, st);   // refresh A X against drift
            have_res = false;                   // and the residuals with it
            if ((rc = lap(info->t_op))) return rc;
        }
    }
    // final pairs: exact A X, Rayleigh quotients, residuals
    rspmm(A, S, ldS, AS, ldS, m, st);
    std::vector<double> xax, xx, rr;
    if ((rc = dgram(W, S, ldS, m, AS, ldS, m, xax))) return rc;
    if ((rc = dgram(W, S, ldS, m, S, ldS, m, xx))) return rc;
    std::vector<double> lam(m);
    for (int t = 0; t < m; t++) lam[t] = xax[(size_t)t * m + t] / xx[(size_t)t * m + t];
    if ((rc = to_dev(W, lam_dev, lam.data(), (size_t)m))) return rc;
    if ((rc = resid_norms(W, AS, S, ldS, lam_dev, m, (int)k, S + m, rr))) return rc;
    if ((rc = lap(info->t_resid))) return rc;
    nconv = 0;
    for (int t = 0; t < k; t++) {
        const double r = std::sqrt(std::max(rr[t], 0.0)) / std::sqrt(xx[(size_t)t * m + t]);
        lambda[t] = lam[t];
        if (resid) resid[t] = r;
        if (r <= rtol) nconv++;
    }
    if (Vout) {
        rto_cols(S, ldS, (int)k, n, T2, st);   // T2 holds n x 2m >= n x k doubles
        touch.join();
        if ((rc = copy_to_host(Vout, T2, (size_t)n * k * sizeof(double), st))) return rc;
    }
    SEHIP_CHECK(hipStreamSynchronize(st));
    info->nconv = nconv;
    info->restarts = it;
    info->nops = solves * m;
    info->flag = nconv >= k ? 0 : 1;
    if (hinfo) hinfo->solves = solves;
    return SPEIGS_OK;
}

}  // namespace sehip
