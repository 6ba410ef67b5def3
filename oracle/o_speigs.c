/*
 * o_speigs.c -- CPU oracle: restatement of the reference speigs() driver and
 * its thick-restart Lanczos (TEST INFRASTRUCTURE ONLY; never linked by the
 * product library).
 *
 * Follows moatas19m/sparse-eigs-mex src/speigs_core.c:
 *   rng / rand_vec          :49-54
 *   operator (A x / solve)  :57-116   (CHOLMOD -> o_chol.c, UMFPACK -> dense LU,
 *                                      dpbtrs -> oracle_pbtrs_u)
 *   orth_block (CGS2)       :122-132
 *   lanczos                 :146-315  (including the breakdown quirk at :221 where
 *                                      H keeps its old leading dimension)
 *   dense_path              :318-353
 *   check_symmetry          :363-401
 *   diag_path               :404-428
 *   components / union-find :436-466
 *   rcm_order, bandwidth_of :471-523
 *   blocks_path             :532-632
 *   speigs_z                :660-768
 *   speigs driver           :783-1024 (banded gate :869-931, shift escalation
 *                                      :953-970, LU fallback :974-984)
 *
 * Pinning (tests/test_oracle.py): the 13 reference tests of test/test_speigs.c
 * restated against closed-form Laplacian spectra, xorshift64* known answers
 * (Vigna 2016 generator, checked by an independent numpy uint64 restatement),
 * the dense eigensolver against numpy.linalg.eigh, the sparse Cholesky against
 * numpy dense factorizations, and subspace angles of multiplets against the
 * separable sine eigenbases.
 * The ordering differs from CHOLMOD's AMD/METIS (absent here); an ordering
 * changes rounding and operation counts, not the eigenpairs.
 */
#include "oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define IMAX(a, b) ((a) > (b) ? (a) : (b))
#define IMIN(a, b) ((a) < (b) ? (a) : (b))

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* ------------------------------------------------------------- rng ----- */
uint64_t oracle_rng_next(uint64_t *s)
{
    uint64_t x = *s;
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}

void oracle_rand_vec(uint64_t *s, int64_t n, double *v)
{
    for (int64_t i = 0; i < n; i++)
        v[i] = (double)(oracle_rng_next(s) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

/* ---------------------------------------------------------- vectors ---- */
/* Pairwise (cascade) summation: error O(u log n) instead of O(u n), which is
 * what an optimised BLAS ddot/dnrm2 delivers with its blocked accumulators.
 * With naive summation the Rayleigh quotient error alone (~sqrt(n) u |lambda|)
 * sits at the tol*||A|| = 1e-14*||A|| convergence floor. */
static double vdot(int64_t n, const double *x, const double *y)
{
    if (n <= 32) {
        double s = 0.0;
        for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
        return s;
    }
    int64_t h = n / 2;
    return vdot(h, x, y) + vdot(n - h, x + h, y + h);
}
static double vnrm2(int64_t n, const double *x)
{
    return sqrt(vdot(n, x, x));
}
static void vscal(int64_t n, double a, double *x)
{
    for (int64_t i = 0; i < n; i++) x[i] *= a;
}
/* y = B s  (B n x m column-major) */
static void gemv_n(int64_t n, int64_t m, const double *B, const double *s, double *y)
{
    for (int64_t i = 0; i < n; i++) y[i] = 0.0;
    for (int64_t c = 0; c < m; c++) {
        double a = s[c];
        const double *b = B + (size_t)c * n;
        for (int64_t i = 0; i < n; i++) y[i] += a * b[i];
    }
}

/* CGS2: two passes of t = B^T w, w -= B t; accumulate t into hout */
static void cgs2(int64_t n, int64_t m, const double *B, double *w, double *hout, double *t)
{
    if (m <= 0) return;
    for (int pass = 0; pass < 2; pass++) {
        for (int64_t c = 0; c < m; c++) t[c] = vdot(n, B + (size_t)c * n, w);
        for (int64_t c = 0; c < m; c++) {
            double a = t[c];
            const double *b = B + (size_t)c * n;
            for (int64_t i = 0; i < n; i++) w[i] -= a * b[i];
        }
        if (hout) for (int64_t c = 0; c < m; c++) hout[c] += t[c];
    }
}

/* ---------------------------------------------------------- operator ---- */
enum { OP_MULT = 0, OP_CHOL = 1, OP_LU = 2, OP_BAND = 3 };
typedef struct {
    int mode;
    int64_t n;
    const int64_t *Ap, *Ai; const double *Ax;
    oracle_chol *chol;
    double *lu; int64_t *piv;
    double *ab; int64_t kd; int64_t *perm; double *bw;
    int64_t nops;
} oop_t;

static void op_mult(const oop_t *op, const double *x, double *y)
{
    oracle_spmv_upper(op->n, op->Ap, op->Ai, op->Ax, x, y);
}

static int op_apply(oop_t *op, const double *x, double *y)
{
    op->nops++;
    switch (op->mode) {
    case OP_MULT: op_mult(op, x, y); return 0;
    case OP_CHOL: oracle_chol_solve(op->chol, x, y); return 0;
    case OP_BAND:
        for (int64_t i = 0; i < op->n; i++) op->bw[i] = x[op->perm[i]];
        oracle_pbtrs_u(op->n, op->kd, op->ab, op->kd + 1, op->bw);
        for (int64_t i = 0; i < op->n; i++) y[op->perm[i]] = op->bw[i];
        return 0;
    default:
        memcpy(y, x, (size_t)op->n * sizeof(double));
        oracle_lu_solve(op->n, op->lu, op->piv, y);
        return 0;
    }
}

typedef struct { double key; int64_t idx; } okv;
static int okv_desc(const void *a, const void *b)
{
    double x = ((const okv *)a)->key, y = ((const okv *)b)->key;
    return (x < y) - (x > y);
}
static int okv_asc(const void *a, const void *b)
{
    double x = ((const okv *)a)->key, y = ((const okv *)b)->key;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------------ lanczos ---- */
static int lanczos(oop_t *op, int64_t n, int64_t k, const speigs_opts *o, double anorm,
                   double *lambda, double *Vout, double *resid, speigs_info *info)
{
    int64_t ncv = o->ncv ? o->ncv : IMAX(2 * k + 1, 20);
    if (ncv > n) ncv = n;
    if (ncv <= k) ncv = IMIN(n, k + 1);
    if (ncv < 2) ncv = IMIN(n, 2);
    const int64_t ldh = ncv;                 /* H keeps this leading dimension */
    int64_t maxit = o->maxit ? o->maxit : 300;
    double tol = o->tol > 0 ? o->tol : 1e-14;
    double rtol = tol * (anorm > 1.0 ? anorm : 1.0);
    double small = 1e-13 * (anorm > 1.0 ? anorm : 1.0);

    double *V = (double *)calloc((size_t)n * (ncv + 1), sizeof(double));
    double *Vn = (double *)calloc((size_t)n * (ncv + 1), sizeof(double));
    double *H = (double *)calloc((size_t)ncv * ncv, sizeof(double));
    double *Hs = (double *)calloc((size_t)ncv * ncv, sizeof(double));
    double *theta = (double *)calloc((size_t)ncv, sizeof(double));
    double *keep = (double *)calloc((size_t)ncv, sizeof(double));
    double *t = (double *)calloc((size_t)IMAX(ncv, 4), sizeof(double));
    double *w = (double *)calloc((size_t)n, sizeof(double));
    double *Ax = (double *)calloc((size_t)n, sizeof(double));
    double *x = (double *)calloc((size_t)n, sizeof(double));
    okv *ord = (okv *)calloc((size_t)ncv, sizeof(okv));
    int rc = SPEIGS_OK;
    if (!V || !Vn || !H || !Hs || !theta || !keep || !t || !w || !Ax || !x || !ord) {
        rc = SPEIGS_ERR_MEM; goto out;
    }
    uint64_t seed = o->seed ? o->seed : 88172645463325252ULL;
    oracle_rand_vec(&seed, n, w);
    {
        double nr = vnrm2(n, w);
        if (nr == 0.0) { rc = SPEIGS_ERR_INTERNAL; goto out; }
        vscal(n, 1.0 / nr, w);
        memcpy(V, w, (size_t)n * sizeof(double));
    }
    int64_t j = 0, restarts = 0, nconv = 0;
    for (;;) {
        /* extend the Krylov basis to ncv vectors */
        for (; j < ncv; j++) {
            double ts = now_s();
            rc = op_apply(op, V + (size_t)j * n, w);
            info->t_op += now_s() - ts;
            if (rc) goto out;
            ts = now_s();
            double *hc = H + (size_t)j * ldh;
            for (int64_t i = 0; i <= j; i++) hc[i] = 0.0;
            cgs2(n, j + 1, V, w, hc, t);
            info->t_ortho += now_s() - ts;
            double beta = vnrm2(n, w), beta_rec = beta;
            if (beta <= small) {
                /* invariant subspace: re-inject random directions */
                beta_rec = 0.0;
                for (int att = 0; att < 3 && beta <= small; att++) {
                    oracle_rand_vec(&seed, n, w);
                    cgs2(n, j + 1, V, w, NULL, t);
                    beta = vnrm2(n, w);
                }
                if (beta <= small) { ncv = j + 1; break; }
            }
            vscal(n, 1.0 / beta, w);
            memcpy(V + (size_t)(j + 1) * n, w, (size_t)n * sizeof(double));
            if (j + 1 < ncv) {
                H[(j + 1) + (size_t)j * ldh] = beta_rec;
                H[j + (size_t)(j + 1) * ldh] = beta_rec;
            }
        }
        /* projected problem; NOTE the reference indexes H with the current ncv
         * as leading dimension (src/speigs_core.c:233-234), which differs from
         * ldh only after a breakdown shrank ncv -- reproduced on purpose */
        int64_t m = ncv;
        for (int64_t c = 0; c < m; c++)
            for (int64_t r = 0; r < m; r++) Hs[r + c * m] = H[r + c * ncv];
        if (oracle_syev_jacobi(m, Hs, theta)) { rc = SPEIGS_ERR_LAPACK; goto out; }
        for (int64_t i = 0; i < m; i++) { ord[i].key = fabs(theta[i]); ord[i].idx = i; }
        qsort(ord, (size_t)m, sizeof(okv), okv_desc);

        /* true residuals of the wanted Ritz pairs against A */
        nconv = 0;
        int64_t nwant = IMIN(k, m);
        double tr = now_s();
        for (int64_t q = 0; q < nwant; q++) {
            int64_t i = ord[q].idx;
            gemv_n(n, m, V, Hs + (size_t)i * m, x);
            double xn = vnrm2(n, x);
            if (xn > 0) vscal(n, 1.0 / xn, x);
            op_mult(op, x, Ax);
            double lam = vdot(n, x, Ax);
            for (int64_t p = 0; p < n; p++) Ax[p] -= lam * x[p];
            double r = vnrm2(n, Ax);
            if (Vout) memcpy(Vout + (size_t)q * n, x, (size_t)n * sizeof(double));
            lambda[q] = lam;
            if (resid) resid[q] = r;
            if (r <= rtol) nconv++;
        }
        info->t_resid += now_s() - tr;
        if (nconv >= IMIN(k, m) || restarts >= maxit || m < ncv) break;
        restarts++;

        /* thick restart with the l dominant Ritz vectors */
        int64_t l = IMIN(k + 5, ncv - 1);
        if (l < 1) l = 1;
        int64_t nret = 0;
        double tk = now_s();
        for (int64_t q = 0; q < m && nret < l; q++) {
            int64_t i = ord[q].idx;
            double *dst = Vn + (size_t)nret * n;
            gemv_n(n, m, V, Hs + (size_t)i * m, dst);
            cgs2(n, nret, Vn, dst, NULL, t);
            double nr = vnrm2(n, dst);
            if (nr < 1e-10) continue;
            vscal(n, 1.0 / nr, dst);
            keep[nret++] = theta[i];
        }
        info->t_ritz += now_s() - tk;
        memcpy(w, V + (size_t)m * n, (size_t)n * sizeof(double));
        cgs2(n, nret, Vn, w, NULL, t);
        {
            double nr = vnrm2(n, w);
            if (nr < 1e-10) {
                oracle_rand_vec(&seed, n, w);
                cgs2(n, nret, Vn, w, NULL, t);
                nr = vnrm2(n, w);
            }
            if (nr < 1e-10) break;
            vscal(n, 1.0 / nr, w);
        }
        memcpy(Vn + (size_t)nret * n, w, (size_t)n * sizeof(double));
        memcpy(V, Vn, (size_t)n * (nret + 1) * sizeof(double));
        memset(H, 0, (size_t)ldh * ldh * sizeof(double));
        for (int64_t c = 0; c < nret; c++) H[c + c * ldh] = keep[c];
        j = nret;
    }
    info->nconv = nconv;
    info->restarts = restarts;
    info->nops = op->nops;
    info->flag = (nconv >= IMIN(k, ncv)) ? 0 : 1;
out:
    free(V); free(Vn); free(H); free(Hs); free(theta); free(keep);
    free(t); free(w); free(Ax); free(x); free(ord);
    return rc;
}

/* ------------------------------------------------------ short-circuits ---- */
static int dense_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                      int64_t k, speigs_which which, double *lambda, double *V,
                      double *resid, speigs_info *info)
{
    double *D = (double *)calloc((size_t)n * n, sizeof(double));
    double *ev = (double *)calloc((size_t)n, sizeof(double));
    okv *ord = (okv *)calloc((size_t)n, sizeof(okv));
    if (!D || !ev || !ord) { free(D); free(ev); free(ord); return SPEIGS_ERR_MEM; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
            D[Ai[p] + (size_t)c * n] = Ax[p];
            D[c + (size_t)Ai[p] * n] = Ax[p];
        }
    if (oracle_syev(n, D, ev)) { free(D); free(ev); free(ord); return SPEIGS_ERR_LAPACK; }
    for (int64_t i = 0; i < n; i++) { ord[i].key = fabs(ev[i]); ord[i].idx = i; }
    qsort(ord, (size_t)n, sizeof(okv), which == SPEIGS_SM ? okv_asc : okv_desc);
    for (int64_t q = 0; q < k; q++) {
        int64_t i = ord[q].idx;
        lambda[q] = ev[i];
        if (V) memcpy(V + (size_t)q * n, D + (size_t)i * n, (size_t)n * sizeof(double));
        if (resid) resid[q] = 0.0;
    }
    info->nconv = k; info->flag = 0; info->path = SPEIGS_PATH_DENSE;
    free(D); free(ev); free(ord);
    return SPEIGS_OK;
}

/* 1 full symmetric, 2 upper only, 3 lower only, 0 non-symmetric */
static int symmetry_kind(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax)
{
    int up = 0, lo = 0;
    for (int64_t c = 0; c < n && !(up && lo); c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            if (Ai[q] < c) up = 1; else if (Ai[q] > c) lo = 1;
        }
    if (up && !lo) return 2;
    if (lo && !up) return 3;
    if (!up && !lo) return 1;
    int64_t nnz = Ap[n];
    int64_t *Tp = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t *Ti = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    double *Tx = (double *)malloc((size_t)(nnz ? nnz : 1) * sizeof(double));
    int64_t *nx = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    if (!Tp || !Ti || !Tx || !nx) { free(Tp); free(Ti); free(Tx); free(nx); return 1; }
    for (int64_t q = 0; q < nnz; q++) nx[Ai[q]]++;
    for (int64_t i = 0; i < n; i++) { Tp[i + 1] = Tp[i] + nx[i]; nx[i] = Tp[i]; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            int64_t d = nx[Ai[q]]++;
            Ti[d] = c; Tx[d] = Ax[q];
        }
    int ok = 1;
    for (int64_t i = 0; i <= n && ok; i++) if (Tp[i] != Ap[i]) ok = 0;
    for (int64_t q = 0; q < nnz && ok; q++) {
        if (Ti[q] != Ai[q]) { ok = 0; break; }
        double a = Ax[q], b = Tx[q], mx = fabs(a) > fabs(b) ? fabs(a) : fabs(b);
        if (fabs(a - b) > 1e-12 * (mx > 0 ? mx : 1.0)) ok = 0;
    }
    free(Tp); free(Ti); free(Tx); free(nx);
    return ok ? 1 : 0;
}

static int diag_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                     int64_t k, speigs_which which, double *lambda, double *V,
                     double *resid, speigs_info *info)
{
    double *d = (double *)calloc((size_t)n, sizeof(double));
    okv *ord = (okv *)calloc((size_t)n, sizeof(okv));
    if (!d || !ord) { free(d); free(ord); return SPEIGS_ERR_MEM; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) if (Ai[p] == c) d[c] = Ax[p];
    for (int64_t i = 0; i < n; i++) { ord[i].key = fabs(d[i]); ord[i].idx = i; }
    qsort(ord, (size_t)n, sizeof(okv), which == SPEIGS_SM ? okv_asc : okv_desc);
    for (int64_t q = 0; q < k; q++) {
        int64_t i = ord[q].idx;
        lambda[q] = d[i];
        if (V) { memset(V + (size_t)q * n, 0, (size_t)n * sizeof(double)); V[(size_t)q * n + i] = 1.0; }
        if (resid) resid[q] = 0.0;
    }
    info->nconv = k; info->flag = 0; info->path = SPEIGS_PATH_DIAG;
    free(d); free(ord);
    return SPEIGS_OK;
}

static int64_t uf_root(int64_t *p, int64_t x)
{
    while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
    return x;
}

static int64_t components(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                          int64_t *comp)
{
    int64_t *p = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *lab = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!p || !lab) { free(p); free(lab); return -1; }
    for (int64_t i = 0; i < n; i++) { p[i] = i; lab[i] = -1; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++)
            if (Ax[q] != 0.0) {
                int64_t a = uf_root(p, Ai[q]), b = uf_root(p, c);
                if (a != b) p[a] = b;
            }
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t r = uf_root(p, i);
        if (lab[r] < 0) lab[r] = nc++;
        comp[i] = lab[r];
    }
    free(p); free(lab);
    return nc;
}

/* reverse Cuthill-McKee over the stored column lists (as the reference does) */
static void rcm(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *perm)
{
    int64_t *deg = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    char *seen = (char *)calloc((size_t)n, 1);
    int64_t *q = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!deg || !seen || !q) {
        for (int64_t i = 0; i < n; i++) perm[i] = i;
        free(deg); free(seen); free(q);
        return;
    }
    for (int64_t c = 0; c < n; c++) deg[c] = Ap[c + 1] - Ap[c];
    int64_t head = 0, tail = 0;
    for (int64_t st = 0; st < n; st++) {
        if (seen[st]) continue;
        int64_t s = st, best = deg[st];
        for (int64_t i = st; i < n; i++)
            if (!seen[i] && deg[i] < best) { best = deg[i]; s = i; }
        seen[s] = 1; q[tail++] = s;
        while (head < tail) {
            int64_t v = q[head++], b0 = tail;
            for (int64_t p = Ap[v]; p < Ap[v + 1]; p++) {
                int64_t u = Ai[p];
                if (!seen[u]) { seen[u] = 1; q[tail++] = u; }
            }
            for (int64_t a = b0 + 1; a < tail; a++) {   /* stable insertion by degree */
                int64_t key = q[a], dk = deg[key], b = a - 1;
                while (b >= b0 && deg[q[b]] > dk) { q[b + 1] = q[b]; b--; }
                q[b + 1] = key;
            }
        }
    }
    /* a start the BFS never reached is skipped by the loop above (as in the
     * reference, which then reverses a partly filled queue: undefined); the
     * unreached nodes are appended in index order to keep a permutation */
    for (int64_t c = 0; c < n && tail < n; c++) if (!seen[c]) q[tail++] = c;
    for (int64_t i = 0; i < n; i++) perm[i] = q[n - 1 - i];
    free(deg); free(seen); free(q);
}

void oracle_rcm(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *perm) { rcm(n, Ap, Ai, perm); }

static int64_t bandwidth(int64_t n, const int64_t *Ap, const int64_t *Ai,
                         const int64_t *perm, int64_t *iperm)
{
    for (int64_t i = 0; i < n; i++) iperm[perm[i]] = i;
    int64_t kd = 0;
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            int64_t d = iperm[Ai[q]] - iperm[c];
            if (d < 0) d = -d;
            if (d > kd) kd = d;
        }
    return kd;
}

static int blocks_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                       int64_t k, speigs_which which, const speigs_opts *opts,
                       const int64_t *comp, int64_t ncomp, double *lambda, double *V,
                       double *resid, speigs_info *info)
{
    int rc = SPEIGS_OK;
    int64_t *sz = (int64_t *)calloc((size_t)ncomp, sizeof(int64_t));
    int64_t *off = (int64_t *)calloc((size_t)ncomp + 1, sizeof(int64_t));
    int64_t *gidx = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *loc = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *fill = (int64_t *)calloc((size_t)ncomp, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) sz[comp[i]]++;
    for (int64_t c = 0; c < ncomp; c++) off[c + 1] = off[c] + sz[c];
    for (int64_t i = 0; i < n; i++) {
        int64_t c = comp[i], slot = off[c] + fill[c]++;
        gidx[slot] = i; loc[i] = slot - off[c];
    }
    int64_t maxc = 0;
    for (int64_t c = 0; c < ncomp; c++) maxc += IMIN(k, sz[c]);
    double **cv = (double **)calloc((size_t)ncomp, sizeof(double *));
    double *clam = (double *)calloc((size_t)maxc, sizeof(double));
    double *cres = (double *)calloc((size_t)maxc, sizeof(double));
    int64_t *ccmp = (int64_t *)calloc((size_t)maxc, sizeof(int64_t));
    int64_t *cslt = (int64_t *)calloc((size_t)maxc, sizeof(int64_t));
    okv *ord = (okv *)calloc((size_t)maxc, sizeof(okv));
    int64_t nz = Ap[n];
    int64_t *sAp = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t *sAi = (int64_t *)malloc((size_t)(nz ? nz : 1) * sizeof(int64_t));
    double *sAx = (double *)malloc((size_t)(nz ? nz : 1) * sizeof(double));
    int64_t ncand = 0;
    for (int64_t c = 0; c < ncomp; c++) {
        int64_t nc = sz[c], cnt = 0;
        for (int64_t jj = 0; jj < nc; jj++) {
            int64_t g = gidx[off[c] + jj];
            sAp[jj] = cnt;
            for (int64_t q = Ap[g]; q < Ap[g + 1]; q++)
                if (comp[Ai[q]] == c) { sAi[cnt] = loc[Ai[q]]; sAx[cnt] = Ax[q]; cnt++; }
        }
        sAp[nc] = cnt;
        int64_t kc = IMIN(k, nc);
        double *lv = (double *)calloc((size_t)kc, sizeof(double));
        double *rv = (double *)calloc((size_t)kc, sizeof(double));
        double *vv = V ? (double *)calloc((size_t)nc * kc, sizeof(double)) : NULL;
        speigs_info si;
        int r2 = oracle_speigs(nc, sAp, sAi, sAx, kc, which, opts, lv, vv, rv, &si);
        if (r2) { free(lv); free(rv); free(vv); rc = r2; goto out; }
        info->nops += si.nops;
        if (si.flag) info->flag = 1;
        cv[c] = vv;
        for (int64_t q = 0; q < kc; q++) {
            clam[ncand] = lv[q]; cres[ncand] = rv[q]; ccmp[ncand] = c; cslt[ncand] = q; ncand++;
        }
        free(lv); free(rv);
    }
    for (int64_t i = 0; i < ncand; i++) { ord[i].key = fabs(clam[i]); ord[i].idx = i; }
    qsort(ord, (size_t)ncand, sizeof(okv), which == SPEIGS_SM ? okv_asc : okv_desc);
    for (int64_t q = 0; q < k && q < ncand; q++) {
        int64_t i = ord[q].idx, c = ccmp[i], sl = cslt[i], nc = sz[c];
        lambda[q] = clam[i];
        if (resid) resid[q] = cres[i];
        if (V) {
            memset(V + (size_t)q * n, 0, (size_t)n * sizeof(double));
            const double *src = cv[c] + (size_t)sl * nc;
            for (int64_t jj = 0; jj < nc; jj++) V[(size_t)q * n + gidx[off[c] + jj]] = src[jj];
        }
    }
    info->nconv = IMIN(k, ncand);
    info->ncomp = ncomp;
    info->path = SPEIGS_PATH_BLOCKS;
out:
    if (cv) { for (int64_t c = 0; c < ncomp; c++) free(cv[c]); free(cv); }
    free(sz); free(off); free(gidx); free(loc); free(fill);
    free(clam); free(cres); free(ccmp); free(cslt); free(ord);
    free(sAp); free(sAi); free(sAx);
    return rc;
}

void oracle_default_opts(speigs_opts *o)
{
    memset(o, 0, sizeof *o);
    o->dense_max = -1;
    o->detect = 1;
    o->check_sym = 1;
}

/* ||A||_1 of the symmetric matrix whose upper triangle is stored */
static double norm1_upper(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax)
{
    double *cs = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i > j) continue;
            cs[j] += fabs(Ax[q]);
            if (i != j) cs[i] += fabs(Ax[q]);
        }
    double m = 0.0;
    for (int64_t j = 0; j < n; j++) if (cs[j] > m) m = cs[j];
    free(cs);
    return m;
}

/* try the banded route; returns 1 when it produced an operator */
static int try_banded(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
                      const speigs_opts *opts, double anorm, oop_t *op, speigs_info *info)
{
    int bmax = opts->band_max ? opts->band_max : 512;
    if (bmax <= 0) return 0;
    double tb = now_s();
    int64_t *pm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *ip = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!pm || !ip) { free(pm); free(ip); return 0; }
    rcm(n, Ap, Ai, pm);
    int64_t kd = bandwidth(n, Ap, Ai, pm, ip);
    double cells = (double)(kd + 1) * (double)n;
    double occ = kd > 0 ? (double)Ap[n] / ((double)n * (double)kd) : 1.0;
    if (!(kd <= bmax && cells <= 2.0e7 && occ >= 0.05)) { free(pm); free(ip); return 0; }
    int64_t ld = kd + 1;
    double *AB = (double *)calloc((size_t)cells, sizeof(double));
    double *W = (double *)malloc((size_t)cells * sizeof(double));
    if (!AB || !W) { free(AB); free(W); free(pm); free(ip); return 0; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            int64_t r = ip[Ai[q]], cc = ip[c];
            int64_t lo = r < cc ? r : cc, hi = r < cc ? cc : r;
            AB[(kd + lo - hi) + (size_t)hi * ld] = Ax[q];
        }
    memcpy(W, AB, (size_t)cells * sizeof(double));
    int bad = oracle_pbtrf_u(n, kd, W, ld);
    double dl = 0.0;
    if (bad) {
        /* same capped escalation; the loop's step runs once more after the
         * successful factorization, so the reported shift is 10x the one used
         * (src/speigs_core.c:908-912) -- reproduced */
        double d0 = anorm * (opts->shift0 > 0 ? opts->shift0 : 1e-6);
        double dcap = anorm * 1e-3;
        if (d0 > dcap) dcap = d0;
        for (dl = d0; dl <= dcap * 1.0000001 && bad; dl *= 10.0) {
            memcpy(W, AB, (size_t)cells * sizeof(double));
            for (int64_t jj = 0; jj < n; jj++) W[kd + (size_t)jj * ld] += dl;
            bad = oracle_pbtrf_u(n, kd, W, ld);
        }
        if (bad) dl = 0.0;
    }
    free(AB);
    if (bad) { free(W); free(pm); free(ip); return 0; }
    op->mode = OP_BAND; op->ab = W; op->kd = kd; op->perm = pm;
    op->bw = (double *)calloc((size_t)n, sizeof(double));
    info->shift = dl; info->path = SPEIGS_PATH_BANDED;
    info->t_factor = now_s() - tb;
    free(ip);
    return 1;
}

int oracle_speigs(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                  const double *Ax, speigs_int k, speigs_which which,
                  const speigs_opts *opts, double *lambda, double *V,
                  double *resid, speigs_info *info)
{
    speigs_opts dflt;
    if (!opts) { oracle_default_opts(&dflt); opts = &dflt; }
    speigs_info local;
    if (!info) info = &local;
    memset(info, 0, sizeof *info);
    if (n <= 0 || !Ap || !Ai || !Ax || !lambda) return SPEIGS_ERR_ARG;
    if (k <= 0 || k > n) return SPEIGS_ERR_ARG;
    if (opts->check_sym != 0 && symmetry_kind(n, Ap, Ai, Ax) == 0) return SPEIGS_ERR_NOTSYM;

    double t0 = now_s();
    int dmax = opts->dense_max < 0 ? 800 : opts->dense_max;
    if (opts->detect) {
        int diag = 1;
        for (int64_t c = 0; c < n && diag; c++)
            for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) if (Ai[p] != c) { diag = 0; break; }
        if (diag) return diag_path(n, Ap, Ai, Ax, k, which, lambda, V, resid, info);
    }
    if (dmax > 0 && (n <= dmax || 4 * k > n)) {
        int rc = dense_path(n, Ap, Ai, Ax, k, which, lambda, V, resid, info);
        info->t_iter = now_s() - t0;
        return rc;
    }
    if (opts->detect) {
        int64_t *comp = (int64_t *)malloc((size_t)n * sizeof(int64_t));
        if (comp) {
            int64_t nc = components(n, Ap, Ai, Ax, comp);
            if (nc > 1) {
                int rc = blocks_path(n, Ap, Ai, Ax, k, which, opts, comp, nc, lambda, V, resid, info);
                free(comp);
                info->t_iter = now_s() - t0;
                return rc;
            }
            free(comp);
        }
    }
    double anorm = norm1_upper(n, Ap, Ai, Ax);
    if (anorm <= 0) anorm = 1.0;
    info->anorm = anorm;

    oop_t op;
    memset(&op, 0, sizeof op);
    op.n = n; op.Ap = Ap; op.Ai = Ai; op.Ax = Ax;
    int rc = SPEIGS_OK;
    if (which == SPEIGS_LM) {
        op.mode = OP_MULT;
        info->path = SPEIGS_PATH_DIRECT;
    } else if (!try_banded(n, Ap, Ai, Ax, opts, anorm, &op, info)) {
        double ta = now_s();
        int st = 0;
        op.chol = oracle_chol_factor(n, Ap, Ai, Ax, 0.0, &st);
        info->t_analyze = now_s() - ta;  /* ordering + symbolic + first numeric */
        if (!op.chol) { rc = SPEIGS_ERR_MEM; goto done; }
        double tf = now_s();
        double delta = 0.0;
        if (st == 1) {
            double d0 = anorm * (opts->shift0 > 0 ? opts->shift0 : 1e-6);
            double dcap = anorm * 1e-3;
            if (d0 > dcap) dcap = d0;
            for (delta = d0; delta <= dcap * 1.0000001; delta *= 10.0) {
                st = oracle_chol_refactor(op.chol, Ap, Ai, Ax, delta);
                if (st != 1) break;
            }
            if (st == 1) delta = 0.0;
        }
        if (st == 0) {
            op.mode = OP_CHOL; info->shift = delta; info->path = SPEIGS_PATH_SHIFTINV;
        } else {
            /* indefinite: LU at sigma = 0 (dense stand-in for UMFPACK) */
            oracle_chol_free(op.chol); op.chol = NULL;
            if (n > 6000) { rc = SPEIGS_ERR_FACTOR; goto done; }
            op.lu = (double *)calloc((size_t)n * n, sizeof(double));
            op.piv = (int64_t *)calloc((size_t)n, sizeof(int64_t));
            if (!op.lu || !op.piv) { rc = SPEIGS_ERR_MEM; goto done; }
            for (int64_t c = 0; c < n; c++)
                for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
                    int64_t r = Ai[p];
                    if (r > c) continue;
                    op.lu[r + (size_t)c * n] = Ax[p];
                    op.lu[c + (size_t)r * n] = Ax[p];
                }
            if (oracle_lu_factor(n, op.lu, op.piv)) { rc = SPEIGS_ERR_FACTOR; goto done; }
            op.mode = OP_LU; info->shift = 0.0; info->path = SPEIGS_PATH_LU;
        }
        info->t_factor = now_s() - tf;
    }
    {
        double ti = now_s();
        rc = lanczos(&op, n, k, opts, anorm, lambda, V, resid, info);
        info->t_iter = now_s() - ti;
    }
    if (rc == SPEIGS_OK) {
        okv *ord = (okv *)calloc((size_t)k, sizeof(okv));
        double *lc = (double *)malloc((size_t)k * sizeof(double));
        double *rcpy = resid ? (double *)malloc((size_t)k * sizeof(double)) : NULL;
        double *Vc = V ? (double *)malloc((size_t)n * k * sizeof(double)) : NULL;
        for (int64_t i = 0; i < k; i++) { ord[i].key = fabs(lambda[i]); ord[i].idx = i; }
        qsort(ord, (size_t)k, sizeof(okv), which == SPEIGS_SM ? okv_asc : okv_desc);
        memcpy(lc, lambda, (size_t)k * sizeof(double));
        if (rcpy) memcpy(rcpy, resid, (size_t)k * sizeof(double));
        if (Vc) memcpy(Vc, V, (size_t)n * k * sizeof(double));
        for (int64_t q = 0; q < k; q++) {
            int64_t i = ord[q].idx;
            lambda[q] = lc[i];
            if (resid) resid[q] = rcpy[i];
            if (V) memcpy(V + (size_t)q * n, Vc + (size_t)i * n, (size_t)n * sizeof(double));
        }
        free(ord); free(lc); free(rcpy); free(Vc);
    }
done:
    oracle_chol_free(op.chol);
    free(op.lu); free(op.piv); free(op.ab); free(op.perm); free(op.bw);
    return rc;
}

/* ------------------------------------------------- complex Hermitian ---- */
int oracle_speigs_z(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                    const double *Axr, const double *Axi, speigs_int k,
                    speigs_which which, const speigs_opts *opts, double *lambda,
                    double *Vr, double *Vi, double *resid, speigs_info *info)
{
    if (!Axi) return oracle_speigs(n, Ap, Ai, Axr, k, which, opts, lambda, Vr, resid, info);
    if (n <= 0 || k <= 0 || k > n || !Ap || !Ai || !Axr || !lambda) return SPEIGS_ERR_ARG;
    int64_t n2 = 2 * n, nnz = Ap[n], nr = 0, ns = 0;
    for (int64_t q = 0; q < nnz; q++) { if (Axr[q] != 0.0) nr++; if (Axi[q] != 0.0) ns++; }
    int64_t nz2 = 2 * (nr + ns);
    int64_t *Bp = (int64_t *)malloc(((size_t)n2 + 1) * sizeof(int64_t));
    int64_t *Bi = (int64_t *)malloc((size_t)(nz2 ? nz2 : 1) * sizeof(int64_t));
    double *Bx = (double *)malloc((size_t)(nz2 ? nz2 : 1) * sizeof(double));
    int64_t t = 0;
    /* [[R, -S], [S, R]]: column j is [R(:,j); S(:,j)], column n+j is [-S(:,j); R(:,j)] */
    for (int64_t jj = 0; jj < n; jj++) {
        Bp[jj] = t;
        for (int64_t q = Ap[jj]; q < Ap[jj + 1]; q++) if (Axr[q] != 0.0) { Bi[t] = Ai[q]; Bx[t++] = Axr[q]; }
        for (int64_t q = Ap[jj]; q < Ap[jj + 1]; q++) if (Axi[q] != 0.0) { Bi[t] = n + Ai[q]; Bx[t++] = Axi[q]; }
    }
    for (int64_t jj = 0; jj < n; jj++) {
        Bp[n + jj] = t;
        for (int64_t q = Ap[jj]; q < Ap[jj + 1]; q++) if (Axi[q] != 0.0) { Bi[t] = Ai[q]; Bx[t++] = -Axi[q]; }
        for (int64_t q = Ap[jj]; q < Ap[jj + 1]; q++) if (Axr[q] != 0.0) { Bi[t] = n + Ai[q]; Bx[t++] = Axr[q]; }
    }
    Bp[n2] = t;
    int64_t kk = IMIN(2 * k + 2, n2);
    double *l2 = (double *)calloc((size_t)kk, sizeof(double));
    double *r2 = (double *)calloc((size_t)kk, sizeof(double));
    double *V2 = (double *)calloc((size_t)n2 * kk, sizeof(double));
    double *ar = (double *)calloc((size_t)n * k, sizeof(double));
    double *ai = (double *)calloc((size_t)n * k, sizeof(double));
    double *wr = (double *)malloc((size_t)n * sizeof(double));
    double *wi = (double *)malloc((size_t)n * sizeof(double));
    int rc = oracle_speigs(n2, Bp, Bi, Bx, kk, which, opts, l2, V2, r2, info);
    if (rc) goto out;
    int64_t nacc = 0;
    for (int64_t c = 0; c < kk && nacc < k; c++) {
        const double *a = V2 + (size_t)c * n2, *b = a + n;
        double nn = 0.0;
        for (int64_t i = 0; i < n; i++) nn += a[i] * a[i] + b[i] * b[i];
        if (nn <= 0.0) continue;
        nn = 1.0 / sqrt(nn);
        for (int64_t i = 0; i < n; i++) { wr[i] = a[i] * nn; wi[i] = b[i] * nn; }
        /* complex Gram-Schmidt (two passes) against accepted directions */
        for (int pass = 0; pass < 2; pass++)
            for (int64_t u = 0; u < nacc; u++) {
                const double *ur = ar + (size_t)u * n, *ui = ai + (size_t)u * n;
                double re = 0.0, im = 0.0;
                for (int64_t i = 0; i < n; i++) {
                    re += ur[i] * wr[i] + ui[i] * wi[i];
                    im += ur[i] * wi[i] - ui[i] * wr[i];
                }
                for (int64_t i = 0; i < n; i++) {
                    wr[i] -= re * ur[i] - im * ui[i];
                    wi[i] -= re * ui[i] + im * ur[i];
                }
            }
        double wn = 0.0;
        for (int64_t i = 0; i < n; i++) wn += wr[i] * wr[i] + wi[i] * wi[i];
        wn = sqrt(wn);
        if (wn <= 0.1) continue;        /* an i-multiple of an accepted vector */
        for (int64_t i = 0; i < n; i++) {
            ar[(size_t)nacc * n + i] = wr[i] / wn;
            ai[(size_t)nacc * n + i] = wi[i] / wn;
        }
        lambda[nacc] = l2[c];
        if (resid) resid[nacc] = r2[c];
        nacc++;
    }
    if (Vr) memcpy(Vr, ar, (size_t)n * k * sizeof(double));
    if (Vi) memcpy(Vi, ai, (size_t)n * k * sizeof(double));
    info->nconv = nacc;
    if (nacc < k) info->flag = 1;
out:
    free(Bp); free(Bi); free(Bx); free(l2); free(r2); free(V2);
    free(ar); free(ai); free(wr); free(wi);
    return rc;
}
