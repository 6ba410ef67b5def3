/*
 * o_mf.c -- multifrontal supernodal Cholesky on the host cores, for the CPU
 * baseline (TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline times the
 * oracle's restatement of the driver and Lanczos loop, o_speigs.c, built over
 * this factor into build/liboracle_mf.so; no product code loads it).
 *
 * Stands in for the supernodal CHOLMOD path the reference calls
 * (cholmod_l_analyze / factorize_p / solve2, src/speigs_core.c:859-970 and
 * :86-100), with the same contract as o_chol.c's simplicial factor:
 *   ordering : nested dissection by breadth-first level structures (George and
 *              Liu 1981), as in o_chol.c, keeping the separator tree: every
 *              separator and every leaf region is one supernode whose columns
 *              are contiguous in the new order;
 *   symbolic : the rows of a supernode below its diagonal block are the later-
 *              ordered neighbours of its columns and its children's rows;
 *   numeric  : multifrontal: per front, A's entries plus the children's update
 *              matrices (extend-add), then dpotrf / dtrsm / dsyrk (BLAS-3,
 *              MKL); disjoint subtrees in parallel (OpenMP, one BLAS thread
 *              each), the top of the tree one front at a time with all threads
 *              in BLAS;
 *   solve    : multifrontal forward (dtrsv, dgemv into the update vector) and
 *              backward (gather, dgemv, dtrsv) sweeps, scheduled the same way.
 * Pivot test as o_chol.c: d <= 1e-11 (A_jj + delta) is "not positive
 * definite" (the driver's shift rescue then applies).
 * Pinned in tests/test_oracle.py against numpy (factor and solve) and against
 * o_chol.c's speigs on the same inputs.
 */
#include "oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* BLAS / LAPACK (Fortran interface, 32-bit integers) */
void dpotrf_(const char *uplo, const int *n, double *a, const int *lda, int *info);
void dtrsm_(const char *side, const char *uplo, const char *ta, const char *diag, const int *m, const int *n,
            const double *alpha, const double *a, const int *lda, double *b, const int *ldb);
void dsyrk_(const char *uplo, const char *trans, const int *n, const int *k, const double *alpha, const double *a,
            const int *lda, const double *beta, double *c, const int *ldc);
void dtrsv_(const char *uplo, const char *trans, const char *diag, const int *n, const double *a, const int *lda,
            double *x, const int *incx);
void dgemv_(const char *trans, const int *m, const int *n, const double *alpha, const double *a, const int *lda,
            const double *x, const int *incx, const double *beta, double *y, const int *incy);
void MKL_Set_Num_Threads(int n);   /* the C interface (by value) */

typedef struct {
    int64_t n, ns;
    int64_t *perm, *iperm;       /* perm[new] = old */
    int64_t *c0, *c1;            /* columns of supernode s: [c0, c1), postorder */
    int64_t *parent;             /* -1 for a root */
    int64_t *chptr, *chl;        /* children */
    int64_t *rptr, *rows;        /* rows below the diagonal block, sorted */
    int64_t *relptr;             /* positions of s's rows in its parent's front */
    int32_t *rel;
    int64_t *loff;               /* panel [L11; L21] of s: F x c at Lx + loff[s] (ld F) */
    double *Lx;
    int64_t *aptr, *aq, *apos;   /* A's stored entries by supernode: index in Ax, front position */
    double *dA;                  /* A's diagonal (new order), for the pivot test */
    double **upd;                /* update matrices between factor steps */
    int64_t *uoff;               /* solve: update vector of s at uvec + uoff[s] */
    double *uvec;
    int64_t *sub, nsub;          /* parallel subtrees (their roots), largest first */
    int64_t *top, ntop;          /* the supernodes above them, in postorder */
    int64_t *first;              /* first supernode of the subtree rooted at s (postorder range) */
    int threads;
    double flops;
    int64_t nnzL;
} mf_t;

/* ---------------------------------------------------------------- graph -- */
typedef struct { int64_t n; int64_t *xadj, *adj; } mgraph_t;

static int mgraph(int64_t n, const int64_t *Ap, const int64_t *Ai, mgraph_t *g)
{
    int64_t *deg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!deg) return -1;
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
            if (Ai[q] < j) { deg[Ai[q]]++; deg[j]++; }
    g->n = n;
    g->xadj = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t t = 0;
    for (int64_t i = 0; i < n; i++) { g->xadj[i] = t; t += deg[i]; }
    g->xadj[n] = t;
    g->adj = (int64_t *)malloc((size_t)(t ? t : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) deg[i] = g->xadj[i];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i < j) { g->adj[deg[i]++] = j; g->adj[deg[j]++] = i; }
        }
    free(deg);
    return 0;
}

/* ------------------------------------------------ dissection with tree -- */
typedef struct {
    const mgraph_t *g;
    int64_t *region, *level, *queue;
    int64_t next_region, next_pos;
    int64_t *perm;
    /* supernodes as they are closed (postorder) */
    int64_t ns, cap;
    int64_t *c0, *c1, *kid0, *kidn;   /* children: kids[kid0 .. kid0 + kidn) */
    int64_t *kids, nkids, kcap;
} nd_w;

static int64_t nd_bfs(nd_w *w, int64_t r, int64_t s, int64_t *height)
{
    const mgraph_t *g = w->g;
    int64_t head = 0, tail = 0;
    w->queue[tail++] = s;
    w->level[s] = 0;
    w->region[s] = -r;
    while (head < tail) {
        const int64_t v = w->queue[head++];
        for (int64_t p = g->xadj[v]; p < g->xadj[v + 1]; p++) {
            const int64_t u = g->adj[p];
            if (w->region[u] == r) {
                w->region[u] = -r;
                w->level[u] = w->level[v] + 1;
                w->queue[tail++] = u;
            }
        }
    }
    for (int64_t i = 0; i < tail; i++) w->region[w->queue[i]] = r;
    *height = w->level[w->queue[tail - 1]];
    return tail;
}

/* a supernode of the given nodes (numbered now) with the given children */
static int64_t nd_close(nd_w *w, const int64_t *nodes, int64_t cnt, const int64_t *ch, int64_t nch)
{
    if (w->ns == w->cap) {
        w->cap = w->cap ? 2 * w->cap : 1024;
        w->c0 = (int64_t *)realloc(w->c0, (size_t)w->cap * sizeof(int64_t));
        w->c1 = (int64_t *)realloc(w->c1, (size_t)w->cap * sizeof(int64_t));
        w->kid0 = (int64_t *)realloc(w->kid0, (size_t)w->cap * sizeof(int64_t));
        w->kidn = (int64_t *)realloc(w->kidn, (size_t)w->cap * sizeof(int64_t));
    }
    while (w->nkids + nch > w->kcap) {
        w->kcap = w->kcap ? 2 * w->kcap : 1024;
        w->kids = (int64_t *)realloc(w->kids, (size_t)w->kcap * sizeof(int64_t));
    }
    const int64_t s = w->ns++;
    w->c0[s] = w->next_pos;
    for (int64_t i = 0; i < cnt; i++) { w->perm[w->next_pos++] = nodes[i]; w->region[nodes[i]] = 0; }
    w->c1[s] = w->next_pos;
    w->kid0[s] = w->nkids;
    w->kidn[s] = nch;
    for (int64_t i = 0; i < nch; i++) w->kids[w->nkids++] = ch[i];
    return s;
}

/* dissect the region r (nodes[0..cnt)); appends the ids of its tree roots
 * to roots[*nroots] */
static void nd_tree(nd_w *w, int64_t *nodes, int64_t cnt, int64_t r, int64_t *roots, int64_t *nroots)
{
    if (cnt <= 64) {
        roots[(*nroots)++] = nd_close(w, nodes, cnt, NULL, 0);
        return;
    }
    int64_t s = nodes[0], h = 0, reached = nd_bfs(w, r, s, &h);
    if (reached < cnt) {
        /* disconnected: the reached component, then the rest */
        int64_t *buf = (int64_t *)malloc((size_t)cnt * sizeof(int64_t));
        const int64_t ra = w->next_region++, rb = w->next_region++;
        int64_t b = reached;
        for (int64_t i = 0; i < reached; i++) { buf[i] = w->queue[i]; w->region[w->queue[i]] = ra; }
        for (int64_t i = 0; i < cnt; i++)
            if (w->region[nodes[i]] == r) { buf[b++] = nodes[i]; w->region[nodes[i]] = rb; }
        nd_tree(w, buf, reached, ra, roots, nroots);
        nd_tree(w, buf + reached, cnt - reached, rb, roots, nroots);
        free(buf);
        return;
    }
    for (int it = 0; it < 4; it++) {        /* pseudo-peripheral root */
        int64_t best = -1, bd = 0;
        for (int64_t i = 0; i < reached; i++) {
            const int64_t v = w->queue[i];
            if (w->level[v] == h) {
                const int64_t d = w->g->xadj[v + 1] - w->g->xadj[v];
                if (best < 0 || d < bd) { best = v; bd = d; }
            }
        }
        int64_t h2 = 0;
        nd_bfs(w, r, best, &h2);
        if (h2 <= h) { nd_bfs(w, r, s, &h); break; }
        s = best;
        h = h2;
    }
    if (h < 2) {
        roots[(*nroots)++] = nd_close(w, nodes, cnt, NULL, 0);
        return;
    }
    int64_t *lsz = (int64_t *)calloc((size_t)h + 1, sizeof(int64_t));
    for (int64_t i = 0; i < cnt; i++) lsz[w->level[w->queue[i]]]++;
    int64_t cum = 0, sep = -1, med = -1, bsz = 0;
    for (int64_t l = 0; l <= h; l++) {
        const int64_t before = cum;
        cum += lsz[l];
        if (med < 0 && 2 * cum >= cnt) med = l;
        if (l > 0 && l < h && 20 * before >= 7 * cnt && 20 * (cnt - cum) >= 7 * cnt)
            if (sep < 0 || lsz[l] < bsz) { sep = l; bsz = lsz[l]; }
    }
    free(lsz);
    if (sep < 0) sep = med;
    if (sep <= 0) sep = 1;
    if (sep >= h) sep = h - 1;
    int64_t na = 0, nb = 0;
    for (int64_t i = 0; i < cnt; i++) {
        const int64_t l = w->level[w->queue[i]];
        na += l < sep;
        nb += l > sep;
    }
    int64_t *buf = (int64_t *)malloc((size_t)cnt * sizeof(int64_t));
    const int64_t ra = w->next_region++, rb = w->next_region++;
    int64_t ia = 0, ib = na, is = na + nb;
    for (int64_t i = 0; i < cnt; i++) {
        const int64_t v = w->queue[i], l = w->level[v];
        if (l < sep) { buf[ia++] = v; w->region[v] = ra; }
        else if (l > sep) { buf[ib++] = v; w->region[v] = rb; }
        else { buf[is++] = v; w->region[v] = -1; }   /* held out of both halves */
    }
    int64_t *kids = (int64_t *)malloc((size_t)(cnt + 1) * sizeof(int64_t));
    int64_t nk = 0;
    nd_tree(w, buf, na, ra, kids, &nk);
    nd_tree(w, buf + na, nb, rb, kids, &nk);
    roots[(*nroots)++] = nd_close(w, buf + na + nb, cnt - na - nb, kids, nk);
    free(kids);
    free(buf);
}

/* -------------------------------------------------------------- analyze -- */
static int cmp64(const void *a, const void *b)
{
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

static void mf_free(mf_t *F)
{
    if (!F) return;
    free(F->perm); free(F->iperm); free(F->c0); free(F->c1); free(F->parent); free(F->chptr); free(F->chl);
    free(F->rptr); free(F->rows); free(F->relptr); free(F->rel); free(F->loff); free(F->Lx); free(F->aptr);
    free(F->aq); free(F->apos); free(F->dA); free(F->uoff); free(F->uvec); free(F->sub); free(F->top);
    free(F->first);
    if (F->upd) {
        for (int64_t s = 0; s < F->ns; s++) free(F->upd[s]);
        free(F->upd);
    }
    free(F);
}

static mf_t *mf_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai)
{
    mf_t *F = (mf_t *)calloc(1, sizeof(mf_t));
    mgraph_t g;
    if (!F || mgraph(n, Ap, Ai, &g)) { free(F); return NULL; }
    F->n = n;
    F->threads = omp_get_max_threads();
    F->perm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    F->iperm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    nd_w w;
    memset(&w, 0, sizeof w);
    w.g = &g;
    w.region = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    w.level = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    w.queue = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    w.perm = F->perm;
    w.next_region = 2;
    int64_t *nodes = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *roots = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t)), nroots = 0;
    for (int64_t i = 0; i < n; i++) { nodes[i] = i; w.region[i] = 1; }
    nd_tree(&w, nodes, n, 1, roots, &nroots);
    free(nodes); free(roots); free(w.region); free(w.level); free(w.queue);
    for (int64_t i = 0; i < n; i++) F->iperm[F->perm[i]] = i;
    const int64_t ns = F->ns = w.ns;
    F->c0 = w.c0;
    F->c1 = w.c1;
    F->parent = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
    F->chptr = (int64_t *)malloc(((size_t)ns + 1) * sizeof(int64_t));
    F->chl = (int64_t *)malloc((size_t)(w.nkids ? w.nkids : 1) * sizeof(int64_t));
    for (int64_t s = 0; s < ns; s++) F->parent[s] = -1;
    F->chptr[0] = 0;
    for (int64_t s = 0; s < ns; s++) {
        F->chptr[s + 1] = F->chptr[s] + w.kidn[s];
        for (int64_t k = 0; k < w.kidn[s]; k++) {
            F->chl[F->chptr[s] + k] = w.kids[w.kid0[s] + k];
            F->parent[w.kids[w.kid0[s] + k]] = s;
        }
    }
    free(w.kid0); free(w.kidn); free(w.kids);
    /* rows: later neighbours of the columns, and the children's later rows */
    int64_t *snode = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t s = 0; s < ns; s++)
        for (int64_t j = F->c0[s]; j < F->c1[s]; j++) snode[j] = s;
    int64_t *mark = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) mark[i] = -1;
    int64_t cap = 4 * n + 16, tot = 0;
    F->rptr = (int64_t *)malloc(((size_t)ns + 1) * sizeof(int64_t));
    F->rows = (int64_t *)malloc((size_t)cap * sizeof(int64_t));
    F->rptr[0] = 0;
    for (int64_t s = 0; s < ns; s++) {
        const int64_t hi = F->c1[s], start = tot;
        for (int64_t k = F->chptr[s]; k < F->chptr[s + 1]; k++) {
            const int64_t ch = F->chl[k];
            for (int64_t q = F->rptr[ch]; q < F->rptr[ch + 1]; q++) {
                const int64_t v = F->rows[q];
                if (v >= hi && mark[v] != s) {
                    mark[v] = s;
                    if (tot == cap) { cap *= 2; F->rows = (int64_t *)realloc(F->rows, (size_t)cap * sizeof(int64_t)); }
                    F->rows[tot++] = v;
                }
            }
        }
        for (int64_t j = F->c0[s]; j < hi; j++) {
            const int64_t old = F->perm[j];
            for (int64_t p = g.xadj[old]; p < g.xadj[old + 1]; p++) {
                const int64_t v = F->iperm[g.adj[p]];
                if (v >= hi && mark[v] != s) {
                    mark[v] = s;
                    if (tot == cap) { cap *= 2; F->rows = (int64_t *)realloc(F->rows, (size_t)cap * sizeof(int64_t)); }
                    F->rows[tot++] = v;
                }
            }
        }
        qsort(F->rows + start, (size_t)(tot - start), sizeof(int64_t), cmp64);
        F->rptr[s + 1] = tot;
    }
    free(mark);
    free(g.xadj); free(g.adj);
    /* positions of each supernode's rows in its parent's front */
    F->relptr = (int64_t *)malloc(((size_t)ns + 1) * sizeof(int64_t));
    F->relptr[0] = 0;
    for (int64_t s = 0; s < ns; s++) F->relptr[s + 1] = F->relptr[s] + (F->rptr[s + 1] - F->rptr[s]);
    F->rel = (int32_t *)malloc((size_t)(F->relptr[ns] ? F->relptr[ns] : 1) * sizeof(int32_t));
    for (int64_t s = 0; s < ns; s++) {
        const int64_t p = F->parent[s];
        if (p < 0) continue;
        const int64_t cp = F->c1[p] - F->c0[p];
        int64_t k = F->rptr[p];
        for (int64_t q = F->rptr[s]; q < F->rptr[s + 1]; q++) {
            const int64_t v = F->rows[q];
            if (v < F->c1[p]) {
                F->rel[F->relptr[s] + q - F->rptr[s]] = (int32_t)(v - F->c0[p]);
            } else {
                while (F->rows[k] < v) k++;
                F->rel[F->relptr[s] + q - F->rptr[s]] = (int32_t)(cp + (k - F->rptr[p]));
            }
        }
    }
    /* panel offsets, counts */
    F->loff = (int64_t *)malloc(((size_t)ns + 1) * sizeof(int64_t));
    F->uoff = (int64_t *)malloc(((size_t)ns + 1) * sizeof(int64_t));
    F->loff[0] = F->uoff[0] = 0;
    double *work = (double *)malloc((size_t)ns * sizeof(double));
    for (int64_t s = 0; s < ns; s++) {
        const int64_t c = F->c1[s] - F->c0[s], r = F->rptr[s + 1] - F->rptr[s], Fr = c + r;
        F->loff[s + 1] = F->loff[s] + Fr * c;
        F->uoff[s + 1] = F->uoff[s] + r;
        F->nnzL += c * (c + 1) / 2 + r * c;
        double fl = 0.0;
        for (int64_t k = 0; k < c; k++) fl += (double)(Fr - k) * (double)(Fr - k);
        F->flops += fl;
        work[s] = fl;
    }
    /* A's entries grouped by supernode (counting sort), each with its front position */
    const int64_t nnz = Ap[n];
    F->aptr = (int64_t *)calloc((size_t)ns + 1, sizeof(int64_t));
    F->aq = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    F->apos = (int64_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int64_t));
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            if (Ai[q] > j) continue;
            const int64_t a = F->iperm[Ai[q]], b = F->iperm[j];
            F->aptr[snode[a < b ? a : b] + 1]++;
        }
    for (int64_t s = 0; s < ns; s++) F->aptr[s + 1] += F->aptr[s];
    int64_t *fill = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
    for (int64_t s = 0; s < ns; s++) fill[s] = F->aptr[s];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            if (Ai[q] > j) continue;
            const int64_t a = F->iperm[Ai[q]], b = F->iperm[j];
            const int64_t col = a < b ? a : b, row = a < b ? b : a, s = snode[col];
            const int64_t c = F->c1[s] - F->c0[s], Fr = c + F->rptr[s + 1] - F->rptr[s];
            int64_t lr;
            if (row < F->c1[s]) {
                lr = row - F->c0[s];
            } else {
                const int64_t *R = F->rows + F->rptr[s];
                int64_t lo = 0, hi2 = F->rptr[s + 1] - F->rptr[s];
                while (lo < hi2) {
                    const int64_t mid = (lo + hi2) / 2;
                    if (R[mid] < row) lo = mid + 1;
                    else hi2 = mid;
                }
                lr = c + lo;
            }
            F->aq[fill[s]] = q;
            F->apos[fill[s]++] = lr + (col - F->c0[s]) * Fr;
        }
    free(fill);
    free(snode);
    /* schedule: subtrees of at most total / (8 threads) work run in parallel */
    double *sw = (double *)malloc((size_t)ns * sizeof(double));
    F->first = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
    for (int64_t s = 0; s < ns; s++) {
        sw[s] = work[s];
        F->first[s] = s;
        for (int64_t k = F->chptr[s]; k < F->chptr[s + 1]; k++) {
            const int64_t ch = F->chl[k];
            sw[s] += sw[ch];
            if (F->first[ch] < F->first[s]) F->first[s] = F->first[ch];
        }
    }
    double total = 0.0;
    for (int64_t s = 0; s < ns; s++)
        if (F->parent[s] < 0) total += sw[s];
    const double lim = total / (8.0 * (F->threads > 0 ? F->threads : 1));
    char *istop = (char *)calloc((size_t)ns, 1);
    for (int64_t s = ns - 1; s >= 0; s--) {     /* parents before children */
        const int64_t p = F->parent[s];
        istop[s] = (p < 0 || istop[p]) && sw[s] > lim;
    }
    F->sub = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
    F->top = (int64_t *)malloc((size_t)ns * sizeof(int64_t));
    for (int64_t s = 0; s < ns; s++) {
        const int64_t p = F->parent[s];
        if (istop[s]) F->top[F->ntop++] = s;
        else if (p < 0 || istop[p]) F->sub[F->nsub++] = s;
    }
    for (int64_t i = 1; i < F->nsub; i++) {     /* largest subtrees first */
        const int64_t v = F->sub[i];
        int64_t j = i;
        while (j > 0 && sw[F->sub[j - 1]] < sw[v]) { F->sub[j] = F->sub[j - 1]; j--; }
        F->sub[j] = v;
    }
    free(istop); free(sw); free(work);
    F->Lx = (double *)malloc((size_t)(F->loff[ns] ? F->loff[ns] : 1) * sizeof(double));
    F->dA = (double *)malloc((size_t)n * sizeof(double));
    F->uvec = (double *)malloc((size_t)(F->uoff[ns] ? F->uoff[ns] : 1) * sizeof(double));
    F->upd = (double **)calloc((size_t)ns, sizeof(double *));
    if (!F->Lx || !F->dA || !F->uvec || !F->upd) { mf_free(F); return NULL; }
    return F;
}

/* -------------------------------------------------------------- numeric -- */
/* one front: 0 ok, 1 not positive definite, -1 out of memory.  par: the
 * front is alone (the top of the tree): its zero fill and extend-add run on
 * all threads, by columns (a child's columns land on distinct columns of the
 * front, so no two threads write one entry) */
static int mf_front(mf_t *F, int64_t s, const double *Ax, double delta, int par)
{
    const int64_t c = F->c1[s] - F->c0[s], r = F->rptr[s + 1] - F->rptr[s], Fr = c + r;
    double *P = F->Lx + F->loff[s];                    /* [L11; L21], F x c, ld F */
    double *U = NULL;                                  /* F22, r x r, ld r */
    if (r > 0) {
        U = (double *)malloc((size_t)(r * r) * sizeof(double));
        if (!U) return -1;
    }
#pragma omp parallel for schedule(static) if (par)
    for (int64_t j = 0; j < Fr; j++) {
        if (j < c) memset(P + j * Fr, 0, (size_t)Fr * sizeof(double));
        else memset(U + (j - c) * r, 0, (size_t)r * sizeof(double));
    }
    for (int64_t k = 0; k < c; k++) P[k + k * Fr] = delta;
    for (int64_t e = F->aptr[s]; e < F->aptr[s + 1]; e++) P[F->apos[e]] += Ax[F->aq[e]];
    for (int64_t k = 0; k < c; k++) F->dA[F->c0[s] + k] = P[k + k * Fr];
    /* extend-add of the children's update matrices, children in order */
    for (int64_t k = F->chptr[s]; k < F->chptr[s + 1]; k++) {
        const int64_t ch = F->chl[k], rc = F->rptr[ch + 1] - F->rptr[ch];
        const int32_t *rl = F->rel + F->relptr[ch];
        double *Uc = F->upd[ch];
#pragma omp parallel for schedule(dynamic, 16) if (par)
        for (int64_t jj = 0; jj < rc; jj++) {
            const int64_t pj = rl[jj];
            for (int64_t ii = jj; ii < rc; ii++) {
                const int64_t pi = rl[ii];
                const double v = Uc[ii + jj * rc];
                if (pj < c) P[pi + pj * Fr] += v;
                else U[(pi - c) + (pj - c) * r] += v;
            }
        }
        free(Uc);
        F->upd[ch] = NULL;
    }
    const int ic = (int)c, ir = (int)r, iF = (int)Fr;
    int info = 0;
    dpotrf_("L", &ic, P, &iF, &info);
    if (info != 0) { free(U); return 1; }
    for (int64_t k = 0; k < c; k++) {
        const double d = P[k + k * Fr] * P[k + k * Fr];
        if (d <= 1e-11 * fabs(F->dA[F->c0[s] + k])) { free(U); return 1; }
    }
    if (r > 0) {
        const double one = 1.0, mone = -1.0;
        dtrsm_("R", "L", "T", "N", &ir, &ic, &one, P, &iF, P + c, &iF);
        dsyrk_("L", "N", &ir, &ic, &mone, P + c, &iF, &one, U, &ir);
        F->upd[s] = U;
    }
    return 0;
}

static int mf_numeric(mf_t *F, const double *Ax, double delta)
{
    int bad = 0;
    for (int64_t s = 0; s < F->ns; s++) { free(F->upd[s]); F->upd[s] = NULL; }
    /* disjoint subtrees in parallel, one BLAS thread each */
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int64_t i = 0; i < F->nsub; i++) {
        const int64_t root = F->sub[i];
        for (int64_t s = F->first[root]; s <= root && !bad; s++) {
            const int st = mf_front(F, s, Ax, delta, 0);
            if (st) bad |= st < 0 ? 2 : 1;
        }
    }
    /* the top of the tree, one front at a time, all threads in BLAS */
    MKL_Set_Num_Threads(F->threads);
    for (int64_t i = 0; i < F->ntop && !bad; i++) {
        const int st = mf_front(F, F->top[i], Ax, delta, 1);
        if (st) bad |= st < 0 ? 2 : 1;
    }
    for (int64_t s = 0; s < F->ns; s++) { free(F->upd[s]); F->upd[s] = NULL; }
    return bad & 2 ? -1 : bad ? 1 : 0;
}

/* ---------------------------------------------------------------- solve -- */
static void mf_fwd_front(const mf_t *F, int64_t s, double *y, double *w)
{
    const int64_t c = F->c1[s] - F->c0[s], r = F->rptr[s + 1] - F->rptr[s], Fr = c + r;
    const double *P = F->Lx + F->loff[s];
    double *yc = y + F->c0[s], *u = F->uvec + F->uoff[s];
    for (int64_t k = 0; k < r; k++) u[k] = 0.0;
    (void)w;
    for (int64_t k = F->chptr[s]; k < F->chptr[s + 1]; k++) {
        const int64_t ch = F->chl[k], rc = F->rptr[ch + 1] - F->rptr[ch];
        const int32_t *rl = F->rel + F->relptr[ch];
        const double *uc = F->uvec + F->uoff[ch];
        for (int64_t q = 0; q < rc; q++) {
            if (rl[q] < c) yc[rl[q]] += uc[q];
            else u[rl[q] - c] += uc[q];
        }
    }
    const int ic = (int)c, ir = (int)r, iF = (int)Fr, one_i = 1;
    dtrsv_("L", "N", "N", &ic, P, &iF, yc, &one_i);
    if (r > 0) {
        const double mone = -1.0, one = 1.0;
        dgemv_("N", &ir, &ic, &mone, P + c, &iF, yc, &one_i, &one, u, &one_i);
    }
}

static void mf_bwd_front(const mf_t *F, int64_t s, double *x, double *xr)
{
    const int64_t c = F->c1[s] - F->c0[s], r = F->rptr[s + 1] - F->rptr[s], Fr = c + r;
    const double *P = F->Lx + F->loff[s];
    double *xc = x + F->c0[s];
    const int ic = (int)c, ir = (int)r, iF = (int)Fr, one_i = 1;
    if (r > 0) {
        const int64_t *R = F->rows + F->rptr[s];
        for (int64_t q = 0; q < r; q++) xr[q] = x[R[q]];
        const double mone = -1.0, one = 1.0;
        dgemv_("T", &ir, &ic, &mone, P + c, &iF, xr, &one_i, &one, xc, &one_i);
    }
    dtrsv_("L", "T", "N", &ic, P, &iF, xc, &one_i);
}

static void mf_solve(const mf_t *F, const double *b, double *x)
{
    const int64_t n = F->n;
    double *y = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; i++) y[i] = b[F->perm[i]];
    /* forward: y holds W1 then Y in place, update vectors in uvec */
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < F->nsub; i++) {
        const int64_t root = F->sub[i];
        for (int64_t s = F->first[root]; s <= root; s++) mf_fwd_front(F, s, y, NULL);
    }
    MKL_Set_Num_Threads(F->threads);
    for (int64_t i = 0; i < F->ntop; i++) mf_fwd_front(F, F->top[i], y, NULL);
    /* backward: the top first (reverse postorder), then the subtrees */
    int64_t rmax = 1;
    for (int64_t s = 0; s < F->ns; s++)
        if (F->rptr[s + 1] - F->rptr[s] > rmax) rmax = F->rptr[s + 1] - F->rptr[s];
    double *xr = (double *)malloc((size_t)rmax * sizeof(double));
    for (int64_t i = F->ntop - 1; i >= 0; i--) mf_bwd_front(F, F->top[i], y, xr);
    free(xr);
#pragma omp parallel
    {
        double *xr2 = (double *)malloc((size_t)rmax * sizeof(double));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < F->nsub; i++) {
            const int64_t root = F->sub[i];
            for (int64_t s = root; s >= F->first[root]; s--) mf_bwd_front(F, s, y, xr2);
        }
        free(xr2);
    }
    for (int64_t i = 0; i < n; i++) x[F->perm[i]] = y[i];
    free(y);
}

/* ---------------------------------------- the oracle_chol-shaped entry -- */
oracle_chol *oracle_mf_factor(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, double delta,
                              int *status)
{
    mf_t *F = mf_analyze(n, Ap, Ai);
    if (!F) { *status = -1; return NULL; }
    *status = mf_numeric(F, Ax, delta);
    return (oracle_chol *)F;
}

int oracle_mf_refactor(oracle_chol *Fc, const int64_t *Ap, const int64_t *Ai, const double *Ax, double delta)
{
    (void)Ap;
    (void)Ai;
    return mf_numeric((mf_t *)Fc, Ax, delta);
}

void oracle_mf_solve(const oracle_chol *Fc, const double *b, double *x) { mf_solve((const mf_t *)Fc, b, x); }

int64_t oracle_mf_nnzL(const oracle_chol *Fc) { return ((const mf_t *)Fc)->nnzL; }
double oracle_mf_flops(const oracle_chol *Fc) { return ((const mf_t *)Fc)->flops; }
void oracle_mf_free(oracle_chol *Fc) { mf_free((mf_t *)Fc); }

/* threads of the next factorizations and solves (OpenMP and BLAS) */
void oracle_mf_set_threads(int t)
{
    if (t < 1) t = 1;
    omp_set_num_threads(t);
    MKL_Set_Num_Threads(t);
}
