/*
 * o_chol.c -- sparse Cholesky of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Stands in for CHOLMOD's analyze/factorize_p/solve2 as the reference calls
 * them (src/speigs_core.c:933-970 and :86-100): x = P^T L^-T L^-1 P b with
 * L L^T = P (A + delta I) P^T, A read from its upper triangle.
 *   ordering : nested dissection by breadth-first level structures -- the
 *              middle level of a rooted level structure from a pseudo-
 *              peripheral node is a vertex separator (George and Liu 1981);
 *   symbolic : elimination tree (Liu's algorithm with path compression) and
 *              row patterns as elimination-tree reaches;
 *   numeric  : up-looking (row-by-row) Cholesky, one sparse triangular solve
 *              per row (Davis 2006, ch. 4).
 * A pivot d <= 1e-11 * C(k,k) is reported as "not positive definite"; the
 * relative floor catches the rounding-level last pivot of a singular PSD matrix
 * (a Neumann Laplacian), which the driver then rescues with a shift exactly as
 * src/speigs_core.c:953-970 does.
 * Pinned in tests/test_oracle.py: ||P(A+dI)P^T - L L^T|| and solve residuals
 * against numpy on small matrices, and the known-answer solve on the
 * separable Laplacian eigenvector.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

struct oracle_chol {
    int64_t n;
    int64_t *perm, *iperm;     /* perm[new] = old */
    int64_t *parent;           /* elimination tree */
    int64_t *Lp, *Li;          /* CSC of L, diagonal first in each column */
    double  *Lx;
    int64_t *Cp, *Ci;          /* permuted upper triangle pattern of A */
    int64_t *Cmap;             /* Cmap[q]: position in A's arrays of C entry q */
    double  *Cx;
    double   flops;
};

/* ------------------------------------------------------------ graph ----- */
typedef struct { int64_t n; int64_t *xadj, *adj; } graph_t;

static int build_graph(int64_t n, const int64_t *Ap, const int64_t *Ai, graph_t *g)
{
    int64_t *deg = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    if (!deg) return -1;
    /* symmetric adjacency from the upper triangle only (i < j), like stype=1 */
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i < j) { deg[i]++; deg[j]++; }
        }
    g->n = n;
    g->xadj = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t tot = 0;
    for (int64_t i = 0; i < n; i++) { g->xadj[i] = tot; tot += deg[i]; }
    g->xadj[n] = tot;
    g->adj = (int64_t *)malloc((size_t)(tot ? tot : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) deg[i] = g->xadj[i];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i < j) { g->adj[deg[i]++] = j; g->adj[deg[j]++] = i; }
        }
    free(deg);
    return 0;
}

/* ------------------------------------------------ nested dissection ----- */
typedef struct {
    const graph_t *g;
    int64_t *region;   /* region stamp per node */
    int64_t *level;    /* BFS level scratch */
    int64_t *queue;
    int64_t *perm;
    int64_t stamp;
} nd_t;

/* BFS inside region `r` from `s`; returns number reached, fills queue, level */
static int64_t bfs(nd_t *w, int64_t r, int64_t s, int64_t *height)
{
    const graph_t *g = w->g;
    int64_t head = 0, tail = 0;
    w->queue[tail++] = s; w->level[s] = 0;
    int64_t stampv = ++w->stamp;   /* visited marker lives in level via -stamp trick */
    (void)stampv;
    /* use region value r+1 temporarily negative to mark visited */
    w->region[s] = -(r + 1);
    int64_t h = 0;
    while (head < tail) {
        int64_t v = w->queue[head++];
        for (int64_t p = g->xadj[v]; p < g->xadj[v + 1]; p++) {
            int64_t u = g->adj[p];
            if (w->region[u] == r) {
                w->region[u] = -(r + 1);
                w->level[u] = w->level[v] + 1;
                if (w->level[u] > h) h = w->level[u];
                w->queue[tail++] = u;
            }
        }
    }
    for (int64_t i = 0; i < tail; i++) w->region[w->queue[i]] = r;
    *height = h;
    return tail;
}

static int64_t next_region = 1;

/* order the `cnt` nodes in nodes[] (all carry region id r) into perm[pos..] */
static void nd_rec(nd_t *w, int64_t *nodes, int64_t cnt, int64_t r, int64_t pos)
{
    const graph_t *g = w->g;
    if (cnt <= 48) {
        for (int64_t i = 0; i < cnt; i++) { w->perm[pos + i] = nodes[i]; w->region[nodes[i]] = 0; }
        return;
    }
    /* pseudo-peripheral node: repeat BFS from a min-degree node of the last level */
    int64_t s = nodes[0], h = 0, reached = bfs(w, r, s, &h);
    for (int it = 0; it < 4 && reached == cnt; it++) {
        int64_t best = -1, bdeg = 0;
        for (int64_t i = 0; i < reached; i++) {
            int64_t v = w->queue[i];
            if (w->level[v] == h) {
                int64_t d = g->xadj[v + 1] - g->xadj[v];
                if (best < 0 || d < bdeg) { best = v; bdeg = d; }
            }
        }
        int64_t h2 = 0;
        int64_t got = bfs(w, r, best, &h2);
        (void)got;
        if (h2 <= h) { /* no gain: redo BFS from the best root found so far */
            reached = bfs(w, r, s, &h);
            break;
        }
        s = best; h = h2;
    }
    int64_t *buf = (int64_t *)malloc((size_t)cnt * sizeof(int64_t));
    if (reached < cnt) {
        /* disconnected region: the reached component first, the rest after */
        int64_t a = 0, b = reached;
        int64_t ra = next_region++, rb = next_region++;
        for (int64_t i = 0; i < reached; i++) { buf[a++] = w->queue[i]; }
        for (int64_t i = 0; i < reached; i++) w->region[w->queue[i]] = ra;
        for (int64_t i = 0; i < cnt; i++)
            if (w->region[nodes[i]] == r) { buf[b++] = nodes[i]; w->region[nodes[i]] = rb; }
        nd_rec(w, buf, reached, ra, pos);
        nd_rec(w, buf + reached, cnt - reached, rb, pos + reached);
        free(buf);
        return;
    }
    if (h < 2) {
        for (int64_t i = 0; i < cnt; i++) { w->perm[pos + i] = nodes[i]; w->region[nodes[i]] = 0; }
        free(buf);
        return;
    }
    /* level sizes, then the separator level: among levels whose cumulative
     * split is within [0.35, 0.65], the smallest; else the median level */
    int64_t *lsz = (int64_t *)calloc((size_t)h + 1, sizeof(int64_t));
    for (int64_t i = 0; i < cnt; i++) lsz[w->level[w->queue[i]]]++;
    int64_t cum = 0, sep = -1, med = -1, bestsz = 0;
    for (int64_t l = 0; l <= h; l++) {
        int64_t before = cum; cum += lsz[l];
        if (med < 0 && 2 * cum >= cnt) med = l;
        if (l > 0 && l < h && 20 * before >= 7 * cnt && 20 * (cnt - cum) >= 7 * cnt) {
            if (sep < 0 || lsz[l] < bestsz) { sep = l; bestsz = lsz[l]; }
        }
    }
    if (sep < 0) sep = med;
    if (sep <= 0) sep = 1;
    if (sep >= h) sep = h - 1;
    int64_t na = 0, nb = 0, ns = 0;
    for (int64_t i = 0; i < cnt; i++) {
        int64_t l = w->level[w->queue[i]];
        if (l < sep) na++; else if (l > sep) nb++; else ns++;
    }
    int64_t ra = next_region++, rb = next_region++;
    int64_t ia = 0, ib = na, is = na + nb;
    for (int64_t i = 0; i < cnt; i++) {
        int64_t v = w->queue[i], l = w->level[v];
        if (l < sep) { buf[ia++] = v; w->region[v] = ra; }
        else if (l > sep) { buf[ib++] = v; w->region[v] = rb; }
        else { buf[is++] = v; w->region[v] = 0; }
    }
    free(lsz);
    for (int64_t i = 0; i < ns; i++) w->perm[pos + na + nb + i] = buf[na + nb + i];
    nd_rec(w, buf, na, ra, pos);
    nd_rec(w, buf + na, nb, rb, pos + na);
    free(buf);
}

int oracle_nd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *perm)
{
    graph_t g;
    if (build_graph(n, Ap, Ai, &g)) return -1;
    nd_t w;
    w.g = &g;
    w.region = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    w.level = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    w.queue = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    w.perm = perm;
    w.stamp = 0;
    int64_t *nodes = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    next_region = 2;
    for (int64_t i = 0; i < n; i++) { nodes[i] = i; w.region[i] = 1; }
    nd_rec(&w, nodes, n, 1, 0);
    free(nodes); free(w.region); free(w.level); free(w.queue);
    free(g.xadj); free(g.adj);
    return 0;
}

/* ------------------------------------------------------ symbolic -------- */
static int build_C(oracle_chol *F, const int64_t *Ap, const int64_t *Ai)
{
    int64_t n = F->n;
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i > j) continue;               /* upper triangle only */
            int64_t pi = F->iperm[i], pj = F->iperm[j];
            cnt[pi > pj ? pi : pj]++;
        }
    F->Cp = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t tot = 0;
    for (int64_t j = 0; j < n; j++) { F->Cp[j] = tot; tot += cnt[j]; }
    F->Cp[n] = tot;
    F->Ci = (int64_t *)malloc((size_t)(tot ? tot : 1) * sizeof(int64_t));
    F->Cmap = (int64_t *)malloc((size_t)(tot ? tot : 1) * sizeof(int64_t));
    F->Cx = (double *)malloc((size_t)(tot ? tot : 1) * sizeof(double));
    for (int64_t j = 0; j < n; j++) cnt[j] = F->Cp[j];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i > j) continue;
            int64_t pi = F->iperm[i], pj = F->iperm[j];
            int64_t col = pi > pj ? pi : pj, row = pi > pj ? pj : pi;
            F->Ci[cnt[col]] = row; F->Cmap[cnt[col]] = q; cnt[col]++;
        }
    free(cnt);
    return 0;
}

static void etree(oracle_chol *F)
{
    int64_t n = F->n;
    int64_t *anc = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    for (int64_t k = 0; k < n; k++) {
        F->parent[k] = -1; anc[k] = -1;
        for (int64_t q = F->Cp[k]; q < F->Cp[k + 1]; q++) {
            int64_t i = F->Ci[q];
            while (i != -1 && i < k) {
                int64_t nxt = anc[i];
                anc[i] = k;
                if (nxt == -1) { F->parent[i] = k; break; }
                i = nxt;
            }
        }
    }
    free(anc);
}

/* row pattern of L(k, 0:k-1) = union of etree paths from C(0:k-1,k) */
static int64_t ereach(const oracle_chol *F, int64_t k, int64_t *mark, int64_t *pat)
{
    int64_t np = 0;
    mark[k] = k;
    for (int64_t q = F->Cp[k]; q < F->Cp[k + 1]; q++) {
        int64_t i = F->Ci[q];
        while (i != -1 && i < k && mark[i] != k) { mark[i] = k; pat[np++] = i; i = F->parent[i]; }
    }
    return np;
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

/* ------------------------------------------------------- numeric -------- */
static int numeric(oracle_chol *F, const double *Ax, double delta)
{
    int64_t n = F->n;
    for (int64_t q = 0; q < F->Cp[n]; q++) F->Cx[q] = Ax[F->Cmap[q]];
    int64_t *mark = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *pat = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *fill = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    double *x = (double *)calloc((size_t)n, sizeof(double));
    for (int64_t i = 0; i < n; i++) { mark[i] = -1; fill[i] = F->Lp[i] + 1; }
    int status = 0;
    double fl = 0.0;
    for (int64_t k = 0; k < n && !status; k++) {
        int64_t np = ereach(F, k, mark, pat);
        qsort(pat, (size_t)np, sizeof(int64_t), cmp_i64);
        double ckk = 0.0;
        for (int64_t q = F->Cp[k]; q < F->Cp[k + 1]; q++) {
            int64_t i = F->Ci[q];
            if (i == k) ckk += F->Cx[q]; else x[i] += F->Cx[q];
        }
        ckk += delta;
        double d = ckk;
        for (int64_t t = 0; t < np; t++) {
            int64_t j = pat[t];
            double lkj = x[j] / F->Lx[F->Lp[j]];
            x[j] = 0.0;
            for (int64_t p = F->Lp[j] + 1; p < fill[j]; p++) x[F->Li[p]] -= F->Lx[p] * lkj;
            fl += 2.0 * (double)(fill[j] - F->Lp[j] - 1) + 2.0;
            d -= lkj * lkj;
            F->Li[fill[j]] = k; F->Lx[fill[j]] = lkj; fill[j]++;
        }
        if (!(d > 1e-11 * fabs(ckk))) { status = 1; break; }
        F->Li[F->Lp[k]] = k; F->Lx[F->Lp[k]] = sqrt(d);
    }
    F->flops = fl;
    free(mark); free(pat); free(fill); free(x);
    return status;
}

oracle_chol *oracle_chol_factor(int64_t n, const int64_t *Ap, const int64_t *Ai,
                                const double *Ax, double delta, int *status)
{
    oracle_chol *F = (oracle_chol *)calloc(1, sizeof(oracle_chol));
    if (!F) { *status = -1; return NULL; }
    F->n = n;
    F->perm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    F->iperm = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    F->parent = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!F->perm || !F->iperm || !F->parent) { *status = -1; oracle_chol_free(F); return NULL; }
    oracle_nd_order(n, Ap, Ai, F->perm);
    for (int64_t i = 0; i < n; i++) F->iperm[F->perm[i]] = i;
    build_C(F, Ap, Ai);
    etree(F);
    /* column counts from the row patterns */
    int64_t *mark = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *pat = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *cc = (int64_t *)calloc((size_t)n, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) { mark[i] = -1; cc[i] = 1; }
    for (int64_t k = 0; k < n; k++) {
        int64_t np = ereach(F, k, mark, pat);
        for (int64_t t = 0; t < np; t++) cc[pat[t]]++;
    }
    F->Lp = (int64_t *)malloc(((size_t)n + 1) * sizeof(int64_t));
    int64_t tot = 0;
    for (int64_t j = 0; j < n; j++) { F->Lp[j] = tot; tot += cc[j]; }
    F->Lp[n] = tot;
    free(mark); free(pat); free(cc);
    F->Li = (int64_t *)malloc((size_t)tot * sizeof(int64_t));
    F->Lx = (double *)malloc((size_t)tot * sizeof(double));
    if (!F->Li || !F->Lx) { *status = -1; oracle_chol_free(F); return NULL; }
    *status = numeric(F, Ax, delta);
    return F;
}

int oracle_chol_refactor(oracle_chol *F, const int64_t *Ap, const int64_t *Ai,
                         const double *Ax, double delta)
{
    (void)Ap; (void)Ai;
    return numeric(F, Ax, delta);
}

void oracle_chol_solve(const oracle_chol *F, const double *b, double *x)
{
    int64_t n = F->n;
    double *y = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; i++) y[i] = b[F->perm[i]];
    for (int64_t j = 0; j < n; j++) {
        y[j] /= F->Lx[F->Lp[j]];
        double yj = y[j];
        for (int64_t p = F->Lp[j] + 1; p < F->Lp[j + 1]; p++) y[F->Li[p]] -= F->Lx[p] * yj;
    }
    for (int64_t j = n - 1; j >= 0; j--) {
        double s = y[j];
        for (int64_t p = F->Lp[j] + 1; p < F->Lp[j + 1]; p++) s -= F->Lx[p] * y[F->Li[p]];
        y[j] = s / F->Lx[F->Lp[j]];
    }
    for (int64_t i = 0; i < n; i++) x[F->perm[i]] = y[i];
    free(y);
}

int64_t oracle_chol_nnzL(const oracle_chol *F) { return F->Lp[F->n]; }
double  oracle_chol_flops(const oracle_chol *F) { return F->flops; }
const int64_t *oracle_chol_perm(const oracle_chol *F) { return F->perm; }

void oracle_chol_free(oracle_chol *F)
{
    if (!F) return;
    free(F->perm); free(F->iperm); free(F->parent);
    free(F->Lp); free(F->Li); free(F->Lx);
    free(F->Cp); free(F->Ci); free(F->Cmap); free(F->Cx);
    free(F);
}

void oracle_spmv_upper(int64_t n, const int64_t *Ap, const int64_t *Ai,
                       const double *Ax, const double *x, double *y)
{
    for (int64_t i = 0; i < n; i++) y[i] = 0.0;
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            int64_t i = Ai[q];
            if (i > j) continue;
            y[i] += Ax[q] * x[j];
            if (i != j) y[j] += Ax[q] * x[i];
        }
}
