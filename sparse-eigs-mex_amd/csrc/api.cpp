// api.cpp -- the public C ABI of libspeigs_hip.so: speigs(), speigs_z(),
// speigs_default_opts(), speigs_errmsg() (reference include/speigs.h:70-99,
// driver src/speigs_core.c:783-1024) and the speigs_hip_* extension.
//
// The driver keeps the reference's host-side structure analysis -- symmetry
// check, diagonal / dense short-circuits, connected components, banded gate,
// shift escalation -- and hands the Krylov loop a device operator:
//   'lm'        -> SELL-64 SpMV on the GPU            (path DIRECT)
//   'sm'        -> supernodal Cholesky on the GPU     (path SHIFTINV)
//   banded / LU -> host band Cholesky / dense LU solve behind the same seam
// There is no CPU fallback for the GPU paths: without a HIP device they fail
// with SPEIGS_ERR_INTERNAL.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dev.h"
#include "factor_impl.h"
#include "host_par.h"
#include "batch.h"
#include "mpnn.h"
#include "op_impl.h"

using namespace sehip;

namespace {

struct KeyIdx {
    double key;
    int64_t idx;
};
int key_asc(const void *a, const void *b)
{
    double x = ((const KeyIdx *)a)->key, y = ((const KeyIdx *)b)->key;
    return (x > y) - (x < y);
}
int key_desc(const void *a, const void *b) { return -key_asc(a, b); }

// Storage class of A (src/speigs_core.c:363-401): 1 full symmetric (or
// diagonal), 2 upper triangle only, 3 lower triangle only, 0 not symmetric.
int symmetry_kind(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax)
{
    const int nt = host_threads();
    // which triangles hold entries: a parallel scan (chunks stop once both
    // have been seen)
    std::atomic<int> seen{0};     // bit 0: above the diagonal, bit 1: below
    parallel_for(n, nt, 65536, [&](int64_t c, int) {
        if (seen.load(std::memory_order_relaxed) == 3) return;
        int f = 0;
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            if (Ai[q] < c) f |= 1;
            else if (Ai[q] > c) f |= 2;
        }
        if (f & ~seen.load(std::memory_order_relaxed)) seen.fetch_or(f, std::memory_order_relaxed);
    });
    const int tri = seen.load();
    if (tri == 1) return 2;
    if (tri == 2) return 3;
    if (tri == 0) return 1;
    // Both triangles stored: the reference requires the CSC arrays of A to
    // equal those of its transpose entry for entry (values to 1e-12 relative).
    // The same test without forming the transpose, every column in parallel:
    // a transpose lists each column's rows in non-decreasing order, and lists
    // the entries (c, i) of column i -- a run of equal rows when c repeats --
    // in the order the entries (i, c) stand in column c.  So every column must
    // be non-decreasing, and every run of row i in column c must meet a run of
    // row c of the same length in column i (found by binary search), with the
    // values pairwise equal.  Applied to all runs in both columns, the run
    // lengths also make the column counts equal the row counts.
    std::atomic<bool> sym{true};
    parallel_for(n, nt, 4096, [&](int64_t c, int) {
        if (!sym.load(std::memory_order_relaxed)) return;
        const int64_t e = Ap[c + 1];
        for (int64_t q = Ap[c]; q < e;) {
            const int64_t i = Ai[q];
            int64_t r = q + 1;
            while (r < e && Ai[r] == i) r++;
            if (r < e && Ai[r] < i) { sym.store(false); return; }
            const int64_t *col = Ai + Ap[i], *cend = Ai + Ap[i + 1];
            const int64_t *lo = std::lower_bound(col, cend, c), *hi = std::upper_bound(lo, cend, c);
            if (hi - lo != r - q) { sym.store(false); return; }
            const double *mirror = Ax + (lo - Ai);
            for (int64_t t = 0; t < r - q; t++) {
                const double a = Ax[q + t], b = mirror[t];
                const double mag = std::max(std::fabs(a), std::fabs(b));
                if (std::fabs(a - b) > 1e-12 * (mag > 0.0 ? mag : 1.0)) { sym.store(false); return; }
            }
            q = r;
        }
    });
    return sym.load() ? 1 : 0;
}

// Diagonal A (src/speigs_core.c:404-428): the eigenvalues are the diagonal
// entries (the last stored one of a column, zero where none is stored) and the
// eigenvectors unit vectors.  The k of smallest ('sm') or largest ('lm')
// magnitude are selected without sorting all n, equal magnitudes in index
// order (the order a stable sort of all n gives).
int diag_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int64_t k, speigs_which which,
              double *lambda, double *V, double *resid, speigs_info *info)
{
    std::vector<double> d(n, 0.0);
    parallel_for(n, host_threads(), 65536, [&](int64_t c, int) {
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++)
            if (Ai[q] == c) d[c] = Ax[q];
    });
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    const bool small = which == SPEIGS_SM;
    std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](int64_t a, int64_t b) {
        const double x = std::fabs(d[a]), y = std::fabs(d[b]);
        if (x != y) return small ? x < y : x > y;
        return a < b;
    });
    for (int64_t t = 0; t < k; t++) {
        const int64_t i = idx[t];
        lambda[t] = d[i];
        if (resid) resid[t] = 0.0;
        if (V) {
            double *col = V + t * n;
            std::fill(col, col + n, 0.0);
            col[i] = 1.0;
        }
    }
    info->nconv = k;
    info->flag = 0;
    info->path = SPEIGS_PATH_DIAG;
    return SPEIGS_OK;
}

int dense_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int64_t k, speigs_which which,
               double *lambda, double *V, double *resid, speigs_info *info)
{
    std::vector<double> D((size_t)n * n, 0.0);
    for (int64_t c = 0; c < n; c++)
        for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
            D[Ai[p] + c * n] = Ax[p];
            D[c + Ai[p] * n] = Ax[p];
        }
    // eigenvalues first, the k wanted by |lambda| (the reference sorts all of
    // them, src/speigs_core.c:336-350), then only those k vectors
    DenseTri T;
    if (dense_tri_values(n, D.data(), T)) return SPEIGS_ERR_LAPACK;
    std::vector<KeyIdx> ord(n);
    for (int64_t i = 0; i < n; i++) ord[i] = {std::fabs(T.w[i]), i};
    qsort(ord.data(), (size_t)n, sizeof(KeyIdx), which == SPEIGS_SM ? key_asc : key_desc);
    std::vector<int64_t> sel(k);
    for (int64_t t = 0; t < k; t++) sel[t] = ord[t].idx;
    if (V && dense_tri_vectors(T, k, sel.data(), V, n)) return SPEIGS_ERR_LAPACK;
    for (int64_t t = 0; t < k; t++) {
        lambda[t] = T.w[sel[t]];
        if (resid) resid[t] = 0.0;
    }
    info->nconv = k;
    info->flag = 0;
    info->path = SPEIGS_PATH_DENSE;
    return SPEIGS_OK;
}

// Connected components of the nonzero pattern (reference splits them at
// src/speigs_core.c:577-607), numbered by their lowest node.  Lock-free
// union-find over the columns in parallel: a union links the larger root
// under the smaller, so every root is its set's minimum node, and components
// numbered in the order of their roots are numbered by first appearance.
int64_t components(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, std::vector<int64_t> &comp)
{
    std::unique_ptr<std::atomic<int64_t>[]> p(new std::atomic<int64_t>[n]);
    const int nt = host_threads();
    parallel_for(n, nt, 65536, [&](int64_t i, int) { p[i].store(i, std::memory_order_relaxed); });
    auto find = [&](int64_t x) {
        for (;;) {
            int64_t px = p[x].load(std::memory_order_relaxed);
            if (px == x) return x;
            const int64_t ppx = p[px].load(std::memory_order_relaxed);
            if (ppx != px) p[x].compare_exchange_weak(px, ppx, std::memory_order_relaxed);
            x = ppx;
        }
    };
    parallel_for(n, nt, 4096, [&](int64_t c, int) {
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            if (Ax[q] == 0.0) continue;
            int64_t a = Ai[q], b = c;
            for (;;) {
                a = find(a);
                b = find(b);
                if (a == b) break;
                if (a < b) std::swap(a, b);   // link root a under the smaller root b
                int64_t expect = a;
                if (p[a].compare_exchange_strong(expect, b, std::memory_order_relaxed)) break;
            }
        }
    });
    comp.resize(n);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { comp[i] = find(i); });
    std::vector<int64_t> lab(n);
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) lab[i] = comp[i] == i ? nc++ : -1;
    parallel_for(n, nt, 65536, [&](int64_t i, int) { comp[i] = lab[comp[i]]; });
    return nc;
}

// RCM over the stored column lists (src/speigs_core.c:471-503).  Same order as
// the reference: an outer index `start` runs over 0..n-1; when it is unvisited
// a BFS is seeded at the lowest-index node of minimum degree among the
// unvisited nodes >= start, neighbours are appended in column order and
// insertion-sorted (stably) by degree, and the queue is reversed.  The
// reference finds each seed by a scan to n, O(n^2) on upper-triangle storage
// (column lists reach only lower-numbered neighbours, so nearly every node
// seeds a BFS); degree buckets with cursors that only move forward (past
// visited nodes and past indices < start) give the identical seeds in
// O(n + nnz).  A `start` the BFS never reaches is skipped, as in the reference;
// the reference then reverses a partly filled queue (undefined), here such
// nodes are appended in index order so the result stays a permutation.
void rcm(int64_t n, const int64_t *Ap, const int64_t *Ai, std::vector<int64_t> &perm)
{
    std::vector<int64_t> deg(n), q(n);
    std::vector<char> seen(n, 0);
    int64_t maxdeg = 0;
    for (int64_t c = 0; c < n; c++) {
        deg[c] = Ap[c + 1] - Ap[c];
        maxdeg = std::max(maxdeg, deg[c]);
    }
    std::vector<int64_t> bptr(maxdeg + 2, 0), bnode(n), cur(maxdeg + 1);
    for (int64_t c = 0; c < n; c++) bptr[deg[c] + 1]++;
    for (int64_t d = 0; d <= maxdeg; d++) bptr[d + 1] += bptr[d];
    for (int64_t d = 0; d <= maxdeg; d++) cur[d] = bptr[d];
    {
        std::vector<int64_t> fillp(bptr.begin(), bptr.end() - 1);
        for (int64_t c = 0; c < n; c++) bnode[fillp[deg[c]]++] = c;
    }
    int64_t dmin = 0, head = 0, tail = 0;
    for (int64_t start = 0; start < n; start++) {
        if (seen[start]) continue;
        int64_t s = -1;
        for (int64_t d = dmin; d <= maxdeg; d++) {
            while (cur[d] < bptr[d + 1] && (seen[bnode[cur[d]]] || bnode[cur[d]] < start)) cur[d]++;
            if (cur[d] < bptr[d + 1]) { s = bnode[cur[d]]; break; }
            if (d == dmin) dmin++;           // bucket exhausted for good
        }
        seen[s] = 1;
        q[tail++] = s;
        while (head < tail) {
            int64_t v = q[head++], b0 = tail;
            for (int64_t p = Ap[v]; p < Ap[v + 1]; p++) {
                int64_t u = Ai[p];
                if (!seen[u]) { seen[u] = 1; q[tail++] = u; }
            }
            for (int64_t a = b0 + 1; a < tail; a++) {
                int64_t key = q[a], dk = deg[key], b = a - 1;
                while (b >= b0 && deg[q[b]] > dk) { q[b + 1] = q[b]; b--; }
                q[b + 1] = key;
            }
        }
    }
    for (int64_t c = 0; c < n && tail < n; c++)
        if (!seen[c]) q[tail++] = c;
    perm.resize(n);
    for (int64_t i = 0; i < n; i++) perm[i] = q[n - 1 - i];
}

double norm1_upper(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax)
{
    // |A| column sums from the upper triangle: column j adds its entries to
    // cs[j] and, mirrored, to cs[i] of each row i < j.  Threads take column
    // ranges; a mirrored entry whose row lies in an earlier range is queued and
    // added after all ranges are done, in range order, so each cs[i] sums its
    // terms in the order of the serial loop and the result is bitwise the same.
    std::vector<double> cs(n, 0.0);
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), n / 65536));
    const int64_t per = (n + T - 1) / T;
    std::vector<std::vector<std::pair<int64_t, double>>> late(T);
    parallel_for(T, T, 1, [&](int64_t t, int) {
        const int64_t c0 = t * per, c1 = std::min(n, c0 + per);
        for (int64_t j = c0; j < c1; j++)
            for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
                const int64_t i = Ai[q];
                if (i > j) continue;
                cs[j] += std::fabs(Ax[q]);
                if (i == j) continue;
                if (i >= c0)
                    cs[i] += std::fabs(Ax[q]);
                else
                    late[t].push_back({i, std::fabs(Ax[q])});
            }
    });
    for (int t = 0; t < T; t++)
        for (const auto &e : late[t]) cs[e.first] += e.second;
    double m = 0.0;
    for (double v : cs) m = std::max(m, v);
    return m;
}

// ------------------------------------------------------------- operators
struct MultOp : DevOperator {
    SellA A;
    int apply(const double *x, double *y) override
    {
        spmv(A, x, y, stream);
        return 0;
    }
    int multA(const double *X, int64_t ldx, double *Y, int64_t ldy, int nv) override
    {
        spmv_multi(A, X, ldx, Y, ldy, nv, stream);
        return 0;
    }
};

// host-side solve behind the device seam (banded Cholesky or dense LU short-circuits)
struct HostSolveOp : MultOp {
    int kind = 0;                        // 0 banded, 1 LU
    std::vector<double> fac, bw;
    std::vector<int64_t> perm, piv;
    int64_t kd = 0;
    int apply(const double *x, double *y) override
    {
        std::vector<double> hx(n);
        SEHIP_CHECK(hipMemcpyAsync(hx.data(), x, n * 8, hipMemcpyDeviceToHost, stream));
        SEHIP_CHECK(hipStreamSynchronize(stream));
        if (kind == 0) {
            for (int64_t i = 0; i < n; i++) bw[i] = hx[perm[i]];
            band_solve_u(n, kd, fac.data(), kd + 1, bw.data());
            for (int64_t i = 0; i < n; i++) hx[perm[i]] = bw[i];
        } else {
            lu_solve(n, fac.data(), piv.data(), hx.data());
        }
        SEHIP_CHECK(hipMemcpyAsync(y, hx.data(), n * 8, hipMemcpyHostToDevice, stream));
        SEHIP_CHECK(hipStreamSynchronize(stream));
        return 0;
    }
};

// supernodal Cholesky on the device
struct FactorOp : MultOp {
    Factor *F = nullptr;
    ~FactorOp() override { factor_destroy(F); }
    int apply(const double *x, double *y) override { return factor_apply(F, x, y); }
    int check() override { return factor_check(F); }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

// A lower bound on the bandwidth of every ordering of the stored pattern: a
// stored entry (r, c) puts r within kd positions of c, so the nodes reachable
// from v in d steps along the column lists lie within d kd positions of v and
// kd >= (|ball(v, d)| - 1) / (2 d).  The BFS from v stops after about `budget`
// nodes.
int64_t band_lower_bound(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t v, int64_t budget)
{
    std::vector<uint8_t> seen(n, 0);
    std::vector<int64_t> q;
    q.reserve(budget + 64);
    q.push_back(v);
    seen[v] = 1;
    int64_t lb = 0, head = 0;
    for (int64_t d = 1; head < (int64_t)q.size() && (int64_t)q.size() < budget; d++) {
        const int64_t end = (int64_t)q.size();
        for (; head < end; head++) {
            const int64_t c = q[head];
            for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
                const int64_t r = Ai[p];
                if (r >= 0 && r < n && !seen[r]) {
                    seen[r] = 1;
                    q.push_back(r);
                }
            }
        }
        lb = std::max(lb, ((int64_t)q.size() - 1 + 2 * d - 1) / (2 * d));
    }
    return lb;
}

int try_banded(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, const speigs_opts *o, double anorm,
               HostSolveOp *op, speigs_info *info)
{
    int bmax = o->band_max ? o->band_max : 512;
    if (bmax <= 0) return 0;
    // the gate below needs kd <= bmax and (kd + 1) n <= 2e7; a lower bound on
    // kd above either limit settles it without the RCM ordering (the BFS starts
    // at the last column, which on upper-triangle storage reaches the most)
    const int64_t kcap = std::min<int64_t>(bmax, (int64_t)(2.0e7 / (double)n) - 1);
    if (n > 1 && band_lower_bound(n, Ap, Ai, n - 1, 1 << 14) > kcap) return 0;
    double tb = wall();
    std::vector<int64_t> pm, ip(n);
    rcm(n, Ap, Ai, pm);
    for (int64_t i = 0; i < n; i++) ip[pm[i]] = i;
    int64_t kd = 0;
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) kd = std::max<int64_t>(kd, std::llabs(ip[Ai[q]] - ip[c]));
    double cells = (double)(kd + 1) * (double)n;
    double occ = kd > 0 ? (double)Ap[n] / ((double)n * (double)kd) : 1.0;
    if (!(kd <= bmax && cells <= 2.0e7 && occ >= 0.05)) return 0;
    const int64_t ld = kd + 1;
    std::vector<double> AB((size_t)cells, 0.0);
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            int64_t r = ip[Ai[q]], cc = ip[c];
            int64_t lo = std::min(r, cc), hi = std::max(r, cc);
            AB[(kd + lo - hi) + (size_t)hi * ld] = Ax[q];
        }
    std::vector<double> Wk(AB);
    int bad = band_chol_u(n, kd, Wk.data(), ld);
    double dl = 0.0;
    if (bad) {
        // the reference's loop steps the shift once more after success
        // (src/speigs_core.c:908-912), so the reported shift is 10x the one used
        double d0 = anorm * (o->shift0 > 0 ? o->shift0 : 1e-6), dcap = anorm * 1e-3;
        if (d0 > dcap) dcap = d0;
        for (dl = d0; dl <= dcap * 1.0000001 && bad; dl *= 10.0) {
            Wk = AB;
            for (int64_t j = 0; j < n; j++) Wk[kd + (size_t)j * ld] += dl;
            bad = band_chol_u(n, kd, Wk.data(), ld);
        }
        if (bad) dl = 0.0;
    }
    if (bad) return 0;
    op->kind = 0;
    op->fac.swap(Wk);
    op->perm = pm;
    op->kd = kd;
    op->bw.assign(n, 0.0);
    info->shift = dl;
    info->path = SPEIGS_PATH_BANDED;
    info->t_factor = wall() - tb;
    return 1;
}

int blocks_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int64_t k, speigs_which which,
                const speigs_opts *opts, const speigs_hip_opts *ho, const std::vector<int64_t> &comp, int64_t ncomp,
                double *lambda, double *V, double *resid, speigs_info *info, speigs_hip_info *hinfo);

}  // namespace

extern "C" {

void speigs_default_opts(speigs_opts *o)
{
    std::memset(o, 0, sizeof *o);
    o->dense_max = -1;
    o->detect = 1;
    o->check_sym = 1;
}

void speigs_hip_default_opts(speigs_hip_opts *o)
{
    std::memset(o, 0, sizeof *o);
    o->precision = SPEIGS_HIP_FP64;
    o->start = SPEIGS_HIP_START_UNIFORM;
}

const char *speigs_errmsg(int c)
{
    switch (c) {
    case SPEIGS_OK: return "ok";
    case SPEIGS_ERR_ARG: return "invalid argument";
    case SPEIGS_ERR_MEM: return "out of memory";
    case SPEIGS_ERR_FACTOR: return "factorization failed";
    case SPEIGS_ERR_LAPACK: return "LAPACK failed";
    case SPEIGS_ERR_NOTSYM: return "matrix is not symmetric (this solver is for symmetric/Hermitian matrices only)";
    default: return "internal error";
    }
}

const char *speigs_hip_last_error(void) { return last_error_text(); }

int speigs_hip_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return c;
}

int speigs_ex(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, const double *Ax, speigs_int k,
              speigs_which which, const speigs_opts *opts, const speigs_hip_opts *hopts, double *lambda, double *V,
              double *resid, speigs_info *info, speigs_hip_info *hinfo)
{
    clear_error();
    // declared first, destroyed last: with SPEIGS_HIP_TRACE=1 it reports the
    // time the other locals' destructors take (device memory, streams)
    struct TraceRelease {
        double t = 0.0;
        ~TraceRelease()
        {
            if (t > 0.0) trace_phase("release (memory, streams)", t);
        }
    } trace_release;
    speigs_opts dflt;
    if (!opts) {
        speigs_default_opts(&dflt);
        opts = &dflt;
    }
    speigs_hip_opts hdflt;
    if (!hopts) {
        speigs_hip_default_opts(&hdflt);
        hopts = &hdflt;
    }
    speigs_info li;
    if (!info) info = &li;
    std::memset(info, 0, sizeof *info);
    speigs_hip_info lhi;
    if (!hinfo) hinfo = &lhi;
    std::memset(hinfo, 0, sizeof *hinfo);

    if (n <= 0 || !Ap || !Ai || !Ax || !lambda) return SPEIGS_ERR_ARG;
    if (k <= 0 || k > n) return SPEIGS_ERR_ARG;
    if (opts->check_sym != 0 && symmetry_kind(n, Ap, Ai, Ax) == 0) return SPEIGS_ERR_NOTSYM;

    const double t0 = wall();
    double tp = t0;
    const int dmax = opts->dense_max < 0 ? 800 : opts->dense_max;
    if (opts->detect) {
        bool diag = true;
        for (int64_t c = 0; c < n && diag; c++)
            for (int64_t p = Ap[c]; p < Ap[c + 1]; p++)
                if (Ai[p] != c) { diag = false; break; }
        if (diag) return diag_path(n, Ap, Ai, Ax, k, which, lambda, V, resid, info);
    }
    if (dmax > 0 && (n <= dmax || 4 * k > n)) {
        int rc = dense_path(n, Ap, Ai, Ax, k, which, lambda, V, resid, info);
        info->t_iter = wall() - t0;
        return rc;
    }
    if (opts->detect) {
        std::vector<int64_t> comp;
        int64_t nc = components(n, Ap, Ai, Ax, comp);
        if (nc > 1) {
            int rc = blocks_path(n, Ap, Ai, Ax, k, which, opts, hopts, comp, nc, lambda, V, resid, info, hinfo);
            info->t_iter = wall() - t0;
            return rc;
        }
    }
    trace_phase("components", tp);
    double anorm = norm1_upper(n, Ap, Ai, Ax);
    if (anorm <= 0) anorm = 1.0;
    info->anorm = anorm;

    // ---- the GPU part: needs a device, no CPU fallback
    if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
    if (hipSetDevice(hopts->device) != hipSuccess) return SPEIGS_ERR_ARG;
    StreamGuard sg;
    SEHIP_CHECK(hipStreamCreate(&sg.s));
    DevMem mem;
    // the SELL matrix of the SpMVs is built on the host and uploaded on a
    // stream of its own while the device factorizes (joined before use, and
    // on every early return before `mem` and the streams go away)
    SellA sell_pre;
    int sell_rc = 0;
    StreamGuard sell_sg;
    std::thread sell_th;
    struct Joiner {
        std::thread &t;
        ~Joiner()
        {
            if (t.joinable()) t.join();
        }
    } sell_join{sell_th};
    std::unique_ptr<MultOp> op;
    int rc = SPEIGS_OK;

    if (which == SPEIGS_LM) {
        op.reset(new MultOp());
        info->path = SPEIGS_PATH_DIRECT;
    } else {
        auto *hs = new HostSolveOp();
        trace_phase("device, stream, norm", tp);
        const bool banded = try_banded(n, Ap, Ai, Ax, opts, anorm, hs, info);
        trace_phase("banded gate", tp);
        if (banded) {
            op.reset(hs);
        } else {
            delete hs;
            auto *fo = new FactorOp();
            op.reset(fo);
            FactorOptions fopt;
            fopt.precision = hopts->precision;
            fopt.ir_tol = hopts->ir_tol > 0 ? hopts->ir_tol : 1e-13;
            fopt.ir_maxit = hopts->ir_maxit > 0 ? hopts->ir_maxit : 12;
            fopt.ordering = opts->ordering;
            fopt.use_graph = hopts->use_graph;
            // an apply error e makes the Krylov residual floor ~ e ||A||: the
            // FP64 solves by inverted blocks are refined to 1 % of the tolerance
            fopt.apply_tol = 1e-2 * (opts->tol > 0 ? opts->tol : 1e-14);
            double ta = wall();
            rc = factor_create(n, Ap, Ai, Ax, fopt, sg.s, &fo->F, hinfo);
            if (rc) return rc;
            info->t_analyze = wall() - ta;
            SEHIP_CHECK(hipStreamCreate(&sell_sg.s));
            sell_th = std::thread([&]() { sell_rc = sell_build(n, Ap, Ai, Ax, &sell_pre, mem, sell_sg.s); });
            // the block eigensolver's solve plan is built while the device factorizes
            if (hopts->precision == SPEIGS_HIP_FP32_BLOCK) rowsolve_prepare_async(fo->F);
            double tf = wall();
            double delta = 0.0;
            int st = factor_numeric(fo->F, 0.0);
            if (st < 0) return st;
            if (st == 1) {
                double d0 = anorm * (opts->shift0 > 0 ? opts->shift0 : 1e-6), dcap = anorm * 1e-3;
                if (d0 > dcap) dcap = d0;
                for (delta = d0; delta <= dcap * 1.0000001; delta *= 10.0) {
                    st = factor_numeric(fo->F, delta);
                    if (st < 0) return st;
                    if (st != 1) break;
                }
                if (st == 1) delta = 0.0;
            }
            bool indefinite = false;
            if (st == 0) {
                info->shift = delta;
                info->path = SPEIGS_PATH_SHIFTINV;
            } else {
                indefinite = true;
                // Indefinite: no shift up to the cap makes A + delta I positive
                // definite.  The reference factors A itself at sigma = 0 with an
                // LU (src/speigs_core.c:974-984); here the same supernodal
                // structure holds the signed factorization A = L S L^T
                // (S = diag(+-1), pivots below sqrt(eps) ||A||_1 replaced), and
                // every apply is refined against A to FP64 accuracy.
                st = factor_numeric_signed(fo->F, 1.4901161193847656e-08 * anorm);
                if (st < 0 && st != SPEIGS_ERR_FACTOR) return st;
            }
            if (st == 0 && indefinite) {
                info->shift = 0.0;
                info->path = SPEIGS_PATH_LU;
                hinfo->perturbed_pivots = (double)factor_perturbed(fo->F);
            } else if (st != 0) {
                // the signed factorization broke down (a pivot that is not a
                // number): a dense host LU for small n, else the reference's
                // factorization error
                factor_destroy(fo->F);
                fo->F = nullptr;
                if (n > 6000) return SPEIGS_ERR_FACTOR;
                auto *lu = new HostSolveOp();
                op.reset(lu);
                lu->kind = 1;
                lu->fac.assign((size_t)n * n, 0.0);
                lu->piv.assign(n, 0);
                for (int64_t c = 0; c < n; c++)
                    for (int64_t p = Ap[c]; p < Ap[c + 1]; p++) {
                        int64_t r = Ai[p];
                        if (r > c) continue;
                        lu->fac[r + c * n] = Ax[p];
                        lu->fac[c + r * n] = Ax[p];
                    }
                if (lu_factor(n, lu->fac.data(), lu->piv.data())) return SPEIGS_ERR_FACTOR;
                info->shift = 0.0;
                info->path = SPEIGS_PATH_LU;
            }
            info->t_factor = wall() - tf;
            hinfo->t_numeric = info->t_factor;
            tp = wall();
        }
    }
    op->n = n;
    op->stream = sg.s;
    double tu = wall();
    if (sell_th.joinable()) {
        sell_th.join();
        if (sell_rc) return sell_rc;
        op->A = sell_pre;
    } else if ((rc = sell_build(n, Ap, Ai, Ax, &op->A, mem, sg.s))) {
        return rc;
    }
    hinfo->t_upload += wall() - tu;
    trace_phase("SELL build + upload", tp);
    hinfo->bytes_spmv = spmv_bytes(op->A);
    if (auto *fo = dynamic_cast<FactorOp *>(op.get())) {
        if (fo->F) factor_attach_matrix(fo->F, &op->A);
    }

    double ti = wall();
    auto *fop = dynamic_cast<FactorOp *>(op.get());
    bool block = fop && fop->F && hopts->precision == SPEIGS_HIP_FP32_BLOCK && info->path == SPEIGS_PATH_SHIFTINV;
    if (block && lobpcg_block_size(n, k) == 0) {
        // k beyond one block (k > 24): Lanczos over the same FP32 factor, each
        // apply refined to FP64 (the FP32_IR path)
        block = false;
        fop->F->opt.precision = SPEIGS_HIP_FP32_IR;
    }
    // learned start: the network's block for the block eigensolver, the sum of
    // its first min(k, 32) columns as the Lanczos start vector
    float *xl = nullptr;
    speigs_hip_opts hl = *hopts;
    std::vector<double> lstart;
    if (hopts->start == SPEIGS_HIP_START_LEARNED && hopts->model) {
        if ((rc = mem.alloc(&xl, (size_t)n * MP_WIDTH))) return rc;
        const uint64_t seed = opts->seed ? opts->seed : 88172645463325252ULL;
        auto *mdl = const_cast<speigs_hip_model *>(hopts->model);
        if ((rc = mp_infer_sell(mdl, op->A, seed, xl))) return rc;
        if (!block) {
            std::vector<float> X((size_t)n * MP_WIDTH);
            SEHIP_CHECK(hipMemcpy(X.data(), xl, X.size() * sizeof(float), hipMemcpyDeviceToHost));
            const int nc = (int)std::min<int64_t>(k, MP_WIDTH);
            lstart.assign(n, 0.0);
            for (int64_t i = 0; i < n; i++)
                for (int c = 0; c < nc; c++) lstart[i] += X[(size_t)i * MP_WIDTH + c];
            double nr = 0.0;
            for (double v : lstart) nr += v * v;
            if (nr > 0) {
                hl.start = SPEIGS_HIP_START_VECTOR;
                hl.start_vec = lstart.data();
            }
        }
    }
    if (block)
        rc = lobpcg_dev(op->A, fop->F, sg.s, n, k, opts, &hl, anorm, lambda, V, resid, info, hinfo, xl);
    else
        rc = lanczos_dev(op.get(), n, k, opts, &hl, anorm, lambda, V, resid, info);
    info->t_iter = wall() - ti;
    tp = wall();
    if (auto *fo = dynamic_cast<FactorOp *>(op.get())) {
        if (fo->F) factor_stats(fo->F, hinfo);
    }
    if (rc == SPEIGS_OK) {
        std::vector<KeyIdx> ord(k);
        for (int64_t i = 0; i < k; i++) ord[i] = {std::fabs(lambda[i]), i};
        qsort(ord.data(), (size_t)k, sizeof(KeyIdx), which == SPEIGS_SM ? key_asc : key_desc);
        bool ident = true;
        for (int64_t t = 0; t < k; t++) ident = ident && ord[t].idx == t;
        if (!ident) {
            std::vector<double> lc(lambda, lambda + k), rc2(k), Vc;
            if (resid) rc2.assign(resid, resid + k);
            if (V) Vc.assign(V, V + (size_t)n * k);
            for (int64_t t = 0; t < k; t++) {
                int64_t i = ord[t].idx;
                lambda[t] = lc[i];
                if (resid) resid[t] = rc2[i];
                if (V) std::memcpy(V + t * n, Vc.data() + i * n, (size_t)n * sizeof(double));
            }
        }
    }
    trace_phase("final order", tp);
    if (trace_on()) {
        if (auto *fo = dynamic_cast<FactorOp *>(op.get())) {
            if (fo->F) {
                factor_block_free(fo->F);
                trace_phase("release solve state", tp);
                fo->F->mem.release();
                trace_phase("release factor memory", tp);
            }
        }
        op.reset();
        trace_phase("release operator", tp);
        mem.release();
        trace_phase("release workspace", tp);
        trace_release.t = wall();
    }
    return rc;
}

int speigs(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, const double *Ax, speigs_int k,
           speigs_which which, const speigs_opts *opts, double *lambda, double *V, double *resid, speigs_info *info)
{
    return speigs_ex(n, Ap, Ai, Ax, k, which, opts, nullptr, lambda, V, resid, info, nullptr);
}

// speigs_z (complex Hermitian through the real embedding, src/speigs_core.c:660-768): hermitian.hip

}  // extern "C"

namespace {

// Components of the pattern (reference src/speigs_core.c:532-632): every
// component is solved for min(k, n_c) pairs and the k best candidates over all
// components are returned.  Components that the reference's recursive call
// would send to its diagonal or dense short-circuits are solved on the host,
// spread over threads; all the others are solved together on the device: one
// block-diagonal operator (one factorization of all of them for 'sm') and
// lockstep Lanczos recurrences (batched.cpp).  SPEIGS_HIP_BATCH=0 solves the
// device components one recursive call at a time instead.
struct CompMatrix {
    std::vector<int64_t> Ap, Ai;
    std::vector<double> Ax;
};

void extract_comp(int64_t c, int64_t nc, const std::vector<int64_t> &gidx, int64_t off, const std::vector<int64_t> &comp,
                  const std::vector<int64_t> &loc, const int64_t *Ap, const int64_t *Ai, const double *Ax, CompMatrix &M)
{
    M.Ap.assign(nc + 1, 0);
    M.Ai.clear();
    M.Ax.clear();
    for (int64_t jj = 0; jj < nc; jj++) {
        const int64_t g = gidx[off + jj];
        M.Ap[jj] = (int64_t)M.Ai.size();
        for (int64_t q = Ap[g]; q < Ap[g + 1]; q++)
            if (comp[Ai[q]] == c) {
                M.Ai.push_back(loc[Ai[q]]);
                M.Ax.push_back(Ax[q]);
            }
    }
    M.Ap[nc] = (int64_t)M.Ai.size();
}

int blocks_path(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int64_t k, speigs_which which,
                const speigs_opts *opts, const speigs_hip_opts *ho, const std::vector<int64_t> &comp, int64_t ncomp,
                double *lambda, double *V, double *resid, speigs_info *info, speigs_hip_info *hinfo)
{
    double tp = wall();
    std::vector<int64_t> sz(ncomp, 0), off(ncomp + 1, 0), gidx(n), loc(n), fill(ncomp, 0);
    for (int64_t i = 0; i < n; i++) sz[comp[i]]++;
    for (int64_t c = 0; c < ncomp; c++) off[c + 1] = off[c] + sz[c];
    for (int64_t i = 0; i < n; i++) {
        int64_t c = comp[i], slot = off[c] + fill[c]++;
        gidx[slot] = i;
        loc[i] = slot - off[c];
    }
    const int dmax = opts->dense_max < 0 ? 800 : opts->dense_max;
    const char *be = getenv("SPEIGS_HIP_BATCH");
    const bool batch = !(be && atoi(be) == 0);
    // per component: its k, whether it is solved by the batch, its candidates
    std::vector<int64_t> kc(ncomp);
    std::vector<char> onhost(ncomp, 1);
    for (int64_t c = 0; c < ncomp; c++) {
        kc[c] = std::min(k, sz[c]);
        const bool shortcut = (dmax > 0 && (sz[c] <= dmax || 4 * kc[c] > sz[c])) || sz[c] == 1;
        onhost[c] = !batch || shortcut;
    }
    std::vector<std::vector<double>> cv(ncomp), cl(ncomp), cr(ncomp);
    std::vector<int> crc(ncomp, 0);
    std::vector<speigs_info> cinfo(ncomp);
    // Components the recursion would send to the dense short-circuit
    // (src/speigs_core.c:318-353) and that are not diagonal go to the device
    // in one batch (dense_batch.hip) on a helper thread of their own, beside
    // the other host components and the device batch; the tiny ones (below
    // kDenseDevMin rows, microseconds on the host) stay in the recursion.
    // SPEIGS_HIP_DENSE_DEVICE=0 keeps them all on the host.
    constexpr int64_t kDenseDevMin = 64;
    const char *dde = getenv("SPEIGS_HIP_DENSE_DEVICE");
    const bool dense_dev = batch && !(dde && atoi(dde) == 0);
    std::vector<int64_t> densec;
    std::vector<std::vector<double>> dmat;
    std::vector<DenseJob> djobs;
    if (dense_dev) {
        std::vector<int64_t> cand;
        for (int64_t c = 0; c < ncomp; c++)
            if (onhost[c] && dmax > 0 && (sz[c] <= dmax || 4 * kc[c] > sz[c]) && sz[c] >= kDenseDevMin &&
                sz[c] <= kDenseBatchMax)
                cand.push_back(c);
        std::vector<char> isdiag(cand.size(), 0);
        dmat.resize(cand.size());
        parallel_for((int64_t)cand.size(), host_threads(), 1, [&](int64_t i, int) {
            const int64_t c = cand[i], nc = sz[c];
            CompMatrix M;
            extract_comp(c, nc, gidx, off[c], comp, loc, Ap, Ai, Ax, M);
            bool diag = true;
            for (int64_t jj = 0; jj < nc && diag; jj++)
                for (int64_t q = M.Ap[jj]; q < M.Ap[jj + 1]; q++)
                    if (M.Ai[q] != jj) {
                        diag = false;
                        break;
                    }
            isdiag[i] = diag && opts->detect;
            if (isdiag[i]) return;
            std::vector<double> &D = dmat[i];
            D.assign((size_t)nc * nc, 0.0);
            for (int64_t jj = 0; jj < nc; jj++)
                for (int64_t q = M.Ap[jj]; q < M.Ap[jj + 1]; q++) {
                    D[M.Ai[q] + jj * nc] = M.Ax[q];
                    D[jj + M.Ai[q] * nc] = M.Ax[q];
                }
        });
        for (size_t i = 0; i < cand.size(); i++) {
            if (isdiag[i]) continue;
            const int64_t c = cand[i];
            densec.push_back(c);
            onhost[c] = 0;
            cl[c].assign(kc[c], 0.0);
            cr[c].assign(kc[c], 0.0);
            if (V) cv[c].assign((size_t)sz[c] * kc[c], 0.0);
            DenseJob J;
            J.n = sz[c];
            J.k = kc[c];
            J.A = dmat[i].data();
            J.largest = which == SPEIGS_LM;
            J.lambda = cl[c].data();
            J.V = V ? cv[c].data() : nullptr;
            djobs.push_back(std::move(J));
        }
    }
    int dense_rc = 0;
    std::string dense_msg;
    std::thread dense_th;
    struct DenseJoin {
        std::thread &t;
        ~DenseJoin()
        {
            if (t.joinable()) t.join();
        }
    } dense_join{dense_th};
    if (!djobs.empty()) {
        if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
        const int dev = ho->device;
        dense_th = std::thread([&, dev] {
            auto run = [&]() -> int {
                SEHIP_CHECK(hipSetDevice(dev));
                StreamGuard ds;
                SEHIP_CHECK(hipStreamCreate(&ds.s));
                return dense_batch(djobs, ds.s);
            };
            dense_rc = run();
            if (dense_rc) dense_msg = last_error_text();
        });
    }
    // host components (each a recursive call: diagonal, dense, or -- when
    // batching is off -- the full single-component device path), in parallel
    // when they stay on the host
    {
        std::vector<int64_t> hostc;
        for (int64_t c = 0; c < ncomp; c++)
            if (onhost[c] && std::find(densec.begin(), densec.end(), c) == densec.end()) hostc.push_back(c);
        const int nt = batch ? host_threads() : 1;
        std::vector<CompMatrix> mats(nt);
        parallel_for((int64_t)hostc.size(), nt, 1, [&](int64_t i, int tid) {
            const int64_t c = hostc[i], nc = sz[c];
            CompMatrix &M = mats[tid];
            extract_comp(c, nc, gidx, off[c], comp, loc, Ap, Ai, Ax, M);
            cl[c].assign(kc[c], 0.0);
            cr[c].assign(kc[c], 0.0);
            if (V) cv[c].assign((size_t)nc * kc[c], 0.0);
            crc[c] = speigs_ex(nc, M.Ap.data(), M.Ai.data(), M.Ax.data(), kc[c], which, opts, ho, cl[c].data(),
                               V ? cv[c].data() : nullptr, cr[c].data(), &cinfo[c], nullptr);
        });
        for (int64_t c : hostc) {
            if (crc[c]) return crc[c];
            info->nops += cinfo[c].nops;
            if (cinfo[c].flag) info->flag = 1;
        }
        trace_phase("host components", tp);
    }
    // device components: one block-diagonal matrix
    std::vector<int64_t> devc;
    for (int64_t c = 0; c < ncomp; c++)
        if (!onhost[c] && std::find(densec.begin(), densec.end(), c) == densec.end()) devc.push_back(c);
    if (!devc.empty()) {
        if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
        if (hipSetDevice(ho->device) != hipSuccess) return SPEIGS_ERR_ARG;
        std::vector<BatchComp> bc(devc.size());
        int64_t nb = 0, nnzb = 0;
        for (size_t i = 0; i < devc.size(); i++) {
            bc[i].off = nb;
            bc[i].n = sz[devc[i]];
            bc[i].k = kc[devc[i]];
            nb += bc[i].n;
        }
        std::vector<int64_t> bAp(nb + 1, 0), bAi;
        std::vector<double> bAx;
        {
            // components in parallel: extract, norm, then copy into place
            const int nt = host_threads();
            std::vector<CompMatrix> mats(devc.size());
            parallel_for((int64_t)devc.size(), nt, 1, [&](int64_t i, int) {
                const int64_t c = devc[i], nc = sz[c];
                extract_comp(c, nc, gidx, off[c], comp, loc, Ap, Ai, Ax, mats[i]);
                bc[i].anorm = norm1_upper(nc, mats[i].Ap.data(), mats[i].Ai.data(), mats[i].Ax.data());
                if (bc[i].anorm <= 0) bc[i].anorm = 1.0;
            });
            std::vector<int64_t> zoff(devc.size() + 1, 0);
            for (size_t i = 0; i < devc.size(); i++) zoff[i + 1] = zoff[i] + (int64_t)mats[i].Ai.size();
            nnzb = zoff[devc.size()];
            bAi.resize(std::max<int64_t>(nnzb, 1));
            bAx.resize(std::max<int64_t>(nnzb, 1));
            parallel_for((int64_t)devc.size(), nt, 1, [&](int64_t i, int) {
                const CompMatrix &M = mats[i];
                for (int64_t jj = 0; jj < bc[i].n; jj++) bAp[bc[i].off + jj] = zoff[i] + M.Ap[jj];
                for (size_t q = 0; q < M.Ai.size(); q++) {
                    bAi[zoff[i] + q] = M.Ai[q] + bc[i].off;
                    bAx[zoff[i] + q] = M.Ax[q];
                }
            });
            bAp[nb] = nnzb;
        }
        trace_phase("device block matrix", tp);
        StreamGuard sg;
        SEHIP_CHECK(hipStreamCreate(&sg.s));
        DevMem mem;
        std::unique_ptr<MultOp> op;
        int rc = SPEIGS_OK;
        bool fallback = false;
        if (which == SPEIGS_LM) {
            op.reset(new MultOp());
        } else {
            auto *fo = new FactorOp();
            op.reset(fo);
            FactorOptions fopt;
            // the block eigensolver is per matrix, not batched: the batched
            // components run lockstep Lanczos, where the FP64 factor is the
            // faster operator (an FP32 one refined to FP64 costs ~3 solves per
            // apply: 512 components, 1.50 s in FP64 vs 2.77 s), so the fast
            // path takes FP64 here; FP32_IR asked for explicitly stays FP32_IR
            fopt.precision = ho->precision == SPEIGS_HIP_FP32_BLOCK ? SPEIGS_HIP_FP64 : ho->precision;
            fopt.ir_tol = ho->ir_tol > 0 ? ho->ir_tol : 1e-13;
            fopt.ir_maxit = ho->ir_maxit > 0 ? ho->ir_maxit : 12;
            fopt.ordering = opts->ordering;
            fopt.apply_tol = 1e-2 * (opts->tol > 0 ? opts->tol : 1e-14);
            double ta = wall();
            if ((rc = factor_create(nb, bAp.data(), bAi.data(), bAx.data(), fopt, sg.s, &fo->F, hinfo))) return rc;
            info->t_analyze += wall() - ta;
            double tf = wall();
            const int st = factor_numeric(fo->F, 0.0);
            if (st < 0) return st;
            info->t_factor += wall() - tf;
            if (hinfo) hinfo->t_numeric = info->t_factor;
            // a component that is not positive definite needs its own shift
            // (or LU): solve the device components one call at a time
            fallback = st == 1;
        }
        if (!fallback) {
            op->n = nb;
            op->stream = sg.s;
            if ((rc = sell_build(nb, bAp.data(), bAi.data(), bAx.data(), &op->A, mem, sg.s))) return rc;
            if (auto *fo = dynamic_cast<FactorOp *>(op.get())) factor_attach_matrix(fo->F, &op->A);
            trace_phase("factor, SELL build", tp);
            double ti = wall();
            if ((rc = lanczos_batched(op.get(), nb, bc, opts, V != nullptr))) return rc;
            info->t_iter += wall() - ti;
            trace_phase("batched Lanczos", tp);
            if (auto *fo = dynamic_cast<FactorOp *>(op.get())) factor_stats(fo->F, hinfo);
            for (size_t i = 0; i < devc.size(); i++) {
                const int64_t c = devc[i];
                cl[c] = bc[i].lambda;
                cr[c] = bc[i].resid;
                if (V) cv[c] = std::move(bc[i].V);
                info->nops += bc[i].nops;
                if (bc[i].flag) info->flag = 1;
            }
        } else {
            op.reset();
            CompMatrix M;
            for (int64_t c : devc) {
                const int64_t nc = sz[c];
                extract_comp(c, nc, gidx, off[c], comp, loc, Ap, Ai, Ax, M);
                cl[c].assign(kc[c], 0.0);
                cr[c].assign(kc[c], 0.0);
                if (V) cv[c].assign((size_t)nc * kc[c], 0.0);
                speigs_info si;
                const int r2 = speigs_ex(nc, M.Ap.data(), M.Ai.data(), M.Ax.data(), kc[c], which, opts, ho,
                                         cl[c].data(), V ? cv[c].data() : nullptr, cr[c].data(), &si, nullptr);
                if (r2) return r2;
                info->nops += si.nops;
                if (si.flag) info->flag = 1;
            }
        }
    }
    if (dense_th.joinable()) {
        dense_th.join();
        if (dense_rc) {
            note_error(dense_msg.c_str());
            return dense_rc;
        }
        trace_phase("dense components (device batch)", tp);
    }
    // candidates in component order, then the k best by |lambda|
    std::vector<double> clam, cres;
    std::vector<int64_t> ccmp, cslt;
    for (int64_t c = 0; c < ncomp; c++)
        for (size_t t = 0; t < cl[c].size(); t++) {
            clam.push_back(cl[c][t]);
            cres.push_back(cr[c][t]);
            ccmp.push_back(c);
            cslt.push_back((int64_t)t);
        }
    const int64_t ncand = (int64_t)clam.size();
    std::vector<KeyIdx> ord(ncand);
    for (int64_t i = 0; i < ncand; i++) ord[i] = {std::fabs(clam[i]), i};
    qsort(ord.data(), (size_t)ncand, sizeof(KeyIdx), which == SPEIGS_SM ? key_asc : key_desc);
    for (int64_t t = 0; t < k && t < ncand; t++) {
        int64_t i = ord[t].idx, c = ccmp[i], sl = cslt[i], nc = sz[c];
        lambda[t] = clam[i];
        if (resid) resid[t] = cres[i];
        if (V) {
            std::memset(V + t * n, 0, (size_t)n * sizeof(double));
            const double *src = cv[c].data() + sl * nc;
            for (int64_t jj = 0; jj < nc; jj++) V[t * n + gidx[off[c] + jj]] = src[jj];
        }
    }
    info->nconv = std::min(k, ncand);
    info->ncomp = ncomp;
    info->path = SPEIGS_PATH_BLOCKS;
    return SPEIGS_OK;
}

}  // namespace

// ---------------------------------------------------------------- operator ABI
// (struct speigs_hip_op: op_impl.h)

extern "C" {

int speigs_hip_op_create(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, const double *Ax, int mode,
                         int precision, double delta, speigs_hip_op **out, speigs_hip_info *hinfo)
{
    clear_error();
    if (!out || n <= 0 || !Ap || !Ai || !Ax) return SPEIGS_ERR_ARG;
    *out = nullptr;
    if (mode != SPEIGS_HIP_OP_MULT && mode != SPEIGS_HIP_OP_SOLVE) return SPEIGS_ERR_ARG;
    if (speigs_hip_device_count() <= 0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "no HIP device visible");
    speigs_hip_info lhi;
    if (!hinfo) hinfo = &lhi;
    std::memset(hinfo, 0, sizeof *hinfo);
    std::unique_ptr<speigs_hip_op> op(new speigs_hip_op());
    op->mode = mode;
    op->n = n;
    op->precision = precision;
    SEHIP_CHECK(hipGetDevice(&op->dev));
    SEHIP_CHECK(hipStreamCreate(&op->stream));
    int rc = sell_build(n, Ap, Ai, Ax, &op->A, op->mem, op->stream);
    if (rc) return rc;
    hinfo->bytes_spmv = spmv_bytes(op->A);
    if (mode == SPEIGS_HIP_OP_SOLVE) {
        FactorOptions fo;
        fo.precision = precision;
        rc = factor_create(n, Ap, Ai, Ax, fo, op->stream, &op->F, hinfo);
        if (rc) return rc;
        factor_attach_matrix(op->F, &op->A);
        int st = factor_numeric(op->F, delta);
        if (st < 0) return st;
        if (st == 1) return SPEIGS_ERR_FACTOR;
        if (precision == SPEIGS_HIP_FP32_BLOCK && (rc = factor_invert_l11(op->F))) return rc;
        factor_stats(op->F, hinfo);
    }
    *out = op.release();
    return SPEIGS_OK;
}

int speigs_hip_op_refactor(speigs_hip_op *op, double delta)
{
    clear_error();
    if (!op || !op->F) return SPEIGS_ERR_ARG;
    int st = factor_numeric(op->F, delta);
    if (st < 0) return st;
    if (st == 1) return SPEIGS_ERR_FACTOR;
    if (op->precision == SPEIGS_HIP_FP32_BLOCK) return factor_invert_l11(op->F);
    return SPEIGS_OK;
}

static int op_buffers(speigs_hip_op *op, int64_t nvec)
{
    if (op->cap >= nvec) return 0;
    int64_t ld = op->n;
    int rc;
    if ((rc = op->mem.alloc(&op->x, (size_t)ld * nvec))) return rc;
    if ((rc = op->mem.alloc(&op->y, (size_t)ld * nvec))) return rc;
    op->cap = nvec;
    return 0;
}

int speigs_hip_op_multA(speigs_hip_op *op, const double *x, double *y, speigs_int nvec)
{
    clear_error();
    if (!op || !x || !y || nvec <= 0) return SPEIGS_ERR_ARG;
    int rc = op_buffers(op, nvec);
    if (rc) return rc;
    const size_t bytes = (size_t)op->n * nvec * sizeof(double);
    SEHIP_CHECK(hipMemcpyAsync(op->x, x, bytes, hipMemcpyHostToDevice, op->stream));
    spmv_multi(op->A, op->x, op->n, op->y, op->n, (int)nvec, op->stream);
    SEHIP_CHECK(hipMemcpyAsync(y, op->y, bytes, hipMemcpyDeviceToHost, op->stream));
    SEHIP_CHECK(hipStreamSynchronize(op->stream));
    return SPEIGS_OK;
}

int speigs_hip_op_apply(speigs_hip_op *op, const double *x, double *y, speigs_int nvec)
{
    clear_error();
    if (!op) return SPEIGS_ERR_ARG;
    if (op->mode == SPEIGS_HIP_OP_MULT) return speigs_hip_op_multA(op, x, y, nvec);
    if (!x || !y || nvec <= 0) return SPEIGS_ERR_ARG;
    int rc = op_buffers(op, nvec);
    if (rc) return rc;
    const size_t bytes = (size_t)op->n * nvec * sizeof(double);
    SEHIP_CHECK(hipMemcpyAsync(op->x, x, bytes, hipMemcpyHostToDevice, op->stream));
    if (op->precision == SPEIGS_HIP_FP32_BLOCK) {
        // one unrefined FP32 solve per block of <= 32 columns (the block
        // eigensolver's preconditioner); SPEIGS_HIP_BLOCK_GENERIC=1 selects the
        // generic FMA tile kernels with FP64 vectors instead of the MFMA ones
        const char *g = getenv("SPEIGS_HIP_BLOCK_GENERIC");
        const bool generic = g && atoi(g) == 1;
        for (int64_t v0 = 0; v0 < nvec; v0 += 32) {
            const int nb = (int)std::min<int64_t>(32, nvec - v0);
            double *xb = op->x + v0 * op->n, *yb = op->y + v0 * op->n;
            if (generic) {
                if ((rc = factor_block_solve_d(op->F, xb, yb, nb))) return rc;
            } else {
                if ((rc = factor_block_solve_rows(op->F, xb, 1, op->n, yb, 1, op->n, nb))) return rc;
            }
        }
    } else {
        for (int64_t v = 0; v < nvec; v++)
            if ((rc = factor_apply(op->F, op->x + v * op->n, op->y + v * op->n))) return rc;
    }
    SEHIP_CHECK(hipMemcpyAsync(y, op->y, bytes, hipMemcpyDeviceToHost, op->stream));
    SEHIP_CHECK(hipStreamSynchronize(op->stream));
    return factor_check(op->F);
}

void speigs_hip_op_destroy(speigs_hip_op *op) { delete op; }

}  // extern "C"

extern "C" int speigs_hip_analyze(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, speigs_int *perm,
                                  speigs_int *snode, speigs_hip_info *hinfo)
{
    if (n <= 0 || !Ap || !Ai) return SPEIGS_ERR_ARG;
    Symbolic S;
    SymbolicOptions so;
    int rc = symbolic_analyze(n, Ap, Ai, so, &S);
    if (rc) return rc;
    if (perm) std::memcpy(perm, S.perm.data(), (size_t)n * sizeof(int64_t));
    if (snode)
        for (int64_t s = 0; s < S.ns; s++)
            for (int64_t j = S.c0[s]; j < S.c1[s]; j++) snode[j] = s;
    if (hinfo) {
        std::memset(hinfo, 0, sizeof *hinfo);
        hinfo->nnz_L = S.nnzL;
        hinfo->factor_flops = S.flops;
        hinfo->nsuper = S.ns;
        hinfo->nlevels = S.nlevels;
        hinfo->t_order = S.t_order;
        hinfo->t_symbolic = S.t_symbolic;
    }
    return SPEIGS_OK;
}

extern "C" int speigs_hip_rcm(speigs_int n, const speigs_int *Ap, const speigs_int *Ai, speigs_int *perm)
{
    if (n <= 0 || !Ap || !Ai || !perm) return SPEIGS_ERR_ARG;
    std::vector<int64_t> pm;
    rcm(n, Ap, Ai, pm);
    std::memcpy(perm, pm.data(), (size_t)n * sizeof(int64_t));
    return SPEIGS_OK;
}
