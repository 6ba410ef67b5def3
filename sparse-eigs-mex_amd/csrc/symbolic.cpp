// symbolic.cpp -- host ordering and supernodal symbolic analysis.
//
// Replaces cholmod_l_analyze with its AMD/METIS ordering (reference
// src/speigs_core.c:859-863, :933-936), which are absent here.
//
// Ordering: nested dissection.  A region is split by the middle level of a
// rooted level structure grown from a pseudo-peripheral node (any level of a
// breadth-first search separates the levels before it from those after it);
// the separator is then thinned -- a separator node with no neighbour on one
// side moves to the other side.  Regions of at most `leaf` nodes stop the
// recursion, and a disconnected region splits into its components.
//
// Supernodes: every separator and every leaf region is one supernode.  Its
// columns are contiguous in the new order (children first, separator last),
// and the rows of its front below the diagonal block are the boundary of its
// region: the later-ordered nodes adjacent to the region.  For a region D with
// separator S and child regions D_i,  B(D) = (N(S) u B(D_1) u ... ) minus D.
// Because a region is eliminated before its boundary, these are exactly the
// nonzero rows of the region's last column in L, and every child's rows are a
// subset of its parent's front -- the extend-add of the multifrontal method.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>

#include "factor.h"
#include "host_par.h"

namespace sehip {

namespace {

struct Graph {
    int64_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj;
};

// Adjacency of the stored strict upper triangle, both directions.  Row x
// lists its lower neighbours (column x's entries above the diagonal, in stored
// order), then its higher neighbours j > x in ascending j -- the order a serial
// pass over the columns produces; built in parallel, the second half of each
// row filled through atomic cursors and then sorted.
void build_graph(int64_t n, const int64_t *Ap, const int64_t *Ai, Graph &g, int nt)
{
    std::vector<int64_t> lo(n, 0), up(n, 0);
    parallel_for(n, nt, 8192, [&](int64_t j, int) {
        int64_t c = 0;
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i < j) {
                c++;
                __atomic_fetch_add(&up[i], (int64_t)1, __ATOMIC_RELAXED);
            }
        }
        lo[j] = c;
    });
    g.n = n;
    g.xadj.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) g.xadj[i + 1] = g.xadj[i] + lo[i] + up[i];
    g.adj.resize(std::max<int64_t>(g.xadj[n], 1));
    // cursors into the upper halves
    parallel_for(n, nt, 65536, [&](int64_t i, int) { up[i] = g.xadj[i] + lo[i]; });
    parallel_for(n, nt, 8192, [&](int64_t j, int) {
        int64_t p = g.xadj[j];
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i < j) {
                g.adj[p++] = (int32_t)i;
                g.adj[__atomic_fetch_add(&up[i], (int64_t)1, __ATOMIC_RELAXED)] = (int32_t)j;
            }
        }
    });
    parallel_for(n, nt, 8192, [&](int64_t i, int) {
        std::sort(g.adj.begin() + g.xadj[i] + lo[i], g.adj.begin() + g.xadj[i + 1]);
    });
}

// Region ids live in one array shared by the threads that dissect disjoint
// regions; a thread reads the id of a neighbour that may belong to another
// thread's region (only to see that it is not its own), so the accesses are
// relaxed atomics.
inline int32_t rid_ld(const int32_t *a, int64_t i) { return __atomic_load_n(a + i, __ATOMIC_RELAXED); }
inline void rid_st(int32_t *a, int64_t i, int32_t v) { __atomic_store_n(a + i, v, __ATOMIC_RELAXED); }

// dissection tree before numbering: a separator (or leaf region) and its children
struct NdNode {
    std::vector<int32_t> cols;
    std::vector<std::unique_ptr<NdNode>> kids;
};
typedef std::vector<std::unique_ptr<NdNode>> NdList;

class Dissection {
public:
    Dissection(const Graph &g, int64_t leaf, std::vector<int64_t> &perm, int threads)
        : g_(g), leaf_(leaf), perm_(perm), rid_(g.n, 1), lev_(g.n, 0), lev2_(g.n, 0), free_threads_(threads - 1) {}

    std::vector<int64_t> c0, c1;
    std::vector<std::vector<int64_t>> kids;

    void run()
    {
        std::vector<int32_t> all(g_.n);
        std::iota(all.begin(), all.end(), 0);
        NdList roots;
        split_components(all, 1, roots);
        // number the tree in postorder (children in order, then the node): the
        // columns of a node follow all of its descendants'
        for (auto &r : roots) number(*r);
    }

private:
    // region sizes below which a split is not handed to another thread
    static constexpr int64_t kSpawnMin = 16384;
    struct Scratch {
        std::vector<int32_t> q, qs;
    };

    const Graph &g_;
    int64_t leaf_;
    std::vector<int64_t> &perm_;
    std::vector<int32_t> rid_, lev_, lev2_;     // lev2_: a candidate's levels (pseudo-peripheral search)
    std::atomic<int32_t> next_rid_{2};
    std::atomic<int> free_threads_;
    int64_t next_pos_ = 0;

    int64_t deg(int32_t v) const { return g_.xadj[v + 1] - g_.xadj[v]; }

    static Scratch &scratch()
    {
        thread_local Scratch sc;
        return sc;
    }

    // BFS inside region r from s; fills q[0..cnt), lev, returns cnt and height
    int64_t bfs(int32_t r, int32_t s, int32_t &h, std::vector<int32_t> &q, int32_t *lev)
    {
        int32_t *rid = rid_.data(), *qq = q.data();
        const int64_t *xadj = g_.xadj.data();
        const int32_t *adj = g_.adj.data();
        const int32_t vis = -r;       // temporary mark
        int64_t head = 0, tail = 0;
        qq[tail++] = s;
        lev[s] = 0;
        rid_st(rid, s, vis);
        // the queue is in level order: the nodes being expanded are at level
        // cur until head reaches lend.  The memory a later expansion touches
        // is prefetched in three stages (its xadj entry, adjacency list and
        // neighbours' ids), which does not change the visiting order.
        int32_t cur = 0;
        int64_t lend = 1;
        constexpr int64_t kA = 16;
        while (head < tail) {
            if (head == lend) {
                cur++;
                lend = tail;
            }
            if (head + 3 * kA < tail) __builtin_prefetch(xadj + qq[head + 3 * kA]);
            if (head + 2 * kA < tail) __builtin_prefetch(adj + xadj[qq[head + 2 * kA]]);
            if (head + kA < tail) {
                const int32_t w = qq[head + kA];
                for (int64_t p = xadj[w]; p < xadj[w + 1]; p++) __builtin_prefetch(rid + adj[p], 1);
            }
            const int32_t v = qq[head++];
            const int32_t lv = cur + 1;
            for (int64_t p = xadj[v]; p < xadj[v + 1]; p++) {
                const int32_t u = adj[p];
                if (rid_ld(rid, u) == r) {
                    rid_st(rid, u, vis);
                    lev[u] = lv;
                    qq[tail++] = u;
                }
            }
        }
        h = lev[qq[tail - 1]];        // the last node queued is on the last level
        for (int64_t i = 0; i < tail; i++) rid_st(rid, qq[i], r);
        return tail;
    }

    void number(NdNode &nd)
    {
        std::vector<int64_t> ch;
        for (auto &k : nd.kids) {
            number(*k);
            ch.push_back((int64_t)c0.size() - 1);
        }
        c0.push_back(next_pos_);
        for (int32_t v : nd.cols) perm_[next_pos_++] = v;
        c1.push_back(next_pos_);
        kids.push_back(std::move(ch));
        NdList().swap(nd.kids);
        std::vector<int32_t>().swap(nd.cols);
    }

    static std::unique_ptr<NdNode> leaf_node(const int32_t *cols, int64_t nc)
    {
        std::unique_ptr<NdNode> nd(new NdNode());
        nd->cols.assign(cols, cols + nc);
        return nd;
    }

    // region r (nodes) may be disconnected: dissect each component
    void split_components(std::vector<int32_t> &nodes, int32_t r, NdList &roots)
    {
        if ((int64_t)nodes.size() <= leaf_) {
            roots.push_back(leaf_node(nodes.data(), (int64_t)nodes.size()));
            return;
        }
        std::vector<int32_t> &q = scratch().q;
        if (q.size() < nodes.size()) q.resize(nodes.size());
        int32_t h = 0;
        int64_t cnt = bfs(r, nodes[0], h, q, lev_.data());
        if (cnt == (int64_t)nodes.size()) {
            dissect(nodes, r, roots, h);    // q and lev_ hold the BFS from nodes[0]
            return;
        }
        // peel components off one at a time
        std::vector<std::vector<int32_t>> comps;
        std::vector<int32_t> rids;
        size_t start = 0;
        int32_t *rid = rid_.data();
        for (;;) {
            while (start < nodes.size() && rid_ld(rid, nodes[start]) != r) start++;
            if (start == nodes.size()) break;
            cnt = bfs(r, nodes[start], h, q, lev_.data());
            const int32_t nr = next_rid_.fetch_add(1);
            comps.emplace_back(q.begin(), q.begin() + cnt);
            for (int32_t v : comps.back()) rid_st(rid, v, nr);
            rids.push_back(nr);
        }
        std::vector<int32_t>().swap(nodes);
        // components are independent: the large ones are dissected on free
        // threads (each into its own slot), the rest by this thread; the slots
        // are collected in component order
        std::vector<NdList> slot(comps.size());
        auto one = [&](size_t c) {
            if ((int64_t)comps[c].size() <= leaf_)
                slot[c].push_back(leaf_node(comps[c].data(), (int64_t)comps[c].size()));
            else
                dissect(comps[c], rids[c], slot[c], -1);
            std::vector<int32_t>().swap(comps[c]);
        };
        // many components with enough work in total (a block-diagonal input):
        // all of them go to the pool; otherwise only the large ones
        int64_t total = 0;
        for (auto &cc : comps) total += (int64_t)cc.size();
        const bool many = comps.size() >= 16 && total >= 2 * kSpawnMin;
        std::vector<size_t> large, small;
        for (size_t c = 0; c < comps.size(); c++)
            (many || (int64_t)comps[c].size() >= kSpawnMin ? large : small).push_back(c);
        std::atomic<size_t> next{0};
        auto work = [&] {
            for (;;) {
                const size_t q = next.fetch_add(1);
                if (q >= large.size()) break;
                one(large[q]);
            }
        };
        std::vector<std::thread> th;
        for (size_t t = 1; t < large.size(); t++) {
            int f = free_threads_.load();
            while (f > 0 && !free_threads_.compare_exchange_weak(f, f - 1)) {}
            if (f <= 0) break;
            th.emplace_back([&] {
                work();
                free_threads_.fetch_add(1);
            });
        }
        work();
        for (size_t c : small) one(c);
        for (auto &t : th) t.join();
        for (auto &sl : slot)
            for (auto &r : sl) roots.push_back(std::move(r));
    }

    // connected region r whose nodes are listed in the order of a BFS from
    // nodes[0], lev_ holding the levels (height h): no component search, no
    // first BFS
    void split_ordered(std::vector<int32_t> &nodes, int32_t r, NdList &roots, int32_t h)
    {
        if ((int64_t)nodes.size() <= leaf_) {
            roots.push_back(leaf_node(nodes.data(), (int64_t)nodes.size()));
            return;
        }
        std::vector<int32_t> &q = scratch().q;
        if (q.size() < nodes.size()) q.resize(nodes.size());
        std::copy(nodes.begin(), nodes.end(), q.begin());
        dissect(nodes, r, roots, h);
    }

    // connected region r with nodes.size() > leaf; h0 >= 0: the thread's queue
    // and lev_ already hold the BFS from nodes[0], of height h0
    void dissect(std::vector<int32_t> &nodes, int32_t r, NdList &roots, int32_t h0)
    {
        Scratch &sc = scratch();
        const int64_t cnt = (int64_t)nodes.size();
        if ((int64_t)sc.q.size() < cnt) sc.q.resize(cnt);
        std::vector<int32_t> &q = sc.q;
        int32_t h = h0;
        int32_t s = nodes[0];
        if (h0 < 0) bfs(r, s, h, q, lev_.data());
        for (int it = 0; it < 3; it++) {         // pseudo-peripheral root
            int32_t best = -1;
            int64_t bd = 0;
            for (int64_t i = cnt - 1; i >= 0 && lev_[q[i]] == h; i--) {
                const int32_t v = q[i];
                if (best < 0 || deg(v) < bd) { best = v; bd = deg(v); }
            }
            // the candidate's level structure goes to a queue and level array
            // of its own, so the current one stays in place if it is no better
            if ((int64_t)sc.qs.size() < cnt) sc.qs.resize(cnt);
            int32_t h2 = 0;
            bfs(r, best, h2, sc.qs, lev2_.data());
            if (h2 > h) {
                s = best;
                h = h2;
                std::copy(sc.qs.begin(), sc.qs.begin() + cnt, q.begin());
                for (int64_t i = 0; i < cnt; i++) lev_[q[i]] = lev2_[q[i]];
            } else {
                break;
            }
        }
        if (h < 2) {
            roots.push_back(leaf_node(q.data(), cnt));
            return;
        }
        // choose the separator level among those leaving both sides >= 30%:
        // the smallest size x (1 + |A - B| / 2n), so that a slightly larger
        // level that halves the region wins over a smaller lopsided one (on
        // 3-D grids a corner-rooted structure has near-equal middle levels;
        // balanced splits there cut factor flops by ~10%)
        std::vector<int64_t> lsz(h + 1, 0);
        for (int64_t i = 0; i < cnt; i++) lsz[lev_[q[i]]]++;
        int32_t sep = -1, med = -1;
        int64_t cum = 0, best = 0;
        for (int32_t l = 0; l <= h; l++) {
            const int64_t before = cum;
            cum += lsz[l];
            if (med < 0 && 2 * cum >= cnt) med = l;
            if (l > 0 && l < h && 10 * before >= 3 * cnt && 10 * (cnt - cum) >= 3 * cnt) {
                // compare lsz (1 + |before - after| / 2 cnt) in integers
                const int64_t imb = before > cnt - cum ? before - (cnt - cum) : (cnt - cum) - before;
                const int64_t sc = lsz[l] * (2 * cnt + imb);
                if (sep < 0 || sc < best) { sep = l; best = sc; }
            }
        }
        if (sep < 0) sep = med;
        sep = std::min(std::max(sep, 1), h - 1);
        // thin the separator: no neighbour after it -> before it; then the reverse
        const int32_t *rid = rid_.data();
        std::vector<int32_t> S;
        for (int64_t i = 0; i < cnt; i++)
            if (lev_[q[i]] == sep) S.push_back(q[i]);
        for (int32_t v : S) {
            bool touches_b = false;
            for (int64_t p = g_.xadj[v]; p < g_.xadj[v + 1] && !touches_b; p++) {
                const int32_t u = g_.adj[p];
                touches_b = rid_ld(rid, u) == r && lev_[u] > sep;
            }
            if (!touches_b) lev_[v] = sep - 1;
        }
        for (int32_t v : S) {
            if (lev_[v] != sep) continue;
            bool touches_a = false;
            for (int64_t p = g_.xadj[v]; p < g_.xadj[v + 1] && !touches_a; p++) {
                const int32_t u = g_.adj[p];
                touches_a = rid_ld(rid, u) == r && lev_[u] < sep;
            }
            if (!touches_a) lev_[v] = sep + 1;
        }
        std::vector<int32_t> A, B;
        std::unique_ptr<NdNode> nd(new NdNode());
        A.reserve(cnt / 2 + 1);
        B.reserve(cnt / 2 + 1);
        const int32_t ra = next_rid_.fetch_add(2), rb = ra + 1;
        for (int64_t i = 0; i < cnt; i++) {
            const int32_t v = q[i], l = lev_[v];
            if (l < sep) A.push_back(v);
            else if (l > sep) B.push_back(v);
            else nd->cols.push_back(v);
        }
        for (int32_t v : A) rid_st(rid_.data(), v, ra);
        for (int32_t v : B) rid_st(rid_.data(), v, rb);
        // A (levels below sep, and the separator nodes moved to it) lists its
        // nodes in the order of a BFS from s inside A, at the same levels --
        // a shortest path from s to a node below sep stays below sep, and a
        // moved node is reached from its BFS parent at sep - 1 -- once the
        // moved nodes are back at level sep; A is connected
        int32_t ha = sep - 1;
        for (int32_t v : S)
            if (lev_[v] == sep - 1) {
                lev_[v] = sep;
                ha = sep;
            }
        std::vector<int32_t>().swap(nodes);
        NdList ka, kb;
        bool spawned = false;
        std::thread th;
        if ((int64_t)A.size() >= kSpawnMin && (int64_t)B.size() >= kSpawnMin) {
            int f = free_threads_.load();
            while (f > 0 && !free_threads_.compare_exchange_weak(f, f - 1)) {}
            if (f > 0) {
                th = std::thread([this, &A, &ka, ra, ha] {
                    split_ordered(A, ra, ka, ha);
                    free_threads_.fetch_add(1);
                });
                spawned = true;
            }
        }
        if (!spawned) split_ordered(A, ra, ka, ha);
        split_components(B, rb, kb);
        if (spawned) th.join();
        std::vector<int32_t>().swap(A);
        std::vector<int32_t>().swap(B);
        for (auto &k : ka) nd->kids.push_back(std::move(k));
        for (auto &k : kb) nd->kids.push_back(std::move(k));
        roots.push_back(std::move(nd));
    }
};

inline double sum_sq(double x) { return x * (x + 1.0) * (2.0 * x + 1.0) / 6.0; }

}  // namespace

int symbolic_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, const SymbolicOptions &o, Symbolic *S)
{
    if (n >= (int64_t)1 << 31) return SPEIGS_ERR_ARG;
    double t0 = wall();
    const int nt = host_threads();
    Graph g;
    build_graph(n, Ap, Ai, g, nt);
    S->n = n;
    S->perm.assign(n, 0);
    Dissection nd(g, std::max<int64_t>(o.leaf, 1), S->perm, nt);
    nd.run();
    S->iperm.resize(n);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { S->iperm[S->perm[i]] = i; });
    double t1 = wall();
    S->t_order = t1 - t0;

    const int64_t ns = (int64_t)nd.c0.size();
    S->ns = ns;
    S->c0 = nd.c0;
    S->c1 = nd.c1;
    S->parent.assign(ns, -1);
    S->chptr.assign(ns + 1, 0);
    for (int64_t s = 0; s < ns; s++) S->chptr[s + 1] = S->chptr[s] + (int64_t)nd.kids[s].size();
    S->chl.resize(S->chptr[ns]);
    for (int64_t s = 0; s < ns; s++)
        for (size_t i = 0; i < nd.kids[s].size(); i++) {
            S->chl[S->chptr[s] + i] = nd.kids[s][i];
            S->parent[nd.kids[s][i]] = s;
        }
    // levels (a node's level exceeds its children's)
    S->level.assign(ns, 0);
    S->nlevels = 0;
    for (int64_t s = 0; s < ns; s++) {
        int32_t lv = 0;
        for (int64_t p = S->chptr[s]; p < S->chptr[s + 1]; p++) lv = std::max(lv, S->level[S->chl[p]] + 1);
        S->level[s] = lv;
        S->nlevels = std::max(S->nlevels, lv + 1);
    }
    std::vector<std::vector<int64_t>> bylevel(S->nlevels);
    for (int64_t s = 0; s < ns; s++) bylevel[S->level[s]].push_back(s);
    // boundary rows, one level at a time (children's rows are complete before
    // their parent's level starts), nodes of a level in parallel
    std::vector<std::vector<int64_t>> rowset(ns);
    std::vector<std::vector<int32_t>> mark(nt);
    // the inverse permutation in 32 bits (half the cache footprint of the
    // lookups below)
    std::vector<int32_t> ip32(n);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { ip32[i] = (int32_t)S->iperm[i]; });
    for (int32_t L = 0; L < S->nlevels; L++) {
        const std::vector<int64_t> &nodes = bylevel[L];
        parallel_for((int64_t)nodes.size(), nt, 16, [&](int64_t i, int tid) {
            const int64_t s = nodes[i];
            std::vector<int32_t> &mk = mark[tid];
            if (mk.empty()) mk.assign(n, -1);
            std::vector<int64_t> &cand = rowset[s];
            const int64_t hi = S->c1[s];
            for (int64_t p = S->chptr[s]; p < S->chptr[s + 1]; p++)
                for (int64_t v : rowset[S->chl[p]])
                    if (v >= hi && mk[v] != (int32_t)s) { mk[v] = (int32_t)s; cand.push_back(v); }
            for (int64_t col = S->c0[s]; col < hi; col++) {
                const int32_t old = (int32_t)S->perm[col];
                for (int64_t p = g.xadj[old]; p < g.xadj[old + 1]; p++) {
                    const int64_t v = ip32[g.adj[p]];
                    if (v >= hi && mk[v] != (int32_t)s) { mk[v] = (int32_t)s; cand.push_back(v); }
                }
            }
            std::sort(cand.begin(), cand.end());
        });
    }
    S->rptr.assign(ns + 1, 0);
    for (int64_t s = 0; s < ns; s++) S->rptr[s + 1] = S->rptr[s] + (int64_t)rowset[s].size();
    S->rows.resize(S->rptr[ns]);
    parallel_for(ns, nt, 256, [&](int64_t s, int) {
        std::copy(rowset[s].begin(), rowset[s].end(), S->rows.begin() + S->rptr[s]);
        std::vector<int64_t>().swap(rowset[s]);      // freed in parallel too
    });
    std::vector<std::vector<int64_t>>().swap(rowset);
    // front sizes, offsets, counts
    S->F.assign(ns, 0);
    S->foff.assign(ns, 0);
    S->pool = 0;
    S->flops = 0.0;
    S->nnzL = 0;
    for (int64_t s = 0; s < ns; s++) {
        const int64_t c = S->c1[s] - S->c0[s], r = S->rptr[s + 1] - S->rptr[s], F = c + r;
        S->F[s] = F;
        S->foff[s] = S->pool;
        S->pool += F * F;
        S->nnzL += c * F - c * (c - 1) / 2;
        S->flops += sum_sq((double)F) - sum_sq((double)(F - c));
    }
    // relative positions of each supernode's rows in its parent's front
    S->relptr.assign(ns + 1, 0);
    for (int64_t s = 0; s < ns; s++) S->relptr[s + 1] = S->relptr[s] + (S->rptr[s + 1] - S->rptr[s]);
    S->rel.resize(std::max<int64_t>(S->relptr[ns], 1));
    S->rel[0] = 0;
    std::atomic<int> bad{0};
    parallel_for(ns, nt, 256, [&](int64_t s, int) {
        const int64_t p = S->parent[s];
        if (p < 0) {
            if (S->rptr[s + 1] != S->rptr[s]) bad.store(1);
            return;
        }
        const int64_t cp = S->c1[p] - S->c0[p];
        const int64_t *Rp = S->rows.data() + S->rptr[p], nRp = S->rptr[p + 1] - S->rptr[p];
        int64_t k = 0;
        for (int64_t q = S->rptr[s]; q < S->rptr[s + 1]; q++) {
            const int64_t v = S->rows[q];
            int64_t pos;
            if (v < S->c1[p]) {
                if (v < S->c0[p]) {
                    bad.store(1);
                    return;
                }
                pos = v - S->c0[p];
            } else {
                while (k < nRp && Rp[k] < v) k++;
                if (k >= nRp || Rp[k] != v) {
                    bad.store(1);
                    return;
                }
                pos = cp + k;
            }
            S->rel[S->relptr[s] + (q - S->rptr[s])] = (int32_t)pos;
        }
    });
    if (bad.load()) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "symbolic: a supernode row missing from its parent front");
    S->t_symbolic = wall() - t1;
    return SPEIGS_OK;
}

}  // namespace sehip
