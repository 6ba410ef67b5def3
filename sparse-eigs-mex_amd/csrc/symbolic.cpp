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
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>

#include "factor.h"
#include "host_par.h"

namespace sehip {

namespace {

// An array whose elements start uninitialized (each is written before it is
// read): std::vector's fill would be a serial pass over megabytes.
template <class T>
struct RawArray {
    std::unique_ptr<T[]> p;
    void alloc(int64_t m) { p.reset(new T[std::max<int64_t>(m, 1)]); }
    T &operator[](int64_t i) { return p[i]; }
    const T &operator[](int64_t i) const { return p[i]; }
    T *data() { return p.get(); }
    const T *data() const { return p.get(); }
};

struct Graph {
    int64_t n = 0;
    RawArray<int64_t> xadj;
    RawArray<int32_t> adj;
};

// Adjacency of the stored strict upper triangle, both directions.  Row x
// lists its lower neighbours (column x's entries above the diagonal, in stored
// order), then its higher neighbours j > x in ascending j -- the order a serial
// pass over the columns produces; built in parallel, the second half of each
// row filled through atomic cursors and then sorted.
void build_graph(int64_t n, const int64_t *Ap, const int64_t *Ai, Graph &g, int nt)
{
    RawArray<int64_t> lo, up;
    lo.alloc(n);
    up.alloc(n);
    parallel_for(n, nt, 65536, [&](int64_t i, int) { up[i] = 0; });
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
    g.xadj.alloc(n + 1);
    // row starts: an exclusive scan of the row lengths, by blocks
    const int nb = std::max(1, std::min(nt, (int)(n / 65536) + 1));
    std::vector<int64_t> bsum(nb + 1, 0);
    parallel_for(nb, nb, 1, [&](int64_t b, int) {
        int64_t t = 0;
        for (int64_t i = n * b / nb; i < n * (b + 1) / nb; i++) t += lo[i] + up[i];
        bsum[b + 1] = t;
    });
    for (int b = 0; b < nb; b++) bsum[b + 1] += bsum[b];
    parallel_for(nb, nb, 1, [&](int64_t b, int) {
        int64_t t = bsum[b];
        for (int64_t i = n * b / nb; i < n * (b + 1) / nb; i++) {
            g.xadj[i] = t;
            t += lo[i] + up[i];
        }
    });
    g.xadj[n] = bsum[nb];
    g.adj.alloc(g.xadj[n]);
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
        std::sort(g.adj.data() + g.xadj[i] + lo[i], g.adj.data() + g.xadj[i + 1]);
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
        : g_(g), leaf_(leaf), perm_(perm), rid_(g.n), lev_(g.n), lev2_(g.n), threads_(threads)
    {
        int32_t *rid = rid_.data();
        parallel_for(g.n, threads, 65536, [&](int64_t i, int) { rid[i] = 1; });
    }

    std::vector<int64_t> c0, c1;
    std::vector<std::vector<int64_t>> kids;

    void run()
    {
        std::vector<int32_t> all(g_.n);
        std::iota(all.begin(), all.end(), 0);
        NdList roots;
        // thread 0 dissects from the top; the others take forked regions
        auto work = [&](int t) {
            if (t > 0) {
                worker();
                return;
            }
            split_components(all, 1, roots);
            {
                std::lock_guard<std::mutex> lk(jm_);
                finished_ = true;
            }
            jcv_.notify_all();
        };
        if (threads_ <= 1) {
            work(0);
        } else if (!HostPool::get().run(threads_, work)) {
            std::vector<std::thread> th;
            for (int t = 1; t < threads_; t++) th.emplace_back(work, t);
            work(0);
            for (auto &x : th) x.join();
        }
        // number the tree in postorder (children in order, then the node): the
        // columns of a node follow all of its descendants'
        for (auto &r : roots) number(*r);
    }

private:
    // Fork-join of the recursion: a forked region waits in a shared list until
    // an idle thread takes it, or until its parent, done with its other side,
    // runs it itself; a parent whose forked side is running elsewhere helps
    // with other waiting regions meanwhile.  Regions smaller than kForkMin
    // are not forked.  Which thread dissects a region does not change the
    // result.
    static constexpr int64_t kForkMin = 8192;
    struct Job {
        std::function<void()> fn;
        int state = 0;               // 0 waiting, 1 running, 2 done (under jm_)
    };
    void fork(Job &j)
    {
        if (threads_ <= 1) return;   // joined (run) in place
        {
            std::lock_guard<std::mutex> lk(jm_);
            jobs_.push_back(&j);
        }
        jcv_.notify_one();
    }
    void run_job(Job &j, std::unique_lock<std::mutex> &lk)
    {
        j.state = 1;
        lk.unlock();
        j.fn();
        lk.lock();
        j.state = 2;
        jcv_.notify_all();
    }
    void join(Job &j)
    {
        std::unique_lock<std::mutex> lk(jm_);
        if (j.state == 0) {
            auto it = std::find(jobs_.begin(), jobs_.end(), &j);
            if (it != jobs_.end()) jobs_.erase(it);
            run_job(j, lk);
            return;
        }
        while (j.state != 2) {
            if (!jobs_.empty()) {
                Job *o = jobs_.back();
                jobs_.pop_back();
                run_job(*o, lk);
                continue;
            }
            jcv_.wait(lk);
        }
    }
    void worker()
    {
        std::unique_lock<std::mutex> lk(jm_);
        for (;;) {
            if (!jobs_.empty()) {
                Job *o = jobs_.front();      // the oldest: the largest regions
                jobs_.pop_front();
                run_job(*o, lk);
                continue;
            }
            if (finished_) return;
            jcv_.wait(lk);
        }
    }

    struct Scratch {
        std::vector<int32_t> q, qs;
        std::vector<int64_t> lo, los;    // level starts of q, of qs
    };

    const Graph &g_;
    int64_t leaf_;
    std::vector<int64_t> &perm_;
    std::vector<int32_t> rid_, lev_, lev2_;     // lev2_: a candidate's levels (pseudo-peripheral search)
    std::atomic<int32_t> next_rid_{2};
    const int threads_;
    std::mutex jm_;
    std::condition_variable jcv_;
    std::deque<Job *> jobs_;
    bool finished_ = false;
    int64_t next_pos_ = 0;

    int64_t deg(int32_t v) const { return g_.xadj[v + 1] - g_.xadj[v]; }

    static Scratch &scratch()
    {
        thread_local Scratch sc;
        return sc;
    }

    // BFS inside region r from s; fills q[0..cnt) in level order, lev, and
    // lofs[l] = the first position of level l (lofs[h + 1] = cnt); returns cnt
    // and the height h
    int64_t bfs(int32_t r, int32_t s, int32_t &h, std::vector<int32_t> &q, int32_t *lev, std::vector<int64_t> &lofs)
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
        lofs.clear();
        lofs.push_back(0);
        constexpr int64_t kA = 16;
        while (head < tail) {
            if (head == lend) {
                cur++;
                lend = tail;
                lofs.push_back(head);
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
        h = cur;                      // the level of the last node queued
        lofs.push_back(tail);
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
        Scratch &sc = scratch();
        std::vector<int32_t> &q = sc.q;
        if (q.size() < nodes.size()) q.resize(nodes.size());
        int32_t h = 0;
        int64_t cnt = bfs(r, nodes[0], h, q, lev_.data(), sc.lo);
        if (cnt == (int64_t)nodes.size()) {
            dissect(nodes, r, roots, h);    // q, lev_ and sc.lo hold the BFS from nodes[0]
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
            cnt = bfs(r, nodes[start], h, q, lev_.data(), sc.lo);
            const int32_t nr = next_rid_.fetch_add(1);
            comps.emplace_back(q.begin(), q.begin() + cnt);
            for (int32_t v : comps.back()) rid_st(rid, v, nr);
            rids.push_back(nr);
        }
        std::vector<int32_t>().swap(nodes);
        // components are independent: the large ones are forked, the rest
        // dissected here; the slots are collected in component order
        std::vector<NdList> slot(comps.size());
        auto one = [&](size_t c) {
            if ((int64_t)comps[c].size() <= leaf_)
                slot[c].push_back(leaf_node(comps[c].data(), (int64_t)comps[c].size()));
            else
                dissect(comps[c], rids[c], slot[c], -1);
            std::vector<int32_t>().swap(comps[c]);
        };
        // (which ones are forked is fixed first: a forked component's list is
        // released by whichever thread dissects it)
        std::vector<char> big(comps.size());
        for (size_t c = 0; c < comps.size(); c++) big[c] = (int64_t)comps[c].size() >= kForkMin;
        std::deque<Job> forked;
        for (size_t c = 0; c < comps.size(); c++)
            if (big[c]) {
                forked.emplace_back();
                forked.back().fn = [&one, c] { one(c); };
                fork(forked.back());
            }
        for (size_t c = 0; c < comps.size(); c++)
            if (!big[c]) one(c);
        for (auto &j : forked) join(j);
        for (auto &sl : slot)
            for (auto &r : sl) roots.push_back(std::move(r));
    }

    // connected region r whose nodes are listed in the order of a BFS from
    // nodes[0], lev_ holding the levels (height h, level starts lofs): no
    // component search, no first BFS
    void split_ordered(std::vector<int32_t> &nodes, int32_t r, NdList &roots, int32_t h, std::vector<int64_t> &lofs)
    {
        if ((int64_t)nodes.size() <= leaf_) {
            roots.push_back(leaf_node(nodes.data(), (int64_t)nodes.size()));
            return;
        }
        Scratch &sc = scratch();
        if (sc.q.size() < nodes.size()) sc.q.resize(nodes.size());
        std::copy(nodes.begin(), nodes.end(), sc.q.begin());
        sc.lo.swap(lofs);
        dissect(nodes, r, roots, h);
    }

    // connected region r with nodes.size() > leaf; h0 >= 0: the thread's queue,
    // level starts and lev_ already hold the BFS from nodes[0], of height h0
    void dissect(std::vector<int32_t> &nodes, int32_t r, NdList &roots, int32_t h0)
    {
        Scratch &sc = scratch();
        const int64_t cnt = (int64_t)nodes.size();
        if ((int64_t)sc.q.size() < cnt) sc.q.resize(cnt);
        std::vector<int32_t> &q = sc.q;
        std::vector<int64_t> &lo = sc.lo;
        int32_t h = h0;
        int32_t s = nodes[0];
        if (h0 < 0) bfs(r, s, h, q, lev_.data(), lo);
        for (int it = 0; it < 3; it++) {         // pseudo-peripheral root
            int32_t best = -1;
            int64_t bd = 0;
            for (int64_t i = cnt - 1; i >= lo[h]; i--) {     // the last level
                const int32_t v = q[i];
                if (best < 0 || deg(v) < bd) { best = v; bd = deg(v); }
            }
            // the candidate's level structure goes to a queue and level array
            // of its own, so the current one stays in place if it is no better
            if ((int64_t)sc.qs.size() < cnt) sc.qs.resize(cnt);
            int32_t h2 = 0;
            bfs(r, best, h2, sc.qs, lev2_.data(), sc.los);
            if (h2 > h) {
                s = best;
                h = h2;
                std::copy(sc.qs.begin(), sc.qs.begin() + cnt, q.begin());
                for (int64_t i = 0; i < cnt; i++) lev_[q[i]] = lev2_[q[i]];
                lo.swap(sc.los);
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
        for (int32_t l = 0; l <= h; l++) lsz[l] = lo[l + 1] - lo[l];
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
        const int64_t a0 = lo[sep], b0 = lo[sep + 1];     // the level structure: [A | S | B]
        std::vector<int32_t> S(q.begin() + a0, q.begin() + b0);
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
        // A = the levels before sep, then the separator nodes moved to it; B =
        // the nodes moved to it, then the levels after sep (the order of a scan
        // of the level structure).  A and B keep the region's id: they are not
        // adjacent, so a search inside one never reaches the other; only the
        // separator's nodes leave the region (id 0, no region's)
        std::vector<int32_t> A(q.begin(), q.begin() + a0), B;
        std::unique_ptr<NdNode> nd(new NdNode());
        B.reserve(cnt - b0 + S.size());
        for (int32_t v : S) {
            const int32_t l = lev_[v];
            if (l < sep) A.push_back(v);
            else if (l > sep) B.push_back(v);
            else nd->cols.push_back(v);
        }
        B.insert(B.end(), q.begin() + b0, q.begin() + cnt);
        for (int32_t v : nd->cols) rid_st(rid_.data(), v, 0);
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
        std::vector<int64_t> alo(lo.begin(), lo.begin() + sep + 1);    // A's level starts
        if (ha == sep) alo.push_back((int64_t)A.size());
        std::vector<int32_t>().swap(nodes);
        NdList ka, kb;
        Job ja;
        ja.fn = [this, &A, &ka, r, ha, &alo] { split_ordered(A, r, ka, ha, alo); };
        const bool forked = (int64_t)A.size() >= kForkMin && (int64_t)B.size() >= kForkMin;
        if (forked) fork(ja);
        else ja.fn();
        split_components(B, r, kb);
        if (forked) join(ja);
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
    double t0 = wall(), tp = t0;
    const int nt = host_threads();
    Graph g;
    build_graph(n, Ap, Ai, g, nt);
    trace_phase("ordering: graph", tp);
    S->n = n;
    S->perm.assign(n, 0);
    Dissection nd(g, std::max<int64_t>(o.leaf, 1), S->perm, nt);
    nd.run();
    trace_phase("ordering: dissection", tp);
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
        // a few large fronts at the top levels: one per task
        const int64_t grain = std::max<int64_t>(1, std::min<int64_t>(16, (int64_t)nodes.size() / (4 * nt)));
        parallel_for((int64_t)nodes.size(), nt, grain, [&](int64_t i, int tid) {
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
