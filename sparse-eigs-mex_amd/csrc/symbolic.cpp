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
#include <cstring>
#include <numeric>

#include "factor.h"

namespace sehip {

namespace {

struct Graph {
    int64_t n = 0;
    std::vector<int64_t> xadj;
    std::vector<int32_t> adj;
};

void build_graph(int64_t n, const int64_t *Ap, const int64_t *Ai, Graph &g)
{
    std::vector<int64_t> deg(n + 1, 0);
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i < j) { deg[i]++; deg[j]++; }
        }
    g.n = n;
    g.xadj.assign(n + 1, 0);
    for (int64_t i = 0; i < n; i++) g.xadj[i + 1] = g.xadj[i] + deg[i];
    g.adj.resize(std::max<int64_t>(g.xadj[n], 1));
    for (int64_t i = 0; i < n; i++) deg[i] = g.xadj[i];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (i < j) {
                g.adj[deg[i]++] = (int32_t)j;
                g.adj[deg[j]++] = (int32_t)i;
            }
        }
}

class Dissection {
public:
    Dissection(const Graph &g, int64_t leaf, std::vector<int64_t> &perm)
        : g_(g), leaf_(leaf), perm_(perm), rid_(g.n, 1), lev_(g.n, 0), q_(g.n) {}

    std::vector<int64_t> c0, c1;
    std::vector<std::vector<int64_t>> kids;

    void run()
    {
        std::vector<int32_t> all(g_.n);
        std::iota(all.begin(), all.end(), 0);
        std::vector<int64_t> roots;
        split_components(all, 1, roots);
    }

private:
    const Graph &g_;
    int64_t leaf_;
    std::vector<int64_t> &perm_;
    std::vector<int32_t> rid_, lev_, q_;
    int64_t next_pos_ = 0;
    int32_t next_rid_ = 2;

    int64_t deg(int32_t v) const { return g_.xadj[v + 1] - g_.xadj[v]; }

    // BFS inside region r from s; fills q_[0..cnt), lev_, returns cnt and height
    int64_t bfs(int32_t r, int32_t s, int32_t &h)
    {
        int64_t head = 0, tail = 0;
        const int32_t vis = -r;       // temporary mark
        q_[tail++] = s;
        lev_[s] = 0;
        rid_[s] = vis;
        h = 0;
        while (head < tail) {
            const int32_t v = q_[head++];
            const int32_t lv = lev_[v] + 1;
            for (int64_t p = g_.xadj[v]; p < g_.xadj[v + 1]; p++) {
                const int32_t u = g_.adj[p];
                if (rid_[u] == r) {
                    rid_[u] = vis;
                    lev_[u] = lv;
                    if (lv > h) h = lv;
                    q_[tail++] = u;
                }
            }
        }
        for (int64_t i = 0; i < tail; i++) rid_[q_[i]] = r;
        return tail;
    }

    int64_t new_node(const int32_t *cols, int64_t nc, std::vector<int64_t> &&children)
    {
        const int64_t id = (int64_t)c0.size();
        c0.push_back(next_pos_);
        for (int64_t i = 0; i < nc; i++) {
            perm_[next_pos_++] = cols[i];
            rid_[cols[i]] = 0;
        }
        c1.push_back(next_pos_);
        kids.push_back(std::move(children));
        return id;
    }

    // region r (nodes) may be disconnected: dissect each component
    void split_components(std::vector<int32_t> &nodes, int32_t r, std::vector<int64_t> &roots)
    {
        if ((int64_t)nodes.size() <= leaf_) {
            roots.push_back(new_node(nodes.data(), (int64_t)nodes.size(), {}));
            return;
        }
        int32_t h = 0;
        int64_t cnt = bfs(r, nodes[0], h);
        if (cnt == (int64_t)nodes.size()) {
            dissect(nodes, r, roots);
            return;
        }
        // peel components off one at a time
        std::vector<std::vector<int32_t>> comps;
        std::vector<int32_t> rids;
        size_t start = 0;
        for (;;) {
            while (start < nodes.size() && rid_[nodes[start]] != r) start++;
            if (start == nodes.size()) break;
            cnt = bfs(r, nodes[start], h);
            const int32_t nr = next_rid_++;
            comps.emplace_back(q_.begin(), q_.begin() + cnt);
            for (int32_t v : comps.back()) rid_[v] = nr;
            rids.push_back(nr);
        }
        nodes.clear();
        nodes.shrink_to_fit();
        for (size_t c = 0; c < comps.size(); c++) {
            if ((int64_t)comps[c].size() <= leaf_)
                roots.push_back(new_node(comps[c].data(), (int64_t)comps[c].size(), {}));
            else
                dissect(comps[c], rids[c], roots);
            std::vector<int32_t>().swap(comps[c]);
        }
    }

    // connected region r with nodes.size() > leaf
    void dissect(std::vector<int32_t> &nodes, int32_t r, std::vector<int64_t> &roots)
    {
        const int64_t cnt = (int64_t)nodes.size();
        int32_t h = 0;
        int32_t s = nodes[0];
        bfs(r, s, h);
        for (int it = 0; it < 3; it++) {         // pseudo-peripheral root
            int32_t best = -1;
            int64_t bd = 0;
            for (int64_t i = cnt - 1; i >= 0 && lev_[q_[i]] == h; i--) {
                const int32_t v = q_[i];
                if (best < 0 || deg(v) < bd) { best = v; bd = deg(v); }
            }
            int32_t h2 = 0;
            bfs(r, best, h2);
            if (h2 > h) {
                s = best;
                h = h2;
            } else {
                bfs(r, s, h);
                break;
            }
        }
        if (h < 2) {
            roots.push_back(new_node(q_.data(), cnt, {}));
            return;
        }
        // choose the separator level: smallest level with both sides >= 30%
        std::vector<int64_t> lsz(h + 1, 0);
        for (int64_t i = 0; i < cnt; i++) lsz[lev_[q_[i]]]++;
        int32_t sep = -1, med = -1;
        int64_t cum = 0, best = 0;
        for (int32_t l = 0; l <= h; l++) {
            const int64_t before = cum;
            cum += lsz[l];
            if (med < 0 && 2 * cum >= cnt) med = l;
            if (l > 0 && l < h && 10 * before >= 3 * cnt && 10 * (cnt - cum) >= 3 * cnt)
                if (sep < 0 || lsz[l] < best) { sep = l; best = lsz[l]; }
        }
        if (sep < 0) sep = med;
        sep = std::min(std::max(sep, 1), h - 1);
        // thin the separator: no neighbour after it -> before it; then the reverse
        std::vector<int32_t> S;
        for (int64_t i = 0; i < cnt; i++)
            if (lev_[q_[i]] == sep) S.push_back(q_[i]);
        for (int32_t v : S) {
            bool touches_b = false;
            for (int64_t p = g_.xadj[v]; p < g_.xadj[v + 1] && !touches_b; p++) {
                const int32_t u = g_.adj[p];
                touches_b = rid_[u] == r && lev_[u] > sep;
            }
            if (!touches_b) lev_[v] = sep - 1;
        }
        for (int32_t v : S) {
            if (lev_[v] != sep) continue;
            bool touches_a = false;
            for (int64_t p = g_.xadj[v]; p < g_.xadj[v + 1] && !touches_a; p++) {
                const int32_t u = g_.adj[p];
                touches_a = rid_[u] == r && lev_[u] < sep;
            }
            if (!touches_a) lev_[v] = sep + 1;
        }
        std::vector<int32_t> A, B, Sep;
        A.reserve(cnt / 2 + 1);
        B.reserve(cnt / 2 + 1);
        const int32_t ra = next_rid_++, rb = next_rid_++;
        for (int64_t i = 0; i < cnt; i++) {
            const int32_t v = q_[i], l = lev_[v];
            if (l < sep) A.push_back(v);
            else if (l > sep) B.push_back(v);
            else Sep.push_back(v);
        }
        for (int32_t v : A) rid_[v] = ra;
        for (int32_t v : B) rid_[v] = rb;
        nodes.clear();
        nodes.shrink_to_fit();
        std::vector<int64_t> children;
        split_components(A, ra, children);
        std::vector<int32_t>().swap(A);
        split_components(B, rb, children);
        std::vector<int32_t>().swap(B);
        roots.push_back(new_node(Sep.data(), (int64_t)Sep.size(), std::move(children)));
    }
};

inline double sum_sq(double x) { return x * (x + 1.0) * (2.0 * x + 1.0) / 6.0; }

}  // namespace

int symbolic_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, const SymbolicOptions &o, Symbolic *S)
{
    if (n >= (int64_t)1 << 31) return SPEIGS_ERR_ARG;
    double t0 = wall();
    Graph g;
    build_graph(n, Ap, Ai, g);
    S->n = n;
    S->perm.assign(n, 0);
    Dissection nd(g, std::max<int64_t>(o.leaf, 1), S->perm);
    nd.run();
    S->iperm.assign(n, 0);
    for (int64_t i = 0; i < n; i++) S->iperm[S->perm[i]] = i;
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
    // boundary rows, children before parents (ids are a postorder)
    std::vector<int64_t> mark(n, -1), cand;
    S->rptr.assign(ns + 1, 0);
    S->rows.clear();
    for (int64_t s = 0; s < ns; s++) {
        cand.clear();
        const int64_t hi = S->c1[s];
        for (int64_t p = S->chptr[s]; p < S->chptr[s + 1]; p++) {
            const int64_t c = S->chl[p];
            for (int64_t q = S->rptr[c]; q < S->rptr[c + 1]; q++) {
                const int64_t v = S->rows[q];
                if (v >= hi && mark[v] != s) { mark[v] = s; cand.push_back(v); }
            }
        }
        for (int64_t col = S->c0[s]; col < hi; col++) {
            const int32_t old = (int32_t)S->perm[col];
            for (int64_t p = g.xadj[old]; p < g.xadj[old + 1]; p++) {
                const int64_t v = S->iperm[g.adj[p]];
                if (v >= hi && mark[v] != s) { mark[v] = s; cand.push_back(v); }
            }
        }
        std::sort(cand.begin(), cand.end());
        S->rows.insert(S->rows.end(), cand.begin(), cand.end());
        S->rptr[s + 1] = (int64_t)S->rows.size();
    }
    // levels, front sizes, offsets, counts
    S->level.assign(ns, 0);
    S->nlevels = 0;
    S->F.assign(ns, 0);
    S->foff.assign(ns, 0);
    S->pool = 0;
    S->flops = 0.0;
    S->nnzL = 0;
    for (int64_t s = 0; s < ns; s++) {
        int32_t lv = 0;
        for (int64_t p = S->chptr[s]; p < S->chptr[s + 1]; p++) lv = std::max(lv, S->level[S->chl[p]] + 1);
        S->level[s] = lv;
        S->nlevels = std::max(S->nlevels, lv + 1);
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
    S->rel.assign(std::max<int64_t>(S->relptr[ns], 1), 0);
    for (int64_t s = 0; s < ns; s++) {
        const int64_t p = S->parent[s];
        if (p < 0) {
            if (S->rptr[s + 1] != S->rptr[s]) return SPEIGS_ERR_INTERNAL;
            continue;
        }
        const int64_t cp = S->c1[p] - S->c0[p];
        const int64_t *Rp = S->rows.data() + S->rptr[p], nRp = S->rptr[p + 1] - S->rptr[p];
        int64_t k = 0;
        for (int64_t q = S->rptr[s]; q < S->rptr[s + 1]; q++) {
            const int64_t v = S->rows[q];
            int64_t pos;
            if (v < S->c1[p]) {
                if (v < S->c0[p]) return SPEIGS_ERR_INTERNAL;
                pos = v - S->c0[p];
            } else {
                while (k < nRp && Rp[k] < v) k++;
                if (k >= nRp || Rp[k] != v) return SPEIGS_ERR_INTERNAL;
                pos = cp + k;
            }
            S->rel[S->relptr[s] + (q - S->rptr[s])] = (int32_t)pos;
        }
    }
    S->t_symbolic = wall() - t1;
    return SPEIGS_OK;
}

}  // namespace sehip
