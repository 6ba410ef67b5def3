// batched.cpp -- thick-restart Lanczos for many components in lockstep.
//
// Every component follows the reference recurrence (src/speigs_core.c:146-315)
// with its own basis dimension, restart count, breakdown handling, random
// stream and convergence test, exactly as the single-matrix driver
// (lanczos.cpp) does; what is shared is the device work.  The components are
// the diagonal blocks of one block-diagonal operator whose vectors hold all
// components' rows, so one operator apply (one solve with the factor of the
// whole block-diagonal matrix) serves the Krylov step of every component that
// is extending, and the CGS2 passes, norms, Ritz products and residuals are
// segmented kernels over all components (batch.hip).  A component that
// converges drops out of the masks; the others go on.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "batch.h"

namespace sehip {

namespace {

struct KeyIdx {
    double key;
    int64_t idx;
};
int by_key_desc(const void *a, const void *b)
{
    const double x = ((const KeyIdx *)a)->key, y = ((const KeyIdx *)b)->key;
    return (x < y) - (x > y);
}

double host_sumsq(int64_t n, const double *x)
{
    if (n <= 32) {
        double s = 0.0;
        for (int64_t i = 0; i < n; i++) s += x[i] * x[i];
        return s;
    }
    return host_sumsq(n / 2, x) + host_sumsq(n - n / 2, x + n / 2);
}

enum { EXTEND = 0, EVAL = 1, DONE = 2 };

struct CompState {
    int64_t ncv = 0, ldh = 0, kcap = 0, maxit = 0, j = 0, m = 0, nwant = 0, l = 0;
    int64_t nret = 0, computed = 0, base = 0;
    double rtol = 0.0, small = 0.0, beta = 0.0, beta_rec = 0.0;
    uint64_t seed = 0;
    int state = EXTEND;
    bool restart = false;
    std::vector<double> H, Hs, Z, theta, keep;
    std::vector<KeyIdx> ord;
};

struct BWork {
    hipStream_t st = nullptr;
    int64_t n = 0, ld = 0, nc = 0;
    int ps = 0;
    int64_t ss = 0;
    SegLayout L;
    double *V = nullptr, *Vn = nullptr, *w = nullptr, *x = nullptr, *X = nullptr, *AX = nullptr, *T = nullptr;
    double *partial = nullptr, *h1 = nullptr, *h2 = nullptr, *nr = nullptr, *lam = nullptr, *rr = nullptr;
    double *S = nullptr, *sc = nullptr, *sc2 = nullptr;
    int32_t *act = nullptr, *mc = nullptr, *jc = nullptr, *one = nullptr;
    std::vector<double> hh1, hh2, hnr, hS, hsc;
    std::vector<int32_t> iact, im, ij, ione;
};

int up_i32(BWork &W, int32_t *d, const std::vector<int32_t> &h)
{
    SEHIP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice, W.st));
    return 0;
}

int up_f64(BWork &W, double *d, const std::vector<double> &h)
{
    SEHIP_CHECK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, W.st));
    return 0;
}

int down_f64(BWork &W, std::vector<double> &h, const double *d)
{
    SEHIP_CHECK(hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, W.st));
    return 0;
}

// CGS2 of w against B[:, 0:m_c] for the components with act_c (W.iact, W.im
// set by the caller); |w|^2 per component into W.hnr, the pass coefficients
// into W.hh1 / W.hh2 (rows of W.ps)
int seg_cgs2(BWork &W, const double *B, double *w)
{
    int rc;
    std::vector<int32_t> mm(W.nc);
    for (int64_t c = 0; c < W.nc; c++) mm[c] = W.iact[c] ? W.im[c] : 0;
    if ((rc = up_i32(W, W.act, W.iact))) return rc;
    if ((rc = up_i32(W, W.mc, mm))) return rc;
    for (int64_t c = 0; c < W.nc; c++) W.ione[c] = W.iact[c] ? 1 : 0;
    if ((rc = up_i32(W, W.one, W.ione))) return rc;
    seg_dots(W.L, B, W.ld, W.mc, w, W.partial, W.ps, W.st);
    seg_reduce(W.L, W.partial, W.ps, W.mc, W.h1, W.st);
    seg_update(W.L, B, W.ld, W.act, W.mc, W.h1, w, W.partial, W.ps, 1, W.st);
    seg_reduce(W.L, W.partial, W.ps, W.mc, W.h2, W.st);
    seg_update(W.L, B, W.ld, W.act, W.mc, W.h2, w, W.partial, W.ps, 2, W.st);
    seg_reduce(W.L, W.partial, W.ps, W.one, W.nr, W.st);
    if ((rc = down_f64(W, W.hh1, W.h1))) return rc;
    if ((rc = down_f64(W, W.hh2, W.h2))) return rc;
    if ((rc = down_f64(W, W.hnr, W.nr))) return rc;
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    return 0;
}

// the next n_c values of component c's xorshift64* stream into rows of w
int reinject(BWork &W, const BatchComp &bc, CompState &cs, double *w, std::vector<double> &buf)
{
    buf.resize((size_t)bc.n);
    uniform_vec(&cs.seed, bc.n, buf.data());
    SEHIP_CHECK(hipMemcpyAsync(w + bc.off, buf.data(), (size_t)bc.n * sizeof(double), hipMemcpyHostToDevice, W.st));
    return 0;
}

}  // namespace

int lanczos_batched(DevOperator *op, int64_t n, std::vector<BatchComp> &comps, const speigs_opts *o,
                    bool want_vectors)
{
    const int64_t nc = (int64_t)comps.size();
    if (nc == 0) return 0;
    BWork W;
    DevMem mem;
    W.st = op->stream;
    W.n = n;
    W.ld = pad_ld(n);
    W.nc = nc;
    std::vector<CompState> cs(nc);
    int64_t NCV = 2, KCAP = 1;
    const uint64_t seed0 = o->seed ? o->seed : 88172645463325252ULL;
    for (int64_t c = 0; c < nc; c++) {
        const BatchComp &b = comps[c];
        CompState &s = cs[c];
        int64_t ncv = o->ncv ? o->ncv : std::max<int64_t>(2 * b.k + 1, 20);
        if (ncv > b.n) ncv = b.n;
        if (ncv <= b.k) ncv = std::min<int64_t>(b.n, b.k + 1);
        if (ncv < 2) ncv = std::min<int64_t>(b.n, 2);
        if (ncv > 4000 || ncv < 1) return SPEIGS_ERR_ARG;
        s.ncv = s.ldh = ncv;
        s.kcap = std::max<int64_t>({b.k, ncv, 1});
        s.maxit = o->maxit ? o->maxit : 300;
        const double tol = o->tol > 0 ? o->tol : 1e-14, scale = b.anorm > 1.0 ? b.anorm : 1.0;
        s.rtol = tol * scale;
        s.small = 1e-13 * scale;
        s.seed = seed0;
        s.H.assign((size_t)ncv * ncv, 0.0);
        s.Hs.assign((size_t)ncv * ncv, 0.0);
        s.Z.assign((size_t)ncv * ncv, 0.0);
        s.theta.assign(ncv, 0.0);
        s.keep.assign(ncv, 0.0);
        s.ord.resize(ncv);
        NCV = std::max(NCV, ncv);
        KCAP = std::max(KCAP, s.kcap);
    }
    W.ps = (int)std::max<int64_t>(NCV + 2, KCAP);
    W.ss = (NCV + 1) * KCAP;
    int rc;
    if ((rc = seg_layout_build([&] {
             std::vector<int64_t> off(nc + 1);
             for (int64_t c = 0; c < nc; c++) off[c] = comps[c].off;
             off[nc] = n;
             return off;
         }(), W.L, mem, W.st)))
        return rc;
    const size_t col = (size_t)W.ld;
    if ((rc = mem.alloc(&W.V, col * (NCV + 1)))) return rc;
    if ((rc = mem.alloc(&W.Vn, col * (NCV + 1)))) return rc;
    if ((rc = mem.alloc(&W.w, col))) return rc;
    if ((rc = mem.alloc(&W.x, col))) return rc;
    if ((rc = mem.alloc(&W.X, col * KCAP))) return rc;
    if ((rc = mem.alloc(&W.AX, col * KCAP))) return rc;
    if ((rc = mem.alloc(&W.T, col * KCAP))) return rc;
    if ((rc = mem.alloc(&W.partial, (size_t)W.L.nchunks * W.ps))) return rc;
    const size_t cps = (size_t)nc * W.ps;
    if ((rc = mem.alloc(&W.h1, cps))) return rc;
    if ((rc = mem.alloc(&W.h2, cps))) return rc;
    if ((rc = mem.alloc(&W.nr, cps))) return rc;
    if ((rc = mem.alloc(&W.lam, cps))) return rc;
    if ((rc = mem.alloc(&W.rr, cps))) return rc;
    if ((rc = mem.alloc(&W.sc2, cps))) return rc;
    if ((rc = mem.alloc(&W.S, (size_t)nc * W.ss))) return rc;
    if ((rc = mem.alloc(&W.sc, (size_t)nc))) return rc;
    if ((rc = mem.alloc(&W.act, (size_t)nc))) return rc;
    if ((rc = mem.alloc(&W.mc, (size_t)nc))) return rc;
    if ((rc = mem.alloc(&W.jc, (size_t)nc))) return rc;
    if ((rc = mem.alloc(&W.one, (size_t)nc))) return rc;
    W.hh1.assign(cps, 0.0);
    W.hh2.assign(cps, 0.0);
    W.hnr.assign(cps, 0.0);
    W.hS.assign((size_t)nc * W.ss, 0.0);
    W.hsc.assign((size_t)nc, 0.0);
    W.iact.assign(nc, 0);
    W.im.assign(nc, 0);
    W.ij.assign(nc, -1);
    W.ione.assign(nc, 0);
    SEHIP_CHECK(hipMemsetAsync(W.V, 0, col * (NCV + 1) * sizeof(double), W.st));
    SEHIP_CHECK(hipMemsetAsync(W.Vn, 0, col * (NCV + 1) * sizeof(double), W.st));
    SEHIP_CHECK(hipMemsetAsync(W.X, 0, col * KCAP * sizeof(double), W.st));

    // start vectors: each component's own copy of the reference stream
    {
        std::vector<double> hv((size_t)n, 0.0);
        for (int64_t c = 0; c < nc; c++) {
            double *seg = hv.data() + comps[c].off;
            uniform_vec(&cs[c].seed, comps[c].n, seg);
            const double nr = std::sqrt(host_sumsq(comps[c].n, seg));
            if (nr == 0.0) return SEHIP_FAIL(SPEIGS_ERR_INTERNAL, "re-injected start vector has norm 0");
            const double s = 1.0 / nr;
            for (int64_t i = 0; i < comps[c].n; i++) seg[i] *= s;
        }
        SEHIP_CHECK(hipMemcpyAsync(W.V, hv.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, W.st));
        SEHIP_CHECK(hipStreamSynchronize(W.st));
    }
    std::vector<double> buf, sel, lamh(cps), rrh(cps), xxh(cps), sc2h(cps, 1.0);
    std::vector<int32_t> pc(nc, 0), mm(nc, 0), jd(nc, -1);
    auto active = [&](int st2) {
        for (int64_t c = 0; c < nc; c++)
            if (cs[c].state == st2) return true;
        return false;
    };
    double t_apply = 0.0, t_orth = 0.0, t_eval = 0.0, t_rest = 0.0;
    int64_t steps = 0, evals = 0;
    while (active(EXTEND) || active(EVAL)) {
        // ---- one Krylov step of every extending component
        if (active(EXTEND)) {
            steps++;
            double ta = wall();
            for (int64_t c = 0; c < nc; c++) {
                const bool e = cs[c].state == EXTEND;
                W.ij[c] = e ? (int32_t)cs[c].j : -1;
                W.iact[c] = e;
                W.im[c] = e ? (int32_t)(cs[c].j + 1) : 0;
            }
            if ((rc = up_i32(W, W.jc, W.ij))) return rc;
            seg_gather_col(W.L, W.V, W.ld, W.jc, W.x, W.st);
            if ((rc = op->apply(W.x, W.w))) return rc;
            for (int64_t c = 0; c < nc; c++)
                if (W.iact[c]) comps[c].nops++;
            if (trace_on()) {
                SEHIP_CHECK(hipStreamSynchronize(W.st));
                t_apply += wall() - ta;
                ta = wall();
            }
            if ((rc = seg_cgs2(W, W.V, W.w))) return rc;
            std::vector<int32_t> brk(nc, 0);
            bool anybrk = false;
            for (int64_t c = 0; c < nc; c++) {
                if (!W.iact[c]) continue;
                CompState &s = cs[c];
                double *hc = s.H.data() + (size_t)s.j * s.ldh;
                for (int64_t i = 0; i <= s.j; i++) hc[i] = (0.0 + W.hh1[c * W.ps + i]) + W.hh2[c * W.ps + i];
                s.beta = std::sqrt(W.hnr[c * W.ps]);
                s.beta_rec = s.beta;
                if (s.beta <= s.small) {
                    s.beta_rec = 0.0;
                    brk[c] = 1;
                    anybrk = true;
                }
            }
            for (int att = 0; att < 3 && anybrk; att++) {
                anybrk = false;
                for (int64_t c = 0; c < nc; c++) {
                    W.iact[c] = brk[c];
                    W.im[c] = brk[c] ? (int32_t)(cs[c].j + 1) : 0;
                    if (brk[c] && (rc = reinject(W, comps[c], cs[c], W.w, buf))) return rc;
                }
                if ((rc = seg_cgs2(W, W.V, W.w))) return rc;
                for (int64_t c = 0; c < nc; c++) {
                    if (!brk[c]) continue;
                    cs[c].beta = std::sqrt(W.hnr[c * W.ps]);
                    if (cs[c].beta <= cs[c].small) anybrk = true;
                    else brk[c] = 0;
                }
            }
            for (int64_t c = 0; c < nc; c++) {
                CompState &s = cs[c];
                jd[c] = -1;
                if (s.state != EXTEND) continue;
                if (brk[c]) {                 // invariant subspace: the basis stops here
                    s.ncv = s.j + 1;
                    s.state = EVAL;
                    continue;
                }
                jd[c] = (int32_t)(s.j + 1);
                W.hsc[c] = 1.0 / s.beta;
            }
            if ((rc = up_i32(W, W.jc, jd))) return rc;
            if ((rc = up_f64(W, W.sc, W.hsc))) return rc;
            seg_store_col(W.L, W.w, W.sc, W.jc, W.V, W.ld, W.st);
            for (int64_t c = 0; c < nc; c++) {
                CompState &s = cs[c];
                if (jd[c] < 0) continue;
                if (s.j + 1 < s.ncv) {
                    s.H[(s.j + 1) + (size_t)s.j * s.ldh] = s.beta_rec;
                    s.H[s.j + (size_t)(s.j + 1) * s.ldh] = s.beta_rec;
                }
                s.j++;
                if (s.j >= s.ncv) s.state = EVAL;
            }
            t_orth += wall() - ta;
            continue;   // extend until no component is extending
        }
        evals++;
        double te = wall();

        // ---- projected problems and residuals of every component at a restart point
        for (int64_t c = 0; c < nc; c++) {
            pc[c] = 0;
            mm[c] = 0;
            CompState &s = cs[c];
            if (s.state != EVAL) continue;
            // leading-dimension quirk of src/speigs_core.c:233-234: after a
            // breakdown shrank ncv, H is read with the new ncv as its stride
            const int64_t m = s.ncv;
            s.m = m;
            for (int64_t cc = 0; cc < m; cc++)
                for (int64_t r = 0; r < m; r++) s.Hs[r + cc * m] = s.H[r + cc * s.ncv];
            if (jacobi_eig(m, s.Hs.data(), m, s.Z.data(), s.theta.data())) return SPEIGS_ERR_LAPACK;
            for (int64_t i = 0; i < m; i++) s.ord[i] = {std::fabs(s.theta[i]), i};
            qsort(s.ord.data(), (size_t)m, sizeof(KeyIdx), by_key_desc);
            s.nwant = std::min<int64_t>(comps[c].k, m);
            double *Sc = W.hS.data() + (size_t)c * W.ss;
            for (int64_t t = 0; t < s.nwant; t++)
                std::memcpy(Sc + (size_t)t * m, s.Z.data() + (size_t)s.ord[t].idx * m, (size_t)m * sizeof(double));
            pc[c] = (int32_t)s.nwant;
            mm[c] = (int32_t)m;
        }
        int nwmax = 0;
        for (int64_t c = 0; c < nc; c++) nwmax = std::max(nwmax, pc[c]);
        if ((rc = up_f64(W, W.S, W.hS))) return rc;
        if ((rc = up_i32(W, W.mc, mm))) return rc;
        if ((rc = up_i32(W, W.act, pc))) return rc;     // act holds p_c here
        seg_gemm(W.L, W.V, W.ld, W.mc, W.act, W.S, W.ss, W.X, W.ld, W.st);
        seg_pairs(W.L, W.X, W.ld, W.X, W.ld, W.act, W.lam, W.partial, W.ps, 0, W.st);
        seg_reduce(W.L, W.partial, W.ps, W.act, W.rr, W.st);
        if ((rc = down_f64(W, xxh, W.rr))) return rc;
        SEHIP_CHECK(hipStreamSynchronize(W.st));
        for (int64_t c = 0; c < nc; c++)
            for (int t = 0; t < pc[c]; t++) {
                const double xn = std::sqrt(xxh[c * W.ps + t]);
                sc2h[c * W.ps + t] = xn > 0 ? 1.0 / xn : 1.0;
            }
        if ((rc = up_f64(W, W.sc2, sc2h))) return rc;
        seg_scale_cols(W.L, W.X, W.ld, W.act, W.sc2, W.ps, W.st);
        if ((rc = op->multA(W.X, W.ld, W.AX, W.ld, nwmax))) return rc;
        seg_pairs(W.L, W.X, W.ld, W.AX, W.ld, W.act, W.lam, W.partial, W.ps, 0, W.st);
        seg_reduce(W.L, W.partial, W.ps, W.act, W.lam, W.st);
        seg_pairs(W.L, W.X, W.ld, W.AX, W.ld, W.act, W.lam, W.partial, W.ps, 1, W.st);
        seg_reduce(W.L, W.partial, W.ps, W.act, W.rr, W.st);
        if ((rc = down_f64(W, lamh, W.lam))) return rc;
        if ((rc = down_f64(W, rrh, W.rr))) return rc;
        SEHIP_CHECK(hipStreamSynchronize(W.st));
        if ((rc = op->check())) return rc;
        bool anyrestart = false;
        for (int64_t c = 0; c < nc; c++) {
            CompState &s = cs[c];
            if (s.state != EVAL) continue;
            BatchComp &b = comps[c];
            b.lambda.assign(s.nwant, 0.0);
            b.resid.assign(s.nwant, 0.0);
            int64_t nconv = 0;
            for (int64_t t = 0; t < s.nwant; t++) {
                const double r = std::sqrt(rrh[c * W.ps + t]);
                b.lambda[t] = lamh[c * W.ps + t];
                b.resid[t] = r;
                if (r <= s.rtol) nconv++;
            }
            b.nconv = nconv;
            if (nconv >= std::min<int64_t>(b.k, s.m) || b.restarts >= s.maxit || s.m < s.ncv) {
                s.state = DONE;
                continue;
            }
            b.restarts++;
            s.restart = true;
            anyrestart = true;
            s.l = std::max<int64_t>(1, std::min<int64_t>(b.k + 5, s.ncv - 1));
            s.nret = s.computed = s.base = 0;
        }
        t_eval += wall() - te;
        te = wall();
        if (!anyrestart) continue;

        // ---- thick restart: l dominant Ritz vectors per component, orthonormalised one by one
        int64_t mmax = 0;
        for (int64_t c = 0; c < nc; c++)
            if (cs[c].restart) mmax = std::max(mmax, cs[c].m);
        for (int64_t t = 0; t < mmax; t++) {
            bool any = false, batch = false;
            for (int64_t c = 0; c < nc; c++) {
                CompState &s = cs[c];
                const bool a = s.restart && t < s.m && s.nret < s.l;
                W.iact[c] = a;
                W.im[c] = a ? (int32_t)s.nret : 0;
                pc[c] = 0;
                mm[c] = 0;
                if (!a) continue;
                any = true;
                if (t >= s.computed) {          // next batch of candidate Ritz vectors into T
                    int64_t p = std::min<int64_t>(s.l, s.m - t);
                    p = std::min<int64_t>(p, s.kcap);
                    double *Sc = W.hS.data() + (size_t)c * W.ss;
                    for (int64_t q = 0; q < p; q++)
                        std::memcpy(Sc + (size_t)q * s.m, s.Z.data() + (size_t)s.ord[t + q].idx * s.m,
                                    (size_t)s.m * sizeof(double));
                    s.base = t;
                    s.computed = t + p;
                    pc[c] = (int32_t)p;
                    mm[c] = (int32_t)s.m;
                    batch = true;
                }
            }
            if (!any) break;
            if (batch) {
                if ((rc = up_f64(W, W.S, W.hS))) return rc;
                if ((rc = up_i32(W, W.mc, mm))) return rc;
                if ((rc = up_i32(W, W.one, pc))) return rc;
                seg_gemm(W.L, W.V, W.ld, W.mc, W.one, W.S, W.ss, W.T, W.ld, W.st);
            }
            for (int64_t c = 0; c < nc; c++) W.ij[c] = W.iact[c] ? (int32_t)(t - cs[c].base) : -1;
            if ((rc = up_i32(W, W.jc, W.ij))) return rc;
            seg_gather_col(W.L, W.T, W.ld, W.jc, W.x, W.st);
            if ((rc = seg_cgs2(W, W.Vn, W.x))) return rc;
            for (int64_t c = 0; c < nc; c++) {
                jd[c] = -1;
                if (!W.iact[c]) continue;
                CompState &s = cs[c];
                const double nr = std::sqrt(W.hnr[c * W.ps]);
                if (nr < 1e-10) continue;
                jd[c] = (int32_t)s.nret;
                W.hsc[c] = 1.0 / nr;
                s.keep[s.nret] = s.theta[s.ord[t].idx];
                s.nret++;
            }
            if ((rc = up_i32(W, W.jc, jd))) return rc;
            if ((rc = up_f64(W, W.sc, W.hsc))) return rc;
            seg_store_col(W.L, W.x, W.sc, W.jc, W.Vn, W.ld, W.st);
        }
        // continuation vector: the last basis vector, orthogonalised against the kept ones
        for (int64_t c = 0; c < nc; c++) {
            W.iact[c] = cs[c].restart;
            W.im[c] = cs[c].restart ? (int32_t)cs[c].nret : 0;
            W.ij[c] = cs[c].restart ? (int32_t)cs[c].m : -1;
        }
        if ((rc = up_i32(W, W.jc, W.ij))) return rc;
        seg_gather_col(W.L, W.V, W.ld, W.jc, W.w, W.st);
        if ((rc = seg_cgs2(W, W.Vn, W.w))) return rc;
        std::vector<double> nrv(nc, 0.0);
        bool anysmall = false;
        for (int64_t c = 0; c < nc; c++) {
            if (!cs[c].restart) continue;
            nrv[c] = std::sqrt(W.hnr[c * W.ps]);
            if (nrv[c] < 1e-10) anysmall = true;
        }
        if (anysmall) {
            for (int64_t c = 0; c < nc; c++) {
                const bool a = cs[c].restart && nrv[c] < 1e-10;
                W.iact[c] = a;
                W.im[c] = a ? (int32_t)cs[c].nret : 0;
                if (a && (rc = reinject(W, comps[c], cs[c], W.w, buf))) return rc;
            }
            if ((rc = seg_cgs2(W, W.Vn, W.w))) return rc;
            for (int64_t c = 0; c < nc; c++)
                if (W.iact[c]) nrv[c] = std::sqrt(W.hnr[c * W.ps]);
        }
        for (int64_t c = 0; c < nc; c++) {
            jd[c] = -1;
            mm[c] = 0;
            CompState &s = cs[c];
            if (!s.restart) continue;
            s.restart = false;
            if (nrv[c] < 1e-10) {             // no new direction: stop with the last Ritz pairs
                s.state = DONE;
                continue;
            }
            jd[c] = (int32_t)s.nret;
            W.hsc[c] = 1.0 / nrv[c];
            mm[c] = (int32_t)(s.nret + 1);
            std::fill(s.H.begin(), s.H.end(), 0.0);
            for (int64_t q = 0; q < s.nret; q++) s.H[q + (size_t)q * s.ldh] = s.keep[q];
            s.j = s.nret;
            s.state = EXTEND;
        }
        if ((rc = up_i32(W, W.jc, jd))) return rc;
        if ((rc = up_f64(W, W.sc, W.hsc))) return rc;
        seg_store_col(W.L, W.w, W.sc, W.jc, W.Vn, W.ld, W.st);
        if ((rc = up_i32(W, W.mc, mm))) return rc;
        seg_copy_cols(W.L, W.Vn, W.V, W.ld, W.mc, W.st);
        t_rest += wall() - te;
    }
    if (trace_on())
        fprintf(stderr, "[speigs_hip] batched Lanczos: %lld steps (apply %.1f ms, CGS2 %.1f ms), %lld evaluations "
                        "(%.1f ms), restarts %.1f ms\n",
                (long long)steps, 1e3 * t_apply, 1e3 * t_orth, (long long)evals, 1e3 * t_eval, 1e3 * t_rest);
    for (int64_t c = 0; c < nc; c++) {
        BatchComp &b = comps[c];
        b.flag = b.nconv >= std::min<int64_t>(b.k, cs[c].ncv) ? 0 : 1;
    }
    if (want_vectors) {
        // the Ritz vectors of the last evaluation of every component, its rows of X
        std::vector<double> col((size_t)n);
        for (int64_t t = 0; t < KCAP; t++) {
            bool need = false;
            for (int64_t c = 0; c < nc; c++) need = need || t < (int64_t)comps[c].lambda.size();
            if (!need) break;
            SEHIP_CHECK(hipMemcpyAsync(col.data(), W.X + (size_t)t * W.ld, (size_t)n * sizeof(double),
                                       hipMemcpyDeviceToHost, W.st));
            SEHIP_CHECK(hipStreamSynchronize(W.st));
            for (int64_t c = 0; c < nc; c++) {
                BatchComp &b = comps[c];
                const int64_t nw = (int64_t)b.lambda.size();
                if (t >= nw) continue;
                if (b.V.size() != (size_t)b.n * nw) b.V.assign((size_t)b.n * nw, 0.0);
                std::memcpy(b.V.data() + (size_t)t * b.n, col.data() + b.off, (size_t)b.n * sizeof(double));
            }
        }
    }
    SEHIP_CHECK(hipStreamSynchronize(W.st));
    return SPEIGS_OK;
}

}  // namespace sehip
