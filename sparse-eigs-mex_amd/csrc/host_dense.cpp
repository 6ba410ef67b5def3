// host_dense.cpp -- small dense kernels that stay on the host CPU:
//   * jacobi_eig: the m x m projected Lanczos problem (the reference calls
//     dsyev_("V","U"), src/speigs_core.c:232-237).  Cyclic two-sided Jacobi
//     with the relative stopping test |a_pq| <= eps sqrt(|a_pp a_qq|) keeps the
//     small Ritz values of a graded positive definite projection to relative
//     accuracy (Demmel and Veselic 1992); m <= ncv is at most a few hundred.
//   * dense_syev: Householder tridiagonalisation, implicit QL with shifts,
//     ascending order, all vectors (small projected problems);
//     dense_tri_values / dense_tri_vectors: the dense short-circuit
//     (src/speigs_core.c:318-353), all eigenvalues but only the k returned vectors.
//   * band_chol_u / band_solve_u: the banded short-circuit (dpbtrf/dpbtrs "U",
//     src/speigs_core.c:869-931, :101-110), LAPACK's upper band storage.
//   * lu_factor / lu_solve: dense partial-pivoting LU for the indefinite
//     fallback at small n (the reference uses UMFPACK there, :974-984).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "dev.h"

namespace sehip {

static const double kEps = 2.220446049250313e-16;

static void sort_ascending(int64_t m, double *w, double *Z, int64_t ldz)
{
    for (int64_t i = 0; i + 1 < m; i++) {
        int64_t b = i;
        for (int64_t j = i + 1; j < m; j++)
            if (w[j] < w[b]) b = j;
        if (b != i) {
            std::swap(w[i], w[b]);
            for (int64_t r = 0; r < m; r++) std::swap(Z[r + i * ldz], Z[r + b * ldz]);
        }
    }
}

int jacobi_eig(int64_t m, const double *Hu, int64_t ldh, double *Z, double *w)
{
    std::vector<double> a((size_t)m * m);
    for (int64_t c = 0; c < m; c++)
        for (int64_t r = 0; r < m; r++) a[r + c * m] = (r <= c) ? Hu[r + c * ldh] : Hu[c + r * ldh];
    for (int64_t c = 0; c < m; c++)
        for (int64_t r = 0; r < m; r++) Z[r + c * m] = (r == c) ? 1.0 : 0.0;
    int sweep = 0;
    for (; sweep < 100; sweep++) {
        bool any = false;
        for (int64_t p = 0; p + 1 < m; p++)
            for (int64_t q = p + 1; q < m; q++) {
                const double apq = a[p + q * m], app = a[p + p * m], aqq = a[q + q * m];
                if (apq == 0.0 || std::fabs(apq) <= kEps * std::sqrt(std::fabs(app * aqq))) continue;
                any = true;
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int64_t r = 0; r < m; r++) {
                    const double x = a[r + p * m], y = a[r + q * m];
                    a[r + p * m] = c * x - s * y;
                    a[r + q * m] = s * x + c * y;
                }
                for (int64_t r = 0; r < m; r++) {
                    const double x = a[p + r * m], y = a[q + r * m];
                    a[p + r * m] = c * x - s * y;
                    a[q + r * m] = s * x + c * y;
                }
                a[p + p * m] = app - t * apq;
                a[q + q * m] = aqq + t * apq;
                a[p + q * m] = a[q + p * m] = 0.0;
                for (int64_t r = 0; r < m; r++) {
                    const double x = Z[r + p * m], y = Z[r + q * m];
                    Z[r + p * m] = c * x - s * y;
                    Z[r + q * m] = s * x + c * y;
                }
            }
        if (!any) break;
    }
    if (sweep >= 100) return 1;
    for (int64_t i = 0; i < m; i++) w[i] = a[i + i * m];
    sort_ascending(m, w, Z, m);
    return 0;
}

int dense_syev(int64_t n, double *A, double *d)
{
    if (n <= 0) return 0;
    if (n == 1) { d[0] = A[0]; A[0] = 1.0; return 0; }
    for (int64_t c = 0; c < n; c++)
        for (int64_t r = c + 1; r < n; r++) A[r + c * n] = A[c + r * n];
    std::vector<double> Q((size_t)n * n, 0.0), v(n), p(n), e(n, 0.0);
    for (int64_t i = 0; i < n; i++) Q[i + i * n] = 1.0;
    for (int64_t k = 0; k + 2 < n; k++) {
        double xn = 0.0;
        for (int64_t i = k + 1; i < n; i++) xn += A[i + k * n] * A[i + k * n];
        xn = std::sqrt(xn);
        const double x0 = A[(k + 1) + k * n];
        if (xn == 0.0) { e[k] = 0.0; continue; }
        const double alpha = x0 > 0 ? -xn : xn;
        for (int64_t i = k + 1; i < n; i++) v[i] = A[i + k * n];
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int64_t i = k + 1; i < n; i++) vn += v[i] * v[i];
        vn = std::sqrt(vn);
        if (vn == 0.0) { e[k] = x0; continue; }
        for (int64_t i = k + 1; i < n; i++) v[i] /= vn;
        double K = 0.0;
        for (int64_t i = k + 1; i < n; i++) p[i] = 0.0;
        for (int64_t j = k + 1; j < n; j++) {      // p = A22 v, column sweep
            const double vj = v[j];
            const double *aj = A + j * n;
            for (int64_t i = k + 1; i < n; i++) p[i] += aj[i] * vj;
        }
        for (int64_t i = k + 1; i < n; i++) K += v[i] * p[i];
        for (int64_t i = k + 1; i < n; i++) p[i] = 2.0 * p[i] - 2.0 * K * v[i];
        for (int64_t j = k + 1; j < n; j++) {
            double *aj = A + j * n;
            const double vj = v[j], pj = p[j];
            for (int64_t i = k + 1; i < n; i++) aj[i] -= v[i] * pj + p[i] * vj;
        }
        e[k] = alpha;
        // Q <- Q (I - 2 v v^T): t = Q v by column sweeps, then each column of Q
        // minus 2 v_i t (contiguous, vectorisable)
        for (int64_t r = 0; r < n; r++) p[r] = 0.0;
        for (int64_t i = k + 1; i < n; i++) {
            const double vi = v[i];
            const double *__restrict__ qi = Q.data() + i * n;
            double *__restrict__ t = p.data();
            for (int64_t r = 0; r < n; r++) t[r] += qi[r] * vi;
        }
        for (int64_t i = k + 1; i < n; i++) {
            const double vi2 = 2.0 * v[i];
            double *__restrict__ qi = Q.data() + i * n;
            const double *__restrict__ t = p.data();
            for (int64_t r = 0; r < n; r++) qi[r] -= vi2 * t[r];
        }
    }
    for (int64_t i = 0; i < n; i++) d[i] = A[i + i * n];
    if (n >= 2) {
        // e[k] for k <= n-3 were set above; the last sub-diagonal entry is read off
        e[n - 2] = A[(n - 1) + (n - 2) * n];
    }
    e[n - 1] = 0.0;
    // implicit QL
    for (int64_t l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int64_t mm = l;
            for (; mm < n - 1; mm++)
                if (std::fabs(e[mm]) <= kEps * (std::fabs(d[mm]) + std::fabs(d[mm + 1]))) break;
            if (mm == l) break;
            if (++iter > 60) return (int)(l + 1);
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            bool under = false;
            for (int64_t i = mm - 1; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= pp; e[mm] = 0.0; under = true; break; }
                s = f / r;
                c = g / r;
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
                double *__restrict__ zi = Q.data() + i * n;
                double *__restrict__ zi1 = Q.data() + (i + 1) * n;
                for (int64_t q = 0; q < n; q++) {
                    const double zf = zi1[q];
                    zi1[q] = s * zi[q] + c * zf;
                    zi[q] = c * zi[q] - s * zf;
                }
            }
            if (under) continue;
            d[l] -= pp;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    std::memcpy(A, Q.data(), (size_t)n * n * sizeof(double));
    sort_ascending(n, d, A, n);
    return 0;
}

// ---- selected eigenpairs of a dense symmetric matrix (the dense short-circuit
// needs k of n): Householder tridiagonalisation with the reflectors kept in
// place, implicit QL for all eigenvalues (the same iteration as dense_syev, so
// the same values), inverse iteration on the tridiagonal for the selected
// ones (Wilkinson; vectors of a cluster are reorthogonalised against each
// other, modified Gram-Schmidt), and the back-transformation of just those
// vectors: O(n^3 * 4/3) + O(n^2 k) instead of O(n^3) per accumulated sweep.
namespace {

// A (n x n, full symmetric) -> tridiagonal d, e; reflector k (unit vector,
// H_k = I - 2 v v^T acting on rows k+1..n-1) left in A[k+1.., k]
void tridiag_inplace(int64_t n, double *A, double *d, double *e)
{
    std::vector<double> v(n), p(n);
    for (int64_t k = 0; k + 2 < n; k++) {
        double xn = 0.0;
        for (int64_t i = k + 1; i < n; i++) xn += A[i + k * n] * A[i + k * n];
        xn = std::sqrt(xn);
        const double x0 = A[(k + 1) + k * n];
        for (int64_t i = k + 1; i < n; i++) v[i] = 0.0;
        if (xn == 0.0) {
            e[k] = 0.0;
            for (int64_t i = k + 1; i < n; i++) A[i + k * n] = 0.0;
            continue;
        }
        const double alpha = x0 > 0 ? -xn : xn;
        for (int64_t i = k + 1; i < n; i++) v[i] = A[i + k * n];
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int64_t i = k + 1; i < n; i++) vn += v[i] * v[i];
        vn = std::sqrt(vn);
        if (vn == 0.0) {
            e[k] = x0;
            for (int64_t i = k + 1; i < n; i++) A[i + k * n] = 0.0;
            continue;
        }
        for (int64_t i = k + 1; i < n; i++) v[i] /= vn;
        double K = 0.0;
        for (int64_t i = k + 1; i < n; i++) p[i] = 0.0;
        for (int64_t j = k + 1; j < n; j++) {
            const double vj = v[j];
            const double *aj = A + j * n;
            for (int64_t i = k + 1; i < n; i++) p[i] += aj[i] * vj;
        }
        for (int64_t i = k + 1; i < n; i++) K += v[i] * p[i];
        for (int64_t i = k + 1; i < n; i++) p[i] = 2.0 * p[i] - 2.0 * K * v[i];
        for (int64_t j = k + 1; j < n; j++) {
            double *aj = A + j * n;
            const double vj = v[j], pj = p[j];
            for (int64_t i = k + 1; i < n; i++) aj[i] -= v[i] * pj + p[i] * vj;
        }
        e[k] = alpha;
        for (int64_t i = k + 1; i < n; i++) A[i + k * n] = v[i];
    }
    for (int64_t i = 0; i < n; i++) d[i] = A[i + i * n];
    if (n >= 2) e[n - 2] = A[(n - 1) + (n - 2) * n];
    e[n - 1] = 0.0;
}

// the same reduction on the lower triangle of A (column-major, entries above
// the diagonal unused): p = A v as a symmetric product over the lower part and
// the rank-2 update on the lower part only, half the flops of the full one;
// the same reflectors (H_k = I - 2 v v^T, v in A[k+1.., k]) up to rounding
void tridiag_lower(int64_t n, double *A, double *d, double *e)
{
    std::vector<double> vv(n), pv(n);
    double *__restrict__ v = vv.data();
    double *__restrict__ p = pv.data();
    for (int64_t k = 0; k + 2 < n; k++) {
        double *__restrict__ ak = A + k * n;
        double xn = 0.0;
        for (int64_t i = k + 1; i < n; i++) xn += ak[i] * ak[i];
        xn = std::sqrt(xn);
        const double x0 = ak[k + 1];
        if (xn == 0.0) {
            e[k] = 0.0;
            for (int64_t i = k + 1; i < n; i++) ak[i] = 0.0;
            continue;
        }
        const double alpha = x0 > 0 ? -xn : xn;
        for (int64_t i = k + 1; i < n; i++) v[i] = ak[i];
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int64_t i = k + 1; i < n; i++) vn += v[i] * v[i];
        vn = std::sqrt(vn);
        if (vn == 0.0) {
            e[k] = x0;
            for (int64_t i = k + 1; i < n; i++) ak[i] = 0.0;
            continue;
        }
        const double iv = 1.0 / vn;
        for (int64_t i = k + 1; i < n; i++) v[i] *= iv;
        for (int64_t i = k + 1; i < n; i++) p[i] = 0.0;
        for (int64_t j = k + 1; j < n; j++) {
            const double *__restrict__ aj = A + j * n;
            const double vj = v[j];
            double s = aj[j] * vj;
            for (int64_t i = j + 1; i < n; i++) {
                p[i] += aj[i] * vj;
                s += aj[i] * v[i];
            }
            p[j] += s;
        }
        double K = 0.0;
        for (int64_t i = k + 1; i < n; i++) K += v[i] * p[i];
        for (int64_t i = k + 1; i < n; i++) p[i] = 2.0 * p[i] - 2.0 * K * v[i];
        for (int64_t j = k + 1; j < n; j++) {
            double *__restrict__ aj = A + j * n;
            const double vj = v[j], pj = p[j];
            for (int64_t i = j; i < n; i++) aj[i] -= v[i] * pj + p[i] * vj;
        }
        e[k] = alpha;
        for (int64_t i = k + 1; i < n; i++) ak[i] = v[i];
    }
    for (int64_t i = 0; i < n; i++) d[i] = A[i + i * n];
    if (n >= 2) e[n - 2] = A[(n - 1) + (n - 2) * n];
    e[n - 1] = 0.0;
}

// sqrt(f^2 + g^2): the plain formula where neither square can overflow or
// underflow to zero, hypot (scaled) otherwise
inline double rot_norm(double f, double g)
{
    const double m = std::max(std::fabs(f), std::fabs(g));
    if (m > 1e-150 && m < 1e150) return std::sqrt(f * f + g * g);
    return std::hypot(f, g);
}

// eigenvalues of the tridiagonal (d, e) by implicit QL, no vectors; ascending.
// FAST: the rotations' norms by rot_norm instead of hypot
template <bool FAST>
int ql_values_t(int64_t n, double *d, double *e)
{
    for (int64_t l = 0; l < n; l++) {
        int iter = 0;
        for (;;) {
            int64_t mm = l;
            for (; mm < n - 1; mm++)
                if (std::fabs(e[mm]) <= kEps * (std::fabs(d[mm]) + std::fabs(d[mm + 1]))) break;
            if (mm == l) break;
            if (++iter > 60) return (int)(l + 1);
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = FAST ? rot_norm(g, 1.0) : std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, pp = 0.0;
            bool under = false;
            for (int64_t i = mm - 1; i >= l; i--) {
                const double f = s * e[i], b = c * e[i];
                r = FAST ? rot_norm(f, g) : std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= pp; e[mm] = 0.0; under = true; break; }
                if (FAST) {
                    const double ir = 1.0 / r;
                    s = f * ir;
                    c = g * ir;
                } else {
                    s = f / r;
                    c = g / r;
                }
                g = d[i + 1] - pp;
                r = (d[i] - g) * s + 2.0 * c * b;
                pp = s * r;
                d[i + 1] = g + pp;
                g = c * r - b;
            }
            if (under) continue;
            d[l] -= pp;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    std::sort(d, d + n);
    return 0;
}

int ql_values(int64_t n, double *d, double *e) { return ql_values_t<false>(n, d, e); }

// workspace of the shifted tridiagonal solves, allocated once per
// dense_tri_vectors call: the factorization of T - sigma I (rows u0, u1, u2 of
// U, multipliers l, row swaps sw) and, for the fast solves, U's reciprocal
// pivots
struct TriWork {
    std::vector<double> u0, u1, u2, l, rp;
    std::vector<char> sw;
    explicit TriWork(int64_t n) : u0(n), u1(n), u2(n), l(n), rp(n), sw(n) {}
};

// T - sigma I = P L U, T = tridiag(e, d, e), Gaussian elimination with partial
// pivoting; tiny pivots replaced by eps3
void tri_shift_factor(int64_t n, const double *d, const double *e, double sigma, double eps3, TriWork &w)
{
    double *u0 = w.u0.data(), *u1 = w.u1.data(), *u2 = w.u2.data(), *l = w.l.data();
    char *sw = w.sw.data();
    std::fill(sw, sw + n, 0);
    // row i of the working matrix: (a_i,i, a_i,i+1, a_i,i+2)
    double a0 = d[0] - sigma, a1 = n > 1 ? e[0] : 0.0, a2 = 0.0;
    for (int64_t i = 0; i + 1 < n; i++) {
        // next row: (e_i, d_{i+1} - sigma, e_{i+1})
        const double b0 = e[i], b1 = d[i + 1] - sigma, b2 = i + 2 < n ? e[i + 1] : 0.0;
        if (std::fabs(b0) > std::fabs(a0)) {      // swap rows
            sw[i] = 1;
            u0[i] = b0; u1[i] = b1; u2[i] = b2;
            const double m = a0 / b0;
            l[i] = m;
            a0 = a1 - m * b1;
            a1 = a2 - m * b2;
            a2 = 0.0;
        } else {
            if (a0 == 0.0) a0 = eps3;
            u0[i] = a0; u1[i] = a1; u2[i] = a2;
            const double m = b0 / a0;
            l[i] = m;
            a0 = b1 - m * a1;
            a1 = b2 - m * a2;
            a2 = 0.0;
        }
    }
    if (a0 == 0.0) a0 = eps3;
    u0[n - 1] = a0; u1[n - 1] = 0.0; u2[n - 1] = 0.0;
    for (int64_t i = 0; i < n; i++) {
        double piv = u0[i];
        if (std::fabs(piv) < eps3) piv = piv < 0 ? -eps3 : eps3;
        w.rp[i] = 1.0 / piv;
    }
}

// x <- (T - sigma I)^-1 x with the factorization in w; FAST: the back
// substitution multiplies by the reciprocal pivots instead of dividing
void tri_shift_apply(int64_t n, double eps3, double *x, const TriWork &w, bool fast)
{
    const double *u0 = w.u0.data(), *u1 = w.u1.data(), *u2 = w.u2.data(), *l = w.l.data();
    const char *sw = w.sw.data();
    for (int64_t i = 0; i + 1 < n; i++) {       // forward: apply the row operations to x
        if (sw[i]) {
            const double t = x[i];
            x[i] = x[i + 1];
            x[i + 1] = t - l[i] * x[i];
        } else {
            x[i + 1] -= l[i] * x[i];
        }
    }
    for (int64_t i = n - 1; i >= 0; i--) {      // back substitution
        double s = x[i];
        if (i + 1 < n) s -= u1[i] * x[i + 1];
        if (i + 2 < n) s -= u2[i] * x[i + 2];
        if (fast) {
            x[i] = s * w.rp[i];
        } else {
            double piv = u0[i];
            if (std::fabs(piv) < eps3) piv = piv < 0 ? -eps3 : eps3;
            x[i] = s / piv;
        }
    }
}

}  // namespace

int dense_tri_values_from(int64_t n, const double *d, const double *e, DenseTri &T)
{
    T.n = n;
    T.A.clear();
    T.d0.assign(d, d + n);
    T.e0.assign(e, e + n);
    T.w = T.d0;
    if (n <= 1) return 0;
    std::vector<double> e2 = T.e0;
    return ql_values(n, T.w.data(), e2.data());
}

int dense_tri_values(int64_t n, const double *Afull, DenseTri &T)
{
    T.n = n;
    T.A.assign(Afull, Afull + (size_t)n * n);
    T.w.assign(n, 0.0);
    if (n <= 0) return 0;
    if (n == 1) {
        T.w[0] = T.A[0];
        return 0;
    }
    for (int64_t c = 0; c < n; c++)
        for (int64_t r = c + 1; r < n; r++) T.A[r + c * n] = T.A[c + r * n];
    T.d0.assign(n, 0.0);
    T.e0.assign(n, 0.0);
    tridiag_inplace(n, T.A.data(), T.d0.data(), T.e0.data());
    std::vector<double> e = T.e0;
    T.w = T.d0;
    return ql_values(n, T.w.data(), e.data());
}

int dense_tri_values_rr(int64_t n, const double *Afull, DenseTri &T)
{
    T.n = n;
    T.A.assign(Afull, Afull + (size_t)n * n);
    T.w.assign(n, 0.0);
    T.iters = 3;
    T.fast = true;
    if (n <= 0) return 0;
    if (n == 1) {
        T.w[0] = T.A[0];
        return 0;
    }
    T.d0.assign(n, 0.0);
    T.e0.assign(n, 0.0);
    tridiag_lower(n, T.A.data(), T.d0.data(), T.e0.data());
    std::vector<double> e = T.e0;
    T.w = T.d0;
    return ql_values_t<true>(n, T.w.data(), e.data());
}

// the selected eigenvectors of the tridiagonal (d0, e0): X[t] for sel[t]
static int tri_vectors(const DenseTri &T, int64_t k, const int64_t *sel, std::vector<std::vector<double>> &X)
{
    const int64_t n = T.n;
    const std::vector<double> &d0 = T.d0, &e0 = T.e0, &w = T.w;
    double tnorm = 0.0;
    for (int64_t i = 0; i < n; i++)
        tnorm = std::max(tnorm, std::fabs(d0[i]) + (i > 0 ? std::fabs(e0[i - 1]) : 0.0) + std::fabs(e0[i]));
    if (tnorm == 0.0) tnorm = 1.0;
    const double eps3 = kEps * tnorm, gaptol = 1e-3 * tnorm;
    // selected eigenvalues in ascending order, so clusters are contiguous
    std::vector<int64_t> pos(k);
    std::iota(pos.begin(), pos.end(), 0);
    std::sort(pos.begin(), pos.end(), [&](int64_t a, int64_t b) { return sel[a] < sel[b]; });
    X.assign(k, std::vector<double>(n));
    TriWork tw(n);
    int64_t cstart = 0;
    double prev_sigma = 0.0;
    for (int64_t q = 0; q < k; q++) {
        const int64_t t = pos[q];
        const double lam = w[sel[t]];
        if (q > 0 && lam - w[sel[pos[q - 1]]] > gaptol) cstart = q;
        double sigma = lam;
        if (q > 0 && sigma <= prev_sigma) sigma = prev_sigma + eps3;   // separate equal shifts
        prev_sigma = sigma;
        std::vector<double> &x = X[t];
        uint64_t z = 0x9E3779B97F4A7C15ULL * (uint64_t)(q + 1);
        for (int64_t i = 0; i < n; i++) {          // deterministic start vector
            z ^= z >> 12; z ^= z << 25; z ^= z >> 27;
            x[i] = (double)((z * 0x2545F4914F6CDD1DULL) >> 11) * (1.0 / 9007199254740992.0) - 0.5;
        }
        // one factorization of T - sigma I for all of the vector's steps
        tri_shift_factor(n, d0.data(), e0.data(), sigma, eps3, tw);
        for (int it = 0; it < T.iters; it++) {
            tri_shift_apply(n, eps3, x.data(), tw, T.fast);
            for (int64_t p2 = cstart; p2 < q; p2++) {     // reorthogonalise within the cluster
                const std::vector<double> &y = X[pos[p2]];
                double dot = 0.0;
                for (int64_t i = 0; i < n; i++) dot += x[i] * y[i];
                for (int64_t i = 0; i < n; i++) x[i] -= dot * y[i];
            }
            double nr = 0.0;
            for (int64_t i = 0; i < n; i++) nr += x[i] * x[i];
            nr = std::sqrt(nr);
            if (nr == 0.0) return 2;
            for (int64_t i = 0; i < n; i++) x[i] /= nr;
        }
    }
    return 0;
}

int dense_tri_zvectors(const DenseTri &T, int64_t k, const int64_t *sel, double *Z, int64_t ldz)
{
    const int64_t n = T.n;
    if (k <= 0) return 0;
    if (n == 1) {
        Z[0] = 1.0;
        return 0;
    }
    std::vector<std::vector<double>> X;
    const int rc = tri_vectors(T, k, sel, X);
    if (rc) return rc;
    for (int64_t t = 0; t < k; t++) std::copy(X[t].begin(), X[t].end(), Z + (size_t)t * ldz);
    return 0;
}

int dense_tri_vectors(const DenseTri &T, int64_t k, const int64_t *sel, double *Vsel, int64_t ldv)
{
    const int64_t n = T.n;
    if (k <= 0) return 0;
    if (n == 1) {
        Vsel[0] = 1.0;
        return 0;
    }
    std::vector<std::vector<double>> X;
    const int rc = tri_vectors(T, k, sel, X);
    if (rc) return rc;
    // back-transformation: x <- H_0 H_1 ... H_{n-3} x for all k vectors at
    // once, the vectors interleaved (Xr[i k + t]) so that the loops over t
    // vectorize; each vector's dot product still sums over i in order
    std::vector<double> Xr((size_t)n * k), sv(k);
    for (int64_t t = 0; t < k; t++)
        for (int64_t i = 0; i < n; i++) Xr[(size_t)i * k + t] = X[t][i];
    for (int64_t kk = n - 3; kk >= 0; kk--) {
        const double *v = T.A.data() + kk * n;
        std::fill(sv.begin(), sv.end(), 0.0);
        for (int64_t i = kk + 1; i < n; i++) {
            const double vi = v[i];
            const double *xr = Xr.data() + (size_t)i * k;
            for (int64_t t = 0; t < k; t++) sv[t] += vi * xr[t];
        }
        for (int64_t t = 0; t < k; t++) sv[t] *= 2.0;
        for (int64_t i = kk + 1; i < n; i++) {
            const double vi = v[i];
            double *xr = Xr.data() + (size_t)i * k;
            for (int64_t t = 0; t < k; t++)
                xr[t] -= sv[t] * vi;     // exact for sv[t] = 0
        }
    }
    for (int64_t t = 0; t < k; t++)
        for (int64_t i = 0; i < n; i++) Vsel[(size_t)t * ldv + i] = Xr[(size_t)i * k + t];
    return 0;
}

#define UB(i, j) AB[kd + (i) - (j) + (size_t)(j) * ld]
int band_chol_u(int64_t n, int64_t kd, double *AB, int64_t ld)
{
    for (int64_t j = 0; j < n; j++) {
        const int64_t i0 = std::max<int64_t>(0, j - kd);
        for (int64_t i = i0; i < j; i++) {
            double s = UB(i, j);
            for (int64_t p = i0; p < i; p++) s -= UB(p, i) * UB(p, j);
            UB(i, j) = s / UB(i, i);
        }
        double dd = UB(j, j);
        for (int64_t p = i0; p < j; p++) dd -= UB(p, j) * UB(p, j);
        if (!(dd > 0.0)) return (int)(j + 1);
        UB(j, j) = std::sqrt(dd);
    }
    return 0;
}

void band_solve_u(int64_t n, int64_t kd, const double *AB, int64_t ld, double *b)
{
    for (int64_t j = 0; j < n; j++) {
        const int64_t i0 = std::max<int64_t>(0, j - kd);
        double s = b[j];
        for (int64_t i = i0; i < j; i++) s -= UB(i, j) * b[i];
        b[j] = s / UB(j, j);
    }
    for (int64_t j = n - 1; j >= 0; j--) {
        b[j] /= UB(j, j);
        const int64_t i0 = std::max<int64_t>(0, j - kd);
        const double bj = b[j];
        for (int64_t i = i0; i < j; i++) b[i] -= UB(i, j) * bj;
    }
}
#undef UB

int lu_factor(int64_t n, double *A, int64_t *piv)
{
    for (int64_t k = 0; k < n; k++) {
        int64_t p = k;
        double best = std::fabs(A[k + k * n]);
        for (int64_t i = k + 1; i < n; i++)
            if (std::fabs(A[i + k * n]) > best) { best = std::fabs(A[i + k * n]); p = i; }
        piv[k] = p;
        if (best == 0.0) return (int)(k + 1);
        if (p != k)
            for (int64_t j = 0; j < n; j++) std::swap(A[k + j * n], A[p + j * n]);
        const double inv = 1.0 / A[k + k * n];
        for (int64_t i = k + 1; i < n; i++) A[i + k * n] *= inv;
        for (int64_t j = k + 1; j < n; j++) {
            const double a = A[k + j * n];
            if (a == 0.0) continue;
            for (int64_t i = k + 1; i < n; i++) A[i + j * n] -= A[i + k * n] * a;
        }
    }
    return 0;
}

void lu_solve(int64_t n, const double *LU, const int64_t *piv, double *b)
{
    for (int64_t k = 0; k < n; k++)
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
    for (int64_t j = 0; j < n; j++) {
        const double bj = b[j];
        if (bj != 0.0)
            for (int64_t i = j + 1; i < n; i++) b[i] -= LU[i + j * n] * bj;
    }
    for (int64_t j = n - 1; j >= 0; j--) {
        b[j] /= LU[j + j * n];
        const double bj = b[j];
        if (bj != 0.0)
            for (int64_t i = 0; i < j; i++) b[i] -= LU[i + j * n] * bj;
    }
}

}  // namespace sehip
