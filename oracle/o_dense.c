/*
 * o_dense.c -- dense kernels of the CPU oracle (TEST INFRASTRUCTURE ONLY).
 *
 * These stand in for the LAPACK routines the reference calls:
 *   oracle_syev    for dsyev_("V","U")   (src/speigs_core.c:180, 236, 332-335)
 *   oracle_pbtrf_u for dpbtrf_("U")      (src/speigs_core.c:904, 911)
 *   oracle_pbtrs_u for dpbtrs_("U")      (src/speigs_core.c:106)
 *   oracle_lu_*    for the UMFPACK LU fallback (src/speigs_core.c:974-984, 111-115),
 *                  dense, so only for small n.
 * Methods: Householder reduction to tridiagonal form followed by the implicit
 * QL iteration with Wilkinson-type shifts (Bowdler, Martin, Reinsch and
 * Wilkinson 1968); Gaussian elimination with partial pivoting; column-wise
 * banded Cholesky A = U^T U.
 * Pinned in tests/test_oracle.py against numpy.linalg.eigh / solve and against
 * closed-form tridiagonal spectra.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* A (m x m, column-major, upper triangle read) -> Q T Q^T.
 * d: diagonal of T, e[i] = T(i+1,i), Q overwrites A. */
static void householder_tridiag(int64_t m, double *A, double *d, double *e)
{
    /* symmetrise from the upper triangle */
    for (int64_t c = 0; c < m; c++)
        for (int64_t r = c + 1; r < m; r++) A[r + c * m] = A[c + r * m];
    double *Q = (double *)calloc((size_t)m * m, sizeof(double));
    double *v = (double *)calloc((size_t)m, sizeof(double));
    double *p = (double *)calloc((size_t)m, sizeof(double));
    for (int64_t i = 0; i < m; i++) Q[i + i * m] = 1.0;

    for (int64_t k = 0; k + 2 < m; k++) {
        /* x = A[k+1:, k] */
        double xn = 0.0;
        for (int64_t i = k + 1; i < m; i++) xn += A[i + k * m] * A[i + k * m];
        xn = sqrt(xn);
        double x0 = A[(k + 1) + k * m];
        if (xn == 0.0) { e[k] = 0.0; continue; }
        double alpha = (x0 > 0) ? -xn : xn;
        for (int64_t i = k + 1; i < m; i++) v[i] = A[i + k * m];
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int64_t i = k + 1; i < m; i++) vn += v[i] * v[i];
        vn = sqrt(vn);
        if (vn == 0.0) { e[k] = x0; continue; }
        for (int64_t i = k + 1; i < m; i++) v[i] /= vn;
        /* p = A22 v ; K = v.p ; w = 2p - 2K v ; A22 -= v w^T + w v^T */
        double K = 0.0;
        for (int64_t i = k + 1; i < m; i++) {
            double s = 0.0;
            for (int64_t j = k + 1; j < m; j++) s += A[i + j * m] * v[j];
            p[i] = s;
            K += v[i] * s;
        }
        for (int64_t i = k + 1; i < m; i++) p[i] = 2.0 * p[i] - 2.0 * K * v[i];
        for (int64_t j = k + 1; j < m; j++)
            for (int64_t i = k + 1; i < m; i++)
                A[i + j * m] -= v[i] * p[j] + p[i] * v[j];
        e[k] = alpha;
        for (int64_t i = k + 2; i < m; i++) { A[i + k * m] = 0.0; A[k + i * m] = 0.0; }
        A[(k + 1) + k * m] = alpha; A[k + (k + 1) * m] = alpha;
        /* Q <- Q H,  H = I - 2 v v^T on indices k+1.. */
        for (int64_t r = 0; r < m; r++) {
            double s = 0.0;
            for (int64_t i = k + 1; i < m; i++) s += Q[r + i * m] * v[i];
            s *= 2.0;
            for (int64_t i = k + 1; i < m; i++) Q[r + i * m] -= s * v[i];
        }
    }
    for (int64_t i = 0; i < m; i++) d[i] = A[i + i * m];
    if (m >= 2) e[m - 2] = A[(m - 1) + (m - 2) * m];
    e[m - 1] = 0.0;
    memcpy(A, Q, (size_t)m * m * sizeof(double));
    free(Q); free(v); free(p);
}

/* implicit QL on (d, e), rotations accumulated into the columns of Z */
static int tridiag_ql(int64_t m, double *d, double *e, double *Z)
{
    const double eps = 2.220446049250313e-16;
    for (int64_t l = 0; l < m; l++) {
        int iter = 0;
        int64_t mm;
        for (;;) {
            for (mm = l; mm < m - 1; mm++) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= eps * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) return (int)(l + 1);
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int64_t i;
            int underflow = 0;
            for (i = mm - 1; i >= l; i--) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= p; e[mm] = 0.0; underflow = 1; break; }
                s = f / r; c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                for (int64_t q = 0; q < m; q++) {
                    double zf = Z[q + (i + 1) * m];
                    Z[q + (i + 1) * m] = s * Z[q + i * m] + c * zf;
                    Z[q + i * m]       = c * Z[q + i * m] - s * zf;
                }
            }
            if (underflow) continue;
            d[l] -= p; e[l] = g; e[mm] = 0.0;
        }
    }
    return 0;
}

int oracle_syev(int64_t m, double *A, double *w)
{
    if (m <= 0) return 0;
    if (m == 1) { w[0] = A[0]; A[0] = 1.0; return 0; }
    double *e = (double *)calloc((size_t)m, sizeof(double));
    householder_tridiag(m, A, w, e);
    int info = tridiag_ql(m, w, e, A);
    free(e);
    if (info) return info;
    /* ascending order, like dsyev */
    for (int64_t i = 0; i < m - 1; i++) {
        int64_t kmin = i;
        for (int64_t j = i + 1; j < m; j++) if (w[j] < w[kmin]) kmin = j;
        if (kmin != i) {
            double t = w[i]; w[i] = w[kmin]; w[kmin] = t;
            for (int64_t q = 0; q < m; q++) {
                double z = A[q + i * m]; A[q + i * m] = A[q + kmin * m]; A[q + kmin * m] = z;
            }
        }
    }
    return 0;
}

int oracle_lu_factor(int64_t n, double *A, int64_t *piv)
{
    for (int64_t k = 0; k < n; k++) {
        int64_t p = k; double best = fabs(A[k + k * n]);
        for (int64_t i = k + 1; i < n; i++)
            if (fabs(A[i + k * n]) > best) { best = fabs(A[i + k * n]); p = i; }
        piv[k] = p;
        if (best == 0.0) return (int)(k + 1);
        if (p != k)
            for (int64_t j = 0; j < n; j++) {
                double t = A[k + j * n]; A[k + j * n] = A[p + j * n]; A[p + j * n] = t;
            }
        double inv = 1.0 / A[k + k * n];
        for (int64_t i = k + 1; i < n; i++) A[i + k * n] *= inv;
        for (int64_t j = k + 1; j < n; j++) {
            double a = A[k + j * n];
            if (a == 0.0) continue;
            for (int64_t i = k + 1; i < n; i++) A[i + j * n] -= A[i + k * n] * a;
        }
    }
    return 0;
}

void oracle_lu_solve(int64_t n, const double *LU, const int64_t *piv, double *b)
{
    for (int64_t k = 0; k < n; k++) {
        int64_t p = piv[k];
        if (p != k) { double t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int64_t j = 0; j < n; j++) {
        double bj = b[j];
        if (bj != 0.0) for (int64_t i = j + 1; i < n; i++) b[i] -= LU[i + j * n] * bj;
    }
    for (int64_t j = n - 1; j >= 0; j--) {
        b[j] /= LU[j + j * n];
        double bj = b[j];
        if (bj != 0.0) for (int64_t i = 0; i < j; i++) b[i] -= LU[i + j * n] * bj;
    }
}

/* A(i,j), max(0,j-kd) <= i <= j, lives at AB[kd + i - j + j*ldab]; A = U^T U */
int oracle_pbtrf_u(int64_t n, int64_t kd, double *AB, int64_t ldab)
{
#define UB(i, j) AB[kd + (i) - (j) + (size_t)(j) * ldab]
    for (int64_t j = 0; j < n; j++) {
        int64_t i0 = j - kd > 0 ? j - kd : 0;
        for (int64_t i = i0; i < j; i++) {
            double s = UB(i, j);
            int64_t p0 = (i - kd > i0) ? i - kd : i0;
            for (int64_t p = p0; p < i; p++) s -= UB(p, i) * UB(p, j);
            UB(i, j) = s / UB(i, i);
        }
        double d = UB(j, j);
        for (int64_t p = i0; p < j; p++) d -= UB(p, j) * UB(p, j);
        if (!(d > 0.0)) return (int)(j + 1);
        UB(j, j) = sqrt(d);
    }
    return 0;
}

void oracle_pbtrs_u(int64_t n, int64_t kd, const double *AB, int64_t ldab, double *b)
{
    /* U^T y = b */
    for (int64_t j = 0; j < n; j++) {
        int64_t i0 = j - kd > 0 ? j - kd : 0;
        double s = b[j];
        for (int64_t i = i0; i < j; i++) s -= UB(i, j) * b[i];
        b[j] = s / UB(j, j);
    }
    /* U x = y */
    for (int64_t j = n - 1; j >= 0; j--) {
        b[j] /= UB(j, j);
        int64_t i0 = j - kd > 0 ? j - kd : 0;
        double bj = b[j];
        for (int64_t i = i0; i < j; i++) b[i] -= UB(i, j) * bj;
    }
#undef UB
}

int oracle_syev_jacobi(int64_t m, double *A, double *w)
{
    if (m <= 0) return 0;
    double *a = (double *)malloc((size_t)m * m * sizeof(double));
    for (int64_t c = 0; c < m; c++)
        for (int64_t r = 0; r < m; r++)
            a[r + c * m] = (r <= c) ? A[r + c * m] : A[c + r * m];
    for (int64_t c = 0; c < m; c++)
        for (int64_t r = 0; r < m; r++) A[r + c * m] = (r == c) ? 1.0 : 0.0;
    const double eps = 2.220446049250313e-16;
    int sweep;
    for (sweep = 0; sweep < 100; sweep++) {
        int rotated = 0;
        for (int64_t p = 0; p < m - 1; p++)
            for (int64_t q = p + 1; q < m; q++) {
                double apq = a[p + q * m];
                double app = a[p + p * m], aqq = a[q + q * m];
                if (fabs(apq) <= eps * sqrt(fabs(app * aqq)) || apq == 0.0) {
                    continue;
                }
                rotated = 1;
                double th = (aqq - app) / (2.0 * apq);
                double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int64_t r = 0; r < m; r++) {
                    double arp = a[r + p * m], arq = a[r + q * m];
                    a[r + p * m] = c * arp - s * arq;
                    a[r + q * m] = s * arp + c * arq;
                }
                for (int64_t r = 0; r < m; r++) {
                    double apr = a[p + r * m], aqr = a[q + r * m];
                    a[p + r * m] = c * apr - s * aqr;
                    a[q + r * m] = s * apr + c * aqr;
                }
                a[p + p * m] = app - t * apq;
                a[q + q * m] = aqq + t * apq;
                a[p + q * m] = 0.0; a[q + p * m] = 0.0;
                for (int64_t r = 0; r < m; r++) {
                    double vrp = A[r + p * m], vrq = A[r + q * m];
                    A[r + p * m] = c * vrp - s * vrq;
                    A[r + q * m] = s * vrp + c * vrq;
                }
            }
        if (!rotated) break;
    }
    for (int64_t i = 0; i < m; i++) w[i] = a[i + i * m];
    free(a);
    if (sweep >= 100) return 1;
    for (int64_t i = 0; i < m - 1; i++) {
        int64_t kmin = i;
        for (int64_t j = i + 1; j < m; j++) if (w[j] < w[kmin]) kmin = j;
        if (kmin != i) {
            double t = w[i]; w[i] = w[kmin]; w[kmin] = t;
            for (int64_t q = 0; q < m; q++) {
                double z = A[q + i * m]; A[q + i * m] = A[q + kmin * m]; A[q + kmin * m] = z;
            }
        }
    }
    return 0;
}
