/*
 * oracle.h -- CPU oracle for the speigs hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (moatas19m/sparse-eigs-mex src/speigs_core.c) used as the checker
 * for the MI355X build.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * The reference's third-party numerics (CHOLMOD, METIS/AMD, UMFPACK, LAPACK,
 * BLAS) are absent from this image, so the oracle carries its own:
 *   - o_dense.c : symmetric eigensolver (Householder tridiagonalisation +
 *                 implicit QL), LU with partial pivoting, banded Cholesky;
 *   - o_chol.c  : nested-dissection ordering by BFS level sets, elimination
 *                 tree, up-looking sparse Cholesky and its solve;
 *   - o_speigs.c: the driver, Lanczos loop and short-circuits.
 * Pinning of each function is listed in o_speigs.c's header.
 */
#ifndef SPEIGS_ORACLE_H
#define SPEIGS_ORACLE_H

#include <stdint.h>
#include "../include/speigs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* same contract as speigs()/speigs_z()/speigs_default_opts() */
void oracle_default_opts(speigs_opts *o);
int  oracle_speigs(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                   const double *Ax, speigs_int k, speigs_which which,
                   const speigs_opts *opts, double *lambda, double *V,
                   double *resid, speigs_info *info);
int  oracle_speigs_z(speigs_int n, const speigs_int *Ap, const speigs_int *Ai,
                     const double *Axr, const double *Axi, speigs_int k,
                     speigs_which which, const speigs_opts *opts,
                     double *lambda, double *Vr, double *Vi, double *resid,
                     speigs_info *info);

/* xorshift64* and the uniform start vector (reference speigs_core.c:49-54) */
uint64_t oracle_rng_next(uint64_t *s);
void     oracle_rand_vec(uint64_t *s, int64_t n, double *v);

/* dense symmetric eigenproblem from the UPPER triangle of column-major A (m x m,
 * lda m); on return A holds orthonormal eigenvectors, w ascending eigenvalues.
 * Returns 0, or >0 if the QL iteration failed to converge. */
int oracle_syev(int64_t m, double *A, double *w);

/* same contract, cyclic two-sided Jacobi: eigenvalues and vectors of a
 * positive definite A to relative accuracy (Demmel and Veselic 1992) -- used
 * for the small projected Lanczos problem, whose largest Ritz value 1/delta can
 * exceed the wanted ones by 1e4-1e7 */
int oracle_syev_jacobi(int64_t m, double *A, double *w);

/* y = A x where A is symmetric and only its upper triangle is read */
void oracle_spmv_upper(int64_t n, const int64_t *Ap, const int64_t *Ai,
                       const double *Ax, const double *x, double *y);

/* sparse Cholesky of P (A + delta I) P^T, reading the upper triangle of A */
typedef struct oracle_chol oracle_chol;
/* status: 0 ok, 1 not positive definite, <0 out of memory */
oracle_chol *oracle_chol_factor(int64_t n, const int64_t *Ap, const int64_t *Ai,
                                const double *Ax, double delta, int *status);
int     oracle_chol_refactor(oracle_chol *F, const int64_t *Ap, const int64_t *Ai,
                             const double *Ax, double delta);
void    oracle_chol_solve(const oracle_chol *F, const double *b, double *x);
int64_t oracle_chol_nnzL(const oracle_chol *F);
double  oracle_chol_flops(const oracle_chol *F);
const int64_t *oracle_chol_perm(const oracle_chol *F);
void    oracle_chol_free(oracle_chol *F);

/* nested-dissection ordering (perm[new] = old) of the graph of A (upper or full) */
int oracle_nd_order(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *perm);

/* reverse Cuthill-McKee of the banded gate, the reference's seed scan kept (O(n^2)
 * on upper-triangle storage) */
void oracle_rcm(int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *perm);

/* dense helpers used by the short-circuit paths */
int  oracle_lu_factor(int64_t n, double *A, int64_t *piv);         /* 0 ok, >0 singular */
void oracle_lu_solve(int64_t n, const double *LU, const int64_t *piv, double *b);
int  oracle_pbtrf_u(int64_t n, int64_t kd, double *AB, int64_t ldab); /* 0 ok, >0 not PD */
void oracle_pbtrs_u(int64_t n, int64_t kd, const double *AB, int64_t ldab, double *b);

#ifdef __cplusplus
}
#endif
#endif
