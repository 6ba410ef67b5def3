/*
 * speigs.h -- public C API of the MI355X build of speigs: the k smallest ('sm')
 * or largest ('lm') magnitude eigenpairs of a sparse real symmetric (or complex
 * Hermitian) matrix given in compressed-sparse-column form.
 *
 * This header is ABI-identical to the reference interface it replaces
 * (include/speigs.h:21-107 of moatas19m/sparse-eigs-mex): the same index type,
 * enums, struct layouts, entry points and error codes, so the MATLAB MEX
 * gateway (src/speigs_mex.c:153-155) and other existing callers link against
 * libspeigs_hip.so unchanged.  The struct layouts are never extended here:
 * MI355X-specific knobs and telemetry live in speigs_hip.h.
 */
#ifndef SPEIGS_H
#define SPEIGS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 64-bit indices throughout (MATLAB mwIndex with -largeArrayDims).
 * Reference: include/speigs.h:21. */
typedef int64_t speigs_int;

/* Which end of the spectrum, by magnitude.  Reference: include/speigs.h:23-26. */
typedef enum {
    SPEIGS_SM = 0,   /* smallest |lambda|: shift-invert Lanczos over a Cholesky factor */
    SPEIGS_LM = 1    /* largest  |lambda|: Lanczos directly on A (SpMV operator)      */
} speigs_which;

/* The route the structure analysis took.  Reference: include/speigs.h:30-38. */
typedef enum {
    SPEIGS_PATH_DIAG      = 0,  /* diagonal matrix, eigenvalues read off the diagonal */
    SPEIGS_PATH_DENSE     = 1,  /* small n or large k/n: dense symmetric eigensolver */
    SPEIGS_PATH_SHIFTINV  = 2,  /* supernodal Cholesky shift-invert Lanczos (GPU)    */
    SPEIGS_PATH_DIRECT    = 3,  /* Lanczos on A itself, 'lm' (GPU SpMV)              */
    SPEIGS_PATH_BLOCKS    = 4,  /* split into connected components                   */
    SPEIGS_PATH_LU        = 5,  /* indefinite matrix: LU shift-invert                */
    SPEIGS_PATH_BANDED    = 6   /* narrow band after RCM: banded Cholesky            */
} speigs_path;

/* Solver options.  Field order and types are fixed by the reference
 * (include/speigs.h:40-55); zero means "use the default" as noted. */
typedef struct {
    double      tol;        /* converged when ||Av - lambda v||_2 <= tol*max(||A||_1,1); 0 -> 1e-14 */
    speigs_int  maxit;      /* restart cap; 0 -> 300                                          */
    speigs_int  ncv;        /* Krylov basis size; 0 -> max(2k+1, 20)                          */
    int         dense_max;  /* dense path when n <= dense_max; -1 -> 800, 0 -> never          */
    int         detect;     /* 1 -> structure analysis (diag / components), 0 -> skip         */
    uint64_t    seed;       /* start-vector seed; 0 -> 88172645463325252                      */
    int         check_sym;  /* 1 -> reject non-symmetric input (O(nnz) check), 0 -> trust it */
    int         band_max;   /* banded path if RCM bandwidth <= band_max; 0 -> 512, <0 -> off  */
    int         ordering;   /* 0 auto by size, 1 minimum-degree class, 2 nested dissection   */
    double      shift0;     /* first shift for a semi-definite A, relative to ||A||_1; 0 -> 1e-6 */
} speigs_opts;

/* Telemetry.  Reference: include/speigs.h:57-68. */
typedef struct {
    int         flag;       /* 0 all wanted pairs converged, 1 partial        */
    speigs_int  nconv;
    speigs_int  restarts;
    speigs_int  nops;       /* operator applications (solves or products)    */
    double      shift;      /* delta used in A + delta*I                     */
    double      anorm;      /* ||A||_1                                       */
    speigs_path path;
    speigs_int  ncomp;      /* connected components found                    */
    double      t_analyze, t_factor, t_iter;
    double      t_op, t_ortho, t_resid, t_ritz;
} speigs_info;

/* Reference: src/speigs_core.c:635-639. */
void speigs_default_opts(speigs_opts *o);

/* A is n-by-n CSC with sorted row indices per column; full symmetric storage
 * or the upper triangle only (only the upper triangle is read).
 *   lambda[k]  eigenvalues sorted by |lambda| (ascending 'sm', descending 'lm')
 *   V[n*k]     eigenvectors, column-major; NULL skips them
 *   resid[k]   ||A v - lambda v||_2; NULL skips them
 * Returns SPEIGS_OK or a negative SPEIGS_ERR_* code; partial convergence is
 * SPEIGS_OK with info->flag == 1.  Reference: include/speigs.h:70-83,
 * src/speigs_core.c:783-1024. */
int speigs(speigs_int n,
           const speigs_int *Ap, const speigs_int *Ai, const double *Ax,
           speigs_int k, speigs_which which, const speigs_opts *opts,
           double *lambda, double *V, double *resid, speigs_info *info);

/* Complex Hermitian A = Axr + i*Axi (split storage).  Axi == NULL is the real
 * case.  Reference: include/speigs.h:85-97, src/speigs_core.c:660-768. */
int speigs_z(speigs_int n,
             const speigs_int *Ap, const speigs_int *Ai,
             const double *Axr, const double *Axi,
             speigs_int k, speigs_which which, const speigs_opts *opts,
             double *lambda, double *Vr, double *Vi, double *resid,
             speigs_info *info);

/* Reference: src/speigs_core.c:770-781. */
const char *speigs_errmsg(int code);

/* Reference: include/speigs.h:101-107. */
#define SPEIGS_OK             0
#define SPEIGS_ERR_ARG      (-1)
#define SPEIGS_ERR_MEM      (-2)
#define SPEIGS_ERR_FACTOR   (-3)
#define SPEIGS_ERR_LAPACK   (-4)
#define SPEIGS_ERR_INTERNAL (-5)
#define SPEIGS_ERR_NOTSYM   (-6)

#ifdef __cplusplus
}
#endif
#endif /* SPEIGS_H */
