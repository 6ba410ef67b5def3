// factor.h -- supernodal (multifrontal) Cholesky of P (A + delta I) P^T on the
// device, and the triangular solves that apply (A + delta I)^-1.
// Replaces the reference's CHOLMOD calls cholmod_l_analyze / factorize_p /
// solve2 (src/speigs_core.c:933-970, :86-100).
#pragma once
#include <cstdint>
#include <memory>
#include <new>
#include <type_traits>
#include <utility>
#include <vector>

#include "dev.h"

namespace sehip {

// ------------------------------------------------------------ symbolic
// Nested dissection by breadth-first level structures; every separator (and
// every leaf region) is one supernode whose columns are contiguous in the new
// order and whose row structure below the diagonal block is the boundary of
// its region -- the set of later (ancestor-separator) nodes adjacent to it.
// Allocator that leaves elements default-initialised (no zero fill): for large
// index arrays that are written in parallel right after they are sized.
template <class T>
struct NoInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAlloc<U>;
    };
    NoInitAlloc() = default;
    template <class U>
    NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
    template <class U>
    void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value)
    {
        ::new ((void *)p) U;
    }
    template <class U, class... Args>
    void construct(U *p, Args &&...args)
    {
        ::new ((void *)p) U(std::forward<Args>(args)...);
    }
};

struct Symbolic {
    int64_t n = 0;
    std::vector<int64_t> perm, iperm;          // perm[new] = old
    int64_t ns = 0;                            // supernodes, in postorder
    std::vector<int64_t> c0, c1;               // column range (new indices)
    std::vector<int64_t> parent;               // -1 for roots
    std::vector<int64_t> chptr, chl;           // children lists (CSR)
    std::vector<int64_t> rptr;                 // boundary rows, sorted, CSR
    std::vector<int64_t, NoInitAlloc<int64_t>> rows;
    std::vector<int32_t> level;                // 0 = leaves
    int32_t nlevels = 0;
    std::vector<int64_t> F;                    // front order c + r
    std::vector<int64_t> foff;                 // front offset (elements) in the pool
    int64_t pool = 0;
    std::vector<int64_t> relptr;               // per supernode: offset of its rel map
    std::vector<int32_t, NoInitAlloc<int32_t>> rel;   // positions of its rows in the parent front
    double flops = 0.0;
    int64_t nnzL = 0;
    double t_order = 0.0, t_symbolic = 0.0;
};

struct SymbolicOptions {
    int64_t leaf = 64;          // regions of at most this many nodes become leaf fronts
};

int symbolic_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, const SymbolicOptions &o, Symbolic *S);

// ------------------------------------------------------------ numeric + solve
struct FactorOptions {
    int precision = SPEIGS_HIP_FP64;
    double ir_tol = 1e-13;
    int ir_maxit = 12;
    // FP64 solves by the inverted diagonal blocks (solve mode 3): refine an
    // apply against A until ||b - (A + delta I) y|| <= apply_tol ||b|| (or the
    // residual stops falling); the solve's own error grows with cond(L11)
    double apply_tol = 1e-16;
    int ordering = 0;
    int use_graph = 0;
};

struct Factor;
int factor_create(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, const FactorOptions &o,
                  hipStream_t st, Factor **F, speigs_hip_info *hinfo);
// 0 ok, 1 not positive definite, <0 error
int factor_numeric(Factor *F, double delta);
// The indefinite fallback: A = L S L^T (S = diag(+-1)) without pivoting, a pivot
// of magnitude below tau replaced by +-tau; factor_apply then refines every
// solve against A.  0 ok (factor_perturbed() pivots replaced), <0 error.
int factor_numeric_signed(Factor *F, double tau);
int factor_perturbed(const Factor *F);
void factor_attach_matrix(Factor *F, const SellA *A);
// y = (A + delta I)^-1 x on device vectors (FP64 in/out).  A solve whose
// bounded waits gave up is reported by factor_check once the factor's stream
// has been synchronized after it (and by the next factor_apply).
int factor_apply(Factor *F, const double *x, double *y);
int factor_check(const Factor *F);
// true when factor_apply allocates nothing: the single-vector solves, whose
// workspace is allocated with the factor (not the inverse-based debug mode)
bool factor_apply_preallocated(const Factor *F);
void factor_stats(const Factor *F, speigs_hip_info *hinfo);
void factor_destroy(Factor *F);
const Symbolic *factor_symbolic(const Factor *F);

}  // namespace sehip
