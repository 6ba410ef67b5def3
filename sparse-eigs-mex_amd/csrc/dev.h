// dev.h -- internal C++ interface of libspeigs_hip (not installed; the public
// ABI is include/speigs.h + include/speigs_hip.h).
//
// Layout in HBM (one speigs() call, n rows):
//   * A as SELL-64: rows in slices of 64, each slice stored column-major
//     (entry k of all 64 rows contiguous), FP64 values + int32 columns; the
//     full symmetric matrix is expanded from the stored upper triangle.
//   * Krylov basis V and restart buffer Vn: n x (ncv+1) FP64, column-major,
//     leading dimension padded to 32 doubles (256 B) so every column starts on
//     a 256-byte boundary.
//   * supernodal factor: dense fronts (column-major F x F) in one pool, FP64
//     or FP32; the first c columns of a front are the factor panel read by the
//     triangular solves.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>

#include "../../include/speigs.h"
#include "../../include/speigs_hip.h"

namespace sehip {

// Every internal failure records where it happened (source file and line, the
// failing call and, for HIP calls, the runtime's error name) in a per-thread
// record that speigs_hip_last_error() returns; the error code stays the
// reference's (SPEIGS_ERR_*).
#define SEHIP_CHECK(expr)                                                     \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) return sehip::hip_fail(e_, __FILE__, __LINE__, #expr); \
    } while (0)
#define SEHIP_FAIL(code, what) sehip::fail((code), __FILE__, __LINE__, (what))

int hip_error_code(hipError_t e);
// records the site and returns the SPEIGS_ERR_* code of e (clears HIP's last error)
int hip_fail(hipError_t e, const char *file, int line, const char *call);
// records the site and returns code
int fail(int code, const char *file, int line, const char *what);
// the calling thread's record ("" if none); a helper thread's record is
// handed to the thread that joins it with note_error()
const char *last_error_text();
void note_error(const char *text);
void clear_error();
double wall();
// SPEIGS_HIP_TRACE=1: phase times of the set-up on stderr; trace_phase prints
// the time since t and resets t to now
bool trace_on();
void trace_phase(const char *what, double &t);

constexpr int kBlock = 256;          // 4 wave64s
constexpr int kRowsPerThread = 4;    // reduction kernels: rows per thread
constexpr int64_t kRowsPerBlock = (int64_t)kBlock * kRowsPerThread;

inline int64_t pad_ld(int64_t n) { return (n + 31) / 32 * 32; }
inline int64_t nblocks_rows(int64_t n) { return (n + kRowsPerBlock - 1) / kRowsPerBlock; }

// ------------------------------------------------------------ device memory
// Device buffers come from a per-device cache of freed blocks (devcache.cpp):
// dev_alloc reuses a kept block of the request's size class or allocates one;
// dev_release waits for the blocks' devices to finish their queued work and
// keeps the blocks for later requests (up to a cap per device).
struct DevBlock {
    void *p;
    size_t bytes;
    int dev;
};
int dev_alloc(void **p, size_t bytes, int *dev, size_t *bytes_out);
void dev_release(std::vector<DevBlock> &blocks);

struct DevMem {
    std::vector<DevBlock> blocks;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { release(); }
    template <class T> int alloc(T **p, size_t count) {
        void *q = nullptr;
        int dev = 0;
        size_t got = 0;
        if (count == 0) count = 1;
        const int rc = dev_alloc(&q, count * sizeof(T), &dev, &got);
        if (rc) { *p = nullptr; return rc; }
        blocks.push_back({q, got, dev});
        *p = static_cast<T *>(q);
        return 0;
    }
    void release() { dev_release(blocks); }
};

// ------------------------------------------------------------ host <-> device copies
// Copies between pageable host memory and the device through pinned staging
// buffers, several host threads in parallel (hostcopy.cpp).  copy_to_host
// returns once dst holds the data (the copy is ordered after the work queued
// on st); copy_to_device returns once src may be reused, the data reaching the
// device in stream order on st.
int copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st);
int copy_to_host_2d(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t rows,
                    hipStream_t st);
int copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t st);

// Writes every page of a host buffer once (each byte written is its own value)
// on helper threads, so that a later copy into memory the caller has just
// allocated does not stop for page faults; join() (or the destructor) waits.
class HostTouch {
public:
    HostTouch() = default;
    HostTouch(const HostTouch &) = delete;
    HostTouch &operator=(const HostTouch &) = delete;
    ~HostTouch() { join(); }
    void start(void *p, size_t bytes);
    void join();

private:
    std::vector<std::thread> th_;
};

// ------------------------------------------------------------ SELL matrix
struct SellA {
    int64_t n = 0, nslices = 0, stored = 0, nnz_full = 0;
    int64_t *sptr = nullptr;   // slice start (entries)
    int32_t *swid = nullptr;   // slice width
    int32_t *col = nullptr;
    double *val = nullptr;
};
// expand the upper triangle of CSC A to full symmetric SELL-64 and upload
int sell_build(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax,
               SellA *S, DevMem &mem, hipStream_t st);
void spmv(const SellA &S, const double *x, double *y, hipStream_t st);
void spmv_multi(const SellA &S, const double *X, int64_t ldx, double *Y, int64_t ldy, int nv,
                hipStream_t st);
double spmv_bytes(const SellA &S);

// ------------------------------------------------------------ BLAS-1/2 on tall V
// partial[b*m + c] = sum over row block b of V[:,c] . w
void coldots_partial(const double *V, int64_t ldv, int m, const double *w, int64_t n,
                     double *partial, hipStream_t st);
// out[c] = sum_b partial[b*m + c] (fixed order); nb = number of row blocks
void reduce_partial(const double *partial, int64_t nb, int m, double *out, hipStream_t st);
// w -= V h; then partial dots V^T w (mode 1) or partial |w|^2 (mode 2) or nothing (0)
void cgs_update(const double *V, int64_t ldv, int m, const double *h, double *w, int64_t n,
                double *partial, int mode, hipStream_t st);
// X[:, t] = V[:, 0:m] S[:, t], t < p (S column-major m x p, device)
void tsgemm(const double *V, int64_t ldv, int m, const double *S, int p, double *X, int64_t ldx,
            int64_t n, hipStream_t st);
// partial[b*p + t] = sum X[:,t] . Y[:,t]
void pairdots_partial(const double *X, int64_t ldx, const double *Y, int64_t ldy, int p, int64_t n,
                      double *partial, hipStream_t st);
// partial[b*p + t] = sum (Y[:,t] - lam[t] X[:,t])^2
void residual_partial(const double *X, int64_t ldx, const double *Y, int64_t ldy, const double *lam,
                      int p, int64_t n, double *partial, hipStream_t st);
// X[:,t] *= s[t]
void scale_cols(double *X, int64_t ldx, const double *s, int p, int64_t n, hipStream_t st);
// y = a x
void scale_copy(const double *x, double a, double *y, int64_t n, hipStream_t st);
// y = b - A x - delta x  (FP64 residual of the shifted system)
void axpby_residual(const double *b, const double *Ax, const double *x, double delta, double *r,
                    int64_t n, hipStream_t st);
// x += d
void vadd(double *x, const double *d, int64_t n, hipStream_t st);

// ------------------------------------------------------------ row-major multivectors
// n x w blocks with element (r, c) at X[r * ld + c] (blockops.hip)
int64_t rgram_blocks(int64_t n);
// G (p x q, column-major, ld p) = A^T B for p, q <= 96; partial must hold
// rgram_blocks(n) * p * q doubles
void rgram(const double *A, int64_t lda, int p, const double *B, int64_t ldb, int q, int64_t n, double *partial,
           double *G, hipStream_t st);
// X = V C + beta X, V n x m, C m x p column-major on the device, m, p <= 96; X may be V
// itself (same stride, p <= m) but must not overlap it otherwise
void rgemm(const double *V, int64_t ldv, int m, const double *C, int p, double *X, int64_t ldx, int64_t n,
           double beta, hipStream_t st);
// the block eigensolver's Ritz step: Sn[:, 0:2m) = S C2, ASn[:, 0:m) = AS C2[:, 0:m),
// Sn[:, 2m:3m) = ASn[:, 0:m) - Sn[:, 0:m) diag(lam), out[c] = ||Sn[:, 2m + c]||^2
// for c < k (deterministic); S, AS, Sn, ASn share the row stride ld >= 3m; d <= 96,
// m <= 32; part must hold rritz_part(n) doubles, out 32
int64_t rritz_blocks(int64_t n);
int64_t rritz_part(int64_t n);
void rritz(const double *S, const double *AS, int64_t ld, int d, const double *C2, int m, const double *lam, int k,
           double *Sn, double *ASn, int64_t n, double *part, double *out, hipStream_t st);
// Y = A X (nv <= 96)
void rspmm(const SellA &S, const double *X, int64_t ldx, double *Y, int64_t ldy, int nv, hipStream_t st);
// R[:, c] = AX[:, c] - lam[c] X[:, c]
void rresid(const double *AX, int64_t ldax, const double *X, int64_t ldx, const double *lam, int m, int64_t n,
            double *R, int64_t ldr, hipStream_t st);
void rcopy(const double *X, int64_t ldx, int m, int64_t n, double *Y, int64_t ldy, hipStream_t st);
// R = AX - X diag(lam), m <= 32, and out[c] = ||R[:, c]||^2 for c < k (deterministic);
// part must hold rresid_norms_part() doubles
void rresid_norms(const double *AX, int64_t ldax, const double *X, int64_t ldx, const double *lam, int m, int k,
                  int64_t n, double *R, int64_t ldr, double *part, double *out, hipStream_t st);
int64_t rresid_norms_part();
// layout changes against column-major (ld n) blocks, m <= 96
void rto_cols(const double *X, int64_t ldx, int m, int64_t n, double *Y, hipStream_t st);
void rfrom_cols(const double *Y, int m, int64_t n, double *X, int64_t ldx, hipStream_t st);
// row-major FP32 (ld ldx) -> row-major FP64 (ld ldy), m columns
void rfrom_f32rows(const float *X, int64_t ldx, int m, int64_t n, double *Y, int64_t ldy, hipStream_t st);
// columns [c0, c1): uniform [-0.5, 0.5) from a counter-based generator
void rrandom(double *X, int64_t ldx, int c0, int c1, int64_t n, uint64_t seed, hipStream_t st);

// ------------------------------------------------------------ operator seam
struct DevOperator {
    virtual ~DevOperator() {}
    // y = M x (device pointers), M = A ('lm') or (A + delta I)^-1 ('sm')
    virtual int apply(const double *x, double *y) = 0;
    // y = A x, nv vectors
    virtual int multA(const double *X, int64_t ldx, double *Y, int64_t ldy, int nv) = 0;
    // after the operator's stream was synchronized: a failure the device
    // recorded during the applies since (0 if none)
    virtual int check() { return 0; }
    int64_t nops = 0;
    int64_t n = 0;
    hipStream_t stream = nullptr;
};

// Thick-restart Lanczos over a device operator (reference src/speigs_core.c:146-315).
int lanczos_dev(DevOperator *op, int64_t n, int64_t k, const speigs_opts *o,
                const speigs_hip_opts *ho, double anorm, double *lambda, double *Vout, double *resid,
                speigs_info *info);

// host dense kernels (host_dense.cpp)
int jacobi_eig(int64_t m, const double *Hupper, int64_t ldh, double *Z, double *w);
int dense_syev(int64_t n, double *A, double *w);   // Householder + implicit QL, ascending
// dense symmetric eigenproblem in two steps: all eigenvalues (ascending, T.w),
// then the eigenvectors of the ascending positions sel[0..k) into Vsel (ld ldv)
struct DenseTri {
    int64_t n = 0;
    std::vector<double> A, d0, e0, w;   // reflectors, tridiagonal, eigenvalues
    int iters = 4;                      // inverse-iteration steps per vector
    bool fast = false;                  // their back substitutions by reciprocal pivots
};
int dense_tri_values(int64_t n, const double *Afull, DenseTri &T);
// the same for the block eigensolver's projected matrices: the reduction on
// the lower triangle, QL rotations normed by sqrt where no scaling is needed,
// three inverse-iteration steps per vector with reciprocal pivots (values and
// vectors agree with dense_tri_values' to rounding, not bitwise)
int dense_tri_values_rr(int64_t n, const double *Afull, DenseTri &T);
// the same from a tridiagonal already formed (d, e; e[n - 1] unused): T.A empty
int dense_tri_values_from(int64_t n, const double *d, const double *e, DenseTri &T);
int dense_tri_vectors(const DenseTri &T, int64_t k, const int64_t *sel, double *Vsel, int64_t ldv);
// the selected eigenvectors of T's tridiagonal only (before the back-transformation)
int dense_tri_zvectors(const DenseTri &T, int64_t k, const int64_t *sel, double *Z, int64_t ldz);

// The dense short-circuit for a batch of small matrices (dense_batch.hip):
// each job's k eigenpairs of smallest (largest) |lambda|, the
// tridiagonalisation and back-transformation on the device (stream st).
constexpr int64_t kDenseBatchMax = 1024;    // largest n of a batched matrix
struct DenseJob {
    int64_t n = 0, k = 0;
    const double *A = nullptr;   // host, n x n column-major, full symmetric
    bool largest = false;        // 'lm': largest |lambda|
    double *lambda = nullptr;    // out: k values, in selection order
    double *V = nullptr;         // out: n x k vectors, or null
    std::vector<int64_t> sel;    // the selected eigenvalue indices (ascending order of T's values)
};
int dense_batch(std::vector<DenseJob> &jobs, hipStream_t st);
int band_chol_u(int64_t n, int64_t kd, double *AB, int64_t ld);
void band_solve_u(int64_t n, int64_t kd, const double *AB, int64_t ld, double *b);
int lu_factor(int64_t n, double *A, int64_t *piv);
void lu_solve(int64_t n, const double *LU, const int64_t *piv, double *b);

uint64_t xorshift_next(uint64_t *s);
void uniform_vec(uint64_t *s, int64_t n, double *v);

}  // namespace sehip
