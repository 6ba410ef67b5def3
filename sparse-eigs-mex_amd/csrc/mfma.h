// mfma.h -- 64 x 64 tile products on the gfx950 matrix cores, shared by the
// factorization and the inverse/solve kernels (device code only).
//
// Operand and result layouts of v_mfma_{f32,f64}_16x16x4 as used here: the
// first source gives, in lane l, element (l%16, k = l/16) of the left 16 x 4
// block, the second source element (k = l/16, l%16) of the right 4 x 16 block;
// the result holds, in lane l and register v, element (i, l%16) with
// i = 4(l/16)+v (f32) or i = l/16+4v (f64).  A product C = A B is formed as
// C^T = B^T A^T (first source from B, second from A), so lanes 0..15 of one
// result register hold 16 consecutive rows of one column of C and a
// column-major store of C is coalesced.
#pragma once
#include <hip/hip_runtime.h>

namespace sehip {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <class T> struct MfmaAcc;
template <> struct MfmaAcc<double> {
    typedef dbl4 V;
    // acc += (column block of B)^T (row block of A)^T; xi from A, xj from B
    static __device__ __forceinline__ V blk(double xi, double xj, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(xj, xi, c, 0, 0, 0);
    }
    // column of C (within the 16-block) held by register v of lane group lk
    static __device__ __forceinline__ int colof(int lk, int v) { return lk + 4 * v; }
};
template <> struct MfmaAcc<float> {
    typedef flt4 V;
    static __device__ __forceinline__ V blk(float xi, float xj, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(xj, xi, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int colof(int lk, int v) { return 4 * lk + v; }
};

// acc (this wave's 32 x 32 quadrant of a 64 x 64 C, as 2 x 2 blocks of 16 x 16)
// += A B over kw (multiple of 4) inner indices; As[k][row] holds A^T, Bs[k][col]
// holds B, both in LDS with row pitch P.  Wave w owns rows 32(w&1).., columns
// 32(w>>1)..; in the result, lane l, block (a, b), register v is
// C[32(w&1) + 16a + l%16][32(w>>1) + 16b + colof(l/16, v)].
template <class T, int P>
__device__ __forceinline__ void tile_mma(const T (*As)[P], const T (*Bs)[P], int kw,
                                         typename MfmaAcc<T>::V acc[2][2])
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int rb = (wv & 1) * 32, cb = (wv >> 1) * 32;
    const int lr = lane & 15, lk = lane >> 4;
    for (int kk = 0; kk < kw; kk += 4) {
        T xi[2], xj[2];
#pragma unroll
        for (int a = 0; a < 2; a++) {
            xi[a] = As[kk + lk][rb + 16 * a + lr];
            xj[a] = Bs[kk + lk][cb + 16 * a + lr];
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) acc[a][b] = MfmaAcc<T>::blk(xi[a], xj[b], acc[a][b]);
    }
}

template <class T>
__device__ __forceinline__ void zero_acc(typename MfmaAcc<T>::V acc[2][2])
{
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = typename MfmaAcc<T>::V{T(0), T(0), T(0), T(0)};
}

}  // namespace sehip
