// Kernels of the direct consumers of the LU factors: the determinant (diagonal product and
// interchange sign) and the explicit inverse (identity right-hand sides, conditional copy-back).
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

// det(A) from packed L\U factors: the diagonal of U multiplied in increasing row order with
// one running product (the rounding of the reference's sequential loops,
// src/matrix_operation/cpu/determinant.cpp:19-28 and
// src/matrix_operation/gpu/hipblas/inversion_hip.hip:36-58), negated when the pivots hold an
// odd number of interchanges (piv[i] != i + 1, 1-based), and zero when `info` reports a zero
// pivot. One workgroup: 256 diagonal entries are staged through LDS per step and one lane
// keeps the running product. piv may be null (no sign), info may be null (not checked).
// square = 1: the Cholesky form, (prod of L's diagonal)^2 (inversion_hip.hip:62-80).
template <typename T>
__global__ __launch_bounds__(256) void lu_det_kernel(const T* __restrict__ LU, long long lda, long long n,
    const long long* __restrict__ piv, const int* __restrict__ info, T* __restrict__ out, int square)
{
    __shared__ T diag[256];
    __shared__ int odd;
    const int t = threadIdx.x;
    if (t == 0) odd = 0;
    T prod = T(1);
    int mine = 0;
    // the next 256 diagonal entries are loaded while lane 0 multiplies the current ones
    T cur = (t < n) ? LU[static_cast<long long>(t) * lda + t] : T(1);
    for (long long i0 = 0; i0 < n; i0 += 256) {
        const long long i = i0 + t;
        const long long inext = i + 256;
        const T nxt = (inext < n) ? LU[inext * lda + inext] : T(1);
        if (i < n && piv != nullptr && piv[i] != i + 1) mine ^= 1;
        __syncthreads();
        diag[t] = cur;
        __syncthreads();
        if (t == 0) {
            const int cnt = static_cast<int>(n - i0 < 256 ? n - i0 : 256);
            for (int k = 0; k < cnt; ++k) prod *= diag[k];
        }
        cur = nxt;
    }
    if (mine) atomicXor(&odd, 1);
    __syncthreads();
    if (t == 0) {
        T r = odd ? -prod : prod;
        if (square) r = r * r;
        if (info != nullptr && *info != 0) r = T(0);
        out[0] = r;
    }
}

// X = P I (n x n, row stride ldx): row i is the unit row e_perm[i], perm[i] being the row of
// the input that lands in row i after the interchanges; grid-stride loop.
template <typename T>
__global__ __launch_bounds__(256) void perm_eye_kernel(T* __restrict__ X, long long n, long long ldx,
    const int* __restrict__ perm)
{
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; idx < n * n; idx += stride) {
        const long long i = idx / n, j = idx - i * n;
        X[i * ldx + j] = (static_cast<long long>(perm[i]) == j) ? T(1) : T(0);
    }
}

constexpr int kStripRows = 64;

// One row strip of a blocked triangular solve with many right-hand sides:
// X[r0 : r0 + w, :] = D^-1 X[r0 : r0 + w, :] for the w x w diagonal block D = A[r0 : r0 + w,
// r0 : r0 + w] (w <= 64), lower (kLower; unit diagonal when kUnit) or non-unit upper. One
// thread per column of X
// (64-thread workgroups, consecutive columns so the row accesses coalesce), the column's w
// values in registers, D staged in LDS and read as broadcasts.
template <typename T, bool kLower, bool kUnit = true>
__global__ __launch_bounds__(64) void strip_trsm_kernel(const T* __restrict__ A, long long lda, long long r0, int w,
    T* __restrict__ X, long long ldx, long long ncols)
{
    __shared__ T tri[kStripRows][kStripRows + 1];
    // loads in batches of 16 per thread, all in flight before any is written to LDS
    constexpr int kBatch = 16;
#pragma unroll
    for (int base = 0; base < kStripRows * kStripRows; base += 64 * kBatch) {
        T v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = base + static_cast<int>(threadIdx.x) + 64 * u;
            const int i = idx / kStripRows, j = idx - i * kStripRows;
            v[u] = (i < w && j < w) ? A[(r0 + i) * lda + r0 + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = base + static_cast<int>(threadIdx.x) + 64 * u;
            tri[idx / kStripRows][idx % kStripRows] = v[u];
        }
    }
    __syncthreads();
    const long long col = static_cast<long long>(blockIdx.x) * 64 + threadIdx.x;
    if (col >= ncols) return;
    T x[kStripRows];
#pragma unroll
    for (int i = 0; i < kStripRows; ++i) x[i] = (i < w) ? X[(r0 + i) * ldx + col] : T(0);
    if constexpr (kLower) {
        // each row's body under a branch on (i < w), as in the upper form: with a select
        // instead, the unit form is scheduled with the LDS reads hoisted and spills to scratch
#pragma unroll
        for (int i = 0; i < kStripRows; ++i) {
            if (i < w) {
                T acc = x[i];
#pragma unroll
                for (int k = 0; k < i; ++k) acc -= tri[i][k] * x[k];
                if constexpr (kUnit)
                    x[i] = acc;
                else
                    x[i] = acc / tri[i][i];
            }
        }
    } else {
#pragma unroll
        for (int i = kStripRows - 1; i >= 0; --i) {
            if (i < w) {
                T acc = x[i];
#pragma unroll
                for (int k = i + 1; k < kStripRows; ++k) acc -= tri[i][k] * x[k];
                x[i] = acc / tri[i][i];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < kStripRows; ++i)
        if (i < w) X[(r0 + i) * ldx + col] = x[i];
}

// X = I (n x n, row stride ldx); grid-stride loop.
template <typename T>
__global__ __launch_bounds__(256) void eye_kernel(T* __restrict__ X, long long n, long long ldx)
{
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; idx < n * n; idx += stride) {
        const long long i = idx / n, j = idx - i * n;
        X[i * ldx + j] = (i == j) ? T(1) : T(0);
    }
}

// dst = src (n x n, both with row stride n) unless *skip is nonzero (a singular factorisation
// leaves the factors in place, as the reference's LU inversion does); grid-stride loop.
template <typename T>
__global__ __launch_bounds__(256) void copy_unless_kernel(const T* __restrict__ src, T* __restrict__ dst,
    long long count, const int* __restrict__ skip, const int* __restrict__ err)
{
    if (skip != nullptr && *skip != 0) return;
    if (err != nullptr && *err != 0) return;  // a bounded wait of the factorisation expired
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; idx < count; idx += stride)
        dst[idx] = src[idx];
}

}  // namespace lrn
