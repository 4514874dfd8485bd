// Kernels of the blocked Cholesky factorisation A = L L^T (row-major, lower triangle read):
// per 64-wide block column, the diagonal block in one workgroup, the rows below it solved
// against that block (also written transposed for the trailing GEMM), then the trailing
// update on the GEMM kernels of the LU path. A final pass writes the triangle layout.
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

constexpr int kCholNB = 64;

// Diagonal block A[k : k + bs, k : k + bs] (bs <= 64) factored right-looking: at step
// j the updated diagonal d = a_jj - sum_{p<j} l_jp^2 fails the factorisation when d <= 0
// (the reference's test, src/matrix_operation/cpu/native/cholesky_decomposition_native.cpp:
// 139-142; NaN passes it), *info = k + j + 1; otherwise l_jj = sqrt(d), the column below is
// divided by it and the rest of the block's lower triangle takes the rank-1 update. Only the
// lower triangle is read or written. Skipped when *info is already nonzero.
template <typename T>
__global__ __launch_bounds__(64) void chol_diag_kernel(T* __restrict__ A, long long lda, long long k, int bs,
    int* __restrict__ info)
{
    // one wavefront: thread i holds row i of the block in registers; per step the diagonal and
    // then column j pass through LDS (the same operations, in the same order, as a
    // right-looking update of the block in place)
    __shared__ T col[kCholNB];
    __shared__ T diag;
    if (*info != 0) return;
    const int t = threadIdx.x;
    T x[kCholNB];
#pragma unroll
    for (int c = 0; c < kCholNB; ++c) x[c] = (t < bs && c <= t) ? A[(k + t) * lda + k + c] : T(0);
    int fail = kCholNB;  // step at which the factorisation failed (uniform: read from LDS)
#pragma unroll
    for (int j = 0; j < kCholNB; ++j) {
        if (j < bs && fail == kCholNB) {
            if (t == j) diag = x[j];
            __syncthreads();
            const T d = diag;
            if (d <= T(0)) {
                if (t == 0) info[0] = static_cast<int>(k + j + 1);
                fail = j;
            } else {
                const T l = sqrt(d);
                if (t == j) x[j] = l;
                if (t > j) x[j] = x[j] / l;
                col[t] = x[j];
                __syncthreads();
                // lower triangle of the rest: (t, c) with j < c <= t
#pragma unroll
                for (int c = j + 1; c < kCholNB; ++c)
                    if (c <= t && t < bs) x[c] = x[c] - x[j] * col[c];
                __syncthreads();
            }
        }
    }
    // on failure at step f: what the reference's row-by-row kernel has written when it stops at
    // row f (cholesky_decomposition_native.cpp:124-142), i.e. the completed rows and row f's
    // multipliers; row f's diagonal and the rows after it stay as they were
    if (t < bs && t <= fail) {
#pragma unroll
        for (int c = 0; c < kCholNB; ++c)
            if (c <= t && c < fail) A[(k + t) * lda + k + c] = x[c];
    }
}

// Rows r in [k + bs, n) of the block column: x L11^T = a, i.e. x_j = (a_j - sum_{p<j} x_p l_jp)
// / l_jj in increasing j (cholesky_decomposition_native.cpp:278-286), written back in place and
// transposed into Wt (bs x m, row stride ldw, column r - k - bs) for the trailing GEMM. 64
// rows per workgroup, staged through LDS so the row reads and writes coalesce; L11 read as
// broadcasts. Skipped when *info is nonzero.
template <typename T>
__global__ __launch_bounds__(64) void chol_rows_kernel(T* __restrict__ A, long long lda, long long k, int bs,
    long long n, T* __restrict__ Wt, long long ldw, const int* __restrict__ info)
{
    __shared__ T l11[kCholNB][kCholNB + 1];
    __shared__ T tile[kCholNB][kCholNB + 1];
    if (*info != 0) return;
    const int t = threadIdx.x;
    const long long r0 = k + bs + static_cast<long long>(blockIdx.x) * 64;
    const int rows = static_cast<int>(n - r0 < 64 ? n - r0 : 64);
    // loads in batches of 16 per thread, all in flight before any is written to LDS (a rolled
    // loop would wait for each load in turn: 64 round trips per thread)
    constexpr int kBatch = 16;
#pragma unroll
    for (int base = 0; base < kCholNB * kCholNB; base += 64 * kBatch) {
        T lv[kBatch], tv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = base + t + 64 * u;
            const int i = idx / kCholNB, j = idx - i * kCholNB;
            lv[u] = (i < bs && j <= i) ? A[(k + i) * lda + k + j] : T(0);
            tv[u] = (i < rows && j < bs) ? A[(r0 + i) * lda + k + j] : T(0);
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = base + t + 64 * u;
            const int i = idx / kCholNB, j = idx - i * kCholNB;
            l11[i][j] = lv[u];
            tile[i][j] = tv[u];
        }
    }
    __syncthreads();
    if (t < rows) {
        // the row in registers: the dependent chain reads only the broadcast l11 entries
        T x[kCholNB];
#pragma unroll
        for (int p = 0; p < kCholNB; ++p) x[p] = tile[t][p];
#pragma unroll
        for (int j = 0; j < kCholNB; ++j) {
            if (j < bs) {
                T acc = x[j];
#pragma unroll
                for (int p = 0; p < j; ++p) acc = acc - x[p] * l11[j][p];
                x[j] = acc / l11[j][j];
            }
        }
#pragma unroll
        for (int p = 0; p < kCholNB; ++p) tile[t][p] = x[p];
    }
    __syncthreads();
#pragma unroll 16
    for (int idx = t; idx < kCholNB * kCholNB; idx += 64) {
        const int i = idx / kCholNB, j = idx - i * kCholNB;
        if (i < rows && j < bs) A[(r0 + i) * lda + k + j] = tile[i][j];
    }
    // transposed copy: row j of Wt holds column j of the solved rows
#pragma unroll 16
    for (int idx = t; idx < kCholNB * kCholNB; idx += 64) {
        const int j = idx / kCholNB, i = idx - j * kCholNB;
        if (i < rows && j < bs) Wt[j * ldw + (r0 - k - bs) + i] = tile[i][j];
    }
}

// Final layout: upper = 0, zero the strict upper triangle; upper = 1, U[j][i] = L[i][j] and the
// strict lower triangle zeroed (cholesky_decomposition_native.cpp:262-265,312-319); upper = 2,
// L^T copied into the strict upper triangle with L kept (the symmetric storage the inversion
// solves with). One thread per pair (i > j), grid-stride. Skipped when *info is nonzero (the
// factors are incomplete).
template <typename T>
__global__ __launch_bounds__(256) void chol_layout_kernel(T* __restrict__ A, long long n, int upper,
    const int* __restrict__ info)
{
    if (*info != 0) return;
    const long long pairs = n * (n - 1) / 2;
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long q = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; q < pairs; q += stride) {
        // q -> (i, j), i > j, rows in order: row i holds pairs [i(i-1)/2, i(i+1)/2)
        long long i = static_cast<long long>((1.0 + sqrt(1.0 + 8.0 * static_cast<double>(q))) * 0.5);
        while (i * (i - 1) / 2 > q) --i;
        while ((i + 1) * i / 2 <= q) ++i;
        const long long j = q - i * (i - 1) / 2;
        if (upper == 2) {
            A[j * n + i] = A[i * n + j];
        } else if (upper) {
            A[j * n + i] = A[i * n + j];
            A[i * n + j] = T(0);
        } else {
            A[j * n + i] = T(0);
        }
    }
}

}  // namespace lrn
