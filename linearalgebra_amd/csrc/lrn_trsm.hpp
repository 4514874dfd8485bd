// Blocked triangular solve with many right-hand sides: op(A) X = X in place, op(A) n x n
// triangular, X n x nrhs row-major.
//
// The rows are cut into 64-row strips. First every 64 x 64 diagonal block of op(A) is inverted
// (all blocks at once, trsm_diag_inv_kernel). Then one launch per strip (trsm_step_kernel):
// strip s is final when launch s starts; every row tile still to be solved receives
//     X_t -= op(A)[t][s] X_s
// on the matrix cores, and the workgroup of the next strip finishes that strip as
//     X_t = inv(D_t) X_t
// in the same launch, so launch s + 1 starts from a final strip again. There is no wait between
// workgroups: the launches on the stream order the strips.
//
// The reference solves row by row (src/matrix_operation/cpu/native/triangular_solver_native.cpp:64-153,
// src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113); this form reaches the same
// solution up to rounding (products of the diagonal-block inverses in place of the
// substitution inside each 64-row strip, blocked sums across strips), the blocked arrangement
// used by GPU triangular solvers for many right-hand sides.
#pragma once

#include <hip/hip_runtime.h>

#include "lrn_device.hpp"

namespace lrn {

constexpr int kTsRows = 64;               // rows per strip, columns per workgroup
constexpr int kTsInvElems = kTsRows * kTsRows;  // elements of one stored diagonal-block inverse

// Entry (i, j) of the operator: A[i][j], or A[j][i] when transposed.
template <typename T, bool kTrans>
__device__ __forceinline__ T op_coef(const T* __restrict__ A, long long lda, long long i, long long j)
{
    return kTrans ? A[j * lda + i] : A[i * lda + j];
}

// Tile loads: the 16 elements of each thread are loaded into registers by a fully unrolled
// loop before any of them is written to LDS, so the 16 loads are in flight together (a rolled
// loop issues one load, waits for it and writes LDS, sixteen times in a row).
constexpr int kTsPer = kTsRows * kTsRows / 256;

// dst[k][i] = op(A)[r0 + i][k0 + k] (zero outside rows x krows), read along A's contiguous index
template <typename T, bool kTrans>
__device__ __forceinline__ void ts_load_op_tile(T (*dst)[kTsRows + 2], const T* __restrict__ A, long long lda,
    long long r0, int rows, long long k0, int krows, int tid)
{
    T v[kTsPer];
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) {
        const int idx = tid + 256 * q;
        const int i = kTrans ? (idx & 63) : (idx >> 6);
        const int k = kTrans ? (idx >> 6) : (idx & 63);
        v[q] = (i < rows && k < krows) ? op_coef<T, kTrans>(A, lda, r0 + i, k0 + k) : T(0);
    }
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) {
        const int idx = tid + 256 * q;
        const int i = kTrans ? (idx & 63) : (idx >> 6);
        const int k = kTrans ? (idx >> 6) : (idx & 63);
        dst[k][i] = v[q];
    }
}

// dst[k][j] = X[k0 + k][j] (zero outside krows x cols); kCoherent: agent-coherent loads
template <typename T, bool kCoherent>
__device__ __forceinline__ void ts_load_x_tile(T (*dst)[kTsRows + 2], const T* __restrict__ X, long long ldx,
    long long k0, int krows, int cols, int tid)
{
    T v[kTsPer];
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) {
        const int idx = tid + 256 * q;
        const int k = idx >> 6, j = idx & 63;
        const bool in = k < krows && j < cols;
        if constexpr (kCoherent)
            v[q] = in ? ld_agent(&X[(k0 + k) * ldx + j]) : T(0);
        else
            v[q] = in ? X[(k0 + k) * ldx + j] : T(0);
    }
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) {
        const int idx = tid + 256 * q;
        dst[idx >> 6][idx & 63] = v[q];
    }
}

// Inverse of every 64 x 64 diagonal block D_b of op(A) (lower triangular when kLower, with an
// implicit unit diagonal when kUnit; only that triangle is read), stored transposed:
// dinvT[b * 4096 + j * 64 + i] = inv(D_b)[i][j]. Padding rows and columns of the last block
// are zero. grid (nblk, 4): workgroup (b, g) computes columns 16 g .. 16 g + 15 of inv(D_b),
// wave w of it four of them, one column at a time with lane = row: forward (backward)
// substitution on the unit vector, each solved entry broadcast with a lane read. The entries of
// A (type TA) are widened to the arithmetic type T as they are read (FP32 factors inverted in
// FP64 for the mixed solver's refinement).
template <typename TA, typename T, bool kLower, bool kUnit, bool kTrans>
__global__ __launch_bounds__(256) void trsm_diag_inv_kernel(const TA* __restrict__ A, long long lda, long long n,
    T* __restrict__ dinvT, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    __shared__ T Ds[kTsRows][kTsRows + 1];
    const long long b = blockIdx.x;
    const long long r0 = b * kTsRows;
    const int rows = static_cast<int>(min(static_cast<long long>(kTsRows), n - r0));
    {
        // consecutive threads read consecutive addresses of A in either orientation; all 16
        // loads of a thread in flight before any is written to LDS
        T v[kTsPer];
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) {
            const int idx = threadIdx.x + 256 * q;
            const int r = kTrans ? (idx & 63) : (idx >> 6);
            const int c = kTrans ? (idx >> 6) : (idx & 63);
            const bool in = r < rows && c < rows && (kLower ? (c < r || (!kUnit && c == r)) : (c > r || (!kUnit && c == r)));
            v[q] = in ? static_cast<T>(op_coef<TA, kTrans>(A, lda, r0 + r, r0 + c)) : T(0);
        }
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) {
            const int idx = threadIdx.x + 256 * q;
            Ds[kTrans ? (idx & 63) : (idx >> 6)][kTrans ? (idx >> 6) : (idx & 63)] = v[q];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T drow[kTsRows];
#pragma unroll
    for (int p = 0; p < kTsRows; ++p) drow[p] = Ds[lane][p];
    for (int q = 0; q < 4; ++q) {
        const int j = static_cast<int>(blockIdx.y) * 16 + wid * 4 + q;
        T x = (lane == j) ? T(1) : T(0);
        if constexpr (kLower) {
#pragma unroll
            for (int p = 0; p < kTsRows; ++p) {
                if (p < rows) {
                    if (!kUnit && lane == p) x = x / drow[p];
                    const T xp = read_lane_t<T>(x, p);
                    if (lane > p) x = x - drow[p] * xp;
                }
            }
        } else {
#pragma unroll
            for (int p = kTsRows - 1; p >= 0; --p) {
                if (p < rows) {
                    if (!kUnit && lane == p) x = x / drow[p];
                    const T xp = read_lane_t<T>(x, p);
                    if (lane < p) x = x - drow[p] * xp;
                }
            }
        }
        dinvT[b * kTsInvElems + j * kTsRows + lane] = (lane < rows && j < rows) ? x : T(0);
    }
}

// M_b = inv(D_b) A[b][p] for every 64-row block b with a chain predecessor p at distance dist
// (p = b - dist when kLower, else b + dist), from the stored inverses (dinvT, as written by
// trsm_diag_inv_kernel) and
// the off-diagonal block of A (type TA, widened to T), stored in the same transposed layout:
// MT[b * 4096 + j * 64 + i] = M_b[i][j]; padding rows and columns are zero; blocks without a
// predecessor are not written. One workgroup per block, both operands in LDS, 16 results per
// thread.
template <typename TA, typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_pred_kernel(const TA* __restrict__ A, long long lda, long long n,
    const T* __restrict__ dinvT, T* __restrict__ MT, int dist, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    __shared__ T Di[kTsRows][kTsRows + 1];  // Di[k][i] = inv(D_b)[i][k]
    __shared__ T Ap[kTsRows][kTsRows + 1];  // Ap[k][j] = A[r0 + k][c0 + j]
    const long long nblk = (n + kTsRows - 1) / kTsRows;
    const long long b = blockIdx.x;
    const long long p = kLower ? b - dist : b + dist;
    if (p < 0 || p >= nblk) return;
    const long long r0 = b * kTsRows, c0 = p * kTsRows;
    const int rows = static_cast<int>(min(static_cast<long long>(kTsRows), n - r0));
    const int cols = static_cast<int>(min(static_cast<long long>(kTsRows), n - c0));
    const int tid = threadIdx.x;
    {
        T v[kTsPer], w[kTsPer];
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) {
            const int idx = tid + 256 * q;
            const int k = idx >> 6, j = idx & 63;
            v[q] = dinvT[b * kTsInvElems + idx];
            w[q] = (k < rows && j < cols) ? static_cast<T>(A[(r0 + k) * lda + c0 + j]) : T(0);
        }
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) {
            const int idx = tid + 256 * q;
            Di[idx >> 6][idx & 63] = v[q];
            Ap[idx >> 6][idx & 63] = w[q];
        }
    }
    __syncthreads();
    const int i = tid & 63, j0 = tid >> 6;
    T acc[kTsPer];
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) acc[q] = T(0);
    for (int k = 0; k < kTsRows; ++k) {
        const T d = Di[k][i];
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) acc[q] = fma(d, Ap[k][j0 + 4 * q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) MT[b * kTsInvElems + (j0 + 4 * q) * kTsRows + i] = acc[q];
}

// Y[i][c] = B[src(i)][c] for i < n, c < nc (src(i) = perm[i], or i when perm is null);
// grid-stride loop.
template <typename T>
__global__ __launch_bounds__(256) void gather_rows_kernel(const T* __restrict__ B, long long ldb,
    const int* __restrict__ perm, T* __restrict__ Y, long long ldy, long long n, long long nc,
    const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    const long long total = n * nc;
    const long long stride = static_cast<long long>(gridDim.x) * 256;
    for (long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; idx < total; idx += stride) {
        const long long i = idx / nc, c = idx - i * nc;
        const long long src = perm != nullptr ? static_cast<long long>(perm[i]) : i;
        Y[i * ldy + c] = B[src * ldb + c];
    }
}

// One strip step (see the top of this file). Workgroup (x, y): row tile t = s + 1 + x forward
// (kLower), s - 1 - x backward, right-hand-side columns 64 y .. 64 y + 63. s = -1 forward
// (nblk backward) means no strip is final yet: the launch has one workgroup row, which only
// finishes the first strip. Four waves in a 2 x 2 arrangement, each a 32 x 32 block of
// 16 x 16 MFMA tiles; op(A)'s block and X_s staged k-major in LDS so that each lane's MFMA
// operand is one LDS read (the operand layout of gemm_sub_kernel).
template <typename T, bool kLower, bool kTrans>
__global__ __launch_bounds__(256) void trsm_step_kernel(const T* __restrict__ A, long long lda, long long n,
    const T* __restrict__ dinvT, T* __restrict__ X, long long ldx, long long nrhs, long long s,
    const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    constexpr int PAD = 2;
    __shared__ T As[kTsRows][kTsRows + PAD];  // As[k][i]: left operand, k-major
    __shared__ T Bs[kTsRows][kTsRows + PAD];  // Bs[k][j]: right operand

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const long long nblk = (n + kTsRows - 1) / kTsRows;
    const long long t = kLower ? s + 1 + blockIdx.x : s - 1 - blockIdx.x;
    const long long r0 = t * kTsRows;
    const int rows = static_cast<int>(min(static_cast<long long>(kTsRows), n - r0));
    const long long c0 = static_cast<long long>(blockIdx.y) * kTsRows;
    const int cols = static_cast<int>(min(static_cast<long long>(kTsRows), nrhs - c0));
    const bool has_src = kLower ? (s >= 0) : (s < nblk);
    const bool diag = blockIdx.x == 0;

    acc_t acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};

    auto mma_pass = [&]() {
#pragma unroll
        for (int kk = 0; kk < kTsRows; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[2], b[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) a[mi] = As[kr][wm * 32 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) b[ni] = Bs[kr][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = MF::mma(a[mi], b[ni], acc[mi][ni]);
        }
    };

    if (has_src) {
        const long long k0 = s * kTsRows;
        const int krows = static_cast<int>(min(static_cast<long long>(kTsRows), n - k0));
        ts_load_op_tile<T, kTrans>(As, A, lda, r0, rows, k0, krows, tid);
        ts_load_x_tile<T, false>(Bs, X + c0, ldx, k0, krows, cols, tid);
        __syncthreads();
        mma_pass();
    }

    if (!diag) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                    const int j = wn * 32 + ni * 16 + (lane & 15);
                    if (i < rows && j < cols) {
                        T* p = X + (r0 + i) * ldx + c0 + j;
                        *p = *p - acc[mi][ni][v];
                    }
                }
        return;
    }

    // the next strip: its updated rows become the right operand, inv(D_t) the left one
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                const int j = wn * 32 + ni * 16 + (lane & 15);
                Bs[i][j] = (i < rows && j < cols) ? X[(r0 + i) * ldx + c0 + j] - acc[mi][ni][v] : T(0);
            }
    {
        const T* inv = dinvT + t * kTsInvElems;
        T inv_v[kTsPer];
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) inv_v[q] = inv[tid + 256 * q];
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) {
            const int idx = tid + 256 * q;
            As[idx >> 6][idx & 63] = inv_v[q];
        }
    }
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};
    mma_pass();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                const int j = wn * 32 + ni * 16 + (lane & 15);
                if (i < rows && j < cols) X[(r0 + i) * ldx + c0 + j] = acc[mi][ni][v];
            }
}

// The same solve in one launch: workgroup b owns row tile t (t = b forward, nblk - 1 - b
// backward) and applies the strips it depends on in the order they become final, waiting for
// each on its flag (bounded), then finishes its own tile with its diagonal-block inverse and
// raises its flag. op(A)'s tile for the next strip and inv(D_t) are loaded before the wait, so
// the critical path per strip is one flag hand-off, the strip's 64 x 64 x 64 product and the
// inverse's. Every workgroup must be resident at once (the host checks the occupancy); a wait
// that expires raises *err and the caller poisons the result. X_s is read with
// agent-coherent loads after the flag, and each tile is stored with agent-coherent stores
// that complete before the flag is raised (no fences: an agent-scope acquire would invalidate
// the L2 of the reading XCD on every poll).
template <typename T, bool kLower, bool kTrans>
__global__ __launch_bounds__(256) void trsm_chain_kernel(const T* __restrict__ A, long long lda, long long n,
    const T* __restrict__ dinvT, T* __restrict__ X, long long ldx, long long nrhs, unsigned* __restrict__ flags,
    unsigned spin_limit, int* __restrict__ err, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    constexpr int PAD = 2;
    __shared__ T As[kTsRows][kTsRows + PAD];
    __shared__ T Bs[kTsRows][kTsRows + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const long long nblk = (n + kTsRows - 1) / kTsRows;
    const long long t = kLower ? static_cast<long long>(blockIdx.x) : nblk - 1 - blockIdx.x;
    const long long r0 = t * kTsRows;
    const int rows = static_cast<int>(min(static_cast<long long>(kTsRows), n - r0));
    const int cols = static_cast<int>(min(static_cast<long long>(kTsRows), nrhs));

    acc_t acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};

    auto mma_pass = [&]() {
#pragma unroll
        for (int kk = 0; kk < kTsRows; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[2], b[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) a[mi] = As[kr][wm * 32 + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) b[ni] = Bs[kr][wn * 32 + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = MF::mma(a[mi], b[ni], acc[mi][ni]);
        }
    };

    // this tile's right-hand side (in the product's element layout) and its inverse, loaded
    // before any wait
    T xt[2][2][4];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                const int j = wn * 32 + ni * 16 + (lane & 15);
                xt[mi][ni][v] = (i < rows && j < cols) ? X[(r0 + i) * ldx + j] : T(0);
            }
    T inv_v[kTsPer];
    {
        const T* inv = dinvT + t * kTsInvElems;
#pragma unroll
        for (int q = 0; q < kTsPer; ++q) inv_v[q] = inv[tid + 256 * q];
    }

    const long long deps = kLower ? t : nblk - 1 - t;
    for (long long q = 0; q < deps; ++q) {
        const long long sidx = kLower ? q : nblk - 1 - q;
        const long long k0 = sidx * kTsRows;
        const int krows = static_cast<int>(min(static_cast<long long>(kTsRows), n - k0));
        __syncthreads();  // the previous strip's operands are consumed
        ts_load_op_tile<T, kTrans>(As, A, lda, r0, rows, k0, krows, tid);
        if (tid == 0) wait_flag_coherent(&flags[sidx], 1u, spin_limit, err);
        __syncthreads();
        ts_load_x_tile<T, true>(Bs, X, ldx, k0, krows, cols, tid);
        __syncthreads();
        mma_pass();
    }

    // own tile: Y = X_t - acc, then X_t = inv(D_t) Y
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                const int j = wn * 32 + ni * 16 + (lane & 15);
                Bs[i][j] = (i < rows && j < cols) ? xt[mi][ni][v] - acc[mi][ni][v] : T(0);
            }
#pragma unroll
    for (int q = 0; q < kTsPer; ++q) {
        const int idx = tid + 256 * q;
        As[idx >> 6][idx & 63] = inv_v[q];
    }
    __syncthreads();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};
    mma_pass();
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = wm * 32 + mi * 16 + MF::row(lane, v);
                const int j = wn * 32 + ni * 16 + (lane & 15);
                if (i < rows && j < cols) st_agent(&X[(r0 + i) * ldx + j], static_cast<T>(acc[mi][ni][v]));
            }
    wait_own_stores();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[t], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lrn
