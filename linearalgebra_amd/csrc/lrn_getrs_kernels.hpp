// Kernels of the LU solve (getrs): row permutation from LAPACK-style pivots, then blocked
// forward (unit lower) and backward (upper) substitution on B (row-major, n x nrhs).
//
// Algorithm followed: src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113
// (sequential interchanges, unit-L forward pass, U backward pass with a true division by
// U[i][i]), extended to several right-hand sides.
#pragma once

#include "lrn_device.hpp"
#include "lrn_getrf_kernels.hpp"

namespace lrn {

constexpr int kSolveTB = 64;          // rows per block of the substitution kernels
constexpr int kSolveMaxCols = 64;     // right-hand sides handled per launch
constexpr int kPermLdsMax = 36864;    // largest n whose permutation is built in LDS

// perm[i] = row of the input that lands in row i after the interchanges
// (i, piv[i]-1), i = 0..n-1, applied in order. One wave walks the pivots in blocks of 64:
// each block's net moves come from panel_rowmap and are applied to the label array.
__global__ __launch_bounds__(64) void perm_from_pivots_kernel(const long long* __restrict__ piv, long long n,
    int* __restrict__ perm_out, int* __restrict__ perm_tmp)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const bool in_lds = n <= kPermLdsMax;
    int* perm = in_lds ? reinterpret_cast<int*>(lds_raw) : perm_tmp;
    __shared__ long long pivs[64];
    __shared__ int rowmap[1 + 4 * 64];
    const int lane = threadIdx.x;
    for (long long i = lane; i < n; i += 64) perm[i] = static_cast<int>(i);
    __syncthreads();
    for (long long k = 0; k < n; k += 64) {
        const int w = static_cast<int>(min(64ll, n - k));
        if (lane < w) pivs[lane] = piv[k + lane] - 1;
        __syncthreads();
        panel_rowmap(pivs, k, w, 128, rowmap);
        __syncthreads();
        const int cnt = rowmap[0];
        int v0 = 0, v1 = 0;
        if (lane < cnt) v0 = perm[rowmap[1 + 128 + lane]];
        if (lane + 64 < cnt) v1 = perm[rowmap[1 + 128 + lane + 64]];
        __syncthreads();
        if (lane < cnt) perm[rowmap[1 + lane]] = v0;
        if (lane + 64 < cnt) perm[rowmap[1 + lane + 64]] = v1;
        __syncthreads();
    }
    for (long long i = lane; i < n; i += 64) perm_out[i] = perm[i];
}

// One workgroup per 64-row block; block b waits (bounded) for the blocks it depends on,
// which always have a lower workgroup index, so the kernel drains in dispatch order.
//
// Forward:  Y[b] = L[b][b]^-1 ( B[perm][b] - sum_{j<b} L[b][j] Y[j] )   (unit lower)
// Backward: X[b] = U[b][b]^-1 ( Y[b]       - sum_{j>b} U[b][j] X[j] )   (B receives X)
// `nc` right-hand-side columns starting at column c0 of B / Y (row strides ldb / ldy).
// The factors have type TF and are converted to the working type T as they are staged
// (TF = float, T = double applies low-precision factors in FP64 arithmetic).
template <typename TF, typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_blocked_kernel(const TF* __restrict__ LU, long long lda, long long n,
    const int* __restrict__ perm, const T* __restrict__ Bin, long long ldb_in, T* __restrict__ Out,
    long long ld_out, int nc, unsigned* __restrict__ flags, unsigned spin_limit, int* __restrict__ err,
    const int* __restrict__ skip_if_nonzero)
{
    if (skip_if_nonzero != nullptr && *skip_if_nonzero != 0) return;
    constexpr int TB = kSolveTB;
    __shared__ T Ls[TB][TB + 1];
    __shared__ T Xs[TB][kSolveMaxCols + 1];
    __shared__ T Acc[TB][kSolveMaxCols + 1];

    const int tid = threadIdx.x;
    const long long nblk = (n + TB - 1) / TB;
    const long long b = kLower ? blockIdx.x : (nblk - 1 - blockIdx.x);
    const long long r0 = b * TB;
    const int rows = static_cast<int>(min(static_cast<long long>(TB), n - r0));

    for (int idx = tid; idx < TB * kSolveMaxCols; idx += 256) Acc[idx / kSolveMaxCols][idx % kSolveMaxCols] = T(0);

    // off-diagonal contributions in the order the producing blocks complete: increasing j
    // for the forward pass, decreasing j for the backward pass (block nblk-1 first), so a
    // workgroup only ever waits for the block finished just before its own
    const long long jcount = kLower ? b : (nblk - 1 - b);
    for (long long q = 0; q < jcount; ++q) {
        const long long j = kLower ? q : (nblk - 1 - q);
        const long long c0 = j * TB;
        const int cols = static_cast<int>(min(static_cast<long long>(TB), n - c0));
        // stage the L/U block while the producer may still be running
        for (int idx = tid; idx < TB * TB; idx += 256) {
            const int r = idx / TB, c = idx % TB;
            Ls[r][c] = (r < rows && c < cols) ? static_cast<T>(LU[(r0 + r) * lda + c0 + c]) : T(0);
        }
        if (tid == 0) wait_counter(&flags[j], 1u, spin_limit, err);
        __syncthreads();
        for (int idx = tid; idx < TB * nc; idx += 256) {
            const int r = idx / nc, c = idx % nc;
            Xs[r][c] = (r < cols) ? ld_agent(&Out[(c0 + r) * ld_out + c]) : T(0);
        }
        __syncthreads();
        for (int idx = tid; idx < TB * nc; idx += 256) {
            const int r = idx / nc, c = idx % nc;
            T s = Acc[r][c];
            for (int p = 0; p < TB; ++p) s = fma(Ls[r][p], Xs[p][c], s);
            Acc[r][c] = s;
        }
        __syncthreads();
    }

    // right-hand side of this block, minus the accumulated contributions
    for (int idx = tid; idx < TB * TB; idx += 256) {
        const int r = idx / TB, c = idx % TB;
        Ls[r][c] = (r < rows && c < rows) ? static_cast<T>(LU[(r0 + r) * lda + r0 + c]) : T(0);
    }
    for (int idx = tid; idx < TB * nc; idx += 256) {
        const int r = idx / nc, c = idx % nc;
        T v = T(0);
        if (r < rows) {
            const long long src = kLower ? static_cast<long long>(perm[r0 + r]) : (r0 + r);
            v = Bin[src * ldb_in + c] - Acc[r][c];
        }
        Xs[r][c] = v;
    }
    __syncthreads();

    // diagonal block: one wave per right-hand side column, lane = row
    const int lane = tid & 63, wid = tid >> 6;
    for (int c = wid; c < nc; c += 4) {
        T x = (lane < rows) ? Xs[lane][c] : T(0);
        if (kLower) {
            for (int p = 0; p < rows; ++p) {
                const T xp = __shfl(x, p, 64);
                if (lane > p) x = x - Ls[lane][p] * xp;
            }
        } else {
            for (int p = rows - 1; p >= 0; --p) {
                if (lane == p) x = x / Ls[p][p];
                const T xp = __shfl(x, p, 64);
                if (lane < p) x = x - Ls[lane][p] * xp;
            }
        }
        if (lane < rows) Xs[lane][c] = x;
    }
    __syncthreads();
    for (int idx = tid; idx < rows * nc; idx += 256) {
        const int r = idx / nc, c = idx % nc;
        st_agent(&Out[(r0 + r) * ld_out + c], Xs[r][c]);
    }
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[b], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lrn
