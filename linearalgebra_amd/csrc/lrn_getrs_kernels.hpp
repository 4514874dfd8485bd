// Kernels of the LU solve (getrs): row permutation from LAPACK-style pivots, then blocked
// forward (unit lower) and backward (upper) substitution on B (row-major, n x nrhs).
//
// Algorithm followed: src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113
// (sequential interchanges, unit-L forward pass, U backward pass with a true division by
// U[i][i]), extended to several right-hand sides.
#pragma once

#include "lrn_device.hpp"
#include "lrn_getrf_kernels.hpp"
#include "lrn_panel.hpp"

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

template <typename T> __device__ __forceinline__ T read_lane_t(T v, int lane);
template <> __device__ __forceinline__ double read_lane_t<double>(double v, int lane) { return read_lane_f64(v, lane); }
template <> __device__ __forceinline__ float read_lane_t<float>(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One workgroup per 64-row block; block b waits (bounded) for the blocks it depends on,
// which always have a lower workgroup index, so the kernel drains in dispatch order.
//
// Forward:  Y[b] = L[b][b]^-1 ( B[perm][b] - sum_{j<b} L[b][j] Y[j] )   (unit lower)
// Backward: X[b] = U[b][b]^-1 ( Y[b]       - sum_{j>b} U[b][j] X[j] )   (B receives X)
// `nc` right-hand-side columns starting at column c0 of B / Y (row strides ldb / ldy).
// The factors have type TF and are converted to the working type T as they are staged
// (TF = float, T = double applies low-precision factors in FP64 arithmetic).
//
// Critical path: block b finishes one step after block b-1 (b+1 backward), so everything
// that does not depend on the previous block -- the diagonal block, this block's right-hand
// side, the earlier off-diagonal blocks -- is loaded or applied before the wait, and the
// diagonal solve keeps the block row in registers and broadcasts each solved value with a
// lane read instead of an LDS round trip.
template <typename TF, typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv_blocked_kernel(const TF* __restrict__ LU, long long lda, long long n,
    const int* __restrict__ perm, const T* __restrict__ Bin, long long ldb_in, T* __restrict__ Out,
    long long ld_out, int nc, unsigned* __restrict__ flags, unsigned spin_limit, int* __restrict__ err,
    const int* __restrict__ skip_if_nonzero)
{
    if (skip_if_nonzero != nullptr && *skip_if_nonzero != 0) return;
    constexpr int TB = kSolveTB;
    constexpr int kRhsRegs = TB * kSolveMaxCols / 256;
    __shared__ T Ls[TB][TB + 1];
    __shared__ T Ds[TB][TB + 1];
    __shared__ T Xs[TB][kSolveMaxCols + 1];
    __shared__ T Acc[TB][kSolveMaxCols + 1];

    const int tid = threadIdx.x;
    const long long nblk = (n + TB - 1) / TB;
    const long long b = kLower ? blockIdx.x : (nblk - 1 - blockIdx.x);
    const long long r0 = b * TB;
    const int rows = static_cast<int>(min(static_cast<long long>(TB), n - r0));

    // independent of every other block: diagonal block and right-hand side
    for (int idx = tid; idx < TB * TB; idx += 256) {
        const int r = idx / TB, c = idx % TB;
        Ds[r][c] = (r < rows && c < rows) ? static_cast<T>(LU[(r0 + r) * lda + r0 + c]) : T(0);
    }
    T rhs[kRhsRegs];
#pragma unroll
    for (int k = 0; k < kRhsRegs; ++k) {
        const int idx = tid + 256 * k;
        const int r = nc > 0 ? idx / nc : 0, c = nc > 0 ? idx % nc : 0;
        rhs[k] = T(0);
        if (idx < TB * nc && r < rows) {
            const long long src = kLower ? static_cast<long long>(perm[r0 + r]) : (r0 + r);
            rhs[k] = Bin[src * ldb_in + c];
        }
    }
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

    // right-hand side of this block minus the accumulated contributions
#pragma unroll
    for (int k = 0; k < kRhsRegs; ++k) {
        const int idx = tid + 256 * k;
        if (idx < TB * nc) {
            const int r = idx / nc, c = idx % nc;
            Xs[r][c] = (r < rows) ? rhs[k] - Acc[r][c] : T(0);
        }
    }
    __syncthreads();

    // diagonal block: one wave per right-hand-side column, lane = row, the lane's row of the
    // block in registers
    const int lane = tid & 63, wid = tid >> 6;
    if (wid < nc) {
        T drow[TB];
#pragma unroll
        for (int p = 0; p < TB; ++p) drow[p] = Ds[lane][p];
        for (int c = wid; c < nc; c += 4) {
            T x = (lane < rows) ? Xs[lane][c] : T(0);
            if (kLower) {
#pragma unroll
                for (int p = 0; p < TB; ++p) {
                    if (p < rows) {
                        const T xp = read_lane_t<T>(x, p);
                        if (lane > p) x = x - drow[p] * xp;
                    }
                }
            } else {
#pragma unroll
                for (int p = TB - 1; p >= 0; --p) {
                    if (p < rows) {
                        if (lane == p) x = x / drow[p];
                        const T xp = read_lane_t<T>(x, p);
                        if (lane < p) x = x - drow[p] * xp;
                    }
                }
            }
            if (lane < rows) Xs[lane][c] = x;
        }
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

// Single right-hand side (the refinement's and the plain solve's case), latency-optimised.
// One workgroup per 64-row block as above, but:
//  * thread t owns row t / 4 of the block and a quarter (16 entries) of every off-diagonal
//    block row: the off-diagonal block is read straight into registers (the next one while
//    waiting for the current one) and each thread keeps its partial dot product across all
//    blocks, so no LDS staging of the factors and no per-block reduction;
//  * a solved block is published as tagged 16-byte chunks {x, row, tag} (xch, one per
//    row): the consumer's poll and its data read are the same load, with no flag or fence;
//  * the diagonal block row lives in registers of wave 0 (lane = row) and solved values are
//    broadcast with lane reads; the backward pass divides as q = x * (1/d) with one
//    residual correction q += (x - q*d) * (1/d), the reciprocal computed off the chain.
// Bin / Out as in trsv_blocked_kernel with nc = 1; xch holds n chunks whose tags differ
// from `tag` when the launch starts.
template <typename TF, typename T, int kPer>
__device__ __forceinline__ void trsv1_load_row(T (&lreg)[kPer], const TF* __restrict__ lrow, long long c0, long long n,
    bool row_ok)
{
#pragma unroll
    for (int e = 0; e < kPer; ++e) lreg[e] = (row_ok && c0 + e < n) ? static_cast<T>(lrow[c0 + e]) : T(0);
}

template <typename TF, typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv1_kernel(const TF* __restrict__ LU, long long lda, long long n,
    const int* __restrict__ perm, const T* __restrict__ Bin, long long ldb_in, T* __restrict__ Out, long long ld_out,
    Chunk* __restrict__ xch, unsigned tag, unsigned spin_limit, int* __restrict__ err,
    const int* __restrict__ skip_if_nonzero)
{
    if (skip_if_nonzero != nullptr && *skip_if_nonzero != 0) return;
    constexpr int TB = kSolveTB;
    constexpr int kPer = TB / 4;
    __shared__ T Xs[2][TB];
    __shared__ T Xv[TB];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const long long nblk = (n + TB - 1) / TB;
    const long long b = kLower ? blockIdx.x : (nblk - 1 - blockIdx.x);
    const long long r0 = b * TB;
    const int rows = static_cast<int>(min(static_cast<long long>(TB), n - r0));
    const int r = tid >> 2, part = tid & 3;
    const bool row_ok = r < rows;
    const TF* lrow = LU + (r0 + r) * lda;

    // wave 0: its lane's row of the diagonal block and the reciprocal of its diagonal entry
    T drow[TB];
    T d_self = T(1), rcp_self = T(1);
    if (tid < 64) {
#pragma unroll
        for (int p = 0; p < TB; ++p)
            drow[p] = (lane < rows && p < rows) ? static_cast<T>(LU[(r0 + lane) * lda + r0 + p]) : T(0);
        if (!kLower && lane < rows) {
            d_self = static_cast<T>(LU[(r0 + lane) * lda + r0 + lane]);
            rcp_self = T(1) / d_self;
        }
    }
    T rhs = T(0);
    if (part == 0 && row_ok) rhs = Bin[(kLower ? static_cast<long long>(perm[r0 + r]) : (r0 + r)) * ldb_in];

    const long long jcount = kLower ? b : (nblk - 1 - b);
    T part_sum = T(0);
    T lreg[kPer];
    if (jcount > 0) trsv1_load_row<TF, T, kPer>(lreg, lrow, (kLower ? 0 : nblk - 1) * TB + part * kPer, n, row_ok);
    for (long long q = 0; q < jcount; ++q) {
        const long long j = kLower ? q : (nblk - 1 - q);
        const int buf = static_cast<int>(q & 1);
        if (tid < 64) {
            const long long g = j * TB + lane;
            T v = T(0);
            if (g < n) {
                const Chunk* const pc[1] = {xch + g};
                u32x4 c[1];
                unsigned spins = 0;
                for (;;) {
                    ld_chunks<1>(c, pc);
                    if (c[0].w == tag) break;
                    ++spins;
                    if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(err))) {
                        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                v = static_cast<T>(chunk_value(c[0]));
            }
            Xs[buf][lane] = v;
        }
        __syncthreads();
        T cur[kPer];
#pragma unroll
        for (int e = 0; e < kPer; ++e) cur[e] = lreg[e];
        if (q + 1 < jcount)
            trsv1_load_row<TF, T, kPer>(lreg, lrow, (kLower ? q + 1 : nblk - 2 - q) * TB + part * kPer, n, row_ok);
#pragma unroll
        for (int e = 0; e < kPer; ++e) part_sum = fma(cur[e], Xs[buf][part * kPer + e], part_sum);
    }
    // the four partial sums of a row sit in adjacent lanes
    part_sum += __shfl_xor(part_sum, 1, 64);
    part_sum += __shfl_xor(part_sum, 2, 64);
    if (part == 0) Xv[r] = row_ok ? rhs - part_sum : T(0);
    __syncthreads();

    if (tid < 64) {
        T x = Xv[lane];
        if (kLower) {
#pragma unroll
            for (int p = 0; p < TB; ++p) {
                if (p < rows) {
                    const T xp = read_lane_t<T>(x, p);
                    if (lane > p) x = x - drow[p] * xp;
                }
            }
        } else {
#pragma unroll
            for (int p = TB - 1; p >= 0; --p) {
                if (p < rows) {
                    if (lane == p) {
                        const T q0 = x * rcp_self;
                        x = fma(fma(-q0, d_self, x), rcp_self, q0);
                    }
                    const T xp = read_lane_t<T>(x, p);
                    if (lane < p) x = x - drow[p] * xp;
                }
            }
        }
        if (lane < rows) {
            Out[(r0 + lane) * ld_out] = x;
            st_chunk(xch + r0 + lane, static_cast<double>(x), static_cast<int>(r0 + lane), tag);
        }
    }
}

}  // namespace lrn
