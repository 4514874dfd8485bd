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
// (i, piv[i]-1), i = 0..n-1, applied in order, computed in two parallel steps. Step 1: the net
// row moves of every block of 64 pivots (panel_rowmap), all blocks at once, one wave each;
// maps[blk] holds {count, dst[128], src[128]}. Step 2: one workgroup composes the block maps
// in order on a label array (LDS for n <= kPermLdsMax), prefetching the next block's moves.
constexpr int kPermMapInts = 1 + 4 * 64;

static __global__ __launch_bounds__(64) void pivot_block_maps_kernel(const long long* __restrict__ piv, long long n,
    int* __restrict__ maps)
{
    __shared__ long long pivs[64];
    const int lane = threadIdx.x;
    const long long k = static_cast<long long>(blockIdx.x) * 64;
    const int w = static_cast<int>(min(64ll, n - k));
    if (lane < w) pivs[lane] = piv[k + lane] - 1;
    __syncthreads();
    panel_rowmap(pivs, k, w, 128, maps + static_cast<size_t>(blockIdx.x) * kPermMapInts);
}

static __global__ __launch_bounds__(256) void perm_compose_kernel(const int* __restrict__ maps, long long n, long long nblocks,
    int* __restrict__ perm_out, int* __restrict__ perm_tmp)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const bool in_lds = n <= kPermLdsMax;
    int* perm = in_lds ? reinterpret_cast<int*>(lds_raw) : perm_tmp;
    const int t = threadIdx.x;
    for (long long i = t; i < n; i += 256) perm[i] = static_cast<int>(i);
    // thread t < 128 handles move t of a block
    int cnt = 0, dst = 0, src = 0;
    if (nblocks > 0) {
        cnt = maps[0];
        if (t < 128) {
            dst = maps[1 + t];
            src = maps[1 + 128 + t];
        }
    }
    __syncthreads();
    for (long long blk = 0; blk < nblocks; ++blk) {
        int ncnt = 0, ndst = 0, nsrc = 0;
        if (blk + 1 < nblocks) {
            const int* rm = maps + static_cast<size_t>(blk + 1) * kPermMapInts;
            ncnt = rm[0];
            if (t < 128) {
                ndst = rm[1 + t];
                nsrc = rm[1 + 128 + t];
            }
        }
        int v = 0;
        if (t < cnt) v = perm[src];
        __syncthreads();
        if (t < cnt) perm[dst] = v;
        __syncthreads();
        cnt = ncnt;
        dst = ndst;
        src = nsrc;
    }
    for (long long i = t; i < n; i += 256) perm_out[i] = perm[i];
}

// One workgroup per 64-row block; block b waits (bounded) for the blocks it depends on,
// which always have a lower workgroup index, so the kernel drains in dispatch order.
//
// Forward:  Y[b] = L[b][b]^-1 ( B[perm][b] - sum_{j<b} L[b][j] Y[j] )   (unit lower)
// Backward: X[b] = U[b][b]^-1 ( Y[b]       - sum_{j>b} U[b][j] X[j] )   (B receives X)
// In general kLower picks the direction (increasing rows), kUnit an implicit unit diagonal
// (otherwise a true division), kTrans reads the operator transposed (entry (i, j) is
// A[j][i]), and perm == nullptr the identity: the triangular solver of the library uses it
// with the reference's four flag combinations.
// `nc` right-hand-side columns starting at column c0 of B / Y (row strides ldb / ldy).
// The factors have type TF and are converted to the working type T as they are staged
// (TF = float, T = double applies low-precision factors in FP64 arithmetic).
//
// Critical path: block b finishes one step after block b-1 (b+1 backward), so everything
// that does not depend on the previous block -- the diagonal block, this block's right-hand
// side, the earlier off-diagonal blocks -- is loaded or applied before the wait, and the
// diagonal solve keeps the block row in registers and broadcasts each solved value with a
// lane read instead of an LDS round trip.
template <typename TF, typename T, bool kLower, bool kUnit = kLower, bool kTrans = false>
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

    // entry (i, j) of the triangular operator: A[i][j], or A[j][i] when transposed
    auto coef = [&](long long i, long long j) { return static_cast<T>(kTrans ? LU[j * lda + i] : LU[i * lda + j]); };
    // independent of every other block: diagonal block (its own triangle only; a unit
    // diagonal is not read) and right-hand side
    for (int idx = tid; idx < TB * TB; idx += 256) {
        const int r = idx / TB, c = idx % TB;
        const bool in = (r < rows && c < rows) && (kLower ? (c < r || (!kUnit && c == r)) : (c > r || (!kUnit && c == r)));
        Ds[r][c] = in ? coef(r0 + r, r0 + c) : T(0);
    }
    T rhs[kRhsRegs];
#pragma unroll
    for (int k = 0; k < kRhsRegs; ++k) {
        const int idx = tid + 256 * k;
        const int r = nc > 0 ? idx / nc : 0, c = nc > 0 ? idx % nc : 0;
        rhs[k] = T(0);
        if (idx < TB * nc && r < rows) {
            const long long src = (kLower && perm != nullptr) ? static_cast<long long>(perm[r0 + r]) : (r0 + r);
            rhs[k] = Bin[src * ldb_in + c];
        }
    }
    for (int idx = tid; idx < TB * kSolveMaxCols; idx += 256) Acc[idx / kSolveMaxCols][idx % kSolveMaxCols] = T(0);
    // Acc is cleared with a different thread-to-entry mapping than the one that reads it
    // below; with no earlier block (jcount == 0) no other barrier separates the two
    __syncthreads();

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
            Ls[r][c] = (r < rows && c < cols) ? coef(r0 + r, c0 + c) : T(0);
        }
        if (tid == 0) wait_flag_coherent(&flags[j], 1u, spin_limit, err);
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
                        if (!kUnit && lane == p) x = x / drow[p];
                        const T xp = read_lane_t<T>(x, p);
                        if (lane > p) x = x - drow[p] * xp;
                    }
                }
            } else {
#pragma unroll
                for (int p = TB - 1; p >= 0; --p) {
                    if (p < rows) {
                        if (!kUnit && lane == p) x = x / drow[p];
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
    wait_own_stores();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&flags[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
                if (spin_limit == 0u) raise_error(err);  // forced expiry (tests)
                for (;;) {
                    ld_chunks<1>(c, pc);
                    if (c[0].w == tag) break;
                    ++spins;
                    if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(err))) {
                        raise_error(err);
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

// Single right-hand side in the inverse form (the mixed solver's refinement, FP32 factors TF
// applied in the working type T). Same 64-row blocks, thread layout and tagged-chunk hand-off
// as trsv1_kernel, with the inverses of the diagonal blocks (dinvT, stored transposed as by
// trsm_diag_inv_kernel: element b * 4096 + j * 64 + i is inv(D_b)[i][j]) and the products
// M1_b = inv(D_b) A[b][p1], M2_b = inv(D_b) A[b][p2] with the two chain predecessors
// p1 = b -+ 1, p2 = b -+ 2 (pred1T, pred2T, same layout, trsv_pred_kernel) computed once per
// factorisation. Block b is
//     x_b = z_b - M1_b x_p1 - M2_b x_p2,
//     z_b = inv(D_b) (rhs_b - sum of A[b][j] x_j over the blocks before p2)
//  * the off-diagonal blocks before the last three are taken kTrsvGroup at a time: the factor
//    loads of a group and the polls of its solved values (one chunk per block per lane of
//    wave 0) are all in flight together, one memory round trip per group;
//  * the block before p2 is taken alone, so z_b is ready two hand-offs before it is needed;
//  * x_p1 (and x_p2, arrived one hand-off earlier): each wave polls the values itself, so the
//    hand-off is followed by 32 products, two lane exchanges and the store, with no workgroup
//    barrier and no substitution.
// Thread (r, part) holds row r, columns 16 part .. 16 part + 15, of inv(D_b), M1_b and M2_b in
// registers from the start. Products of the blocks before p2 are accumulated in block order.
// xcd_per > 0: workgroup i solves chain position (i % 8) * xcd_per + i / 8 (grid 8 * xcd_per,
// positions >= nblk exit), so each run of xcd_per consecutive blocks is solved on one XCD
// (workgroup i runs on XCD i % 8); every workgroup must then be resident at once (checked by
// the host). xcd_per = 0: position i (the dispatch order, no residency needed).
template <typename TF, int kPer>
__device__ __forceinline__ void trsv1_load_raw(TF (&lreg)[kPer], const TF* __restrict__ lrow, long long c0, long long n,
    bool row_ok)
{
    if (row_ok && c0 + kPer <= n) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) lreg[e] = lrow[c0 + e];
    } else {
#pragma unroll
        for (int e = 0; e < kPer; ++e) lreg[e] = (row_ok && c0 + e < n) ? lrow[c0 + e] : TF(0);
    }
}

constexpr int kTrsvGroup = 8;  // off-diagonal blocks per group: 8 x 16 factor entries per thread

template <typename TF, typename T, bool kLower>
__global__ __launch_bounds__(256) void trsv1_inv_kernel(const TF* __restrict__ LU, long long lda, long long n,
    const int* __restrict__ perm, const T* __restrict__ Bin, long long ldb_in, T* __restrict__ Out, long long ld_out,
    Chunk* __restrict__ xch, unsigned tag, unsigned spin_limit, int* __restrict__ err,
    const int* __restrict__ skip_if_nonzero, const T* __restrict__ dinvT, const T* __restrict__ pred1T,
    const T* __restrict__ pred2T, int xcd_per)
{
    if (skip_if_nonzero != nullptr && *skip_if_nonzero != 0) return;
    constexpr int TB = kSolveTB;
    constexpr int kPer = TB / 4;
    constexpr int G = kTrsvGroup;
    __shared__ T Xs[TB];
    __shared__ T Xg[2][G][TB];
    __shared__ T Xv[TB];
    __shared__ T Xw[4][2][TB];  // each wave's copy of the two predecessors' values

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const long long nblk = (n + TB - 1) / TB;
    const long long chain_pos = xcd_per > 0
        ? static_cast<long long>(blockIdx.x % kXcds) * xcd_per + blockIdx.x / kXcds
        : static_cast<long long>(blockIdx.x);
    if (chain_pos >= nblk) return;
    const long long b = kLower ? chain_pos : (nblk - 1 - chain_pos);
    const long long r0 = b * TB;
    const int rows = static_cast<int>(min(static_cast<long long>(TB), n - r0));
    const int r = tid >> 2, part = tid & 3;
    const bool row_ok = r < rows;
    const TF* lrow = LU + (r0 + r) * lda;

    // number of earlier blocks in the chain order; the q-th of them is jblk(q)
    const long long jcount = kLower ? b : (nblk - 1 - b);
    auto jblk = [&](long long q) -> long long { return kLower ? q : (nblk - 1 - q); };

    T dinv_reg[kPer], m1_reg[kPer], m2_reg[kPer];
    {
        const T* dp = dinvT + b * (TB * TB) + r;
        const T* m1 = pred1T + b * (TB * TB) + r;
        const T* m2 = pred2T + b * (TB * TB) + r;
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            dinv_reg[e] = dp[(part * kPer + e) * TB];
            m1_reg[e] = jcount > 0 ? m1[(part * kPer + e) * TB] : T(0);
            m2_reg[e] = jcount > 1 ? m2[(part * kPer + e) * TB] : T(0);
        }
    }
    T rhs = T(0);
    if (part == 0 && row_ok) rhs = Bin[(kLower ? static_cast<long long>(perm[r0 + r]) : (r0 + r)) * ldb_in];

    T part_sum = T(0);
    const long long q_alone = jcount - 3;  // the block before the two predecessors (none when < 0)
    TF lfin[kPer];
    if (q_alone >= 0) trsv1_load_raw<TF, kPer>(lfin, lrow, jblk(q_alone) * TB + part * kPer, n, row_ok);
    int gbuf = 0;
    for (long long q0 = 0; q0 < q_alone; q0 += G, gbuf ^= 1) {
        const int cnt = static_cast<int>(min(static_cast<long long>(G), q_alone - q0));
        TF lg[G][kPer];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (g < cnt) {
                trsv1_load_raw<TF, kPer>(lg[g], lrow, jblk(q0 + g) * TB + part * kPer, n, row_ok);
            } else {
#pragma unroll
                for (int e = 0; e < kPer; ++e) lg[g][e] = TF(0);
            }
        }
        if (tid < 64) {
            // the chunks this lane needs; padding entries repeat a needed chunk (row 0 of the
            // group's first block always exists)
            const Chunk* pc[G];
            const Chunk* const first = xch + jblk(q0) * TB;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const long long row = jblk(q0 + g) * TB + lane;
                pc[g] = (g < cnt && row < n) ? xch + row : first;
            }
            u32x4 c[G];
            spin_chunks<G>(c, pc, tag, spin_limit, err);
#pragma unroll
            for (int g = 0; g < G; ++g) Xg[gbuf][g][lane] = static_cast<T>(chunk_value(c[g]));
        }
        __syncthreads();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (g < cnt) {
#pragma unroll
                for (int e = 0; e < kPer; ++e)
                    part_sum = fma(static_cast<T>(lg[g][e]), Xg[gbuf][g][part * kPer + e], part_sum);
            }
        }
    }
    if (q_alone >= 0) {
        if (tid < 64) {
            const long long row = jblk(q_alone) * TB + lane;
            T v = T(0);
            if (row < n) {
                u32x4 c;
                spin_chunk1(c, xch + row, tag, spin_limit, err);
                v = static_cast<T>(chunk_value(c));
            }
            Xs[lane] = v;
        }
        __syncthreads();
#pragma unroll
        for (int e = 0; e < kPer; ++e) part_sum = fma(static_cast<T>(lfin[e]), Xs[part * kPer + e], part_sum);
    }
    // the four partial sums of a row sit in adjacent lanes
    part_sum += __shfl_xor(part_sum, 1, 64);
    part_sum += __shfl_xor(part_sum, 2, 64);
    if (part == 0) Xv[r] = row_ok ? rhs - part_sum : T(0);
    __syncthreads();

    // z_b (the inverse's padding rows and columns are zero)
    T x = T(0);
#pragma unroll
    for (int e = 0; e < kPer; ++e) x = fma(dinv_reg[e], Xv[part * kPer + e], x);
    x += __shfl_xor(x, 1, 64);
    x += __shfl_xor(x, 2, 64);
    if (jcount > 0) {
        // x_b = z_b - M1_b x_p1 - M2_b x_p2: every wave polls the 2 x 64 values itself (x_p2
        // has normally arrived long before)
        const long long prow1 = jblk(jcount - 1) * TB + lane;
        const long long prow2 = jcount > 1 ? jblk(jcount - 2) * TB + lane : n;
        const bool has2 = prow2 < n;
        T v1 = T(0), v2 = T(0);
        if (prow1 < n) {
            const Chunk* pc[2] = {xch + prow1, xch + (has2 ? prow2 : prow1)};
            u32x4 c[2];
            spin_chunks<2>(c, pc, tag, spin_limit, err);
            v1 = static_cast<T>(chunk_value(c[0]));
            if (has2) v2 = static_cast<T>(chunk_value(c[1]));
        } else if (has2) {
            u32x4 c;
            spin_chunk1(c, xch + prow2, tag, spin_limit, err);
            v2 = static_cast<T>(chunk_value(c));
        }
        Xw[wid][0][lane] = v1;
        Xw[wid][1][lane] = v2;
        // the wave's own LDS writes, visible to its other lanes
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        T d1 = T(0), d2 = T(0);
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            d1 = fma(m1_reg[e], Xw[wid][0][part * kPer + e], d1);
            d2 = fma(m2_reg[e], Xw[wid][1][part * kPer + e], d2);
        }
        T d = d1 + d2;
        d += __shfl_xor(d, 1, 64);
        d += __shfl_xor(d, 2, 64);
        x -= d;
    }
    if (part == 0 && row_ok) {
        Out[(r0 + r) * ld_out] = x;
        st_chunk(xch + r0 + r, static_cast<double>(x), static_cast<int>(r0 + r), tag);
    }
}

// Single right-hand side with FP32 factors applied in the working type T (the refinement of
// the mixed solver): 256-row workgroups, so the chain of hand-offs between workgroups has
// n/256 links instead of n/64.
//  * off-diagonal part: the solved values arrive as tagged chunks in 64-row sub-blocks
//    (wave 0 polls one chunk per lane). Wave w covers rows 64w..64w+63 of the block; lane l
//    reads columns 4(l%16)..+3 of rows 4i + l/16 (i < 16), so every load instruction reads
//    four full 256-byte row pieces (coalesced), one sub-block ahead of its use; each lane
//    keeps 16 running partial sums, combined through LDS once all sub-blocks are in;
//  * diagonal block: four 64-row sub-blocks solved in turn, each by one wave with its row of
//    the diagonal sub-block in registers and lane reads for the solved values; the six
//    off-diagonal 64 x 64 pieces inside the block are staged in LDS at the start, and after
//    each sub-block the waves of the later sub-blocks take its contribution at once; every
//    solved sub-block is published immediately, so the next workgroup starts on it while
//    this one solves the rest.
// Lower pass: rows in increasing order, unit diagonal. Upper pass: decreasing order,
// q = x * (1/d) with one residual correction.
constexpr int kChainSB = 64;

// Slot of the off-diagonal piece (s, s') inside an NS-sub-block workgroup block; lower:
// s' < s; upper: s' > s, mirrored onto the same slots.
template <int NS> __device__ __forceinline__ int chain_piece(int s, int sp, bool lower)
{
    const int a = lower ? s : NS - 1 - s, c = lower ? sp : NS - 1 - sp;
    return a * (a - 1) / 2 + c;
}

// NS sub-blocks of 64 rows per workgroup (NS * 64 threads, one row each): 4 (256 rows) or
// 2 (128 rows), which halves the per-workgroup stream of factor rows for the same halving
// of hand-offs.
template <typename T, bool kLower, bool kVec4, int NS>
__global__ __launch_bounds__(kChainSB * NS) void trsv_chain_kernel(const float* __restrict__ LU, long long lda,
    long long n, const int* __restrict__ perm, const T* __restrict__ Bin, T* __restrict__ Out,
    Chunk* __restrict__ xch, unsigned tag, unsigned spin_limit, int* __restrict__ err,
    const int* __restrict__ skip_if_nonzero)
{
    if (skip_if_nonzero != nullptr && *skip_if_nonzero != 0) return;
    constexpr int SB = kChainSB, TB = NS * SB, NP = NS * (NS - 1) / 2;
    static_assert(NS >= 2 && NS <= 4, "2 to 4 sub-blocks");
    __shared__ float Lp[NP][SB][SB + 1];
    __shared__ T Xs[2][SB];
    __shared__ T Xd[NS][SB];
    __shared__ T Pp[TB][17];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const long long nblk = (n + TB - 1) / TB;
    const long long b = kLower ? blockIdx.x : (nblk - 1 - blockIdx.x);
    const long long r0 = b * TB;
    const int rows = static_cast<int>(min(static_cast<long long>(TB), n - r0));
    const int nsub = (rows + SB - 1) / SB;
    const long long nsub_total = (n + SB - 1) / SB;
    const bool row_ok = tid < rows;
    const long long row = r0 + tid;
    const float* lrow = LU + (row_ok ? row : r0) * lda;

    for (int idx = tid; idx < NP * SB * SB; idx += TB) {
        const int piece = idx / (SB * SB);
        const int rem = idx - piece * SB * SB;
        const int rr = rem / SB, cc = rem - rr * SB;
        // piece -> (a, c) with c < a in the mirrored numbering
        const int a = piece < 1 ? 1 : (piece < 3 ? 2 : 3);
        const int c = piece - a * (a - 1) / 2;
        const int s = kLower ? a : NS - 1 - a, sp = kLower ? c : NS - 1 - c;
        const long long gr = r0 + s * SB + rr, gc = r0 + sp * SB + cc;
        Lp[piece][rr][cc] = (gr < n && gc < n) ? LU[gr * lda + gc] : 0.f;
    }
    float drow[SB];
#pragma unroll
    for (int p = 0; p < SB; ++p) {
        const long long gc = r0 + wid * SB + p;
        drow[p] = (row_ok && gc < n) ? lrow[gc] : 0.f;
    }
    T d_self = T(1), rcp_self = T(1);
    if (!kLower && row_ok) {
        d_self = static_cast<T>(lrow[row]);
        rcp_self = T(1) / d_self;
    }
    T rhs = T(0);
    if (row_ok) rhs = Bin[kLower ? static_cast<long long>(perm[row]) : row];

    // off-diagonal sub-blocks, in the order they are solved; the row pieces alternate
    // between two register buffers (one in use, the next one loading)
    const long long jcount = kLower ? b * NS : (nsub_total - (b + 1) * NS);
    const int rq = lane >> 4, cq = lane & 15;
    T acc[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = T(0);
    float la[16][4], lb[16][4];
    auto sub_at = [&](long long q) { return kLower ? q : (nsub_total - 1 - q); };
    auto load_piece = [&](float (&dst)[16][4], long long j) {
        const long long c0 = j * SB + 4 * cq;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int lr = wid * SB + 4 * i + rq;
            const bool ok = lr < rows;
            const float* src = LU + (r0 + (ok ? lr : 0)) * lda + c0;
            if (kVec4) {
                // c0 + 3 < n whenever c0 < n: n is a multiple of 4 for this variant
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (ok && c0 < n) v = *reinterpret_cast<const float4*>(src);
                dst[i][0] = v.x;
                dst[i][1] = v.y;
                dst[i][2] = v.z;
                dst[i][3] = v.w;
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) dst[i][k] = (ok && c0 + k < n) ? src[k] : 0.f;
            }
        }
    };
    auto step = [&](long long q, const float (&cur)[16][4], float (&nxt)[16][4]) {
        const long long j = sub_at(q);
        const int buf = static_cast<int>(q & 1);
        if (wid == 0) {
            const long long g = j * SB + lane;
            T v = T(0);
            if (g < n) {
                const Chunk* const pc[1] = {xch + g};
                u32x4 c[1];
                unsigned spins = 0;
                if (spin_limit == 0u) raise_error(err);  // forced expiry (tests)
                for (;;) {
                    ld_chunks<1>(c, pc);
                    if (c[0].w == tag) break;
                    ++spins;
                    if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(err))) {
                        raise_error(err);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
                v = static_cast<T>(chunk_value(c[0]));
            }
            Xs[buf][lane] = v;
        }
        __syncthreads();
        if (q + 1 < jcount) load_piece(nxt, sub_at(q + 1));
        T xv[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) xv[k] = Xs[buf][4 * cq + k];
#pragma unroll
        for (int i = 0; i < 16; ++i)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[i] = fma(static_cast<T>(cur[i][k]), xv[k], acc[i]);
    };
    if (jcount > 0) load_piece(la, sub_at(0));
    for (long long q = 0; q < jcount; q += 2) {
        step(q, la, lb);
        if (q + 1 < jcount) step(q + 1, lb, la);
    }
    // combine the 16 partial sums of every row
#pragma unroll
    for (int i = 0; i < 16; ++i) Pp[wid * SB + 4 * i + rq][cq] = acc[i];
    __syncthreads();
    T part = T(0);
#pragma unroll
    for (int k = 0; k < 16; ++k) part += Pp[tid][k];

    // diagonal block, sub-block by sub-block
    for (int si = 0; si < NS; ++si) {
        const int s = kLower ? si : (NS - 1 - si);
        if (s >= nsub) continue;
        __syncthreads();
        if (wid == s) {
            T x = row_ok ? rhs - part : T(0);
            const int live = rows - s * SB;  // rows of this sub-block
            if (kLower) {
#pragma unroll
                for (int p = 0; p < SB; ++p) {
                    if (p < live) {
                        const T xp = read_lane_t<T>(x, p);
                        if (lane > p) x = x - static_cast<T>(drow[p]) * xp;
                    }
                }
            } else {
#pragma unroll
                for (int p = SB - 1; p >= 0; --p) {
                    if (p < live) {
                        if (lane == p) {
                            const T q0 = x * rcp_self;
                            x = fma(fma(-q0, d_self, x), rcp_self, q0);
                        }
                        const T xp = read_lane_t<T>(x, p);
                        if (lane < p) x = x - static_cast<T>(drow[p]) * xp;
                    }
                }
            }
            Xd[s][lane] = row_ok ? x : T(0);
            if (row_ok) {
                Out[row] = x;
                st_chunk(xch + row, static_cast<double>(x), static_cast<int>(row), tag);
            }
        }
        __syncthreads();
        const bool later = kLower ? (wid > s) : (wid < s);
        if (later && wid < nsub) {
            const int piece = chain_piece<NS>(wid, s, kLower);
#pragma unroll
            for (int e0 = 0; e0 < SB; e0 += 16) {
#pragma unroll
                for (int e = e0; e < e0 + 16; ++e) part = fma(static_cast<T>(Lp[piece][lane][e]), Xd[s][e], part);
                asm volatile("" ::: "memory");
            }
        }
    }
}

}  // namespace lrn
