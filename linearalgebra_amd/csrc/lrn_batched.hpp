// Batched LU solve of many small independent systems of uneven size (n <= 256, nrhs <= 8)
// in one persistent kernel.
//
// batch_order_kernel sorts the systems by decreasing order (counting sort over n) and
// resets the work counter, so it can sit in a captured graph without a memset node.
// batched_gesv_kernel: every workgroup (256 threads, two per compute unit) repeatedly takes the
// next system from the sorted list (one atomic per system) and solves it alone, right-looking
// in 32-column steps:
//   row map       rows are never moved while the system is factored: an LDS map gives the
//                 physical row of every logical row, each pivot swaps two map entries, and one
//                 pass at the end writes the rows in logical order (the LAPACK layout);
//   panel         one row per thread in registers, never handed over: an interchange only
//                 swaps the logical positions of two rows; per column a DPP wave argmax, the
//                 wave's candidate row staged in double-buffered LDS slots, ONE workgroup
//                 barrier, then every thread picks the winner of the four waves and applies the
//                 rank-1 update (the reference's pivot rule: first maximum of |a| in logical
//                 row order wins, NaN never chosen, exactly-zero pivots skipped and reported,
//                 reciprocal multiplier unless |d| < T_MIN);
//   U12           one column per thread, forward substitution with L11 in the reference's TRSM
//                 row order;
//   A22 -= L21 U12  L21 staged transposed in LDS, U12 in 64-column chunks; each wave owns
//                 32 x 32 tiles of the update, on the FP64 matrix cores (4 x 16 x 4 updates of
//                 v_mfma_f64_4x4x4_4b, as in the trailing-update GEMM) or, in FP32, vector FMAs;
//                 the tile's old values are loaded before its products and the product is
//                 subtracted once (delta-then-subtract);
//   solve         only when the factorisation is nonsingular (the reference's LU mode): B rows
//                 permuted by the pivots, blocked unit-lower forward and upper backward
//                 substitution (true division by U[i][i]) in LDS.
// Algorithm followed: src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232,
// 247-271, 329-429 and src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113,
// 160-171 (paths relative to the reference repository).
#pragma once

#include <type_traits>

#include "lrn_device.hpp"
#include "lrn_gemm_f64.hpp"

namespace lrn {

constexpr int kBatMaxN = 256;
constexpr int kBatMaxRhs = 8;
constexpr int kBatThreads = 256;
constexpr int kBatDiag = 16;   // diagonal block of the substitutions
constexpr int kBatNB = 32;     // panel width
constexpr int kBatUC = 64;     // columns of U12 staged per chunk of the update
constexpr int kBatLtLd = kBatMaxN - kBatNB + 1;  // row stride of the transposed L21 (odd)

// Bucket of a system for the counting sort: 0 for n = 256 ... 255 for n = 1; 256 = invalid.
__device__ __forceinline__ int batch_bucket(int n, int nrhs)
{
    if (n < 1 || n > kBatMaxN || nrhs < 0 || nrhs > kBatMaxRhs) return kBatMaxN;
    return kBatMaxN - n;
}

__global__ __launch_bounds__(1024) void batch_order_kernel(const int* __restrict__ nv, const int* __restrict__ nrhsv,
    int count, int* __restrict__ order, unsigned* __restrict__ ctr, int* __restrict__ info)
{
    __shared__ int hist[kBatMaxN + 1];
    __shared__ int base[kBatMaxN + 1];
    const int tid = threadIdx.x;
    for (int i = tid; i <= kBatMaxN; i += 1024) hist[i] = 0;
    __syncthreads();
    for (int i = tid; i < count; i += 1024) atomicAdd(&hist[batch_bucket(nv[i], nrhsv[i])], 1);
    __syncthreads();
    if (tid == 0) {
        int s = 0;
        for (int k = 0; k <= kBatMaxN; ++k) {
            base[k] = s;
            s += hist[k];
        }
        *ctr = 0u;
    }
    __syncthreads();
    for (int i = tid; i < count; i += 1024) {
        const int key = batch_bucket(nv[i], nrhsv[i]);
        order[atomicAdd(&base[key], 1)] = i;
        if (key == kBatMaxN) info[i] = -1;
    }
}

// LDS of one workgroup. The solve's right-hand sides reuse the transposed-L21 area.
template <typename T> struct BatLds {
    int phys[kBatMaxN];               // logical row -> physical row
    T lt[kBatNB][kBatLtLd];           // L21 transposed: lt[p][r] = L[k + w + r][k + p]
    union {
        T l11[kBatNB][kBatNB + 1];    // the panel's unit-lower block (U12 phase)
        T us[kBatNB][kBatUC + 1];     // a 64-column chunk of U12 (update phase)
    };
    T cand[2][kBatThreads / kWave][kBatNB];  // each wave's candidate row, by column parity
    double key[2][kBatThreads / kWave];
    int kidx[2][kBatThreads / kWave];
    int sys;
    int info;
};

// One 32 x 32 tile of A22 -= L21 U12 for one wave: rows r0.. of the update (logical rows
// k + w + r), columns c0.. (logical columns k + w + c; column cc0 of the staged U12 chunk).
// Result element (g, c) of a lane: tile row 4g + (lane >> 4), tile column 16c + (lane & 15).
template <typename T>
__device__ __forceinline__ void bat_update_tile(T* __restrict__ A, int n, int kw, const BatLds<T>& S, int r0, int c0,
    int cc0, int mr, int nc, int lane)
{
    T acc[8][2], cv[8][2];
    T* rowp[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        const int r = r0 + 4 * g + (lane >> 4);
        rowp[g] = r < mr ? A + static_cast<long long>(S.phys[kw + r]) * n + kw : nullptr;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int col = c0 + 16 * c + (lane & 15);
            acc[g][c] = T(0);
            cv[g][c] = (rowp[g] != nullptr && col < nc) ? rowp[g][col] : T(0);
        }
    }
    if constexpr (std::is_same<T, double>::value) {
        // matrix cores: A operand L[row 4g + (lane & 3)][k + (lane >> 4)], B operand
        // U[k + (lane >> 4)][column 16c + (lane & 15)]
#pragma unroll
        for (int kk = 0; kk < kBatNB; kk += 4) {
            const int kl = kk + (lane >> 4);
            double b[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) b[c] = S.us[kl][cc0 + 16 * c + (lane & 15)];
#pragma unroll
            for (int g = 0; g < 8; ++g) {
                const double a = S.lt[kl][r0 + 4 * g + (lane & 3)];
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[g][c] = mfma4x4(a, b[c], acc[g][c]);
            }
        }
    } else {
        // vector FMAs, each lane its own 8 x 2 elements, k increasing
#pragma unroll 2
        for (int kk = 0; kk < kBatNB; ++kk) {
            T b[2];
#pragma unroll
            for (int c = 0; c < 2; ++c) b[c] = S.us[kk][cc0 + 16 * c + (lane & 15)];
#pragma unroll
            for (int g = 0; g < 8; ++g) {
                const T a = S.lt[kk][r0 + 4 * g + (lane >> 4)];
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[g][c] = fma(a, b[c], acc[g][c]);
            }
            asm volatile("" ::: "memory");  // one k of operands in flight at a time
        }
    }
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        if (rowp[g] == nullptr) continue;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int col = c0 + 16 * c + (lane & 15);
            if (col < nc) rowp[g][col] = cv[g][c] - acc[g][c];
        }
    }
}

// The panel of one step. Thread tid holds the row it loaded (logical row k + tid at the start
// of the step) and never hands it over: an interchange only changes which logical position
// each row holds (pos, tracked by every thread for its own row, and the LDS row map), so no
// row data moves between threads. The w columns are eliminated in groups of eight: a runtime
// loop over the groups, the eight columns of a group by a compile-time recursion (step<J>), so
// every index into the register window v is a constant and the code stays small. v[c] is
// panel column col0 + c; after each group its first eight columns are final for every row,
// are stored to the row, and the window shifts by eight.
template <typename T> struct BatPanel {
    T v[kBatNB];
    BatLds<T>* S;
    long long* pv;
    int k, w, m, tid, lane, wid;
    int pos;   // logical position (relative to k) of this thread's row; >= m for no row
    int col0;  // panel column of v[0]

    template <int J> __device__ __forceinline__ void step()
    {
        constexpr int NB = kBatNB;
        constexpr int NW = kBatThreads / kWave;
        constexpr int par = J & 1;  // groups have an even number of columns
        const int j = col0 + J;     // panel column
        // candidates: the rows not yet chosen (logical position >= j), ordered by |a| and then
        // by logical position (the reference's first-maximum rule); lanes are not in row order,
        // so the lowest position among the maxima comes from a second reduction
        const double key0 = (pos >= j && pos < m) ? pivot_key(v[J], pos == j) : -2.0;
        const double key = wave_max_f64(key0);
        const double negpos = wave_max_f64(key0 == key ? -static_cast<double>(pos) : -1e300);
        const int ri = static_cast<int>(-negpos);
        if (key >= -1.0 && key0 == key && pos == ri) {
#pragma unroll
            for (int c = 0; c < NB; ++c) S->cand[par][wid][c] = v[c];
        }
        if (lane == 0) {
            S->key[par][wid] = key;
            S->kidx[par][wid] = ri;
        }
        __syncthreads();
        double bk = S->key[par][0];
        int bi = S->kidx[par][0], bw = 0;
#pragma unroll
        for (int ww = 1; ww < NW; ++ww) {
            const double kv = S->key[par][ww];
            const int ki = S->kidx[par][ww];
            if (argmax_better(kv, ki, bk, bi)) {
                bk = kv;
                bi = ki;
                bw = ww;
            }
        }
        const int p = bi;  // logical position of the pivot row (m >= w > j, so one exists)
        const T* prow = S->cand[par][bw];
        // the interchange of logical rows j and p: positions only
        const int was = pos;
        if (was == p) pos = j;
        if (was == j) pos = p;
        if (tid == 0) {
            pv[k + j] = static_cast<long long>(k + p + 1);
            if (p != j) {
                const int t2 = S->phys[k + j];
                S->phys[k + j] = S->phys[k + p];
                S->phys[k + p] = t2;
            }
        }
        const T d = prow[J];
        if (d == T(0)) {
            if (tid == 0 && S->info == 0) S->info = k + j + 1;
        } else if (pos > j && pos < m) {
            const bool use_recip = fabs(static_cast<double>(d)) >= static_cast<double>(TypeMin<T>::v);
            const T rcp = use_recip ? T(1) / d : T(1);
            const T l = use_recip ? v[J] * rcp : v[J] / d;
            v[J] = l;
            // columns in groups of eight, so only eight pivot-row values are live
#pragma unroll
            for (int c0 = (J + 1) / 8 * 8; c0 < NB; c0 += 8) {
#pragma unroll
                for (int c = c0; c < c0 + 8; ++c)
                    if (c > J) v[c] = v[c] - l * prow[c];
                asm volatile("" ::: "memory");
            }
        }
    }

    template <int J> __device__ __forceinline__ void group()
    {
        if constexpr (J < 8) {
            if (col0 + J < w) {
                step<J>();
                group<J + 1>();
            }
        }
    }

    // Eliminates all w columns; `row` is this thread's row at panel column 0 (null for no row).
    __device__ __forceinline__ void run(T* row)
    {
        for (col0 = 0; col0 < w; col0 += 8) {
            asm volatile("v_mov_b32 %0, %0" : "+v"(tid));  // keep the lane masks inside the loop
            group<0>();
            if (row != nullptr) {
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    if (col0 + c < w) row[col0 + c] = v[c];
            }
#pragma unroll
            for (int c = 0; c < kBatNB - 8; ++c) v[c] = v[c + 8];
#pragma unroll
            for (int c = kBatNB - 8; c < kBatNB; ++c) v[c] = T(0);
        }
    }
};

// U12 column c: u <- L11^-1 u, rows in increasing order, each row's products in increasing
// order (the reference's TRSM order); a compile-time recursion over the rows keeps u in
// registers.
template <typename T, int I>
__device__ __forceinline__ void bat_trsm_rows(T (&u)[kBatNB], const T (&l11)[kBatNB][kBatNB + 1], int w)
{
    if constexpr (I < kBatNB) {
        if (I < w) {
#pragma unroll
            for (int pp = 0; pp < I; ++pp) u[I] = u[I] - l11[I][pp] * u[pp];
            asm volatile("" ::: "memory");  // one row of L11 in flight at a time
            bat_trsm_rows<T, I + 1>(u, l11, w);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBatThreads, 2) void batched_gesv_kernel(T* __restrict__ a,
    const long long* __restrict__ a_off, long long* __restrict__ piv, const long long* __restrict__ piv_off,
    T* __restrict__ b, const long long* __restrict__ b_off, const int* __restrict__ nv, const int* __restrict__ nrhsv,
    int* __restrict__ info, const int* __restrict__ order, unsigned* __restrict__ ctr, int count)
{
    constexpr int NB = kBatNB;
    constexpr int NW = kBatThreads / kWave;
    __shared__ BatLds<T> S;
    T (*const Bs)[kBatMaxRhs + 1] = reinterpret_cast<T (*)[kBatMaxRhs + 1]>(&S.lt[0][0]);
    static_assert(sizeof(T) * kBatMaxN * (kBatMaxRhs + 1) <= sizeof(S.lt), "B block must fit in the L21 area");

    const int tid0 = threadIdx.x;

    for (;;) {
        // the thread indices pass through an opaque move per system, so the lane masks and LDS
        // addresses derived from them are formed where they are used instead of being hoisted
        // out of the persistent loop (where they would occupy registers throughout)
        int tid = tid0, lane, wid;
        auto launder = [&]() {
            asm volatile("v_mov_b32 %0, %0" : "+v"(tid));
            lane = tid & (kWave - 1);
            wid = tid / kWave;
        };
        launder();
        if (tid == 0) S.sys = static_cast<int>(atomicAdd(ctr, 1u));
        __syncthreads();
        const int q = S.sys;
        if (q >= count) break;
        const int sidx = order[q];
        const int n = nv[sidx], nrhs = nrhsv[sidx];
        if (batch_bucket(n, nrhs) == kBatMaxN) {
            __syncthreads();
            continue;
        }
        T* const A = a + a_off[sidx];
        long long* const pv = piv + piv_off[sidx];
        T* const B = b + b_off[sidx];
        if (tid == 0) S.info = 0;
        S.phys[tid] = tid;  // kBatThreads == kBatMaxN
        __syncthreads();

        for (int k = 0; k < n; k += NB) {
            launder();  // per step as well: the per-column masks are formed inside the step
            const int w = min(NB, n - k);
            const int m = n - k;
            // ---------------------------------------------------------------- panel
            BatPanel<T> P;
            P.S = &S;
            P.pv = pv;
            P.k = k;
            P.w = w;
            P.m = m;
            P.tid = tid;
            P.lane = lane;
            P.wid = wid;
            P.pos = tid < m ? tid : kBatMaxN;
            T* const myrow = A + static_cast<long long>(S.phys[tid < m ? k + tid : k]) * n + k;
#pragma unroll
            for (int c = 0; c < NB; ++c) P.v[c] = (tid < m && c < w) ? myrow[c] : T(0);
            // rows never move: each finished group of columns goes back where the row came from
            P.run(tid < m ? myrow : nullptr);
            __syncthreads();  // the row map is final for this step
            if (k + w >= n) break;
            const int mr = m - w, nc = n - k - w, kw = k + w;
            if (tid < m) {
                // the finished row (its own stores, read back) into the U12 / update operands;
                // zeros beyond w, so the products over all 32 k never see unused columns
                T rw[NB];
#pragma unroll
                for (int c = 0; c < NB; ++c) rw[c] = c < w ? myrow[c] : T(0);
                if (P.pos < w) {
#pragma unroll
                    for (int c = 0; c < NB; ++c) S.l11[P.pos][c] = rw[c];
                } else {
#pragma unroll
                    for (int c = 0; c < NB; ++c) S.lt[c][P.pos - w] = rw[c];
                }
            }
            __syncthreads();
            // ---------------------------------------------------------------- U12
            if (tid < nc) {
                T u[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    u[i] = (i < w) ? A[static_cast<long long>(S.phys[k + i]) * n + kw + tid] : T(0);
                bat_trsm_rows<T, 1>(u, S.l11, w);
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (i < w) A[static_cast<long long>(S.phys[k + i]) * n + kw + tid] = u[i];
            }
            // ---------------------------------------------------------------- A22 -= L21 U12
            const int rtiles = (mr + 31) / 32;
            for (int cb = 0; cb < nc; cb += kBatUC) {
                __syncthreads();  // U12 written (first chunk) / the previous chunk's tiles done
                {
                    // the thread's loads all in flight before any is written to LDS
                    constexpr int kPer = NB * kBatUC / kBatThreads;
                    T uv[kPer];
#pragma unroll
                    for (int e = 0; e < kPer; ++e) {
                        const int idx = tid + kBatThreads * e;
                        const int pp = idx / kBatUC, cc = idx - pp * kBatUC;
                        uv[e] = (pp < w && cb + cc < nc) ? A[static_cast<long long>(S.phys[k + pp]) * n + kw + cb + cc]
                                                         : T(0);
                    }
#pragma unroll
                    for (int e = 0; e < kPer; ++e) {
                        const int idx = tid + kBatThreads * e;
                        S.us[idx / kBatUC][idx % kBatUC] = uv[e];
                    }
                }
                __syncthreads();
                const int ctiles = min(2, (nc - cb + 31) / 32);
                for (int t = wid; t < rtiles * ctiles; t += NW) {
                    const int rt = t / ctiles, ct = t - rt * ctiles;
                    bat_update_tile<T>(A, n, kw, S, 32 * rt, cb + 32 * ct, 32 * ct, mr, nc, lane);
                }
            }
            __syncthreads();  // the update is complete before the next panel reads its rows
        }
        __syncthreads();
        // -------------------------------------------------------------------- row order
        // write the rows in logical order, 32 columns at a time (all reads of a block before
        // any of its writes); lane groups of 32 cover the 32 columns of one row, so each wave
        // instruction moves two 32-element row pieces
        {
            constexpr int kRowsPer = kBatMaxN / (kBatThreads / NB);  // rows per thread per block
            const int cc = tid & (NB - 1), rg = tid / NB;
            for (int c0 = 0; c0 < n; c0 += NB) {
                T rv[kRowsPer];
                const bool col_ok = c0 + cc < n;
#pragma unroll
                for (int q = 0; q < kRowsPer; ++q) {
                    const int i = rg + (kBatThreads / NB) * q;
                    rv[q] = (i < n && col_ok) ? A[static_cast<long long>(S.phys[i]) * n + c0 + cc] : T(0);
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < kRowsPer; ++q) {
                    const int i = rg + (kBatThreads / NB) * q;
                    if (i < n && col_ok) A[static_cast<long long>(i) * n + c0 + cc] = rv[q];
                }
                __syncthreads();
            }
        }
        const int inf = S.info;

        // ------------------------------------------------------------ solve (nonsingular only)
        launder();
        if (inf == 0 && nrhs > 0) {
            for (int idx = tid; idx < n * nrhs; idx += kBatThreads) {
                const int r = idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                Bs[r][c] = B[static_cast<long long>(S.phys[r]) * nrhs + c];
            }
            __syncthreads();
            const int groups = (nrhs + 3) / 4;  // waves with four right-hand sides each
            const int dr = lane >> 2, dc = lane & 3;
            // forward, unit lower
            for (int blk = 0; blk < n; blk += kBatDiag) {
                const int bw = min(kBatDiag, n - blk);
                if (wid < groups) {
                    const int c = wid * 4 + dc;
                    T lrow[kBatDiag];
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp)
                        lrow[pp] = (dr < bw && pp < dr) ? A[static_cast<long long>(blk + dr) * n + blk + pp] : T(0);
                    T x = (dr < bw && c < nrhs) ? Bs[blk + dr][c] : T(0);
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp) {
                        const T xp = __shfl(x, pp * 4 + dc, kWave);
                        if (pp < dr) x = x - lrow[pp] * xp;
                    }
                    if (dr < bw && c < nrhs) Bs[blk + dr][c] = x;
                }
                __syncthreads();
                const int rest = n - blk - bw;
                for (int idx = tid; idx < rest * nrhs; idx += kBatThreads) {
                    const int i = blk + bw + idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                    const T* lr = A + static_cast<long long>(i) * n + blk;
                    T lv[kBatDiag];  // all loads issued before the products
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp) lv[pp] = pp < bw ? lr[pp] : T(0);
                    T sacc = Bs[i][c];
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp)
                        if (pp < bw) sacc = sacc - lv[pp] * Bs[blk + pp][c];
                    Bs[i][c] = sacc;
                }
                __syncthreads();
            }
            // backward, upper with division
            for (int blk = ((n - 1) / kBatDiag) * kBatDiag; blk >= 0; blk -= kBatDiag) {
                const int bw = min(kBatDiag, n - blk);
                if (wid < groups) {
                    const int c = wid * 4 + dc;
                    T urow[kBatDiag];
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp)
                        urow[pp] = (dr < bw && pp >= dr && pp < bw) ? A[static_cast<long long>(blk + dr) * n + blk + pp] : T(1);
                    T x = (dr < bw && c < nrhs) ? Bs[blk + dr][c] : T(0);
#pragma unroll
                    for (int pp = kBatDiag - 1; pp >= 0; --pp) {
                        if (pp < bw) {
                            if (dr == pp) x = x / urow[pp];
                            const T xp = __shfl(x, pp * 4 + dc, kWave);
                            if (dr < pp) x = x - urow[pp] * xp;
                        }
                    }
                    if (dr < bw && c < nrhs) Bs[blk + dr][c] = x;
                }
                __syncthreads();
                for (int idx = tid; idx < blk * nrhs; idx += kBatThreads) {
                    const int i = idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                    const T* ur = A + static_cast<long long>(i) * n + blk;
                    T uv[kBatDiag];  // all loads issued before the products
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp) uv[pp] = pp < bw ? ur[pp] : T(0);
                    T sacc = Bs[i][c];
#pragma unroll
                    for (int pp = 0; pp < kBatDiag; ++pp)
                        if (pp < bw) sacc = sacc - uv[pp] * Bs[blk + pp][c];
                    Bs[i][c] = sacc;
                }
                __syncthreads();
            }
            for (int idx = tid; idx < n * nrhs; idx += kBatThreads) {
                const int r = idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                B[static_cast<long long>(r) * nrhs + c] = Bs[r][c];
            }
        }
        if (tid == 0) info[sidx] = inf;
        __syncthreads();
    }
}

}  // namespace lrn
