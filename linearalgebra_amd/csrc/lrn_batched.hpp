// Batched LU solve of many small independent systems of uneven size (n <= 256, nrhs <= 8)
// in one persistent kernel.
//
// batch_order_kernel sorts the systems by decreasing order (counting sort over n) and
// resets the work counter, so it can sit in a captured graph without a memset node.
// batched_gesv_kernel: every workgroup repeatedly takes the next system from the sorted
// list (one atomic per system) and solves it alone:
//   factorisation  right-looking, blocked by NB columns: the NB-column panel is factored
//                  with one panel row per thread, in registers, with partial pivoting (the reference's pivot rule: first maximum of
//                  |a| wins, NaN never chosen, exactly-zero pivots skipped and reported,
//                  reciprocal multiplier unless |d| < T_MIN), its interchanges are applied
//                  to the other columns, U12 = L11^-1 A12 (row order of the reference's
//                  TRSM), and A22 -= L21 U12 with register-blocked FMAs;
//   solve          only when the factorisation is nonsingular (the reference's LU mode):
//                  B rows permuted by the pivots, blocked unit-lower forward and upper
//                  backward substitution (true division by U[i][i]) in LDS, one wave per
//                  group of right-hand sides for the 16-row diagonal blocks.
// The matrix is factored in place in global memory (it stays in L2 / the MALL while the
// workgroup works on it); panel, U strip and B live in LDS.
// Algorithm followed: src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232,
// 247-271, 329-429 and src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113,
// 160-171 (paths relative to the reference repository).
#pragma once

#include "lrn_device.hpp"
#include "lrn_getrf_kernels.hpp"

namespace lrn {

constexpr int kBatMaxN = 256;
constexpr int kBatMaxRhs = 8;
constexpr int kBatThreads = 256;
constexpr int kBatDiag = 16;  // diagonal block of the substitutions

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

template <typename T, int NB>
__global__ __launch_bounds__(kBatThreads, 3) void batched_gesv_kernel(T* __restrict__ a, const long long* __restrict__ a_off,
    long long* __restrict__ piv, const long long* __restrict__ piv_off, T* __restrict__ b,
    const long long* __restrict__ b_off, const int* __restrict__ nv, const int* __restrict__ nrhsv,
    int* __restrict__ info, const int* __restrict__ order, unsigned* __restrict__ ctr, int count)
{
    constexpr int UC = 64;
    __shared__ T P[kBatMaxN][NB + 1];
    __shared__ T Us[NB][UC + 1];
    __shared__ int perm[kBatMaxN];
    __shared__ double red_v[kBatThreads / kWave];
    __shared__ long long red_i[kBatThreads / kWave];
    __shared__ long long pivs[NB];
    __shared__ T prow[NB];
    __shared__ T jrow[NB];
    __shared__ T prcp;
    __shared__ int rowmap[1 + 4 * NB];
    __shared__ int s_sys;
    __shared__ int s_info;

    // the right-hand sides of the solve reuse the panel's LDS (the solve starts after the
    // factorisation's last barrier)
    static_assert(kBatMaxRhs + 1 <= NB + 1, "B block must fit in the panel buffer");
    T (*const Bs)[kBatMaxRhs + 1] = reinterpret_cast<T (*)[kBatMaxRhs + 1]>(&P[0][0]);

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;

    for (;;) {
        if (tid == 0) s_sys = static_cast<int>(atomicAdd(ctr, 1u));
        __syncthreads();
        const int q = s_sys;
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
        if (tid == 0) s_info = 0;
        for (int r = tid; r < n; r += kBatThreads) perm[r] = r;
        __syncthreads();

        // ------------------------------------------------------------ factorisation
        for (int k = 0; k < n; k += NB) {
            const int w = min(NB, n - k);
            const int m = n - k;
            // panel: thread t holds panel row t (m <= 256 = threads) in registers; per column
            // a DPP wave argmax, one barrier to combine the four waves, one to publish the
            // pivot row and the diagonal row, then the interchange and the rank-1 update in
            // registers
            T rowv[NB];
#pragma unroll
            for (int c = 0; c < NB; ++c)
                rowv[c] = (tid < m && c < w) ? A[static_cast<long long>(k + tid) * n + k + c] : T(0);
#pragma unroll
            for (int j = 0; j < NB; ++j) {
                if (j < w) {
                    double key = (tid >= j && tid < m) ? pivot_key(rowv[j], tid == j) : -2.0;
                    long long ri = tid;
                    wave_argmax_ordered(key, ri);
                    if (lane == 0) {
                        red_v[wid] = key;
                        red_i[wid] = ri;
                    }
                    __syncthreads();
                    double bv = red_v[0];
                    long long bi = red_i[0];
#pragma unroll
                    for (int qq = 1; qq < kBatThreads / kWave; ++qq)
                        if (argmax_better(red_v[qq], red_i[qq], bv, bi)) {
                            bv = red_v[qq];
                            bi = red_i[qq];
                        }
                    const int p = static_cast<int>(bi);
                    if (tid == p) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) prow[c] = rowv[c];
                        // the multiplier's reciprocal, computed once (|d| >= T_MIN: x * (1/d))
                        const T d0 = rowv[j];
                        prcp = (d0 != T(0) && fabs(static_cast<double>(d0)) >= static_cast<double>(TypeMin<T>::v))
                                   ? T(1) / d0
                                   : T(0);
                    }
                    if (tid == j) {
#pragma unroll
                        for (int c = 0; c < NB; ++c) jrow[c] = rowv[c];
                    }
                    if (tid == 0) {
                        pv[k + j] = static_cast<long long>(k + p + 1);
                        pivs[j] = static_cast<long long>(k + p);
                        if (p != j) {
                            const int t = perm[k + j];
                            perm[k + j] = perm[k + p];
                            perm[k + p] = t;
                        }
                    }
                    __syncthreads();
                    if (p != j) {
                        if (tid == j) {
#pragma unroll
                            for (int c = 0; c < NB; ++c) rowv[c] = prow[c];
                        } else if (tid == p) {
#pragma unroll
                            for (int c = 0; c < NB; ++c) rowv[c] = jrow[c];
                        }
                    }
                    const T d = prow[j];
                    if (d == T(0)) {
                        if (tid == 0 && s_info == 0) s_info = k + j + 1;
                    } else if (tid > j && tid < m) {
                        const T rcp = prcp;
                        const T l = (rcp != T(0)) ? rowv[j] * rcp : rowv[j] / d;
                        rowv[j] = l;
#pragma unroll
                        for (int c = j + 1; c < NB; ++c)
                            if (c < w) rowv[c] = rowv[c] - l * prow[c];
                    }
                }
            }
            if (tid < m) {
#pragma unroll
                for (int c = 0; c < NB; ++c)
                    if (c < w) {
                        P[tid][c] = rowv[c];
                        A[static_cast<long long>(k + tid) * n + k + c] = rowv[c];
                    }
            }
            __syncthreads();
            if (wid == 0) panel_rowmap(pivs, k, w, 2 * NB, rowmap);
            __syncthreads();
            // the panel's interchanges on every other column (one column per thread): the
            // net row moves (rowmap: row dst receives row src) in one gather and one scatter
            {
                const int cnt = rowmap[0];
                const int* dst = rowmap + 1;
                const int* src = rowmap + 1 + 2 * NB;
                const int c = tid;  // n <= 256 = threads
                if (c < n && !(c >= k && c < k + w)) {
                    // all reads before any write (the moves form a permutation of the rows)
                    T v[2 * NB];
#pragma unroll
                    for (int q2 = 0; q2 < 2 * NB; ++q2)
                        v[q2] = (q2 < cnt) ? A[static_cast<long long>(src[q2]) * n + c] : T(0);
#pragma unroll
                    for (int q2 = 0; q2 < 2 * NB; ++q2)
                        if (q2 < cnt) A[static_cast<long long>(dst[q2]) * n + c] = v[q2];
                }
            }
            __syncthreads();
            if (k + w >= n) break;
            // U12 = L11^-1 A12: one column per thread (n <= 256 = threads), the column in
            // registers, rows in the reference's TRSM order
            const int mr = m - w, nc = n - k - w;
            if (tid < nc) {
                T* const col = A + static_cast<long long>(k) * n + k + w + tid;
                T u[NB];
#pragma unroll
                for (int i = 0; i < NB; ++i) u[i] = (i < w) ? col[static_cast<long long>(i) * n] : T(0);
#pragma unroll
                for (int i = 1; i < NB; ++i) {
                    if (i < w) {
#pragma unroll
                        for (int p = 0; p < i; ++p) u[i] = u[i] - P[i][p] * u[p];
                    }
                    // one row of L11 in flight at a time: the scheduler would otherwise issue
                    // all NB^2/2 LDS reads up front and run out of registers
                    asm volatile("" ::: "memory");
                }
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (i < w) col[static_cast<long long>(i) * n] = u[i];
            }
            __syncthreads();
            // A22 -= L21 U12: 64 x 64 per pass, 4 x 4 per thread
            const int ty = tid >> 4, tx = tid & 15;
            for (int cb = 0; cb < nc; cb += UC) {
                for (int idx = tid; idx < NB * UC; idx += kBatThreads) {
                    const int p = idx / UC, cc = idx - (idx / UC) * UC;
                    Us[p][cc] = (p < w && cb + cc < nc) ? A[static_cast<long long>(k + p) * n + k + w + cb + cc] : T(0);
                }
                __syncthreads();
                for (int rb = 0; rb < mr; rb += 64) {
                    T acc[4][4], cv[4][4];
                    const int r0 = rb + ty * 4;
                    // C first: its global latency overlaps the products below
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const T* crow = A + static_cast<long long>(k + w + r0 + i) * n + k + w + cb + tx * 4;
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) {
                            acc[i][jj] = T(0);
                            cv[i][jj] = (r0 + i < mr && cb + tx * 4 + jj < nc) ? crow[jj] : T(0);
                        }
                    }
                    // partial unroll: a full one hoists every p's LDS operands and spills
#pragma unroll 2
                    for (int p = 0; p < NB; ++p) {
                        T av[4], bw4[4];
#pragma unroll
                        for (int i = 0; i < 4; ++i) av[i] = (p < w && r0 + i < mr) ? P[w + r0 + i][p] : T(0);
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj) bw4[jj] = Us[p][tx * 4 + jj];
#pragma unroll
                        for (int i = 0; i < 4; ++i)
#pragma unroll
                            for (int jj = 0; jj < 4; ++jj) acc[i][jj] = fma(av[i], bw4[jj], acc[i][jj]);
                    }
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (r0 + i >= mr) continue;
                        T* crow = A + static_cast<long long>(k + w + r0 + i) * n + k + w + cb + tx * 4;
#pragma unroll
                        for (int jj = 0; jj < 4; ++jj)
                            if (cb + tx * 4 + jj < nc) crow[jj] = cv[i][jj] - acc[i][jj];
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
        const int inf = s_info;

        // ------------------------------------------------------------ solve (nonsingular only)
        if (inf == 0 && nrhs > 0) {
            for (int idx = tid; idx < n * nrhs; idx += kBatThreads) {
                const int r = idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                Bs[r][c] = B[static_cast<long long>(perm[r]) * nrhs + c];
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
                    for (int p = 0; p < kBatDiag; ++p)
                        lrow[p] = (dr < bw && p < dr) ? A[static_cast<long long>(blk + dr) * n + blk + p] : T(0);
                    T x = (dr < bw && c < nrhs) ? Bs[blk + dr][c] : T(0);
#pragma unroll
                    for (int p = 0; p < kBatDiag; ++p) {
                        const T xp = __shfl(x, p * 4 + dc, kWave);
                        if (p < dr) x = x - lrow[p] * xp;
                    }
                    if (dr < bw && c < nrhs) Bs[blk + dr][c] = x;
                }
                __syncthreads();
                const int rest = n - blk - bw;
                for (int idx = tid; idx < rest * nrhs; idx += kBatThreads) {
                    const int i = blk + bw + idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                    const T* lr = A + static_cast<long long>(i) * n + blk;
                    T s = Bs[i][c];
                    for (int p = 0; p < bw; ++p) s = s - lr[p] * Bs[blk + p][c];
                    Bs[i][c] = s;
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
                    for (int p = 0; p < kBatDiag; ++p)
                        urow[p] = (dr < bw && p >= dr && p < bw) ? A[static_cast<long long>(blk + dr) * n + blk + p] : T(1);
                    T x = (dr < bw && c < nrhs) ? Bs[blk + dr][c] : T(0);
#pragma unroll
                    for (int p = kBatDiag - 1; p >= 0; --p) {
                        if (p < bw) {
                            if (dr == p) x = x / urow[p];
                            const T xp = __shfl(x, p * 4 + dc, kWave);
                            if (dr < p) x = x - urow[p] * xp;
                        }
                    }
                    if (dr < bw && c < nrhs) Bs[blk + dr][c] = x;
                }
                __syncthreads();
                for (int idx = tid; idx < blk * nrhs; idx += kBatThreads) {
                    const int i = idx / nrhs, c = idx - (idx / nrhs) * nrhs;
                    const T* ur = A + static_cast<long long>(i) * n + blk;
                    T s = Bs[i][c];
                    for (int p = 0; p < bw; ++p) s = s - ur[p] * Bs[blk + p][c];
                    Bs[i][c] = s;
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
