// Kernels of the right-looking blocked LU factorisation (row-major, partial pivoting).
//
//   getf2_panel      cooperative panel factorisation of an m x w column block: pivot search by
//                    wave-shuffle argmax, row swaps and rank-1 updates on an LDS-resident tile,
//                    one bounded inter-workgroup barrier per column
//   laswp_trsm       one panel's row interchanges on the rest of its block column, and its
//                    U12 rows
//   trsm_llnu        U12 = L11^-1 A12 for a unit-lower w x w diagonal block (w <= 64)
//   gemm_sub         C -= A * B on the matrix cores (FP64 / FP32 MFMA 16x16x4)
//
// Algorithm followed: src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232
// (panel), :247-271 (deferred interchanges), :364-380 (triangular solve), :382-422 (trailing
// update as delta = L21*U12 followed by A22 -= delta).
#pragma once

#include <float.h>

#include "lrn_device.hpp"

namespace lrn {

constexpr int kPanelThreads = 256;
constexpr int kPanelMaxWG = 256;  // upper bound on workgroups of one panel launch

template <typename T> struct PanelWs {
    unsigned* ctr;     // barrier counter (monotonic over one factorisation)
    double* sval;      // [2][kPanelMaxWG] published local maxima
    long long* sidx;   // [2][kPanelMaxWG] their row indices
    T* srow;           // [2][kPanelMaxWG][W] the candidate rows
    T* sdiag;          // [2][W] the current diagonal row
    int* info;         // first exactly-zero pivot, 1-based; 0 = none
    int* err;          // raised when a bounded wait times out
    int* rowmap;       // [1 + 4W]: count, dst[2W], src[2W] of the panel's net row moves
    void* chunks;      // chunk slots of the one-round-trip panel kernel (lrn_panel.hpp)
    void* dchunks;
};


// Net row movement of one panel's interchanges. The swaps (i, piv[i]-1), i = k..k+w-1, are
// replayed on position labels by one wave: lane l tracks position k+l, and positions at or
// beyond k+w are tracked as "extras" in the lanes in order of first use. Output: pairs
// (dst, src) meaning "row dst of the result is row src of the input".
__device__ inline void panel_rowmap(const long long* pivs, long long k, int w, int W2, int* rowmap)
{
    const int lane = threadIdx.x;  // called by wave 0 only
    long long c = k + lane;        // content label at position k+lane
    long long e = -1, ec = -1;     // extra position handled by this lane, and its content
    int ne = 0;
    for (int ii = 0; ii < w; ++ii) {
        const long long p = pivs[ii];
        const long long ci = __shfl(c, ii, kWave);
        long long cp;
        if (p < k + w) {
            const int lp = static_cast<int>(p - k);
            cp = __shfl(c, lp, kWave);
            if (lane == lp) c = ci;
        } else {
            const unsigned long long mask = __ballot(lane < ne && e == p);
            if (mask) {
                const int l2 = __ffsll(static_cast<long long>(mask)) - 1;
                cp = __shfl(ec, l2, kWave);
                if (lane == l2) ec = ci;
            } else {
                cp = p;
                if (lane == ne) {
                    e = p;
                    ec = ci;
                }
                ++ne;
            }
        }
        if (lane == ii) c = cp;
    }
    const bool moved = lane < w && c != k + lane;
    const unsigned long long m1 = __ballot(moved);
    const unsigned long long m2 = __ballot(lane < ne);
    const unsigned long long lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int n1 = __popcll(m1);
    if (moved) {
        const int q = __popcll(m1 & lower);
        rowmap[1 + q] = static_cast<int>(k + lane);
        rowmap[1 + W2 + q] = static_cast<int>(c);
    }
    if (lane < ne) {
        const int q = n1 + __popcll(m2 & lower);
        rowmap[1 + q] = static_cast<int>(e);
        rowmap[1 + W2 + q] = static_cast<int>(ec);
    }
    if (lane == 0) rowmap[0] = n1 + __popcll(m2);
}

// Cooperative panel factorisation of A[k:n, k:k+w] (row-major, stride lda).
// Workgroup b owns rows k + b*R ... k + b*R + R - 1, holds them in LDS for the whole launch,
// and all P workgroups agree on each column's pivot through one bounded barrier per column.
// Requires R >= w (so workgroup 0 owns every diagonal row of the panel) and every one of the
// P workgroups resident at once (the host sizes P from the occupancy of this kernel).
template <typename T, int W>
__global__ __launch_bounds__(kPanelThreads) void getf2_panel_kernel(T* __restrict__ A, long long lda,
    long long k, long long m, int w, int R, int P, unsigned base, unsigned spin_limit,
    long long* __restrict__ piv, PanelWs<T> ws)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int LD = W + 1;  // odd stride in 8-byte words: column reads hit distinct banks
    T* tile = reinterpret_cast<T*>(lds_raw);
    __shared__ T prow[W];
    __shared__ T drow[W];
    __shared__ double red_v[kPanelThreads / kWave];
    __shared__ long long red_i[kPanelThreads / kWave];
    __shared__ long long win_idx;
    __shared__ long long pivs[W];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int b = blockIdx.x;
    const long long row0 = k + static_cast<long long>(b) * R;  // first global row of this WG
    const long long n = k + m;
    const int rows_here = static_cast<int>(min(static_cast<long long>(R), n - row0));

    // Load the workgroup's rows of the panel.
    for (int idx = tid; idx < rows_here * w; idx += kPanelThreads) {
        const int r = idx / w;
        const int c = idx - r * w;
        tile[r * LD + c] = A[(row0 + r) * lda + k + c];
    }
    __syncthreads();

    // threads per row in the elimination step
    const int tpr = (R >= kPanelThreads) ? 1 : (kPanelThreads / R);
    const int rstep = kPanelThreads / tpr;

    for (int j = 0; j < w; ++j) {
        const long long i = k + j;  // global index of the diagonal row
        const int par = j & 1;

        // 1. local argmax over this workgroup's rows at or below the diagonal
        double bv = -2.0;
        long long bi = 0x7fffffffffffffffll;
        for (int r = tid; r < rows_here; r += kPanelThreads) {
            const long long g = row0 + r;
            if (g >= i) {
                const double v = pivot_key(tile[r * LD + j], g == i);
                if (argmax_better(v, g, bv, bi)) {
                    bv = v;
                    bi = g;
                }
            }
        }
        wave_argmax(bv, bi);
        if (lane == 0) {
            red_v[wid] = bv;
            red_i[wid] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            for (int q = 1; q < kPanelThreads / kWave; ++q)
                if (argmax_better(red_v[q], red_i[q], bv, bi)) {
                    bv = red_v[q];
                    bi = red_i[q];
                }
            red_v[0] = bv;
            red_i[0] = bi;
        }
        __syncthreads();
        bv = red_v[0];
        bi = red_i[0];

        long long p;  // global pivot row
        if (P > 1) {
            // 2. publish the local candidate (and, from workgroup 0, the diagonal row)
            if (tid == 0) {
                st_agent(&ws.sval[par * kPanelMaxWG + b], bv);
                st_agent(&ws.sidx[par * kPanelMaxWG + b], bi);
            }
            if (tid < w && bv >= -1.0) {
                const int rl = static_cast<int>(bi - row0);
                st_agent(&ws.srow[(static_cast<size_t>(par) * kPanelMaxWG + b) * W + tid], tile[rl * LD + tid]);
            }
            if (b == 0 && tid < w) st_agent(&ws.sdiag[par * W + tid], tile[j * LD + tid]);
            __threadfence();
            __syncthreads();
            // 3. one barrier per column
            if (tid == 0) {
                signal_counter(ws.ctr);
                wait_counter(ws.ctr, base + static_cast<unsigned>(j + 1) * static_cast<unsigned>(P), spin_limit, ws.err);
            }
            __syncthreads();
            // 4. global winner
            double gv = -2.0;
            long long gi = 0x7fffffffffffffffll;
            if (tid < P) {
                gv = ld_agent(&ws.sval[par * kPanelMaxWG + tid]);
                gi = ld_agent(&ws.sidx[par * kPanelMaxWG + tid]);
            }
            wave_argmax(gv, gi);
            if (lane == 0) {
                red_v[wid] = gv;
                red_i[wid] = gi;
            }
            __syncthreads();
            if (tid == 0) {
                for (int q = 1; q < kPanelThreads / kWave; ++q)
                    if (argmax_better(red_v[q], red_i[q], gv, gi)) {
                        gv = red_v[q];
                        gi = red_i[q];
                    }
                win_idx = gi;
            }
            __syncthreads();
            p = win_idx;
            if (p < i || p >= k + m) {
                // only an expired wait can get here: report it and keep the rows in place
                if (tid == 0) raise_error(ws.err);
                p = i;
            }
            const int wp = static_cast<int>((p - k) / R);
            if (tid < w) {
                prow[tid] = (wp == b) ? tile[static_cast<int>(p - row0) * LD + tid]
                                      : ld_agent(&ws.srow[(static_cast<size_t>(par) * kPanelMaxWG + wp) * W + tid]);
                drow[tid] = (b == 0) ? tile[j * LD + tid] : ld_agent(&ws.sdiag[par * W + tid]);
            }
        } else {
            p = bi;
            if (tid < w) {
                prow[tid] = tile[static_cast<int>(p - row0) * LD + tid];
                drow[tid] = tile[j * LD + tid];
            }
        }
        __syncthreads();

        // 5. interchange rows i and p inside the panel columns
        if (p != i) {
            if (b == 0 && tid < w) tile[j * LD + tid] = prow[tid];
            const long long pl = p - row0;
            if (pl >= 0 && pl < rows_here && tid < w) tile[static_cast<int>(pl) * LD + tid] = drow[tid];
        }
        if (b == 0 && tid == 0) {
            piv[i] = p + 1;
            pivs[j] = p;
        }
        __syncthreads();

        // 6. multipliers and rank-1 update of the rows below the diagonal
        const T d = prow[j];
        if (d == T(0)) {
            if (b == 0 && tid == 0 && *ws.info == 0) *ws.info = static_cast<int>(i + 1);
        } else {
            const bool use_recip = fabs(static_cast<double>(d)) >= static_cast<double>(TypeMin<T>::v);
            const T rcp = use_recip ? T(1) / d : T(1);
            for (int r = tid / tpr; r < rows_here; r += rstep) {
                if (row0 + r <= i) continue;
                T* rowp = tile + r * LD;
                const T l = use_recip ? rowp[j] * rcp : rowp[j] / d;
                for (int c = j + 1 + (tid % tpr); c < w; c += tpr) rowp[c] = rowp[c] - l * prow[c];
                if ((tid % tpr) == 0) rowp[j] = l;
            }
        }
        __syncthreads();
    }

    // Write the factored rows back.
    for (int idx = tid; idx < rows_here * w; idx += kPanelThreads) {
        const int r = idx / w;
        const int c = idx - r * w;
        A[(row0 + r) * lda + k + c] = tile[r * LD + c];
    }
    if (b == 0 && tid < kWave) panel_rowmap(pivs, k, w, 2 * W, ws.rowmap);
}

// Row moves are applied in strips of kSwapCols columns per workgroup: gather every source
// row into LDS, then scatter to the destination rows, so a row that is both read and written
// is never raced.
constexpr int kSwapCols = 64;
// After an inner panel (rows k.., columns k..k+w-1) of block column [K, KB): its row moves on
// the block column's other columns and U12 = L11^-1 A12 of rows k..k+w-1, in one launch
// (a row-move kernel followed by trsm_llnu_kernel would be two latency-bound launches on
// little data).
// Workgroups [0, na) take 64-column strips of [K, k) (moves only), workgroups [na, ..) strips
// of [rs, KB) (usually rs = k + w; the next block column's columns when the factorisation
// updates it inner panel by inner panel). All loads of a workgroup are issued together: the moved rows, and for the
// right strips L11 and the strip's rows k..k+w-1 as they were; A12 after the moves is then
// assembled in LDS (a moved row's values from the gathered rows, the others as loaded), so
// the kernel makes one memory round trip before its stores. One thread per column then solves
// its column in registers in increasing pivot order with every product rounded before its
// subtraction -- bitwise the result of trsm_llnu_kernel.
template <typename T, int W>
__global__ __launch_bounds__(256) void laswp_trsm_kernel(T* __restrict__ A, long long lda,
    const int* __restrict__ rowmap, int W2, long long K, long long k, int w, long long rs, long long KB, int na)
{
    __shared__ T buf[2 * W][kSwapCols];
    __shared__ T L11[W][W + 1];
    __shared__ T X[W][kSwapCols + 1];
    const int tid = threadIdx.x;
    const bool right = static_cast<int>(blockIdx.x) >= na;
    const long long cb = right ? rs + static_cast<long long>(blockIdx.x - na) * kSwapCols
                               : K + static_cast<long long>(blockIdx.x) * kSwapCols;
    const long long ce = right ? KB : k;
    const int ncols = static_cast<int>(min(static_cast<long long>(kSwapCols), ce - cb));
    const int cnt = rowmap[0];
    const int* dst = rowmap + 1;
    const int* src = rowmap + 1 + W2;
    // every load is issued before the first of them is used: fully unrolled loops over the
    // thread's share, through registers (a rolled loop waits for each load in turn)
    constexpr int kMoves = 2 * W * kSwapCols / 256;
    constexpr int kL = (W * W + 255) / 256, kX = (W * kSwapCols + 255) / 256;
    T v[kMoves], lv[kL], xv[kX];
#pragma unroll
    for (int t = 0; t < kMoves; ++t) {
        const int idx = tid + 256 * t;
        const int q = idx / kSwapCols, c = idx - q * kSwapCols;
        v[t] = (q < cnt && c < ncols) ? A[static_cast<long long>(src[q]) * lda + cb + c] : T(0);
    }
    if (right) {
#pragma unroll
        for (int t = 0; t < kL; ++t) {
            const int idx = tid + 256 * t;
            const int r = idx / W, c = idx % W;
            lv[t] = (idx < W * W && r < w && c < w) ? A[(k + r) * lda + k + c] : T(0);
        }
#pragma unroll
        for (int t = 0; t < kX; ++t) {
            const int idx = tid + 256 * t;
            const int r = idx / kSwapCols, c = idx % kSwapCols;
            xv[t] = (idx < W * kSwapCols && r < w && c < ncols) ? A[(k + r) * lda + cb + c] : T(0);
        }
    }
#pragma unroll
    for (int t = 0; t < kMoves; ++t) {
        const int idx = tid + 256 * t;
        buf[idx / kSwapCols][idx % kSwapCols] = v[t];
    }
    if (right) {
#pragma unroll
        for (int t = 0; t < kL; ++t) {
            const int idx = tid + 256 * t;
            if (idx < W * W) L11[idx / W][idx % W] = lv[t];
        }
#pragma unroll
        for (int t = 0; t < kX; ++t) {
            const int idx = tid + 256 * t;
            if (idx < W * kSwapCols) X[idx / kSwapCols][idx % kSwapCols] = xv[t];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kMoves; ++t) {
        const int idx = tid + 256 * t;
        const int q = idx / kSwapCols, c = idx - q * kSwapCols;
        if (q < cnt && c < ncols) A[static_cast<long long>(dst[q]) * lda + cb + c] = buf[q][c];
    }
    if (!right) return;  // uniform over the workgroup
    // rows k..k+w-1 that receive a moved row take it from the gathered rows (each row is the
    // destination of at most one move)
#pragma unroll
    for (int t = 0; t < kMoves; ++t) {
        const int idx = tid + 256 * t;
        const int q = idx / kSwapCols, c = idx - q * kSwapCols;
        if (q < cnt && c < ncols) {
            const long long d = static_cast<long long>(dst[q]) - k;
            if (d >= 0 && d < w) X[d][c] = buf[q][c];
        }
    }
    __syncthreads();
    if (tid < ncols) {
        // the column in registers (every index a compile-time constant); the L11 entries are
        // broadcast reads off the chain of products and subtractions
        const int c = tid;
        T u[W];
#pragma unroll
        for (int i = 0; i < W; ++i) u[i] = (i < w) ? X[i][c] : T(0);
#pragma unroll
        for (int i = 1; i < W; ++i) {
            if (i < w) {
                T sum = u[i];
#pragma unroll
                for (int p = 0; p < i; ++p) sum = sum - L11[i][p] * u[p];
                u[i] = sum;
            }
        }
#pragma unroll
        for (int i = 1; i < W; ++i)
            if (i < w) A[(k + i) * lda + cb + c] = u[i];
    }
}

// Applies nmaps consecutive row maps (each the net moves of one inner panel, stride
// map_stride ints) in order to the columns [c0a, c1a) and [c0b, c1b): one launch for a whole
// block column's interchanges. Each workgroup moves a strip of kSwapCols columns; the maps
// are applied one after the other with a block-wide fence between them.
template <typename T>
__global__ __launch_bounds__(256) void laswp_multi_kernel(T* __restrict__ A, long long lda,
    const int* __restrict__ rowmaps, int nmaps, int map_stride, int W2, long long c0a, long long c1a, long long c0b,
    long long c1b)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* buf = reinterpret_cast<T*>(lds_raw);  // [W2][kSwapCols]
    const long long la = c1a - c0a;
    const long long total = la + (c1b - c0b);
    const long long s = static_cast<long long>(blockIdx.x) * kSwapCols;
    for (int mi = 0; mi < nmaps; ++mi) {
        const int* rm = rowmaps + static_cast<size_t>(mi) * map_stride;
        const int cnt = rm[0];
        const int* dst = rm + 1;
        const int* src = rm + 1 + W2;
        for (int idx = threadIdx.x; idx < cnt * kSwapCols; idx += blockDim.x) {
            const int q = idx / kSwapCols;
            const int cc = idx - q * kSwapCols;
            const long long t = s + cc;
            if (t < total) {
                const long long col = (t < la) ? c0a + t : c0b + (t - la);
                buf[idx] = A[static_cast<long long>(src[q]) * lda + col];
            }
        }
        __syncthreads();
        for (int idx = threadIdx.x; idx < cnt * kSwapCols; idx += blockDim.x) {
            const int q = idx / kSwapCols;
            const int cc = idx - q * kSwapCols;
            const long long t = s + cc;
            if (t < total) {
                const long long col = (t < la) ? c0a + t : c0b + (t - la);
                A[static_cast<long long>(dst[q]) * lda + col] = buf[idx];
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// U12 = L11^-1 * A12: rows k..k+w-1, columns [c0, c1). One thread per column; L11 (unit
// lower, from the factored diagonal block) is broadcast from LDS. The rows are processed in
// chunks of 16 held in registers; earlier chunks are re-read from memory (just written, so
// cache hits). Each entry is reduced in increasing p with every product rounded before its
// subtraction, the order of the reference's row-wise forward substitution.
template <typename T, int W>
__global__ __launch_bounds__(256) void trsm_llnu_kernel(T* __restrict__ A, long long lda, long long k, int w,
    long long c0, long long c1)
{
    constexpr int CH = 16;
    __shared__ T L[W][W + 1];
    for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
        const int r = idx / w;
        const int c = idx - r * w;
        L[r][c] = A[(k + r) * lda + k + c];
    }
    __syncthreads();
    const long long col = c0 + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= c1) return;
    T* colp = A + k * lda + col;
    for (int c = 0; c * CH < w; ++c) {
        T x[CH];
#pragma unroll
        for (int q = 0; q < CH; ++q) x[q] = (c * CH + q < w) ? colp[(c * CH + q) * lda] : T(0);
        for (int e = 0; e < c; ++e) {
            T y[CH];
#pragma unroll
            for (int q = 0; q < CH; ++q) y[q] = colp[(e * CH + q) * lda];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
                T s = x[q];
                const T* lrow = &L[(c * CH + q) < W ? (c * CH + q) : 0][e * CH];
#pragma unroll
                for (int p = 0; p < CH; ++p) s = s - lrow[p] * y[p];
                x[q] = s;
            }
        }
#pragma unroll
        for (int q = 1; q < CH; ++q) {
            T s = x[q];
            const T* lrow = &L[(c * CH + q) < W ? (c * CH + q) : 0][c * CH];
#pragma unroll
            for (int p = 0; p < q; ++p) s = s - lrow[p] * x[p];
            x[q] = s;
        }
#pragma unroll
        for (int q = 0; q < CH; ++q)
            if (c * CH + q < w) colp[(c * CH + q) * lda] = x[q];
    }
}

// Update of the rest of a block column after one of its inner panels (rows k..k+w-1 and the
// columns below them factored, U12 = L11^-1 A12 already solved in place by trsm_llnu_kernel):
//   A22 -= L21 * U12   (U12: rows k..k+w-1, columns [c0, c1); A22: rows k+w..n-1).
// The skinny GEMM of a block column (K = w, N <= the block width) as many small workgroups:
// workgroup (x, y) owns the 64-row block x and the 64-column tile y of A22, stages its L21
// rows and U12 tile in LDS and updates its block with 4 x 4 register tiles, each entry's
// products accumulated from zero and subtracted once. It keeps its speed on the few compute
// units of a look-ahead panel stream, where 128 x 128 matrix-core tiles leave most lanes
// idle. (The same update on the matrix cores with 64 x 64 tiles is gemm_sub_kernel<T, 64,
// 64, 32>; the host picks one.)
constexpr int kBcRows = 64, kBcCols = 64;
template <typename T, int W>
__global__ __launch_bounds__(256) void bc_update_kernel(T* __restrict__ A, long long lda, long long n, long long k,
    int w, long long c0, long long c1)
{
    __shared__ T U[W][kBcCols + 1];
    __shared__ T Lb[kBcRows][W + 1];
    const int tid = threadIdx.x;
    const long long cb = c0 + static_cast<long long>(blockIdx.y) * kBcCols;
    const int ncols = static_cast<int>(min(static_cast<long long>(kBcCols), c1 - cb));
    const long long rb = k + w + static_cast<long long>(blockIdx.x) * kBcRows;
    const int nrows = static_cast<int>(max(0ll, min(static_cast<long long>(kBcRows), n - rb)));
    for (int idx = tid; idx < kBcRows * W; idx += 256) {
        const int r = idx / W, c = idx % W;
        Lb[r][c] = (r < nrows && c < w) ? A[(rb + r) * lda + k + c] : T(0);
    }
    for (int idx = tid; idx < W * kBcCols; idx += 256) {
        const int r = idx / kBcCols, c = idx % kBcCols;
        U[r][c] = (r < w && c < ncols) ? A[(k + r) * lda + cb + c] : T(0);
    }
    __syncthreads();
    // 16 x 16 threads, 4 x 4 entries each: rows 4*ty.., columns 4*tx..
    const int tx = tid & 15, ty = tid >> 4;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
#pragma unroll 4
    for (int p = 0; p < w; ++p) {
        T l[4], u[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) l[i] = Lb[4 * ty + i][p];
#pragma unroll
        for (int j = 0; j < 4; ++j) u[j] = U[p][4 * tx + j];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = fma(l[i], u[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = 4 * ty + i;
        if (r >= nrows) continue;
        T* crow = A + (rb + r) * lda + cb;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (4 * tx + j < ncols) crow[4 * tx + j] = crow[4 * tx + j] - acc[i][j];
    }
}

// All row maps of a block column (nmaps maps, each the net moves of one inner panel, at most
// W2 moves each, applied in order) on the columns [c0a, c1a) and [c0b, c1b) in one pass:
// each workgroup takes a strip of kSwapBlkCols columns, gives every distinct row named by the
// maps a slot (LDS hash table keyed by row), gathers those rows once, applies the maps one
// after the other inside LDS, and scatters the rows back once: two memory round trips instead
// of two per map. The result equals applying the maps in order (laswp_multi_kernel).
constexpr int kSwapBlkCols = 16;
constexpr int kSwapBlkRows = 512;  // distinct rows a block column's maps can name (2 * 256)
template <typename T>
__global__ __launch_bounds__(256) void laswp_block_kernel(T* __restrict__ A, long long lda,
    const int* __restrict__ rowmaps, int nmaps, int map_stride, int W2, long long c0a, long long c1a, long long c0b,
    long long c1b)
{
    constexpr int SC = kSwapBlkCols, HS = 2048, MAXM = 16;
    __shared__ T buf[kSwapBlkRows][SC + 1];
    __shared__ int hkey[HS];
    __shared__ int hslot[HS];
    __shared__ int rowof[kSwapBlkRows];
    __shared__ short sdst[MAXM][64];
    __shared__ short ssrc[MAXM][64];
    __shared__ int cnts[MAXM];
    __shared__ int nslots;
    const int tid = threadIdx.x;
    const long long la = c1a - c0a;
    const long long total = la + (c1b - c0b);
    const long long s0 = static_cast<long long>(blockIdx.x) * SC;
    for (int i = tid; i < HS; i += 256) hkey[i] = -1;
    if (tid == 0) nslots = 0;
    if (tid < nmaps) cnts[tid] = rowmaps[static_cast<size_t>(tid) * map_stride];
    __syncthreads();
    auto slot_of = [&](int row, bool insert) -> int {
        unsigned h = (static_cast<unsigned>(row) * 2654435761u) >> 21;  // 11 bits
        for (;;) {
            const int k = hkey[h];
            if (k == row) return h;
            if (k == -1) {
                if (!insert) return -1;
                const int old = atomicCAS(&hkey[h], -1, row);
                if (old == -1 || old == row) return h;
            }
            h = (h + 1) & (HS - 1);
        }
    };
    // every row a map names: dst and src sets of one map are the same set
    for (int idx = tid; idx < nmaps * 64; idx += 256) {
        const int m = idx >> 6, q = idx & 63;
        if (q < cnts[m] && q < W2) {
            const int row = rowmaps[static_cast<size_t>(m) * map_stride + 1 + q];
            const int h = slot_of(row, true);
            (void)h;
        }
    }
    __syncthreads();
    for (int h = tid; h < HS; h += 256)
        if (hkey[h] != -1) {
            const int sl = atomicAdd(&nslots, 1);
            hslot[h] = sl;
            rowof[sl] = hkey[h];
        }
    __syncthreads();
    for (int idx = tid; idx < nmaps * 64; idx += 256) {
        const int m = idx >> 6, q = idx & 63;
        if (q < cnts[m] && q < W2) {
            const int* rm = rowmaps + static_cast<size_t>(m) * map_stride;
            sdst[m][q] = static_cast<short>(hslot[slot_of(rm[1 + q], false)]);
            ssrc[m][q] = static_cast<short>(hslot[slot_of(rm[1 + W2 + q], false)]);
        }
    }
    const int ns = nslots;
    {
        // all of the thread's gathers in flight at once (unrolled, through registers)
        constexpr int kG = kSwapBlkRows * SC / 256;
        T v[kG];
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            const int idx = tid + 256 * g;
            const int sl = idx / SC, c = idx - sl * SC;
            const long long t = s0 + c;
            v[g] = T(0);
            if (sl < ns && t < total) {
                const long long col = (t < la) ? c0a + t : c0b + (t - la);
                v[g] = A[static_cast<long long>(rowof[sl]) * lda + col];
            }
        }
#pragma unroll
        for (int g = 0; g < kG; ++g) {
            const int idx = tid + 256 * g;
            const int sl = idx / SC, c = idx - sl * SC;
            if (sl < ns) buf[sl][c] = v[g];
        }
    }
    __syncthreads();
    for (int m = 0; m < nmaps; ++m) {
        const int cnt = cnts[m];
        T v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = tid + 256 * k;
            const int q = idx / SC, c = idx - q * SC;
            if (q < cnt) v[k] = buf[ssrc[m][q]][c];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int idx = tid + 256 * k;
            const int q = idx / SC, c = idx - q * SC;
            if (q < cnt) buf[sdst[m][q]][c] = v[k];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < ns * SC; idx += 256) {
        const int sl = idx / SC, c = idx - sl * SC;
        const long long t = s0 + c;
        if (t < total) {
            const long long col = (t < la) ? c0a + t : c0b + (t - la);
            A[static_cast<long long>(rowof[sl]) * lda + col] = buf[sl][c];
        }
    }
}

// U12 = L11^-1 A12 for a whole block column in one launch: rows K..K+B-1 (B <= 256), columns
// [c0, c1), L11 the unit-lower B x B block at (K, K). Each workgroup owns 64 columns and keeps
// the finished U rows of its columns in LDS. The rows are done in 32-row groups: first every
// row of the group takes the contributions of all earlier groups (one row and 8 columns per
// thread), then one thread per column finishes the group by forward substitution against
// the 32 x 32 diagonal block. Every entry receives its updates in increasing pivot order with
// each product rounded before the subtraction, which is exactly the row-wise forward
// substitution of src/matrix_operation/cpu/native/lu_decomposition_native.cpp:364-380, so the
// result does not depend on the grouping, nor on TN (columns per workgroup). Wide column
// ranges: 32 (FP32) or 16 (FP64), the most columns that still leave room for several
// workgroups per compute unit (LDS: 256 x (TN + 1) entries of U); on 224 compute units, the
// far columns [768, 16384) of n = 16384 take 114 us with TN = 32 against 207 us with 64 in
// FP32, and 283 us with TN = 16 against 524 us with 64 in FP64. One block column's 256
// columns alone: 8, for enough workgroups.
template <typename T> constexpr int u12_tn_wide() { return sizeof(T) == 4 ? 32 : 16; }
constexpr int kU12TNNarrow = 8;
template <typename T, int TN>
__global__ __launch_bounds__(256) void u12_block_kernel(T* __restrict__ A, long long lda, long long K, int B,
    long long c0, long long c1)
{
    static_assert(TN % 8 == 0 && TN <= 64, "8 threads per row, TN / 8 columns each");
    constexpr int CPT = TN / 8;
    __shared__ T U[256][TN + 1];
    __shared__ T Ld[32][33];
    __shared__ T Lc[2][32][33];  // 32 x 32 pieces of the rows' L, double-buffered
    const int tid = threadIdx.x;
    const long long cb = c0 + static_cast<long long>(blockIdx.x) * TN;
    const int ncols = static_cast<int>(min(static_cast<long long>(TN), c1 - cb));
    const int nin = (B + 31) / 32;
    const int r = tid >> 3, cg = (tid & 7) * CPT;
    const int lr = tid >> 3, lc = (tid & 7) * 4;  // staging: 4 consecutive L entries per thread
    for (int p = 0; p < nin; ++p) {
        const long long r0 = K + 32 * p;
        const int rp = min(32, B - 32 * p);
        T x[CPT];
        if (r < rp) {
            const T* arow = A + (r0 + r) * lda + cb + cg;
#pragma unroll
            for (int e = 0; e < CPT; ++e) x[e] = (cg + e < ncols) ? arow[e] : T(0);
        }
        // contributions of the earlier groups, 32 pivots at a time
        T st[4];
        auto fetch = [&](int qc) {
            const T* src = A + (r0 + lr) * lda + K + 32 * qc + lc;
#pragma unroll
            for (int e = 0; e < 4; ++e) st[e] = (lr < rp) ? src[e] : T(0);
        };
        if (p > 0) fetch(0);
        for (int qc = 0; qc < p; ++qc) {
            const int buf = qc & 1;
#pragma unroll
            for (int e = 0; e < 4; ++e) Lc[buf][lr][lc + e] = st[e];
            __syncthreads();
            if (qc + 1 < p) fetch(qc + 1);
            if (r < rp) {
#pragma unroll 8
                for (int qq = 0; qq < 32; ++qq) {
                    const T l = Lc[buf][r][qq];
#pragma unroll
                    for (int e = 0; e < CPT; ++e) x[e] = x[e] - l * U[32 * qc + qq][cg + e];
                }
            }
        }
        if (r < rp) {
#pragma unroll
            for (int e = 0; e < CPT; ++e) U[32 * p + r][cg + e] = x[e];
        }
        {
            // the thread's four loads in flight together
            T dv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int idx = tid + 256 * e;
                const int i = idx >> 5, q = idx & 31;
                dv[e] = (i < rp && q < rp) ? A[(r0 + i) * lda + r0 + q] : T(0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int idx = tid + 256 * e;
                Ld[idx >> 5][idx & 31] = dv[e];
            }
        }
        __syncthreads();
        if (tid < TN) {
            // the column's 32 entries in registers (fully unrolled, so every index is a
            // constant); the diagonal-block entries are broadcast reads that do not depend on
            // the chain, so only the products and subtractions are on it
            const int c = tid;
            T u[32];
#pragma unroll
            for (int i = 0; i < 32; ++i) u[i] = (i < rp) ? U[32 * p + i][c] : T(0);
#pragma unroll
            for (int i = 1; i < 32; ++i) {
                if (i < rp) {
                    T sum = u[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) sum = sum - Ld[i][q] * u[q];
                    u[i] = sum;
                }
            }
#pragma unroll
            for (int i = 0; i < 32; ++i)
                if (i < rp) U[32 * p + i][c] = u[i];
            if (c < ncols) {
#pragma unroll
                for (int i = 0; i < 32; ++i)
                    if (i < rp) A[(r0 + i) * lda + cb + c] = u[i];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------- GEMM

// C[M x N] -= A[M x K] * B[K x N], all row-major (leading dimensions lda, ldb, ldc).
// 256 threads = 4 waves in a 2 x 2 arrangement; each wave owns a (BM/2) x (BN/2) block of
// 16 x 16 MFMA tiles. A and B tiles are staged k-major in LDS so each lane's MFMA operand
// (A[i][k] at lane i + 16k, B[k][j] at lane j + 16k) is one conflict-free LDS read.
// The product is accumulated from zero and subtracted from C once at the end, the
// reference's delta-then-subtract structure.
// kOneStage: K <= BK (the block-column update, K = the inner panel width): no K loop, so the
// staging registers are free before the products, and the register budget is set for five
// (FP32) or four (FP64) workgroups per compute unit (no spills): the update runs on the 32
// compute units of the panel stream, where resident workgroups, not arithmetic, set its time.
template <typename T, int BM, int BN, int BK, bool kOneStage = false>
__global__ __launch_bounds__(256, kOneStage ? (sizeof(T) == 4 ? 5 : 4) : 1) void gemm_sub_kernel(const T* __restrict__ A, long long lda,
    const T* __restrict__ B, long long ldb, T* __restrict__ C, long long ldc, int M, int N, int K,
    int tiles_n)
{
    using MF = Mfma16<T>;
    using acc_t = typename MF::acc_t;
    constexpr int WM = BM / 2, WN = BN / 2;
    constexpr int MI = WM / 16, NI = WN / 16;
    constexpr int PAD = 2;
    __shared__ T As[BK][BM + PAD];
    __shared__ T Bs[BK][BN + PAD];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int tile = blockIdx.x;
    const int tm = tile / tiles_n, tn = tile - tm * tiles_n;
    const long long row0 = static_cast<long long>(tm) * BM;
    const long long col0 = static_cast<long long>(tn) * BN;

    acc_t acc[MI][NI];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = acc_t{0, 0, 0, 0};
    // C's old values are loaded with the first operand tiles, so the tile costs one memory
    // round trip instead of two (the block-column update has K = 32: a single operand stage)
    T cpre[MI][NI][4];
#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long gr = row0 + wm * WM + mi * 16 + MF::row(lane, v);
                const long long gc = col0 + wn * WN + ni * 16 + (lane & 15);
                cpre[mi][ni][v] = (gr < M && gc < N) ? C[gr * ldc + gc] : T(0);
            }

    constexpr int A_PER = BM * BK / 256;
    constexpr int B_PER = BK * BN / 256;
    T ra[A_PER], rb[B_PER];

    auto load_tiles = [&](int k0) {
#pragma unroll
        for (int q = 0; q < A_PER; ++q) {
            const int idx = tid + q * 256;
            const int r = idx / BK, kk = idx - (idx / BK) * BK;
            const long long gr = row0 + r;
            const int gk = k0 + kk;
            ra[q] = (gr < M && gk < K) ? A[gr * lda + gk] : T(0);
        }
#pragma unroll
        for (int q = 0; q < B_PER; ++q) {
            const int idx = tid + q * 256;
            const int kk = idx / BN, c = idx - (idx / BN) * BN;
            const long long gc = col0 + c;
            const int gk = k0 + kk;
            rb[q] = (gc < N && gk < K) ? B[static_cast<long long>(gk) * ldb + gc] : T(0);
        }
    };
    auto store_tiles = [&]() {
#pragma unroll
        for (int q = 0; q < A_PER; ++q) {
            const int idx = tid + q * 256;
            const int r = idx / BK, kk = idx - (idx / BK) * BK;
            As[kk][r] = ra[q];
        }
#pragma unroll
        for (int q = 0; q < B_PER; ++q) {
            const int idx = tid + q * 256;
            const int kk = idx / BN, c = idx - (idx / BN) * BN;
            Bs[kk][c] = rb[q];
        }
    };

    auto products = [&]() {
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T a[MI], bb[NI];
            const int kr = kk + (lane >> 4);
#pragma unroll
            for (int mi = 0; mi < MI; ++mi) a[mi] = As[kr][wm * WM + mi * 16 + (lane & 15)];
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) bb[ni] = Bs[kr][wn * WN + ni * 16 + (lane & 15)];
#pragma unroll
            for (int mi = 0; mi < MI; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) acc[mi][ni] = MF::mma(a[mi], bb[ni], acc[mi][ni]);
        }
    };
    load_tiles(0);
    if constexpr (kOneStage) {
        store_tiles();
        __syncthreads();
        products();
    } else {
        for (int k0 = 0; k0 < K; k0 += BK) {
            __syncthreads();
            store_tiles();
            __syncthreads();
            if (k0 + BK < K) load_tiles(k0 + BK);
            products();
        }
    }

#pragma unroll
    for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long gr = row0 + wm * WM + mi * 16 + MF::row(lane, v);
                const long long gc = col0 + wn * WN + ni * 16 + (lane & 15);
                if (gr < M && gc < N) C[gr * ldc + gc] = cpre[mi][ni][v] - acc[mi][ni][v];
            }
}

// Zero-fill of small control areas (a kernel rather than hipMemsetAsync, so the sequence
// stays valid inside a captured graph).
static __global__ void fill_u32_kernel(unsigned* p, unsigned v, long long count)
{
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// Gate of a solve that follows an asynchronous factorisation: *gate = 1 when the factors are
// singular (*info != 0; the reference's LU mode then leaves x unchanged) or when a bounded wait
// of the factorisation expired (*err != 0). In the second case the right-hand sides (n x nrhs,
// row stride ldx) are also overwritten with NaN, so a timed-out factorisation can never be
// read as a solution.
template <typename T>
__global__ void solve_gate_kernel(const int* __restrict__ info, const int* __restrict__ err, int* __restrict__ gate,
    T* __restrict__ x, long long n, long long nrhs, long long ldx)
{
    const bool expired = *err != 0;
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) *gate = (expired || *info != 0) ? 1 : 0;
    if (!expired) return;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long e = i; e < n * nrhs; e += stride) x[(e / nrhs) * ldx + e % nrhs] = static_cast<T>(__builtin_nan(""));
}

// NaN over x (n x nrhs, row stride ldx) when a bounded wait of the substitution expired.
template <typename T>
__global__ void poison_on_error_kernel(const int* __restrict__ err, T* __restrict__ x, long long n, long long nrhs,
    long long ldx)
{
    if (*err == 0) return;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < n * nrhs; e += stride)
        x[(e / nrhs) * ldx + e % nrhs] = static_cast<T>(__builtin_nan(""));
}

}  // namespace lrn
