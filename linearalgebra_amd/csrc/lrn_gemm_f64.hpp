// FP64 trailing-update GEMM, C[M x N] -= A[M x K] * B[K x N] (row-major), on the 4-block
// 4x4x4 FP64 matrix instruction.
//
// On gfx950 v_mfma_f64_4x4x4_4b sustains a far higher FP64 rate than the 16x16x4 form
// (measured with register-only loops: ~74 TF/s against ~48 TF/s). One instruction computes
// four independent 4x4x4 products; block g (lanes with (lane >> 2) & 3 == g) of the
// instruction reads A[i][k] from lane 16k + 4g + i and B[k][j] from lane 16k + 4g + j and
// adds the product to lane 16i + 4g + j. Feeding every block the same four A rows and a
// different group of four B columns turns one instruction into a 4 x 16 x 4 update:
//   A operand: A[lane & 3][lane >> 4]      (the same value in the four blocks)
//   B operand: B[lane >> 4][lane & 15]
//   result   : D[lane >> 4][lane & 15]
//
// Tiling: 512 threads = 8 waves (4 x 2); a wave owns 32 x 64 of C as 8 x 4 such 4 x 16
// tiles (32 FP64 accumulators per lane), small enough for four waves per SIMD. A and B tiles of
// BK = 8 are staged k-major in LDS, double-buffered with one barrier per K tile; global loads of
// tile t+1 are issued before the MFMAs of tile t. (BK = 16 needs more registers than four waves
// per SIMD leave and spilled 44 bytes per lane: 52.5 against 54.4 TF/s for BK = 8 at
// C[16128 x 16128] -= A[16128 x 256] B[256 x 16128], with the same products in the same order.) Workgroup-to-tile mapping is XCD-aware: consecutive workgroup
// ids go round-robin to the 8 XCDs, so each XCD receives a contiguous run of tiles, walked
// in groups of 8 tile rows so the A and B panels a group shares stay in that XCD's L2.
// The product is accumulated from zero and subtracted from C once (the delta-then-subtract
// structure of the reference's trailing update, lu_decomposition_native.cpp:392-422).
#pragma once

#include "lrn_device.hpp"

namespace lrn {

constexpr int kG64BM = 128, kG64BN = 128, kG64BK = 8;
constexpr int kG64Ld = kG64BK / 8;  // 16-byte pairs of A and of B each thread stages per K tile
// tri argument of the trailing-update GEMMs that disables the lower-triangle mode
constexpr long long kNoTriangle = -(1ll << 62);
typedef double g64v2 __attribute__((ext_vector_type(2)));  // native vector: stays in registers
constexpr int kGroupRows = 8;

__device__ __forceinline__ double mfma4x4(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// Maps a workgroup id to a C tile (tm x tn tiles); returns false for padding workgroups.
__device__ __forceinline__ bool xcd_tile(int wg, int tm, int tn, int per_xcd, int& tile_m, int& tile_n)
{
    const int t = (wg % kXcds) * per_xcd + wg / kXcds;
    if (t >= tm * tn) return false;
    const int span = kGroupRows * tn;
    const int group = t / span;
    const int first = group * kGroupRows;
    const int gm = min(tm - first, kGroupRows);
    const int r = t - group * span;
    tile_m = first + r % gm;
    tile_n = r / gm;
    return true;
}

// Global -> register staging of one K tile (BK = 8): one pair of A and one pair of B per thread.
// kCheck selects the bounds-checked form for edge tiles; interior tiles use plain 16-byte loads.
template <bool kCheck, bool kVec>
__device__ __forceinline__ void g64_load(g64v2 (&ra)[kG64Ld], g64v2 (&rb)[kG64Ld], const double* __restrict__ A,
    long long lda, const double* __restrict__ B, long long ldb, long long row0, long long col0, int k0, int M, int N, int K,
    int tid)
{
#pragma unroll
    for (int q = 0; q < kG64Ld; ++q) {
        const int idx = tid + q * 512;  // 128 rows x BK/2 pairs
        const int r = idx / (kG64BK / 2), c = (idx % (kG64BK / 2)) * 2;
        const long long gr = row0 + r;
        const int gk = k0 + c;
        if (!kCheck) {
            ra[q] = *reinterpret_cast<const g64v2*>(A + gr * lda + gk);
        } else {
            g64v2 v = {0.0, 0.0};
            if (gr < M) {
                const double* p = A + gr * lda + gk;
                if (kVec && gk + 1 < K) {
                    v = *reinterpret_cast<const g64v2*>(p);
                } else {
                    if (gk < K) v.x = p[0];
                    if (gk + 1 < K) v.y = p[1];
                }
            }
            ra[q] = v;
        }
    }
#pragma unroll
    for (int q = 0; q < kG64Ld; ++q) {
        const int idx = tid + q * 512;  // BK rows x 64 pairs
        const int r = idx >> 6, c = (idx & 63) * 2;
        const int gk = k0 + r;
        const long long gc = col0 + c;
        if (!kCheck) {
            rb[q] = *reinterpret_cast<const g64v2*>(B + static_cast<long long>(gk) * ldb + gc);
        } else {
            g64v2 v = {0.0, 0.0};
            if (gk < K) {
                const double* p = B + static_cast<long long>(gk) * ldb + gc;
                if (kVec && gc + 1 < N) {
                    v = *reinterpret_cast<const g64v2*>(p);
                } else {
                    if (gc < N) v.x = p[0];
                    if (gc + 1 < N) v.y = p[1];
                }
            }
            rb[q] = v;
        }
    }
}

// 512 threads. kCheck = false requires M, N multiples of 128, K a multiple of 8 and 16-byte
// aligned rows (the host launches interior tiles this way and edge strips with kCheck = true).
template <bool kCheck>
__global__ __launch_bounds__(512, 4) void gemm_sub_f64_kernel(const double* __restrict__ A, long long lda,
    const double* __restrict__ B, long long ldb, double* __restrict__ C, long long ldc, int M, int N, int K, int tm,
    int tn, int per_xcd, const int* __restrict__ skip, long long tri)
{
    if (skip != nullptr && *skip != 0) return;  // e.g. a Cholesky factorisation that has failed
    constexpr int BM = kG64BM, BN = kG64BN, BK = kG64BK;
    constexpr int PADA = 4, PADB = 4;
    constexpr int RG = 8, CG = 4;  // 4-row groups and 16-column groups per wave (32 x 64 of C)
    __shared__ double As[2][BK][BM + PADA];
    __shared__ double Bs[2][BK][BN + PADB];

    // persistent: workgroup g handles virtual tiles g, g + grid, ... (all workgroups see the
    // same trip count, so the barriers inside the loop are uniform)
    const int nvirt = per_xcd * kXcds;
    for (int vt = blockIdx.x; vt < nvirt; vt += gridDim.x) {
    int tile_m, tile_n;
    if (!xcd_tile(vt, tm, tn, per_xcd, tile_m, tile_n)) continue;
    // lower-triangle mode (tri != kNoTriangle): a tile whose columns all lie right of its rows
    // (global column = local column + tri, relative to the same diagonal) is skipped
    if (tri != kNoTriangle &&
        static_cast<long long>(tile_n) * kG64BN + tri > static_cast<long long>(tile_m) * kG64BM + kG64BM - 1)
        continue;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const long long row0 = static_cast<long long>(tile_m) * BM;
    const long long col0 = static_cast<long long>(tile_n) * BN;

    double acc[RG][CG];
#pragma unroll
    for (int a = 0; a < RG; ++a)
#pragma unroll
        for (int b = 0; b < CG; ++b) acc[a][b] = 0.0;

    g64v2 ra[kG64Ld], rb[kG64Ld];
    const int nk = (K + BK - 1) / BK;
    g64_load<kCheck, false>(ra, rb, A, lda, B, ldb, row0, col0, 0, M, N, K, tid);

    // LDS store of the staged tile: A transposed to k-major, B as is
#define LRN_G64_STORE(buf)                                                              \
    do {                                                                                \
        _Pragma("unroll") for (int q = 0; q < kG64Ld; ++q)                              \
        {                                                                               \
            const int idx = tid + q * 512;                                              \
            const int r = idx / (kG64BK / 2), c = (idx % (kG64BK / 2)) * 2;             \
            As[buf][c][r] = ra[q].x;                                                    \
            As[buf][c + 1][r] = ra[q].y;                                                \
        }                                                                               \
        _Pragma("unroll") for (int q = 0; q < kG64Ld; ++q)                              \
        {                                                                               \
            const int idx = tid + q * 512;                                              \
            const int r = idx >> 6, c = (idx & 63) * 2;                                 \
            *reinterpret_cast<g64v2*>(&Bs[buf][r][c]) = rb[q];                        \
        }                                                                               \
    } while (0)

    LRN_G64_STORE(0);
    __syncthreads();
    const int arow = wm * 32 + (lane & 3);
    const int bcol = wn * 64 + (lane & 15);
    const int kl = lane >> 4;
    for (int t = 0; t < nk; ++t) {
        const int buf = t & 1;
        const int kn = (t + 1) * BK;
        if (t + 1 < nk) g64_load<kCheck, false>(ra, rb, A, lda, B, ldb, row0, col0, kn, M, N, K, tid);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double b[CG];
#pragma unroll
            for (int c = 0; c < CG; ++c) b[c] = Bs[buf][kk + kl][bcol + 16 * c];
#pragma unroll
            for (int g = 0; g < RG; ++g) {
                const double a = As[buf][kk + kl][arow + 4 * g];
#pragma unroll
                for (int c = 0; c < CG; ++c) acc[g][c] = mfma4x4(a, b[c], acc[g][c]);
            }
        }
        if (t + 1 < nk) LRN_G64_STORE(buf ^ 1);
        __syncthreads();
    }
#undef LRN_G64_STORE

    // result of tile (g, c) in this lane: row 4g + (lane >> 4), column 16c + (lane & 15)
#pragma unroll
    for (int g = 0; g < RG; ++g) {
        const long long gr = row0 + wm * 32 + 4 * g + (lane >> 4);
        if (kCheck && gr >= M) continue;
#pragma unroll
        for (int c = 0; c < CG; ++c) {
            const long long gc = col0 + wn * 64 + 16 * c + (lane & 15);
            if (!kCheck || gc < N) {
                double* cp = C + gr * ldc + gc;
                *cp = *cp - acc[g][c];
            }
        }
    }
    __syncthreads();  // LDS is reused by the next tile
    }
}

}  // namespace lrn
