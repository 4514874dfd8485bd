// FP32 trailing-update GEMM, C[M x N] -= A[M x K] * B[K x N] (row-major), on
// v_mfma_f32_16x16x4_f32 (the FP32 factorisation of the mixed solver).
//
// Tiling: 256 threads = 4 waves (2 x 2); a wave owns 64 x 64 of C as 4 x 4 tiles of 16 x 16
// (64 FP32 accumulators per lane). K tiles of BK = 16 are staged k-major in LDS (A
// transposed on the way in, B as is), double-buffered with one barrier per K tile; the
// float4 global loads of tile t+1 are issued before the 64 matrix instructions of tile t.
// BK = 16 keeps the LDS at 33 KB per workgroup, so three workgroups share a compute unit and
// hide each other's barriers and C traffic (measured at C[16128 x 16128] -= A[16128 x 256]
// B[256 x 16128]: 97 TF/s with three workgroups per unit against 91 TF/s for BK = 32 with two;
// BK = 64 with one: 67 TF/s; the same products in the same order either way).
// Persistent, XCD-aware grid (xcd_tile of lrn_gemm_f64.hpp): each XCD walks a contiguous run
// of C tiles in groups of tile rows, so the A and B panels a group shares stay in its L2.
// The product is accumulated from zero and subtracted from C once (delta-then-subtract, as
// the reference's trailing update, lu_decomposition_native.cpp:392-422).
#pragma once

#include "lrn_gemm_f64.hpp"

namespace lrn {

constexpr int kG32BM = 128, kG32BN = 128, kG32BK = 16;
constexpr int kG32PerCu = 3;  // workgroups per compute unit the persistent grid is sized for
typedef float g32v4 __attribute__((ext_vector_type(4)));  // native vector: stays in registers

// Global -> register staging of one K tile: 2 float4 of A and 2 float4 of B per thread.
constexpr int kG32Ld = kG32BK / 8;
template <bool kCheck>
__device__ __forceinline__ void g32_load(g32v4 (&ra)[kG32Ld], g32v4 (&rb)[kG32Ld], const float* __restrict__ A,
    long long lda, const float* __restrict__ B, long long ldb, long long row0, long long col0, int k0, int M, int N, int K,
    int tid)
{
#pragma unroll
    for (int q = 0; q < kG32Ld; ++q) {
        const int idx = tid + q * 256;  // 128 rows x 4 float4
        const int r = idx >> 2, c = (idx & 3) * 4;
        const long long gr = row0 + r;
        const int gk = k0 + c;
        if (!kCheck) {
            ra[q] = *reinterpret_cast<const g32v4*>(A + gr * lda + gk);
        } else {
            g32v4 v = {0.f, 0.f, 0.f, 0.f};
            if (gr < M) {
                const float* p = A + gr * lda + gk;
                if (gk < K) v.x = p[0];
                if (gk + 1 < K) v.y = p[1];
                if (gk + 2 < K) v.z = p[2];
                if (gk + 3 < K) v.w = p[3];
            }
            ra[q] = v;
        }
    }
#pragma unroll
    for (int q = 0; q < kG32Ld; ++q) {
        const int idx = tid + q * 256;  // 16 rows x 32 float4
        const int r = idx >> 5, c = (idx & 31) * 4;
        const int gk = k0 + r;
        const long long gc = col0 + c;
        if (!kCheck) {
            rb[q] = *reinterpret_cast<const g32v4*>(B + static_cast<long long>(gk) * ldb + gc);
        } else {
            g32v4 v = {0.f, 0.f, 0.f, 0.f};
            if (gk < K) {
                const float* p = B + static_cast<long long>(gk) * ldb + gc;
                if (gc < N) v.x = p[0];
                if (gc + 1 < N) v.y = p[1];
                if (gc + 2 < N) v.z = p[2];
                if (gc + 3 < N) v.w = p[3];
            }
            rb[q] = v;
        }
    }
}

// kCheck = false requires M, N multiples of 128, K a multiple of 16 and 16-byte aligned rows
// (the host launches interior tiles this way and the ragged strips with kCheck = true).
template <bool kCheck>
__global__ __launch_bounds__(256, kG32PerCu) void gemm_sub_f32_kernel(const float* __restrict__ A, long long lda,
    const float* __restrict__ B, long long ldb, float* __restrict__ C, long long ldc, int M, int N, int K, int tm,
    int tn, int per_xcd, const int* __restrict__ skip, long long tri)
{
    if (skip != nullptr && *skip != 0) return;  // e.g. a Cholesky factorisation that has failed
    constexpr int BM = kG32BM, BN = kG32BN, BK = kG32BK;
    constexpr int PAD = 4;
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __shared__ float As[2][BK][BM + PAD];
    __shared__ float Bs[2][BK][BN + PAD];

    const int nvirt = per_xcd * kXcds;
    for (int vt = blockIdx.x; vt < nvirt; vt += gridDim.x) {
        int tile_m, tile_n;
        if (!xcd_tile(vt, tm, tn, per_xcd, tile_m, tile_n)) continue;
        // lower-triangle mode (tri != kNoTriangle): a tile whose columns all lie right of its
        // rows (global column = local column + tri, relative to the same diagonal) is skipped
        if (tri != kNoTriangle &&
            static_cast<long long>(tile_n) * kG32BN + tri > static_cast<long long>(tile_m) * kG32BM + kG32BM - 1)
            continue;
        const int tid = threadIdx.x;
        const int lane = tid & 63;
        const int wid = tid >> 6;
        const int wm = wid >> 1, wn = wid & 1;
        const long long row0 = static_cast<long long>(tile_m) * BM;
        const long long col0 = static_cast<long long>(tile_n) * BN;

        acc_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = acc_t{0.f, 0.f, 0.f, 0.f};

        g32v4 ra[kG32Ld], rb[kG32Ld];
        const int nk = (K + BK - 1) / BK;
        g32_load<kCheck>(ra, rb, A, lda, B, ldb, row0, col0, 0, M, N, K, tid);

#define LRN_G32_STORE(buf)                                                              \
    do {                                                                                \
        _Pragma("unroll") for (int q = 0; q < kG32Ld; ++q)                              \
        {                                                                               \
            const int idx = tid + q * 256;                                              \
            const int r = idx >> 2, c = (idx & 3) * 4;                                  \
            As[buf][c][r] = ra[q].x;                                                    \
            As[buf][c + 1][r] = ra[q].y;                                                \
            As[buf][c + 2][r] = ra[q].z;                                                \
            As[buf][c + 3][r] = ra[q].w;                                                \
        }                                                                               \
        _Pragma("unroll") for (int q = 0; q < kG32Ld; ++q)                              \
        {                                                                               \
            const int idx = tid + q * 256;                                              \
            const int r = idx >> 5, c = (idx & 31) * 4;                                 \
            *reinterpret_cast<g32v4*>(&Bs[buf][r][c]) = rb[q];                         \
        }                                                                               \
    } while (0)
        LRN_G32_STORE(0);
        __syncthreads();
        const int am = wm * 64 + (lane & 15);
        const int bn = wn * 64 + (lane & 15);
        const int kl = lane >> 4;
        for (int t = 0; t < nk; ++t) {
            const int buf = t & 1;
            if (t + 1 < nk) g32_load<kCheck>(ra, rb, A, lda, B, ldb, row0, col0, (t + 1) * BK, M, N, K, tid);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                float a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) a[i] = As[buf][kk + kl][am + 16 * i];
#pragma unroll
                for (int j = 0; j < 4; ++j) b[j] = Bs[buf][kk + kl][bn + 16 * j];
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
            }
            if (t + 1 < nk) LRN_G32_STORE(buf ^ 1);
            __syncthreads();
        }
#undef LRN_G32_STORE

        // result element v of tile (i, j) in this lane: row 16i + 4 (lane >> 4) + v,
        // column 16j + (lane & 15)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long gr = row0 + wm * 64 + 16 * i + 4 * (lane >> 4) + v;
                if (kCheck && gr >= M) continue;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const long long gc = col0 + wn * 64 + 16 * j + (lane & 15);
                    if (!kCheck || gc < N) {
                        float* cp = C + gr * ldc + gc;
                        *cp = *cp - acc[i][j][v];
                    }
                }
            }
        __syncthreads();  // LDS is reused by the next tile
    }
}

}  // namespace lrn
