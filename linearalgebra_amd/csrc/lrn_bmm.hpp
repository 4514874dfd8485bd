// Batched square multiplication C_i = A_i B_i (count row-major dim x dim matrices stored back
// to back). One workgroup per 64 x 64 output tile of one matrix: 256 threads, a 4 x 4 register
// tile each, the k range in 16-deep steps staged through LDS (A transposed so both operands are
// read as rows). Each output is accumulated with fma in increasing k, the order of the
// reference's one-thread-per-element loop (src/matrix_operation/gpu/hipblas/
// multiplication_batched_hip.hip:12-37). On gfx950 the FP64 and FP32 vector FMA rates equal the
// dense matrix-core rates, so this form is bounded by the same peak.
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

constexpr int kBmmTile = 64;
constexpr int kBmmK = 16;

template <typename T>
__global__ __launch_bounds__(256) void batched_gemm_kernel(int dim, long long count, int tiles_n, long long tiles,
    const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C)
{
    __shared__ T As[kBmmK][kBmmTile + 1];
    __shared__ T Bs[kBmmK][kBmmTile + 1];
    const long long wg = blockIdx.x;
    const long long mat = wg / tiles;
    if (mat >= count) return;
    const int tile = static_cast<int>(wg - mat * tiles);
    const int row0 = (tile / tiles_n) * kBmmTile, col0 = (tile % tiles_n) * kBmmTile;
    const long long off = mat * static_cast<long long>(dim) * dim;
    const T* a = A + off;
    const T* b = B + off;
    T* c = C + off;
    const int t = threadIdx.x;
    const int tr = (t / 16) * 4, tc = (t % 16) * 4;
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
    for (int k0 = 0; k0 < dim; k0 += kBmmK) {
        // A tile 64 x 16 and B tile 16 x 64: 1024 elements each, 4 per thread
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int e = t + 256 * q;
            const int ar = e / kBmmK, ak = e % kBmmK;
            const int gr = row0 + ar, gk = k0 + ak;
            As[ak][ar] = (gr < dim && gk < dim) ? a[static_cast<long long>(gr) * dim + gk] : T(0);
            const int bk = e / kBmmTile, bc = e % kBmmTile;
            const int hk = k0 + bk, hc = col0 + bc;
            Bs[bk][bc] = (hk < dim && hc < dim) ? b[static_cast<long long>(hk) * dim + hc] : T(0);
        }
        __syncthreads();
        const int kk = dim - k0 < kBmmK ? dim - k0 : kBmmK;
        for (int k = 0; k < kk; ++k) {
            T av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) av[i] = As[k][tr + i];
#pragma unroll
            for (int j = 0; j < 4; ++j) bv[j] = Bs[k][tc + j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fma(av[i], bv[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gr = row0 + tr + i;
        if (gr >= dim) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gc = col0 + tc + j;
            if (gc < dim) c[static_cast<long long>(gr) * dim + gc] = acc[i][j];
        }
    }
}

// Small matrices (dim <= kBmmSmall): one thread per output element over the whole batch, the
// operands read straight from memory (each row and column is reused from the caches by the
// dim threads that share it); the same fma chain in increasing k as the tiled kernel, so the
// two agree bit for bit.
constexpr int kBmmSmall = 16;

template <typename T>
__global__ __launch_bounds__(256) void batched_gemm_small_kernel(int dim, long long total, const T* __restrict__ A,
    const T* __restrict__ B, T* __restrict__ C)
{
    const long long idx = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const long long d2 = static_cast<long long>(dim) * dim;
    const long long mat = idx / d2;
    const int rem = static_cast<int>(idx - mat * d2);
    const int i = rem / dim, j = rem - i * dim;
    const T* a = A + mat * d2 + static_cast<long long>(i) * dim;
    const T* b = B + mat * d2 + j;
    T acc = T(0);
    for (int k = 0; k < dim; ++k) acc = fma(a[k], b[static_cast<long long>(k) * dim], acc);
    C[idx] = acc;
}

}  // namespace lrn
