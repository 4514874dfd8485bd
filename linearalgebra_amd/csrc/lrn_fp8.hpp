// Trailing update on the FP8 matrix cores with per-block scaling:
//     C[M x N] -= A[M x K] * B[K x N]      (FP32 data, row-major, FP32 accumulation)
// with each operand split into `P` planes of FP8 e4m3 values (OCP e4m3fn: 3 mantissa bits,
// largest finite 448) and one power-of-two (E8M0) scale per 32 consecutive K elements of a row of
// A or of a column of B, the block-scaled format that v_mfma_scale_f32_16x16x128_f8f6f4 consumes:
//     block scale  2^e with e = E - 8 for the block's largest magnitude m 2^E (m in [0.5, 1)), so
//                  the scaled block lies within [-256, 256] and never reaches e4m3's NaN encoding;
//     plane 0      q0 = rne_e4m3(x / 2^e0), plane p the same of the residual x - sum_{i<p} deq(qi),
//                  each plane with its own block scale (the subtraction is exact in FP32);
//     product      sum over plane pairs (i, j) with i + j < P (1, 3 or 6 products), accumulated in
//                  FP32 by the matrix cores, subtracted from C once (delta-then-subtract).
// One plane keeps about 4 significant bits of each element, two about 8, three about 12.
//
// Operand layout of v_mfma_scale_f32_16x16x128_f8f6f4 on gfx950 (measured with one-hot FP8
// operands): lane l holds row l % 16 of A (column l % 16 of B); with g = l / 16 its bytes 0-15
// are k = 16 g + 0..15 and bytes 16-31 are k = 64 + 16 g + 0..15; the scale of the 32-element
// block k in [32 kb, 32 kb + 32) of row r is read from lane r + 16 kb (byte 0); the FP32 result
// is D[4 (l / 16) + v][l % 16] in element v of lane l. A block's 32 elements are therefore the
// first (or second) halves of lanes l and l ^ 16, which exchange their partial maxima.
//
// Workgroup: 256 threads, a 64 x 64 tile of C, four waves of 32 x 32 (2 x 2 MFMA tiles); K in
// steps of 128 staged through LDS as FP32 (A row-major, B transposed), quantised by each wave
// into registers, then P (P + 1) / 2 scaled MFMAs per tile pair.
#pragma once

#include <hip/hip_runtime.h>

#include "lrn_device.hpp"

namespace lrn {

constexpr int kF8BM = 64, kF8BN = 64, kF8BK = 128;
constexpr int kF8Pad = 4;  // LDS row padding (floats): conflict-free 16-byte reads of 16 rows

typedef int f8x32_t __attribute__((ext_vector_type(8)));   // 32 e4m3 values
typedef float f8acc_t __attribute__((ext_vector_type(4)));

// Exponent e of a block's E8M0 scale from its largest magnitude: the scaled block lies within
// [-256, 256]. A zero block gets e = -127 (its values stay zero); NaN and infinity propagate.
__device__ __forceinline__ int fp8_block_exp(float amax)
{
    if (!(amax > 0.0f)) return -127;
    if (!(amax <= 3.4e38f)) return 127;
    int E = 0;
    (void)frexpf(amax, &E);
    const int e = E - 8;
    return e < -127 ? -127 : (e > 127 ? 127 : e);
}

// One plane of a lane's 32 operand values (bytes 0-15: r[0..15], bytes 16-31: r[16..31]):
// quantises r, leaves the residual in r, and returns the lane's scale register (the exponent
// of the block this lane supplies the scale for, as E8M0 in byte 0).
__device__ __forceinline__ int fp8_plane(float (&r)[32], f8x32_t& q, int lane)
{
    float lo = 0.0f, hi = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        lo = fmaxf(lo, fabsf(r[j]));
        hi = fmaxf(hi, fabsf(r[16 + j]));
    }
    // NaN survives fmaxf only as the second operand; keep it explicitly
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (r[j] != r[j]) lo = r[j];
        if (r[16 + j] != r[16 + j]) hi = r[16 + j];
    }
    const float plo = __shfl_xor(lo, 16, kWave), phi = __shfl_xor(hi, 16, kWave);
    lo = (lo != lo || plo != plo) ? __builtin_nanf("") : fmaxf(lo, plo);
    hi = (hi != hi || phi != phi) ? __builtin_nanf("") : fmaxf(hi, phi);
    const int elo = fp8_block_exp(lo), ehi = fp8_block_exp(hi);
#pragma unroll
    for (int d = 0; d < 8; ++d) {
        const int j = 4 * d;
        const int e = d < 4 ? elo : ehi;
        const float s = ldexpf(1.0f, -e);
        int w = __builtin_amdgcn_cvt_pk_fp8_f32(r[j] * s, r[j + 1] * s, 0, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(r[j + 2] * s, r[j + 3] * s, w, true);
        q[d] = w;
        const float back = ldexpf(1.0f, e);
        r[j] -= __builtin_amdgcn_cvt_f32_fp8(w, 0) * back;
        r[j + 1] -= __builtin_amdgcn_cvt_f32_fp8(w, 1) * back;
        r[j + 2] -= __builtin_amdgcn_cvt_f32_fp8(w, 2) * back;
        r[j + 3] -= __builtin_amdgcn_cvt_f32_fp8(w, 3) * back;
    }
    // lane group g supplies the scale of block g: g = 0 its own low half, g = 1 the low half of
    // group 2 (lane + 16), g = 2 the high half of group 0 (lane - 32), g = 3 its own high half
    const int g = lane >> 4;
    const int src = g == 1 ? lane + 16 : (g == 2 ? lane - 32 : lane);
    const int both = (elo + 127) | ((ehi + 127) << 8);
    const int got = __shfl(both, src, kWave);
    return (g <= 1) ? (got & 0xff) : ((got >> 8) & 0xff);
}

template <int P>
__global__ __launch_bounds__(256) void gemm_sub_fp8_kernel(const float* __restrict__ A, long long lda,
    const float* __restrict__ B, long long ldb, float* __restrict__ C, long long ldc, int M, int N, int K, int tiles_n,
    const int* __restrict__ skip)
{
    if (skip != nullptr && *skip != 0) return;
    __shared__ float As[kF8BM][kF8BK + kF8Pad];
    __shared__ float Bs[kF8BN][kF8BK + kF8Pad];  // transposed: Bs[c][k] = B[k][c]

    const int tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int g = lane >> 4;
    const long long row0 = static_cast<long long>(blockIdx.x / tiles_n) * kF8BM;
    const long long col0 = static_cast<long long>(blockIdx.x % tiles_n) * kF8BN;

    f8acc_t acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = f8acc_t{0.0f, 0.0f, 0.0f, 0.0f};

    for (int k0 = 0; k0 < K; k0 += kF8BK) {
        __syncthreads();
        // operand tiles in batches of 16 loads per thread, all in flight before any is written
        // to LDS (a rolled loop would wait for each load in turn)
        constexpr int kBatch = 16;
#pragma unroll
        for (int base = 0; base < kF8BM * kF8BK; base += 256 * kBatch) {
            float v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int idx = base + tid + 256 * u;
                const int r = idx / kF8BK, k = idx % kF8BK;
                const long long gr = row0 + r;
                v[u] = (gr < M && k0 + k < K) ? A[gr * lda + k0 + k] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int idx = base + tid + 256 * u;
                As[idx / kF8BK][idx % kF8BK] = v[u];
            }
        }
#pragma unroll
        for (int base = 0; base < kF8BK * kF8BN; base += 256 * kBatch) {
            float v[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int idx = base + tid + 256 * u;
                const int k = idx / kF8BN, c = idx % kF8BN;
                const long long gc = col0 + c;
                v[u] = (gc < N && k0 + k < K) ? B[static_cast<long long>(k0 + k) * ldb + gc] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kBatch; ++u) {
                const int idx = base + tid + 256 * u;
                Bs[idx % kF8BN][idx / kF8BN] = v[u];
            }
        }
        __syncthreads();

        // this wave's two row tiles of A, quantised
        f8x32_t qa[2][P];
        int sa[2][P];
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            float r[32];
            const float* src = &As[wm * 32 + mi * 16 + (lane & 15)][0];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                r[j] = src[16 * g + j];
                r[16 + j] = src[64 + 16 * g + j];
            }
#pragma unroll
            for (int p = 0; p < P; ++p) sa[mi][p] = fp8_plane(r, qa[mi][p], lane);
        }
#pragma unroll
        for (int ni = 0; ni < 2; ++ni) {
            f8x32_t qb[P];
            int sb[P];
            float r[32];
            const float* src = &Bs[wn * 32 + ni * 16 + (lane & 15)][0];
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                r[j] = src[16 * g + j];
                r[16 + j] = src[64 + 16 * g + j];
            }
#pragma unroll
            for (int p = 0; p < P; ++p) sb[p] = fp8_plane(r, qb[p], lane);
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int i = 0; i < P; ++i)
#pragma unroll
                    for (int j = 0; j + i < P; ++j)
                        acc[mi][ni] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(qa[mi][i], qb[j], acc[mi][ni],
                            0, 0, 0, sa[mi][i], 0, sb[j]);
        }
    }

#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long gr = row0 + wm * 32 + mi * 16 + 4 * g + v;
                const long long gc = col0 + wn * 32 + ni * 16 + (lane & 15);
                if (gr < M && gc < N) {
                    float* cp = C + gr * ldc + gc;
                    *cp = *cp - acc[mi][ni][v];
                }
            }
}

}  // namespace lrn
