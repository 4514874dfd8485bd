// Low-precision trailing update of the mixed-precision factorisation on the bf16 matrix
// cores, A, B, C FP32 row-major, FP32 accumulation. Every operand x is split into NP bf16
// planes, x ~ p0 + p1 + p2 with p0 = bf16(x), p1 = bf16(x - p0), p2 = bf16(x - p0 - p1), and
// C -= sum of the plane products that matter at that accuracy:
//   NP = 1: A0 B0                                   (bf16 operands, ~2^-8 relative)
//   NP = 2: A0 B0 + A0 B1 + A1 B0                   (~2^-16)
//   NP = 3: the NP = 2 terms + A0 B2 + A2 B0 + A1 B1 (~2^-24, FP32-level factors)
// Products are accumulated smallest first.
//
// v_mfma_f32_16x16x32_bf16 (gfx950): lane l supplies A[l % 16][8*(l/16) .. 8*(l/16)+7] and
// B[8*(l/16) .. +7][l % 16] as 8 packed bf16; the FP32 result element v of lane l is
// D[4*(l/16) + v][l % 16] (checked on gfx950 with one-hot operands).
// Tiling: 256 threads = 4 waves (2 x 2), 64 x 64 of C per wave (4 x 4 MFMA tiles), K staged in
// steps of 32; operands are rounded to bf16 (round to nearest even) as they are staged in LDS,
// A row-major and B transposed, so every operand fragment is one 16-byte LDS read.
#pragma once

#include "lrn_device.hpp"

namespace lrn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kBfBM = 128, kBfBN = 128, kBfBK = 32, kBfPad = 8;

template <int NP>
__global__ __launch_bounds__(256) void gemm_sub_bf16_kernel(const float* __restrict__ A, long long lda,
    const float* __restrict__ B, long long ldb, float* __restrict__ C, long long ldc, int M, int N, int K, int tm,
    int tn, int per_xcd)
{
    static_assert(NP >= 1 && NP <= 3, "one to three bf16 planes");
    __shared__ __align__(16) __bf16 As[NP][2][kBfBM][kBfBK + kBfPad];
    __shared__ __align__(16) __bf16 Bs[NP][2][kBfBN][kBfBK + kBfPad];

    const int nvirt = per_xcd * 8;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    for (int vt = blockIdx.x; vt < nvirt; vt += gridDim.x) {
        const int t = (vt % 8) * per_xcd + vt / 8;
        if (t >= tm * tn) continue;
        const int span = 8 * tn;
        const int group = t / span;
        const int first = group * 8;
        const int gm = min(tm - first, 8);
        const int r0 = t - group * span;
        const long long row0 = static_cast<long long>(first + r0 % gm) * kBfBM;
        const long long col0 = static_cast<long long>(r0 / gm) * kBfBN;

        f32x4_t acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = f32x4_t{0.f, 0.f, 0.f, 0.f};

        // staging: A 128 x 32 (thread: row tid/2, 16 consecutive k), B 32 x 128 (thread: k tid/8, 16 consecutive j)
        float ra[16], rb[16];
        const int ar = tid >> 1, ak = (tid & 1) * 16;
        const int bk = tid >> 3, bj = (tid & 7) * 16;
        auto load = [&](int k0) {
            const long long gr = row0 + ar;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const int gk = k0 + ak + e;
                ra[e] = (gr < M && gk < K) ? A[gr * lda + gk] : 0.f;
            }
            const int gk = k0 + bk;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                const long long gc = col0 + bj + e;
                rb[e] = (gk < K && gc < N) ? B[static_cast<long long>(gk) * ldb + gc] : 0.f;
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float rem[8];
#pragma unroll
                for (int e = 0; e < 8; ++e) rem[e] = ra[8 * h + e];
#pragma unroll
                for (int pl = 0; pl < NP; ++pl) {
                    bf16x8_t v;
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        v[e] = static_cast<__bf16>(rem[e]);
                        rem[e] = rem[e] - static_cast<float>(v[e]);
                    }
                    *reinterpret_cast<bf16x8_t*>(&As[pl][buf][ar][ak + 8 * h]) = v;
                }
            }
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                float rem = rb[e];
#pragma unroll
                for (int pl = 0; pl < NP; ++pl) {
                    const __bf16 h = static_cast<__bf16>(rem);
                    Bs[pl][buf][bj + e][bk] = h;
                    rem = rem - static_cast<float>(h);
                }
            }
        };

        const int nk = (K + kBfBK - 1) / kBfBK;
        load(0);
        store(0);
        __syncthreads();
        for (int s = 0; s < nk; ++s) {
            const int buf = s & 1;
            if (s + 1 < nk) load((s + 1) * kBfBK);
            bf16x8_t a[NP][4], b[NP][4];
#pragma unroll
            for (int pl = 0; pl < NP; ++pl) {
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
                    a[pl][mi] = *reinterpret_cast<const bf16x8_t*>(
                        &As[pl][buf][wm * 64 + mi * 16 + (lane & 15)][8 * (lane >> 4)]);
#pragma unroll
                for (int ni = 0; ni < 4; ++ni)
                    b[pl][ni] = *reinterpret_cast<const bf16x8_t*>(
                        &Bs[pl][buf][wn * 64 + ni * 16 + (lane & 15)][8 * (lane >> 4)]);
            }
            // plane pairs (pa, pb), smallest products first
#pragma unroll
            for (int q = 0; q < (NP == 1 ? 1 : (NP == 2 ? 3 : 6)); ++q) {
                constexpr int kPa3[6] = {2, 0, 1, 1, 0, 0}, kPb3[6] = {0, 2, 1, 0, 1, 0};
                constexpr int kPa2[3] = {1, 0, 0}, kPb2[3] = {0, 1, 0};
                const int pa = NP == 3 ? kPa3[q] : (NP == 2 ? kPa2[q] : 0);
                const int pb = NP == 3 ? kPb3[q] : (NP == 2 ? kPb2[q] : 0);
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                    for (int ni = 0; ni < 4; ++ni)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[pa][mi], b[pb][ni], acc[mi][ni], 0, 0, 0);
            }
            if (s + 1 < nk) store(buf ^ 1);
            __syncthreads();
        }
#pragma unroll
        for (int mi = 0; mi < 4; ++mi)
#pragma unroll
            for (int ni = 0; ni < 4; ++ni)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const long long gr = row0 + wm * 64 + mi * 16 + 4 * (lane >> 4) + v;
                    const long long gc = col0 + wn * 64 + ni * 16 + (lane & 15);
                    if (gr < M && gc < N) {
                        float* cp = C + gr * ldc + gc;
                        *cp = *cp - acc[mi][ni][v];
                    }
                }
    }
}

}  // namespace lrn
