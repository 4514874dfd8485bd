// Tournament-pivoted panel (communication-avoiding LU) for the FP32 factorisation of the
// mixed-precision solver, where the pivots only have to give a stable factorisation, not
// the reference's exact sequence.
//
// calu_select_kernel: every wave takes 64*Q consecutive rows of the m x w panel, holds them
//   in registers and runs partial-pivoting elimination on them alone, which nominates w
//   candidate rows. The candidates of a workgroup's four waves are reduced the same way
//   (elimination on their original values) to w rows per workgroup; the last workgroup to
//   arrive (one atomic counter, no waiting) reduces all workgroups' candidates in a tree of
//   128-row groups to the final w pivot rows. The final elimination also gives L11\U11 of
//   the pivot rows. That workgroup writes the LAPACK pivots (sequential interchanges), the
//   net row moves, L11\U11 and the original values of rows k..k+w-1 to the workspace.
// calu_apply_kernel: rows k..k+w-1 of the panel receive L11\U11; every other row r
//   (after the interchanges) receives L(r) = A(r) U11^-1.
// No workgroup ever waits for another one: the panel costs two launches and no per-column
// exchange between workgroups.
#pragma once

#include "lrn_getrf_kernels.hpp"

namespace lrn {

constexpr int kCaluThreads = 256;
constexpr int kCaluMaxP = 128;  // workgroups of one select launch
constexpr int kCaluLeafQ = 4;   // rows per lane at the leaves: 256 rows per wave, 1024 per workgroup
constexpr int kCaluTreeQ = 2;   // rows per lane in the tree: groups of 128 candidate rows
constexpr int kCaluRowsPerWG = 4 * 64 * kCaluLeafQ;
constexpr int kCaluGroup = 64 * kCaluTreeQ;
static_assert(4 * 32 <= kCaluGroup, "a workgroup's four waves nominate at most 128 rows");

// Order-preserving key of a candidate: 0 = not a candidate, 1 = NaN, 2 + bits(|a|) otherwise.
__device__ __forceinline__ unsigned calu_key(float a)
{
    const float v = fabsf(a);
    return (v != v) ? 1u : 2u + __float_as_uint(v);
}

// Wave-wide maximum of an unsigned value (DPP within rows of 16 lanes, then the four rows).
__device__ __forceinline__ unsigned wave_max_u32(unsigned v)
{
    unsigned o;
    o = static_cast<unsigned>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0xB1, 0xF, 0xF, false));
    v = o > v ? o : v;
    o = static_cast<unsigned>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x4E, 0xF, 0xF, false));
    v = o > v ? o : v;
    o = static_cast<unsigned>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x141, 0xF, 0xF, false));
    v = o > v ? o : v;
    o = static_cast<unsigned>(__builtin_amdgcn_mov_dpp(static_cast<int>(v), 0x140, 0xF, 0xF, false));
    v = o > v ? o : v;
    const unsigned a = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 0));
    const unsigned b = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 16));
    const unsigned c = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 32));
    const unsigned d = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(v), 48));
    const unsigned ab = a > b ? a : b, cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}

template <int Q> __device__ __forceinline__ float pick_slot(const float (&x)[Q], int t)
{
    float r = x[0];
#pragma unroll
    for (int q = 1; q < Q; ++q)
        if (t == q) r = x[q];
    return r;
}
template <int Q> __device__ __forceinline__ int pick_slot(const int (&x)[Q], int t)
{
    int r = x[0];
#pragma unroll
    for (int q = 1; q < Q; ++q)
        if (t == q) r = x[q];
    return r;
}

// Partial-pivoting elimination of up to 64*Q rows of the panel columns [c0, c0 + w) by one
// wave; lane l holds rows l, l + 64, ... of the set, whose global indices come from rowof(i).
// sel[j] (written by lane 0) = global row chosen at step j, for j < min(w, nrows). With fac,
// fac[j * W + c] = the chosen row j after elimination (multipliers for c < j, U for c >= j).
// *zero_step = first step whose pivot is exactly zero (left unchanged if none).
// bcast: W floats of LDS private to the wave (the pivot row of the current step).
// kVec4: w == W and every row start 16-byte aligned (rows read as float4).
// Returns min(w, nrows).
template <int W, int Q, bool kVec4, typename RowOf>
__device__ int wave_gepp(const float* __restrict__ A, long long lda, long long c0, int w, int nrows, RowOf rowof,
    int* sel, float* fac, int* zero_step, float* bcast)
{
    const int lane = threadIdx.x & (kWave - 1);
    float v[Q][W];
    int gr[Q];
    int step_of[Q];
    unsigned avail = 0u;
#pragma unroll
    for (int t = 0; t < Q; ++t) {
        const int i = lane + kWave * t;
        const bool ok = i < nrows;
        gr[t] = ok ? rowof(i) : 0;
        step_of[t] = -1;
        if (ok) avail |= 1u << t;
        const float* src = A + static_cast<long long>(gr[t]) * lda + c0;
        if (kVec4) {
#pragma unroll
            for (int c = 0; c < W; c += 4) {
                float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
                if (ok) q4 = *reinterpret_cast<const float4*>(src + c);
                v[t][c] = q4.x;
                v[t][c + 1] = q4.y;
                v[t][c + 2] = q4.z;
                v[t][c + 3] = q4.w;
            }
        } else {
#pragma unroll
            for (int c = 0; c < W; ++c) v[t][c] = (ok && c < w) ? src[c] : 0.f;
        }
    }
    const int steps = min(w, nrows);
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (j < steps) {
            unsigned bk = 0u;
            int bt = 0;
#pragma unroll
            for (int t = 0; t < Q; ++t) {
                const unsigned kk = ((avail >> t) & 1u) ? calu_key(v[t][j]) : 0u;
                if (kk > bk) {
                    bk = kk;
                    bt = t;
                }
            }
            const unsigned mk = wave_max_u32(bk);
            const unsigned long long hits = __ballot(bk == mk);
            const int L = __ffsll(static_cast<long long>(hits)) - 1;
            const int ts = __builtin_amdgcn_readlane(bt, L);
            // the owner lane writes the pivot row to LDS; every lane reads it back
            if (lane == L) {
#pragma unroll
                for (int t = 0; t < Q; ++t)
                    if (t == ts) {
#pragma unroll
                        for (int c = j; c < W; ++c) bcast[c] = v[t][c];
                        avail &= ~(1u << t);
                        step_of[t] = j;
                    }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            float prow[W];
#pragma unroll
            for (int c = j; c < W; ++c) prow[c] = bcast[c];
            const int grow = __builtin_amdgcn_readlane(pick_slot<Q>(gr, ts), L);
            if (lane == 0) sel[j] = grow;
            const float d = prow[j];
            if (d == 0.f) {
                if (zero_step != nullptr && lane == 0 && *zero_step < 0) *zero_step = j;
            } else {
                const bool use_recip = fabsf(d) >= FLT_MIN;
                const float rcp = use_recip ? 1.f / d : 1.f;
#pragma unroll
                for (int t = 0; t < Q; ++t) {
                    if ((avail >> t) & 1u) {
                        const float l = use_recip ? v[t][j] * rcp : v[t][j] / d;
                        v[t][j] = l;
#pragma unroll
                        for (int c = j + 1; c < W; ++c) v[t][c] = __builtin_fmaf(-l, prow[c], v[t][c]);
                    }
                }
            }
            // every lane has read the pivot row before the next step overwrites it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (fac != nullptr) {
#pragma unroll
        for (int t = 0; t < Q; ++t)
            if (step_of[t] >= 0) {
#pragma unroll
                for (int c = 0; c < W; ++c) fac[step_of[t] * W + c] = v[t][c];
            }
    }
    return steps;
}

// U11^-1 of the factored pivot block (fac, W x W in LDS, U on and above the diagonal) by one
// wave: lane c computes column c by back substitution. Columns and rows >= w are zero.
template <int W>
__device__ void calu_u_inverse(const float* fac, int w, float* uinv)
{
    const int c = threadIdx.x & (kWave - 1);
    float x[W];
#pragma unroll
    for (int i = W - 1; i >= 0; --i) {
        float r = 0.f;
        if (c < w && i <= c) {
            float sum = (i == c) ? 1.f : 0.f;
#pragma unroll
            for (int p = i + 1; p < W; ++p)
                if (p <= c) sum = __builtin_fmaf(-fac[i * W + p], x[p], sum);
            r = sum / fac[i * W + i];
        }
        x[i] = r;
    }
    if (c < W) {
#pragma unroll
        for (int i = 0; i < W; ++i) uinv[i * W + c] = x[i];
    }
}

template <typename T> struct CaluWs {
    int* cand;      // [kCaluMaxP][W] candidates of each workgroup (global rows)
    unsigned* ctr;  // arrival counter, monotonic over one factorisation
    T* lu11;        // [W][W] L11\U11 of the pivot rows
    T* top;         // [W][W] original values of rows k..k+w-1, panel columns
    T* uinv;        // [W][W] U11^-1
    int* info;      // first exactly-zero pivot, 1-based
    int* rowmap;    // [1 + 4W] net row moves of the panel
};

// Sequential-interchange pivots (LAPACK ipiv positions, 0-based) from the pivot rows chosen
// in order: position k+j is swapped with the current position of row win[j]. Lane l tracks
// the content of position k+l; positions at or beyond k+w that receive a displaced row are
// tracked as extras. Called by one whole wave; pivs written by lane 0.
__device__ inline void calu_positions(const int* win, long long k, int w, long long* pivs)
{
    const int lane = threadIdx.x & (kWave - 1);
    long long c = k + lane;      // content of position k+lane
    long long e = -1, ec = -1;   // extra position of this lane and its content
    int ne = 0;
    for (int j = 0; j < w; ++j) {
        const long long g = win[j];
        const unsigned long long mt = __ballot(lane < w && c == g);
        const unsigned long long me = __ballot(lane < ne && ec == g);
        const long long cj = __shfl(c, j, kWave);
        long long p;
        if (mt) {
            const int lt = __ffsll(static_cast<long long>(mt)) - 1;
            p = k + lt;
            if (lane == lt && lt != j) c = cj;
        } else if (me) {
            const int lx = __ffsll(static_cast<long long>(me)) - 1;
            p = __shfl(e, lx, kWave);
            if (lane == lx) ec = cj;
        } else {
            p = g;
            if (lane == ne) {
                e = g;
                ec = cj;
            }
            ++ne;
        }
        if (lane == j) c = g;
        if (lane == 0) pivs[j] = p;
    }
}

// The last stage of the tournament, by wave 0 of one workgroup: elimination on the final
// candidates (cnt <= 128 rows), then the LAPACK pivots, the net row moves, info, L11\U11,
// U11^-1 and the original top rows.
template <int W, bool kVec4>
__device__ void calu_finish(const float* __restrict__ A, long long lda, long long k, int w, int cnt, const int* list,
    long long* __restrict__ piv, const CaluWs<float>& ws, float* fac, int* win, long long* pivs, int* s_zero,
    float* bcast)
{
    const int tid = threadIdx.x;
    wave_gepp<W, kCaluTreeQ, kVec4>(A, lda, k, w, cnt, [&](int i) { return list[i]; }, win, fac, s_zero, bcast);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    calu_positions(win, k, w, pivs);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    panel_rowmap(pivs, k, w, 2 * W, ws.rowmap);
    if (tid < w) piv[k + tid] = pivs[tid] + 1;
    if (tid == 0 && *s_zero >= 0 && *ws.info == 0) *ws.info = static_cast<int>(k + *s_zero + 1);
    calu_u_inverse<W>(fac, w, ws.uinv);
    for (int i = tid; i < W * W; i += kWave) {
        const int r = i / W, c = i - r * W;
        ws.lu11[i] = fac[i];
        ws.top[i] = (r < w && c < w) ? A[(k + r) * lda + k + c] : 0.f;
    }
}

template <int W, bool kVec4>
__global__ __launch_bounds__(kCaluThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void calu_select_kernel(const float* __restrict__ A, long long lda,
    long long k, long long m, int w, int P, unsigned base, long long* __restrict__ piv, CaluWs<float> ws)
{
    __shared__ int wsel[4][W];
    __shared__ int list_a[kCaluMaxP * W];
    __shared__ int list_b[kCaluMaxP * W];
    __shared__ float fac[W * W];
    __shared__ float bcast[4][W];
    __shared__ long long pivs[W];
    __shared__ int win[W];
    __shared__ int wcnt[4];
    __shared__ int s_last;
    __shared__ int s_zero;

    const int tid = threadIdx.x;
    const int wid = tid / kWave;
    const int b = blockIdx.x;
    const long long wg_row0 = k + static_cast<long long>(b) * kCaluRowsPerWG;
    const long long n = k + m;
    constexpr int RW = kWave * kCaluLeafQ;
    if (tid == 0) s_zero = -1;

    // leaves: each wave over its own rows
    const long long r0 = wg_row0 + static_cast<long long>(wid) * RW;
    const int nr = static_cast<int>(max(0ll, min(static_cast<long long>(RW), n - r0)));
    int nw = 0;
    if (nr > 0)
        nw = wave_gepp<W, kCaluLeafQ, kVec4>(A, lda, k, w, nr, [&](int i) { return static_cast<int>(r0 + i); },
            wsel[wid], nullptr, nullptr, bcast[wid]);
    if ((tid & (kWave - 1)) == 0) wcnt[wid] = nw;
    __syncthreads();

    // workgroup level: wave 0 over the four waves' candidates; with a single workgroup this
    // is the final stage
    if (wid == 0) {
        int tot = 0;
        for (int q = 0; q < 4; ++q) {
            for (int i = tid; i < wcnt[q]; i += kWave) list_a[tot + i] = wsel[q][i];
            tot += wcnt[q];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (P == 1)
            calu_finish<W, kVec4>(A, lda, k, w, tot, list_a, piv, ws, fac, win, pivs, &s_zero, bcast[0]);
        else
            wave_gepp<W, kCaluTreeQ, kVec4>(A, lda, k, w, tot, [&](int i) { return list_a[i]; }, &ws.cand[b * W],
                nullptr, nullptr, bcast[0]);
    }
    if (P == 1) return;
    __syncthreads();
    if (tid == 0) {
        __threadfence();
        const unsigned old = atomicAdd(ws.ctr, 1u);
        s_last = (old == base + static_cast<unsigned>(P) - 1u) ? 1 : 0;
        __threadfence();
    }
    __syncthreads();
    if (!s_last) return;

    // last workgroup: all candidates; every workgroup nominated min(w, its rows)
    const long long rows_last = n - (k + static_cast<long long>(P - 1) * kCaluRowsPerWG);
    int cnt = (P - 1) * w + static_cast<int>(min(static_cast<long long>(w), rows_last));
    for (int i = tid; i < cnt; i += kCaluThreads) {
        const int bb = i / w, jj = i - bb * w;
        list_a[i] = __hip_atomic_load(&ws.cand[bb * W + jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    int* cur = list_a;
    int* nxt = list_b;
    while (cnt > kCaluGroup) {
        const int groups = (cnt + kCaluGroup - 1) / kCaluGroup;
        for (int g = wid; g < groups; g += 4) {
            const int gn = min(kCaluGroup, cnt - g * kCaluGroup);
            const int* gl = cur + g * kCaluGroup;
            wave_gepp<W, kCaluTreeQ, kVec4>(A, lda, k, w, gn, [&](int i) { return gl[i]; }, nxt + g * w, nullptr,
                nullptr, bcast[wid]);
        }
        __syncthreads();
        cnt = (groups - 1) * w + min(w, cnt - (groups - 1) * kCaluGroup);
        int* t = cur;
        cur = nxt;
        nxt = t;
    }
    if (wid == 0) calu_finish<W, kVec4>(A, lda, k, w, cnt, cur, piv, ws, fac, win, pivs, &s_zero, bcast[0]);
}

// Rows k..k+w-1 of the panel receive L11\U11; every other row r (after the interchanges)
// receives L(r) = A(r) U11^-1, each entry an independent dot product with the precomputed
// inverse (uniform across the rows: scalar loads). One thread per row.
template <int W, bool kVec4>
__global__ __launch_bounds__(256) void calu_apply_kernel(float* __restrict__ A, long long lda, long long k, long long m,
    int w, const float* __restrict__ lu11, const float* __restrict__ top, const float* __restrict__ uinv,
    const int* __restrict__ rowmap)
{
    __shared__ int rdst[2 * W], rsrc[2 * W];
    __shared__ int rcnt;
    const int tid = threadIdx.x;
    if (tid == 0) rcnt = rowmap[0];
    if (tid < 2 * W) {
        rdst[tid] = rowmap[1 + tid];
        rsrc[tid] = rowmap[1 + 2 * W + tid];
    }
    __syncthreads();
    const long long i = k + static_cast<long long>(blockIdx.x) * 256 + tid;
    if (i >= k + m) return;
    float* dst = A + i * lda + k;
    if (i < k + w) {
        const float* f = lu11 + (i - k) * W;
        for (int c = 0; c < w; ++c) dst[c] = f[c];
        return;
    }
    const float* src = dst;
    const int cnt = rcnt;
    for (int q = 0; q < cnt; ++q)
        if (rdst[q] == static_cast<int>(i)) src = top + static_cast<long long>(rsrc[q] - k) * W;
    float a[W];
    if (kVec4) {
#pragma unroll
        for (int c = 0; c < W; c += 4) {
            const float4 q4 = *reinterpret_cast<const float4*>(src + c);
            a[c] = q4.x;
            a[c + 1] = q4.y;
            a[c + 2] = q4.z;
            a[c + 3] = q4.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < W; ++c) a[c] = (c < w) ? src[c] : 0.f;
    }
    float x[W];
#pragma unroll
    for (int c = 0; c < W; ++c) {
        float sum = 0.f;
#pragma unroll
        for (int p = 0; p <= c; ++p) sum = __builtin_fmaf(a[p], uinv[p * W + c], sum);
        x[c] = sum;
    }
    if (kVec4) {
#pragma unroll
        for (int c = 0; c < W; c += 4) *reinterpret_cast<float4*>(dst + c) = make_float4(x[c], x[c + 1], x[c + 2], x[c + 3]);
    } else {
#pragma unroll
        for (int c = 0; c < W; ++c)
            if (c < w) dst[c] = x[c];
    }
}

}  // namespace lrn
