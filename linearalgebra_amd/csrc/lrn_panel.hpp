// Cooperative panel factorisation with one memory round trip per column.
//
// P workgroups each hold R rows of the m x w panel in LDS. For every column each workgroup
// publishes its candidate pivot row (its local argmax row, all w columns) as self-validating
// 16-byte chunks {value, row index, tag}; workgroup 0 also publishes the current diagonal
// row. Every workgroup then reads all P candidate rows and the diagonal row in one batch of
// agent-coherent 16-byte loads, repeating the batch (bounded) until every chunk carries the
// column's tag. With the data and its readiness in the same 16-byte access there is no
// separate flag, counter, fence or second round trip.
//
// Pivot rule, interchange and elimination order are those of
// src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232.
#pragma once

#include "lrn_getrf_kernels.hpp"

namespace lrn {

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

constexpr int kPanel2MaxWG = 64;  // workgroups of one launch
constexpr int kPanel2CPT = 9;     // chunks read per thread per column: ceil((64 * 32 + 32) / 256)

struct alignas(16) Chunk {
    double v;
    int idx;
    unsigned tag;
};

template <typename T> struct Panel2Ws {
    Chunk* slots;   // [2][kPanel2MaxWG][W]; zero before the first launch of a factorisation
    Chunk* dslots;  // [2][W]
    int* info;
    int* err;
    int* rowmap;
};

// Agent-coherent 16-byte store (the same cache policy the compiler uses for agent-scope
// atomic stores: sc1). One lane writes one chunk in one instruction.
__device__ __forceinline__ void st_chunk(Chunk* p, double v, int idx, unsigned tag)
{
    const unsigned long long bits = static_cast<unsigned long long>(__double_as_longlong(v));
    u32x4 d;
    d.x = static_cast<unsigned>(bits);
    d.y = static_cast<unsigned>(bits >> 32);
    d.z = static_cast<unsigned>(idx);
    d.w = tag;
    asm volatile("global_store_dwordx4 %0, %1, off sc1" : : "v"(p), "v"(d) : "memory");
}

// Nine agent-coherent 16-byte loads in flight at once; the wait is inside the same asm
// statement, so no output register is live while its load is outstanding.
__device__ __forceinline__ void ld_chunks9(u32x4 (&r)[kPanel2CPT], const Chunk* const (&p)[kPanel2CPT])
{
    asm volatile(
        "global_load_dwordx4 %0, %9, off sc1\n\t"
        "global_load_dwordx4 %1, %10, off sc1\n\t"
        "global_load_dwordx4 %2, %11, off sc1\n\t"
        "global_load_dwordx4 %3, %12, off sc1\n\t"
        "global_load_dwordx4 %4, %13, off sc1\n\t"
        "global_load_dwordx4 %5, %14, off sc1\n\t"
        "global_load_dwordx4 %6, %15, off sc1\n\t"
        "global_load_dwordx4 %7, %16, off sc1\n\t"
        "global_load_dwordx4 %8, %17, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]), "=&v"(r[7]),
          "=&v"(r[8])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]), "v"(p[8])
        : "memory");
}

__device__ __forceinline__ double chunk_value(const u32x4& c)
{
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(c.y) << 32) | c.x));
}

template <typename T, int W> __host__ __device__ constexpr size_t panel2_tile_bytes(int R)
{
    return (static_cast<size_t>(R) * (W + 1) * sizeof(T) + 15) / 16 * 16;
}
template <typename T, int W> __host__ __device__ constexpr size_t panel2_lds_bytes(int R, int P)
{
    return panel2_tile_bytes<T, W>(R) + static_cast<size_t>(P) * W * sizeof(double);
}

template <typename T, int W>
__global__ __launch_bounds__(kPanelThreads) void getf2_panel2_kernel(T* __restrict__ A, long long lda, long long k,
    long long m, int w, int R, int P, unsigned tag_base, unsigned spin_limit, long long* __restrict__ piv,
    Panel2Ws<T> ws)
{
    static_assert(kPanel2MaxWG * W + W <= kPanel2CPT * kPanelThreads, "chunk budget");
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int LD = W + 1;
    T* tile = reinterpret_cast<T*>(lds_raw);
    double* cand = reinterpret_cast<double*>(lds_raw + panel2_tile_bytes<T, W>(R));  // [P][W]
    __shared__ int cand_idx[kPanel2MaxWG];
    __shared__ double drow[W];
    __shared__ double red_v[kPanelThreads / kWave];
    __shared__ long long red_i[kPanelThreads / kWave];
    __shared__ long long pivs[W];
    __shared__ int win_b;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int b = blockIdx.x;
    const long long row0 = k + static_cast<long long>(b) * R;
    const long long n = k + m;
    const int rows_here = static_cast<int>(min(static_cast<long long>(R), n - row0));

    for (int idx = tid; idx < rows_here * w; idx += kPanelThreads) {
        const int r = idx / w;
        const int c = idx - r * w;
        tile[r * LD + c] = A[(row0 + r) * lda + k + c];
    }
    __syncthreads();

    const int tpr = (R >= kPanelThreads) ? 1 : (kPanelThreads / R);
    const int rstep = kPanelThreads / tpr;
    const int total_chunks = P * w + w;

    for (int j = 0; j < w; ++j) {
        const long long i = k + j;
        const int par = j & 1;
        const unsigned tag = tag_base + static_cast<unsigned>(j) + 1u;

        // 1. local argmax
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
        bv = red_v[0];
        bi = red_i[0];
#pragma unroll
        for (int q = 1; q < kPanelThreads / kWave; ++q)
            if (argmax_better(red_v[q], red_i[q], bv, bi)) {
                bv = red_v[q];
                bi = red_i[q];
            }

        long long p;
        const double* prow;
        if (P > 1) {
            // 2. publish the candidate row; workgroup 0 also publishes the diagonal row
            const int bl = static_cast<int>(bi - row0);
            if (tid < w)
                st_chunk(&ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + b) * W + tid],
                    static_cast<double>(tile[bl * LD + tid]), static_cast<int>(bi), tag);
            if (b == 0 && tid >= kWave && tid - kWave < w)
                st_chunk(&ws.dslots[par * W + (tid - kWave)], static_cast<double>(tile[j * LD + (tid - kWave)]),
                    static_cast<int>(i), tag);
            // 3. gather every candidate row and the diagonal row, repeating until all tags match
            const Chunk* addr[kPanel2CPT];
#pragma unroll
            for (int s = 0; s < kPanel2CPT; ++s) {
                const int q = tid + s * kPanelThreads;
                const Chunk* a = ws.slots;
                if (q < P * w)
                    a = &ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + q / w) * W + (q % w)];
                else if (q < total_chunks)
                    a = &ws.dslots[par * W + (q - P * w)];
                addr[s] = a;
            }
            u32x4 r[kPanel2CPT];
            unsigned spins = 0;
            for (;;) {
                ld_chunks9(r, addr);
                bool ok = true;
#pragma unroll
                for (int s = 0; s < kPanel2CPT; ++s)
                    if (tid + s * kPanelThreads < total_chunks && r[s].w != tag) ok = false;
                if (ok) break;
                ++spins;
                if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(ws.err))) {
                    __hip_atomic_store(ws.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
#pragma unroll
            for (int s = 0; s < kPanel2CPT; ++s) {
                const int q = tid + s * kPanelThreads;
                if (q < P * w) {
                    const int bb = q / w, c = q % w;
                    cand[bb * W + c] = chunk_value(r[s]);
                    if (c == 0) cand_idx[bb] = static_cast<int>(r[s].z);
                } else if (q < total_chunks) {
                    drow[q - P * w] = chunk_value(r[s]);
                }
            }
            __syncthreads();
            // 4. global winner over the P candidates
            double gv = -2.0;
            long long gi = 0x7fffffffffffffffll;
            if (tid < P) {
                gi = cand_idx[tid];
                gv = pivot_key(cand[tid * W + j], gi == i);
            }
            wave_argmax(gv, gi);
            if (lane == 0) {
                red_v[wid] = gv;
                red_i[wid] = gi;
            }
            __syncthreads();
            gv = red_v[0];
            gi = red_i[0];
#pragma unroll
            for (int q = 1; q < kPanelThreads / kWave; ++q)
                if (argmax_better(red_v[q], red_i[q], gv, gi)) {
                    gv = red_v[q];
                    gi = red_i[q];
                }
            p = gi;
            if (tid == 0) win_b = static_cast<int>((p - k) / R);
            __syncthreads();
            prow = cand + win_b * W;
        } else {
            p = bi;
            if (tid < w) {
                cand[tid] = static_cast<double>(tile[static_cast<int>(p - row0) * LD + tid]);
                drow[tid] = static_cast<double>(tile[j * LD + tid]);
            }
            __syncthreads();
            prow = cand;
        }

        // 5. interchange rows i and p inside the panel columns
        if (p != i) {
            if (b == 0 && tid < w) tile[j * LD + tid] = static_cast<T>(prow[tid]);
            const long long pl = p - row0;
            if (pl >= 0 && pl < rows_here && tid < w) tile[static_cast<int>(pl) * LD + tid] = static_cast<T>(drow[tid]);
        }
        if (b == 0 && tid == 0) {
            piv[i] = p + 1;
            pivs[j] = p;
        }
        __syncthreads();

        // 6. multipliers and rank-1 update below the diagonal
        const T d = static_cast<T>(prow[j]);
        if (d == T(0)) {
            if (b == 0 && tid == 0 && *ws.info == 0) *ws.info = static_cast<int>(i + 1);
        } else {
            const bool use_recip = fabs(static_cast<double>(d)) >= static_cast<double>(TypeMin<T>::v);
            const T rcp = use_recip ? T(1) / d : T(1);
            for (int rr = tid / tpr; rr < rows_here; rr += rstep) {
                if (row0 + rr <= i) continue;
                T* rowp = tile + rr * LD;
                const T l = use_recip ? rowp[j] * rcp : rowp[j] / d;
                for (int c = j + 1 + (tid % tpr); c < w; c += tpr) rowp[c] = rowp[c] - l * static_cast<T>(prow[c]);
                if ((tid % tpr) == 0) rowp[j] = l;
            }
        }
        __syncthreads();
    }

    for (int idx = tid; idx < rows_here * w; idx += kPanelThreads) {
        const int r = idx / w;
        const int c = idx - r * w;
        A[(row0 + r) * lda + k + c] = tile[r * LD + c];
    }
    if (b == 0 && tid < kWave) panel_rowmap(pivs, k, w, 2 * W, ws.rowmap);
}

}  // namespace lrn
