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
inline int panel_cpt(int P) { return P < 8 ? 1 : P < 16 ? 2 : P < 24 ? 3 : P < 32 ? 4 : P < 40 ? 5 : 9; }

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
    // s_nop: a store of more than 8 bytes needs one wait state before a VALU instruction may
    // overwrite its data registers, which the compiler does not insert after inline asm
    asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 1" : : "v"(p), "v"(d) : "memory");
}

// N agent-coherent 16-byte loads in flight at once; the wait is inside the same asm
// statement, so no output register is live while its load is outstanding.
template <int N> __device__ __forceinline__ void ld_chunks(u32x4 (&r)[N], const Chunk* const (&p)[N]);
template <> __device__ __forceinline__ void ld_chunks<1>(u32x4 (&r)[1], const Chunk* const (&p)[1])
{
    asm volatile(
        "global_load_dwordx4 %0, %1, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0])
        : "v"(p[0])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<2>(u32x4 (&r)[2], const Chunk* const (&p)[2])
{
    asm volatile(
        "global_load_dwordx4 %0, %2, off sc1\n\t"
        "global_load_dwordx4 %1, %3, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1])
        : "v"(p[0]), "v"(p[1])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<3>(u32x4 (&r)[3], const Chunk* const (&p)[3])
{
    asm volatile(
        "global_load_dwordx4 %0, %3, off sc1\n\t"
        "global_load_dwordx4 %1, %4, off sc1\n\t"
        "global_load_dwordx4 %2, %5, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2])
        : "v"(p[0]), "v"(p[1]), "v"(p[2])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<4>(u32x4 (&r)[4], const Chunk* const (&p)[4])
{
    asm volatile(
        "global_load_dwordx4 %0, %4, off sc1\n\t"
        "global_load_dwordx4 %1, %5, off sc1\n\t"
        "global_load_dwordx4 %2, %6, off sc1\n\t"
        "global_load_dwordx4 %3, %7, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<5>(u32x4 (&r)[5], const Chunk* const (&p)[5])
{
    asm volatile(
        "global_load_dwordx4 %0, %5, off sc1\n\t"
        "global_load_dwordx4 %1, %6, off sc1\n\t"
        "global_load_dwordx4 %2, %7, off sc1\n\t"
        "global_load_dwordx4 %3, %8, off sc1\n\t"
        "global_load_dwordx4 %4, %9, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<8>(u32x4 (&r)[8], const Chunk* const (&p)[8])
{
    asm volatile(
        "global_load_dwordx4 %0, %8, off sc1\n\t"
        "global_load_dwordx4 %1, %9, off sc1\n\t"
        "global_load_dwordx4 %2, %10, off sc1\n\t"
        "global_load_dwordx4 %3, %11, off sc1\n\t"
        "global_load_dwordx4 %4, %12, off sc1\n\t"
        "global_load_dwordx4 %5, %13, off sc1\n\t"
        "global_load_dwordx4 %6, %14, off sc1\n\t"
        "global_load_dwordx4 %7, %15, off sc1\n\t"
        "s_waitcnt vmcnt(0)"
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]), "=&v"(r[7])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7])
        : "memory");
}
template <> __device__ __forceinline__ void ld_chunks<9>(u32x4 (&r)[9], const Chunk* const (&p)[9])
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
        : "=&v"(r[0]), "=&v"(r[1]), "=&v"(r[2]), "=&v"(r[3]), "=&v"(r[4]), "=&v"(r[5]), "=&v"(r[6]), "=&v"(r[7]), "=&v"(r[8])
        : "v"(p[0]), "v"(p[1]), "v"(p[2]), "v"(p[3]), "v"(p[4]), "v"(p[5]), "v"(p[6]), "v"(p[7]), "v"(p[8])
        : "memory");
}

// Bounded poll of N chunks until all carry `tag` (raises *err and returns on expiry).
template <int N>
__device__ __forceinline__ void spin_chunks(u32x4 (&r)[N], const Chunk* const (&p)[N], unsigned tag, unsigned spin_limit,
    int* err)
{
    if (spin_limit == 0u) {
#pragma unroll
        for (int s = 0; s < N; ++s) r[s] = u32x4{0u, 0u, 0u, 0u};
        raise_error(err);
        return;
    }
    unsigned spins = 0;
    for (;;) {
        ld_chunks<N>(r, p);
        bool ok = true;
#pragma unroll
        for (int s = 0; s < N; ++s)
            if (r[s].w != tag) ok = false;
        if (ok) return;
        ++spins;
        if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(err))) {
            raise_error(err);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void spin_chunk1(u32x4& r, const Chunk* p, unsigned tag, unsigned spin_limit, int* err)
{
    if (spin_limit == 0u) {
        r = u32x4{0u, 0u, 0u, 0u};
        raise_error(err);
        return;
    }
    unsigned spins = 0;
    for (;;) {
        const Chunk* const pc[1] = {p};
        u32x4 rr[1];
        ld_chunks<1>(rr, pc);
        r = rr[0];
        if (r.w == tag) return;
        ++spins;
        if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(err))) {
            raise_error(err);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
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

// Version with the trailing part of each elimination deferred. Column j+1 is brought up to
// date right after column j's multipliers are known; the rest of column j's rank-1 update
// (columns j+2..) is applied during step j+1 after this workgroup has published its
// candidate, so it overlaps the time the other workgroups' chunks take to arrive. The
// candidate row and the diagonal row receive that deferred update before they are
// published. Every element still receives its updates one by one in column order, each
// product rounded before the subtraction, so the result is that of the plain order.
// CPT = chunks each thread reads per column (>= ceil((P*W + W) / 256)).
template <typename T, int W, int CPT>
__global__ __launch_bounds__(kPanelThreads) void getf2_panel3_kernel(T* __restrict__ A, long long lda, long long k,
    long long m, int w, int R, int P, unsigned tag_base, unsigned spin_limit, long long* __restrict__ piv,
    Panel2Ws<T> ws)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    constexpr int LD = W + 1;
    T* tile = reinterpret_cast<T*>(lds_raw);
    double* cand = reinterpret_cast<double*>(lds_raw + panel2_tile_bytes<T, W>(R));  // [P][W]
    __shared__ int cand_idx[kPanel2MaxWG];
    __shared__ double drow[W];
    __shared__ double red_v[kPanelThreads / kWave];
    __shared__ long long red_i[kPanelThreads / kWave];
    __shared__ long long pivs[W];
    __shared__ long long win_row;
    __shared__ int win_b;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int b = blockIdx.x;
    const long long row0 = k + static_cast<long long>(b) * R;
    const long long n = k + m;
    const int rows_here = static_cast<int>(min(static_cast<long long>(R), n - row0));
    const int total_chunks = P * w + w;

    for (int idx = tid; idx < rows_here * w; idx += kPanelThreads) {
        const int r = idx / w;
        const int c = idx - r * w;
        tile[r * LD + c] = A[(row0 + r) * lda + k + c];
    }

    // rows of the argmax scan: contiguous per thread, so lane order is row order
    const int rpt = (R + kPanelThreads - 1) / kPanelThreads;
    const int r_lo = tid * rpt;
    const int r_hi = min(r_lo + rpt, rows_here);
    // elimination mapping: tpr threads per row
    const int tpr = (R >= kPanelThreads) ? 1 : (kPanelThreads / R);
    const int rstep = kPanelThreads / tpr;

    // chunk addresses are fixed for the launch except for the parity
    const Chunk* addr0[CPT];
    const Chunk* addr1[CPT];
#pragma unroll
    for (int s = 0; s < CPT; ++s) {
        int q = tid + s * kPanelThreads;
        if (q >= total_chunks) q = q % total_chunks;  // spare loads spread over real chunks
        if (q < P * w) {
            addr0[s] = &ws.slots[(0 * kPanel2MaxWG + q / w) * W + (q % w)];
            addr1[s] = &ws.slots[(1 * kPanel2MaxWG + q / w) * W + (q % w)];
        } else {
            addr0[s] = &ws.dslots[0 * W + (q - P * w)];
            addr1[s] = &ws.dslots[1 * W + (q - P * w)];
        }
    }
    __syncthreads();

    __shared__ int prev_zero;  // column j-1 had an exactly-zero pivot: it has no deferred update
    if (tid == 0) prev_zero = 1;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const long long i = k + j;
        const int par = j & 1;
        const unsigned tag = tag_base + static_cast<unsigned>(j) + 1u;

        // (i) local argmax over column j (current)
        double bv = -2.0;
        long long bi = 0x7fffffffffffffffll;
        for (int r = r_lo; r < r_hi; ++r) {
            const long long g = row0 + r;
            if (g >= i) {
                const double v = pivot_key(tile[r * LD + j], g == i);
                if (v > bv) {
                    bv = v;
                    bi = g;
                }
            }
        }
        // lanes hold consecutive row ranges, so the lowest lane with the maximum holds the
        // lowest such row: DPP maximum + ballot instead of a shuffle butterfly
        wave_argmax_ordered(bv, bi);
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
        const int bl = static_cast<int>(bi - row0);  // local index of the candidate row

        const bool deferred = !prev_zero;  // j == 0 starts with prev_zero set
        // (ii) deferred update of column j-1 on the candidate row and the diagonal row
        if (deferred) {
            const double* pu = cand + win_b * W;  // previous pivot row (still in LDS)
            const int jm = j - 1;
            if (tid < w && tid > j) {
                T* rc = tile + bl * LD;
                rc[tid] = rc[tid] - rc[jm] * static_cast<T>(pu[tid]);
            }
            if (b == 0 && tid >= kWave && tid - kWave < w && tid - kWave > j && bl != j) {
                const int c = tid - kWave;
                T* rd = tile + j * LD;
                rd[c] = rd[c] - rd[jm] * static_cast<T>(pu[c]);
            }
        }
        __syncthreads();

        long long p;
        if (P > 1) {
            // (iii) publish
            if (tid < w)
                st_chunk(&ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + b) * W + tid],
                    static_cast<double>(tile[bl * LD + tid]), static_cast<int>(bi), tag);
            if (b == 0 && tid >= kWave && tid - kWave < w)
                st_chunk(&ws.dslots[par * W + (tid - kWave)], static_cast<double>(tile[j * LD + (tid - kWave)]),
                    static_cast<int>(i), tag);
        }

        // (iv) deferred update of column j-1 on every other row, columns > j
        if (deferred) {
            const double* pu = cand + win_b * W;
            const int jm = j - 1;
            for (int rr = tid / tpr; rr < rows_here; rr += rstep) {
                if (row0 + rr < i || rr == bl || (b == 0 && rr == j)) continue;
                T* rowp = tile + rr * LD;
                const T l = rowp[jm];
                for (int c = j + 1 + (tid % tpr); c < w; c += tpr) rowp[c] = rowp[c] - l * static_cast<T>(pu[c]);
            }
        }

        if (P > 1) {
            // (v) gather all candidates and the diagonal row
            u32x4 r[CPT];
            unsigned spins = 0;
            if (spin_limit == 0u) raise_error(ws.err);  // forced expiry (tests)
            for (;;) {
                if (par)
                    ld_chunks<CPT>(r, addr1);
                else
                    ld_chunks<CPT>(r, addr0);
                bool ok = true;
#pragma unroll
                for (int s = 0; s < CPT; ++s)
                    if (r[s].w != tag) ok = false;
                if (ok) break;
                ++spins;
                if (spins > spin_limit || ((spins & 255u) == 0u && error_raised(ws.err))) {
                    raise_error(ws.err);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __syncthreads();  // every thread is done with cand / win_b of column j-1
#pragma unroll
            for (int s = 0; s < CPT; ++s) {
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
            // (vi) winner over the P candidates (one wave)
            if (wid == 0) {
                double gv = -2.0;
                long long gi = 0x7fffffffffffffffll;
                if (lane < P) {
                    gi = cand_idx[lane];
                    gv = pivot_key(cand[lane * W + j], gi == i);
                }
                wave_argmax_ordered(gv, gi);  // lane b holds workgroup b's row range
                if (lane == 0) {
                    if (gi < i || gi >= k + m) {
                        // only an expired wait can get here: report it, keep the rows in place
                        raise_error(ws.err);
                        gi = i;
                    }
                    win_row = gi;
                    win_b = static_cast<int>((gi - k) / R);
                }
            }
            __syncthreads();
            p = win_row;
        } else {
            __syncthreads();
            if (tid < w) {
                cand[tid] = static_cast<double>(tile[bl * LD + tid]);
                drow[tid] = static_cast<double>(tile[j * LD + tid]);
            }
            if (tid == 0) win_b = 0;
            __syncthreads();
            p = bi;
        }
        const double* prow = cand + win_b * W;

        // (vii) interchange rows i and p inside the panel columns
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

        // (viii) multipliers of column j and the update of column j+1 only
        const T d = static_cast<T>(prow[j]);
        if (tid == 0) prev_zero = (d == T(0)) ? 1 : 0;
        if (d == T(0)) {
            // exactly-zero pivot: the column is skipped (no multipliers, no update)
            if (b == 0 && tid == 0 && *ws.info == 0) *ws.info = static_cast<int>(i + 1);
        } else {
            const bool use_recip = fabs(static_cast<double>(d)) >= static_cast<double>(TypeMin<T>::v);
            const T rcp = use_recip ? T(1) / d : T(1);
            const T u1 = (j + 1 < w) ? static_cast<T>(prow[j + 1]) : T(0);
            for (int rr = tid; rr < rows_here; rr += kPanelThreads) {
                if (row0 + rr <= i) continue;
                T* rowp = tile + rr * LD;
                const T l = use_recip ? rowp[j] * rcp : rowp[j] / d;
                rowp[j] = l;
                if (j + 1 < w) rowp[j + 1] = rowp[j + 1] - l * u1;
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

// Register-resident version. Thread t of workgroup b holds RPT consecutive rows of the panel
// (rows k + b*R + t*RPT + q, R = 256*RPT) with all W columns in registers, so lane order is
// row order for the pivot search and the rank-1 update needs no LDS traffic. Per column:
//   1. argmax over the thread's rows, DPP wave argmax, one barrier to combine the four waves;
//   2. the thread holding the workgroup's candidate (and, in workgroup 0, the thread holding
//      the diagonal row) copies the row to LDS, and 32 lanes of its wave publish it as tagged
//      chunks with one store each;
//   3. phase A: thread t reads workgroup t's candidate entry of the current column (one chunk,
//      repeated until its tag matches); one barrier; every wave picks the winner itself;
//   4. phase B: every workgroup reads the winner's row, and the winner's workgroup (which
//      holds row p) the diagonal row; one barrier;
//   5. the threads holding rows i and p exchange them; every thread forms its multipliers and
//      updates its rows.
// Three barriers per column and 96 chunks read per workgroup. The pivot rule, the
// interchange, the multiplier (x * (1/d) unless |d| < T_MIN, then x / d) and the update order
// (each product rounded before it is subtracted, columns updated left to right) are those of
// src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232.
// With P == 1 the candidate and diagonal rows go through LDS instead of memory.

// State of one thread of getf2_panel4_kernel; step<J>() is column J, with J a compile-time
// constant so every access to the row registers is static.
// Copies row q of the thread's registers to an LDS row (as doubles).
template <typename T, int W, int RPT>
__device__ __forceinline__ void stage_row(const T (&v)[RPT][W], int q, double* dst)
{
#pragma unroll
    for (int qq = 0; qq < RPT; ++qq)
        if (qq == q) {
#pragma unroll
            for (int c = 0; c < W; ++c) dst[c] = static_cast<double>(v[qq][c]);
        }
}

template <typename T, int W, int RPT> struct Panel4 {
    T v[RPT][W];
    double (*stage)[2][W];
    T* A;
    long long lda, k, n, my0, row0;
    int w, P, tid, lane, wid, b;
    unsigned tag_base, spin_limit;
    long long* piv;
    Panel2Ws<T> ws;
    double* red_v;
    long long* red_i;
    double (*cand)[kPanel2MaxWG][W];
    int (*cand_idx)[kPanel2MaxWG];
    double (*drow)[W];
    long long* pivs;

    template <int J> __device__ __forceinline__ void step()
    {
        constexpr int R = kPanelThreads * RPT;
        const long long i = k + J;
        constexpr int par = J & 1;
        const unsigned tag = tag_base + static_cast<unsigned>(J) + 1u;

        // 1. workgroup candidate
        double bv = -2.0;
        long long bi = 0x7fffffffffffffffll;
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const long long g = my0 + q;
            if (g >= i && g < n) {
                const double key = pivot_key(v[q][J], g == i);
                if (key > bv) {
                    bv = key;
                    bi = g;
                }
            }
        }
        wave_argmax_ordered(bv, bi);
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

        // 2. publish (P > 1) or hand over through LDS (P == 1): the owner lane copies the row to
        //    its wave's staging area, then lane c of that wave handles column c
        const bool own_cand = bv >= -1.0 && bi >= my0 && bi < my0 + RPT;
        const bool own_diag = i >= my0 && i < my0 + RPT;
        const bool wave_cand = bv >= -1.0 && bi >= row0 && ((bi - row0) / RPT) / kWave == wid;
        const bool wave_diag = i >= row0 && i < row0 + R && ((i - row0) / RPT) / kWave == wid;
        if (own_cand) stage_row(v, static_cast<int>(bi - my0), stage[wid][0]);
        if (own_diag) stage_row(v, static_cast<int>(i - my0), stage[wid][1]);
        if (wave_cand || wave_diag) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (lane < w) {
                if (wave_cand) {
                    if (P > 1)
                        st_chunk(&ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + b) * W + lane], stage[wid][0][lane],
                            static_cast<int>(bi), tag);
                    else
                        cand[par][0][lane] = stage[wid][0][lane];
                }
                if (wave_diag) {
                    if (P > 1)
                        st_chunk(&ws.dslots[par * W + lane], stage[wid][1][lane], static_cast<int>(i), tag);
                    else
                        drow[par][lane] = stage[wid][1][lane];
                }
            }
            if (P == 1 && wave_cand && lane == 0) cand_idx[par][0] = static_cast<int>(bi);
        }

        // 3. phase A: every workgroup's candidate entry of column J (thread t reads workgroup t's)
        if (P > 1) {
            if (tid < P) {
                u32x4 r;
                spin_chunk1(r, &ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + tid) * W + J], tag, spin_limit,
                    ws.err);
                cand[par][tid][J] = chunk_value(r);
                cand_idx[par][tid] = static_cast<int>(r.z);
            }
            __syncthreads();
        } else {
            __syncthreads();
        }

        // 4. winner (each wave on its own)
        double gv = -2.0;
        long long gi = 0x7fffffffffffffffll;
        if (lane < P) {
            gi = cand_idx[par][lane];
            gv = pivot_key(static_cast<T>(cand[par][lane][J]), gi == i);
        }
        wave_argmax_ordered(gv, gi);  // lane b holds workgroup b's row range
        long long p = gi;
        if (p < i || p >= n) {
            // only a failed exchange can get here: report it and leave every row in place
            if (tid == 0) __hip_atomic_store(ws.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            p = i;
        }
        const int wb = static_cast<int>(min(max((p - k) / R, 0ll), static_cast<long long>(P - 1)));

        // 5. phase B: the winner's row (every workgroup) and the diagonal row (the winner's
        //    workgroup, which holds row p)
        if (P > 1) {
            if (tid < w) {
                u32x4 r;
                spin_chunk1(r, &ws.slots[(static_cast<size_t>(par) * kPanel2MaxWG + wb) * W + tid], tag, spin_limit,
                    ws.err);
                cand[par][wb][tid] = chunk_value(r);
            } else if (b == wb && tid >= kWave && tid - kWave < w) {
                u32x4 r;
                spin_chunk1(r, &ws.dslots[par * W + (tid - kWave)], tag, spin_limit, ws.err);
                drow[par][tid - kWave] = chunk_value(r);
            }
            __syncthreads();
        }
        const double* prow = cand[par][wb];

        // 6. interchange, multipliers, rank-1 update
        if (p != i) {
            if (own_diag) {
                const int qd = static_cast<int>(i - my0);
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (q == qd) {
#pragma unroll
                        for (int c = 0; c < W; ++c)
                            if (c < w) v[q][c] = static_cast<T>(prow[c]);
                    }
            }
            if (p >= my0 && p < my0 + RPT) {
                const int qp = static_cast<int>(p - my0);
#pragma unroll
                for (int q = 0; q < RPT; ++q)
                    if (q == qp) {
#pragma unroll
                        for (int c = 0; c < W; ++c)
                            if (c < w) v[q][c] = static_cast<T>(drow[par][c]);
                    }
            }
        }
        if (b == 0 && tid == 0) {
            piv[i] = p + 1;
            pivs[J] = p;
        }
        const T d = static_cast<T>(prow[J]);
        if (d == T(0)) {
            if (b == 0 && tid == 0 && *ws.info == 0) *ws.info = static_cast<int>(i + 1);
        } else {
            const bool use_recip = fabs(static_cast<double>(d)) >= static_cast<double>(TypeMin<T>::v);
            const T rcp = use_recip ? T(1) / d : T(1);
            T l[RPT];
#pragma unroll
            for (int q = 0; q < RPT; ++q) {
                const long long g = my0 + q;
                l[q] = T(0);
                if (g > i && g < n) {
                    l[q] = use_recip ? v[q][J] * rcp : v[q][J] / d;
                    v[q][J] = l[q];
                }
            }
            // columns in groups of 8, so only eight pivot-row values are live at a time
#pragma unroll
            for (int c0 = (J + 1) / 8 * 8; c0 < W; c0 += 8) {
#pragma unroll
                for (int c = c0; c < c0 + 8; ++c) {
                    if (c > J) {
                        const T u = static_cast<T>(prow[c]);
#pragma unroll
                        for (int q = 0; q < RPT; ++q)
                            if (my0 + q > i && my0 + q < n) v[q][c] = v[q][c] - l[q] * u;
                    }
                }
                asm volatile("" ::: "memory");
            }
        }
    }

    template <int J> __device__ __forceinline__ void run()
    {
        if constexpr (J < W) {
            if (J < w) {
                step<J>();
                run<J + 1>();
            }
        }
    }
};

// Register-resident kernel: see the comment above stage_row / Panel4.
template <typename T, int W, int RPT>
__global__ __launch_bounds__(kPanelThreads) void getf2_panel4_kernel(T* __restrict__ A, long long lda, long long k,
    long long m, int w, int P, unsigned tag_base, unsigned spin_limit, long long* __restrict__ piv, Panel2Ws<T> ws)
{
    constexpr int R = kPanelThreads * RPT;
    __shared__ double red_v[kPanelThreads / kWave];
    __shared__ long long red_i[kPanelThreads / kWave];
    __shared__ double cand[2][kPanel2MaxWG][W];
    __shared__ int cand_idx[2][kPanel2MaxWG];
    __shared__ double drow[2][W];
    __shared__ long long pivs[W];
    __shared__ double stage[kPanelThreads / kWave][2][W];

    Panel4<T, W, RPT> st;
    st.A = A;
    st.lda = lda;
    st.k = k;
    st.n = k + m;
    st.w = w;
    st.P = P;
    st.tid = threadIdx.x;
    st.lane = st.tid & (kWave - 1);
    st.wid = st.tid / kWave;
    st.b = blockIdx.x;
    st.row0 = k + static_cast<long long>(st.b) * R;
    st.my0 = st.row0 + static_cast<long long>(st.tid) * RPT;
    st.tag_base = tag_base;
    st.spin_limit = spin_limit;
    st.piv = piv;
    st.ws = ws;
    st.red_v = red_v;
    st.red_i = red_i;
    st.cand = cand;
    st.cand_idx = cand_idx;
    st.drow = drow;
    st.pivs = pivs;
    st.stage = stage;

#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const long long g = st.my0 + q;
        const T* src = A + (g < st.n ? g : k) * lda + k;
#pragma unroll
        for (int c = 0; c < W; ++c) st.v[q][c] = (g < st.n && c < w) ? src[c] : T(0);
    }

    st.template run<0>();

#pragma unroll
    for (int q = 0; q < RPT; ++q) {
        const long long g = st.my0 + q;
        if (g < st.n) {
            T* dst = A + g * lda + k;
#pragma unroll
            for (int c = 0; c < W; ++c)
                if (c < w) dst[c] = st.v[q][c];
        }
    }
    if (st.b == 0) {
        __syncthreads();
        if (st.tid < kWave) panel_rowmap(pivs, k, w, 2 * W, ws.rowmap);
    }
}

}  // namespace lrn

