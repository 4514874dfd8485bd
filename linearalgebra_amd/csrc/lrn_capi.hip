// Host driver and C ABI of the gfx950 LU solve path.
//
// Factorisation (lu_decomposition): right-looking, blocked twice.
//   outer block NB (default 256) -> the trailing update A22 -= L21 * U12 is one large MFMA GEMM
//   inner panel W (<= 64)        -> cooperative getf2 panel, then laswp / TRSM / small GEMMs
//                                   that finish the NB-wide block column and its U12 rows
// Solve (linear_solver): permutation from the pivots, blocked forward and backward
// substitution in two persistent kernels.
//
// Reference behaviour kept (paths relative to the reference repository):
//   argument checks              src/matrix_operation/cpu/native/lu_decomposition_native.cpp:30-35,
//                                src/matrix_operation/cpu/native/linear_solver_native.cpp:24-28
//   info = first exact zero pivot, factorisation continues   lu_decomposition_native.cpp:147-157,200-208
//   solve skipped on a singular matrix in LU mode            linear_solver_native.cpp:160-171
//   synchronous lu_decomposition, asynchronous solver         docs/gpu_backend_design.md:64-99
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <memory>
#include <type_traits>
#include <vector>

#include "../../include/linalg_rocm_native.h"
#include "lrn_getrf_kernels.hpp"
#include "lrn_getrs_kernels.hpp"
#include "lrn_panel.hpp"
#include "lrn_gemm_f64.hpp"
#include "lrn_gemm_f32.hpp"
#include "lrn_gemm_bf16.hpp"
#include "lrn_mixed.hpp"
#include "lrn_trace.hpp"
#include "lrn_batched_launch.hpp"
#include "lrn_calu.hpp"
#include "lrn_inverse.hpp"
#include "lrn_cholesky.hpp"
#include "lrn_bmm.hpp"
#include "lrn_trsm_launch.hpp"
#include "lrn_fp8_launch.hpp"

namespace lrn {
namespace {

constexpr int kMaxW = 64;
constexpr int kMaxInner = 16;  // inner panels per outer block column (row-map slots)
constexpr int kRowmapSets = 3;  // block columns whose row maps are live at once
// Polls allowed to every bounded inter-workgroup wait (lrn_debug_set_spin_limit; 0 makes every
// wait expire at once, which tests use to force the error paths).
std::atomic<unsigned> g_spin_limit{1u << 20};
unsigned spin_limit() { return g_spin_limit.load(std::memory_order_relaxed); }

std::atomic<int> g_outer_nb{0};
std::atomic<int> g_panel_w{0};
std::atomic<int> g_panel_max_wg{0};

int tuned_outer_nb(long long n)
{
    const int v = g_outer_nb.load();
    if (v > 0) return v;
    return n >= 2048 ? 256 : 128;
}
int tuned_panel_w()
{
    const int v = g_panel_w.load();
    return (v > 0) ? std::min(v, kMaxW) : 32;
}
int tuned_panel_max_wg()
{
    const int v = g_panel_max_wg.load();
    return (v > 0) ? std::min(v, kPanelMaxWG) : 128;
}

// Single-pass carve of a caller buffer, shared by the size query and the real call, in the
// manner of the reference's workspace_arena (include/common/workspace_arena.h:35-73) but
// with 256-byte alignment for device memory.
struct Arena {
    unsigned char* base;
    size_t cap;
    size_t off = 0;
    bool ok = true;
    explicit Arena(void* b = nullptr, size_t c = 0) : base(static_cast<unsigned char*>(b)), cap(c) {}
    template <typename T> T* next(size_t count)
    {
        const size_t a = (off + 255) / 256 * 256;
        off = a + sizeof(T) * count;
        if (base == nullptr) return nullptr;
        if (off > cap) {
            ok = false;
            return nullptr;
        }
        return reinterpret_cast<T*>(base + a);
    }
};

template <typename T> struct GetrfLayout {
    PanelWs<T> ws;
};

template <typename T> GetrfLayout<T> carve_getrf(Arena& ar)
{
    GetrfLayout<T> L;
    L.ws.ctr = ar.next<unsigned>(64);
    L.ws.info = ar.next<int>(64);
    L.ws.err = ar.next<int>(64);
    L.ws.sval = ar.next<double>(2 * kPanelMaxWG);
    L.ws.sidx = ar.next<long long>(2 * kPanelMaxWG);
    L.ws.srow = ar.next<T>(static_cast<size_t>(2) * kPanelMaxWG * kMaxW);
    L.ws.sdiag = ar.next<T>(2 * kMaxW);
    // three sets of row maps: with the look-ahead block column K+2 is factored while block
    // column K's maps may still be applied to the far columns
    L.ws.rowmap = ar.next<int>(static_cast<size_t>(kRowmapSets) * kMaxInner * (1 + 4 * kMaxW));
    L.ws.chunks = ar.next<Chunk>(static_cast<size_t>(2) * kPanel2MaxWG * 32);
    L.ws.dchunks = ar.next<Chunk>(2 * 32);
    return L;
}

template <typename T> struct GetrsLayout {
    int* perm;
    int* perm_tmp;
    unsigned* flags_f;
    unsigned* flags_b;
    int* err;
    T* y;
    Chunk* xch;  // single right-hand side: tagged exchange of solved values, one per row
    int* maps;   // net row moves of every 64-pivot block (permutation from the pivots)
    int* gate;   // nonzero: skip the solve (singular factors, or an expired factorisation wait)
    T* dinv;     // two or more right-hand sides: inverses of the 64 x 64 diagonal blocks of L, then of U
};

long long solve_blocks(long long n) { return (n + kSolveTB - 1) / kSolveTB; }

template <typename T> GetrsLayout<T> carve_getrs(Arena& ar, long long n, long long nrhs)
{
    GetrsLayout<T> L;
    L.perm = ar.next<int>(static_cast<size_t>(n));
    L.perm_tmp = ar.next<int>(n > kPermLdsMax ? static_cast<size_t>(n) : 1);
    L.flags_f = ar.next<unsigned>(static_cast<size_t>(solve_blocks(n)));
    L.flags_b = ar.next<unsigned>(static_cast<size_t>(solve_blocks(n)));
    L.err = ar.next<int>(64);
    L.y = ar.next<T>(static_cast<size_t>(n) * static_cast<size_t>(std::min<long long>(nrhs, kSolveMaxCols)));
    L.xch = ar.next<Chunk>(static_cast<size_t>(n));
    L.maps = ar.next<int>(static_cast<size_t>((n + 63) / 64) * kPermMapInts);
    L.gate = ar.next<int>(64);
    L.dinv = nrhs >= 2 ? ar.next<T>(static_cast<size_t>(2 * trsm_dinv_elems(n))) : nullptr;
    return L;
}

struct DeviceProps {
    int cus = 0;
    int max_lds = 0;
};

DeviceProps& device_props(int dev)
{
    static std::mutex mu;
    static std::vector<DeviceProps> props;
    std::lock_guard<std::mutex> g(mu);
    if (dev >= static_cast<int>(props.size())) props.resize(dev + 1);
    DeviceProps& p = props[dev];
    if (p.cus == 0) {
        if (hipDeviceGetAttribute(&p.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || p.cus <= 0)
            p.cus = 1;
        if (hipDeviceGetAttribute(&p.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            p.max_lds = 65536;
    }
    return p;
}

#define LRN_HIP(expr)                                   \
    do {                                                \
        if ((expr) != hipSuccess) return LRN_E_HIP;     \
    } while (0)

#define LRN_TRY(expr)                         \
    do {                                      \
        const int _st = (expr);               \
        if (_st != LRN_OK) return _st;        \
    } while (0)

int next_pow2(long long v)
{
    int r = 1;
    while (r < v) r <<= 1;
    return r;
}

// ------------------------------------------------------------------------------- launches

// One-time per (kernel, device) set-up: raise the dynamic-LDS limit to the device maximum
// and cache the occupancy for a given LDS size, so launches make no extra runtime calls.
int prepare_kernel(const void* f, size_t lds, int threads, int* per_cu)
{
    struct Entry {
        const void* f;
        size_t lds;
        int threads;
        int dev;
        int per_cu;
    };
    static std::mutex mu;
    static std::vector<Entry> cache;
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    for (const Entry& e : cache)
        if (e.f == f && e.lds >= lds && e.threads == threads && e.dev == dev && e.lds == lds) {
            *per_cu = e.per_cu;
            return LRN_OK;
        }
    hipFuncAttributes fa;
    LRN_HIP(hipFuncGetAttributes(&fa, f));
    const int max_dyn = device_props(dev).max_lds - static_cast<int>(fa.sharedSizeBytes);
    if (static_cast<int>(lds) > max_dyn) return LRN_E_UNSUPPORTED;
    if (lds > 0) LRN_HIP(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    int occ = 0;
    LRN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, threads, lds));
    cache.push_back(Entry{f, lds, threads, dev, occ});
    *per_cu = occ;
    return LRN_OK;
}

template <typename T> size_t panel_lds_bytes(int R, int W) { return static_cast<size_t>(R) * (W + 1) * sizeof(T); }

// Panel kernel of the 32-column panels: 0 = register-resident rows (getf2_panel4_kernel, the
// default), 1 = rows in LDS with deferred updates (getf2_panel3_kernel).
std::atomic<int> g_panel_kind{0};

int g_panel2_target_wg()
{
    const int v = g_panel_max_wg.load();
    return (v > 0) ? std::min(v, kPanel2MaxWG) : 32;
}

// flops of the rank-1 updates and scalings of an m x w panel
double panel_flops(long long m, int w)
{
    double f = 0.0;
    for (int j = 0; j < w && j < m; ++j) f += static_cast<double>(m - j - 1) * (1.0 + 2.0 * (w - j - 1));
    return f;
}

template <typename T, int W>
int launch_panel(T* A, long long lda, long long k, long long m, int w, long long* piv, const PanelWs<T>& ws,
    unsigned& base, hipStream_t s, int* need_cus = nullptr)
{
    // need_cus != nullptr: launch nothing, return the compute units the panel's cooperative
    // workgroups need to be resident at once (workgroups / workgroups per unit)
    auto need = [&](int P, int per_cu) {
        *need_cus = (P > 1 && per_cu > 0) ? (P + per_cu - 1) / per_cu : 1;
        return LRN_OK;
    };
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    const DeviceProps& props = device_props(dev);
    if constexpr (W == 32) {
        // register-resident cooperative panel (lrn_panel.hpp, getf2_panel4_kernel): one row per
        // thread, so up to 64 x 256 = 16384 rows
        const long long P4 = (m + kPanelThreads - 1) / kPanelThreads;
        if (g_panel_kind.load() == 0 && P4 <= kPanel2MaxWG) {
            const int P = static_cast<int>(P4);
            Panel2Ws<T> pw;
            pw.slots = static_cast<Chunk*>(ws.chunks);
            pw.dslots = static_cast<Chunk*>(ws.dchunks);
            pw.info = ws.info;
            pw.err = ws.err;
            pw.rowmap = ws.rowmap;
            auto kern = getf2_panel4_kernel<T, W, 1>;
            int per_cu = 0;
            LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(kern), 0, kPanelThreads, &per_cu));
            if (need_cus != nullptr) return need(P, per_cu);
            if (P > 1 && (per_cu <= 0 || static_cast<long long>(per_cu) * props.cus < P)) return LRN_E_UNSUPPORTED;
            TraceScope ts(kTrPanel, panel_flops(m, w), s);
            hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(P)), dim3(kPanelThreads), 0, s, A, lda, k, m, w, P, base,
                spin_limit(), piv, pw);
            LRN_HIP(hipGetLastError());
            base += static_cast<unsigned>(w);
            return LRN_OK;
        }
        // one-round-trip cooperative panel with the rows in LDS (getf2_panel3_kernel)
        int R, P;
        if (m <= 256) {
            R = std::max(next_pow2(m), 32);
            P = 1;
        } else {
            const int target = g_panel2_target_wg();
            R = next_pow2((m + target - 1) / target);
            R = std::max(R, 64);
            if (R > 512) R = 512;
            P = static_cast<int>((m + R - 1) / R);
        }
        if (P <= kPanel2MaxWG) {
            const size_t lds = panel2_lds_bytes<T, W>(R, std::max(P, 1));
            Panel2Ws<T> pw;
            pw.slots = static_cast<Chunk*>(ws.chunks);
            pw.dslots = static_cast<Chunk*>(ws.dchunks);
            pw.info = ws.info;
            pw.err = ws.err;
            pw.rowmap = ws.rowmap;
            auto go = [&](auto kern) -> int {
                int per_cu = 0;
                LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(kern), lds, kPanelThreads, &per_cu));
                if (need_cus != nullptr) return need(P, per_cu);
                if (P > 1 && (per_cu <= 0 || static_cast<long long>(per_cu) * props.cus < P)) return LRN_E_UNSUPPORTED;
                TraceScope ts(kTrPanel, panel_flops(m, w), s);
                hipLaunchKernelGGL(kern, dim3(P), dim3(kPanelThreads), lds, s, A, lda, k, m, w, R, P, base, spin_limit(),
                    piv, pw);
                LRN_HIP(hipGetLastError());
                return LRN_OK;
            };
            switch (panel_cpt(P)) {
            case 1: LRN_TRY(go(getf2_panel3_kernel<T, W, 1>)); break;
            case 2: LRN_TRY(go(getf2_panel3_kernel<T, W, 2>)); break;
            case 3: LRN_TRY(go(getf2_panel3_kernel<T, W, 3>)); break;
            case 4: LRN_TRY(go(getf2_panel3_kernel<T, W, 4>)); break;
            case 5: LRN_TRY(go(getf2_panel3_kernel<T, W, 5>)); break;
            default: LRN_TRY(go(getf2_panel3_kernel<T, W, 9>)); break;
            }
            base += static_cast<unsigned>(w);
            return LRN_OK;
        }
    }
    const int dev_max_wg = tuned_panel_max_wg();
    int R, P;
    if (m <= 256) {
        R = std::max(next_pow2(m), W < 32 ? 32 : W);
        P = 1;
    } else {
        R = next_pow2((m + dev_max_wg - 1) / dev_max_wg);
        R = std::max(R, std::max(64, W));
        const int rmax = (W <= 32) ? 512 : 256;
        if (R > rmax) R = rmax;
        P = static_cast<int>((m + R - 1) / R);
        if (P > kPanelMaxWG) return LRN_E_UNSUPPORTED;
    }
    const size_t lds = panel_lds_bytes<T>(R, W);
    auto kern = getf2_panel_kernel<T, W>;
    {
        int per_cu = 0;
        LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(kern), lds, kPanelThreads, &per_cu));
        if (need_cus != nullptr) return need(P, per_cu);
        if (P > 1 && (per_cu <= 0 || static_cast<long long>(per_cu) * props.cus < P)) return LRN_E_UNSUPPORTED;
    }
    TraceScope ts(kTrPanel, panel_flops(m, w), s);
    hipLaunchKernelGGL(kern, dim3(P), dim3(kPanelThreads), lds, s, A, lda, k, m, w, R, P, base, spin_limit(), piv, ws);
    LRN_HIP(hipGetLastError());
    if (P > 1) base += static_cast<unsigned>(P) * static_cast<unsigned>(w);
    return LRN_OK;
}

// Tournament-pivoted panel of the FP32 factorisation (lrn_calu.hpp): selection (one launch,
// the last workgroup to arrive picks the pivot rows) and application (one launch).
template <int W>
int launch_calu(float* A, long long lda, long long k, long long m, int w, long long* piv, CaluWs<float> cw,
    int* rowmap, unsigned& cbase, hipStream_t s)
{
    const long long P = (m + kCaluRowsPerWG - 1) / kCaluRowsPerWG;
    if (P > kCaluMaxP || w > W || w > m) return LRN_E_UNSUPPORTED;
    cw.rowmap = rowmap;
    TraceScope ts(kTrPanel, panel_flops(m, w), s);
    const bool vec4 = w == W && lda % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0;
    const dim3 sg(static_cast<unsigned>(P)), ag(static_cast<unsigned>((m + 255) / 256));
    if (vec4)
        hipLaunchKernelGGL((calu_select_kernel<W, true>), sg, dim3(kCaluThreads), 0, s, A, lda, k, m, w,
            static_cast<int>(P), cbase, piv, cw);
    else
        hipLaunchKernelGGL((calu_select_kernel<W, false>), sg, dim3(kCaluThreads), 0, s, A, lda, k, m, w,
            static_cast<int>(P), cbase, piv, cw);
    LRN_HIP(hipGetLastError());
    if (P > 1) cbase += static_cast<unsigned>(P);
    if (vec4)
        hipLaunchKernelGGL((calu_apply_kernel<W, true>), ag, dim3(256), 0, s, A, lda, k, m, w, cw.lu11, cw.top, cw.uinv,
            rowmap);
    else
        hipLaunchKernelGGL((calu_apply_kernel<W, false>), ag, dim3(256), 0, s, A, lda, k, m, w, cw.lu11, cw.top,
            cw.uinv, rowmap);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T>
int launch_laswp_multi(T* A, long long lda, const int* rowmaps, int nmaps, int map_stride, int W2, long long c0a,
    long long c1a, long long c0b, long long c1b, hipStream_t s)
{
    const long long total = (c1a - c0a) + (c1b - c0b);
    if (total <= 0 || nmaps <= 0) return LRN_OK;
    if (W2 <= 64 && nmaps <= 16 && nmaps * W2 <= kSwapBlkRows) {
        // one gather and one scatter for all the maps
        const long long blocks = (total + kSwapBlkCols - 1) / kSwapBlkCols;
        TraceScope ts(kTrLaswp, 2.0 * W2 * nmaps * static_cast<double>(total) * sizeof(T), s);
        hipLaunchKernelGGL(laswp_block_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, A, lda, rowmaps,
            nmaps, map_stride, W2, c0a, c1a, c0b, c1b);
        LRN_HIP(hipGetLastError());
        return LRN_OK;
    }
    const long long blocks = (total + kSwapCols - 1) / kSwapCols;
    const size_t lds = static_cast<size_t>(W2) * kSwapCols * sizeof(T);
    int per_cu = 0;
    LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(laswp_multi_kernel<T>), lds, 256, &per_cu));
    TraceScope ts(kTrLaswp, 2.0 * W2 * nmaps * static_cast<double>(total) * sizeof(T), s);
    hipLaunchKernelGGL(laswp_multi_kernel<T>, dim3(static_cast<unsigned>(blocks)), dim3(256), lds, s, A, lda, rowmaps,
        nmaps, map_stride, W2, c0a, c1a, c0b, c1b);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T, int W>
int launch_trsm(T* A, long long lda, long long k, int w, long long c0, long long c1, hipStream_t s)
{
    if (c1 <= c0) return LRN_OK;
    const long long blocks = (c1 - c0 + 255) / 256;
    TraceScope ts(kTrTrsm, (static_cast<double>(w) * w + 2.0 * w * static_cast<double>(c1 - c0)) * sizeof(T), s);
    hipLaunchKernelGGL((trsm_llnu_kernel<T, W>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, A, lda, k, w, c0,
        c1);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// Inner panel [k, k+w)'s interchanges on the columns [K, k) (moves only) and, with its U12
// rows, on the columns [rs, re).
template <typename T, int W>
int launch_laswp_trsm(T* A, long long lda, const int* rowmap, int W2, long long K, long long k, int w, long long rs,
    long long re, hipStream_t s)
{
    const long long na = (k - K + kSwapCols - 1) / kSwapCols;
    const long long nb = re > rs ? (re - rs + kSwapCols - 1) / kSwapCols : 0;
    if (na + nb <= 0) return LRN_OK;
    const double right = static_cast<double>(std::max(0ll, re - rs));
    TraceScope ts(kTrLaswp, (2.0 * W2 * ((k - K) + right) + static_cast<double>(w) * w + 2.0 * w * right) * sizeof(T), s);
    hipLaunchKernelGGL((laswp_trsm_kernel<T, W>), dim3(static_cast<unsigned>(na + nb)), dim3(256), 0, s, A, lda, rowmap,
        W2, K, k, w, rs, re, static_cast<int>(na));
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// Algorithmic flops of one trailing-update launch: 2 M N K, or in the lower-triangle mode
// (tri != kNoTriangle) those of the tiles it computes.
double gemm_flops(long long m, long long nn, long long K, long long tri, int bm, int bn)
{
    if (tri == kNoTriangle) return 2.0 * m * nn * K;
    double f = 0.0;
    for (long long r0 = 0; r0 < m; r0 += bm) {
        const long long rows = std::min<long long>(bm, m - r0);
        // columns c with c + tri <= r0 + bm - 1, in whole tiles
        const long long last = r0 + bm - 1 - tri;  // largest local column a computed tile may start at
        if (last < 0) continue;
        const long long ntiles = std::min<long long>((last / bn) + 1, (nn + bn - 1) / bn);
        const long long cols = std::min<long long>(ntiles * bn, nn);
        f += 2.0 * rows * cols * K;
    }
    return f;
}

// C[M x N] -= A[M x K] * B[K x N]. lower: C is a square block on the diagonal of a symmetric
// update of which only the lower triangle is read later; tiles entirely above the diagonal
// are not computed (a tile straddling it is computed whole).
template <typename T>
int launch_gemm_sub(const T* A, long long lda, const T* B, long long ldb, T* C, long long ldc, long long M,
    long long N, long long K, hipStream_t s, int cus = 0, const int* skip = nullptr, bool lower = false)
{
    if (M <= 0 || N <= 0 || K <= 0) return LRN_OK;
    if constexpr (std::is_same<T, double>::value) {
        // interior tiles on the unchecked kernel, the ragged right / bottom strips on the checked one
        const bool aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0) &&
                             (lda % 2 == 0) && (ldb % 2 == 0) && (K % kG64BK == 0);
        const long long Mi = aligned ? M / kG64BM * kG64BM : 0;
        const long long Ni = aligned ? N / kG64BN * kG64BN : 0;
        auto launch = [&](bool check, const double* a, const double* b, double* c, long long m, long long nn,
                          long long coff) -> int {
            if (m <= 0 || nn <= 0) return LRN_OK;
            const long long tri = lower ? coff : kNoTriangle;
            const long long tm = (m + kG64BM - 1) / kG64BM, tn = (nn + kG64BN - 1) / kG64BN;
            if (tm * tn > 0x3fffffffll) return LRN_E_UNSUPPORTED;
            const long long per_xcd = (tm * tn + kXcds - 1) / kXcds;
            // persistent grid: two workgroups per compute unit at most
            int dev = 0;
            LRN_HIP(hipGetDevice(&dev));
            const long long cap = 2ll * (cus > 0 ? cus : device_props(dev).cus) / kXcds * kXcds;
            const dim3 grid(static_cast<unsigned>(std::min(per_xcd * kXcds, cap)));
            TraceScope ts(check ? kTrGemmF64Edge : kTrGemmF64, gemm_flops(m, nn, K, tri, kG64BM, kG64BN), s);
            if (check)
                hipLaunchKernelGGL(gemm_sub_f64_kernel<true>, grid, dim3(512), 0, s, a, lda, b, ldb, c, ldc,
                    static_cast<int>(m), static_cast<int>(nn), static_cast<int>(K), static_cast<int>(tm),
                    static_cast<int>(tn), static_cast<int>(per_xcd), skip, tri);
            else
                hipLaunchKernelGGL(gemm_sub_f64_kernel<false>, grid, dim3(512), 0, s, a, lda, b, ldb, c, ldc,
                    static_cast<int>(m), static_cast<int>(nn), static_cast<int>(K), static_cast<int>(tm),
                    static_cast<int>(tn), static_cast<int>(per_xcd), skip, tri);
            LRN_HIP(hipGetLastError());
            return LRN_OK;
        };
        LRN_TRY(launch(false, A, B, C, Mi, Ni, 0));
        LRN_TRY(launch(true, A, B + Ni, C + Ni, M, N - Ni, Ni));                   // right strip, all rows
        LRN_TRY(launch(true, A + Mi * lda, B, C + Mi * ldc, M - Mi, Ni, -Mi));     // bottom strip
        return LRN_OK;
    } else {
        // FP32: the same interior / ragged-strip split on the double-buffered kernel
        const bool aligned = ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0) &&
                             (lda % 4 == 0) && (ldb % 4 == 0) && (K % kG32BK == 0);
        const long long Mi = aligned ? M / kG32BM * kG32BM : 0;
        const long long Ni = aligned ? N / kG32BN * kG32BN : 0;
        auto launch = [&](bool check, const float* a, const float* b, float* c, long long m, long long nn,
                          long long coff) -> int {
            if (m <= 0 || nn <= 0) return LRN_OK;
            const long long tri = lower ? coff : kNoTriangle;
            const long long tm = (m + kG32BM - 1) / kG32BM, tn = (nn + kG32BN - 1) / kG32BN;
            if (tm * tn > 0x3fffffffll) return LRN_E_UNSUPPORTED;
            const long long per_xcd = (tm * tn + kXcds - 1) / kXcds;
            int dev = 0;
            LRN_HIP(hipGetDevice(&dev));
            const long long cap = static_cast<long long>(kG32PerCu) * (cus > 0 ? cus : device_props(dev).cus) / kXcds * kXcds;
            const dim3 grid(static_cast<unsigned>(std::min(per_xcd * kXcds, cap)));
            TraceScope ts(kTrGemmF32, gemm_flops(m, nn, K, tri, kG32BM, kG32BN), s);
            if (check)
                hipLaunchKernelGGL(gemm_sub_f32_kernel<true>, grid, dim3(256), 0, s, a, lda, b, ldb, c, ldc,
                    static_cast<int>(m), static_cast<int>(nn), static_cast<int>(K), static_cast<int>(tm),
                    static_cast<int>(tn), static_cast<int>(per_xcd), skip, tri);
            else
                hipLaunchKernelGGL(gemm_sub_f32_kernel<false>, grid, dim3(256), 0, s, a, lda, b, ldb, c, ldc,
                    static_cast<int>(m), static_cast<int>(nn), static_cast<int>(K), static_cast<int>(tm),
                    static_cast<int>(tn), static_cast<int>(per_xcd), skip, tri);
            LRN_HIP(hipGetLastError());
            return LRN_OK;
        };
        LRN_TRY(launch(false, A, B, C, Mi, Ni, 0));
        LRN_TRY(launch(true, A, B + Ni, C + Ni, M, N - Ni, Ni));
        LRN_TRY(launch(true, A + Mi * lda, B, C + Mi * ldc, M - Mi, Ni, -Mi));
        return LRN_OK;
    }
}

// Trailing update of the low-precision factorisation: 0 = bf16 operands, 1 = two bf16 planes
// (three products), 2 = three bf16 planes (six products, FP32-level), 3 = FP32 MFMA (the
// default: FP32-level factors, and on gfx950 faster than the six-product form of this build),
// 4 / 5 / 6 = one / two / three block-scaled FP8 e4m3 planes (1 / 3 / 6 products).
constexpr int kMixedUpdateModes = 7;
std::atomic<int> g_mixed_update{3};

// Panel of the low-precision factorisation: 1 = tournament pivoting, 0 = partial pivoting
// with the cooperative panel of the plain path (the default: on gfx950 the tournament's
// selection kernel is not yet faster than the cooperative panel).
std::atomic<int> g_mixed_panel{0};

// C[M x N] -= A * B, FP32 data, with `planes` bf16 planes per operand (1..3)
int launch_gemm_sub_bf16(const float* A, long long lda, const float* B, long long ldb, float* C, long long ldc,
    long long M, long long N, long long K, hipStream_t s, int planes, int cus = 0)
{
    if (M <= 0 || N <= 0 || K <= 0) return LRN_OK;
    if (planes < 1 || planes > 3) return LRN_E_INTERNAL;
    const long long tm = (M + kBfBM - 1) / kBfBM, tn = (N + kBfBN - 1) / kBfBN;
    if (tm * tn > 0x3fffffffll) return LRN_E_UNSUPPORTED;
    const long long per_xcd = (tm * tn + 7) / 8;
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    const long long cap = 4ll * (cus > 0 ? cus : device_props(dev).cus) / kXcds * kXcds;
    const dim3 grid(static_cast<unsigned>(std::min(per_xcd * 8, cap)));
    const int kind = planes == 1 ? kTrGemmBf16 : (planes == 2 ? kTrGemmBf16x3 : kTrGemmBf16x6);
    TraceScope ts(kind, 2.0 * M * N * K, s);
    const int iM = static_cast<int>(M), iN = static_cast<int>(N), iK = static_cast<int>(K);
    const int itm = static_cast<int>(tm), itn = static_cast<int>(tn), ipx = static_cast<int>(per_xcd);
    if (planes == 1)
        hipLaunchKernelGGL(gemm_sub_bf16_kernel<1>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itm, itn, ipx);
    else if (planes == 2)
        hipLaunchKernelGGL(gemm_sub_bf16_kernel<2>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itm, itn, ipx);
    else
        hipLaunchKernelGGL(gemm_sub_bf16_kernel<3>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itm, itn, ipx);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// The large trailing update: FP64 / FP32 matrix-core GEMM, or, for the low-precision
// factorisation of the mixed solver (upd = 0, 1, 2), 1-3 bf16 planes with FP32 accumulation.
// upd < 0 or 3: the GEMM of the data type.
template <typename T>
int launch_trailing(const T* A, long long lda, const T* B, long long ldb, T* C, long long ldc, long long M, long long N,
    long long K, int upd, hipStream_t s, int cus = 0)
{
    if constexpr (std::is_same<T, float>::value) {
        if (upd >= 0 && upd <= 2) return launch_gemm_sub_bf16(A, lda, B, ldb, C, ldc, M, N, K, s, upd + 1, cus);
        if (upd >= 4 && upd <= 6) {
            if (M <= 0 || N <= 0 || K <= 0) return LRN_OK;
            const int kind = upd == 4 ? kTrGemmFp8 : (upd == 5 ? kTrGemmFp8x3 : kTrGemmFp8x6);
            TraceScope ts(kind, 2.0 * M * N * K, s);
            return gemm_sub_fp8_launch(A, lda, B, ldb, C, ldc, M, N, K, upd - 3, s, nullptr);
        }
    }
    (void)upd;
    return launch_gemm_sub<T>(A, lda, B, ldb, C, ldc, M, N, K, s, cus);
}

int launch_fill(unsigned* p, unsigned v, long long count, hipStream_t s)
{
    if (count <= 0) return LRN_OK;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s, p, v, count);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

unsigned fill_grid(long long count)
{
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>((count + 255) / 256, 4096)));
}

// x (n x nrhs, stride ldx) <- NaN when *err is raised (an expired substitution wait)
template <typename T> int launch_poison(const int* err, T* x, long long n, long long nrhs, long long ldx, hipStream_t s)
{
    hipLaunchKernelGGL(poison_on_error_kernel<T>, dim3(fill_grid(n * nrhs)), dim3(256), 0, s, err, x, n, nrhs, ldx);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// *gate <- singular or expired factorisation; x <- NaN on an expired wait
template <typename T>
int launch_gate(const int* info, const int* err, int* gate, T* x, long long n, long long nrhs, long long ldx,
    hipStream_t s)
{
    hipLaunchKernelGGL(solve_gate_kernel<T>, dim3(fill_grid(n * nrhs)), dim3(256), 0, s, info, err, gate, x, n, nrhs,
        ldx);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}


// Clears the barrier counter, status words and every chunk tag before a factorisation, so
// no tag left by an earlier call on the same workspace can match.
template <typename T> int reset_panel_state(const PanelWs<T>& ws, hipStream_t s)
{
    LRN_TRY(launch_fill(ws.ctr, 0u, 64, s));
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(ws.info), 0u, 64, s));
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(ws.err), 0u, 64, s));
    LRN_TRY(launch_fill(static_cast<unsigned*>(ws.chunks), 0u, 2ll * kPanel2MaxWG * 32 * 4, s));
    LRN_TRY(launch_fill(static_cast<unsigned*>(ws.dchunks), 0u, 2ll * 32 * 4, s));
    return LRN_OK;
}

// Per-device helper streams for the look-ahead. Every launch of a look-ahead factorisation
// goes to these streams (the caller's stream only hands over and takes back through events):
// the legacy NULL stream, which callers commonly pass, waits for all work on blocking streams
// before each of its launches, so work left on it would serialise the three streams.
//   narrow: the first block column and each inner panel's update of the columns beyond its
//           own block column, restricted to the GEMM stream's compute units (from the panel's
//           units its kernels slowed the panels beside them: mixed n = 16384 solve 114.9 ms
//           against 114.4 ms on the same box);
//   panel : the inner panels and their own block column's updates, restricted to `np`
//           compute units so the cooperative panel's workgroups always find room;
//   gemm  : the rest of each trailing update, restricted to the other compute units.
// Default priority. `mu` keeps one factorisation at a time enqueuing on a set, so the
// events below always refer to the same call.
struct AsideStream {
    int np = 0;
    hipStream_t narrow = nullptr;
    hipStream_t panel = nullptr;
    hipStream_t gemm = nullptr;
    hipEvent_t start = nullptr;      // recorded on the caller's stream at entry
    hipEvent_t ready = nullptr;      // recorded on the narrow stream: next block column updated
    hipEvent_t done = nullptr;       // recorded on the panel stream: next block column factored
    hipEvent_t rest_done = nullptr;  // recorded on the GEMM stream: rest of the trailing update done
    hipEvent_t next_ready = nullptr; // recorded on the GEMM stream: block column i+2 updated by block column i
    hipEvent_t step = nullptr;       // recorded on the panel stream after each inner panel
    hipEvent_t ext_done = nullptr;   // recorded on the narrow stream: next block column's inner updates done
    std::mutex mu;
};

// The helper streams and events live until the process ends; they are released by an exit
// handler registered after the HIP runtime has started (so it runs before the runtime's own
// teardown) rather than left to the runtime's teardown. The table itself is never destroyed.
std::mutex g_aside_mu;
std::vector<std::vector<std::unique_ptr<AsideStream>>>& aside_table()
{
    static auto* t = new std::vector<std::vector<std::unique_ptr<AsideStream>>>();
    return *t;
}

void release_aside_streams()
{
    std::lock_guard<std::mutex> g(g_aside_mu);
    for (auto& dev_list : aside_table())
        for (auto& a : dev_list) {
            (void)hipStreamSynchronize(a->narrow);
            (void)hipStreamSynchronize(a->panel);
            (void)hipStreamSynchronize(a->gemm);
            (void)hipEventDestroy(a->start);
            (void)hipEventDestroy(a->ready);
            (void)hipEventDestroy(a->done);
            (void)hipEventDestroy(a->rest_done);
            (void)hipEventDestroy(a->step);
            (void)hipEventDestroy(a->next_ready);
            (void)hipEventDestroy(a->ext_done);
            (void)hipStreamDestroy(a->narrow);
            (void)hipStreamDestroy(a->panel);
            (void)hipStreamDestroy(a->gemm);
        }
    aside_table().clear();
}

int aside_stream(int np, AsideStream** out)
{
    static std::mutex& mu = g_aside_mu;
    auto& per_dev = aside_table();
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    const int cus = device_props(dev).cus;
    if (np >= cus) return LRN_E_UNSUPPORTED;
    std::lock_guard<std::mutex> g(mu);
    if (dev >= static_cast<int>(per_dev.size())) per_dev.resize(dev + 1);
    for (auto& a : per_dev[dev])
        if (a->np == np) {
            *out = a.get();
            return LRN_OK;
        }
    auto a = std::make_unique<AsideStream>();
    a->np = np;
    const int words = (cus + 31) / 32;
    std::vector<uint32_t> mp(words, 0u), mg(words, 0u);
    for (int c = 0; c < cus; ++c) (c < np ? mp : mg)[c / 32] |= 1u << (c % 32);
    LRN_HIP(hipExtStreamCreateWithCUMask(&a->narrow, static_cast<uint32_t>(words), mg.data()));
    LRN_HIP(hipExtStreamCreateWithCUMask(&a->panel, static_cast<uint32_t>(words), mp.data()));
    LRN_HIP(hipExtStreamCreateWithCUMask(&a->gemm, static_cast<uint32_t>(words), mg.data()));
    LRN_HIP(hipEventCreateWithFlags(&a->start, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->ready, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->done, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->rest_done, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->step, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->next_ready, hipEventDisableTiming));
    LRN_HIP(hipEventCreateWithFlags(&a->ext_done, hipEventDisableTiming));
    *out = a.get();
    per_dev[dev].push_back(std::move(a));
    static const bool at_exit_registered = (std::atexit(release_aside_streams) == 0);
    (void)at_exit_registered;
    return LRN_OK;
}

std::atomic<int> g_lookahead{0};
// Update inside a block column after each inner panel: 0 = bc_update_kernel (64 x 64 vector
// tiles), 1 = the trailing-update GEMM (128 x 128 matrix-core tiles), 2 = gemm_sub_kernel with
// 64 x 64 matrix-core tiles.
std::atomic<int> g_block_update{0};

// A[k+w.., c0..c1) -= L21 U12 after inner panel [k, k+w): the update kind selected by
// g_block_update. Every kind computes each entry independently of the column range, so an
// update split over several ranges is bitwise one update of their union.
template <typename T, int W>
int launch_block_update(T* A, long long lda, long long n, long long k, int w, long long c0, long long c1, hipStream_t s)
{
    const long long rows = n - (k + w), cols = c1 - c0;
    if (rows <= 0 || cols <= 0) return LRN_OK;
    const int kind = g_block_update.load();
    if (kind == 1)
        return launch_gemm_sub<T>(A + (k + w) * lda + k, lda, A + k * lda + c0, lda, A + (k + w) * lda + c0, lda, rows,
            cols, w, s);
    TraceScope ts(kTrBcUpdate, 2.0 * rows * cols * w, s);
    if (kind == 2) {
        // matrix cores, 64 x 64 tiles, the whole K = w in one LDS stage
        const long long tm = (rows + 63) / 64, tn = (cols + 63) / 64;
        if (w <= 32)
            hipLaunchKernelGGL((gemm_sub_kernel<T, 64, 64, 32, true>), dim3(static_cast<unsigned>(tm * tn)), dim3(256), 0,
                s, A + (k + w) * lda + k, lda, A + k * lda + c0, lda, A + (k + w) * lda + c0, lda,
                static_cast<int>(rows), static_cast<int>(cols), w, static_cast<int>(tn));
        else
            hipLaunchKernelGGL((gemm_sub_kernel<T, 64, 64, 32>), dim3(static_cast<unsigned>(tm * tn)), dim3(256), 0, s,
                A + (k + w) * lda + k, lda, A + k * lda + c0, lda, A + (k + w) * lda + c0, lda, static_cast<int>(rows),
                static_cast<int>(cols), w, static_cast<int>(tn));
        LRN_HIP(hipGetLastError());
        return LRN_OK;
    }
    const dim3 grid(static_cast<unsigned>(std::max<long long>(1, (rows + kBcRows - 1) / kBcRows)),
        static_cast<unsigned>((cols + kBcCols - 1) / kBcCols));
    hipLaunchKernelGGL((bc_update_kernel<T, W>), grid, dim3(256), 0, s, A, lda, n, k, w, c0, c1);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// Where a block column's factorisation sends the update of the columns right of it.
//   sx == nullptr: every inner panel's interchanges (on the block column's earlier columns
//     too), U12 rows and update cover [k+w, KE) in one launch each, on s;
//   sx != nullptr: the part [KB, KE) (the next block column) goes to stream sx, inner panel by
//     inner panel: sx waits `step` (recorded on s after each panel) and runs that panel's
//     interchanges, U12 rows and update there while s continues with the next panel. The
//     panel's interchanges on the block column's earlier columns (their L) go to sx as well,
//     in the same launch: sx reads those columns for the earlier panels' updates, and s
//     never reads them again. At the end s waits `ext_done` (recorded on sx).
struct NextColumns {
    long long KE = 0;
    hipStream_t sx = nullptr;
    hipEvent_t step = nullptr;
    hipEvent_t ext_done = nullptr;
};

// Factors block column [K, K+B) (rows K..n-1) inner panel by inner panel: each inner panel's
// interchanges, U12 rows and A22 update are applied to the rest of the block column and to the
// columns [K+B, nc.KE) (the next block column, or none).
template <typename T, int W>
int factor_block_column(T* A, long long n, long long K, long long B, long long* piv, const PanelWs<T>& ws,
    int* rowmaps, unsigned& base, hipStream_t s, const CaluWs<float>* calu, unsigned& cbase, const NextColumns& nc)
{
    const long long lda = n;
    const long long KB = K + B;
    const long long KE = std::max(KB, nc.KE);
    const bool split = nc.sx != nullptr && KE > KB;
    const int rm_stride = 1 + 4 * W;
    int slot = 0;
    for (long long k = K; k < KB; k += W, ++slot) {
        const int w = static_cast<int>(std::min<long long>(W, KB - k));
        PanelWs<T> pw = ws;
        pw.rowmap = rowmaps + slot * rm_stride;
        if constexpr (std::is_same<T, float>::value && W == 32) {
            if (calu != nullptr)
                LRN_TRY(launch_calu<W>(A, lda, k, n - k, w, piv, *calu, pw.rowmap, cbase, s));
            else
                LRN_TRY((launch_panel<T, W>(A, lda, k, n - k, w, piv, pw, base, s)));
        } else {
            (void)calu;
            (void)cbase;
            LRN_TRY((launch_panel<T, W>(A, lda, k, n - k, w, piv, pw, base, s)));
        }
        if (!split) {
            LRN_TRY((launch_laswp_trsm<T, W>(A, lda, pw.rowmap, 2 * W, K, k, w, k + w, KE, s)));
            LRN_TRY((launch_block_update<T, W>(A, lda, n, k, w, k + w, KE, s)));
            continue;
        }
        LRN_HIP(hipEventRecord(nc.step, s));
        LRN_HIP(hipStreamWaitEvent(nc.sx, nc.step, 0));
        LRN_TRY((launch_laswp_trsm<T, W>(A, lda, pw.rowmap, 2 * W, K, k, w, KB, KE, nc.sx)));
        LRN_TRY((launch_block_update<T, W>(A, lda, n, k, w, KB, KE, nc.sx)));
        LRN_TRY((launch_laswp_trsm<T, W>(A, lda, pw.rowmap, 2 * W, k, k, w, k + w, KB, s)));
        LRN_TRY((launch_block_update<T, W>(A, lda, n, k, w, k + w, KB, s)));
    }
    if (split) {
        LRN_HIP(hipEventRecord(nc.ext_done, nc.sx));
        LRN_HIP(hipStreamWaitEvent(s, nc.ext_done, 0));
    }
    return LRN_OK;
}

// U12 rows of block column [K, K+B) for the columns [c0, c1): blocked forward substitution,
// one TRSM per inner panel and a GEMM for the rows below it inside the block.
template <typename T, int W>
int u12_columns(T* A, long long n, long long K, long long B, long long c0, long long c1, hipStream_t s)
{
    const long long lda = n;
    if (c1 <= c0) return LRN_OK;
    if (B > 256) {
        const long long KB = K + B;
        for (long long k = K; k < KB; k += W) {
            const int w = static_cast<int>(std::min<long long>(W, KB - k));
            LRN_TRY((launch_trsm<T, W>(A, lda, k, w, c0, c1, s)));
            LRN_TRY(launch_gemm_sub<T>(A + (k + w) * lda + k, lda, A + k * lda + c0, lda, A + (k + w) * lda + c0, lda,
                KB - (k + w), c1 - c0, w, s));
        }
        return LRN_OK;
    }
    // one launch for the whole block column
    // narrow column ranges (the next block column): 8 columns per workgroup for parallelism
    const bool narrow = c1 - c0 <= 256;
    const int tn = narrow ? kU12TNNarrow : u12_tn_wide<T>();
    const long long blocks = (c1 - c0 + tn - 1) / tn;
    TraceScope ts(kTrTrsm, (static_cast<double>(B) * B + 2.0 * B * static_cast<double>(c1 - c0)) * sizeof(T), s);
    if (narrow)
        hipLaunchKernelGGL((u12_block_kernel<T, kU12TNNarrow>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, A,
            lda, K, static_cast<int>(B), c0, c1);
    else
        hipLaunchKernelGGL((u12_block_kernel<T, u12_tn_wide<T>()>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s,
            A, lda, K, static_cast<int>(B), c0, c1);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// Single-stream order of the blocked factorisation. Block column i is factored together with
// the update of block column i+1 by its inner panels; then block column i+2 and the far
// columns [K_{i+3}, n) receive block column i's interchanges, U12 rows and update, and the
// columns left of it its interchanges.
template <typename T, int W>
int getrf_blocked_plain(T* A, long long n, long long* piv, const PanelWs<T>& ws, hipStream_t s, int upd,
    const CaluWs<float>* calu)
{
    const long long lda = n;
    const long long NB = std::min<long long>(std::max<long long>(tuned_outer_nb(n), W), static_cast<long long>(W) * kMaxInner);
    const int rm_stride = 1 + 4 * W;
    unsigned base = 0;
    unsigned cbase = 0;  // arrival counter of the tournament panels (cleared with the panel state)
    const int set_stride = kMaxInner * (1 + 4 * kMaxW);
    auto maps = [&](long long K0) { return ws.rowmap + ((K0 / NB) % kRowmapSets) * set_stride; };
    LRN_TRY(reset_panel_state(ws, s));
    NextColumns nc;
    nc.KE = std::min(2 * NB, n);
    LRN_TRY((factor_block_column<T, W>(A, n, 0, std::min(NB, n), piv, ws, maps(0), base, s, calu, cbase, nc)));
    for (long long K = 0; K < n; K += NB) {
        const long long B = std::min(NB, n - K);
        const long long KB = K + B;
        const long long K2 = std::min(KB + NB, n);  // end of block column i+1
        const int nmaps = static_cast<int>((B + W - 1) / W);
        LRN_TRY(launch_laswp_multi<T>(A, lda, maps(K), nmaps, rm_stride, 2 * W, 0, K, K2, n, s));
        if (KB >= n) break;
        if (K2 < n) {
            LRN_TRY((u12_columns<T, W>(A, n, K, B, K2, n, s)));
            LRN_TRY(launch_trailing<T>(A + KB * lda + K, lda, A + K * lda + K2, lda, A + KB * lda + K2, lda, n - KB,
                n - K2, B, upd, s));
        }
        nc.KE = std::min(K2 + NB, n);
        LRN_TRY((factor_block_column<T, W>(A, n, KB, K2 - KB, piv, ws, maps(KB), base, s, calu, cbase, nc)));
    }
    return LRN_OK;
}

// The blocked factorisation, right-looking, optionally with a look-ahead. For block column i
// = [K, KB) (the panel stream factors it; its inner panels' updates of block column i+1 run on
// the narrow stream as each panel finishes, so block column i+1 is ready when block column i
// is), with K2, K3 the ends of block columns i+1 and i+2:
//   gemm stream  : block column i+2's interchanges, U12 rows and update by block column i
//                  (event next_ready), then the interchanges on the columns [0, K) and on the
//                  far columns [K3, n), their U12 rows, and the far part of A22 -= L21 * U12;
//   narrow stream: after next_ready, block column i+1's inner-panel updates of block column
//                  i+2 (and their interchanges on block column i+1's earlier columns);
//   panel stream : the factorisation of block column i+1 as soon as block column i is done,
// so the panels of the next block column run while the far update of this one is still going
// and no block column waits for a separate update of its own columns. Nothing waits on the
// host. Results are bitwise those of getrf_blocked_plain (the same kernels on the same
// column ranges, each entry computed independently of how the ranges are split).
template <typename T, int W>
int getrf_blocked(T* A, long long n, long long* piv, const PanelWs<T>& ws, hipStream_t s, int upd = -1,
    const CaluWs<float>* calu = nullptr)
{
    const long long lda = n;
    const long long NB = std::min<long long>(std::max<long long>(tuned_outer_nb(n), W), static_cast<long long>(W) * kMaxInner);
    const int rm_stride = 1 + 4 * W;
    unsigned base = 0;
    unsigned cbase = 0;  // arrival counter of the tournament panels (cleared with the panel state)
    const int set_stride = kMaxInner * (1 + 4 * kMaxW);
    auto maps = [&](long long K0) { return ws.rowmap + ((K0 / NB) % kRowmapSets) * set_stride; };
    if (g_lookahead.load() == 0 || n <= NB) return getrf_blocked_plain<T, W>(A, n, piv, ws, s, upd, calu);
    // compute units of the panel stream: a multiple of 8 (the same number on every XCD),
    // enough to hold every cooperative workgroup of the largest (first) panel at once, and one
    // more per XCD: the narrow stream's work runs beside the panels and may hold slots on
    // these units, so an exact fit can leave a panel unable to become resident (at n = 16384,
    // 64 panel workgroups on 16 units timed out their bounded wait)
    const int la_mode = g_lookahead.load();
    int np = la_mode == 1 ? 32 : la_mode;
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    {
        int need = 0;
        unsigned base_q = 0;
        LRN_TRY((launch_panel<T, W>(A, lda, 0, n, W, piv, ws, base_q, s, &need)));
        np = std::max(np, (need + kXcds - 1) / kXcds * kXcds + kXcds);
    }
    if (np > device_props(dev).cus / 2) return getrf_blocked_plain<T, W>(A, n, piv, ws, s, upd, calu);
    AsideStream* side = nullptr;
    LRN_TRY(aside_stream(np, &side));
    std::lock_guard<std::mutex> side_lock(side->mu);
    const int gemm_cus = device_props(dev).cus - side->np;
    hipStream_t sn = side->narrow;
    LRN_HIP(hipEventRecord(side->start, s));
    LRN_HIP(hipStreamWaitEvent(sn, side->start, 0));
    LRN_HIP(hipStreamWaitEvent(side->panel, side->start, 0));
    LRN_HIP(hipStreamWaitEvent(side->gemm, side->start, 0));
    LRN_TRY(reset_panel_state(ws, sn));
    // the first block column and its update of the second: one stream (all compute units)
    NextColumns first;
    first.KE = std::min(2 * NB, n);
    LRN_TRY((factor_block_column<T, W>(A, n, 0, std::min(NB, n), piv, ws, maps(0), base, sn, calu, cbase, first)));
    LRN_HIP(hipEventRecord(side->done, sn));
    NextColumns nc;
    nc.sx = sn;
    nc.step = side->step;
    nc.ext_done = side->ext_done;
    // the step ending at column KB0 updates the block column after next on the narrow stream
    // (FP64 over the first 30 % of the columns, where the far update is the critical path)
    auto next_on_narrow = [&](long long KB0) { return std::is_same<T, double>::value && 10 * (n - KB0) > 7 * n; };
    for (long long K = 0; K < n; K += NB) {
        const long long B = std::min(NB, n - K);
        const long long KB = K + B;
        const int nmaps = static_cast<int>((B + W - 1) / W);
        int* const rms = maps(K);
        const long long K2 = std::min(KB + NB, n);  // end of block column i+1
        const long long K3 = std::min(K2 + NB, n);  // end of block column i+2
        // block column i+2's interchanges, U12 rows and update by block column i: on the GEMM
        // stream ahead of the far columns (event next_ready; on the GEMM stream's own compute
        // units they do not slow the panels, as they do from the narrow stream, whose kernels
        // share the panel's units while the far GEMM fills the others), except while the FP64
        // far update outlasts the panels, where they would lengthen the critical GEMM stream
        const bool next_on_gemm = !next_on_narrow(KB);
        LRN_HIP(hipStreamWaitEvent(side->gemm, side->done, 0));
        if (next_on_gemm && K3 > K2) {
            LRN_TRY(launch_laswp_multi<T>(A, lda, rms, nmaps, rm_stride, 2 * W, K2, K2, K2, K3, side->gemm));
            LRN_TRY((u12_columns<T, W>(A, n, K, B, K2, K3, side->gemm)));
            LRN_TRY(launch_trailing<T>(A + KB * lda + K, lda, A + K * lda + K2, lda, A + KB * lda + K2, lda, n - KB,
                K3 - K2, B, upd, side->gemm, gemm_cus));
        }
        if (next_on_gemm) LRN_HIP(hipEventRecord(side->next_ready, side->gemm));
        // then the left columns' interchanges and the far columns
        LRN_TRY(launch_laswp_multi<T>(A, lda, rms, nmaps, rm_stride, 2 * W, 0, K, K3, n, side->gemm));
        if (K3 < n) {
            LRN_TRY((u12_columns<T, W>(A, n, K, B, K3, n, side->gemm)));
            LRN_TRY(launch_trailing<T>(A + KB * lda + K, lda, A + K * lda + K3, lda, A + KB * lda + K3, lda, n - KB,
                n - K3, B, upd, side->gemm, gemm_cus));
        }
        LRN_HIP(hipEventRecord(side->rest_done, side->gemm));
        if (KB >= n) break;
        if (next_on_gemm) {
            LRN_HIP(hipStreamWaitEvent(sn, side->next_ready, 0));
        } else {
            // on the narrow stream (it waited for block column i-1's far update, which covered
            // these columns, at the end of the last step)
            LRN_HIP(hipStreamWaitEvent(sn, side->done, 0));
            if (K3 > K2) {
                LRN_TRY(launch_laswp_multi<T>(A, lda, rms, nmaps, rm_stride, 2 * W, K2, K2, K2, K3, sn));
                LRN_TRY((u12_columns<T, W>(A, n, K, B, K2, K3, sn)));
                LRN_TRY(launch_trailing<T>(A + KB * lda + K, lda, A + K * lda + K2, lda, A + KB * lda + K2, lda,
                    n - KB, K3 - K2, B, upd, sn));
            }
        }
        // block column i+1 on the panel stream, its updates of block column i+2 on the narrow one
        LRN_HIP(hipStreamWaitEvent(side->panel, side->done, 0));
        nc.KE = K3;
        LRN_TRY((factor_block_column<T, W>(A, n, KB, K2 - KB, piv, ws, maps(KB), base, side->panel, calu, cbase, nc)));
        LRN_HIP(hipEventRecord(side->done, side->panel));
        // block column i+3's update by block column i+1 needs this step's far update: on the
        // GEMM stream by order, on the narrow stream by this wait
        if (next_on_narrow(K2)) LRN_HIP(hipStreamWaitEvent(sn, side->rest_done, 0));
    }
    // hand back: the caller's stream continues once all three streams are done
    LRN_HIP(hipEventRecord(side->ready, sn));
    LRN_HIP(hipEventRecord(side->done, side->panel));
    LRN_HIP(hipEventRecord(side->rest_done, side->gemm));
    LRN_HIP(hipStreamWaitEvent(s, side->ready, 0));
    LRN_HIP(hipStreamWaitEvent(s, side->done, 0));
    LRN_HIP(hipStreamWaitEvent(s, side->rest_done, 0));
    return LRN_OK;
}

// single_stream: the factorisation stays on `s` (no look-ahead helper streams, so no HIP object
// is created); the results are bitwise those of the look-ahead order either way.
template <typename T>
int getrf_dispatch(T* A, long long n, long long* piv, const PanelWs<T>& ws, hipStream_t s, bool single_stream = false)
{
    const int w = tuned_panel_w();
    if (single_stream)
        return w <= 32 ? getrf_blocked_plain<T, 32>(A, n, piv, ws, s, -1, nullptr)
                       : getrf_blocked_plain<T, 64>(A, n, piv, ws, s, -1, nullptr);
    if (w <= 32) return getrf_blocked<T, 32>(A, n, piv, ws, s);
    return getrf_blocked<T, 64>(A, n, piv, ws, s);
}

template <typename T> size_t getrf_ws_bytes(long long)
{
    Arena ar;
    carve_getrf<T>(ar);
    return ar.off;
}

template <typename T> size_t getrs_ws_bytes(long long n, long long nrhs)
{
    Arena ar;
    carve_getrs<T>(ar, n, nrhs);
    return ar.off;
}

template <typename T> size_t solver_ws_bytes(long long n, long long nrhs, int type);

// Single-RHS substitution with FP32 factors in FP64 arithmetic (the refinement's
// preconditioner): 0 = the 64-row kernel (the default), 1 = 256-row chained kernel (streams
// its factor rows through too few compute units to keep up), 2 = 128-row chained kernel.
std::atomic<int> g_chain_solve{4};  // refinement substitution (lrn_set_refine_solver)
// Refinement of the mixed solver: 0 = LU-IR while each step gains at least 2x, then GMRES-IR;
// 1 = GMRES-IR right after the first LU-IR step.
std::atomic<int> g_refinement{0};
// Solves with at least this many right-hand sides (and at least two) use the blocked form with
// inverted diagonal blocks (lrn_trsm.hpp); fewer use the chained substitution kernels. 0: never.
std::atomic<int> g_multi_rhs_min{2};
bool use_blocked_trsm(long long nrhs)
{
    const int m = g_multi_rhs_min.load(std::memory_order_relaxed);
    return m > 0 && nrhs >= 2 && nrhs >= m;
}
// Kernel of the blocked form: 0 = one launch per 64-row strip (trsm_step_kernel), 1 = one
// launch per 64 columns with the strips chained through flags (trsm_chain_kernel; falls back
// to 0 when its workgroups cannot all be resident).
std::atomic<int> g_multi_rhs_kernel{1};

// X = op(A)^-1 X (X n x nrhs, stride ldx) with the diagonal-block inverses in dinv. flags:
// solve_blocks(n) words, err: the error word (an expired chained wait raises it).
template <typename T>
int trsm_apply(bool lower, bool trans, const T* a, long long lda, long long n, const T* dinv, T* x, long long ldx,
    long long nrhs, unsigned* flags, int* err, const int* skip, hipStream_t s)
{
    if (g_multi_rhs_kernel.load(std::memory_order_relaxed) == 1) {
        int rc = LRN_OK;
        for (long long c0 = 0; c0 < nrhs && rc == LRN_OK; c0 += kSolveMaxCols) {
            const long long nc = std::min<long long>(kSolveMaxCols, nrhs - c0);
            LRN_TRY(launch_fill(flags, 0u, solve_blocks(n), s));
            rc = trsm_chain_launch<T>(lower, trans, a, lda, n, dinv, x + c0, ldx, nc, flags, spin_limit(), err, skip, s);
            if (rc == LRN_E_UNSUPPORTED && c0 != 0) return LRN_E_INTERNAL;  // residency cannot change between chunks
        }
        if (rc != LRN_E_UNSUPPORTED) return rc;
    }
    return trsm_steps_launch<T>(lower, trans, a, lda, n, dinv, x, ldx, nrhs, skip, s);
}

// Solve with factors already in place. skip_dev: if non-null and nonzero, the kernels return
// without touching B (LU mode after a singular factorisation).
// Row permutation of the factorisation (from the pivots), into L.perm.
template <typename T> int getrs_perm(long long n, const long long* piv, const GetrsLayout<T>& L, hipStream_t s)
{
    const size_t perm_lds = (n <= kPermLdsMax) ? static_cast<size_t>(n) * sizeof(int) : 0;
    {
        int per_cu = 0;
        LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(perm_compose_kernel), perm_lds, 256, &per_cu));
    }
    const long long nblocks = (n + 63) / 64;
    hipLaunchKernelGGL(pivot_block_maps_kernel, dim3(static_cast<unsigned>(nblocks)), dim3(64), 0, s, piv, n, L.maps);
    LRN_HIP(hipGetLastError());
    hipLaunchKernelGGL(perm_compose_kernel, dim3(1), dim3(256), perm_lds, s, L.maps, n, nblocks, L.perm, L.perm_tmp);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

// Forward and backward substitution with the permutation already in L.perm; factors of type
// TF, right-hand sides and arithmetic of type T.
// skip_dev: if non-null and nonzero, the kernels return without touching B.
// sticky_err: L.err is neither cleared nor checked here (the caller clears it once and reads it
// after a sequence of solves); otherwise it is cleared first and an expired wait leaves NaN in B.
template <typename T, typename TF = T>
int getrs_subst(const TF* LU, long long n, T* B, long long nrhs, long long ldb, const GetrsLayout<T>& L,
    const int* skip_dev, hipStream_t s, bool sticky_err = false)
{
    const long long nblk = solve_blocks(n);
    if (!sticky_err) LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(L.err), 0u, 64, s));
    if constexpr (std::is_same<T, TF>::value) {
        if (L.dinv != nullptr && use_blocked_trsm(nrhs)) {
            // blocked: P B gathered into y 64 columns at a time, L then U applied in place
            // strip by strip with the inverted diagonal blocks, y copied back into B
            const double tri_bytes = (0.5 * static_cast<double>(n) * n + 2.0 * n * nrhs) * sizeof(T);
            TraceScope ts(kTrTrsv, 2.0 * tri_bytes, s);
            T* dl = L.dinv;
            T* du = L.dinv + trsm_dinv_elems(n);
            LRN_TRY(trsm_inv_launch<T>(true, true, false, LU, n, n, dl, skip_dev, s));
            LRN_TRY(trsm_inv_launch<T>(false, false, false, LU, n, n, du, skip_dev, s));
            for (long long c0 = 0; c0 < nrhs; c0 += kSolveMaxCols) {
                const long long nc = std::min<long long>(kSolveMaxCols, nrhs - c0);
                LRN_TRY(gather_rows_launch<T>(B + c0, ldb, L.perm, L.y, nc, n, nc, skip_dev, s));
                LRN_TRY(trsm_apply<T>(true, false, LU, n, n, dl, L.y, nc, nc, L.flags_f, L.err, skip_dev, s));
                LRN_TRY(trsm_apply<T>(false, false, LU, n, n, du, L.y, nc, nc, L.flags_b, L.err, skip_dev, s));
                LRN_TRY(gather_rows_launch<T>(L.y, nc, nullptr, B + c0, ldb, n, nc, skip_dev, s));
            }
            if (!sticky_err) LRN_TRY(launch_poison<T>(L.err, B, n, nrhs, ldb, s));
            return LRN_OK;
        }
    }
    for (long long c0 = 0; c0 < nrhs; c0 += kSolveMaxCols) {
        const int nc = static_cast<int>(std::min<long long>(kSolveMaxCols, nrhs - c0));
        if (nc == 1) {
            // latency-optimised single right-hand side: tags 1 (forward) and 2 (backward)
            LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(L.xch), 0u, 4 * n, s));
            const double tri_bytes = (0.5 * static_cast<double>(n) * n) * sizeof(TF) + 4.0 * n * sizeof(T);
            TraceScope ts(kTrTrsv, 2.0 * tri_bytes, s);
            if constexpr (std::is_same<TF, float>::value && std::is_same<T, double>::value) {
                const int chain = g_chain_solve.load();
                if ((chain == 3 || chain == 4) && L.dinv != nullptr) {
                    // the diagonal blocks applied through their inverses and the predecessor
                    // through its precomputed product (in L.dinv, written by the caller with
                    // refine_inverses); mode 4 also keeps runs of consecutive blocks on one XCD
                    // when every workgroup can be resident
                    int xcd_per = 0;
                    if (chain == 4 && nblk >= 2 * kXcds) {
                        int per_cu = 0, dev = 0;
                        LRN_TRY(prepare_kernel(reinterpret_cast<const void*>(trsv1_inv_kernel<float, double, true>), 0,
                            256, &per_cu));
                        LRN_HIP(hipGetDevice(&dev));
                        if (nblk <= static_cast<long long>(per_cu) * device_props(dev).cus)
                            xcd_per = static_cast<int>((nblk + kXcds - 1) / kXcds);
                    }
                    const unsigned grid = static_cast<unsigned>(xcd_per > 0 ? kXcds * xcd_per : nblk);
                    const long long e = trsm_dinv_elems(n);
                    const double* dl = L.dinv;
                    const double* du = L.dinv + e;
                    hipLaunchKernelGGL((trsv1_inv_kernel<float, double, true>), dim3(grid), dim3(256), 0, s, LU, n, n,
                        L.perm, B + c0, ldb, L.y, 1ll, L.xch, 1u, spin_limit(), L.err, skip_dev, dl, dl + 2 * e,
                        dl + 4 * e, xcd_per);
                    LRN_HIP(hipGetLastError());
                    hipLaunchKernelGGL((trsv1_inv_kernel<float, double, false>), dim3(grid), dim3(256), 0, s, LU, n,
                        n, L.perm, L.y, 1ll, B + c0, ldb, L.xch, 2u, spin_limit(), L.err, skip_dev, du, du + 2 * e,
                        du + 4 * e, xcd_per);
                    LRN_HIP(hipGetLastError());
                    continue;
                }
                if ((chain == 1 || chain == 2) && ldb == 1) {
                    const bool vec4 = (n % 4 == 0) && (reinterpret_cast<uintptr_t>(LU) % 16 == 0);
                    auto go = [&](auto ns) -> int {
                        constexpr int NS = decltype(ns)::value;
                        constexpr int TB = kChainSB * NS;
                        const unsigned nbk = static_cast<unsigned>((n + TB - 1) / TB);
                        if (vec4) {
                            hipLaunchKernelGGL((trsv_chain_kernel<T, true, true, NS>), dim3(nbk), dim3(TB), 0, s, LU, n,
                                n, L.perm, B + c0, L.y, L.xch, 1u, spin_limit(), L.err, skip_dev);
                            LRN_HIP(hipGetLastError());
                            hipLaunchKernelGGL((trsv_chain_kernel<T, false, true, NS>), dim3(nbk), dim3(TB), 0, s, LU, n,
                                n, L.perm, L.y, B + c0, L.xch, 2u, spin_limit(), L.err, skip_dev);
                        } else {
                            hipLaunchKernelGGL((trsv_chain_kernel<T, true, false, NS>), dim3(nbk), dim3(TB), 0, s, LU, n,
                                n, L.perm, B + c0, L.y, L.xch, 1u, spin_limit(), L.err, skip_dev);
                            LRN_HIP(hipGetLastError());
                            hipLaunchKernelGGL((trsv_chain_kernel<T, false, false, NS>), dim3(nbk), dim3(TB), 0, s, LU,
                                n, n, L.perm, L.y, B + c0, L.xch, 2u, spin_limit(), L.err, skip_dev);
                        }
                        LRN_HIP(hipGetLastError());
                        return LRN_OK;
                    };
                    if (chain == 2)
                        LRN_TRY(go(std::integral_constant<int, 2>{}));
                    else
                        LRN_TRY(go(std::integral_constant<int, 4>{}));
                    continue;
                }
            }
            hipLaunchKernelGGL((trsv1_kernel<TF, T, true>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, LU, n,
                n, L.perm, B + c0, ldb, L.y, 1ll, L.xch, 1u, spin_limit(), L.err, skip_dev);
            LRN_HIP(hipGetLastError());
            hipLaunchKernelGGL((trsv1_kernel<TF, T, false>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, LU, n,
                n, L.perm, L.y, 1ll, B + c0, ldb, L.xch, 2u, spin_limit(), L.err, skip_dev);
            LRN_HIP(hipGetLastError());
            continue;
        }
        LRN_TRY(launch_fill(L.flags_f, 0u, nblk, s));
        LRN_TRY(launch_fill(L.flags_b, 0u, nblk, s));
        const double tri_bytes = (0.5 * static_cast<double>(n) * n + 2.0 * n * nc) * sizeof(T);
        TraceScope ts(kTrTrsv, 2.0 * tri_bytes, s);
        hipLaunchKernelGGL((trsv_blocked_kernel<TF, T, true>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, LU, n, n,
            L.perm, B + c0, ldb, L.y, static_cast<long long>(nc), nc, L.flags_f, spin_limit(), L.err, skip_dev);
        LRN_HIP(hipGetLastError());
        hipLaunchKernelGGL((trsv_blocked_kernel<TF, T, false>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, LU, n, n,
            L.perm, L.y, static_cast<long long>(nc), B + c0, ldb, nc, L.flags_b, spin_limit(), L.err, skip_dev);
        LRN_HIP(hipGetLastError());
    }
    if (!sticky_err) LRN_TRY(launch_poison<T>(L.err, B, n, nrhs, ldb, s));
    return LRN_OK;
}

// Solve with factors already in place.
template <typename T>
int getrs_run(const T* LU, long long n, const long long* piv, T* B, long long nrhs, long long ldb,
    const GetrsLayout<T>& L, const int* skip_dev, hipStream_t s)
{
    LRN_TRY(getrs_perm<T>(n, piv, L, s));
    return getrs_subst<T>(LU, n, B, nrhs, ldb, L, skip_dev, s);
}

constexpr long long kMaxOrder = 1ll << 17;
// ------------------------------------------------------------- triangular solver
// op(A) X = B with A triangular (reference: include/matrix_operation/triangular_solver.h:11-38,
// algorithm src/matrix_operation/cpu/native/triangular_solver_native.cpp:64-153): rows are
// solved in 64-row blocks chained between workgroups (trsv_blocked_kernel with the
// direction, unit-diagonal and transpose flags), 64 right-hand sides per launch.
// Two or more right-hand sides: also the inverses of the diagonal blocks (blocked form).
template <typename T> size_t trsm_ws_bytes(long long n, long long nrhs)
{
    Arena ar;
    ar.next<unsigned>(static_cast<size_t>(solve_blocks(n)));
    ar.next<int>(64);
    if (nrhs >= 2) ar.next<T>(static_cast<size_t>(trsm_dinv_elems(n)));
    return ar.off;
}

template <typename T, bool kFwd, bool kUnit, bool kTrans>
int launch_trsm_chain(const T* a, long long lda, long long n, const T* b, long long ldb, T* x, long long ldx, int nc,
    unsigned* flags, int* err, hipStream_t s, const int* skip = nullptr)
{
    hipLaunchKernelGGL((trsv_blocked_kernel<T, T, kFwd, kUnit, kTrans>), dim3(static_cast<unsigned>(solve_blocks(n))),
        dim3(256), 0, s, a, lda, n, nullptr, b, ldb, x, ldx, nc, flags, spin_limit(), err, skip);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T>
int triangular_solver_impl(int upper, int transpose, int unit, long long n, long long nrhs, const T* a, long long lda,
    const T* b, long long ldb, T* x, long long ldx, void* ws, size_t ws_size, hipStream_t s)
{
    // the reference's checks (src/common/triangular_solver_checks.h:11-35), in argument order
    if (n <= 0) return -4;
    if (nrhs <= 0) return -5;
    if (a == nullptr) return -6;
    if (lda < n) return -7;
    if (b == nullptr || b == a) return -8;
    if (ldb < nrhs) return -9;
    if (x == nullptr || x == a || x == b) return -10;
    if (ldx < nrhs) return -11;
    if (n > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (ws == nullptr || ws_size < trsm_ws_bytes<T>(n, nrhs)) return -12;
    Arena ar(ws, ws_size);
    unsigned* flags = ar.next<unsigned>(static_cast<size_t>(solve_blocks(n)));
    int* err = ar.next<int>(64);
    T* dinv = nrhs >= 2 ? ar.next<T>(static_cast<size_t>(trsm_dinv_elems(n))) : nullptr;
    // rows in increasing order when op(A) is lower triangular
    const bool fwd = (upper != 0) == (transpose != 0);
    const bool u = unit != 0, t = transpose != 0;
    TraceScope ts(kTrTrsv, (0.5 * static_cast<double>(n) * n + 2.0 * static_cast<double>(n) * nrhs) * sizeof(T), s);
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(err), 0u, 64, s));
    if (dinv != nullptr && use_blocked_trsm(nrhs)) {
        // blocked: X = B, then X = op(A)^-1 X strip by strip with the inverted diagonal blocks
        LRN_TRY(trsm_inv_launch<T>(fwd, u, t, a, lda, n, dinv, nullptr, s));
        LRN_TRY(gather_rows_launch<T>(b, ldb, nullptr, x, ldx, n, nrhs, nullptr, s));
        LRN_TRY(trsm_apply<T>(fwd, t, a, lda, n, dinv, x, ldx, nrhs, flags, err, nullptr, s));
        return launch_poison<T>(err, x, n, nrhs, ldx, s);
    }
    for (long long c0 = 0; c0 < nrhs; c0 += kSolveMaxCols) {
        const int nc = static_cast<int>(std::min<long long>(kSolveMaxCols, nrhs - c0));
        LRN_TRY(launch_fill(flags, 0u, solve_blocks(n), s));
        const T* bb = b + c0;
        T* xx = x + c0;
        int rc;
        if (fwd)
            rc = u ? (t ? launch_trsm_chain<T, true, true, true>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s)
                        : launch_trsm_chain<T, true, true, false>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s))
                   : (t ? launch_trsm_chain<T, true, false, true>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s)
                        : launch_trsm_chain<T, true, false, false>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s));
        else
            rc = u ? (t ? launch_trsm_chain<T, false, true, true>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s)
                        : launch_trsm_chain<T, false, true, false>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s))
                   : (t ? launch_trsm_chain<T, false, false, true>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s)
                        : launch_trsm_chain<T, false, false, false>(a, lda, n, bb, ldb, xx, ldx, nc, flags, err, s));
        LRN_TRY(rc);
    }
    // an expired wait between row blocks leaves NaN, never a wrong x
    return launch_poison<T>(err, x, n, nrhs, ldx, s);
}


// ------------------------------------------------------------- Cholesky
// A = L L^T, row-major, lower triangle read (reference: include/matrix_operation/
// cholesky_decomposition.h; algorithm src/matrix_operation/cpu/native/
// cholesky_decomposition_native.cpp:245-320): 64-wide block columns, each the diagonal block
// (chol_diag_kernel), the rows below it (chol_rows_kernel, also written transposed into Wt)
// and the trailing update A22 -= L21 L21^T on the GEMM kernels (the whole square: the upper
// part it writes is never read and is overwritten by the final layout pass).
long long chol_ldw(long long n) { return (n + 3) / 4 * 4; }

template <typename T> struct CholLayout {
    int* info = nullptr;
    T* wt = nullptr;
};

template <typename T> CholLayout<T> carve_chol(Arena& ar, long long n)
{
    CholLayout<T> L;
    L.info = ar.next<int>(64);
    L.wt = ar.next<T>(static_cast<size_t>(kCholNB) * static_cast<size_t>(chol_ldw(n)));
    return L;
}

template <typename T> size_t chol_ws_bytes(long long n)
{
    Arena ar;
    carve_chol<T>(ar, n);
    return ar.off;
}

// Factor into the lower triangle (info on the device, zero on success), then write the
// layout (upper = 1: U = L^T in the upper triangle).
template <typename T> int chol_factor(T* A, long long n, int upper, const CholLayout<T>& L, hipStream_t s)
{
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(L.info), 0u, 64, s));
    const long long ldw = chol_ldw(n);
    for (long long k = 0; k < n; k += kCholNB) {
        const int bs = static_cast<int>(std::min<long long>(kCholNB, n - k));
        hipLaunchKernelGGL(chol_diag_kernel<T>, dim3(1), dim3(64), 0, s, A, n, k, bs, L.info);
        LRN_HIP(hipGetLastError());
        const long long m = n - k - bs;
        if (m <= 0) break;
        hipLaunchKernelGGL(chol_rows_kernel<T>, dim3(static_cast<unsigned>((m + 63) / 64)), dim3(64), 0, s, A, n, k,
            bs, n, L.wt, ldw, L.info);
        LRN_HIP(hipGetLastError());
        // skipped once a diagonal block has failed (the reference stops at the failing block)
        LRN_TRY(launch_gemm_sub<T>(A + (k + bs) * n + k, n, L.wt, ldw, A + (k + bs) * n + (k + bs), n, m, m, bs, s, 0,
            L.info, true));
    }
    const long long pairs = n * (n - 1) / 2;
    const unsigned grid = static_cast<unsigned>(std::max<long long>(1, std::min<long long>((pairs + 255) / 256, 1ll << 16)));
    hipLaunchKernelGGL(chol_layout_kernel<T>, dim3(grid), dim3(256), 0, s, A, n, upper, L.info);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T> size_t solver_ws_bytes(long long n, long long nrhs, int type)
{
    Arena ar;
    if (type == LRN_LU_LINEAR_SOLVER) carve_getrf<T>(ar);
    if (type == LRN_CHOLESKY_LINEAR_SOLVER) carve_chol<T>(ar, n);
    carve_getrs<T>(ar, n, nrhs);
    return ar.off;
}

// Asynchronous form: info to device memory, no host synchronisation (the reference's GPU
// specialisation, src/matrix_operation/gpu/hipblas/cholesky_decomposition_hip.cpp:69-100).
template <typename T>
int cholesky_async_impl(T* c, int64_t lda, int upper, void* workspace, size_t ws_size, int* info_dev, void* stream)
{
    if (c == nullptr) return -1;
    if (lda <= 0) return -2;
    if (upper != 0 && upper != 1) return -3;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < chol_ws_bytes<T>(lda)) return -4;
    if (info_dev == nullptr) return -6;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const auto L = carve_chol<T>(ar, lda);
    LRN_TRY(chol_factor<T>(c, lda, upper, L, s));
    LRN_HIP(hipMemcpyAsync(info_dev, L.info, sizeof(int), hipMemcpyDeviceToDevice, s));
    return LRN_OK;
}

template <typename T>
int cholesky_impl(T* c, int64_t lda, int upper, void* workspace, size_t ws_size, int* info, void* stream)
{
    if (c == nullptr) return -1;
    if (lda <= 0) return -2;
    if (upper != 0 && upper != 1) return -3;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < chol_ws_bytes<T>(lda)) return -4;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const auto L = carve_chol<T>(ar, lda);
    LRN_TRY(chol_factor<T>(c, lda, upper, L, s));
    int host = 0;
    LRN_HIP(hipMemcpyAsync(&host, L.info, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipStreamSynchronize(s));
    if (info != nullptr) *info = host;
    return LRN_OK;
}

// L L^T X = B in place (B row-major n x nrhs, stride ldb) with the lower factor in m: forward
// with L, backward with L^T (transposed reads of the same triangle), 64 columns per launch
// through G.y; skipped while *skip is nonzero.
template <typename T>
int chol_solve(const T* m, long long n, T* B, long long nrhs, long long ldb, const GetrsLayout<T>& G,
    const int* skip, hipStream_t s)
{
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(G.err), 0u, 64, s));
    if (G.dinv != nullptr && use_blocked_trsm(nrhs)) {
        // blocked, in place on B: L (lower, non-unit) then L^T (the same triangle read transposed)
        const double tri_bytes = (0.5 * static_cast<double>(n) * n + 2.0 * n * nrhs) * sizeof(T);
        TraceScope ts(kTrTrsv, 2.0 * tri_bytes, s);
        T* dl = G.dinv;
        T* dlt = G.dinv + trsm_dinv_elems(n);
        LRN_TRY(trsm_inv_launch<T>(true, false, false, m, n, n, dl, skip, s));
        LRN_TRY(trsm_inv_launch<T>(false, false, true, m, n, n, dlt, skip, s));
        LRN_TRY(trsm_apply<T>(true, false, m, n, n, dl, B, ldb, nrhs, G.flags_f, G.err, skip, s));
        LRN_TRY(trsm_apply<T>(false, true, m, n, n, dlt, B, ldb, nrhs, G.flags_b, G.err, skip, s));
        return launch_poison<T>(G.err, B, n, nrhs, ldb, s);
    }
    for (long long c0 = 0; c0 < nrhs; c0 += kSolveMaxCols) {
        const int nc = static_cast<int>(std::min<long long>(kSolveMaxCols, nrhs - c0));
        LRN_TRY(launch_fill(G.flags_f, 0u, solve_blocks(n), s));
        LRN_TRY(launch_fill(G.flags_b, 0u, solve_blocks(n), s));
        LRN_TRY((launch_trsm_chain<T, true, false, false>(m, n, n, B + c0, ldb, G.y, nc, nc, G.flags_f, G.err, s, skip)));
        LRN_TRY((launch_trsm_chain<T, false, false, true>(m, n, n, G.y, nc, B + c0, ldb, nc, G.flags_b, G.err, s, skip)));
    }
    return launch_poison<T>(G.err, B, n, nrhs, ldb, s);
}

// ------------------------------------------------------------- batched multiplication
// C_i = A_i B_i (include/matrix_operation/multiplication_batched.h:7-42): pointer checks first,
// then the sizes (multiplication_batched_hip.hip:46-57); asynchronous on the stream.
template <typename T>
int batched_mm_impl(int64_t dim, int64_t count, const T* a, const T* b, T* c, void* stream)
{
    if (a == nullptr) return -3;
    if (b == nullptr) return -4;
    if (c == nullptr) return -5;
    if (dim <= 0) return -1;
    if (count <= 0) return -2;
    if (dim > (1ll << 16)) return LRN_E_UNSUPPORTED;
    const long long tn = (dim + kBmmTile - 1) / kBmmTile;
    const long long tiles = tn * tn;
    if (count > 0x7fffffffll / tiles) return LRN_E_UNSUPPORTED;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dim <= kBmmSmall) {
        const long long total = dim * dim * count;
        if ((total + 255) / 256 > 0x7fffffffll) return LRN_E_UNSUPPORTED;
        hipLaunchKernelGGL(batched_gemm_small_kernel<T>, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0,
            s, static_cast<int>(dim), total, a, b, c);
        LRN_HIP(hipGetLastError());
        return LRN_OK;
    }
    hipLaunchKernelGGL(batched_gemm_kernel<T>, dim3(static_cast<unsigned>(tiles * count)), dim3(256), 0, s,
        static_cast<int>(dim), static_cast<long long>(count), static_cast<int>(tn), tiles, a, b, c);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T>
int lu_decomposition_impl(T* m, int64_t lda, int64_t* pivot, void* workspace, size_t ws_size, int* info, void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -2;
    if (pivot == nullptr) return -3;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < getrf_ws_bytes<T>(lda)) return -4;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const auto L = carve_getrf<T>(ar);
    LRN_TRY(getrf_dispatch<T>(m, lda, reinterpret_cast<long long*>(pivot), L.ws, s));
    int host[2] = {0, 0};
    LRN_HIP(hipMemcpyAsync(&host[0], L.ws.info, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipMemcpyAsync(&host[1], L.ws.err, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipStreamSynchronize(s));
    if (host[1] != 0) return LRN_E_TIMEOUT;
    if (info != nullptr) *info = host[0];
    return LRN_OK;
}

template <typename T>
int getrf_async_impl(T* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace, size_t ws_size, void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -2;
    if (pivot == nullptr) return -3;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < getrf_ws_bytes<T>(lda)) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const auto L = carve_getrf<T>(ar);
    LRN_TRY(getrf_dispatch<T>(m, lda, reinterpret_cast<long long*>(pivot), L.ws, s));
    if (info_dev != nullptr) LRN_HIP(hipMemcpyAsync(info_dev, L.ws.info, sizeof(int), hipMemcpyDeviceToDevice, s));
    return LRN_OK;
}

struct TournamentLayout {
    GetrfLayout<float> f;
    CaluWs<float> calu;
};

TournamentLayout carve_tournament(Arena& ar)
{
    TournamentLayout L;
    L.f = carve_getrf<float>(ar);
    L.calu.cand = ar.next<int>(static_cast<size_t>(kCaluMaxP) * kMaxW);
    L.calu.lu11 = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.top = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.uinv = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.ctr = L.f.ws.ctr == nullptr ? nullptr : L.f.ws.ctr + 32;
    L.calu.info = L.f.ws.info;
    L.calu.rowmap = nullptr;
    return L;
}

size_t tournament_ws_bytes()
{
    Arena ar;
    carve_tournament(ar);
    return ar.off;
}

int getrf_tournament_impl(float* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace, size_t ws_size,
    void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -2;
    if (pivot == nullptr) return -3;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < tournament_ws_bytes()) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const TournamentLayout L = carve_tournament(ar);
    LRN_TRY((getrf_blocked<float, 32>(m, lda, reinterpret_cast<long long*>(pivot), L.f.ws, s, -1, &L.calu)));
    if (info_dev != nullptr) LRN_HIP(hipMemcpyAsync(info_dev, L.f.ws.info, sizeof(int), hipMemcpyDeviceToDevice, s));
    return LRN_OK;
}

template <typename T>
int solver_impl(T* m, int64_t* pivot, int64_t lda, T* b, int64_t nrhs, int64_t ldb, int type, void* workspace,
    size_t ws_size, void* stream, bool single_stream = false)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -3;
    if (b == nullptr) return -4;
    if (nrhs < 0) return -5;
    if (ldb < std::max<int64_t>(1, nrhs)) return -6;
    const bool chol = type == LRN_CHOLESKY_LINEAR_SOLVER || type == LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER;
    if (!chol && type != LRN_LU_LINEAR_SOLVER && type != LRN_LU_UPFRONT_LINEAR_SOLVER) return -7;
    if (pivot == nullptr && !chol) return -2;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < solver_ws_bytes<T>(lda, std::max<int64_t>(nrhs, 1), type)) return -8;
    if (nrhs == 0) return LRN_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    if (chol) {
        // linear_solver_native.cpp:177-189: factor the lower triangle, solve only on success
        const int* skip = nullptr;
        if (type == LRN_CHOLESKY_LINEAR_SOLVER) {
            const auto C = carve_chol<T>(ar, lda);
            LRN_TRY(chol_factor<T>(m, lda, 0, C, s));
            skip = C.info;
        }
        const auto G = carve_getrs<T>(ar, lda, nrhs);
        return chol_solve<T>(m, lda, b, nrhs, ldb, G, skip, s);
    }
    const int* skip = nullptr;
    GetrfLayout<T> F{};
    if (type == LRN_LU_LINEAR_SOLVER) {
        F = carve_getrf<T>(ar);
        LRN_TRY(getrf_dispatch<T>(m, lda, reinterpret_cast<long long*>(pivot), F.ws, s, single_stream));
    }
    const auto G = carve_getrs<T>(ar, lda, nrhs);
    if (type == LRN_LU_LINEAR_SOLVER) {
        // solve only nonsingular factors (linear_solver_native.cpp:160-171); a factorisation
        // whose bounded waits expired leaves NaN in b instead of a wrong solution
        LRN_TRY(launch_gate<T>(F.ws.info, F.ws.err, G.gate, b, lda, nrhs, ldb, s));
        skip = G.gate;
    }
    return getrs_run<T>(m, lda, reinterpret_cast<const long long*>(pivot), b, nrhs, ldb, G, skip, s);
}

// ------------------------------------------------------------- determinant and inversion
// Direct consumers of the factors (include/matrix_operation/determinant.h:13-55,
// include/matrix_operation/inversion.h:13-53). LU mode factors m in place first, into the
// caller's pivots or, when pivot is null (the GPU reference ignores that argument,
// src/matrix_operation/gpu/hipblas/inversion_hip.hip:662-700), into pivots carved from the
// workspace; LU_UPFRONT reads the caller's factors and pivots.
bool is_lu_type(int type) { return type == LRN_LU_LINEAR_SOLVER || type == LRN_LU_UPFRONT_LINEAR_SOLVER; }
bool is_chol_type(int type)
{
    return type == LRN_CHOLESKY_LINEAR_SOLVER || type == LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER;
}

template <typename T> struct ConsumerLayout {
    GetrfLayout<T> F;
    CholLayout<T> C;
    long long* piv = nullptr;
    int* info = nullptr;
    T* result = nullptr;
};

template <typename T> ConsumerLayout<T> carve_consumer(Arena& ar, long long n, int type)
{
    ConsumerLayout<T> L;
    if (type == LRN_LU_LINEAR_SOLVER) {
        L.F = carve_getrf<T>(ar);
        L.piv = ar.next<long long>(static_cast<size_t>(n));
    }
    if (type == LRN_CHOLESKY_LINEAR_SOLVER) L.C = carve_chol<T>(ar, n);
    L.info = ar.next<int>(64);
    L.result = ar.next<T>(1);
    return L;
}

template <typename T> size_t det_ws_bytes(long long n, int type)
{
    Arena ar;
    carve_consumer<T>(ar, n, type);
    return ar.off;
}

template <typename T> size_t inverse_ws_bytes(long long n, int type)
{
    Arena ar;
    carve_consumer<T>(ar, n, type);
    carve_getrs<T>(ar, n, 1);
    ar.next<T>(static_cast<size_t>(n) * static_cast<size_t>(n));
    return ar.off;
}

// LU X = X in place for all n columns of the n x n block X (row stride n; L unit lower, or
// non-unit lower for a Cholesky factor stored with L^T above it), blocked like the factorisation: 64-row strips solved by strip_trsm_kernel, the rest of each 256-row block
// updated by a GEMM of depth 64, the rows beyond the block by one GEMM of depth 256.
constexpr long long kInvNB = 256;

template <typename T>
int launch_strip(const T* LU, long long n, long long r0, int w, T* X, bool lower, hipStream_t s, bool unit = true)
{
    const dim3 grid(static_cast<unsigned>((n + 63) / 64));
    if (lower && unit)
        hipLaunchKernelGGL((strip_trsm_kernel<T, true, true>), grid, dim3(64), 0, s, LU, n, r0, w, X, n, n);
    else if (lower)
        hipLaunchKernelGGL((strip_trsm_kernel<T, true, false>), grid, dim3(64), 0, s, LU, n, r0, w, X, n, n);
    else
        hipLaunchKernelGGL((strip_trsm_kernel<T, false>), grid, dim3(64), 0, s, LU, n, r0, w, X, n, n);
    LRN_HIP(hipGetLastError());
    return LRN_OK;
}

template <typename T> int blocked_lu_solve_all(const T* LU, long long n, T* X, hipStream_t s, bool unit_lower = true)
{
    // forward: unit lower L, top to bottom
    for (long long b0 = 0; b0 < n; b0 += kInvNB) {
        const long long b1 = std::min(n, b0 + kInvNB);
        for (long long s0 = b0; s0 < b1; s0 += kStripRows) {
            const long long s1 = std::min(b1, s0 + kStripRows);
            LRN_TRY(launch_strip<T>(LU, n, s0, static_cast<int>(s1 - s0), X, true, s, unit_lower));
            LRN_TRY(launch_gemm_sub<T>(LU + s1 * n + s0, n, X + s0 * n, n, X + s1 * n, n, b1 - s1, n, s1 - s0, s));
        }
        LRN_TRY(launch_gemm_sub<T>(LU + b1 * n + b0, n, X + b0 * n, n, X + b1 * n, n, n - b1, n, b1 - b0, s));
    }
    // backward: non-unit upper U, bottom to top
    for (long long b0 = (n - 1) / kInvNB * kInvNB; b0 >= 0; b0 -= kInvNB) {
        const long long b1 = std::min(n, b0 + kInvNB);
        for (long long s0 = b0 + (b1 - 1 - b0) / kStripRows * kStripRows; s0 >= b0; s0 -= kStripRows) {
            const long long s1 = std::min(b1, s0 + kStripRows);
            LRN_TRY(launch_strip<T>(LU, n, s0, static_cast<int>(s1 - s0), X, false, s));
            LRN_TRY(launch_gemm_sub<T>(LU + b0 * n + s0, n, X + s0 * n, n, X + b0 * n, n, s0 - b0, n, s1 - s0, s));
        }
        LRN_TRY(launch_gemm_sub<T>(LU + b0, n, X + b0 * n, n, X, n, b0, n, b1 - b0, s));
    }
    return LRN_OK;
}

// Factor (LU or Cholesky mode) and return the pivots to use (none for Cholesky); info ends up
// in L.info on the device. chol_layout: the Cholesky factor's layout pass (0 = L with zeros
// above, 2 = L with L^T mirrored above), applied to the caller's factor in CHOLESKY_UPFRONT.
template <typename T>
int consumer_factor(T* m, int64_t* pivot, long long n, int type, const ConsumerLayout<T>& L, hipStream_t s,
    const long long** piv_out, int chol_layout = 0)
{
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(L.info), 0u, 64, s));
    if (is_chol_type(type)) {
        *piv_out = nullptr;
        if (type == LRN_CHOLESKY_LINEAR_SOLVER) {
            LRN_TRY(chol_factor<T>(m, n, chol_layout, L.C, s));
            LRN_HIP(hipMemcpyAsync(L.info, L.C.info, sizeof(int), hipMemcpyDeviceToDevice, s));
        } else if (chol_layout == 2) {
            const long long pairs = n * (n - 1) / 2;
            const unsigned grid =
                static_cast<unsigned>(std::max<long long>(1, std::min<long long>((pairs + 255) / 256, 1ll << 16)));
            hipLaunchKernelGGL(chol_layout_kernel<T>, dim3(grid), dim3(256), 0, s, m, n, 2, L.info);
            LRN_HIP(hipGetLastError());
        }
        return LRN_OK;
    }
    if (type == LRN_LU_LINEAR_SOLVER) {
        long long* piv = pivot != nullptr ? reinterpret_cast<long long*>(pivot) : L.piv;
        LRN_TRY(getrf_dispatch<T>(m, n, piv, L.F.ws, s));
        LRN_HIP(hipMemcpyAsync(L.info, L.F.ws.info, sizeof(int), hipMemcpyDeviceToDevice, s));
        *piv_out = piv;
    } else {
        *piv_out = reinterpret_cast<const long long*>(pivot);
    }
    return LRN_OK;
}

// Status of the cooperative panel's bounded waits (LU mode), read after a synchronisation.
template <typename T> int consumer_status(const ConsumerLayout<T>& L, int type, hipStream_t s)
{
    int err = 0;
    if (type == LRN_LU_LINEAR_SOLVER) LRN_HIP(hipMemcpyAsync(&err, L.F.ws.err, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipStreamSynchronize(s));
    return err != 0 ? LRN_E_TIMEOUT : LRN_OK;
}

template <typename T>
int determinant_impl(T* m, int64_t* pivot, int64_t lda, int type, void* workspace, size_t ws_size, T* det,
    void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -3;
    if (!is_lu_type(type) && !is_chol_type(type)) return -4;
    if (type == LRN_LU_UPFRONT_LINEAR_SOLVER && pivot == nullptr) return -2;
    if (det == nullptr) return -7;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < det_ws_bytes<T>(lda, type)) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, ws_size);
    const auto L = carve_consumer<T>(ar, lda, type);
    const long long* piv = nullptr;
    LRN_TRY(consumer_factor<T>(m, pivot, lda, type, L, s, &piv));
    hipLaunchKernelGGL(lu_det_kernel<T>, dim3(1), dim3(256), 0, s, m, static_cast<long long>(lda),
        static_cast<long long>(lda), piv, L.info, L.result, is_chol_type(type) ? 1 : 0);
    LRN_HIP(hipGetLastError());
    T host = T(0);
    LRN_HIP(hipMemcpyAsync(&host, L.result, sizeof(T), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipStreamSynchronize(s));
    LRN_TRY(consumer_status<T>(L, type, s));
    *det = host;
    return LRN_OK;
}

// A^-1 in place: X = P I, L U X = X by the blocked strip solve and GEMM updates, then X is
// copied over m unless the factorisation found a zero pivot (m then keeps the factors,
// src/matrix_operation/cpu/native/inversion_native.cpp:421-436).
template <typename T>
int inversion_impl(T* m, int64_t* pivot, int64_t lda, int type, void* workspace, size_t ws_size, void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -3;
    if (!is_lu_type(type) && !is_chol_type(type)) return -4;
    if (type == LRN_LU_UPFRONT_LINEAR_SOLVER && pivot == nullptr) return -2;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || ws_size < inverse_ws_bytes<T>(lda, type)) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = lda;
    Arena ar(workspace, ws_size);
    const auto L = carve_consumer<T>(ar, n, type);
    const auto G = carve_getrs<T>(ar, n, 1);
    T* X = ar.next<T>(static_cast<size_t>(n) * static_cast<size_t>(n));
    const long long* piv = nullptr;
    const bool chol = is_chol_type(type);
    LRN_TRY(consumer_factor<T>(m, pivot, n, type, L, s, &piv, 2));
    const unsigned grid = static_cast<unsigned>(std::min<long long>((n * n + 255) / 256, 1ll << 16));
    if (chol) {
        // L L^T X = I with L^T mirrored into the upper triangle
        hipLaunchKernelGGL(eye_kernel<T>, dim3(grid), dim3(256), 0, s, X, n, n);
    } else {
        LRN_TRY(getrs_perm<T>(n, piv, G, s));
        hipLaunchKernelGGL(perm_eye_kernel<T>, dim3(grid), dim3(256), 0, s, X, n, n, G.perm);
    }
    LRN_HIP(hipGetLastError());
    LRN_TRY(blocked_lu_solve_all<T>(m, n, X, s, !chol));
    hipLaunchKernelGGL(copy_unless_kernel<T>, dim3(grid), dim3(256), 0, s, X, m, n * n, L.info,
        type == LRN_LU_LINEAR_SOLVER ? L.F.ws.err : nullptr);
    LRN_HIP(hipGetLastError());
    return consumer_status<T>(L, type, s);
}

template <typename T>
int debug_panel_impl(T* a, int64_t lda, int64_t k, int w, int64_t* pivot, void* ws, size_t ws_size, void* stream)
{
    if (ws_size < getrf_ws_bytes<T>(lda)) return -8;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(ws, ws_size);
    const auto L = carve_getrf<T>(ar);
    LRN_TRY(reset_panel_state(L.ws, s));
    unsigned base = 0;
    if (w <= 32) return launch_panel<T, 32>(a, lda, k, lda - k, w, reinterpret_cast<long long*>(pivot), L.ws, base, s);
    return launch_panel<T, kMaxW>(a, lda, k, lda - k, w, reinterpret_cast<long long*>(pivot), L.ws, base, s);
}


// ------------------------------------------------------------------- mixed-precision solve

// For the refinement's substitution modes 3 and 4 (lrn_set_refine_solver), from FP32 factors,
// in FP64: the inverted 64 x 64 diagonal blocks of unit L and of U, then their products with
// the off-diagonal blocks one and two blocks away (trsv_pred_kernel), in G.dinv
// (6 x trsm_dinv_elems(n), interleaved: L, U, L dist 1, U dist 1, L dist 2, U dist 2); nothing
// to do for the other modes. The factors are fixed during the refinement, so this runs
// once per factorisation.
int refine_inverses(const float* lu, long long n, const GetrsLayout<double>& G, hipStream_t s)
{
    const int chain = g_chain_solve.load();
    if ((chain != 3 && chain != 4) || G.dinv == nullptr) return LRN_OK;
    const long long e = trsm_dinv_elems(n);
    double* const dl = G.dinv;
    double* const du = G.dinv + e;
    LRN_TRY(trsm_inv_widen_launch(true, true, lu, n, n, dl, nullptr, s));
    LRN_TRY(trsm_inv_widen_launch(false, false, lu, n, n, du, nullptr, s));
    for (int dist = 1; dist <= 2; ++dist) {
        LRN_TRY(trsv_pred_widen_launch(true, dist, lu, n, n, dl, dl + 2 * dist * e, nullptr, s));
        LRN_TRY(trsv_pred_widen_launch(false, dist, lu, n, n, du, du + 2 * dist * e, nullptr, s));
    }
    return LRN_OK;
}

// Workspace of the stand-alone preconditioner (lrn_getrs_f32f64): the single-RHS buffers and
// room for the inverted diagonal blocks.
GetrsLayout<double> carve_f32f64(Arena& ar, long long n)
{
    auto L = carve_getrs<double>(ar, n, 1);
    L.dinv = ar.next<double>(static_cast<size_t>(6 * trsm_dinv_elems(n)));  // see refine_inverses
    return L;
}

size_t f32f64_ws_bytes(long long n)
{
    Arena ar;
    carve_f32f64(ar, n);
    return ar.off;
}

struct MixedLayout {
    float* a32;
    long long* piv;
    GetrfLayout<float> f;
    CaluWs<float> calu;     // tournament-panel buffers of the FP32 factorisation
    GetrsLayout<double> g;  // FP64 substitution buffers (the FP32 factors are applied in FP64)
    double* r;
    double* w;
    double* d;
    double* xbest;
    double* rowsum;
    double* V;     // (m + 1) vectors of length n
    double* h1;
    double* h2;
    double* scal;  // 8 scalars
    double* afb;   // FP64 copy for the fallback (null when not sized)
    long long* pivfb;
    GetrfLayout<double> ffb;
    GetrsLayout<double> gfb;
};

MixedLayout carve_mixed(Arena& ar, long long n, int m, bool fallback)
{
    MixedLayout L;
    L.a32 = ar.next<float>(static_cast<size_t>(n) * static_cast<size_t>(n));
    L.piv = ar.next<long long>(static_cast<size_t>(n));
    L.f = carve_getrf<float>(ar);
    L.calu.cand = ar.next<int>(static_cast<size_t>(kCaluMaxP) * kMaxW);
    L.calu.lu11 = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.top = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.uinv = ar.next<float>(static_cast<size_t>(kMaxW) * kMaxW);
    L.calu.ctr = L.f.ws.ctr == nullptr ? nullptr : L.f.ws.ctr + 32;
    L.calu.info = L.f.ws.info;
    L.calu.rowmap = nullptr;
    L.g = carve_getrs<double>(ar, n, 1);
    L.g.dinv = ar.next<double>(static_cast<size_t>(6 * trsm_dinv_elems(n)));  // see refine_inverses
    L.r = ar.next<double>(static_cast<size_t>(n));
    L.w = ar.next<double>(static_cast<size_t>(n));
    L.d = ar.next<double>(static_cast<size_t>(n));
    L.xbest = ar.next<double>(static_cast<size_t>(n));
    L.rowsum = ar.next<double>(static_cast<size_t>(n));
    L.V = ar.next<double>(static_cast<size_t>(m + 1) * static_cast<size_t>(n));
    L.h1 = ar.next<double>(static_cast<size_t>(m + 2));
    L.h2 = ar.next<double>(static_cast<size_t>(m + 2));
    L.scal = ar.next<double>(8);
    L.afb = nullptr;
    L.pivfb = nullptr;
    if (fallback) {
        L.afb = ar.next<double>(static_cast<size_t>(n) * static_cast<size_t>(n));
        L.pivfb = ar.next<long long>(static_cast<size_t>(n));
        L.ffb = carve_getrf<double>(ar);
        L.gfb = carve_getrs<double>(ar, n, 1);
    }
    return L;
}

size_t mixed_ws_bytes(long long n, int m, bool fallback)
{
    Arena ar;
    carve_mixed(ar, n, m, fallback);
    return ar.off;
}

unsigned grid_for(long long n, int threads = 256)
{
    return static_cast<unsigned>(std::max<long long>(1, std::min<long long>((n + threads - 1) / threads, 1 << 20)));
}

// Page-locked host words for the refinement's two norms and an event recorded after their copy,
// per thread and device, created on first use and kept until the process ends: the copy is
// then asynchronous, and the host waits for it alone rather than for the whole stream.
int norms_buffer(double** words, hipEvent_t* copied)
{
    struct Slot {
        double* p = nullptr;
        hipEvent_t ev = nullptr;
    };
    static thread_local std::vector<Slot> slots;
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    if (dev >= static_cast<int>(slots.size())) slots.resize(dev + 1);
    Slot& sl = slots[dev];
    if (sl.p == nullptr) {
        LRN_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.p), 2 * sizeof(double), hipHostMallocDefault));
        LRN_HIP(hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming));
    }
    *words = sl.p;
    *copied = sl.ev;
    return LRN_OK;
}

struct MixedSolver {
    const double* A;
    long long n;
    const double* b;
    double* x;
    const MixedLayout& L;
    hipStream_t s;
    double anorm = 0.0, bnorm = 0.0;

    int host_scalar(const double* dev, double* out)
    {
        LRN_HIP(hipMemcpyAsync(out, dev, sizeof(double), hipMemcpyDeviceToHost, s));
        LRN_HIP(hipStreamSynchronize(s));
        return LRN_OK;
    }
    int reduce(const double* u, const double* v, int op, double* out)
    {
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, u, v, n, op, L.scal);
        LRN_HIP(hipGetLastError());
        return host_scalar(L.scal, out);
    }
    // out = beta * bb + alpha * A v
    int gemv(const double* v, const double* bb, double* out, double alpha, double beta)
    {
        const bool vec = n % 2 == 0 && ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(v)) % 16 == 0);
        if (vec)
            hipLaunchKernelGGL(gemv_rows_vec_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, s, A, n, v, bb,
                out, alpha, beta);
        else
            hipLaunchKernelGGL(gemv_rows_kernel, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, s, A, n, v, bb,
                out, alpha, beta, 0);
        LRN_HIP(hipGetLastError());
        return LRN_OK;
    }
    // out = M^-1 in, with M the low-precision factors applied in FP64 arithmetic (the factor
    // entries are widened as they are read); in == out allowed
    int precond(const double* in, double* out)
    {
        if (in != out) LRN_TRY(copy(out, in));
        // L.g.err is sticky over the whole refinement and read once at the end
        return getrs_subst<double, float>(L.a32, n, out, 1, 1, L.g, nullptr, s, true);
    }
    // r = b - A x and the backward error (both norms fetched with one synchronisation)
    // next: also enqueue the next correction M^-1 r into L.d after the norms' copy, so the GPU
    // computes it while the host waits for the norms (page-locked words, an event after the copy)
    int residual(double* berr, bool next = false)
    {
        LRN_TRY(gemv(x, b, L.r, -1.0, 1.0));
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, L.r, nullptr, n, 0, L.scal);
        LRN_HIP(hipGetLastError());
        hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, x, nullptr, n, 0, L.scal + 1);
        LRN_HIP(hipGetLastError());
        double* norms = nullptr;
        hipEvent_t copied = nullptr;
        LRN_TRY(norms_buffer(&norms, &copied));
        LRN_HIP(hipMemcpyAsync(norms, L.scal, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
        LRN_HIP(hipEventRecord(copied, s));
        if (next) LRN_TRY(precond(L.r, L.d));
        LRN_HIP(hipEventSynchronize(copied));
        const double rn = norms[0], xn = norms[1];
        const double den = anorm * xn + bnorm;
        *berr = den > 0.0 ? rn / den : rn;
        return LRN_OK;
    }
    int axpby(double* y, const double* xx, double alpha, double beta)
    {
        hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, xx, n, alpha, beta, nullptr, nullptr);
        LRN_HIP(hipGetLastError());
        return LRN_OK;
    }
    int copy(double* dst, const double* src)
    {
        LRN_HIP(hipMemcpyAsync(dst, src, sizeof(double) * static_cast<size_t>(n), hipMemcpyDeviceToDevice, s));
        return LRN_OK;
    }

    // One restart cycle of left-preconditioned GMRES on A d = r (d returned in L.d), classical
    // Gram-Schmidt with one re-orthogonalisation, FP64 vectors and inner products.
    int gmres(int m, double rtol, int* iters)
    {
        const long long ldv = n;
        LRN_TRY(precond(L.r, L.w));
        double beta2 = 0.0;
        LRN_TRY(reduce(L.w, nullptr, 2, &beta2));
        const double beta = std::sqrt(beta2);
        LRN_TRY(axpby(L.d, nullptr, 0.0, 0.0));
        if (!(beta > 0.0)) return LRN_OK;
        std::vector<double> H(static_cast<size_t>(m + 1) * m, 0.0), cs(m, 0.0), sn(m, 0.0), g(m + 1, 0.0);
        g[0] = beta;
        // v0 = w / beta
        hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, L.V, L.w, n, 0.0, 1.0 / beta, nullptr,
            nullptr);
        LRN_HIP(hipGetLastError());
        int j = 0;
        std::vector<double> h(m + 2), h2(m + 2);
        for (; j < m; ++j) {
            double* vj = L.V + static_cast<long long>(j) * ldv;
            LRN_TRY(gemv(vj, nullptr, L.w, 1.0, 0.0));
            LRN_TRY(precond(L.w, L.w));
            for (int pass = 0; pass < 2; ++pass) {
                double* hd = pass == 0 ? L.h1 : L.h2;
                hipLaunchKernelGGL(multi_dot_kernel, dim3(j + 1), dim3(256), 0, s, L.V, ldv, L.w, n, hd);
                LRN_HIP(hipGetLastError());
                hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_for(n)), dim3(256), 0, s, L.V, ldv, hd, j + 1, L.w, n);
                LRN_HIP(hipGetLastError());
            }
            LRN_HIP(hipMemcpyAsync(h.data(), L.h1, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s));
            LRN_HIP(hipMemcpyAsync(h2.data(), L.h2, sizeof(double) * (j + 1), hipMemcpyDeviceToHost, s));
            double wn2 = 0.0;
            LRN_TRY(reduce(L.w, nullptr, 2, &wn2));  // synchronises the copies above too
            for (int i = 0; i <= j; ++i) h[i] += h2[i];
            h[j + 1] = std::sqrt(wn2);
            // apply the previous rotations, then a new one
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double rr = std::hypot(h[j], h[j + 1]);
            cs[j] = rr > 0.0 ? h[j] / rr : 1.0;
            sn[j] = rr > 0.0 ? h[j + 1] / rr : 0.0;
            h[j] = rr;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; ++i) H[static_cast<size_t>(i) * m + j] = h[i];
            const double hn = std::sqrt(wn2);
            ++*iters;
            if (hn > 0.0 && j + 1 < m + 1) {
                hipLaunchKernelGGL(axpby_kernel, dim3(grid_for(n)), dim3(256), 0, s, L.V + static_cast<long long>(j + 1) * ldv,
                    L.w, n, 0.0, 1.0 / hn, nullptr, nullptr);
                LRN_HIP(hipGetLastError());
            }
            if (std::fabs(g[j + 1]) <= rtol * beta || !(hn > 0.0)) {
                ++j;
                break;
            }
        }
        // y = H(0:j, 0:j)^-1 g(0:j); d = V y
        std::vector<double> y(j, 0.0);
        for (int i = j - 1; i >= 0; --i) {
            double t = g[i];
            for (int k = i + 1; k < j; ++k) t -= H[static_cast<size_t>(i) * m + k] * y[k];
            y[i] = t / H[static_cast<size_t>(i) * m + i];
        }
        for (auto& v : y) v = -v;
        if (j > 0) {
            LRN_HIP(hipMemcpyAsync(L.h1, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(multi_axpy_kernel, dim3(grid_for(n)), dim3(256), 0, s, L.V, ldv, L.h1, j, L.d, n);
            LRN_HIP(hipGetLastError());
            LRN_HIP(hipStreamSynchronize(s));  // y (host memory) must outlive the copy
        }
        return LRN_OK;
    }
};

// LRN_MIXED_VERBOSE=1 in the environment prints the backward error of every refinement
// step to stderr (diagnostics only).
bool mixed_verbose()
{
    static const bool v = [] {
        const char* e = std::getenv("LRN_MIXED_VERBOSE");
        return e != nullptr && e[0] == '1';
    }();
    return v;
}

int gesv_mixed_impl(const double* A, long long n, const double* b, double* x, double tol, int max_iter, int m,
    void* ws, size_t ws_size, lrn_mixed_stats* stats, hipStream_t s, int upd = -1)
{
    if (upd < 0 || upd >= kMixedUpdateModes) upd = g_mixed_update.load();
    const bool fallback = ws_size >= mixed_ws_bytes(n, m, true);
    if (ws_size < mixed_ws_bytes(n, m, false)) return -9;
    Arena ar(ws, ws_size);
    const MixedLayout L = carve_mixed(ar, n, m, fallback);
    lrn_mixed_stats st{};
    st.path = 3;
    hipEvent_t e0, e1, e2;
    LRN_HIP(hipEventCreate(&e0));
    LRN_HIP(hipEventCreate(&e1));
    LRN_HIP(hipEventCreate(&e2));
    struct EvGuard {
        hipEvent_t a, b, c;
        ~EvGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            (void)hipEventDestroy(c);
        }
    } guard{e0, e1, e2};
    LRN_HIP(hipEventRecord(e0, s));
    // A in FP32 and the row sums of |A| (for ||A||_inf) in one pass over A
    if (n % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(L.a32) % 8 == 0)
        hipLaunchKernelGGL(convert_rowsum_kernel<true>, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, s, A, n,
            L.a32, L.rowsum);
    else
        hipLaunchKernelGGL(convert_rowsum_kernel<false>, dim3(static_cast<unsigned>((n + 3) / 4)), dim3(256), 0, s, A,
            n, L.a32, L.rowsum);
    LRN_HIP(hipGetLastError());
    LRN_TRY((getrf_blocked<float, 32>(L.a32, n, L.piv, L.f.ws, s, upd, g_mixed_panel.load() ? &L.calu : nullptr)));
    LRN_TRY(getrs_perm<double>(n, L.piv, L.g, s));
    LRN_TRY(refine_inverses(L.a32, n, L.g, s));
    LRN_TRY(launch_fill(reinterpret_cast<unsigned*>(L.g.err), 0u, 64, s));
    LRN_HIP(hipEventRecord(e1, s));
    // ||A||_inf (from the conversion's row sums) and ||b||, fetched with info and err in one
    // synchronisation
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, L.rowsum, nullptr, n, 0, L.scal + 2);
    LRN_HIP(hipGetLastError());
    hipLaunchKernelGGL(reduce_kernel, dim3(1), dim3(1024), 0, s, b, nullptr, n, 0, L.scal + 3);
    LRN_HIP(hipGetLastError());
    int info = 0, err = 0;
    double norms[2] = {0.0, 0.0};
    LRN_HIP(hipMemcpyAsync(&info, L.f.ws.info, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipMemcpyAsync(&err, L.f.ws.err, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipMemcpyAsync(norms, L.scal + 2, sizeof(norms), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipStreamSynchronize(s));
    if (err != 0) return LRN_E_TIMEOUT;
    st.info = info;

    MixedSolver S{A, n, b, x, L, s};
    S.anorm = norms[0];
    S.bnorm = norms[1];

    double berr = 1.0 / 0.0;
    bool done = false;
    if (info == 0) {
        // LU-IR: x = M^-1 b, then x += M^-1 (b - A x) while the backward error keeps falling fast
        LRN_TRY(S.axpby(x, nullptr, 0.0, 0.0));
        LRN_TRY(S.copy(L.r, b));
        double best = 1.0 / 0.0;
        bool have_next = false;                            // L.d already holds the next correction
        double last = 1.0 / 0.0, before = 1.0 / 0.0;  // the last two backward errors
        for (int it = 0; it < max_iter; ++it) {
            if (!have_next) LRN_TRY(S.precond(L.r, L.d));
            LRN_TRY(S.axpby(x, L.d, 1.0, 1.0));
            ++st.ir_iterations;
            // the next correction goes beside the host's wait for the backward error when the
            // last two steps' rate predicts that this one stays above the tolerance (it is
            // computed from the same residual either way, so the iterates do not change; a
            // wasted one costs a pair of substitutions)
            const bool next = it >= 2 && g_refinement.load() != 1 && last * (last / before) > tol;
            LRN_TRY(S.residual(&berr, next));
            have_next = next;
            before = last;
            last = berr;
            if (mixed_verbose()) std::fprintf(stderr, "lrn mixed n=%lld lu-ir %d berr %.3e\n", n, it, berr);
            const double prev_best = best;
            if (berr < best) {
                LRN_TRY(S.copy(L.xbest, x));
                best = berr;
            }
            if (berr <= tol) {
                st.path = 0;
                done = true;
                break;
            }
            // stagnating (less than a 2x gain over the best so far) or diverging: GMRES-IR
            if (it >= 1 && !(berr <= 0.5 * prev_best)) break;
            if (g_refinement.load() == 1) break;  // GMRES-IR from the first iterate
        }
        if (!done) {
            // restart GMRES-IR from the best LU-IR iterate
            LRN_TRY(S.copy(x, L.xbest));
            LRN_TRY(S.residual(&berr));
            double gbest = berr;
            for (int outer = 0; outer < max_iter && !done; ++outer) {
                // reduce the preconditioned residual by what the target asks for, with a
                // factor 10 of margin (within [1e-12, 1e-2])
                const double rtol = std::min(1e-2, std::max(1e-12, 0.1 * tol / berr));
                LRN_TRY(S.gmres(m, rtol, &st.gmres_iterations));
                LRN_TRY(S.axpby(x, L.d, 1.0, 1.0));
                ++st.gmres_outer;
                LRN_TRY(S.residual(&berr));
                if (mixed_verbose())
                    std::fprintf(stderr, "lrn mixed n=%lld gmres-ir %d (inner %d) berr %.3e\n", n, outer,
                        st.gmres_iterations, berr);
                if (berr <= tol) {
                    st.path = 1;
                    done = true;
                    break;
                }
                const double prev = gbest;
                if (berr < gbest) {
                    LRN_TRY(S.copy(L.xbest, x));
                    gbest = berr;
                }
                // a restart cycle that does not halve the best backward error: GMRES-IR has
                // stalled (the low-precision factors are too far from A); stop here
                if (outer >= 1 && !(berr <= 0.5 * prev)) break;
            }
            if (!done) {
                // leave the best iterate in x
                LRN_TRY(S.copy(x, L.xbest));
                berr = gbest;
            }
        }
    }
    int fb_info = 0, fb_err = 0;
    if (!done && fallback) {
        LRN_HIP(hipMemcpyAsync(L.afb, A, sizeof(double) * static_cast<size_t>(n * n), hipMemcpyDeviceToDevice, s));
        LRN_TRY(getrf_dispatch<double>(L.afb, n, L.pivfb, L.ffb.ws, s));
        LRN_TRY(S.copy(x, b));
        LRN_TRY(launch_gate<double>(L.ffb.ws.info, L.ffb.ws.err, L.gfb.gate, x, n, 1, 1, s));
        LRN_TRY(getrs_run<double>(L.afb, n, L.pivfb, x, 1, 1, L.gfb, L.gfb.gate, s));
        LRN_TRY(S.residual(&berr));
        LRN_HIP(hipMemcpyAsync(&fb_info, L.ffb.ws.info, sizeof(int), hipMemcpyDeviceToHost, s));
        LRN_HIP(hipMemcpyAsync(&fb_err, L.ffb.ws.err, sizeof(int), hipMemcpyDeviceToHost, s));
        st.path = 2;
    }
    int subst_err = 0;
    LRN_HIP(hipMemcpyAsync(&subst_err, L.g.err, sizeof(int), hipMemcpyDeviceToHost, s));
    LRN_HIP(hipEventRecord(e2, s));
    LRN_HIP(hipEventSynchronize(e2));
    // an expired bounded wait anywhere (refinement substitutions, fallback factorisation or
    // its solve) invalidates the result
    if (subst_err != 0 || fb_err != 0) return LRN_E_TIMEOUT;
    if (st.path == 2) st.info = fb_info;  // the status of the factors that produced x
    float f_ms = 0.f, r_ms = 0.f;
    LRN_HIP(hipEventElapsedTime(&f_ms, e0, e1));
    LRN_HIP(hipEventElapsedTime(&r_ms, e1, e2));
    st.factor_ms = f_ms;
    st.refine_ms = r_ms;
    st.backward_error = berr;
    if (stats) *stats = st;
    return LRN_OK;
}

// ------------------------------------------------------------------- batched solve

struct BatchLayout {
    int* order;
    unsigned* ctr;
};

BatchLayout carve_batch(Arena& ar, long long count)
{
    BatchLayout L;
    L.order = ar.next<int>(static_cast<size_t>(std::max<long long>(count, 1)));
    L.ctr = ar.next<unsigned>(64);
    return L;
}

size_t batch_ws_bytes(long long count)
{
    Arena ar;
    carve_batch(ar, count);
    return ar.off;
}

std::atomic<int> g_batch_grid{0};  // workgroups of the persistent launch (0 = occupancy x CUs)

template <typename T>
int batched_gesv_impl(T* a, const int64_t* a_off, int64_t* piv, const int64_t* piv_off, T* b, const int64_t* b_off,
    const int32_t* n, const int32_t* nrhs, int64_t count, int32_t* info, void* ws, size_t ws_size, hipStream_t s)
{
    if (count == 0) return LRN_OK;
    Arena ar(ws, ws_size);
    const BatchLayout L = carve_batch(ar, count);
    return batched_solve_launch<T>(a, reinterpret_cast<const long long*>(a_off), reinterpret_cast<long long*>(piv),
        reinterpret_cast<const long long*>(piv_off), b, reinterpret_cast<const long long*>(b_off), n, nrhs, info,
        L.order, L.ctr, static_cast<int>(count), g_batch_grid.load(), s);
}

}  // namespace
}  // namespace lrn

// ------------------------------------------------------------------- allocation-free solve
// Scratch of lrn_linear_solver_noalloc_f64: a file-scope device array that takes the place of
// the caller's workspace (each GPU holds its own copy, resident until the process ends). In
// host code the array's name is only a stand-in for the symbol, never a device address: the
// address comes from hipGetSymbolAddress, once per device with that device current, and only
// that address is handed to kernels.
constexpr size_t kNoAllocScratchBytes = static_cast<size_t>(4) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace lrn {
namespace {

constexpr int kNoAllocMaxDevices = 64;
std::atomic<void*> g_scratch_ptr[kNoAllocMaxDevices];

// Device address of this device's scratch (looked up on the device's first call).
int noalloc_scratch(void** out)
{
    int dev = 0;
    LRN_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) return LRN_E_UNSUPPORTED;
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (p == nullptr) {
        static std::mutex mu;
        std::lock_guard<std::mutex> g(mu);
        p = g_scratch_ptr[dev].load(std::memory_order_acquire);
        if (p == nullptr) {
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return LRN_E_HIP;
            g_scratch_ptr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return LRN_OK;
}

}  // namespace
}  // namespace lrn

using namespace lrn;

extern "C" {

LRN_API const char* lrn_status_string(int status)
{
    switch (status) {
    case LRN_OK: return "ok";
    case LRN_E_HIP: return "HIP runtime call failed";
    case LRN_E_TIMEOUT: return "bounded inter-workgroup wait expired";
    case LRN_E_SCRATCH: return "order too large for the allocation-free entry point's 4 MiB device scratch";
    case LRN_E_UNSUPPORTED: return "unsupported size or mode";
    case LRN_E_INTERNAL: return "internal error";
    default: return status < 0 ? "invalid argument" : "unknown status";
    }
}

LRN_API const char* lrn_version(void) { return "linalg_rocm_native 0.1.0 (gfx950)"; }

LRN_API int lrn_lu_decomposition_workspace_size_f64(int64_t lda, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = getrf_ws_bytes<double>(lda);
    return LRN_OK;
}
LRN_API int lrn_lu_decomposition_workspace_size_f32(int64_t lda, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = getrf_ws_bytes<float>(lda);
    return LRN_OK;
}
LRN_API int lrn_lu_decomposition_f64(double* m, int64_t lda, int64_t* pivot, void* workspace, size_t workspace_size,
    int* info, void* stream)
{
    return lu_decomposition_impl<double>(m, lda, pivot, workspace, workspace_size, info, stream);
}
LRN_API int lrn_lu_decomposition_f32(float* m, int64_t lda, int64_t* pivot, void* workspace, size_t workspace_size,
    int* info, void* stream)
{
    return lu_decomposition_impl<float>(m, lda, pivot, workspace, workspace_size, info, stream);
}

LRN_API int lrn_getrf_async_f64(double* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream)
{
    return getrf_async_impl<double>(m, lda, pivot, info_dev, workspace, workspace_size, stream);
}
LRN_API int lrn_getrf_async_f32(float* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream)
{
    return getrf_async_impl<float>(m, lda, pivot, info_dev, workspace, workspace_size, stream);
}

LRN_API int lrn_getrf_tournament_workspace_size_f32(int64_t lda, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = tournament_ws_bytes();
    return LRN_OK;
}

LRN_API int lrn_getrf_tournament_f32(float* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream)
{
    return getrf_tournament_impl(m, lda, pivot, info_dev, workspace, workspace_size, stream);
}

LRN_API int lrn_linear_solver_workspace_size_f64(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -3;
    *bytes = solver_ws_bytes<double>(lda, 1, type);
    return LRN_OK;
}
LRN_API int lrn_linear_solver_workspace_size_f32(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -3;
    *bytes = solver_ws_bytes<float>(lda, 1, type);
    return LRN_OK;
}
LRN_API int lrn_linear_solver_f64(double* m, int64_t* pivot, int64_t lda, double* x, int type, void* workspace,
    size_t workspace_size, void* stream)
{
    return solver_impl<double>(m, pivot, lda, x, 1, 1, type, workspace, workspace_size, stream);
}
LRN_API int lrn_linear_solver_noalloc_f64(double* m, int64_t* pivot, int64_t lda, double* x, int type, void* stream)
{
    if (m == nullptr) return -1;
    if (lda <= 0) return -3;
    if (x == nullptr) return -4;
    if (type < LRN_LU_LINEAR_SOLVER || type > LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER) return -5;
    if (pivot == nullptr && is_lu_type(type)) return -2;
    if (lda > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (solver_ws_bytes<double>(lda, 1, type) > kNoAllocScratchBytes) return LRN_E_SCRATCH;
    void* ws = nullptr;
    LRN_TRY(noalloc_scratch(&ws));
    return solver_impl<double>(m, pivot, lda, x, 1, 1, type, ws, kNoAllocScratchBytes, stream, true);
}

LRN_API int lrn_linear_solver_noalloc_max_order(int type, int64_t* order)
{
    if (type < LRN_LU_LINEAR_SOLVER || type > LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER) return -1;
    if (order == nullptr) return -2;
    long long lo = 0, hi = kMaxOrder;  // largest order whose workspace fits the scratch
    while (lo < hi) {
        const long long mid = (lo + hi + 1) / 2;
        if (solver_ws_bytes<double>(mid, 1, type) <= kNoAllocScratchBytes)
            lo = mid;
        else
            hi = mid - 1;
    }
    *order = lo;
    return LRN_OK;
}

LRN_API int lrn_linear_solver_f32(float* m, int64_t* pivot, int64_t lda, float* x, int type, void* workspace,
    size_t workspace_size, void* stream)
{
    return solver_impl<float>(m, pivot, lda, x, 1, 1, type, workspace, workspace_size, stream);
}

LRN_API int lrn_triangular_solver_workspace_size_f64(int64_t n, int64_t nrhs, size_t* bytes)
{
    if (n <= 0) return -1;
    if (nrhs <= 0) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::trsm_ws_bytes<double>(n, nrhs);
    return LRN_OK;
}
LRN_API int lrn_triangular_solver_workspace_size_f32(int64_t n, int64_t nrhs, size_t* bytes)
{
    if (n <= 0) return -1;
    if (nrhs <= 0) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::trsm_ws_bytes<float>(n, nrhs);
    return LRN_OK;
}
LRN_API int lrn_triangular_solver_f64(int upper, int transpose, int unit, int64_t n, int64_t nrhs, const double* a,
    int64_t lda, const double* b, int64_t ldb, double* x, int64_t ldx, void* workspace, size_t workspace_size, void* stream)
{
    return lrn::triangular_solver_impl<double>(upper, transpose, unit, n, nrhs, a, lda, b, ldb, x, ldx, workspace,
        workspace_size, static_cast<hipStream_t>(stream));
}
LRN_API int lrn_triangular_solver_f32(int upper, int transpose, int unit, int64_t n, int64_t nrhs, const float* a,
    int64_t lda, const float* b, int64_t ldb, float* x, int64_t ldx, void* workspace, size_t workspace_size, void* stream)
{
    return lrn::triangular_solver_impl<float>(upper, transpose, unit, n, nrhs, a, lda, b, ldb, x, ldx, workspace,
        workspace_size, static_cast<hipStream_t>(stream));
}

LRN_API int lrn_linear_solver_multi_workspace_size_f64(int64_t lda, int64_t nrhs, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (nrhs < 0) return -2;
    if (bytes == nullptr) return -4;
    *bytes = solver_ws_bytes<double>(lda, std::max<int64_t>(nrhs, 1), type);
    return LRN_OK;
}
LRN_API int lrn_linear_solver_multi_workspace_size_f32(int64_t lda, int64_t nrhs, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (nrhs < 0) return -2;
    if (bytes == nullptr) return -4;
    *bytes = solver_ws_bytes<float>(lda, std::max<int64_t>(nrhs, 1), type);
    return LRN_OK;
}
LRN_API int lrn_linear_solver_multi_f64(double* m, int64_t* pivot, int64_t lda, double* b, int64_t nrhs, int64_t ldb,
    int type, void* workspace, size_t workspace_size, void* stream)
{
    return solver_impl<double>(m, pivot, lda, b, nrhs, ldb, type, workspace, workspace_size, stream);
}
LRN_API int lrn_linear_solver_multi_f32(float* m, int64_t* pivot, int64_t lda, float* b, int64_t nrhs, int64_t ldb,
    int type, void* workspace, size_t workspace_size, void* stream)
{
    return solver_impl<float>(m, pivot, lda, b, nrhs, ldb, type, workspace, workspace_size, stream);
}

LRN_API int lrn_batched_matrix_multiplication_f64(int64_t dim, int64_t count, const double* a, const double* b,
    double* c, void* stream)
{
    return lrn::batched_mm_impl<double>(dim, count, a, b, c, stream);
}
LRN_API int lrn_batched_matrix_multiplication_f32(int64_t dim, int64_t count, const float* a, const float* b,
    float* c, void* stream)
{
    return lrn::batched_mm_impl<float>(dim, count, a, b, c, stream);
}

LRN_API int lrn_cholesky_decomposition_workspace_size_f64(int64_t lda, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = lrn::chol_ws_bytes<double>(lda);
    return LRN_OK;
}
LRN_API int lrn_cholesky_decomposition_workspace_size_f32(int64_t lda, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = lrn::chol_ws_bytes<float>(lda);
    return LRN_OK;
}
LRN_API int lrn_cholesky_decomposition_f64(double* c, int64_t lda, int upper, void* workspace, size_t workspace_size,
    int* info, void* stream)
{
    return lrn::cholesky_impl<double>(c, lda, upper, workspace, workspace_size, info, stream);
}
LRN_API int lrn_cholesky_decomposition_f32(float* c, int64_t lda, int upper, void* workspace, size_t workspace_size,
    int* info, void* stream)
{
    return lrn::cholesky_impl<float>(c, lda, upper, workspace, workspace_size, info, stream);
}

LRN_API int lrn_cholesky_decomposition_async_f64(double* c, int64_t lda, int upper, void* workspace,
    size_t workspace_size, int* info_dev, void* stream)
{
    return cholesky_async_impl<double>(c, lda, upper, workspace, workspace_size, info_dev, stream);
}
LRN_API int lrn_cholesky_decomposition_async_f32(float* c, int64_t lda, int upper, void* workspace,
    size_t workspace_size, int* info_dev, void* stream)
{
    return cholesky_async_impl<float>(c, lda, upper, workspace, workspace_size, info_dev, stream);
}

LRN_API int lrn_determinant_workspace_size_f64(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (!lrn::is_lu_type(type) && !lrn::is_chol_type(type)) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::det_ws_bytes<double>(lda, type);
    return LRN_OK;
}
LRN_API int lrn_determinant_workspace_size_f32(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (!lrn::is_lu_type(type) && !lrn::is_chol_type(type)) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::det_ws_bytes<float>(lda, type);
    return LRN_OK;
}
LRN_API int lrn_determinant_f64(double* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, double* det, void* stream)
{
    return lrn::determinant_impl<double>(m, pivot, lda, type, workspace, workspace_size, det, stream);
}
LRN_API int lrn_determinant_f32(float* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, float* det, void* stream)
{
    return lrn::determinant_impl<float>(m, pivot, lda, type, workspace, workspace_size, det, stream);
}

LRN_API int lrn_inversion_workspace_size_f64(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (!lrn::is_lu_type(type) && !lrn::is_chol_type(type)) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::inverse_ws_bytes<double>(lda, type);
    return LRN_OK;
}
LRN_API int lrn_inversion_workspace_size_f32(int64_t lda, int type, size_t* bytes)
{
    if (lda <= 0) return -1;
    if (!lrn::is_lu_type(type) && !lrn::is_chol_type(type)) return -2;
    if (bytes == nullptr) return -3;
    *bytes = lrn::inverse_ws_bytes<float>(lda, type);
    return LRN_OK;
}
LRN_API int lrn_inversion_f64(double* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, void* stream)
{
    return lrn::inversion_impl<double>(m, pivot, lda, type, workspace, workspace_size, stream);
}
LRN_API int lrn_inversion_f32(float* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, void* stream)
{
    return lrn::inversion_impl<float>(m, pivot, lda, type, workspace, workspace_size, stream);
}

LRN_API int lrn_getrs_f32f64_workspace_size(int64_t n, size_t* bytes)
{
    if (n <= 0) return -1;
    if (bytes == nullptr) return -2;
    *bytes = f32f64_ws_bytes(n);
    return LRN_OK;
}

LRN_API int lrn_getrs_f32f64(const float* lu, int64_t n, const int64_t* pivot, double* b, void* workspace,
    size_t workspace_size, void* stream)
{
    if (lu == nullptr) return -1;
    if (n <= 0) return -2;
    if (pivot == nullptr) return -3;
    if (b == nullptr) return -4;
    if (n > kMaxOrder) return LRN_E_UNSUPPORTED;
    if (workspace == nullptr || workspace_size < f32f64_ws_bytes(n)) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Arena ar(workspace, workspace_size);
    const auto L = carve_f32f64(ar, n);
    LRN_TRY(getrs_perm<double>(n, reinterpret_cast<const long long*>(pivot), L, s));
    LRN_TRY(refine_inverses(lu, n, L, s));
    return getrs_subst<double, float>(lu, n, b, 1, 1, L, nullptr, s);
}

LRN_API int lrn_debug_gemm_sub(int dtype, const void* a, int64_t lda, const void* b, int64_t ldb, void* c,
    int64_t ldc, int64_t m, int64_t n, int64_t k, void* stream)
{
    if (a == nullptr || b == nullptr || c == nullptr) return -2;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == 0)
        return launch_gemm_sub<double>(static_cast<const double*>(a), lda, static_cast<const double*>(b), ldb,
            static_cast<double*>(c), ldc, m, n, k, s);
    if (dtype == 1)
        return launch_gemm_sub<float>(static_cast<const float*>(a), lda, static_cast<const float*>(b), ldb,
            static_cast<float*>(c), ldc, m, n, k, s);
    if (dtype >= 2 && dtype <= 4)
        return launch_gemm_sub_bf16(static_cast<const float*>(a), lda, static_cast<const float*>(b), ldb,
            static_cast<float*>(c), ldc, m, n, k, s, dtype - 1);
    if (dtype >= 5 && dtype <= 7)
        return gemm_sub_fp8_launch(static_cast<const float*>(a), lda, static_cast<const float*>(b), ldb,
            static_cast<float*>(c), ldc, m, n, k, dtype - 4, s, nullptr);
    return -1;
}

LRN_API int lrn_debug_trsm(int dtype, void* a, int64_t lda, int64_t k, int w, int64_t c0, int64_t c1, void* stream)
{
    if (a == nullptr) return -2;
    if (w <= 0 || w > kMaxW) return -5;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (dtype == 0) return launch_trsm<double, kMaxW>(static_cast<double*>(a), lda, k, w, c0, c1, s);
    if (dtype == 1) return launch_trsm<float, kMaxW>(static_cast<float*>(a), lda, k, w, c0, c1, s);
    return -1;
}

LRN_API int lrn_debug_panel(int dtype, void* a, int64_t lda, int64_t k, int w, int64_t* pivot, void* workspace,
    size_t workspace_size, void* stream)
{
    if (a == nullptr || pivot == nullptr || workspace == nullptr) return -2;
    if (w <= 0 || w > kMaxW || k < 0 || k + w > lda) return -5;
    if (dtype == 0) return debug_panel_impl<double>(static_cast<double*>(a), lda, k, w, pivot, workspace, workspace_size, stream);
    if (dtype == 1) return debug_panel_impl<float>(static_cast<float*>(a), lda, k, w, pivot, workspace, workspace_size, stream);
    return -1;
}

LRN_API int lrn_gesv_mixed_workspace_size_f64(int64_t n, int gmres_restart, int fallback, size_t* bytes)
{
    if (n <= 0) return -1;
    if (gmres_restart <= 0 || gmres_restart > 200) return -2;
    if (bytes == nullptr) return -4;
    *bytes = mixed_ws_bytes(n, gmres_restart, fallback != 0);
    return LRN_OK;
}

LRN_API int lrn_gesv_mixed_f64(const double* a, int64_t n, const double* b, double* x, double tol, int max_iter,
    int gmres_restart, void* workspace, size_t workspace_size, lrn_mixed_stats* stats, void* stream)
{
    if (a == nullptr) return -1;
    if (n <= 0) return -2;
    if (b == nullptr) return -3;
    if (x == nullptr) return -4;
    if (!(tol > 0.0)) return -5;
    if (max_iter <= 0) return -6;
    if (gmres_restart <= 0 || gmres_restart > 200) return -7;
    if (workspace == nullptr) return -8;
    if (n > kMaxOrder) return LRN_E_UNSUPPORTED;
    return gesv_mixed_impl(a, n, b, x, tol, max_iter, gmres_restart, workspace, workspace_size, stats,
        static_cast<hipStream_t>(stream));
}

LRN_API int lrn_gesv_mixed_batched_f64(const double* a, const int64_t* a_offset, const int64_t* n, const double* b,
    const int64_t* b_offset, double* x, const int64_t* x_offset, int64_t count, const int32_t* update_mode, double tol,
    int max_iter, int gmres_restart, void* workspace, size_t workspace_size, lrn_mixed_stats* stats, void* stream)
{
    if (count < 0) return -8;
    if (count == 0) return LRN_OK;
    if (a == nullptr || a_offset == nullptr) return -1;
    if (n == nullptr) return -3;
    if (b == nullptr || b_offset == nullptr) return -4;
    if (x == nullptr || x_offset == nullptr) return -6;
    if (!(tol > 0.0)) return -10;
    if (max_iter <= 0) return -11;
    if (gmres_restart <= 0 || gmres_restart > 200) return -12;
    if (workspace == nullptr) return -13;
    if (stats == nullptr) return -15;
    for (int64_t i = 0; i < count; ++i)
        if (n[i] <= 0 || n[i] > kMaxOrder || (update_mode != nullptr && (update_mode[i] < 0 || update_mode[i] >= kMixedUpdateModes)))
            return -3;
    for (int64_t i = 0; i < count; ++i) {
        const int st = gesv_mixed_impl(a + a_offset[i], n[i], b + b_offset[i], x + x_offset[i], tol, max_iter,
            gmres_restart, workspace, workspace_size, &stats[i], static_cast<hipStream_t>(stream),
            update_mode != nullptr ? update_mode[i] : -1);
        if (st != LRN_OK) return st;
    }
    return LRN_OK;
}

LRN_API int lrn_batched_gesv_workspace_size(int64_t count, size_t* bytes)
{
    if (count < 0 || count > (1ll << 30)) return -1;
    if (bytes == nullptr) return -2;
    *bytes = batch_ws_bytes(count);
    return LRN_OK;
}

#define LRN_BATCHED_ARGS_CHECK()                                                                   \
    do {                                                                                           \
        if (count < 0 || count > (1ll << 30)) return -9;                                           \
        if (count == 0) return LRN_OK;                                                             \
        if (a == nullptr || a_offset == nullptr) return -1;                                        \
        if (pivot == nullptr || pivot_offset == nullptr) return -3;                                \
        if (b == nullptr || b_offset == nullptr) return -5;                                        \
        if (n == nullptr) return -7;                                                               \
        if (nrhs == nullptr) return -8;                                                            \
        if (info == nullptr) return -10;                                                           \
        if (workspace == nullptr || workspace_size < batch_ws_bytes(count)) return -11;            \
    } while (0)

LRN_API int lrn_batched_gesv_f64(double* a, const int64_t* a_offset, int64_t* pivot, const int64_t* pivot_offset,
    double* b, const int64_t* b_offset, const int32_t* n, const int32_t* nrhs, int64_t count, int32_t* info,
    void* workspace, size_t workspace_size, void* stream)
{
    LRN_BATCHED_ARGS_CHECK();
    return batched_gesv_impl<double>(a, a_offset, pivot, pivot_offset, b, b_offset, n, nrhs, count, info, workspace,
        workspace_size, static_cast<hipStream_t>(stream));
}

LRN_API int lrn_batched_gesv_f32(float* a, const int64_t* a_offset, int64_t* pivot, const int64_t* pivot_offset,
    float* b, const int64_t* b_offset, const int32_t* n, const int32_t* nrhs, int64_t count, int32_t* info,
    void* workspace, size_t workspace_size, void* stream)
{
    LRN_BATCHED_ARGS_CHECK();
    return batched_gesv_impl<float>(a, a_offset, pivot, pivot_offset, b, b_offset, n, nrhs, count, info, workspace,
        workspace_size, static_cast<hipStream_t>(stream));
}

LRN_API int lrn_set_batched_grid(int workgroups)
{
    if (workgroups < 0) return -1;
    g_batch_grid.store(workgroups);
    return LRN_OK;
}

LRN_API int lrn_trace_start(unsigned kinds, int64_t capacity)
{
    if (capacity <= 0 || capacity > (1ll << 22)) return -2;
    return trace_start(kinds, static_cast<long>(capacity)) == 0 ? LRN_OK : LRN_E_HIP;
}

LRN_API int lrn_trace_stop(void)
{
    trace_stop();
    return LRN_OK;
}

LRN_API int lrn_trace_read(int kind, int64_t* launches, double* milliseconds, double* work)
{
    if (kind < 0 || kind >= kTrKinds) return -1;
    if (launches == nullptr || milliseconds == nullptr || work == nullptr) return -2;
    long long c = 0;
    if (trace_read(kind, &c, milliseconds, work) != 0) return LRN_E_HIP;
    *launches = c;
    return LRN_OK;
}

LRN_API int lrn_set_mixed_update(int mode)
{
    if (mode < 0 || mode >= kMixedUpdateModes) return -1;
    g_mixed_update.store(mode);
    return LRN_OK;
}

LRN_API int lrn_set_mixed_panel(int tournament)
{
    g_mixed_panel.store(tournament ? 1 : 0);
    return LRN_OK;
}

LRN_API int lrn_set_panel_kernel(int kind)
{
    if (kind < 0 || kind > 1) return -1;
    g_panel_kind.store(kind);
    return LRN_OK;
}

LRN_API int lrn_debug_set_spin_limit(int64_t limit)
{
    if (limit > 0x7fffffffll) return -1;
    g_spin_limit.store(limit < 0 ? (1u << 20) : static_cast<unsigned>(limit));
    return LRN_OK;
}

LRN_API int lrn_set_refine_solver(int chain)
{
    if (chain < 0 || chain > 4) return -1;
    g_chain_solve.store(chain);
    return LRN_OK;
}

LRN_API int lrn_set_refinement(int mode)
{
    if (mode < 0 || mode > 1) return -1;
    g_refinement.store(mode);
    return LRN_OK;
}

LRN_API int lrn_set_multi_rhs_solver(int min_rhs)
{
    if (min_rhs < 0) return -1;
    g_multi_rhs_min.store(min_rhs);
    return LRN_OK;
}

LRN_API int lrn_set_multi_rhs_kernel(int kind)
{
    if (kind < 0 || kind > 1) return -1;
    g_multi_rhs_kernel.store(kind);
    return LRN_OK;
}

LRN_API int lrn_set_block_update(int kind)
{
    if (kind < 0 || kind > 2) return -1;
    g_block_update.store(kind);
    return LRN_OK;
}

LRN_API int lrn_set_lookahead(int mode)
{
    if (mode < 0 || (mode > 1 && (mode < kXcds || mode % kXcds != 0))) return -1;
    g_lookahead.store(mode);
    return LRN_OK;
}

LRN_API int lrn_set_tuning(int outer_nb, int panel_w, int panel_max_wg)
{
    if (outer_nb > 0 && (outer_nb % 64) != 0) return -1;
    if (panel_w > 0 && panel_w != 32 && panel_w != 64) return -2;
    if (panel_max_wg > kPanelMaxWG) return -3;
    g_outer_nb.store(outer_nb > 0 ? outer_nb : 0);
    g_panel_w.store(panel_w > 0 ? panel_w : 0);
    g_panel_max_wg.store(panel_max_wg > 0 ? panel_max_wg : 0);
    return LRN_OK;
}
LRN_API int lrn_get_tuning(int* outer_nb, int* panel_w, int* panel_max_wg)
{
    if (outer_nb) *outer_nb = tuned_outer_nb(1 << 20);
    if (panel_w) *panel_w = tuned_panel_w();
    if (panel_max_wg) *panel_max_wg = tuned_panel_max_wg();
    return LRN_OK;
}

}  // extern "C"
