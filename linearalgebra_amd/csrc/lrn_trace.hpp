// Optional per-launch timing of the library's kernels with HIP events, for live roofline
// figures: while tracing is on, every launch of a traced kind is bracketed by two events
// recorded on the stream the kernel is launched on, and its algorithmic work (flops for
// the GEMMs, bytes for the memory-bound kernels) is logged beside them. Reading the trace
// synchronises on the events and sums per kind. Off by default; when off a launch pays one
// relaxed atomic load.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace lrn {

enum TraceKind {
    kTrGemmF64 = 0,      // gemm_sub_f64_kernel<false>  (interior tiles), flops
    kTrGemmF64Edge = 1,  // gemm_sub_f64_kernel<true>   (ragged strips), flops
    kTrGemmBf16x3 = 2,   // gemm_sub_bf16_kernel<2>  (two planes), FP32-equivalent flops 2MNK
    kTrGemmBf16 = 3,     // gemm_sub_bf16_kernel<1>, flops
    kTrGemmF32 = 4,      // gemm_sub_f32_kernel, flops
    kTrPanel = 5,        // panel factorisation, flops of the panel (rank-1 updates)
    kTrLaswp = 6,        // row interchanges, bytes (every pivot taken as an interchange)
    kTrTrsm = 7,         // unit-lower TRSM of U12 rows, bytes
    kTrTrsv = 8,         // forward / backward substitution of the solve, bytes
    kTrGemmBf16x6 = 9,   // gemm_sub_bf16_kernel<3>  (three planes), FP32-equivalent flops 2MNK
    kTrBatched = 10,     // batched_gesv_kernel (work not known on the host: 0)
    kTrBcUpdate = 11,    // update inside a block column (bc_update_kernel / 64x64 gemm_sub_kernel), flops
    kTrGemmFp8 = 12,     // gemm_sub_fp8_kernel<1> (one block-scaled e4m3 plane), flops 2MNK
    kTrGemmFp8x3 = 13,   // gemm_sub_fp8_kernel<2> (two planes, three products), FP32-equivalent flops 2MNK
    kTrGemmFp8x6 = 14,   // gemm_sub_fp8_kernel<3> (three planes, six products), FP32-equivalent flops 2MNK
    kTrKinds = 15
};

struct TraceRec {
    int kind;
    double work;
    hipEvent_t a, b;
};

struct Tracer {
    std::atomic<unsigned> mask{0};
    std::mutex mu;
    std::vector<hipEvent_t> pool;
    std::vector<TraceRec> recs;
    size_t used = 0;  // events handed out
};

inline Tracer& tracer()
{
    static Tracer t;
    return t;
}

// Marks the start of a traced launch; returns the record index, or -1 if the kind is not
// traced or the event pool is exhausted (the launch then simply goes untimed).
inline long trace_begin(int kind, double work, hipStream_t s)
{
    Tracer& t = tracer();
    if (!(t.mask.load(std::memory_order_relaxed) & (1u << kind))) return -1;
    std::lock_guard<std::mutex> g(t.mu);
    if (t.used + 2 > t.pool.size()) return -1;
    TraceRec r{kind, work, t.pool[t.used], t.pool[t.used + 1]};
    t.used += 2;
    if (hipEventRecord(r.a, s) != hipSuccess) return -1;
    t.recs.push_back(r);
    return static_cast<long>(t.recs.size() - 1);
}

inline void trace_end(long idx, hipStream_t s)
{
    if (idx < 0) return;
    Tracer& t = tracer();
    std::lock_guard<std::mutex> g(t.mu);
    (void)hipEventRecord(t.recs[static_cast<size_t>(idx)].b, s);
}

// The event pool is released by an exit handler registered after the HIP runtime has
// started, so it runs before the runtime's own teardown.
inline void trace_release()
{
    Tracer& t = tracer();
    std::lock_guard<std::mutex> g(t.mu);
    t.mask.store(0);
    for (hipEvent_t e : t.pool) (void)hipEventDestroy(e);
    t.pool.clear();
    t.recs.clear();
    t.used = 0;
}

// Starts a trace of the kinds in `mask` with room for `capacity` launches (events are
// created here, outside any timed region).
inline int trace_start(unsigned mask, long capacity)
{
    Tracer& t = tracer();  // constructed first, so the exit handler runs before its destructor
    static const bool at_exit_registered = (std::atexit(trace_release) == 0);
    (void)at_exit_registered;
    std::lock_guard<std::mutex> g(t.mu);
    t.mask.store(0);
    const size_t need = static_cast<size_t>(capacity) * 2;
    while (t.pool.size() < need) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return 1;
        t.pool.push_back(e);
    }
    t.recs.clear();
    t.recs.reserve(static_cast<size_t>(capacity));
    t.used = 0;
    t.mask.store(mask);
    return 0;
}

inline void trace_stop() { tracer().mask.store(0); }

// Sums of the trace for one kind: launches, milliseconds and work. Synchronises on the
// events of that kind.
inline int trace_read(int kind, long long* launches, double* ms, double* work)
{
    Tracer& t = tracer();
    std::lock_guard<std::mutex> g(t.mu);
    long long c = 0;
    double m = 0.0, w = 0.0;
    for (const TraceRec& r : t.recs) {
        if (r.kind != kind) continue;
        if (hipEventSynchronize(r.b) != hipSuccess) return 1;
        float e = 0.f;
        if (hipEventElapsedTime(&e, r.a, r.b) != hipSuccess) return 1;
        ++c;
        m += e;
        w += r.work;
    }
    *launches = c;
    *ms = m;
    *work = w;
    return 0;
}

// RAII bracket of one launch
struct TraceScope {
    long idx;
    hipStream_t s;
    TraceScope(int kind, double work, hipStream_t st) : idx(trace_begin(kind, work, st)), s(st) {}
    ~TraceScope() { trace_end(idx, s); }
};

}  // namespace lrn
