// Device-side building blocks shared by the LU kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

namespace lrn {

constexpr int kWave = 64;
// Accelerator complex dies of the MI355X: consecutive workgroup ids are dispatched to them in turn.
constexpr int kXcds = 8;

// Smallest normal number of the type: the reference divides instead of multiplying by the
// reciprocal below it (src/matrix_operation/cpu/native/lu_decomposition_native.cpp:212-228).
template <typename T> struct TypeMin;
template <> struct TypeMin<double> { static constexpr double v = DBL_MIN; };
template <> struct TypeMin<float> { static constexpr float v = FLT_MIN; };

// Pivot-search order: larger magnitude wins; equal magnitudes go to the lower row index,
// which is the "first maximum wins" rule of a strict '>' scan in increasing row order
// (src/matrix_operation/cpu/native/lu_decomposition_native.cpp:174-185).
__device__ __forceinline__ bool argmax_better(double v1, long long i1, double v2, long long i2)
{
    return v1 > v2 || (v1 == v2 && i1 < i2);
}

__device__ __forceinline__ void wave_argmax(double& v, long long& idx)
{
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off, kWave);
        const long long oi = __shfl_xor(idx, off, kWave);
        if (argmax_better(ov, oi, v, idx)) {
            v = ov;
            idx = oi;
        }
    }
}

// Search key of one candidate entry: |a|, except that NaN entries never win, and a NaN on
// the diagonal itself always wins (a sequential scan that starts from |diag| = NaN keeps
// the diagonal, because no comparison against NaN is true).
template <typename T>
__device__ __forceinline__ double pivot_key(T a, bool is_diag)
{
    const double v = fabs(static_cast<double>(a));
    if (v != v) return is_diag ? __builtin_huge_val() : -1.0;
    return v;
}

// Coherent (agent-scope) accesses for data exchanged between workgroups inside one kernel.
__device__ __forceinline__ void st_agent(double* p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(float* p, float v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(long long* p, long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent(const double* p)
{
    return __hip_atomic_load(const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_agent(const float* p)
{
    return __hip_atomic_load(const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long ld_agent(const long long* p)
{
    return __hip_atomic_load(const_cast<long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool error_raised(const int* err)
{
    return __hip_atomic_load(const_cast<int*>(err), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

// Bounded wait until *ctr >= target. Every wait in this library is bounded: after
// `limit` polls it raises *err and returns, and once *err is raised every later wait
// returns at once, so a kernel always exits even if the workgroups it waits for never
// arrive (the host then reports the error instead of a result).
__device__ __forceinline__ void raise_error(int* err) { __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// A wait polls at most `limit` times; limit 0 expires at once without polling (the host sets it
// to force every wait's error path in tests).
__device__ __forceinline__ void wait_counter(const unsigned* ctr, unsigned target, unsigned limit, int* err)
{
    if (limit == 0u) {
        raise_error(err);
        return;
    }
    unsigned spins = 0;
    while (__hip_atomic_load(const_cast<unsigned*>(ctr), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        ++spins;
        if ((spins & 255u) == 0u && error_raised(err)) break;
        if (spins > limit) {
            raise_error(err);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Bounded wait for a flag published with publish_flag, without an acquire fence: the poll
// and the data loads that follow it are agent-coherent (sc1) accesses, and the data loads
// are issued only after the poll has returned the flag, so no cache invalidation is needed
// (an agent-scope acquire invalidates the XCD's L2 on every poll).
__device__ __forceinline__ void wait_flag_coherent(const unsigned* flag, unsigned target, unsigned limit, int* err)
{
    if (limit == 0u) {
        raise_error(err);
        return;
    }
    unsigned spins = 0;
    while (__hip_atomic_load(const_cast<unsigned*>(flag), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        ++spins;
        if ((spins & 255u) == 0u && error_raised(err)) break;
        if (spins > limit) {
            raise_error(err);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Completes this thread's outstanding memory operations (on this GPU family stores are
// counted in vmcnt). Called by every thread after its agent-coherent (sc1) data stores and
// followed by a workgroup barrier, so one thread can then publish a flag that covers them all
// without a release fence (an agent-scope release writes back the XCD's L2).
__device__ __forceinline__ void wait_own_stores() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

__device__ __forceinline__ void signal_counter(unsigned* ctr)
{
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}


}  // namespace lrn

namespace lrn {

// Wave-wide maximum of a double (all 64 lanes active). Within each row of 16 lanes by DPP
// (swap neighbours, swap pairs, half-row mirror, row mirror), then across the four rows by
// reading lanes 0, 16, 32 and 48. Every lane returns the same value.
__device__ __forceinline__ double dpp_max_step(double v, int ctrl_sel)
{
    const long long bits = __double_as_longlong(v);
    int lo = static_cast<int>(bits), hi = static_cast<int>(bits >> 32);
    int lo2, hi2;
    switch (ctrl_sel) {
    case 0:
        lo2 = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);
        hi2 = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
        break;
    case 1:
        lo2 = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);
        hi2 = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
        break;
    case 2:
        lo2 = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xF, 0xF, false);
        hi2 = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xF, 0xF, false);
        break;
    default:
        lo2 = __builtin_amdgcn_mov_dpp(lo, 0x140, 0xF, 0xF, false);
        hi2 = __builtin_amdgcn_mov_dpp(hi, 0x140, 0xF, 0xF, false);
        break;
    }
    const double o = __longlong_as_double((static_cast<long long>(hi2) << 32) | static_cast<unsigned>(lo2));
    return o > v ? o : v;
}

__device__ __forceinline__ double read_lane_f64(double v, int lane)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(static_cast<int>(bits), lane);
    const int hi = __builtin_amdgcn_readlane(static_cast<int>(bits >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}

__device__ __forceinline__ double wave_max_f64(double v)
{
    v = dpp_max_step(v, 0);
    v = dpp_max_step(v, 1);
    v = dpp_max_step(v, 2);
    v = dpp_max_step(v, 3);
    const double a = read_lane_f64(v, 0), b = read_lane_f64(v, 16);
    const double c = read_lane_f64(v, 32), d = read_lane_f64(v, 48);
    const double ab = a > b ? a : b, cd = c > d ? c : d;
    return ab > cd ? ab : cd;
}

// Wave argmax for the pivot search when lanes hold candidates in increasing row order:
// the maximum key, and the row of the lowest lane holding it. Keys are never NaN.
__device__ __forceinline__ void wave_argmax_ordered(double& key, long long& row)
{
    const double m = wave_max_f64(key);
    const unsigned long long hits = __ballot(key == m);
    const int l = __ffsll(static_cast<long long>(hits)) - 1;
    const long long r = (static_cast<long long>(__builtin_amdgcn_readlane(static_cast<int>(row >> 32), l)) << 32) |
                        static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(row), l));
    key = m;
    row = r;
}

template <typename T> __device__ __forceinline__ T read_lane_t(T v, int lane);
template <> __device__ __forceinline__ double read_lane_t<double>(double v, int lane) { return read_lane_f64(v, lane); }
template <> __device__ __forceinline__ float read_lane_t<float>(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}


template <typename T> struct Mfma16;
// Operands of both: A[i][k] at lane i + 16k, B[k][j] at lane j + 16k. Result element v of
// lane l holds D[row(l, v)][l % 16]; the row order differs between the two instructions
// (measured on gfx950 with one-hot operands).
template <> struct Mfma16<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

}  // namespace lrn
