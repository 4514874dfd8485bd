// Device-side building blocks shared by the LU kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lrn {

constexpr int kWave = 64;

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
__device__ __forceinline__ void wait_counter(const unsigned* ctr, unsigned target, unsigned limit, int* err)
{
    unsigned spins = 0;
    while (__hip_atomic_load(const_cast<unsigned*>(ctr), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
        ++spins;
        if ((spins & 255u) == 0u && error_raised(err)) break;
        if (spins > limit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

__device__ __forceinline__ void signal_counter(unsigned* ctr)
{
    __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}


}  // namespace lrn
