// Translation unit of the batched solver's kernels (lrn_batched.hpp) and their launch, compiled
// apart from the rest of the library so the two build in parallel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <mutex>

#include "../../include/linalg_rocm_native.h"
#include "lrn_batched.hpp"
#include "lrn_batched_launch.hpp"
#include "lrn_trace.hpp"

namespace lrn {
namespace {

// Occupancy (workgroups per compute unit) and compute units of the current device, queried once
// per (kernel, device).
template <typename T> int batched_resident(int* per_cu, int* cus)
{
    struct Entry {
        int dev = -1;
        int per_cu = 0;
        int cus = 0;
    };
    static std::mutex mu;
    static Entry cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return LRN_E_HIP;
    std::lock_guard<std::mutex> g(mu);
    Entry& e = cache[dev];
    if (e.dev != dev) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&e.per_cu, reinterpret_cast<const void*>(batched_gesv_kernel<T>),
                kBatThreads, 0) != hipSuccess)
            return LRN_E_HIP;
        if (hipDeviceGetAttribute(&e.cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return LRN_E_HIP;
        e.dev = dev;
    }
    *per_cu = e.per_cu;
    *cus = e.cus;
    return LRN_OK;
}

}  // namespace

template <typename T>
int batched_solve_launch(T* a, const long long* a_off, long long* piv, const long long* piv_off, T* b,
    const long long* b_off, const int* n, const int* nrhs, int* info, int* order, unsigned* ctr, int count, int grid,
    hipStream_t s)
{
    hipLaunchKernelGGL(batch_order_kernel, dim3(1), dim3(1024), 0, s, n, nrhs, count, order, ctr, info);
    if (hipGetLastError() != hipSuccess) return LRN_E_HIP;
    int per_cu = 0, cus = 0;
    const int st = batched_resident<T>(&per_cu, &cus);
    if (st != LRN_OK) return st;
    long long g = grid > 0 ? grid : static_cast<long long>(std::max(per_cu, 1)) * cus;
    g = std::min<long long>(g, count);
    TraceScope ts(kTrBatched, 0.0, s);
    hipLaunchKernelGGL(batched_gesv_kernel<T>, dim3(static_cast<unsigned>(g)), dim3(kBatThreads), 0, s, a, a_off, piv,
        piv_off, b, b_off, n, nrhs, info, order, ctr, count);
    return hipGetLastError() == hipSuccess ? LRN_OK : LRN_E_HIP;
}

template int batched_solve_launch<double>(double*, const long long*, long long*, const long long*, double*,
    const long long*, const int*, const int*, int*, int*, unsigned*, int, int, hipStream_t);
template int batched_solve_launch<float>(float*, const long long*, long long*, const long long*, float*,
    const long long*, const int*, const int*, int*, int*, unsigned*, int, int, hipStream_t);

}  // namespace lrn
