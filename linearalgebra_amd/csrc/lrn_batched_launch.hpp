// Launch interface of the batched solver, whose kernels live in their own translation unit
// (lrn_batched.hip): the ordering kernel and the persistent solver on one stream. Internal to
// the library (hidden visibility); the C ABI wrappers in lrn_capi.hip call it.
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

// Orders the systems by decreasing size into `order`, resets `ctr`, then solves them with a
// persistent grid of `grid` workgroups (0 = occupancy x compute units). Returns 0 or a positive
// status (HIP failure).
template <typename T>
int batched_solve_launch(T* a, const long long* a_off, long long* piv, const long long* piv_off, T* b,
    const long long* b_off, const int* n, const int* nrhs, int* info, int* order, unsigned* ctr, int count, int grid,
    hipStream_t s);

}  // namespace lrn
