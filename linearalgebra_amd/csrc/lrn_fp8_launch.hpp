// Launch interface of the block-scaled FP8 trailing update (lrn_fp8.hpp), whose kernels live in
// their own translation unit (lrn_fp8.hip). Internal to the library (hidden visibility).
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

// C[M x N] -= A[M x K] * B[K x N] (FP32, row-major) with `planes` (1..3) e4m3 planes per
// operand and one scale per 32 K elements; every kernel returns at once when skip is non-null
// and *skip is nonzero. Returns 0 or a positive status.
int gemm_sub_fp8_launch(const float* A, long long lda, const float* B, long long ldb, float* C, long long ldc,
    long long M, long long N, long long K, int planes, hipStream_t s, const int* skip);

}  // namespace lrn
