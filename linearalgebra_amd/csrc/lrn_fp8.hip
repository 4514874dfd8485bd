// Translation unit of the block-scaled FP8 trailing update (lrn_fp8.hpp) and its launch,
// compiled apart from the rest of the library.
#include <hip/hip_runtime.h>

#include "../../include/linalg_rocm_native.h"
#include "lrn_fp8.hpp"
#include "lrn_fp8_launch.hpp"

namespace lrn {

int gemm_sub_fp8_launch(const float* A, long long lda, const float* B, long long ldb, float* C, long long ldc,
    long long M, long long N, long long K, int planes, hipStream_t s, const int* skip)
{
    if (M <= 0 || N <= 0 || K <= 0) return LRN_OK;
    if (planes < 1 || planes > 3) return LRN_E_INTERNAL;
    const long long tm = (M + kF8BM - 1) / kF8BM, tn = (N + kF8BN - 1) / kF8BN;
    if (tm * tn > 0x7fffffffll || M > 0x7fffffffll || N > 0x7fffffffll || K > 0x7fffffffll) return LRN_E_UNSUPPORTED;
    const dim3 grid(static_cast<unsigned>(tm * tn));
    const int iM = static_cast<int>(M), iN = static_cast<int>(N), iK = static_cast<int>(K), itn = static_cast<int>(tn);
    if (planes == 1)
        hipLaunchKernelGGL(gemm_sub_fp8_kernel<1>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itn, skip);
    else if (planes == 2)
        hipLaunchKernelGGL(gemm_sub_fp8_kernel<2>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itn, skip);
    else
        hipLaunchKernelGGL(gemm_sub_fp8_kernel<3>, grid, dim3(256), 0, s, A, lda, B, ldb, C, ldc, iM, iN, iK, itn, skip);
    return hipGetLastError() == hipSuccess ? LRN_OK : LRN_E_HIP;
}

}  // namespace lrn
