// Translation unit of the blocked many-right-hand-side triangular solve (lrn_trsm.hpp) and its
// launches, compiled apart from the rest of the library.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/linalg_rocm_native.h"
#include "lrn_trsm.hpp"
#include "lrn_trsm_launch.hpp"

namespace lrn {

long long trsm_dinv_elems(long long n) { return (n + kTsRows - 1) / kTsRows * kTsInvElems; }

namespace {

template <typename T, bool kLower, bool kUnit, bool kTrans, typename TA = T>
int inv_go(const TA* a, long long lda, long long n, T* dinv, const int* skip, hipStream_t s)
{
    const unsigned nblk = static_cast<unsigned>((n + kTsRows - 1) / kTsRows);
    hipLaunchKernelGGL((trsm_diag_inv_kernel<TA, T, kLower, kUnit, kTrans>), dim3(nblk, 4), dim3(256), 0, s, a, lda, n,
        dinv, skip);
    return hipGetLastError() == hipSuccess ? LRN_OK : LRN_E_HIP;
}

template <typename T, bool kLower, bool kTrans>
int steps_go(const T* a, long long lda, long long n, const T* dinv, T* x, long long ldx, long long nrhs,
    const int* skip, hipStream_t s)
{
    const long long nblk = (n + kTsRows - 1) / kTsRows;
    const unsigned gy = static_cast<unsigned>((nrhs + kTsRows - 1) / kTsRows);
    // forward: strips -1 (none final: first strip only), 0, ..., nblk - 2; backward: nblk, ..., 1
    for (long long q = 0; q < nblk; ++q) {
        const long long st = kLower ? q - 1 : nblk - q;
        const long long tiles = (q == 0) ? 1 : (kLower ? nblk - 1 - st : st);
        hipLaunchKernelGGL((trsm_step_kernel<T, kLower, kTrans>), dim3(static_cast<unsigned>(tiles), gy), dim3(256), 0,
            s, a, lda, n, dinv, x, ldx, nrhs, st, skip);
        if (hipGetLastError() != hipSuccess) return LRN_E_HIP;
    }
    return LRN_OK;
}

}  // namespace

template <typename T>
int trsm_inv_launch(bool lower, bool unit, bool trans, const T* a, long long lda, long long n, T* dinv,
    const int* skip, hipStream_t s)
{
    if (lower)
        return unit ? (trans ? inv_go<T, true, true, true>(a, lda, n, dinv, skip, s)
                             : inv_go<T, true, true, false>(a, lda, n, dinv, skip, s))
                    : (trans ? inv_go<T, true, false, true>(a, lda, n, dinv, skip, s)
                             : inv_go<T, true, false, false>(a, lda, n, dinv, skip, s));
    return unit ? (trans ? inv_go<T, false, true, true>(a, lda, n, dinv, skip, s)
                         : inv_go<T, false, true, false>(a, lda, n, dinv, skip, s))
                : (trans ? inv_go<T, false, false, true>(a, lda, n, dinv, skip, s)
                         : inv_go<T, false, false, false>(a, lda, n, dinv, skip, s));
}

int trsm_inv_widen_launch(bool lower, bool unit, const float* a, long long lda, long long n, double* dinv,
    const int* skip, hipStream_t s)
{
    if (lower)
        return unit ? inv_go<double, true, true, false, float>(a, lda, n, dinv, skip, s)
                    : inv_go<double, true, false, false, float>(a, lda, n, dinv, skip, s);
    return unit ? inv_go<double, false, true, false, float>(a, lda, n, dinv, skip, s)
                : inv_go<double, false, false, false, float>(a, lda, n, dinv, skip, s);
}

int trsv_pred_widen_launch(bool lower, int dist, const float* a, long long lda, long long n, const double* dinv,
    double* mt, const int* skip, hipStream_t s)
{
    const unsigned nblk = static_cast<unsigned>((n + kTsRows - 1) / kTsRows);
    if (lower)
        hipLaunchKernelGGL((trsv_pred_kernel<float, double, true>), dim3(nblk), dim3(256), 0, s, a, lda, n, dinv, mt, dist,
            skip);
    else
        hipLaunchKernelGGL((trsv_pred_kernel<float, double, false>), dim3(nblk), dim3(256), 0, s, a, lda, n, dinv, mt, dist,
            skip);
    return hipGetLastError() == hipSuccess ? LRN_OK : LRN_E_HIP;
}

template <typename T>
int trsm_steps_launch(bool lower, bool trans, const T* a, long long lda, long long n, const T* dinv, T* x,
    long long ldx, long long nrhs, const int* skip, hipStream_t s)
{
    if (lower)
        return trans ? steps_go<T, true, true>(a, lda, n, dinv, x, ldx, nrhs, skip, s)
                     : steps_go<T, true, false>(a, lda, n, dinv, x, ldx, nrhs, skip, s);
    return trans ? steps_go<T, false, true>(a, lda, n, dinv, x, ldx, nrhs, skip, s)
                 : steps_go<T, false, false>(a, lda, n, dinv, x, ldx, nrhs, skip, s);
}

long long trsm_flag_count(long long n, long long nrhs)
{
    return (n + kTsRows - 1) / kTsRows * ((nrhs + kTsRows - 1) / kTsRows);
}

namespace {

template <typename T, bool kLower, bool kTrans>
int chain_go(const T* a, long long lda, long long n, const T* dinv, T* x, long long ldx, long long nrhs,
    unsigned* flags, unsigned spin_limit, int* err, const int* skip, hipStream_t s)
{
    auto kern = trsm_chain_kernel<T, kLower, kTrans>;
    const long long nblk = (n + kTsRows - 1) / kTsRows;
    int dev = 0, per_cu = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LRN_E_HIP;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), 256, 0) != hipSuccess)
        return LRN_E_HIP;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return LRN_E_HIP;
    if (nblk > static_cast<long long>(per_cu) * cus) return LRN_E_UNSUPPORTED;
    for (long long c0 = 0, q = 0; c0 < nrhs; c0 += kTsRows, ++q) {
        const long long nc = std::min<long long>(kTsRows, nrhs - c0);
        hipLaunchKernelGGL((trsm_chain_kernel<T, kLower, kTrans>), dim3(static_cast<unsigned>(nblk)), dim3(256), 0, s, a,
            lda, n, dinv, x + c0, ldx, nc, flags + q * nblk, spin_limit, err, skip);
        if (hipGetLastError() != hipSuccess) return LRN_E_HIP;
    }
    return LRN_OK;
}

}  // namespace

template <typename T>
int trsm_chain_launch(bool lower, bool trans, const T* a, long long lda, long long n, const T* dinv, T* x,
    long long ldx, long long nrhs, unsigned* flags, unsigned spin_limit, int* err, const int* skip, hipStream_t s)
{
    if (lower)
        return trans ? chain_go<T, true, true>(a, lda, n, dinv, x, ldx, nrhs, flags, spin_limit, err, skip, s)
                     : chain_go<T, true, false>(a, lda, n, dinv, x, ldx, nrhs, flags, spin_limit, err, skip, s);
    return trans ? chain_go<T, false, true>(a, lda, n, dinv, x, ldx, nrhs, flags, spin_limit, err, skip, s)
                 : chain_go<T, false, false>(a, lda, n, dinv, x, ldx, nrhs, flags, spin_limit, err, skip, s);
}

template <typename T>
int gather_rows_launch(const T* b, long long ldb, const int* perm, T* y, long long ldy, long long n, long long nc,
    const int* skip, hipStream_t s)
{
    const long long blocks = std::min<long long>((n * nc + 255) / 256, 4096);
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(static_cast<unsigned>(std::max<long long>(blocks, 1))), dim3(256),
        0, s, b, ldb, perm, y, ldy, n, nc, skip);
    return hipGetLastError() == hipSuccess ? LRN_OK : LRN_E_HIP;
}

template int trsm_inv_launch<double>(bool, bool, bool, const double*, long long, long long, double*, const int*,
    hipStream_t);
template int trsm_inv_launch<float>(bool, bool, bool, const float*, long long, long long, float*, const int*,
    hipStream_t);
template int trsm_steps_launch<double>(bool, bool, const double*, long long, long long, const double*, double*,
    long long, long long, const int*, hipStream_t);
template int trsm_steps_launch<float>(bool, bool, const float*, long long, long long, const float*, float*, long long,
    long long, const int*, hipStream_t);
template int trsm_chain_launch<double>(bool, bool, const double*, long long, long long, const double*, double*,
    long long, long long, unsigned*, unsigned, int*, const int*, hipStream_t);
template int trsm_chain_launch<float>(bool, bool, const float*, long long, long long, const float*, float*, long long,
    long long, unsigned*, unsigned, int*, const int*, hipStream_t);
template int gather_rows_launch<double>(const double*, long long, const int*, double*, long long, long long, long long,
    const int*, hipStream_t);
template int gather_rows_launch<float>(const float*, long long, const int*, float*, long long, long long, long long,
    const int*, hipStream_t);

}  // namespace lrn
