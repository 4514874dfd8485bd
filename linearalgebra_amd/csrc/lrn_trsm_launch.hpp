// Launch interface of the blocked many-right-hand-side triangular solve, whose kernels live in
// their own translation unit (lrn_trsm.hip). Internal to the library (hidden visibility).
// Every function returns 0 or a positive status (HIP failure); every kernel returns at once
// when `skip` is non-null and *skip is nonzero.
#pragma once

#include <hip/hip_runtime.h>

namespace lrn {

// Elements of T holding the inverses of the 64 x 64 diagonal blocks of an n x n operator.
long long trsm_dinv_elems(long long n);

// Inverts every 64 x 64 diagonal block of op(A) (lower: op(A) lower triangular; unit: implicit
// unit diagonal; trans: op(A) = A^T) into dinv (trsm_dinv_elems(n) elements).
template <typename T>
int trsm_inv_launch(bool lower, bool unit, bool trans, const T* a, long long lda, long long n, T* dinv,
    const int* skip, hipStream_t s);

// The same inverses of FP32 triangular factors (not transposed), computed and stored in FP64.
int trsm_inv_widen_launch(bool lower, bool unit, const float* a, long long lda, long long n, double* dinv,
    const int* skip, hipStream_t s);

// M_b = inv(D_b) A[b][b -+ dist] of every 64-row block b of FP32 triangular factors (a
// predecessor in a forward / backward single-RHS substitution), from the inverses of
// trsm_inv_widen_launch, in FP64, same layout and size (trsm_dinv_elems(n)).
int trsv_pred_widen_launch(bool lower, int dist, const float* a, long long lda, long long n, const double* dinv,
    double* mt, const int* skip, hipStream_t s);

// X = op(A)^-1 X in place (X n x nrhs, row stride ldx) with the inverses from trsm_inv_launch
// (same lower / trans flags): one launch per 64-row strip.
template <typename T>
int trsm_steps_launch(bool lower, bool trans, const T* a, long long lda, long long n, const T* dinv, T* x,
    long long ldx, long long nrhs, const int* skip, hipStream_t s);

// The same solve as trsm_steps_launch in one launch per 64 columns of X (trsm_chain_kernel):
// workgroups chained through flags[] (trsm_flag_count(n, nrhs) zeroed words), bounded waits
// raising *err. Returns LRN_E_UNSUPPORTED, launching nothing, when the tiles' workgroups
// cannot all be resident at once.
long long trsm_flag_count(long long n, long long nrhs);
template <typename T>
int trsm_chain_launch(bool lower, bool trans, const T* a, long long lda, long long n, const T* dinv, T* x,
    long long ldx, long long nrhs, unsigned* flags, unsigned spin_limit, int* err, const int* skip, hipStream_t s);

// Y[i][c] = B[perm[i]][c] (B[i][c] when perm is null) for i < n, c < nc.
template <typename T>
int gather_rows_launch(const T* b, long long ldb, const int* perm, T* y, long long ldy, long long n, long long nc,
    const int* skip, hipStream_t s);

}  // namespace lrn
