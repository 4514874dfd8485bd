/*
 * C++ face of the MI355X-native LU solve path: the reference's operator templates
 * (lu_decomposition<T, B>, linear_solver<T, B>, triangular_solver<T, B>, determinant<T, B>,
 * inversion<T, B>) specialised for a new backend tag
 * linalg_backend::ROCM_NATIVE. Same names, argument meaning, ownership and error
 * behaviour as the reference; the bodies (csrc/rocm_native_backend.cpp) call the C ABI of
 * linalg_rocm_native.h.
 *
 * Interfaces mirrored (paths relative to the reference repository):
 *   include/common/backend.h:21-29                     enum class linalg_backend (+ ROCM_NATIVE)
 *   include/common/types.h:18-29                       linalg_int (64-bit signed)
 *   include/matrix_operation/lu_decomposition.h:15-54  lu_decomposition<T, B>
 *   include/matrix_operation/linear_solver.h:15-65     linear_solver_type, linear_solver<T, B>
 *   include/matrix_operation/triangular_solver.h:20-38 triangular_solver<T, B>
 *   include/matrix_operation/determinant.h:13-55       determinant<T, B>
 *   include/matrix_operation/inversion.h:13-53         inversion<T, B>
 *   include/matrix_operation/cholesky_decomposition.h:10-94  cholesky_decomposition_enum,
 *                                                      cholesky_decomposition<T, B>
 *
 * Inside the reference tree, define LRN_WITH_REFERENCE_HEADERS and put the reference's
 * include root on the include path: the reference's own declarations are used and only the
 * ROCM_NATIVE enumerator has to be added to its linalg_backend. Standalone, this header
 * declares the same templates itself.
 *
 * Errors: invalid arguments throw std::invalid_argument (the reference's LINALG_CHECK
 * throws); HIP runtime failures and expired inter-workgroup waits throw std::runtime_error.
 * A numerically singular matrix is not an error (apply returns false, *info = first zero
 * pivot), exactly as in the reference.
 */
#pragma once

#include <cstddef>
#include <cstdint>

#if defined(LRN_WITH_REFERENCE_HEADERS)
#include "include/common/backend.h"
#include "include/common/types.h"
#include "include/matrix_operation/linear_solver.h"
#include "include/matrix_operation/lu_decomposition.h"
#include "include/matrix_operation/triangular_solver.h"
#include "include/matrix_operation/determinant.h"
#include "include/matrix_operation/inversion.h"
#include "include/matrix_operation/cholesky_decomposition.h"
#else

#if !defined(linalg_int)
#define linalg_int long long int
#endif

namespace linalg
{
enum class linalg_backend : std::uint8_t
{
    NATIVE,
    BLAS_LAPACKE,
    MKL,
    CUDA_BLAS,
    ROCM_BLAS,
    METAL_BLAS,
    ROCM_NATIVE  // this backend: hand-written gfx950 kernels
};

enum class linear_solver_type : std::uint8_t
{
    LU_LINEAR_SOLVER = 1,
    LU_UPFRONT_LINEAR_SOLVER = 2,
    CHOLESKY_LINEAR_SOLVER = 3,
    CHOLESKY_UPFRONT_LINEAR_SOLVER = 4,
};

template <typename T, linalg_backend Backend> struct lu_decomposition
{
    static bool apply(T* m, linalg_int lda, linalg_int* pivot, void* workspace, size_t workspace_size,
        int* info = nullptr, void* stream = nullptr);
    static size_t workspace_size(linalg_int lda);
};

template <typename T, linalg_backend Backend> struct linear_solver
{
    static void apply(T* m, linalg_int* pivot, linalg_int lda, T* x, linear_solver_type type, void* workspace,
        size_t workspace_size, void* stream = nullptr);
    static size_t workspace_size(linalg_int lda, linear_solver_type type);
};

// include/matrix_operation/triangular_solver.h:20-38
template <typename T, linalg_backend Backend> struct triangular_solver
{
    static void apply(bool upper, bool transpose, bool unit_diagonal, linalg_int n, linalg_int nrhs, const T* a,
        linalg_int lda, const T* b, linalg_int ldb, T* x, linalg_int ldx, void* workspace, size_t workspace_size,
        void* stream = nullptr);
    static size_t workspace_size(linalg_int n, linalg_int nrhs);
};

// include/matrix_operation/determinant.h:13-55
template <typename T, linalg_backend Backend> struct determinant
{
    static T apply(T* m, linalg_int* pivot, linalg_int lda, linear_solver_type type, void* workspace,
        size_t workspace_size, void* stream = nullptr);
    static size_t workspace_size(linalg_int lda, linear_solver_type type);
};

// include/matrix_operation/inversion.h:13-53
template <typename T, linalg_backend Backend> struct inversion
{
    static void apply(T* m, linalg_int* pivot, linalg_int lda, linear_solver_type type, void* workspace,
        size_t workspace_size, void* stream = nullptr);
    static size_t workspace_size(linalg_int lda, linear_solver_type type);
};

// include/matrix_operation/cholesky_decomposition.h:15-19,38-94
enum class cholesky_decomposition_enum : char
{
    LOWER_TRIANGULAR = 'L',
    UPPER_TRIANGULAR = 'U'
};

template <typename T, linalg_backend Backend> struct cholesky_decomposition
{
    static bool apply(T* C, linalg_int lda, cholesky_decomposition_enum type, void* workspace, size_t workspace_size,
        int* info, void* stream = nullptr);
    static bool apply_aad(T* L_aad, const T* L, linalg_int lda, cholesky_decomposition_enum type, T* A_aad);
    static size_t workspace_size(linalg_int lda, cholesky_decomposition_enum type);
};
}  // namespace linalg
#endif

namespace linalg
{
// Explicit specialisations defined in csrc/rocm_native_backend.cpp (T = float, double).
// All pointers are device memory except `info` (host). lu_decomposition::apply is
// synchronous on `stream`; linear_solver::apply is asynchronous on `stream`.
template <> bool lu_decomposition<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int lda,
    linalg_int* pivot, void* workspace, size_t workspace_size, int* info, void* stream);
template <> bool lu_decomposition<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int lda,
    linalg_int* pivot, void* workspace, size_t workspace_size, int* info, void* stream);
template <> size_t lu_decomposition<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda);
template <> size_t lu_decomposition<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda);

template <> void linear_solver<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, double* x, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> void linear_solver<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, float* x, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> size_t linear_solver<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);
template <> size_t linear_solver<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);

// triangular_solver: op(A) X = B, device pointers, asynchronous on `stream`.
template <> void triangular_solver<double, linalg_backend::ROCM_NATIVE>::apply(bool upper, bool transpose,
    bool unit_diagonal, linalg_int n, linalg_int nrhs, const double* a, linalg_int lda, const double* b,
    linalg_int ldb, double* x, linalg_int ldx, void* workspace, size_t workspace_size, void* stream);
template <> void triangular_solver<float, linalg_backend::ROCM_NATIVE>::apply(bool upper, bool transpose,
    bool unit_diagonal, linalg_int n, linalg_int nrhs, const float* a, linalg_int lda, const float* b,
    linalg_int ldb, float* x, linalg_int ldx, void* workspace, size_t workspace_size, void* stream);
template <> size_t triangular_solver<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int n,
    linalg_int nrhs);
template <> size_t triangular_solver<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int n,
    linalg_int nrhs);

// determinant (signed, as the GPU reference's) and in-place inversion; m, pivot and workspace
// are device memory (pivot may be null in LU mode); both synchronous on `stream`.
template <> double determinant<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> float determinant<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> size_t determinant<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);
template <> size_t determinant<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);
template <> void inversion<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> void inversion<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream);
template <> size_t inversion<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);
template <> size_t inversion<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type);

// cholesky_decomposition: as the reference's GPU specialisation, `info` is DEVICE memory (required)
// written asynchronously on `stream`, and apply returns true once the work is enqueued (the
// numerical status lives in *info). C is factored in place: L with zeros above (LOWER) or
// U = L^T with zeros below (UPPER). apply_aad is not provided by this backend.
template <> bool cholesky_decomposition<double, linalg_backend::ROCM_NATIVE>::apply(double* C, linalg_int lda,
    cholesky_decomposition_enum type, void* workspace, size_t workspace_size, int* info, void* stream);
template <> bool cholesky_decomposition<float, linalg_backend::ROCM_NATIVE>::apply(float* C, linalg_int lda,
    cholesky_decomposition_enum type, void* workspace, size_t workspace_size, int* info, void* stream);
template <> size_t cholesky_decomposition<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    cholesky_decomposition_enum type);
template <> size_t cholesky_decomposition<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    cholesky_decomposition_enum type);

// Extensions of this backend (no reference counterpart).
namespace rocm_native
{
// Several right-hand sides: x is a row-major lda x nrhs block with row stride ldx.
template <typename T> size_t linear_solver_workspace_size(linalg_int lda, linalg_int nrhs, linear_solver_type type);
template <typename T>
void linear_solver_apply(T* m, linalg_int* pivot, linalg_int lda, T* x, linalg_int nrhs, linalg_int ldx,
    linear_solver_type type, void* workspace, size_t workspace_size, void* stream = nullptr);

// Mixed-precision solve of A x = b (FP64 A, b, x; A unchanged) to normwise backward error
// <= tol: FP32/bf16 factorisation, LU-IR then GMRES-IR with FP64 residuals. Synchronous.
struct mixed_result
{
    int info;
    int path;  // 0 LU-IR, 1 GMRES-IR, 2 FP64 fallback, 3 not converged
    int ir_iterations;
    int gmres_iterations;
    double backward_error;
};
size_t gesv_mixed_workspace_size(linalg_int n, int gmres_restart = 30, bool fallback = true);
mixed_result gesv_mixed(const double* a, linalg_int n, const double* b, double* x, void* workspace,
    size_t workspace_size, double tol = 1e-14, int max_iter = 20, int gmres_restart = 30, void* stream = nullptr);
// batched_matrix_multiplication (include/matrix_operation/multiplication_batched.h:7-42): the
// reference declares it as a free function of namespace linalg; this backend's version lives
// here so both can be linked into one program. C_i = A_i B_i, device pointers, asynchronous.
void batched_matrix_multiplication(linalg_int dim, linalg_int count, const float* a, const float* b, float* c,
    void* stream = nullptr);
void batched_matrix_multiplication(linalg_int dim, linalg_int count, const double* a, const double* b, double* c,
    void* stream = nullptr);
}  // namespace rocm_native
}  // namespace linalg
