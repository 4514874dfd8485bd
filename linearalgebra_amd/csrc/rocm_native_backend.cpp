// linalg_backend::ROCM_NATIVE specialisations of the reference's lu_decomposition and
// linear_solver templates, as thin C++ wrappers over the C ABI (linalg_rocm_native.h).
//
// Reference behaviour kept (paths relative to the reference repository):
//   argument checks throw       src/matrix_operation/cpu/native/lu_decomposition_native.cpp:30-35,
//                               src/matrix_operation/cpu/native/linear_solver_native.cpp:24-28
//   apply returns info == 0     include/matrix_operation/lu_decomposition.h:30-36
//   workspace_size per type     lu_decomposition_native.cpp:293-326, linear_solver_native.cpp:225-241
//   Cholesky modes              the lower factor (linear_solver_native.cpp:177-189); determinant squares
//                               the diagonal product as the GPU reference does
//   cholesky_decomposition      device info, asynchronous (src/matrix_operation/gpu/hipblas/
//                               cholesky_decomposition_hip.cpp:69-100,135-170)
#include <stdexcept>
#include <string>

#include "../../include/linalg_rocm_native.h"
#include "../../include/linalg_rocm_native.hpp"

namespace
{

void check(int code, const char* what)
{
    if (code == LRN_OK) return;
    std::string msg = std::string(what) + ": " + lrn_status_string(code);
    if (code < 0) throw std::invalid_argument(msg + " (argument " + std::to_string(-code) + ")");
    throw std::runtime_error(msg);
}

int solver_type(linalg::linear_solver_type t)
{
    const int v = static_cast<int>(t);
    if (v < LRN_LU_LINEAR_SOLVER || v > LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER)
        throw std::invalid_argument("linear_solver<ROCM_NATIVE>: unknown linear_solver_type");
    return v;
}



template <typename T> struct ops;
template <> struct ops<double> {
    static constexpr auto lu_ws = lrn_lu_decomposition_workspace_size_f64;
    static constexpr auto lu = lrn_lu_decomposition_f64;
    static constexpr auto sv_ws = lrn_linear_solver_workspace_size_f64;
    static constexpr auto sv = lrn_linear_solver_f64;
    static constexpr auto svm_ws = lrn_linear_solver_multi_workspace_size_f64;
    static constexpr auto svm = lrn_linear_solver_multi_f64;
    static constexpr auto tr_ws = lrn_triangular_solver_workspace_size_f64;
    static constexpr auto tr = lrn_triangular_solver_f64;
    static constexpr auto det_ws = lrn_determinant_workspace_size_f64;
    static constexpr auto det = lrn_determinant_f64;
    static constexpr auto inv_ws = lrn_inversion_workspace_size_f64;
    static constexpr auto inv = lrn_inversion_f64;
    static constexpr auto chol_ws = lrn_cholesky_decomposition_workspace_size_f64;
    static constexpr auto chol = lrn_cholesky_decomposition_async_f64;
};
template <> struct ops<float> {
    static constexpr auto lu_ws = lrn_lu_decomposition_workspace_size_f32;
    static constexpr auto lu = lrn_lu_decomposition_f32;
    static constexpr auto sv_ws = lrn_linear_solver_workspace_size_f32;
    static constexpr auto sv = lrn_linear_solver_f32;
    static constexpr auto svm_ws = lrn_linear_solver_multi_workspace_size_f32;
    static constexpr auto svm = lrn_linear_solver_multi_f32;
    static constexpr auto tr_ws = lrn_triangular_solver_workspace_size_f32;
    static constexpr auto tr = lrn_triangular_solver_f32;
    static constexpr auto det_ws = lrn_determinant_workspace_size_f32;
    static constexpr auto det = lrn_determinant_f32;
    static constexpr auto inv_ws = lrn_inversion_workspace_size_f32;
    static constexpr auto inv = lrn_inversion_f32;
    static constexpr auto chol_ws = lrn_cholesky_decomposition_workspace_size_f32;
    static constexpr auto chol = lrn_cholesky_decomposition_async_f32;
};

template <typename T> bool lu_apply(T* m, linalg_int lda, linalg_int* pivot, void* ws, size_t ws_size, int* info, void* stream)
{
    if (m == nullptr || pivot == nullptr) throw std::invalid_argument("lu_decomposition<ROCM_NATIVE>: null m or pivot");
    if (lda <= 0) throw std::invalid_argument("lu_decomposition<ROCM_NATIVE>: lda must be positive");
    int inf = 0;
    check(ops<T>::lu(m, lda, reinterpret_cast<int64_t*>(pivot), ws, ws_size, &inf, stream), "lu_decomposition<ROCM_NATIVE>::apply");
    if (info != nullptr) *info = inf;
    return inf == 0;
}

template <typename T> size_t lu_ws(linalg_int lda)
{
    size_t b = 0;
    check(ops<T>::lu_ws(lda, &b), "lu_decomposition<ROCM_NATIVE>::workspace_size");
    return b;
}

template <typename T>
void solve_apply(T* m, linalg_int* pivot, linalg_int lda, T* x, linalg::linear_solver_type type, void* ws,
    size_t ws_size, void* stream)
{
    if (m == nullptr || x == nullptr) throw std::invalid_argument("linear_solver<ROCM_NATIVE>: null m or x");
    if (lda <= 0) throw std::invalid_argument("linear_solver<ROCM_NATIVE>: lda must be positive");
    const int t = solver_type(type);
    check(ops<T>::sv(m, reinterpret_cast<int64_t*>(pivot), lda, x, t, ws, ws_size, stream),
        "linear_solver<ROCM_NATIVE>::apply");
}

template <typename T> size_t solve_ws(linalg_int lda, linalg::linear_solver_type type)
{
    size_t b = 0;
    check(ops<T>::sv_ws(lda, solver_type(type), &b), "linear_solver<ROCM_NATIVE>::workspace_size");
    return b;
}

}  // namespace

namespace linalg
{

template <> bool lu_decomposition<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int lda,
    linalg_int* pivot, void* workspace, size_t workspace_size, int* info, void* stream)
{
    return lu_apply<double>(m, lda, pivot, workspace, workspace_size, info, stream);
}
template <> bool lu_decomposition<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int lda,
    linalg_int* pivot, void* workspace, size_t workspace_size, int* info, void* stream)
{
    return lu_apply<float>(m, lda, pivot, workspace, workspace_size, info, stream);
}
template <> size_t lu_decomposition<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda)
{
    return lu_ws<double>(lda);
}
template <> size_t lu_decomposition<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda)
{
    return lu_ws<float>(lda);
}

template <> void linear_solver<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, double* x, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    solve_apply<double>(m, pivot, lda, x, type, workspace, workspace_size, stream);
}
template <> void linear_solver<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, float* x, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    solve_apply<float>(m, pivot, lda, x, type, workspace, workspace_size, stream);
}
template <> size_t linear_solver<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return solve_ws<double>(lda, type);
}
template <> size_t linear_solver<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return solve_ws<float>(lda, type);
}

template <typename T>
void tr_apply(bool upper, bool transpose, bool unit, linalg_int n, linalg_int nrhs, const T* a, linalg_int lda,
    const T* b, linalg_int ldb, T* x, linalg_int ldx, void* ws, size_t ws_size, void* stream)
{
    check(ops<T>::tr(upper ? 1 : 0, transpose ? 1 : 0, unit ? 1 : 0, n, nrhs, a, lda, b, ldb, x, ldx, ws, ws_size, stream),
        "triangular_solver<ROCM_NATIVE>::apply");
}
template <typename T> size_t tr_ws(linalg_int n, linalg_int nrhs)
{
    size_t bytes = 0;
    check(ops<T>::tr_ws(n, nrhs, &bytes), "triangular_solver<ROCM_NATIVE>::workspace_size");
    return bytes;
}

template <> void triangular_solver<double, linalg_backend::ROCM_NATIVE>::apply(bool upper, bool transpose,
    bool unit_diagonal, linalg_int n, linalg_int nrhs, const double* a, linalg_int lda, const double* b,
    linalg_int ldb, double* x, linalg_int ldx, void* workspace, size_t workspace_size, void* stream)
{
    tr_apply<double>(upper, transpose, unit_diagonal, n, nrhs, a, lda, b, ldb, x, ldx, workspace, workspace_size, stream);
}
template <> void triangular_solver<float, linalg_backend::ROCM_NATIVE>::apply(bool upper, bool transpose,
    bool unit_diagonal, linalg_int n, linalg_int nrhs, const float* a, linalg_int lda, const float* b,
    linalg_int ldb, float* x, linalg_int ldx, void* workspace, size_t workspace_size, void* stream)
{
    tr_apply<float>(upper, transpose, unit_diagonal, n, nrhs, a, lda, b, ldb, x, ldx, workspace, workspace_size, stream);
}
template <> size_t triangular_solver<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int n,
    linalg_int nrhs)
{
    return tr_ws<double>(n, nrhs);
}
template <> size_t triangular_solver<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int n,
    linalg_int nrhs)
{
    return tr_ws<float>(n, nrhs);
}

template <typename T>
T det_apply(T* m, linalg_int* pivot, linalg_int lda, linear_solver_type type, void* ws, size_t ws_size, void* stream)
{
    if (m == nullptr) throw std::invalid_argument("determinant<ROCM_NATIVE>: m must not be null");
    if (lda <= 0) throw std::invalid_argument("determinant<ROCM_NATIVE>: lda must be positive");
    T out = T(0);
    check(ops<T>::det(m, reinterpret_cast<int64_t*>(pivot), lda, solver_type(type), ws, ws_size, &out, stream),
        "determinant<ROCM_NATIVE>::apply");
    return out;
}
template <typename T> size_t det_ws(linalg_int lda, linear_solver_type type)
{
    size_t bytes = 0;
    check(ops<T>::det_ws(lda, solver_type(type), &bytes), "determinant<ROCM_NATIVE>::workspace_size");
    return bytes;
}
template <typename T>
void inv_apply(T* m, linalg_int* pivot, linalg_int lda, linear_solver_type type, void* ws, size_t ws_size, void* stream)
{
    if (m == nullptr) throw std::invalid_argument("inversion<ROCM_NATIVE>: m must not be null");
    if (lda <= 0) throw std::invalid_argument("inversion<ROCM_NATIVE>: lda must be positive");
    check(ops<T>::inv(m, reinterpret_cast<int64_t*>(pivot), lda, solver_type(type), ws, ws_size, stream),
        "inversion<ROCM_NATIVE>::apply");
}
template <typename T> size_t inv_ws(linalg_int lda, linear_solver_type type)
{
    size_t bytes = 0;
    check(ops<T>::inv_ws(lda, solver_type(type), &bytes), "inversion<ROCM_NATIVE>::workspace_size");
    return bytes;
}

template <> double determinant<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    return det_apply<double>(m, pivot, lda, type, workspace, workspace_size, stream);
}
template <> float determinant<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    return det_apply<float>(m, pivot, lda, type, workspace, workspace_size, stream);
}
template <> size_t determinant<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return det_ws<double>(lda, type);
}
template <> size_t determinant<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return det_ws<float>(lda, type);
}
template <> void inversion<double, linalg_backend::ROCM_NATIVE>::apply(double* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    inv_apply<double>(m, pivot, lda, type, workspace, workspace_size, stream);
}
template <> void inversion<float, linalg_backend::ROCM_NATIVE>::apply(float* m, linalg_int* pivot,
    linalg_int lda, linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    inv_apply<float>(m, pivot, lda, type, workspace, workspace_size, stream);
}
template <> size_t inversion<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return inv_ws<double>(lda, type);
}
template <> size_t inversion<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    linear_solver_type type)
{
    return inv_ws<float>(lda, type);
}

// cholesky_decomposition: the reference's GPU contract (src/matrix_operation/gpu/hipblas/
// cholesky_decomposition_hip.cpp:69-100,135-170): device info required, asynchronous, true once
// enqueued; the triangle is chosen by the enum ('L' / 'U').
template <typename T>
bool chol_apply(T* C, linalg_int lda, cholesky_decomposition_enum type, void* ws, size_t ws_size, int* info,
    void* stream)
{
    if (info == nullptr) throw std::invalid_argument("cholesky_decomposition<ROCM_NATIVE>: device info must not be null");
    if (C == nullptr) throw std::invalid_argument("cholesky_decomposition<ROCM_NATIVE>: C must not be null");
    if (lda <= 0) throw std::invalid_argument("cholesky_decomposition<ROCM_NATIVE>: lda must be positive");
    int upper = 0;
    if (type == cholesky_decomposition_enum::UPPER_TRIANGULAR)
        upper = 1;
    else if (type != cholesky_decomposition_enum::LOWER_TRIANGULAR)
        throw std::invalid_argument("cholesky_decomposition<ROCM_NATIVE>: unknown triangle");
    check(ops<T>::chol(C, lda, upper, ws, ws_size, info, stream), "cholesky_decomposition<ROCM_NATIVE>::apply");
    return true;
}
template <typename T> size_t chol_ws(linalg_int lda, cholesky_decomposition_enum type)
{
    if (type != cholesky_decomposition_enum::LOWER_TRIANGULAR && type != cholesky_decomposition_enum::UPPER_TRIANGULAR)
        throw std::invalid_argument("cholesky_decomposition<ROCM_NATIVE>: unknown triangle");
    size_t bytes = 0;
    check(ops<T>::chol_ws(lda, &bytes), "cholesky_decomposition<ROCM_NATIVE>::workspace_size");
    return bytes;
}

template <> bool cholesky_decomposition<double, linalg_backend::ROCM_NATIVE>::apply(double* C, linalg_int lda,
    cholesky_decomposition_enum type, void* workspace, size_t workspace_size, int* info, void* stream)
{
    return chol_apply<double>(C, lda, type, workspace, workspace_size, info, stream);
}
template <> bool cholesky_decomposition<float, linalg_backend::ROCM_NATIVE>::apply(float* C, linalg_int lda,
    cholesky_decomposition_enum type, void* workspace, size_t workspace_size, int* info, void* stream)
{
    return chol_apply<float>(C, lda, type, workspace, workspace_size, info, stream);
}
template <> size_t cholesky_decomposition<double, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    cholesky_decomposition_enum type)
{
    return chol_ws<double>(lda, type);
}
template <> size_t cholesky_decomposition<float, linalg_backend::ROCM_NATIVE>::workspace_size(linalg_int lda,
    cholesky_decomposition_enum type)
{
    return chol_ws<float>(lda, type);
}

namespace rocm_native
{

template <typename T> size_t linear_solver_workspace_size(linalg_int lda, linalg_int nrhs, linear_solver_type type)
{
    size_t b = 0;
    check(ops<T>::svm_ws(lda, nrhs, solver_type(type), &b), "rocm_native::linear_solver_workspace_size");
    return b;
}

template <typename T>
void linear_solver_apply(T* m, linalg_int* pivot, linalg_int lda, T* x, linalg_int nrhs, linalg_int ldx,
    linear_solver_type type, void* workspace, size_t workspace_size, void* stream)
{
    if (m == nullptr || x == nullptr) throw std::invalid_argument("rocm_native::linear_solver_apply: null m or x");
    check(ops<T>::svm(m, reinterpret_cast<int64_t*>(pivot), lda, x, nrhs, ldx, solver_type(type), workspace,
              workspace_size, stream),
        "rocm_native::linear_solver_apply");
}

template size_t linear_solver_workspace_size<double>(linalg_int, linalg_int, linear_solver_type);
template size_t linear_solver_workspace_size<float>(linalg_int, linalg_int, linear_solver_type);
template void linear_solver_apply<double>(double*, linalg_int*, linalg_int, double*, linalg_int, linalg_int,
    linear_solver_type, void*, size_t, void*);
template void linear_solver_apply<float>(float*, linalg_int*, linalg_int, float*, linalg_int, linalg_int,
    linear_solver_type, void*, size_t, void*);

size_t gesv_mixed_workspace_size(linalg_int n, int gmres_restart, bool fallback)
{
    size_t b = 0;
    check(lrn_gesv_mixed_workspace_size_f64(n, gmres_restart, fallback ? 1 : 0, &b),
        "rocm_native::gesv_mixed_workspace_size");
    return b;
}

mixed_result gesv_mixed(const double* a, linalg_int n, const double* b, double* x, void* workspace,
    size_t workspace_size, double tol, int max_iter, int gmres_restart, void* stream)
{
    lrn_mixed_stats st{};
    check(lrn_gesv_mixed_f64(a, n, b, x, tol, max_iter, gmres_restart, workspace, workspace_size, &st, stream),
        "rocm_native::gesv_mixed");
    return mixed_result{st.info, st.path, st.ir_iterations, st.gmres_iterations, st.backward_error};
}

void batched_matrix_multiplication(linalg_int dim, linalg_int count, const float* a, const float* b, float* c,
    void* stream)
{
    check(lrn_batched_matrix_multiplication_f32(dim, count, a, b, c, stream), "batched_matrix_multiplication");
}
void batched_matrix_multiplication(linalg_int dim, linalg_int count, const double* a, const double* b, double* c,
    void* stream)
{
    check(lrn_batched_matrix_multiplication_f64(dim, count, a, b, c, stream), "batched_matrix_multiplication");
}

}  // namespace rocm_native
}  // namespace linalg
