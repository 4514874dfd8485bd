/*
 * C ABI of the MI355X-native (gfx950) LU solve path.
 *
 * Plain pointers and sizes only. Every matrix is row-major and square with its order as
 * its row stride (lda == n), exactly like the reference's lu_decomposition / linear_solver
 * contract. All array arguments are DEVICE pointers except `info`, which is host memory.
 * The caller owns every buffer, workspace included: no entry point allocates or frees
 * memory. Pivots are 1-based int64 LAPACK-style sequential row interchanges.
 *
 * Return value: LRN_OK (0) on success; -i when argument i (1-based) is invalid; a positive
 * LRN_E_* code for runtime failures. A numerically singular matrix is NOT an error: the
 * factorisation completes and reports the first zero pivot through `info`.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   include/matrix_operation/lu_decomposition.h:38-53   lu_decomposition<T, B>::apply / workspace_size
 *   include/matrix_operation/linear_solver.h:46-64      linear_solver<T, B>::apply / workspace_size
 *   src/matrix_operation/gpu/hipblas/lu_decomposition_hip.cpp:152-204  (ROCM_BLAS specialisation)
 *   src/matrix_operation/gpu/hipblas/linear_solver_hip.cpp:242-306     (ROCM_BLAS specialisation)
 */
#ifndef LINALG_ROCM_NATIVE_H
#define LINALG_ROCM_NATIVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define LRN_API __declspec(dllexport)
#else
#define LRN_API __attribute__((visibility("default")))
#endif

enum lrn_status {
    LRN_OK = 0,
    LRN_E_HIP = 1,          /* a HIP runtime call failed */
    LRN_E_TIMEOUT = 2,      /* a bounded inter-workgroup wait expired (result invalid) */
    LRN_E_UNSUPPORTED = 3,  /* size outside what this build supports */
    LRN_E_INTERNAL = 4,
    LRN_E_SCRATCH = 5       /* order too large for the allocation-free entry point's scratch */
};

/* linear_solver_type values, identical to the reference enum
 * (include/matrix_operation/linear_solver.h:15-21). */
enum lrn_solver_type {
    LRN_LU_LINEAR_SOLVER = 1,
    LRN_LU_UPFRONT_LINEAR_SOLVER = 2,
    LRN_CHOLESKY_LINEAR_SOLVER = 3,
    LRN_CHOLESKY_UPFRONT_LINEAR_SOLVER = 4
};

/* Human-readable text of a status code (static storage). */
LRN_API const char* lrn_status_string(int status);

/* ---- lu_decomposition<T, ROCM_NATIVE> ------------------------------------------------
 * Replaces lu_decomposition<T, B>::workspace_size (lu_decomposition.h:53) and
 * lu_decomposition<T, B>::apply (lu_decomposition.h:38-45; ROCM_BLAS version
 * lu_decomposition_hip.cpp:152-192). In-place packed L\U (unit L below the diagonal),
 * device pivots, host info. SYNCHRONOUS on `stream` (NULL = the null stream): returns
 * after the factorisation has finished and *info has been written, as the reference's
 * GPU specialisation does. */
LRN_API int lrn_lu_decomposition_workspace_size_f64(int64_t lda, size_t* bytes);
LRN_API int lrn_lu_decomposition_workspace_size_f32(int64_t lda, size_t* bytes);
LRN_API int lrn_lu_decomposition_f64(double* m, int64_t lda, int64_t* pivot, void* workspace,
    size_t workspace_size, int* info, void* stream);
LRN_API int lrn_lu_decomposition_f32(float* m, int64_t lda, int64_t* pivot, void* workspace,
    size_t workspace_size, int* info, void* stream);

/* ---- linear_solver<T, ROCM_NATIVE> ---------------------------------------------------
 * Replaces linear_solver<T, B>::workspace_size (linear_solver.h:64) and ::apply
 * (linear_solver.h:46-53; ROCM_BLAS version linear_solver_hip.cpp:242-292). One right-hand
 * side `x` of lda values, overwritten with the solution. LU mode factors `m` in place
 * (pivots to `pivot`) and solves only if the matrix is nonsingular, leaving x unchanged
 * otherwise (linear_solver_native.cpp:160-171); LU_UPFRONT reuses the caller's factors
 * and pivots. ASYNCHRONOUS on `stream`. Cholesky modes (linear_solver_native.cpp:177-189): the
 * lower factor (CHOLESKY factors m in place first, as lrn_cholesky_decomposition with
 * LOWER, and solves only on success; CHOLESKY_UPFRONT reads the caller's lower factor), then
 * L y = b and L^T x = y; `pivot` is not used and may be NULL. */
LRN_API int lrn_linear_solver_workspace_size_f64(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_linear_solver_workspace_size_f32(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_linear_solver_f64(double* m, int64_t* pivot, int64_t lda, double* x, int type,
    void* workspace, size_t workspace_size, void* stream);
LRN_API int lrn_linear_solver_f32(float* m, int64_t* pivot, int64_t lda, float* x, int type,
    void* workspace, size_t workspace_size, void* stream);

/* Allocation-free variant of lrn_linear_solver_f64 (no counterpart in the reference): for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Same arguments, semantics and results (bitwise) as
 * lrn_linear_solver_f64, minus the workspace: the work areas live in a 4 MiB device array of
 * the library (one per GPU, held until the process ends), and the factorisation runs on
 * `stream` alone. Orders whose workspace does not fit return LRN_E_SCRATCH
 * (lrn_linear_solver_noalloc_max_order gives the largest order per solver type). The call
 * allocates nothing. Calls that overlap on one GPU share that GPU's array and RACE: keep them
 * on one stream or otherwise ordered. Argument errors: -1 m, -2 pivot (LU modes), -3 lda,
 * -4 x, -5 type. ASYNCHRONOUS on stream. */
LRN_API int lrn_linear_solver_noalloc_f64(double* m, int64_t* pivot, int64_t lda, double* x, int type,
    void* stream);
LRN_API int lrn_linear_solver_noalloc_max_order(int type, int64_t* order);

/* ---- triangular_solver<T, ROCM_NATIVE> ------------------------------------------------
 * Replaces triangular_solver<T, B>::apply / workspace_size
 * (include/matrix_operation/triangular_solver.h:20-38): op(A) X = B, A n x n triangular
 * (upper or lower, op = transpose or not, unit diagonal implicit and not read), B and X
 * n x nrhs, all row-major with row strides lda, ldb, ldx; only the selected triangle of A
 * is read; A and B are preserved; A, B, X distinct device buffers. Argument errors as the
 * reference's checks (src/common/triangular_solver_checks.h:11-35), in argument order
 * (-4 n, -5 nrhs, -6 a, -7 lda, -8 b, -9 ldb, -10 x, -11 ldx, -12 workspace). Singular A is
 * not diagnosed. ASYNCHRONOUS on stream. */
LRN_API int lrn_triangular_solver_workspace_size_f64(int64_t n, int64_t nrhs, size_t* bytes);
LRN_API int lrn_triangular_solver_workspace_size_f32(int64_t n, int64_t nrhs, size_t* bytes);
LRN_API int lrn_triangular_solver_f64(int upper, int transpose, int unit, int64_t n, int64_t nrhs, const double* a,
    int64_t lda, const double* b, int64_t ldb, double* x, int64_t ldx, void* workspace, size_t workspace_size,
    void* stream);
LRN_API int lrn_triangular_solver_f32(int upper, int transpose, int unit, int64_t n, int64_t nrhs, const float* a,
    int64_t lda, const float* b, int64_t ldb, float* x, int64_t ldx, void* workspace, size_t workspace_size,
    void* stream);

/* ---- cholesky_decomposition<T, ROCM_NATIVE> ---------------------------------------------
 * Replaces cholesky_decomposition<T, B>::apply / workspace_size
 * (include/matrix_operation/cholesky_decomposition.h; algorithm src/matrix_operation/cpu/
 * native/cholesky_decomposition_native.cpp:245-320). C row-major lda x lda, only its lower
 * triangle read; upper = 0 writes L (A = L L^T) with zeros above the diagonal, upper = 1 writes
 * U = L^T (A = U^T U) with zeros below. *info (host, may be NULL) = 0 on success, else the
 * 1-based order of the first leading minor whose pivot is not positive (the factor is then
 * incomplete and C unspecified). Argument errors: -1 C, -2 lda, -3 upper, -4 workspace.
 * SYNCHRONOUS on stream (info is returned on the host, as lu_decomposition does). */
LRN_API int lrn_cholesky_decomposition_workspace_size_f64(int64_t lda, size_t* bytes);
LRN_API int lrn_cholesky_decomposition_workspace_size_f32(int64_t lda, size_t* bytes);
LRN_API int lrn_cholesky_decomposition_f64(double* c, int64_t lda, int upper, void* workspace, size_t workspace_size,
    int* info, void* stream);
LRN_API int lrn_cholesky_decomposition_f32(float* c, int64_t lda, int upper, void* workspace, size_t workspace_size,
    int* info, void* stream);

/* Asynchronous form used by the C++ cholesky_decomposition<T, ROCM_NATIVE> specialisation:
 * replaces the reference's GPU specialisation (src/matrix_operation/gpu/hipblas/
 * cholesky_decomposition_hip.cpp:69-100,135-170), whose status is caller-owned DEVICE memory
 * written on `stream` without a host synchronisation. Same factorisation and layout as
 * lrn_cholesky_decomposition_*; *info_dev (device, required: -6 when NULL) receives 0 or the
 * failing order. ASYNCHRONOUS on stream. */
LRN_API int lrn_cholesky_decomposition_async_f64(double* c, int64_t lda, int upper, void* workspace,
    size_t workspace_size, int* info_dev, void* stream);
LRN_API int lrn_cholesky_decomposition_async_f32(float* c, int64_t lda, int upper, void* workspace,
    size_t workspace_size, int* info_dev, void* stream);

/* ---- batched_matrix_multiplication -----------------------------------------------------
 * Replaces batched_matrix_multiplication (include/matrix_operation/multiplication_batched.h:
 * 7-42; ROCM_BLAS version src/matrix_operation/gpu/hipblas/multiplication_batched_hip.hip):
 * C_i = A_i B_i for i < count, each dim x dim row-major, matrix i at offset i * dim * dim.
 * Each output is an fma chain in increasing k. Argument errors, pointers first as the
 * reference checks them: -3 a, -4 b, -5 c, then -1 dim, -2 count. ASYNCHRONOUS on stream. */
LRN_API int lrn_batched_matrix_multiplication_f64(int64_t dim, int64_t count, const double* a, const double* b,
    double* c, void* stream);
LRN_API int lrn_batched_matrix_multiplication_f32(int64_t dim, int64_t count, const float* a, const float* b,
    float* c, void* stream);

/* ---- determinant<T, ROCM_NATIVE> -------------------------------------------------------
 * Replaces determinant<T, B>::apply / workspace_size (include/matrix_operation/determinant.h:
 * 13-55; ROCM_BLAS version src/matrix_operation/gpu/hipblas/inversion_hip.hip:36-58,440-482).
 * LU mode factors m in place (as lu_decomposition; pivots to `pivot`, or to the workspace
 * when pivot is NULL, since the GPU reference ignores that argument); LU_UPFRONT reads the
 * caller's factors and pivots (pivot required). *det = product of U's diagonal in increasing
 * row order, negated for an odd number of interchanges, 0 when the factorisation finds a
 * zero pivot: the GPU reference's result. The CPU reference (src/matrix_operation/cpu/
 * determinant.cpp:19-28) omits the interchange sign (determinant.h:35-36), so the two agree
 * up to sign. Cholesky modes: CHOLESKY factors m as lrn_cholesky_decomposition (LOWER) and
 * CHOLESKY_UPFRONT reads the caller's lower factor; *det = (product of L's diagonal)^2, the GPU
 * reference's value (inversion_hip.hip:62-80; the CPU reference returns the unsquared product),
 * 0 when the factorisation fails. Argument errors: -1 m, -2 pivot, -3 lda, -4 type,
 * -5 workspace, -7 det. SYNCHRONOUS on stream (the value is returned on the host, as the
 * reference's apply returns it). */
LRN_API int lrn_determinant_workspace_size_f64(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_determinant_workspace_size_f32(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_determinant_f64(double* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, double* det, void* stream);
LRN_API int lrn_determinant_f32(float* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, float* det, void* stream);

/* ---- inversion<T, ROCM_NATIVE> ---------------------------------------------------------
 * Replaces inversion<T, B>::apply / workspace_size (include/matrix_operation/inversion.h:
 * 13-53; CPU version src/matrix_operation/cpu/native/inversion_native.cpp:229-304,411-436).
 * m (row-major lda x lda) is overwritten with A^-1. LU mode factors first (pivots as for
 * the determinant) and, on a zero pivot, leaves the factors in m; LU_UPFRONT inverts the
 * caller's factors and pivots. Cholesky modes: A^-1 = L^-T L^-1 from the lower factor (CHOLESKY
 * factors first and leaves m incomplete on failure; CHOLESKY_UPFRONT reads the caller's lower
 * factor); the whole symmetric inverse is written, so its lower triangle is the reference's
 * result. Argument errors: -1 m, -2 pivot, -3 lda, -4 type, -5 workspace. Workspace holds one
 * lda x lda matrix besides the factorisation's. SYNCHRONOUS on stream. */
LRN_API int lrn_inversion_workspace_size_f64(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_inversion_workspace_size_f32(int64_t lda, int type, size_t* bytes);
LRN_API int lrn_inversion_f64(double* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, void* stream);
LRN_API int lrn_inversion_f32(float* m, int64_t* pivot, int64_t lda, int type, void* workspace,
    size_t workspace_size, void* stream);

/* ---- multi-RHS extension (no reference counterpart: the reference solves one RHS;
 * docs/backend_roadmap.md:285 lists multi-RHS as future work) ---------------------------
 * Same semantics as linear_solver with B a row-major lda x nrhs block (row stride ldb >=
 * nrhs). Each column is solved as the single-RHS solver would solve it. ASYNCHRONOUS. */
LRN_API int lrn_linear_solver_multi_workspace_size_f64(int64_t lda, int64_t nrhs, int type, size_t* bytes);
LRN_API int lrn_linear_solver_multi_workspace_size_f32(int64_t lda, int64_t nrhs, int type, size_t* bytes);
LRN_API int lrn_linear_solver_multi_f64(double* m, int64_t* pivot, int64_t lda, double* b, int64_t nrhs,
    int64_t ldb, int type, void* workspace, size_t workspace_size, void* stream);
LRN_API int lrn_linear_solver_multi_f32(float* m, int64_t* pivot, int64_t lda, float* b, int64_t nrhs,
    int64_t ldb, int type, void* workspace, size_t workspace_size, void* stream);

/* ---- asynchronous factorisation (no reference counterpart) ---------------------------
 * lu_decomposition without the host synchronisation: info is written to DEVICE memory
 * (`info_dev`, one int) instead. Used by the solver and by callers that pipeline work. */
LRN_API int lrn_getrf_async_f64(double* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream);
LRN_API int lrn_getrf_async_f32(float* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream);

/* ---- FP32 LU with tournament pivoting (no reference counterpart) -----------------------
 * The factorisation used inside the mixed solver: same layout, pivot convention (1-based
 * sequential interchanges) and info as lu_decomposition, but the pivot rows of each
 * 32-column panel are chosen by a reduction tree of partial-pivoting eliminations over
 * blocks of rows (communication-avoiding LU), so they need not equal partial pivoting's.
 * info is written to DEVICE memory. ASYNCHRONOUS. */
LRN_API int lrn_getrf_tournament_workspace_size_f32(int64_t lda, size_t* bytes);
LRN_API int lrn_getrf_tournament_f32(float* m, int64_t lda, int64_t* pivot, int* info_dev, void* workspace,
    size_t workspace_size, void* stream);

/* ---- mixed-precision solve with iterative refinement (no reference counterpart) --------
 * Solves A x = b (FP64 A and b, one right-hand side) to FP64-level normwise backward error
 * with a low-precision factorisation: A is rounded to FP32, factored with FP32 panels and
 * bf16 matrix-core trailing updates accumulated in FP32; refinement then runs with the
 * residual r = b - A x, the solution update and the GMRES inner products in FP64.
 * LU-IR first; when it stops converging, GMRES-IR (GMRES preconditioned by the
 * low-precision factors, restart `gmres_restart`); if that fails too and the workspace was
 * sized with fallback = 1, a plain FP64 LU solve. A is not modified. SYNCHRONOUS.
 * Convergence: ||b - A x||_inf <= tol * (||A||_inf ||x||_inf + ||b||_inf). */
typedef struct lrn_mixed_stats {
    int info;              /* zero pivot of the low-precision factorisation (1-based), 0 = none */
    int path;              /* 0 LU-IR, 1 GMRES-IR, 2 FP64 fallback, 3 not converged */
    int ir_iterations;     /* solves with the low-precision factors during LU-IR */
    int gmres_outer;       /* GMRES-IR refinement steps */
    int gmres_iterations;  /* inner GMRES iterations, all restarts */
    double backward_error; /* final normwise backward error (the convergence measure above) */
    double factor_ms;      /* conversion + low-precision factorisation, device time */
    double refine_ms;      /* refinement (and fallback), device time */
} lrn_mixed_stats;

/* Trailing update of the low-precision factorisation (process-wide): every FP32 operand is
 * split into bf16 planes x ~ p0 + p1 + p2 and the plane products run on the bf16 matrix
 * cores with FP32 accumulation. mode 0 = one plane (bf16 operands), 1 = two planes (three
 * products, ~2^-16 relative), 2 = three planes (six products, FP32-level factors),
 * 3 = FP32 matrix instructions (the default). Modes 2 and 3 give FP32-level factors, which
 * LU-IR needs at kappa ~ 1e6; modes 0 and 1 rely on GMRES-IR there. Modes 4, 5, 6 = one, two,
 * three planes of FP8 e4m3 with one power-of-two scale per 32 elements along K of each row of
 * L21 / column of U12 (1, 3, 6 products on the block-scaled FP8 matrix cores, FP32
 * accumulation; about 4, 8, 12 significant bits per element): for well-conditioned systems. */
LRN_API int lrn_set_mixed_update(int mode);
/* Panel of the FP32 factorisation inside the mixed solver (process-wide, no reference
 * counterpart): 1 = tournament pivoting (the pivot rows of each 32-column panel are chosen
 * by a reduction tree of partial-pivoting eliminations over row blocks, so the panel needs
 * no per-column exchange between workgroups), 0 = the partial-pivoting panel of the plain
 * path (the default). Either way the factors only precondition the FP64 refinement. */
LRN_API int lrn_set_mixed_panel(int tournament);
/* Triangular solves of the mixed solver's refinement (FP32 factors applied in FP64,
 * process-wide, no reference counterpart): 0 = the 64-row substitution kernel of the plain
 * single-RHS solve; 1 = 256-row, 2 = 128-row workgroups chained through tagged hand-offs,
 * each solving its rows in 64-row sub-blocks; 3 = 64-row blocks in the inverse form: block b
 * is x_b = z_b - M1_b x_(b-+1) - M2_b x_(b-+2) with z_b = inv(D_b) (rhs_b - the blocks before),
 * M_b = inv(D_b) A[b][b-+1 or 2], all computed in FP64 from the FP32 factors once per
 * factorisation, so the step after each hand-off is a short dot product instead of a 64-step
 * substitution; 4 = mode 3 with runs of consecutive blocks placed on one XCD when all of its
 * workgroups can be resident at once (the default). */
LRN_API int lrn_set_refine_solver(int chain);
/* Kernel of the 32-column cooperative panel (process-wide, no reference counterpart; same
 * pivots and factors either way): 0 = panel rows held in registers, one row per thread of
 * 256-thread workgroups, two exchanges and three workgroup barriers per column (the default);
 * 1 = panel rows in LDS with deferred updates. */
LRN_API int lrn_set_panel_kernel(int kind);
/* The refinement's preconditioner on its own (no reference counterpart): b <- U^-1 L^-1 P b
 * with FP32 factors and pivots (as from lrn_getrf_tournament_f32 or lrn_getrf_async_f32)
 * applied in FP64 arithmetic to one FP64 right-hand side; the kernel is the one selected by
 * lrn_set_refine_solver. Device pointers. ASYNCHRONOUS. */
LRN_API int lrn_getrs_f32f64_workspace_size(int64_t n, size_t* bytes);
LRN_API int lrn_getrs_f32f64(const float* lu, int64_t n, const int64_t* pivot, double* b, void* workspace,
    size_t workspace_size, void* stream);
LRN_API int lrn_gesv_mixed_workspace_size_f64(int64_t n, int gmres_restart, int fallback, size_t* bytes);
LRN_API int lrn_gesv_mixed_f64(const double* a, int64_t n, const double* b, double* x, double tol, int max_iter,
    int gmres_restart, void* workspace, size_t workspace_size, lrn_mixed_stats* stats, void* stream);

/* ---- batched solve of many small systems (no reference counterpart) --------------------
 * `count` independent systems A_i X_i = B_i solved in one persistent kernel (plus a small
 * ordering kernel; both are stream-ordered launches with no memset and no host sync, so
 * the call can be captured in a hipGraph once the first call has run outside capture).
 * System i: order n[i] in [1, 256], nrhs[i] in [0, 8] right-hand sides; A_i row-major
 * n_i x n_i at a + a_offset[i] (elements), factored in place (packed L\U as
 * lu_decomposition); pivots (1-based, as lu_decomposition) at pivot + pivot_offset[i];
 * B_i row-major n_i x nrhs_i at b + b_offset[i], overwritten with X_i only when the
 * factorisation is nonsingular (linear_solver LU mode); info[i] = 0, the first zero pivot
 * (1-based), or -1 for an n / nrhs outside the supported range (system skipped).
 * Every array is device memory. Asynchronous. Pivot rule, elimination and solve order
 * follow lu_decomposition / linear_solver. */
LRN_API int lrn_batched_gesv_workspace_size(int64_t count, size_t* bytes);
LRN_API int lrn_batched_gesv_f64(double* a, const int64_t* a_offset, int64_t* pivot, const int64_t* pivot_offset,
    double* b, const int64_t* b_offset, const int32_t* n, const int32_t* nrhs, int64_t count, int32_t* info,
    void* workspace, size_t workspace_size, void* stream);
LRN_API int lrn_batched_gesv_f32(float* a, const int64_t* a_offset, int64_t* pivot, const int64_t* pivot_offset,
    float* b, const int64_t* b_offset, const int32_t* n, const int32_t* nrhs, int64_t count, int32_t* info,
    void* workspace, size_t workspace_size, void* stream);
/* Workgroups of the persistent batched launch (0 = occupancy x compute units, the default);
 * used by the autotuner. Process-wide. */
LRN_API int lrn_set_batched_grid(int workgroups);

/* Many mixed-precision solves in one call (the batched mixed-precision form): system i
 * (order n[i], FP64 A_i at a + a_offset[i], b_i at b + b_offset[i], solution to
 * x + x_offset[i]) is solved as lrn_gesv_mixed_f64 would, in order, on one stream, with
 * its own trailing-update mode (update_mode[i] as lrn_set_mixed_update; NULL = the
 * process-wide mode) and its own statistics in stats[i]. `a_offset`, `n`, `b_offset`,
 * `x_offset`, `update_mode` and `stats` are HOST arrays; a, b, x device memory. The
 * workspace is sized for the largest order (lrn_gesv_mixed_workspace_size_f64).
 * SYNCHRONOUS. */
LRN_API int lrn_gesv_mixed_batched_f64(const double* a, const int64_t* a_offset, const int64_t* n, const double* b,
    const int64_t* b_offset, double* x, const int64_t* x_offset, int64_t count, const int32_t* update_mode, double tol,
    int max_iter, int gmres_restart, void* workspace, size_t workspace_size, lrn_mixed_stats* stats, void* stream);

/* Device tuning knobs (outer block, panel width, panel workgroups) for the autotuner. A
 * value <= 0 restores the built-in choice. Not thread-safe against concurrent factorisations. */
LRN_API int lrn_set_tuning(int outer_nb, int panel_w, int panel_max_wg);
LRN_API int lrn_get_tuning(int* outer_nb, int* panel_w, int* panel_max_wg);
/* Look-ahead: the next block column is updated and factored on helper streams while the
 * rest of the trailing update runs (results are bitwise those without it). mode 0 = off,
 * 1 = on with 32 compute units for the panel stream, a multiple of 8 (>= 8) = on with that
 * many; orders above 16384 use at least 64. Process-wide; the autotuner picks the value. */
LRN_API int lrn_set_lookahead(int mode);
/* Update of the rest of a block column after each of its inner panels: 0 = 64 x 64 tiles on
 * the vector units (the default), 1 = the trailing-update GEMM (128 x 128 matrix-core tiles),
 * 2 = 64 x 64 matrix-core tiles. Results differ only in rounding of the update's sums.
 * Process-wide; the autotuner picks the value. */
LRN_API int lrn_set_block_update(int kind);
/* Refinement of lrn_gesv_mixed_*: 0 = LU-IR while every step at least halves the backward
 * error, then GMRES-IR from the best iterate (the default); 1 = GMRES-IR right after the
 * first LU-IR step. Process-wide. */
LRN_API int lrn_set_refinement(int mode);
/* Many-right-hand-side solves (LU and Cholesky solver modes, triangular_solver): solves with at
 * least min_rhs (and at least 2) right-hand sides use the blocked form (64 x 64 diagonal blocks
 * inverted, then one matrix-core update launch per 64-row strip); fewer, or min_rhs = 0, the
 * chained substitution kernels. Default 2. Process-wide. */
LRN_API int lrn_set_multi_rhs_solver(int min_rhs);
/* Kernel of that blocked form (process-wide): 1 = one launch per 64 right-hand sides, the
 * 64-row strips chained through flags between workgroups (bounded waits; the default, used when
 * all of its workgroups can be resident at once), 0 = one launch per 64-row strip. */
LRN_API int lrn_set_multi_rhs_kernel(int kind);

/* ---- kernel test hooks (no reference counterpart) -------------------------------------
 * Single-kernel entry points used by the kernel-level unit tests; dtype 0 = f64, 1 = f32,
 * 2, 3, 4 = f32 data split into 1, 2, 3 bf16 planes, FP32 accumulation; 5, 6, 7 = f32 data
 * split into 1, 2, 3 block-scaled FP8 e4m3 planes (scale per 32 K elements), FP32
 * accumulation (gemm_sub only).
 * gemm_sub: C[M x N] -= A[M x K] * B[K x N] (row-major, leading dims lda/ldb/ldc).
 * trsm: rows k..k+w-1, columns [c0, c1) of the lda x lda matrix a <- L11^-1 a, with L11 the
 *       unit-lower w x w block at (k, k).
 * panel: factor the column block a[k:lda, k:k+w] with partial pivoting (pivot[k..k+w) set);
 *        `workspace` must hold lrn_lu_decomposition_workspace_size bytes. Asynchronous. */
LRN_API int lrn_debug_gemm_sub(int dtype, const void* a, int64_t lda, const void* b, int64_t ldb, void* c,
    int64_t ldc, int64_t m, int64_t n, int64_t k, void* stream);
LRN_API int lrn_debug_trsm(int dtype, void* a, int64_t lda, int64_t k, int w, int64_t c0, int64_t c1, void* stream);
LRN_API int lrn_debug_panel(int dtype, void* a, int64_t lda, int64_t k, int w, int64_t* pivot, void* workspace,
    size_t workspace_size, void* stream);

/* ---- kernel timing trace (no reference counterpart) ------------------------------------
 * While a trace runs, each launch of a traced kernel kind is bracketed by two HIP events on
 * the stream it is launched on, and its algorithmic work is logged. `kinds` is a bit mask
 * of LRN_TRACE_* values; `capacity` bounds the number of traced launches (events are
 * created by lrn_trace_start, outside any timed region; launches beyond it go untimed).
 * lrn_trace_read synchronises on the events of one kind and returns the number of
 * launches, their summed duration (ms) and summed work (flops for GEMM and panel kinds,
 * bytes for the others). Process-wide; for measurement, not for concurrent use. */
enum lrn_trace_kind {
    LRN_TRACE_GEMM_F64 = 0,        /* FP64 trailing GEMM, interior tiles (flops 2MNK) */
    LRN_TRACE_GEMM_F64_EDGE = 1,   /* FP64 trailing GEMM, ragged strips (flops) */
    LRN_TRACE_GEMM_BF16X3 = 2,     /* two-plane bf16 trailing GEMM (FP32-equivalent flops 2MNK) */
    LRN_TRACE_GEMM_BF16 = 3,       /* one-plane bf16 trailing GEMM (flops) */
    LRN_TRACE_GEMM_F32 = 4,        /* FP32 GEMM (flops) */
    LRN_TRACE_PANEL = 5,           /* panel factorisation (flops) */
    LRN_TRACE_LASWP = 6,           /* row interchanges (bytes, every pivot taken as a swap) */
    LRN_TRACE_TRSM = 7,            /* U12 triangular solve (bytes) */
    LRN_TRACE_TRSV = 8,            /* solve substitutions (bytes) */
    LRN_TRACE_GEMM_BF16X6 = 9,     /* three-plane bf16 trailing GEMM (FP32-equivalent flops 2MNK) */
    LRN_TRACE_BATCHED = 10,        /* batched solve kernel (work 0: not known on the host) */
    LRN_TRACE_BLOCK_UPDATE = 11,   /* update inside a block column (flops) */
    LRN_TRACE_GEMM_FP8 = 12,       /* one-plane block-scaled FP8 trailing GEMM (flops 2MNK) */
    LRN_TRACE_GEMM_FP8X3 = 13,     /* two-plane FP8 trailing GEMM (FP32-equivalent flops 2MNK) */
    LRN_TRACE_GEMM_FP8X6 = 14      /* three-plane FP8 trailing GEMM (FP32-equivalent flops 2MNK) */
};
LRN_API int lrn_trace_start(unsigned kinds, int64_t capacity);
LRN_API int lrn_trace_stop(void);
LRN_API int lrn_trace_read(int kind, int64_t* launches, double* milliseconds, double* work);

/* Polls each bounded inter-workgroup wait makes before it gives up and raises the error
 * word (no reference counterpart; for tests of the error paths): limit < 0 restores the
 * default (2^20), 0 makes every wait expire at once, so the factorisation reports
 * LRN_E_TIMEOUT and the asynchronous solvers leave NaN in their outputs. Process-wide. */
LRN_API int lrn_debug_set_spin_limit(int64_t limit);

/* Version string of this library. */
LRN_API const char* lrn_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LINALG_ROCM_NATIVE_H */
