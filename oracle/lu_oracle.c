/*
 * CPU restatement of the reference's native FP64/FP32 LU path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: the checker that GPU results are compared against. Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark load it. The
 * product library never links, calls or falls back to it.
 *
 * It restates, in code of its own, the algorithm and floating-point operation order of
 * the reference's native path built with pivoting ON and vectorization OFF:
 *   - panel factorisation        src/matrix_operation/cpu/native/lu_decomposition_native.cpp:158-232
 *   - deferred row interchanges  lu_decomposition_native.cpp:247-271
 *   - blocked driver (nb = 64)   lu_decomposition_native.cpp:285,329-429
 *   - trailing GEMM, i-k-j order src/matrix_operation/cpu/native/multiplication_native.cpp:1546-1568,1703-1720
 *   - solve                      src/matrix_operation/cpu/native/linear_solver_native.cpp:41-113,149-191
 * Every product is rounded before the following add/subtract: build with
 * -ffp-contract=off and no -march flag so that no fused multiply-add is formed.
 *
 * OpenMP only splits loops whose iterations touch disjoint elements (rows of the GEMM,
 * rows of the panel update, column ranges of the triangular solve), so the result is
 * bit-for-bit the same for every thread count.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef long long lint;

static int g_threads = 1;

void ora_set_threads(int t) { g_threads = t > 0 ? t : 1; }
int ora_get_threads(void) { return g_threads; }

/* Minimum row count below which parallel regions are not worth their overhead. */
#define ORA_PAR_ROWS 512

/* Reference block size: lu_decomposition_native.cpp:285. */
#define ORA_NB 64

#define ORA_DEFINE(T, SFX, ABS, TMIN)                                                         \
    /* y[c] -= a * x[c]: lu_decomposition_native.cpp:116-120 (unvectorised branch). */        \
    static inline void axpy_sub_##SFX(T* restrict y, const T* restrict x, T a, lint len)    \
    {                                                                                        \
        for (lint c = 0; c < len; ++c) y[c] -= a * x[c];                                     \
    }                                                                                        \
                                                                                             \
    /* lu_decomposition_native.cpp:158-232. When forced != NULL the pivot row of step i is  \
       forced[i]-1 instead of the column argmax (replay mode); the true column max and the   \
       magnitude of the chosen entry are stored in colmax[i] / chosen[i] when non-NULL. */   \
    static void panel_factor_##SFX(T* restrict m, lint lda, lint start, lint width,          \
        lint* restrict pivot, lint* first_zero, const lint* forced, double* colmax,          \
        double* chosen)                                                                      \
    {                                                                                        \
        for (lint ii = 0; ii < width; ++ii) {                                                \
            const lint i = start + ii;                                                       \
            T* const row_i = m + i * lda;                                                    \
            /* strict '>' scan, so the first maximum wins; initial max is |diag| */          \
            T maxA = ABS(row_i[i]);                                                          \
            lint imax = i;                                                                   \
            const T* col = m + (i + 1) * lda + i;                                            \
            for (lint k = i + 1; k < lda; ++k, col += lda) {                                 \
                const T a = ABS(*col);                                                       \
                if (a > maxA) {                                                              \
                    maxA = a;                                                                \
                    imax = k;                                                                \
                }                                                                            \
            }                                                                                \
            if (forced) {                                                                    \
                const lint f = forced[i] - 1;                                                \
                if (colmax) colmax[i] = (double)maxA;                                        \
                if (chosen) chosen[i] = (double)ABS(m[f * lda + i]);                         \
                imax = f;                                                                    \
            }                                                                                \
            pivot[i] = imax + 1;                                                             \
            if (imax != i) {                                                                 \
                T* const row_p = m + imax * lda;                                             \
                for (lint j = 0; j < start + width; ++j) {                                   \
                    const T t = row_i[j];                                                    \
                    row_i[j] = row_p[j];                                                     \
                    row_p[j] = t;                                                            \
                }                                                                            \
            }                                                                                \
            const T diag = row_i[i];                                                         \
            if (diag == (T)0) {                                                              \
                if (*first_zero < 0) *first_zero = i;                                        \
                continue;                                                                    \
            }                                                                                \
            /* reciprocal unless |diag| is below the smallest normal, then divide */        \
            const int use_recip = ABS(diag) >= TMIN;                                         \
            const T norm = use_recip ? (T)((T)1 / diag) : (T)1;                              \
            const T* const l_i = row_i + start;                                              \
            const lint nrows = lda - (i + 1);                                                \
            _Pragma("omp parallel for schedule(static) num_threads(g_threads) if(nrows > ORA_PAR_ROWS)") \
            for (lint r = i + 1; r < lda; ++r) {                                             \
                T* const l_r = m + r * lda + start;                                          \
                if (use_recip)                                                               \
                    l_r[ii] *= norm;                                                         \
                else                                                                         \
                    l_r[ii] /= diag;                                                         \
                axpy_sub_##SFX(l_r + ii + 1, l_i + ii + 1, l_r[ii], width - ii - 1);         \
            }                                                                                \
        }                                                                                    \
    }                                                                                        \
                                                                                             \
    /* lu_decomposition_native.cpp:247-271: replay the panel's swaps on [start+width, lda). */ \
    static void apply_pivots_trailing_##SFX(T* restrict m, lint lda, lint start, lint width, \
        const lint* restrict pivot)                                                          \
    {                                                                                        \
        for (lint i = start; i < start + width; ++i) {                                       \
            const lint p = pivot[i] - 1;                                                     \
            if (p != i) {                                                                    \
                T* const ri = m + i * lda;                                                   \
                T* const rp = m + p * lda;                                                   \
                for (lint j = start + width; j < lda; ++j) {                                 \
                    const T t = ri[j];                                                       \
                    ri[j] = rp[j];                                                           \
                    rp[j] = t;                                                               \
                }                                                                            \
            }                                                                                \
        }                                                                                    \
    }                                                                                        \
                                                                                             \
    /* lu_decomposition_native.cpp:329-429. Returns first zero pivot (0-based) or -1. */     \
    static lint getrf_impl_##SFX(T* restrict m, lint lda, lint* restrict pivot,             \
        const lint* forced, double* colmax, double* chosen, int* oom)                        \
    {                                                                                        \
        for (lint i = 0; i < lda; ++i) pivot[i] = i + 1;                                     \
        lint first_zero = -1;                                                                \
        for (lint k = 0; k < lda; k += ORA_NB) {                                             \
            const lint bs = (lda - k) < ORA_NB ? (lda - k) : ORA_NB;                         \
            panel_factor_##SFX(m, lda, k, bs, pivot, &first_zero, forced, colmax, chosen);   \
            const lint rest = lda - k - bs;                                                  \
            if (rest == 0) continue;                                                         \
            apply_pivots_trailing_##SFX(m, lda, k, bs, pivot);                               \
            /* U12 = L11^-1 A12, row by row (:370-380); split by column ranges only */       \
            {                                                                                \
                const lint chunk = 256;                                                      \
                const lint nchunks = (rest + chunk - 1) / chunk;                             \
                _Pragma("omp parallel for schedule(static) num_threads(g_threads) if(rest > ORA_PAR_ROWS)") \
                for (lint c = 0; c < nchunks; ++c) {                                         \
                    const lint c0 = c * chunk;                                               \
                    const lint len = (rest - c0) < chunk ? (rest - c0) : chunk;              \
                    for (lint ii = 0; ii < bs; ++ii) {                                       \
                        T* const row_full = m + (k + ii) * lda;                              \
                        T* const row_i = row_full + k + bs + c0;                             \
                        for (lint p = 0; p < ii; ++p) {                                      \
                            const T l_ip = row_full[k + p];                                  \
                            const T* row_p = m + (k + p) * lda + k + bs + c0;                \
                            axpy_sub_##SFX(row_i, row_p, l_ip, len);                         \
                        }                                                                    \
                    }                                                                        \
                }                                                                            \
            }                                                                                \
            /* delta = L21 * U12 (beta = 0 zero-fill, then i-k-j accumulate, alpha = 1),     \
               then A22 -= delta (:392-422; multiplication_native.cpp:1546-1568,1703-1720).  \
               Each delta row is finished before its A22 row is touched and no delta row     \
               reads A22, so a per-thread row buffer replaces the rest x rest scratch. */    \
            _Pragma("omp parallel num_threads(g_threads) if(rest > ORA_PAR_ROWS)")          \
            {                                                                                \
                T* const c_i = (T*)malloc(sizeof(T) * (size_t)rest);                         \
                if (!c_i) *oom = 1;                                                          \
                _Pragma("omp for schedule(static)")                                          \
                for (lint r = 0; r < rest; ++r) {                                            \
                    if (!c_i) continue;                                                      \
                    const T* const a_i = m + (k + bs + r) * lda + k;                         \
                    for (lint j = 0; j < rest; ++j) c_i[j] = (T)0;                           \
                    for (lint p = 0; p < bs; ++p) {                                          \
                        const T* const b_k = m + (k + p) * lda + k + bs;                     \
                        const T s = (T)1 * a_i[p];                                           \
                        for (lint j = 0; j < rest; ++j) c_i[j] += s * b_k[j];                \
                    }                                                                        \
                    T* const out = m + (k + bs + r) * lda + k + bs;                          \
                    axpy_sub_##SFX(out, c_i, (T)1, rest);                                    \
                }                                                                            \
                free(c_i);                                                                   \
            }                                                                                \
        }                                                                                    \
        return first_zero;                                                                   \
    }                                                                                        \
                                                                                             \
    /* Workspace of the reference's native LU: max over panels of (n-k-64)^2 * sizeof(T)    \
       (lu_decomposition_native.cpp:293-326; the unvectorised GEMM needs none). */          \
    size_t ora_getrf_workspace_##SFX(lint lda)                                               \
    {                                                                                        \
        size_t best = 0;                                                                     \
        for (lint k = 0; k < lda; k += ORA_NB) {                                             \
            const lint bs = (lda - k) < ORA_NB ? (lda - k) : ORA_NB;                         \
            if (k + bs >= lda) continue;                                                     \
            const size_t rest = (size_t)(lda - k - bs);                                      \
            const size_t b = rest * rest * sizeof(T);                                        \
            if (b > best) best = b;                                                          \
        }                                                                                    \
        return best;                                                                         \
    }                                                                                        \
                                                                                             \
    /* Returns info (0 or 1-based first zero pivot); -1 if a row buffer allocation failed. */     \
    int ora_getrf_##SFX(T* m, lint lda, lint* pivot)                                         \
    {                                                                                        \
        int oom = 0;                                                                         \
        const lint fz = getrf_impl_##SFX(m, lda, pivot, NULL, NULL, NULL, &oom);             \
        return oom ? -1 : (int)(fz + 1);                                                     \
    }                                                                                        \
                                                                                             \
    /* Replay mode: the same factorisation with the pivot row of every step forced to        \
       forced[i]; colmax[i] / chosen[i] receive the column max and the chosen magnitude. */  \
    int ora_getrf_forced_##SFX(T* m, lint lda, const lint* forced, lint* pivot,              \
        double* colmax, double* chosen)                                                      \
    {                                                                                        \
        int oom = 0;                                                                         \
        const lint fz = getrf_impl_##SFX(m, lda, pivot, forced, colmax, chosen, &oom);       \
        return oom ? -1 : (int)(fz + 1);                                                     \
    }                                                                                        \
                                                                                             \
    /* lu_dot, unvectorised: sequential sum from zero (linear_solver_native.cpp:41-88). */   \
    static inline T dot_##SFX(const T* a, const T* b, lint count)                            \
    {                                                                                        \
        T sum = (T)0;                                                                        \
        for (lint j = 0; j < count; ++j) sum += a[j] * b[j];                                 \
        return sum;                                                                          \
    }                                                                                        \
                                                                                             \
    /* lu_solve (linear_solver_native.cpp:90-113) on one contiguous vector. */               \
    static void lu_solve_vec_##SFX(const T* m, const lint* pivot, lint lda, T* x)            \
    {                                                                                        \
        for (lint i = 0; i < lda; ++i) {                                                     \
            const lint p = pivot[i] - 1;                                                     \
            const T t = x[i];                                                                \
            x[i] = x[p];                                                                     \
            x[p] = t;                                                                        \
        }                                                                                    \
        for (lint i = 0; i < lda; ++i) x[i] -= dot_##SFX(m + i * lda, x, i);                 \
        for (lint i = lda - 1; i >= 0; --i) {                                                \
            const T* row_i = m + i * lda;                                                    \
            x[i] = (x[i] - dot_##SFX(row_i + i + 1, x + i + 1, lda - i - 1)) / row_i[i];     \
        }                                                                                    \
    }                                                                                        \
                                                                                             \
    /* Multi-RHS extension: B is row-major lda x nrhs with row stride ldb; each column is    \
       solved exactly as the single-RHS lu_solve would solve it. */                          \
    int ora_getrs_##SFX(const T* m, lint lda, const lint* pivot, T* b, lint nrhs, lint ldb)  \
    {                                                                                        \
        _Pragma("omp parallel num_threads(g_threads) if(nrhs > 1)")                         \
        {                                                                                    \
            T* col = (T*)malloc(sizeof(T) * (size_t)lda);                                    \
            _Pragma("omp for schedule(static)")                                              \
            for (lint c = 0; c < nrhs; ++c) {                                                \
                if (!col) continue;                                                          \
                for (lint i = 0; i < lda; ++i) col[i] = b[i * ldb + c];                      \
                lu_solve_vec_##SFX(m, pivot, lda, col);                                      \
                for (lint i = 0; i < lda; ++i) b[i * ldb + c] = col[i];                      \
            }                                                                                \
            free(col);                                                                       \
        }                                                                                    \
        return 0;                                                                            \
    }                                                                                        \
                                                                                             \
    /* linear_solver LU mode (linear_solver_native.cpp:160-171): factor, and solve only when \
       the factorisation is nonsingular; x is left unchanged otherwise. Returns info. */     \
    int ora_gesv_##SFX(T* m, lint lda, lint* pivot, T* b, lint nrhs, lint ldb)               \
    {                                                                                        \
        const int info = ora_getrf_##SFX(m, lda, pivot);                                     \
        if (info == 0) ora_getrs_##SFX(m, lda, pivot, b, nrhs, ldb);                         \
        return info;                                                                         \
    }

ORA_DEFINE(double, d, fabs, DBL_MIN)
ORA_DEFINE(float, s, fabsf, FLT_MIN)
