"""ctypes front end of the C oracle plus the input generators and error metrics the tests use.

TEST INFRASTRUCTURE ONLY: see the package docstring.

Reference correspondence:
  getrf   -> lu_decomposition<T, NATIVE>::apply  (src/matrix_operation/cpu/native/lu_decomposition_native.cpp:447-459)
  getrs   -> lu_solve, one column at a time       (src/matrix_operation/cpu/native/linear_solver_native.cpp:90-113)
  gesv    -> linear_solver<T, NATIVE>::apply, LU  (linear_solver_native.cpp:160-171)
  workspace_size -> lu_workspace_size<T>          (lu_decomposition_native.cpp:293-326)
  trsm    -> triangular_solver<T, NATIVE>::apply  (src/matrix_operation/cpu/native/triangular_solver_native.cpp:64-153),
             restated in numpy (scalar-path order); pinned by the reference's known-answer
             generator (Testing/Cxx/triangular_solver_test_cases.h:12-68, restated as
             trsm_known_answer_cases) and scipy's solve_triangular
  determinant -> determinant<T, B>::apply, LU modes (src/matrix_operation/cpu/determinant.cpp:19-28
             unsigned; src/matrix_operation/gpu/hipblas/inversion_hip.hip:36-58 signed); pinned
             by the reference's known answers (Testing/Cxx/TestDeterminant.cxx:82-126,
             Testing/Cxx/TestGpuMatrixOperations.cxx:1466-1492) and numpy.linalg.det
  inverse_from_factors -> lu_invert (src/matrix_operation/cpu/native/inversion_native.cpp:229-304);
             pinned by A A^-1 = I on the reference's inversion input and numpy.linalg.inv
  cholesky, cholesky_solve -> cholesky_decomposition<T, NATIVE> / cholesky_solve
             (cholesky_decomposition_native.cpp:245-320, linear_solver_native.cpp:116-147);
             pinned by the reference's known answer (the factor of C C^T is C,
             Testing/Cxx/TestCholeskyDecomposition.cxx:170-210) and numpy.linalg.cholesky
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblu_oracle.so")
_lib = None

_i64p = ctypes.POINTER(ctypes.c_longlong)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        for sfx, fp in (("d", _f64p), ("s", _f32p)):
            f = getattr(L, f"ora_getrf_{sfx}")
            f.argtypes = [fp, ctypes.c_longlong, _i64p]
            f.restype = ctypes.c_int
            f = getattr(L, f"ora_getrf_forced_{sfx}")
            f.argtypes = [fp, ctypes.c_longlong, _i64p, _i64p, _f64p, _f64p]
            f.restype = ctypes.c_int
            f = getattr(L, f"ora_getrs_{sfx}")
            f.argtypes = [fp, ctypes.c_longlong, _i64p, fp, ctypes.c_longlong, ctypes.c_longlong]
            f.restype = ctypes.c_int
            f = getattr(L, f"ora_gesv_{sfx}")
            f.argtypes = [fp, ctypes.c_longlong, _i64p, fp, ctypes.c_longlong, ctypes.c_longlong]
            f.restype = ctypes.c_int
            f = getattr(L, f"ora_getrf_workspace_{sfx}")
            f.argtypes = [ctypes.c_longlong]
            f.restype = ctypes.c_size_t
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def set_threads(t: int) -> None:
    lib().ora_set_threads(int(t))


def _sfx(dtype) -> tuple[str, object]:
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "d", _f64p
    if dt == np.float32:
        return "s", _f32p
    raise TypeError(f"oracle supports float32/float64, not {dt}")


def _ptr(a: np.ndarray, p):
    return a.ctypes.data_as(p)


def getrf(a: np.ndarray, threads: int | None = None):
    """Factor a copy of row-major square ``a``. Returns (lu, pivot[1-based int64], info)."""
    a = np.ascontiguousarray(a)
    n = a.shape[0]
    assert a.shape == (n, n)
    sfx, fp = _sfx(a.dtype)
    if threads is not None:
        set_threads(threads)
    lu = a.copy()
    piv = np.empty(n, dtype=np.int64)
    info = getattr(lib(), f"ora_getrf_{sfx}")(_ptr(lu, fp), n, _ptr(piv, _i64p))
    if info < 0:
        raise MemoryError("oracle getrf: row buffer allocation failed")
    return lu, piv, info


def getrf_forced(a: np.ndarray, forced_pivot: np.ndarray):
    """Replay the factorisation with every step's pivot row forced to ``forced_pivot``.

    Returns (lu, pivot, info, colmax, chosen): colmax[i] is the largest |a| in column i at
    step i (rows i..n-1), chosen[i] the magnitude of the forced pivot entry.
    """
    a = np.ascontiguousarray(a)
    n = a.shape[0]
    sfx, fp = _sfx(a.dtype)
    lu = a.copy()
    forced = np.ascontiguousarray(forced_pivot, dtype=np.int64)
    piv = np.empty(n, dtype=np.int64)
    colmax = np.zeros(n)
    chosen = np.zeros(n)
    info = getattr(lib(), f"ora_getrf_forced_{sfx}")(
        _ptr(lu, fp), n, _ptr(forced, _i64p), _ptr(piv, _i64p), _ptr(colmax, _f64p), _ptr(chosen, _f64p))
    if info < 0:
        raise MemoryError("oracle getrf_forced: row buffer allocation failed")
    return lu, piv, info, colmax, chosen


def getrs(lu: np.ndarray, piv: np.ndarray, b: np.ndarray, threads: int | None = None):
    """Solve with existing factors. ``b`` is a vector (n,) or row-major (n, nrhs)."""
    lu = np.ascontiguousarray(lu)
    n = lu.shape[0]
    sfx, fp = _sfx(lu.dtype)
    if threads is not None:
        set_threads(threads)
    x = np.array(b, dtype=lu.dtype, copy=True, order="C")
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    piv = np.ascontiguousarray(piv, dtype=np.int64)
    getattr(lib(), f"ora_getrs_{sfx}")(_ptr(lu, fp), n, _ptr(piv, _i64p), _ptr(x, fp), nrhs, nrhs)
    return x


def gesv(a: np.ndarray, b: np.ndarray, threads: int | None = None):
    """linear_solver LU mode: returns (lu, piv, x, info); x == b unchanged when singular."""
    a = np.ascontiguousarray(a)
    n = a.shape[0]
    sfx, fp = _sfx(a.dtype)
    if threads is not None:
        set_threads(threads)
    lu = a.copy()
    piv = np.empty(n, dtype=np.int64)
    x = np.array(b, dtype=a.dtype, copy=True, order="C")
    nrhs = 1 if x.ndim == 1 else x.shape[1]
    info = getattr(lib(), f"ora_gesv_{sfx}")(_ptr(lu, fp), n, _ptr(piv, _i64p), _ptr(x, fp), nrhs, nrhs)
    return lu, piv, x, info


def reference_workspace_size(n: int, dtype=np.float64) -> int:
    sfx, _ = _sfx(dtype)
    return int(getattr(lib(), f"ora_getrf_workspace_{sfx}")(n))


# ----------------------------------------------------------------------------- generators
# Portable inputs: numpy's PCG64 bit generator with its standard-normal transform is a
# documented, platform-independent stream for a given seed, so host, tests and fixtures
# all see the same matrices.

def gaussian(n: int, m: int | None = None, seed: int = 0, dtype=np.float64) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (n, n) if m is None else (n, m)
    return rng.standard_normal(shape).astype(dtype, copy=False)


def diag_dominant(n: int, seed: int = 0, dtype=np.float64) -> np.ndarray:
    """A + n*I with A iid N(0,1)."""
    a = gaussian(n, seed=seed, dtype=np.float64)
    a[np.diag_indices(n)] += n
    return a.astype(dtype, copy=False)


def haar_orthogonal(n: int, seed: int) -> np.ndarray:
    """Q of the QR of a seeded Gaussian, with R's diagonal made positive (Mezzadri)."""
    g = gaussian(n, seed=seed)
    q, r = np.linalg.qr(g)
    q *= np.sign(np.diag(r))[None, :]
    return q


def conditioned(n: int, kappa: float, seed: int = 0) -> np.ndarray:
    """U diag(s) V^T with Haar U, V and s log-spaced from 1 to 1/kappa."""
    u = haar_orthogonal(n, seed)
    v = haar_orthogonal(n, seed + 1)
    s = np.logspace(0.0, -np.log10(kappa), n)
    return (u * s[None, :]) @ v.T


# ----------------------------------------------------------------------------- pivots / metrics

def pivots_to_perm(piv: np.ndarray) -> np.ndarray:
    """Row permutation p with (P^T A)[i] = A[p[i]] for LAPACK-style sequential swaps."""
    n = piv.shape[0]
    perm = np.arange(n)
    for i in range(n):
        j = int(piv[i]) - 1
        if j != i:
            perm[i], perm[j] = perm[j], perm[i]
    return perm


def unpack_lu(lu: np.ndarray):
    n = lu.shape[0]
    lo = np.tril(lu, -1).astype(np.float64)
    lo[np.diag_indices(n)] = 1.0
    up = np.triu(lu).astype(np.float64)
    return lo, up


def reconstruction_error(a: np.ndarray, lu: np.ndarray, piv: np.ndarray) -> float:
    """max|P^T A - L U| / (max|A| * n * eps) -- dimensionless, O(1) for a stable LU."""
    lo, up = unpack_lu(lu)
    perm = pivots_to_perm(piv)
    n = a.shape[0]
    eps = np.finfo(lu.dtype).eps
    diff = np.abs(a.astype(np.float64)[perm] - lo @ up).max()
    scale = max(np.abs(a).max(), np.finfo(np.float64).tiny)
    return float(diff / (scale * n * eps))


def backward_error(a: np.ndarray, x: np.ndarray, b: np.ndarray) -> float:
    """Normwise backward error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf), per column max."""
    a = a.astype(np.float64)
    x2 = x.reshape(x.shape[0], -1).astype(np.float64)
    b2 = b.reshape(b.shape[0], -1).astype(np.float64)
    r = b2 - a @ x2
    na = np.abs(a).sum(axis=1).max()
    num = np.abs(r).max(axis=0)
    den = na * np.abs(x2).max(axis=0) + np.abs(b2).max(axis=0)
    return float((num / np.where(den > 0, den, 1.0)).max())


def forward_error(x: np.ndarray, x_ref: np.ndarray) -> float:
    """||x - x_ref||_inf / ||x_ref||_inf, per column max."""
    x2 = x.reshape(x.shape[0], -1).astype(np.float64)
    r2 = x_ref.reshape(x_ref.shape[0], -1).astype(np.float64)
    num = np.abs(x2 - r2).max(axis=0)
    den = np.abs(r2).max(axis=0)
    return float((num / np.where(den > 0, den, 1.0)).max())


def rel_diff(a: np.ndarray, b: np.ndarray) -> float:
    """Normwise relative difference ||a - b||_F / ||b||_F."""
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def first_pivot_mismatch(p_gpu: np.ndarray, p_ref: np.ndarray) -> int:
    """Index of the first differing pivot, or -1."""
    d = np.nonzero(np.asarray(p_gpu) != np.asarray(p_ref))[0]
    return int(d[0]) if d.size else -1


__all__ = [
    "build", "lib", "set_threads", "getrf", "getrf_forced", "getrs", "gesv",
    "reference_workspace_size", "gaussian", "diag_dominant", "haar_orthogonal", "conditioned",
    "pivots_to_perm", "unpack_lu", "reconstruction_error", "backward_error", "forward_error",
    "rel_diff", "first_pivot_mismatch",
]


# ------------------------------------------------------------------- triangular solver

_TRSM_BLOCK = 32  # triangular_solver_native.cpp:17


def trsm(upper: bool, transpose: bool, unit: bool, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """op(A) X = B the way triangular_solver_native.cpp:64-153 does it (scalar path): rows in
    32-row blocks from the end op(A)'s triangle starts at; inside a block, row i takes
    x_i -= coefficient(i, k) * x_k for the block's earlier rows k (in increasing k), then
    x_i /= coefficient(i, i) unless unit; after the block, the rows still to do take
    update = op(A)[rows, block] @ X[block] (i-k-j order, multiplication_native.cpp:1703-1720)
    and x -= update. a is n x lda (only the triangle is read), b is n x nrhs."""
    n = a.shape[0]
    dt = a.dtype.type
    x = np.array(b[:, :], dtype=a.dtype, copy=True)
    nrhs = x.shape[1]

    def coef(i, j):
        return a[j, i] if transpose else a[i, j]

    eff_upper = upper != transpose
    done = 0
    while done < n:
        width = min(_TRSM_BLOCK, n - done)
        begin = n - done - width if eff_upper else done
        end = begin + width
        for step in range(width):
            i = end - step - 1 if eff_upper else begin + step
            ks = range(i + 1, end) if eff_upper else range(begin, i)
            for k in ks:
                x[i] = x[i] - coef(i, k) * x[k]
            if not unit:
                x[i] = x[i] / coef(i, i)
        first = 0 if eff_upper else end
        rows = begin if eff_upper else n - end
        if rows > 0:
            update = np.zeros((rows, nrhs), dtype=a.dtype)
            for r in range(rows):
                for k in range(width):
                    update[r] = update[r] + coef(first + r, begin + k) * x[begin + k]
            x[first:first + rows] = x[first:first + rows] - update
        done += width
    return x.astype(dt, copy=False)


def trsm_known_answer_cases(dtype=np.float64):
    """The reference's known-answer generator for the triangular solver
    (Testing/Cxx/triangular_solver_test_cases.h:12-68), restated: every (n, nrhs) of the list,
    padded strides lda = n + 2, ldb = nrhs + 1, ldx = nrhs + 3, all eight flag combinations;
    A is NaN outside the selected triangle (and on the diagonal when unit), the expected X is
    sin(i * nrhs + j + 1), and B = op(A) X summed in increasing k in the working type.
    Yields (upper, transpose, unit, n, nrhs, lda, ldb, ldx, a, b, expected) with a, b,
    expected the padded row-major arrays."""
    T = np.dtype(dtype).type
    for n, nrhs in ((1, 1), (5, 3), (17, 7), (35, 1), (35, 9), (65, 2)):
        lda, ldb, ldx = n + 2, nrhs + 1, nrhs + 3
        for upper in (False, True):
            for transpose in (False, True):
                for unit in (False, True):
                    a = np.full((n, lda), np.nan, dtype=dtype)
                    expected = np.full((n, ldx), T(-123), dtype=dtype)
                    for i in range(n):
                        for j in range(n):
                            if (j >= i) if upper else (j <= i):
                                a[i, j] = (np.nan if unit else T(2.0 + 0.02 * i)) if i == j else T(0.01 * np.sin(float(i * n + j + 1)))
                        for j in range(nrhs):
                            expected[i, j] = T(np.sin(float(i * nrhs + j + 1)))
                    b = np.full((n, ldb), T(-123), dtype=dtype)
                    for i in range(n):
                        for j in range(nrhs):
                            acc = T(0)
                            for k in range(n):
                                row, col = (k, i) if transpose else (i, k)
                                if (col < row) if upper else (col > row):
                                    continue
                                c = T(1) if (unit and row == col) else a[row, col]
                                acc = T(acc + T(c * expected[k, j]))
                            b[i, j] = acc
                    yield upper, transpose, unit, n, nrhs, lda, ldb, ldx, a, b, expected


# ------------------------------------------------------------------- determinant and inverse

def determinant(lu: np.ndarray, piv: np.ndarray | None, info: int = 0, signed: bool = True):
    """det from packed L\\U factors: U's diagonal multiplied in increasing row order with one
    running product in the factors' type (src/matrix_operation/cpu/determinant.cpp:19-28 starts
    from m[0]; src/matrix_operation/gpu/hipblas/inversion_hip.hip:36-58 from 1 * m[0], the same
    value), 0 when the factorisation reported a zero pivot (info != 0). signed=True applies the
    interchange sign as the GPU reference does (one flip per piv[i] != i + 1); signed=False is
    the CPU reference's unsigned product."""
    n = lu.shape[0]
    T = lu.dtype.type
    if info != 0:
        return T(0)
    det = T(1)
    for i in range(n):
        det = T(det * lu[i, i])
    if signed and piv is not None:
        odd = int(np.count_nonzero(np.asarray(piv[:n]) != np.arange(1, n + 1))) & 1
        if odd:
            det = T(-det)
    return det


def inverse_from_factors(lu: np.ndarray, piv: np.ndarray) -> np.ndarray:
    """A^-1 from packed factors the way lu_invert does it
    (src/matrix_operation/cpu/native/inversion_native.cpp:229-304): X = I with rows swapped in
    pivot order, then row i -= l_ik * row k for k < i (forward), row i -= u_ik * row k for k > i
    and row i *= 1 / u_ii (backward, i decreasing), every update a whole row."""
    n = lu.shape[0]
    T = lu.dtype.type
    x = np.eye(n, dtype=lu.dtype)
    for i in range(n):
        p = int(piv[i]) - 1
        if p != i:
            x[[i, p]] = x[[p, i]]
    for i in range(n):
        for k in range(i):
            x[i] = x[i] - lu[i, k] * x[k]
    for i in range(n - 1, -1, -1):
        for k in range(i + 1, n):
            x[i] = x[i] - lu[i, k] * x[k]
        x[i] = x[i] * (T(1) / lu[i, i])
    return x


def inversion_spd_case(dim: int, dtype=np.float64, seed: int = 0) -> np.ndarray:
    """The reference suite's inversion input (Testing/Cxx/TestInversion.cxx:44-64): A = R R^T with
    R unit lower triangular and R[i, j] = 0.4 * u (u uniform in [0, 1]) below the diagonal
    (seeded numpy draws stand in for rand())."""
    rng = np.random.default_rng(seed)
    r = np.eye(dim)
    for i in range(dim):
        for j in range(i):
            r[i, j] = 0.4 * rng.random()
    r = r.astype(dtype)
    return (r @ r.T).astype(dtype)


# ------------------------------------------------------------------- Cholesky

_CHOL_BLOCK = 64  # cholesky_decomposition_native.cpp:165


def cholesky(a: np.ndarray, upper: bool = False):
    """A = L L^T the way cholesky_decomposition_blocked_lower does it
    (src/matrix_operation/cpu/native/cholesky_decomposition_native.cpp:245-320): 64-wide
    blocks; in a block, row i takes l_ij = (a_ij - l_i. . l_j.) / l_jj for j < i and
    l_ii = sqrt(a_ii - l_i. . l_i.), failing at the first row whose sum is <= 0; the rows below
    solve against the block the same way; the trailing lower triangle takes a22 -= l21 l21^T.
    Only the lower triangle of ``a`` is read. Returns (factor, info): L with zeros above the
    diagonal (or U = L^T with zeros below when upper), info the 1-based failing order or 0."""
    n = a.shape[0]
    T = a.dtype.type
    m = np.tril(np.array(a, dtype=a.dtype, copy=True))
    for k in range(0, n, _CHOL_BLOCK):
        bs = min(_CHOL_BLOCK, n - k)
        for i in range(k, k + bs):
            for j in range(k, i):
                m[i, j] = T((m[i, j] - T(np.dot(m[i, k:j], m[j, k:j]))) / m[j, j])
            s = T(m[i, i] - T(np.dot(m[i, k:i], m[i, k:i])))
            if s <= 0:
                return m, k + (i - k) + 1
            m[i, i] = T(np.sqrt(s))
        for r in range(k + bs, n):
            for j in range(k, k + bs):
                m[r, j] = T((m[r, j] - T(np.dot(m[r, k:j], m[j, k:j]))) / m[j, j])
        if k + bs < n:
            l21 = m[k + bs:, k:k + bs]
            m[k + bs:, k + bs:] -= np.tril(l21 @ l21.T).astype(a.dtype)
    return (np.ascontiguousarray(m.T) if upper else m), 0


def cholesky_solve(l: np.ndarray, b: np.ndarray) -> np.ndarray:
    """L L^T x = b with the lower factor (linear_solver_native.cpp:116-147): forward with the
    rows of L, then backward with its columns; b is (n,) or (n, nrhs)."""
    n = l.shape[0]
    x = np.array(b, dtype=l.dtype, copy=True)
    for i in range(n):
        x[i] = (x[i] - l[i, :i] @ x[:i]) / l[i, i]
    for i in range(n - 1, -1, -1):
        x[i] = (x[i] - l[i + 1:, i] @ x[i + 1:]) / l[i, i]
    return x


def cholesky_known_answer(dim: int, dtype=np.float64, seed: int = 0):
    """The reference suite's Cholesky case (Testing/Cxx/TestCholeskyDecomposition.cxx:170-210,
    315-330): C lower triangular with 5 on the diagonal and 0.5 u below (u uniform in [0, 1],
    seeded numpy draws standing in for rand()), A = C C^T in the working type; the factor of A
    must be C to 200 eps (absolute, per entry). Returns (a, c)."""
    rng = np.random.default_rng(seed)
    c = np.zeros((dim, dim), dtype=dtype)
    for i in range(dim):
        c[i, i] = 5.0
        for j in range(i):
            c[i, j] = 0.5 * rng.random()
    return (c @ c.T).astype(dtype), c


# ------------------------------------------------------------------- batched multiplication

def batched_matmul(dim: int, count: int, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """C_i = A_i B_i the way batched_matmul_kernel does it
    (src/matrix_operation/gpu/hipblas/multiplication_batched_hip.hip:12-37): every output an
    accumulation in increasing k, here in float64 whatever the input type (the comparison
    tolerance covers the working type's rounding)."""
    A = np.asarray(a, dtype=np.float64).reshape(count, dim, dim)
    B = np.asarray(b, dtype=np.float64).reshape(count, dim, dim)
    C = np.zeros((count, dim, dim))
    for k in range(dim):
        C += A[:, :, k:k + 1] * B[:, k:k + 1, :]
    return C
