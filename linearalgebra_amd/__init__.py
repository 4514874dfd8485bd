"""MI355X-native (gfx950) LU solve path: lu_decomposition / linear_solver for the
ROCM_NATIVE backend, backed by hand-written HIP kernels behind a C ABI.

The compute lives in ``liblinalg_rocm_native.so`` (built from ``csrc/``); this package only
loads it and passes device pointers and streams. There is no CPU fallback.
"""
from .lu import (  # noqa: F401
    LinalgError,
    cholesky_decomposition,
    cholesky_decomposition_enum,
    get_tuning,
    linear_solver,
    linear_solver_type,
    lu_decomposition,
    set_tuning,
)

from .mixed import gesv_mixed  # noqa: F401,E402
from .triangular import triangular_solver  # noqa: F401,E402
from .inverse import determinant, inversion  # noqa: F401,E402
from .multiplication import batched_matrix_multiplication  # noqa: F401,E402

__all__ = ["batched_matrix_multiplication", "cholesky_decomposition", "cholesky_decomposition_enum", "determinant", "inversion", "gesv_mixed", "lu_decomposition", "linear_solver", "linear_solver_type",
           "triangular_solver", "LinalgError", "set_tuning", "get_tuning"]
