"""batched_matrix_multiplication: C_i = A_i B_i for ``count`` square row-major matrices stored
back to back (include/matrix_operation/multiplication_batched.h:7-42: matrix i at offset
i * dim * dim, device pointers, asynchronous on the stream, float32 and float64). Runs the
HIP kernel of the library; there is no CPU fallback.
"""
from __future__ import annotations

import torch

from . import _lib
from .lu import _check, _ptr, _sfx, _stream_ptr


def batched_matrix_multiplication(dim: int, count: int, a: torch.Tensor, b: torch.Tensor, c: torch.Tensor,
                                  stream=None) -> None:
    if a is None or b is None or c is None:
        raise ValueError("batched_matrix_multiplication: a, b, and c must not be null")
    if dim <= 0 or count <= 0:
        raise ValueError("batched_matrix_multiplication: dim and count must be positive")
    for t, what in ((a, "a"), (b, "b"), (c, "c")):
        if not t.is_cuda or not t.is_contiguous() or t.dtype != a.dtype or t.numel() < count * dim * dim:
            raise ValueError(f"batched_matrix_multiplication: {what} must be a contiguous device tensor of "
                             f"count * dim * dim elements, same dtype as a")
    code = getattr(_lib.load(), f"lrn_batched_matrix_multiplication_{_sfx(a.dtype)}")(
        int(dim), int(count), _ptr(a), _ptr(b), _ptr(c), _stream_ptr(stream))
    _check(code, "batched_matrix_multiplication")
