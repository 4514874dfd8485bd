"""batched_matrix_multiplication on the GPU against the oracle and the reference's GPU test
(Testing/Cxx/TestGpuMatrixOperations.cxx:400-445: dim 9, count 6, U(-5, 5) entries)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bmm(dim, count, a, b, dev):
    import torch
    import linearalgebra_amd as la
    ta = torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    tb = torch.from_numpy(np.ascontiguousarray(b)).to(dev)
    tc = torch.full_like(ta, -7.0)
    la.batched_matrix_multiplication(dim, count, ta, tb, tc)
    torch.cuda.synchronize()
    return tc.cpu().numpy().reshape(count, dim, dim)


@pytest.mark.parametrize("dim,count", [(9, 6), (1, 3), (16, 5), (17, 4), (63, 2), (64, 4), (65, 3), (130, 2),
                                       (257, 1), (5, 100000)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-13), (np.float32, 5e-6)])
def test_batched_matmul_matches_oracle(dim, count, dtype, tol, dev):
    rng = np.random.default_rng(dim * 10 + count)
    a = rng.uniform(-5, 5, count * dim * dim).astype(dtype)
    b = rng.uniform(-5, 5, count * dim * dim).astype(dtype)
    c = _bmm(dim, count, a, b, dev)
    ref = O.batched_matmul(dim, count, a, b)
    scale = 25.0 * dim
    err = np.abs(c - ref).max() / scale
    print(f"bmm dim={dim} count={count} {np.dtype(dtype).name}: max err / (25 dim) {err:.2e}", flush=True)
    assert err <= tol


def test_batched_matmul_invalid_arguments(dev):
    import torch
    import linearalgebra_amd as la
    x = torch.zeros(2 * 2 * 2, dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        la.batched_matrix_multiplication(2, 2, None, x, x)
    with pytest.raises(ValueError):
        la.batched_matrix_multiplication(0, 2, x, x, x)
    with pytest.raises(ValueError):
        la.batched_matrix_multiplication(2, 0, x, x, x)
    with pytest.raises(ValueError):
        la.batched_matrix_multiplication(4, 2, x, x, x)  # too few elements
