"""The oracle's batched multiplication against numpy's matmul and a hand-checked case."""
import numpy as np

from oracle import oracle as O


def test_batched_matmul_known_and_numpy():
    a = np.array([[1.0, 2.0], [3.0, 4.0], [0.0, 1.0], [1.0, 0.0]])
    b = np.array([[5.0, 6.0], [7.0, 8.0], [2.0, 3.0], [4.0, 5.0]])
    c = O.batched_matmul(2, 2, a, b)
    assert np.array_equal(c[0], [[19.0, 22.0], [43.0, 50.0]])
    assert np.array_equal(c[1], [[4.0, 5.0], [2.0, 3.0]])  # a row swap of B
    rng = np.random.default_rng(1)
    x = rng.uniform(-5, 5, (6, 9, 9))
    y = rng.uniform(-5, 5, (6, 9, 9))
    assert np.abs(O.batched_matmul(9, 6, x, y) - x @ y).max() <= 1e-12
    # a transposed operand (a plausible slip) is caught by the same comparison
    assert np.abs(O.batched_matmul(9, 6, x, y) - x @ y.transpose(0, 2, 1)).max() > 1.0
