"""Known answers for polynomial root finding, from the reference's tests (python/tests/test_backend_numpy.py:12-166):
coefficients low to high order, expected roots (nan for the roots a degenerate polynomial lacks) and the decimal
precision each case is checked to there."""
import numpy as np

nan = np.nan
CASES = [
    ([0.0, 0.0, 0.0, -1.0], [0.0, 0.0, 0.0], 6),
    ([1, 0], [nan], 6),
    ([0, 1], [0.0], 6),
    ([-1, 1], [1.0], 6),
    ([-1j, 1], [1j], 6),
    ([-1 - 1j, 1], [1 + 1j], 6),
    ([2, 1, 0], [-2, nan], 6),
    ([1, 0, 0], [nan, nan], 6),
    ([0, 0, 0], [nan, nan], 6),
    ([2, -3, 1], [1, 2], 6),
    ([0, -1, 1], [0, 1], 6),
    ([-1, 0, 1], [-1, 1], 6),
    ([2, 3, 1], [-2, -1], 6),
    ([-1, 0, 1, 0], [-1, 1, nan], 6),
    ([-2.7 ** 3, 2.7 ** 2 * 3, -2.7 * 3, 1], [2.7, 2.7, 2.7], 1),
    ([1, 0, 1, 0], [-1j, 1j, nan], 6),
    ([-6, 11, -6, 1], [1, 2, 3], 6),
    ([1, 0, 1], [-1j, 1j], 6),
    ([-1j, 1 - 1j, 1], [-1, 1j], 6),
    ([0, -1, 0, 1], [-1, 0, 1], 6),
    ([6, 11, 6, 1], [-3, -2, -1], 6),
    ([0, 1e-2, 0, -1], [-0.1, 0.0, 0.1], 6),
    ([0.0, 0.0, 0.100001, -1.0], [0.0, 0.0, 0.100001], 4),
    ([0.0, -0.0064, 0.2, -1], [0.0, 0.04, 0.16], 6),
]
# several polynomials at once (one per column)
MULTI = [
    ([[1, 1], [0, 0]], [[nan, nan]]),
    ([[1], [0], [1], [0]], [[-1j], [1j], [nan]]),
    ([[-1, 1], [0, 0], [1, 1], [0, 0]], [[-1, -1j], [1, 1j], [nan, nan]]),
]
