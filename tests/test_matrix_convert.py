"""Host-side matrix algebra of macromax_amd.matrix (no device needed): scattering <-> transfer conversion, the
regularised inverse and the quarter matrices of a literal scattering matrix.

The conversion cases are the reference's (python/tests/test_matrix.py:383-441): the transfer matrix is built from
the quarters by the closed form [[A - B D^-1 C, B D^-1], [-D^-1 C, D^-1]], conversion is its own inverse, and the
Tikhonov-regularised conversion at noise level 1e-3 agrees to 5 decimals.
"""
import numpy as np
import numpy.testing as npt
import pytest

from macromax_amd import matrix


def _assemble(a, b, c, d):
    return np.array([[a, b], [c, d]]).transpose((0, 2, 1, 3)).reshape(2 * np.asarray(a.shape))


CASES = {
    'scalar': [np.atleast_2d(0.9), np.atleast_2d(0.1), np.atleast_2d(0.2), np.atleast_2d(0.8)],
    'matrix': [np.array([[0.9, 0.3], [-0.2, 0.7]]), np.array([[0.1, 0.2], [0.0, 0.1]]),
               np.array([[0.0, -0.2], [0.1, 0.1]]), np.array([[0.8, 0.2], [-0.2, 0.8]])],
    'complex': [np.array([[0.9, 0.3], [-0.2, 0.7 + 0.2j]]), np.array([[0.1j, 0.2], [0.0, 0.1]]),
                np.array([[0.0, -0.2], [0.2 + 0.1j, 0.1]]), np.array([[0.1 + 0.8j, 0.2], [-0.2, 0.8]])],
}


@pytest.mark.parametrize('case', list(CASES))
def test_convert(case):
    a, c, b, d = CASES[case]  # forward-to-forward, forward-to-backward, backward-to-forward, backward-to-backward
    s = _assemble(a, b, c, d)
    d_inv = np.linalg.inv(d)
    t_ref = _assemble(a - b @ d_inv @ c, b @ d_inv, -d_inv @ c, d_inv)
    npt.assert_array_almost_equal(matrix.convert(s), t_ref)
    npt.assert_array_almost_equal(matrix.convert(s, noise_level=1e-3), t_ref, decimal=5)
    npt.assert_array_almost_equal(matrix.convert(matrix.convert(s)), s)
    npt.assert_array_almost_equal(matrix.convert(matrix.convert(s, noise_level=1e-3), noise_level=1e-3), s)
    # the 4D [2, N/2, 2, N/2] form gives the same numbers
    n = s.shape[0] // 2
    npt.assert_array_almost_equal(matrix.convert(s.reshape(2, n, 2, n)).reshape(s.shape), t_ref)


def test_inv_regularised():
    rng = np.random.default_rng(0)
    m = rng.standard_normal((6, 6)) + 1j * rng.standard_normal((6, 6))
    npt.assert_array_almost_equal(matrix.inv(m) @ m, np.eye(6))
    # Tikhonov: sigma / (sigma^2 + noise^2) on the singular values
    u, sv, vh = np.linalg.svd(m)
    noise = 0.3
    expected = (vh.conj().T * (sv / (sv ** 2 + noise ** 2))) @ u.conj().T
    npt.assert_array_almost_equal(matrix.inv(m, noise), expected)


def test_literal_matrix_quarters_and_transfer():
    rng = np.random.default_rng(1)
    arr = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    m = matrix.LiteralScatteringMatrix(arr)
    assert m.side == 8 and m.size == 64 and len(m) == 8
    npt.assert_array_equal(np.asarray(m), arr)
    q = [m.forward_transmission, m.back_reflection, m.front_reflection, m.backward_transmission]
    npt.assert_array_equal(np.vstack([np.hstack(q[:2]), np.hstack(q[2:])]), arr)
    v = rng.standard_normal(4) + 1j * rng.standard_normal(4)
    npt.assert_array_almost_equal(m.forward_transmission @ v, arr[:4, :4] @ v)
    npt.assert_array_almost_equal(m.backward_transmission @ v, arr[4:, 4:] @ v)
    npt.assert_array_almost_equal(m.front_reflection @ v, arr[4:, :4] @ v)
    npt.assert_array_almost_equal(m.back_reflection @ v, arr[:4, 4:] @ v)
    npt.assert_array_almost_equal(m.transfer(), matrix.convert(arr))
    npt.assert_array_almost_equal(m @ np.eye(8), arr)
    npt.assert_array_almost_equal(m[2], arr[2])
    npt.assert_array_almost_equal(m[1:3, 4], arr[1:3, 4])
    npt.assert_array_almost_equal(m.inv(), np.linalg.inv(arr))
    with pytest.raises(ValueError):
        matrix.SquareMatrix(np.zeros((3, 4)))
    with pytest.raises(TypeError):
        matrix.QuarterMatrix(matrix.SquareMatrix(np.eye(3)), False, False)
