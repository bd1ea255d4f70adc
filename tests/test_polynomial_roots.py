"""Polynomial root finding of the backend (BackEnd.calc_roots_of_low_order_polynomial, ``backend/__init__.py:725-850``)
on the CPU: the torch restatement ``macromax_amd.backend.hip.polynomial_roots`` (which BackEndHip runs on the device)
against the reference's known answers, against numpy.roots on random cubics, quadratics and lines, and with
evaluate_polynomial (Horner) checked against direct evaluation."""
import numpy as np
import numpy.testing as npt
import pytest
import torch

from polynomial_cases import CASES, MULTI

EPS32 = float(np.finfo(np.float32).eps)


def _roots(C):
    from macromax_amd.backend.hip import polynomial_roots
    return polynomial_roots(torch.as_tensor(np.asarray(C)).to(torch.complex128), 2.0 * EPS32).numpy()


@pytest.mark.parametrize('C,expected,decimal', CASES, ids=[str(c[0]) for c in CASES])
def test_known_answers(C, expected, decimal):
    npt.assert_array_almost_equal(_roots(C), np.asarray(expected, dtype=np.complex128), decimal=decimal)


@pytest.mark.parametrize('C,expected', MULTI, ids=['lines', 'column', 'two-columns'])
def test_known_answers_several_polynomials(C, expected):
    npt.assert_array_almost_equal(_roots(C), np.asarray(expected, dtype=np.complex128))


def test_constant_polynomials_have_no_roots():
    assert _roots([1]).shape == (0,)
    assert _roots([0]).shape == (0,)
    assert _roots(np.ones((1, 4, 5))).shape == (0, 4, 5)


def test_random_polynomials_against_numpy_roots():
    """Random complex polynomials of degree 1-3 (leading coefficient of magnitude >= 0.1): the sorted roots agree
    with numpy.roots (companion-matrix eigenvalues) to 1e-9 relative, and each root is a zero of its polynomial."""
    rng = np.random.default_rng(3)
    for degree in (1, 2, 3):
        C = rng.standard_normal((degree + 1, 200)) + 1j * rng.standard_normal((degree + 1, 200))
        C[-1] += np.sign(C[-1].real) * 0.1
        got = _roots(C)
        for i in range(C.shape[1]):
            ref = np.roots(C[::-1, i])
            ref = ref[np.lexsort((ref.imag, ref.real))]
            npt.assert_allclose(got[:, i], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
        from macromax_amd.backend.hip import BackEndHip
        vals = np.stack([BackEndHip.evaluate_polynomial(C, got[r]) for r in range(degree)])
        scale = np.abs(C).sum(axis=0) * np.maximum(1.0, np.abs(got).max(axis=0)) ** degree
        assert np.all(np.abs(vals) <= 1e-12 * scale)


def test_evaluate_polynomial():
    """``python/tests/test_backend_numpy.py:163-166``, and Horner's scheme against the power sum."""
    from macromax_amd.backend.hip import BackEndHip
    C = np.array([1, 2, 3, -4])[:, np.newaxis, np.newaxis, np.newaxis, np.newaxis]
    result = BackEndHip.evaluate_polynomial(C, np.array([0, 1])[:, np.newaxis, np.newaxis, np.newaxis, np.newaxis])
    npt.assert_almost_equal(result, np.array([1, 2])[:, np.newaxis, np.newaxis, np.newaxis, np.newaxis])
    x = np.linspace(-2, 2, 9) + 0.5j
    c = np.array([0.5, -1.0, 2.0, 0.25])
    npt.assert_allclose(BackEndHip.evaluate_polynomial(c, x), sum(c[k] * x ** k for k in range(4)), rtol=1e-13)
