"""Pins of the adjoint-gradient restatements of the oracle (no GPU).

- ``oracle/born_torch.py`` reproduces the NumPy oracle's iterates (same operators, fixed alpha);
- ``born_oracle.adjoint_gradient`` (reciprocity: one forward solve with source beta conj(g)) agrees with central
  finite differences of the converged field at frozen alpha, for isotropic and 3x3 media;
- autograd through the unrolled torch iteration agrees with it;
- a planted mistake (the adjoint source without the conjugate) fails the finite-difference check.
"""
import numpy as np
import pytest
import torch

from oracle import born_oracle as ob
from oracle import born_torch as bt

TOL = 1e-13


def _problem(kind, seed=0):
    rng = np.random.default_rng(seed)
    if kind == 'scalar2d':
        shape, C = (20, 24), 1
        n = 1.0 + 0.6 * rng.random(shape)
        eps = (n ** 2)[None, None]
        J = np.zeros((1, *shape), dtype=complex)
        J[0, 4, 6:18] = 1.0
    else:
        shape, C = (8, 10, 12), 3
        n = 1.0 + 0.5 * rng.random(shape)
        if kind == 'iso3d':
            eps = (n ** 2)[None, None]
        else:
            a = rng.standard_normal((3, *shape))
            a /= np.linalg.norm(a, axis=0)
            eps = 1.4 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.6 ** 2 - 1.4 ** 2) * a[:, None] * a[None, :]
        J = np.zeros((3, *shape), dtype=complex)
        J[0, :, :, 2] = 1.0
        J[1, :, :, 2] = 0.3j
    pitch, wl = 0.1e-6, 0.5e-6
    k0 = 2 * np.pi / wl
    chi, s, alpha, beta, k = ob.setup_problem(shape, pitch, k0, eps, current_density=J, bound_thickness=0.4e-6,
                                              vectorial=(C == 3))
    # eps_total = eps + bound, recovered from chi (solver.py:480-484 inverted)
    eye = np.eye(chi.shape[0]).reshape(chi.shape[0], chi.shape[1], *([1] * len(shape)))
    eps_total = (chi / (1.0j / alpha.imag) + alpha * eye) * beta
    box = tuple(slice(int(0.6 * d), int(0.6 * d) + 2) for d in shape)
    return chi, s, alpha, beta, k, C == 3, eps_total, box


def _loss_grad(E, box):
    """f = sum over the box of |E|^2 and its gradient g = df/dRe(E) + i df/dIm(E) = 2 E on the box."""
    m = np.zeros(E.shape)
    m[(slice(None), slice(None)) + box] = 1.0
    return float(np.sum(m * np.abs(E) ** 2)), 2.0 * m * E


@pytest.mark.parametrize('kind', ['scalar2d', 'aniso3d'])
def test_torch_restatement_matches_numpy_iterates(kind):
    chi, s, alpha, beta, k, vec, eps_total, box = _problem(kind)
    o = ob.BornOracle(chi, s, alpha, beta, k, vec)
    for _ in range(6):
        o.step()
    S = s / (1.0j / alpha.imag / beta)
    E = bt.born_field(torch.as_tensor(eps_total), torch.as_tensor(S), alpha, beta, k, vec, 6)
    np.testing.assert_allclose(E.numpy(), o.E, rtol=1e-12, atol=1e-12 * np.abs(o.E).max())


def _fd(eps_total, s, alpha, beta, k, vec, box, idx, h):
    vals = []
    for sign in (+1, -1):
        e = eps_total.copy()
        e[idx] += sign * h
        o = ob.solve_to(ob.chi_base(e, alpha, beta), s, alpha, beta, k, vec, TOL)
        vals.append(_loss_grad(o.E, box)[0])
    return (vals[0] - vals[1]) / (2 * h)


@pytest.mark.parametrize('kind', ['scalar2d', 'iso3d', 'aniso3d'])
def test_adjoint_gradient_matches_finite_differences(kind):
    chi, s, alpha, beta, k, vec, eps_total, box = _problem(kind, seed=1)
    o = ob.solve_to(chi, s, alpha, beta, k, vec, TOL)
    f, g = _loss_grad(o.E, box)
    grad, adj = ob.adjoint_gradient(chi, alpha, beta, k, vec, o.E, g, TOL)
    rng = np.random.default_rng(2)
    shape = eps_total.shape[2:]
    h = 1e-5
    checked = 0
    for _ in range(4):
        vox = tuple(int(rng.integers(0, d)) for d in shape)
        comps = [(0, 0)] if eps_total.shape[0] == 1 else [(0, 0), (0, 2), (2, 1)]
        for (i, j) in comps:
            idx = (i, j) + vox
            fd = _fd(eps_total, s, alpha, beta, k, vec, box, idx, h)
            assert abs(grad[idx].real - fd) <= 1e-6 * np.abs(grad).max() + 1e-5 * abs(fd), (idx, grad[idx], fd)
            checked += 1
    assert checked >= 4


def test_adjoint_planted_mistake_fails():
    """Without the conjugate on the adjoint source, the gradient disagrees with finite differences."""
    chi, s, alpha, beta, k, vec, eps_total, box = _problem('scalar2d', seed=3)
    o = ob.solve_to(chi, s, alpha, beta, k, vec, TOL)
    f, g = _loss_grad(o.E, box)
    good, _ = ob.adjoint_gradient(chi, alpha, beta, k, vec, o.E, g, TOL)
    bad, _ = ob.adjoint_gradient(chi, alpha, beta, k, vec, o.E, np.conj(g), TOL)  # conj(conj(g)) = g: wrong
    idx = (0, 0, 12, 12)
    fd = _fd(eps_total, s, alpha, beta, k, vec, box, idx, 1e-5)
    assert abs(good[idx].real - fd) < 1e-4 * abs(fd)
    assert abs(bad[idx].real - fd) > 1e-2 * abs(fd)


@pytest.mark.parametrize('kind', ['scalar2d', 'aniso3d'])
def test_autograd_through_torch_iteration_matches_adjoint(kind):
    chi, s, alpha, beta, k, vec, eps_total, box = _problem(kind, seed=4)
    o = ob.solve_to(chi, s, alpha, beta, k, vec, TOL)
    f, g = _loss_grad(o.E, box)
    grad, _ = ob.adjoint_gradient(chi, alpha, beta, k, vec, o.E, g, TOL)
    eps_t = torch.as_tensor(eps_total.real.copy(), dtype=torch.float64).requires_grad_(True)
    S = torch.as_tensor(s / (1.0j / alpha.imag / beta))
    eps_c = eps_t.to(torch.complex128) + 1j * torch.as_tensor(eps_total.imag)
    E = bt.born_field(eps_c, S, alpha, beta, k, vec, o.iteration + 50)
    m = torch.zeros(E.shape, dtype=torch.float64)
    m[(slice(None), slice(None)) + box] = 1.0
    loss = torch.sum(m * E.abs() ** 2)
    loss.backward()
    ga = eps_t.grad.numpy()
    np.testing.assert_allclose(float(loss), f, rtol=1e-9)
    np.testing.assert_allclose(ga, grad.real, rtol=0, atol=1e-7 * np.abs(grad).max())
