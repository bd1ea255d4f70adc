"""Adjoint gradients on the GPU (macromax_amd.adjoint) against the oracle and finite differences.

loss = sum over a target box of |E|^2 (SI units), E the converged field for permittivity eps. The HIP gradient
(one transposed-operator Born solve) is compared with
  - the oracle's reciprocity adjoint (oracle/born_oracle.py adjoint_gradient, complex128 NumPy),
  - autograd through the unrolled torch restatement of the iteration (oracle/born_torch.py, complex128 CPU),
  - central finite differences of the GPU forward solve at frozen alpha.
Tolerances: complex128 relative 1e-6 of the largest gradient entry; complex64 (solves to residue 1e-6, near the
single-precision floor) finite differences at h = 1e-3 within 2e-2 of the largest entry.
"""
import numpy as np
import pytest
import torch

from oracle import born_oracle as ob
from oracle import born_torch as bt

pytestmark = pytest.mark.gpu


def _problem(kind, dtype):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    rng = np.random.default_rng({'scalar2d': 1, 'iso3d': 2, 'aniso3d': 3}[kind])
    if kind == 'scalar2d':
        shape = (48, 40)
        J = np.zeros((1, *shape), dtype=np.complex128)
        J[0, 6, 10:30] = 1.0
        eps = (1.0 + 0.5 * rng.random(shape)) ** 2
    else:
        shape = (16, 20, 12)
        J = np.zeros((3, *shape), dtype=np.complex128)
        J[0, :, :, 3] = 1.0
        J[1, :, :, 3] = 0.4j
        if kind == 'iso3d':
            eps = (1.0 + 0.4 * rng.random(shape)) ** 2
        else:
            a = rng.standard_normal((3, *shape))
            a /= np.linalg.norm(a, axis=0)
            eps = 1.4 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.6 ** 2 - 1.4 ** 2) * a[:, None] * a[None, :]
    grid = mm.Grid(shape=list(shape), step=0.1e-6)
    bound = LinearBound(grid, thickness=0.4e-6)
    box = tuple(slice(int(0.65 * d), int(0.65 * d) + 3) for d in shape)
    return grid, eps, J.astype(dtype), bound, box


def _loss(E, box):
    return torch.sum(torch.abs(E[(slice(None),) + box]) ** 2)


def _gpu_gradient(kind, dtype, tol):
    from macromax_amd.adjoint import solve_field
    grid, eps, J, bound, box = _problem(kind, dtype)
    eps_t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    info = {}
    E = solve_field(eps_t, grid, 0.5e-6, current_density=J, bound=bound, dtype=dtype, tol=tol, max_iter=20000,
                    info=info)
    loss = _loss(E, box)
    loss.backward()
    return eps_t.grad.detach().numpy(), float(loss.detach()), info


@pytest.mark.parametrize('kind', ['scalar2d', 'iso3d', 'aniso3d'])
def test_adjoint_matches_oracle_and_autograd(kind):
    import macromax_amd as mm
    print(f'adjoint vs oracle {kind}', flush=True)
    grad, loss, info = _gpu_gradient(kind, np.complex128, 1e-12)
    assert info['adjoint_iterations'] > 0
    grid, eps, J, bound, box = _problem(kind, np.complex128)
    sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J, bound=bound,
                      dtype=np.complex128)
    chi = sol.chi_effective.cpu().numpy()
    s = sol.source_normalized.cpu().numpy()
    k0 = sol.wavenumber
    k = ob.k_vectors(tuple(grid.shape), np.asarray(grid.step) * k0)
    vec = sol.vectorial
    o = ob.solve_to(chi, s, sol.alpha, sol.beta, k, vec, 1e-13)
    m = np.zeros(o.E.shape)
    m[(slice(None), slice(None)) + box] = 1.0
    np.testing.assert_allclose(loss, np.sum(m * np.abs(o.E) ** 2) / k0 ** 4, rtol=1e-8)
    g_int = 2.0 * m * o.E / k0 ** 4
    og, _ = ob.adjoint_gradient(chi, sol.alpha, sol.beta, k, vec, o.E, g_int, 1e-13)
    og = og.real
    og = og[0, 0] if og.shape[0] == 1 else og
    scale = np.abs(og).max()
    np.testing.assert_allclose(grad, og, rtol=0, atol=1e-6 * scale)
    if kind != 'iso3d':  # autograd through the unrolled iteration (CPU): the 2D and the 3x3 cases
        eye = np.eye(chi.shape[0]).reshape(chi.shape[0], chi.shape[1], *([1] * len(grid.shape)))
        eps_total = (chi / (1.0j / sol.alpha.imag) + sol.alpha * eye) * sol.beta
        et = torch.tensor(eps_total.real.copy(), requires_grad=True)
        S = torch.as_tensor(s / (1.0j / sol.alpha.imag / sol.beta))
        Et = bt.born_field(et.to(torch.complex128) + 1j * torch.as_tensor(eps_total.imag), S, sol.alpha, sol.beta,
                           k, vec, o.iteration + 50)
        mt = torch.as_tensor(m)
        lt = torch.sum(mt * Et.abs() ** 2) / k0 ** 4
        lt.backward()
        ag = et.grad.numpy()
        ag = ag[0, 0] if ag.shape[0] == 1 else ag
        np.testing.assert_allclose(grad, ag, rtol=0, atol=1e-6 * scale)


@pytest.mark.parametrize('kind,dtype,h,rtol', [('scalar2d', np.complex128, 1e-6, 1e-5),
                                                ('iso3d', np.complex128, 1e-6, 1e-5),
                                                ('iso3d', np.complex64, 1e-3, 2e-2)])
def test_adjoint_matches_finite_differences(kind, dtype, h, rtol):
    import macromax_amd as mm
    from macromax_amd.adjoint import fields_at_fixed_alpha
    print(f'adjoint vs finite differences {kind} {np.dtype(dtype).name}', flush=True)
    tol = 1e-12 if dtype == np.complex128 else 1e-6
    grad, loss, info = _gpu_gradient(kind, dtype, tol)
    assert np.isfinite(loss) and info['forward_residue'] <= tol and info['adjoint_residue'] <= tol, info
    grid, eps, J, bound, box = _problem(kind, dtype)
    alpha = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J, bound=bound,
                        dtype=dtype).alpha
    rng = np.random.default_rng(9)
    vox = [tuple(int(rng.integers(4, d - 4)) for d in eps.shape) for _ in range(6)]
    epsilons = []
    for v in vox:
        for sgn in (+1, -1):
            e = eps.copy()
            e[v] += sgn * h
            epsilons.append(e)
    E = fields_at_fixed_alpha(epsilons, grid, 0.5e-6, alpha, current_density=J, bound=bound, dtype=dtype, tol=tol,
                              max_iter=20000)
    losses = [float(_loss(E[i], box)) for i in range(len(epsilons))]
    scale = np.abs(grad).max()
    for i, v in enumerate(vox):
        fd = (losses[2 * i] - losses[2 * i + 1]) / (2 * h)
        assert abs(grad[v] - fd) <= rtol * scale, (v, grad[v], fd, scale)
