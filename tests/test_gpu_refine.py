"""Mixed-precision refinement (Solution.refine) against the complex128 oracle.

- The FP64 residual source s + chi E - G^-1 E vanishes at the oracle's converged field (relative 1e-9), which pins
  the inverse Green's operator (scalar, dyadic with its longitudinal part and the k = 0 term);
- refinement sweeps bring a complex64 solution below the complex64 iteration's own error: two sweeps reach a
  relative L2 error <= 2e-6 against the converged oracle field, and each sweep reduces it at least 10-fold.
"""
import numpy as np
import pytest
import torch

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu


def _case(kind):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    rng = np.random.default_rng(5)
    if kind == 'scalar2d':
        shape = (64, 48)
        eps = (1.0 + 0.5 * rng.random(shape)) ** 2
        J = np.zeros((1, *shape), dtype=np.complex64)
        J[0, 8, 10:38] = 1.0
    else:
        shape = (16, 20, 24)
        a = rng.standard_normal((3, *shape))
        a /= np.linalg.norm(a, axis=0)
        eps = 1.5 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.7 ** 2 - 1.5 ** 2) * a[:, None] * a[None, :]
        J = np.zeros((3, *shape), dtype=np.complex64)
        J[0, :, :, 4] = 1.0
        J[2, :, :, 4] = 0.5j
    grid = mm.Grid(shape=list(shape), step=0.1e-6)
    sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J,
                      bound=LinearBound(grid, thickness=0.6e-6), dtype=np.complex64)
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s = sol.source_normalized.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors(shape, np.asarray(grid.step) * sol.wavenumber)
    o = ob.solve_to(chi, s, sol.alpha, sol.beta, k, sol.vectorial, 1e-13)
    return sol, o


def _err(sol, o):
    E = sol.field_array_fp64.cpu().numpy()
    return np.linalg.norm((E - o.E).ravel()) / np.linalg.norm(o.E.ravel())


@pytest.mark.parametrize('kind', ['scalar2d', 'aniso3d'])
def test_fp64_residual_source_vanishes_at_solution(kind):
    print(f'refine residual source {kind}', flush=True)
    sol, o = _case(kind)
    E = torch.as_tensor(o.E, device=sol.field_array.device)
    s_c, chi = sol._Solution__fp64_residual_source(E)
    s = sol.source_normalized.to(torch.complex128)
    assert float(torch.linalg.vector_norm(s_c) / torch.linalg.vector_norm(s)) < 1e-9


@pytest.mark.parametrize('kind', ['scalar2d', 'aniso3d'])
def test_refinement_reduces_error_below_complex64_floor(kind):
    print(f'refine sweeps {kind}', flush=True)
    sol, o = _case(kind)
    sol.solve(lambda s_: s_.residue > 1e-4 and s_.iteration < 5000)
    e0 = _err(sol, o)
    info1 = sol.refine(tol=1e-4)
    e1 = _err(sol, o)
    info2 = sol.refine(tol=1e-4)
    e2 = _err(sol, o)
    print(kind, e0, e1, e2, info1, info2, flush=True)
    assert e1 < 0.1 * e0 and e2 < 0.1 * e1
    assert e2 < 2e-6
    assert info2['fp64_residue_after'] < info1['fp64_residue_after'] < info1['fp64_residue_before']
