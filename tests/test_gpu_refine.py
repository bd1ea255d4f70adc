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
    # chi_eff as the FP64 refinement applies it (for this uniaxial medium u I + w n n^T from the stored u, w, n in
    # FP64; test_fp64_chi_apply_of_a_uniaxial_medium compares that with the complex64 3x3 chi_eff), read off column by
    # column from constant unit fields
    C = 1 + 2 * sol.vectorial
    dev = sol.field_array.device
    cols = []
    for j in range(C):
        e = torch.zeros((C, 1, *shape), dtype=torch.complex128, device=dev)
        e[j] = 1.0
        cols.append(sol._Solution__fp64_chi_apply(e)[:, 0])
    chi = torch.stack(cols, dim=1).cpu().numpy()
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
    s_c = sol._Solution__fp64_residual_source(E)
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


def test_fp64_chi_apply_of_a_uniaxial_medium():
    """The refinement's FP64 chi_eff X of a uniaxial medium (chi stored as u I + w n n^T) equals the 3x3 chi_eff
    applied entry by entry, also after a rejected update changed alpha and the coefficients."""
    import bench
    print('refine fp64 chi apply uniaxial', flush=True)
    dev = torch.device('cuda:0')
    sol, _ = bench.setup_solution(32, dev)
    assert sol.chi_kind == 4
    rng = np.random.default_rng(3)
    X = torch.as_tensor(rng.standard_normal((3, 1, 32, 32, 32)) + 1j * rng.standard_normal((3, 1, 32, 32, 32)),
                        device=dev)
    for rejected in (False, True):
        if rejected:
            sol._set_alpha(complex(sol.alpha.real, 1.5 * sol.alpha.imag))
        chi = sol.chi_effective.to(torch.complex128)
        ref = torch.stack([sum(chi[i, j] * X[j] for j in range(3)) for i in range(3)])
        got = sol._Solution__fp64_chi_apply(X)
        assert got.shape == ref.shape
        assert float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref)) < 1e-6
