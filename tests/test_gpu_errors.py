"""Error behaviour of the C ABI and the Python layer (the reference's conventions: ValueError for unsupported
shapes and sizes, TypeError for wrong types; the C ABI returns MMX_EINVAL / MMX_EUNSUP with a message and never
touches memory on a rejected call)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(shape, dtype='c64'):
    import torch
    from macromax_amd.backend.hip import Plan
    return Plan(shape, [0.3] * len(shape), torch.complex64 if dtype == 'c64' else torch.complex128,
                torch.device('cuda:0'))


@pytest.mark.parametrize('shape', [(8191,), (64, 2 * 4099), (4096 * 4096 * 2,), (8, 0, 4)])
def test_unsupported_or_invalid_shapes(shape, device):
    from macromax_amd._native import NativeError
    print(f'invalid shape {shape}', flush=True)
    with pytest.raises((ValueError, NativeError)):
        _plan(shape)


def test_born_call_argument_checks(device):
    import torch
    from macromax_amd import _native
    print('born argument checks', flush=True)
    plan = _plan((16, 20, 24))
    V = 16 * 20 * 24
    E = torch.zeros(3, V, dtype=torch.complex64, device=device)
    Eo, s = torch.zeros_like(E), torch.zeros_like(E)
    chi = torch.zeros(6, V, dtype=torch.complex64, device=device)
    coef = torch.tensor([1.0, 0.5, 1.0, 0.0, 0.0, 0.0], dtype=torch.float32, device=device)
    norms = torch.zeros(2, dtype=torch.float64, device=device)
    with pytest.raises(ValueError):  # chi size does not match chi_kind 2
        plan.born_step(E, Eo, s, chi, 3, 2, coef, 1, 0, 0, norms)
    with pytest.raises(TypeError):  # complex128 field on a complex64 plan
        plan.born_step(E.to(torch.complex128), Eo, s, chi, 3, 3, coef, 1, 0, 0, norms)
    with pytest.raises(ValueError):  # norms must be float64
        plan.born_step(E, Eo, s, chi, 3, 3, coef, 1, 0, 0, norms.float())
    # the C ABI itself: invalid chi kind, component count and aliasing are rejected with a message
    L = _native.lib()
    st = torch.cuda.current_stream(device).cuda_stream
    before = Eo.clone()
    for ncomp, kind, out in ((3, 5, Eo), (3, 0, Eo), (2, 1, Eo), (3, 3, E)):
        rc = L.mmx_born_step(plan.handle, E.data_ptr(), out.data_ptr(), s.data_ptr(), chi.data_ptr(), ncomp, kind,
                             coef.data_ptr(), 1, 0, 0, norms.data_ptr(), st)
        assert rc == -1, (ncomp, kind)
        assert L.mmx_last_error().decode()
    torch.cuda.synchronize()
    assert torch.equal(Eo, before)
    # nothing to do for an empty batch
    assert L.mmx_born_step(plan.handle, E.data_ptr(), Eo.data_ptr(), s.data_ptr(), chi.data_ptr(), 3, 3,
                           coef.data_ptr(), 0, 0, 0, norms.data_ptr(), st) == 0


def test_plan_parameter_and_source_box_checks(device):
    from macromax_amd._native import NativeError
    print('parameter checks', flush=True)
    plan = _plan((32, 48))
    with pytest.raises(NativeError):
        plan.set_param('xcd_remap', 300)
    with pytest.raises(NativeError):
        plan.set_source_support([0, 0], [33, 48])  # beyond the grid
    with pytest.raises(NativeError):
        plan.set_source_support([5, 0], [4, 48])  # lo > hi
    plan.set_source_support([0, 0], [0, 0])  # an empty box (a zero source) is valid
    plan.set_source_support(None, None)


def test_batch_and_adjoint_edge_cases(device):
    import torch
    import macromax_amd as mm
    from macromax_amd.adjoint import solve_field
    from macromax_amd.batch import BatchSolution
    print('batch and adjoint edge cases', flush=True)
    a = mm.Solution(mm.Grid(shape=[32, 32], step=0.1e-6), vacuum_wavelength=0.5e-6, epsilon=2.0,
                    current_density=np.ones((1, 32, 32)), dtype=np.complex64)
    b = mm.Solution(mm.Grid(shape=[32, 48], step=0.1e-6), vacuum_wavelength=0.5e-6, epsilon=2.0,
                    current_density=np.ones((1, 32, 48)), dtype=np.complex64)
    with pytest.raises(ValueError):
        BatchSolution([a, b])
    with pytest.raises(ValueError):
        BatchSolution([])
    # a loss that does not depend on the field has a zero gradient (no adjoint solve)
    eps = torch.full((32, 32), 2.0, dtype=torch.float64, requires_grad=True)
    J = np.zeros((1, 32, 32))
    J[0, 4, :] = 1.0
    E = solve_field(eps, mm.Grid(shape=[32, 32], step=0.1e-6), 0.5e-6, current_density=J, tol=1e-5)
    (0.0 * E.abs().sum()).backward()
    assert torch.count_nonzero(eps.grad) == 0
