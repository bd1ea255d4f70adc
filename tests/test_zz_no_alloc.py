"""mmx_born_step_noalloc: the Born iteration without device allocation, against mmx_born_step.

Both entry points run the same kernels; the no-allocation one keeps its per-tile FP64 norm partials in the library's
static device array instead of the plan's scratch buffer. Its fields and norms must equal mmx_born_step's on the same
inputs (and the complex128 oracle's to the tolerance of the other Born-step tests), on consecutive calls whose second
input is smaller than the first, and a launch that needs more partials than the array holds must fail before any
kernel runs.
"""
import numpy as np
import pytest
import torch

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu


def _operands(shape, step, ncomp, chi_kind, nbatch, dtype, seed):
    rng = np.random.default_rng(seed)
    V = int(np.prod(shape))
    cplx = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    E = 0.3 * cplx(nbatch, ncomp, 1, *shape)
    s = 0.2 * cplx(nbatch, ncomp, 1, *shape)
    if chi_kind == 1:
        chi = 0.2 * cplx(1, 1, *shape)
        chi_dev = chi
    else:
        m = 0.2 * cplx(3, 3, *shape)
        chi = 0.5 * (m + m.transpose(1, 0, *range(2, m.ndim)))  # complex-symmetric (chi_kind 3 storage)
        chi_dev = np.stack([chi[0, 0], chi[0, 1], chi[0, 2], chi[1, 1], chi[1, 2], chi[2, 2]])
    alphas = [complex(1.1 + 0.1 * b, 0.35 + 0.05 * b) for b in range(nbatch)]
    return E, s, chi, chi_dev, alphas, V


def _run(plan, E, s, chi_dev, ncomp, chi_kind, alphas, noalloc, dev):
    dt = plan.dtype
    rt = plan.real_dtype
    nbatch = E.shape[0]
    Et = torch.as_tensor(E, dtype=dt, device=dev).contiguous()
    Eo = torch.zeros_like(Et)
    st = torch.as_tensor(s, dtype=dt, device=dev).contiguous()
    ct = torch.as_tensor(chi_dev, dtype=dt, device=dev).contiguous()
    coef = torch.tensor([[a.real, a.imag, 1.0, 0.0, 0.0, 0.0] for a in alphas], dtype=rt, device=dev)
    norms = torch.full((2 * nbatch,), -1.0, dtype=torch.float64, device=dev)
    plan.born_step(Et, Eo, st, ct, ncomp, chi_kind, coef, nbatch, ncomp * plan.V, 0, norms, noalloc=noalloc)
    torch.cuda.synchronize()
    return Eo.cpu(), norms.cpu()


def _check_against_oracle(E, s, chi, alphas, shape, step, ncomp, Eo, norms, tol):
    k = ob.k_vectors(shape, step)
    for b, a in enumerate(alphas):
        o = ob.BornOracle(chi, s[b], a, 1.0, k, ncomp == 3, E0=E[b])
        d_ref = o.update()
        d = Eo[b].numpy() - E[b]
        err = np.linalg.norm((d - d_ref).ravel()) / np.linalg.norm(d_ref.ravel())
        assert err < tol, (b, err)
        np.testing.assert_allclose(norms[2 * b].item(), np.linalg.norm(d_ref.ravel()) ** 2, rtol=tol * 10)


def test_noalloc_equals_born_step_on_consecutive_calls_of_shrinking_size(device):
    from macromax_amd.backend.hip import Plan
    print('no-alloc Born step: 48x40x64 c64 aniso nbatch 2, then nbatch 1, then 32x24 c128 scalar', flush=True)
    shape, step = (48, 40, 64), (0.4, 0.45, 0.5)
    plan = Plan(shape, step, torch.complex64, device)
    assert plan.get_param('scratch_gen') == 0
    E, s, chi, chi_dev, alphas, _ = _operands(shape, step, 3, 3, 2, np.complex64, seed=11)
    Eo_n, n_n = _run(plan, E, s, chi_dev, 3, 3, alphas, True, device)
    assert plan.get_param('scratch_gen') == 0  # the plan allocated nothing
    Eo_r, n_r = _run(plan, E, s, chi_dev, 3, 3, alphas, False, device)
    assert torch.equal(Eo_n, Eo_r)
    assert torch.equal(n_n, n_r)
    _check_against_oracle(E, s, chi, alphas, shape, step, 3, Eo_n, n_n, 1e-5)

    # second call, smaller input (one problem): the array still holds the first call's partials
    E1, s1 = E[1:], s[1:]
    Eo_n1, n_n1 = _run(plan, E1, s1, chi_dev, 3, 3, alphas[1:], True, device)
    Eo_r1, n_r1 = _run(plan, E1, s1, chi_dev, 3, 3, alphas[1:], False, device)
    assert torch.equal(Eo_n1, Eo_r1)
    assert torch.equal(n_n1, n_r1)
    # problem 1 alone gives what it gave in the batch (tile shapes may differ with the batch size)
    assert float(torch.linalg.vector_norm(Eo_n1[0] - Eo_n[1]) / torch.linalg.vector_norm(Eo_n[1])) < 1e-6

    # a smaller grid in complex128
    shape2, step2 = (32, 24), (0.5, 0.4)
    plan2 = Plan(shape2, step2, torch.complex128, device)
    E2, s2, chi2, chi2_dev, alphas2, _ = _operands(shape2, step2, 1, 1, 1, np.complex128, seed=12)
    Eo_n2, n_n2 = _run(plan2, E2, s2, chi2_dev, 1, 1, alphas2, True, device)
    Eo_r2, n_r2 = _run(plan2, E2, s2, chi2_dev, 1, 1, alphas2, False, device)
    assert torch.equal(Eo_n2, Eo_r2)
    assert torch.equal(n_n2, n_r2)
    _check_against_oracle(E2, s2, chi2, alphas2, shape2, step2, 1, Eo_n2, n_n2, 1e-11)


def test_noalloc_rejects_launches_beyond_the_static_array(device):
    from macromax_amd import _native
    from macromax_amd.backend.hip import Plan
    print('no-alloc Born step: 1-D 64 points, 300000 problems (rejected)', flush=True)
    plan = Plan((64,), (0.5,), torch.complex64, device)
    nbatch = 300_000  # one tile per problem: 16 bytes of partials each, 4.8 MB > 4 MiB
    assert 16 * nbatch > _native.NOALLOC_SCRATCH_BYTES
    E = torch.zeros((nbatch, 1, 1, 64), dtype=torch.complex64, device=device)
    Eo = torch.zeros_like(E)
    s = torch.zeros((1, 1, 64), dtype=torch.complex64, device=device)
    chi = torch.zeros((1, 1, 64), dtype=torch.complex64, device=device)
    coef = torch.tensor([1.0, 0.5, 1.0, 0.0, 0.0, 0.0] * nbatch, dtype=torch.float32, device=device)
    norms = torch.full((2 * nbatch,), -1.0, dtype=torch.float64, device=device)
    with pytest.raises(_native.NativeError, match='static scratch array'):
        plan.born_step(E, Eo, s, chi, 1, 1, coef, nbatch, 0, 0, norms, noalloc=True)
    # a batch that fits runs
    plan.born_step(E[:1000], Eo[:1000], s, chi, 1, 1, coef[:6000], 1000, 0, 0, norms[:2000], noalloc=True)
    torch.cuda.synchronize()
    assert torch.all(norms[:2000] == 0)
