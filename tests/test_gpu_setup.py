"""Set-up reductions on the GPU (mmx_setup_sigma, mmx_setup_gain) against numpy.linalg, and the preconditioner
constant they produce against the oracle's (solver.py:313-467; oracle find_alpha_beta).

sigma: max over voxels of the largest eigenvalue of (eps - s I)^H (eps - s I), relative 1e-12;
gain: min / max over voxels of the eigenvalues of -i/2 (eps - eps^H), absolute 1e-12 of the matrix norm;
alpha of a Solution (these kernels inside Nelder-Mead) equals the oracle's (NumPy eigenvalues) to 1e-9.
"""
import numpy as np
import pytest

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu


def _fields(kind, shape, seed):
    rng = np.random.default_rng(seed)
    if kind == 'iso':
        return (1.5 + rng.random(shape) + 0.2j * rng.random(shape) - 0.05j)[None, None]
    m = rng.standard_normal((3, 3, *shape)) + 1j * rng.standard_normal((3, 3, *shape))
    if kind == 'herm':
        m = 0.5 * (m + np.conj(np.swapaxes(m, 0, 1)))
    elif kind == 'diag':  # multiples of the identity and diagonal matrices (degenerate eigenvalues)
        d = rng.standard_normal(shape) + 0.1j
        m = np.eye(3).reshape(3, 3, *([1] * len(shape))) * d
        m[0, 0, ::2] += 0.5
    return m


@pytest.mark.parametrize('kind', ['iso', 'full', 'herm', 'diag'])
def test_setup_reductions_match_numpy(kind, device):
    import torch
    from macromax_amd.backend.hip import Plan
    print(f'setup reductions {kind}', flush=True)
    shape = (12, 10, 14)
    plan = Plan(shape, [0.3] * 3, torch.complex64, device)
    eps = _fields(kind, shape, 3)
    t = torch.as_tensor(np.ascontiguousarray(eps), dtype=torch.complex128, device=device)
    C = eps.shape[0]
    mats = np.moveaxis(eps.reshape(C, C, -1), -1, 0)  # [V, C, C]
    for shift in (0.0, 1.3, -0.7):
        d = mats - shift * np.eye(C)
        ref = np.max(np.linalg.eigvalsh(np.conj(np.swapaxes(d, 1, 2)) @ d))
        got = plan.setup_sigma(t, shift)
        assert abs(got - ref) <= 1e-12 * ref, (shift, got, ref)
    h = -0.5j * (mats - np.conj(np.swapaxes(mats, 1, 2)))
    ev = np.linalg.eigvalsh(h)
    lo, hi = plan.setup_gain(t)
    scale = np.abs(mats).max()
    assert abs(lo - ev.min()) <= 1e-12 * scale and abs(hi - ev.max()) <= 1e-12 * scale


@pytest.mark.parametrize('kind', ['iso', 'aniso'])
def test_solution_alpha_matches_oracle(kind):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'alpha vs oracle {kind}', flush=True)
    shape = (16, 20, 12)
    rng = np.random.default_rng(8)
    if kind == 'iso':
        eps = (1.0 + 0.6 * rng.random(shape)) ** 2
    else:
        a = rng.standard_normal((3, *shape))
        a /= np.linalg.norm(a, axis=0)
        eps = 1.45 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.7 ** 2 - 1.45 ** 2) * a[:, None] * a[None, :]
    grid = mm.Grid(shape=list(shape), step=0.2e-6)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[0, :, :, 2] = 1.0
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J,
                      bound=LinearBound(grid, thickness=0.8e-6), dtype=np.complex128)
    chi, s, alpha, beta, k = ob.setup_problem(shape, 0.2e-6, 2 * np.pi / 1e-6,
                                              eps if kind == 'aniso' else eps[None, None], current_density=J,
                                              bound_thickness=0.8e-6)
    np.testing.assert_allclose(sol.alpha.real, alpha.real, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(sol.alpha.imag, alpha.imag, rtol=1e-9)


@pytest.mark.parametrize('case', ['eps_gain', 'eps_gain_mu_loss', 'mu_gain', 'eps_gain_c64'])
def test_gain_only_media_alpha_and_iterates_match_oracle(case):
    """Gain-only media (no absorbing bound, whose extinction is loss): alpha is conjugated before the bias
    (solver.py:329-336, 464-465) when eps or the constant mu has gain and neither has loss; 10 iterates (with the
    divergence guard) follow the oracle."""
    import torch
    import macromax_amd as mm
    print(f'gain-only medium {case}', flush=True)
    shape = (40, 48)
    rng = np.random.default_rng(21)
    n = 1.0 + 0.4 * rng.random(shape)
    gain = 0.15 * rng.random(shape)
    eps = n ** 2 - (0.0 if case == 'mu_gain' else 1j) * gain
    mu = {'eps_gain': 1.0, 'eps_gain_c64': 1.0, 'eps_gain_mu_loss': 1.0 + 0.01j, 'mu_gain': 1.0 - 0.01j}[case]
    dtype = np.complex64 if case.endswith('c64') else np.complex128
    thr = float(np.sqrt(np.finfo(np.float32 if dtype == np.complex64 else np.float64).eps))
    grid = mm.Grid(shape=list(shape), step=0.1e-6)
    J = np.zeros((1, *shape), dtype=np.complex128)
    J[0, 5, 10:38] = 1.0
    sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, mu=mu, current_density=J.astype(dtype),
                      dtype=dtype)
    # the Solution holds eps in its element type (BackEnd.astype) before the FP64 set-up
    eps_held = eps.astype(dtype).astype(np.complex128)
    chi, s, alpha, beta, k = ob.setup_problem(shape, 0.1e-6, 2 * np.pi / 0.5e-6, eps_held[None, None],
                                              current_density=J, vectorial=False, mu=mu, max_allowed_gain=thr)
    conj = case != 'eps_gain_mu_loss'
    assert (alpha.imag - 0.05 < 0) == conj  # sigma_min > 0.05 here, so the conjugate stays negative
    np.testing.assert_allclose(sol.alpha.real, alpha.real, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(sol.alpha.imag, alpha.imag, rtol=1e-9)
    np.testing.assert_allclose(sol.beta, beta, rtol=1e-12)
    o = ob.BornOracle(chi, s, alpha, beta, k, False)
    tol = 1e-5 if dtype == np.complex64 else 1e-10
    for it in range(10):
        acc = sol.step()
        acc_o = o.step()
        assert acc == acc_o, it
        assert sol.alpha == pytest.approx(o.alpha, rel=1e-9), it
        E = sol.field_array.cpu().numpy()
        err = np.linalg.norm((E - o.E).ravel()) / max(np.linalg.norm(o.E.ravel()), 1e-300)
        assert err < tol, (it, err)
