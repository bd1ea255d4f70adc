"""The Solution API on the HIP backend: the reference's solver tests (tests/test_solution.py) and per-iteration
parity of the field trajectory with the complex128 oracle.

Tolerances: analytic plane waves as in the reference (residue < 1e-6 within 70 / 85 iterations, field error
decimal=3 in the interior and decimal=2 overall); trajectory E_k against the oracle started from the same alpha,
relative L2 <= 1e-5 for complex64 after every one of the first 10 iterations (<= 1e-10 for complex128).
"""
import numpy as np
import numpy.testing as npt
import pytest
import scipy.constants as const

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu


def _plane_wave(kind, dtype=np.complex64):
    import macromax_amd as mm
    wavelength = 500e-9
    k0 = 2 * np.pi / wavelength
    n = 1024
    pitch = wavelength / 16
    bt = 10e-6
    x = pitch * np.arange(n) - bt
    f = np.maximum(0, np.maximum((x[0] + bt) - x, x - (x[-1] - bt)) / bt)
    eps = (1 + 1j * 0.1 * f) ** 2
    J = np.zeros(n)
    J[np.argmin(np.abs(x))] = 1.0
    if kind == 'scalar':
        Jc = J.astype(dtype)
    else:
        Jc = (np.array([0, 1, 0])[:, None] * J[None]).astype(dtype)
        if kind == 'anisotropic':
            eps = eps * np.eye(3)[..., None]
            eps[:, :, f <= 0] = np.array([[1.1, 0, 0.3], [0, 1, 0], [0.3, 0, 1.1]])[..., None]
    sol = mm.solve(x, vacuum_wavelength=wavelength, current_density=Jc, epsilon=eps,
                   callback=lambda s: s.residue > 1e-6 and s.iteration < 1e4)
    ref = const.mu_0 * pitch / 2 * const.c * np.exp(1j * k0 * np.abs(x)) * np.exp(-0.1 * k0 * bt * f ** 2 / 2)
    ref = ref[None] if kind == 'scalar' else np.array([0, 1, 0])[:, None] * ref
    sel = (wavelength * 10 < x) & (x < x[-1] - bt - wavelength * 10)
    return sol, ref, sel


@pytest.mark.parametrize('kind,max_iter', [('scalar', 70), ('vectorial', 70), ('anisotropic', 85)])
def test_analytic_plane_wave(kind, max_iter):
    """tests/test_solution.py:91-358 on the HIP backend (complex64)."""
    print(f'plane wave {kind}', flush=True)
    sol, ref, sel = _plane_wave(kind)
    assert sol.residue < 1e-6
    assert sol.iteration <= max_iter
    E = sol.E
    assert E.dtype == np.complex64
    err = E - ref
    npt.assert_almost_equal(np.sqrt(np.mean(np.abs(err[:, sel]) ** 2)) / np.sqrt(np.mean(np.abs(E[:, sel]) ** 2)), 0, decimal=3)
    npt.assert_almost_equal(np.sqrt(np.mean(np.abs(err) ** 2)) / np.sqrt(np.mean(np.abs(E) ** 2)), 0, decimal=2)
    if kind != 'scalar':
        B = sol.B
        H = sol.H
        npt.assert_array_almost_equal(B[2, sel] * const.c, E[1, sel], decimal=4)
        npt.assert_array_almost_equal(B[2, sel] / const.mu_0, H[2, sel], decimal=10)
        if kind == 'vectorial':
            npt.assert_array_equal(E[[0, 2]], 0)


def test_refractive_index_update_and_resolve():
    """tests/test_solution.py:246-271: setting refractive_index rebuilds the preconditioner; the modified medium
    converges, and restoring the index and re-solving recovers the analytic plane wave."""
    sol, ref, sel = _plane_wave('vectorial')
    n_before = sol.refractive_index
    npt.assert_allclose(sol.epsilon, n_before ** 2, rtol=1e-5)
    alpha_before = sol.alpha
    x = np.asarray(sol.grid[0]).ravel()
    slab = (500e-9 <= x) & (x <= 1000e-9)
    n_new = n_before.copy()
    n_new[..., slab] = 1.5
    sol.refractive_index = n_new
    npt.assert_allclose(sol.epsilon[..., slab], 2.25, rtol=1e-6)
    assert sol.alpha != alpha_before
    sol.solve(lambda s: s.residue > 1e-6 and s.iteration < 1e4)
    assert sol.residue < 1e-6
    sol.refractive_index = n_before
    sol.solve(lambda s: s.residue > 1e-6 and s.iteration < 1e4)
    assert sol.residue < 1e-6
    E = sol.E
    err = E - ref
    npt.assert_almost_equal(np.sqrt(np.mean(np.abs(err[:, sel]) ** 2)) / np.sqrt(np.mean(np.abs(E[:, sel]) ** 2)), 0, decimal=3)


@pytest.fixture()
def sol3d():
    """tests/test_solution.py:12-32: 50x100x200 vectorial, complex128, absorbing permittivity along axis 0."""
    import macromax_amd as mm
    wavelength = 4
    grid = mm.Grid(shape=[50, 100, 200], step=wavelength / 4)
    J = np.zeros([3, *grid.shape])
    bt = 5 * wavelength
    c = np.asarray(J.shape) // 2
    J[:, c[1], c[2], c[3]] = [0.0, 1.0, 0.0]
    d = np.maximum(0.0, np.maximum(grid[0].ravel()[0] + bt - grid[0], grid[0].ravel()[-1] - bt - grid[0]) / bt)
    return mm.Solution(grid, vacuum_wavelength=wavelength, epsilon=1.0 + 0.2j * d, current_density=J,
                       dtype=np.complex128)


def test_grid_wavenumber_frequency_wavelength(sol3d):
    """tests/test_solution.py:34-44, and the same medium specified by wavenumber or angular frequency."""
    import macromax_amd as mm
    wavelength = 4
    assert sol3d.grid == mm.Grid(shape=[50, 100, 200], step=wavelength / 4)
    npt.assert_almost_equal(sol3d.wavenumber, 2 * np.pi / wavelength)
    npt.assert_almost_equal(sol3d.angular_frequency / (2 * np.pi * const.c / wavelength), 1.0)
    npt.assert_almost_equal(sol3d.wavelength, wavelength)
    grid = mm.Grid(shape=[16, 24], step=0.5)
    for kw in ({'wavenumber': 2 * np.pi / wavelength}, {'angular_frequency': 2 * np.pi * const.c / wavelength},
               {'vacuum_wavelength': wavelength}):
        s = mm.Solution(grid, epsilon=1.5, current_density=np.ones([1, 16, 24]), **kw)
        npt.assert_almost_equal(s.wavelength, wavelength, err_msg=str(kw))
        npt.assert_almost_equal(s.wavenumber, 2 * np.pi / wavelength, err_msg=str(kw))


def test_iteration_counter(sol3d):
    """tests/test_solution.py:46-57."""
    assert sol3d.iteration == 0
    it = sol3d.__iter__()
    next(it)
    assert sol3d.iteration == 1
    sol3d.solve(lambda s: s.iteration < 5)
    assert sol3d.iteration == 5
    sol3d.iteration = 1
    next(it)
    assert sol3d.iteration == 2
    next(sol3d.__iter__())
    assert sol3d.iteration == 1


def test_update_norm_and_residue(sol3d):
    """tests/test_solution.py:59-77."""
    E0 = sol3d.E.copy()
    next(sol3d.__iter__())
    E1 = sol3d.E.copy()
    npt.assert_almost_equal(sol3d.previous_update_norm, np.sqrt(np.sum(np.abs(E1 - E0) ** 2)))
    npt.assert_almost_equal(sol3d.residue, np.linalg.norm((E1 - E0).ravel()) / np.linalg.norm(E1.ravel()))
    next(sol3d.__iter__())
    E2 = sol3d.E.copy()
    npt.assert_almost_equal(sol3d.previous_update_norm, np.sqrt(np.sum(np.abs(E2 - E1) ** 2)))
    npt.assert_almost_equal(sol3d.residue, np.linalg.norm((E2 - E1).ravel()) / np.linalg.norm(E2.ravel()))


def test_set_E_and_resolve(sol3d):
    """tests/test_solution.py:79-89."""
    E0 = sol3d.E
    sol3d.E = E0
    npt.assert_almost_equal(sol3d.E, E0)
    sol3d.solve(lambda s: s.iteration < 10)
    E_sol = sol3d.E
    sol3d.E = E0
    npt.assert_almost_equal(sol3d.E, E0)
    sol3d.solve(lambda s: s.iteration < 10)
    npt.assert_almost_equal(sol3d.E, E_sol)


def test_in_place_field_edit_voids_the_iteration_launched_ahead():
    """Solution.step launches the next iteration before it decides on the current one. Scaling the current field in
    place (field_array) between steps must void that iteration: the next update is then about 50 times the last one
    and is rejected (alpha_i x 1.5), and the following iterates equal the oracle's from the scaled field."""
    import macromax_amd as mm
    grid = mm.Grid(shape=[32, 40], step=0.125)
    rng = np.random.default_rng(2)
    eps = (1.0 + 0.3 * rng.random(grid.shape)) ** 2
    J = np.zeros([1, *grid.shape])
    J[0, 8, :] = 1.0
    sol = mm.Solution(grid, vacuum_wavelength=1.0, epsilon=eps, current_density=J, dtype=np.complex128)
    for it in range(6):
        assert sol.step()
    alpha = sol.alpha
    sol.field_array.mul_(50.0)
    assert sol.step() is False
    assert sol.alpha == complex(alpha.real, 1.5 * alpha.imag)
    o = _oracle_from_solution(sol)
    o.previous_update_norm = sol.previous_update_norm * sol.wavenumber ** 2
    for it in range(4):
        assert sol.step() == o.step(), it
        assert sol.alpha == o.alpha, it
        err = np.linalg.norm(sol.field_array.cpu().numpy() - o.E) / np.linalg.norm(o.E)
        assert err <= 1e-10, (it, err)


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128], ids=['c64', 'c128'])
def test_split_norm_reduction_matches_the_launch_stream_reduction(dtype):
    """On problems large enough, Solution reduces each iteration's norms on a second stream beside the next iteration
    (mmx_born_iterate_split, two partial slots). Fields, residues and alpha equal those of the launch-stream
    reduction bit for bit over 14 iterations, through a rejected update (field scaled in place), after which the
    slots are reused while the voided iteration's reduction may still be queued."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    shape = (64, 64, 48)
    rng = np.random.default_rng(6)
    grid = mm.Grid(shape=list(shape), step=0.1e-6)
    eps = (1.0 + 0.4 * rng.random(shape)) ** 2
    J = np.zeros((3, *shape), dtype=np.complex64)
    J[0, :, :, 10] = 1.0
    sols = []
    for split in (False, True):
        sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J,
                          bound=LinearBound(grid, thickness=0.6e-6), dtype=dtype)
        sol._Solution__copy_stream = None if not split else sol._Solution__copy_stream
        sols.append(sol)
    plain, split = sols
    assert split._Solution__copy_stream is not None  # 64 x 64 x 48 x 3 values x 4 > 2^20: the split path
    rejected = 0
    for it in range(14):
        if it == 8:
            for s in sols:
                s.field_array.mul_(50.0)
        acc_p, acc_s = plain.step(), split.step()
        assert acc_p == acc_s, it
        rejected += not acc_s
        npt.assert_array_equal(split.field_array.cpu().numpy(), plain.field_array.cpu().numpy(), err_msg=str(it))
        assert split.residue == plain.residue and split.alpha == plain.alpha, it
    assert rejected >= 1


def test_adaptive_stop_criterion():
    """macromax_amd.utils.Adaptive (utils/stop_criterion.py:7-52) as the solve callback: the solve stops at a reading
    below the target residue, within a few iterations of where the every-iteration rule stops."""
    import bench
    from macromax_amd.utils import Adaptive
    # the 128^2 disc medium of the batched benchmark (absorbing boundary; about 900 iterations to 1e-6)
    plain = bench.disc_problem(128, 7).solve(lambda s: s.residue > 1e-6 and s.iteration < 10000)
    adaptive = bench.disc_problem(128, 7).solve(Adaptive(1e-6, 10000))
    assert plain.residue < 1e-6 and plain.iteration < 10000
    assert adaptive.residue < 1e-6
    assert plain.iteration <= adaptive.iteration <= int(1.25 * plain.iteration) + 2, (plain.iteration,
                                                                                       adaptive.iteration)


def _oracle_from_solution(sol):
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s = sol.source_normalized.cpu().numpy().astype(np.complex128)
    E = sol.field_array.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors(tuple(sol.grid.shape), np.asarray(sol.grid.step) * sol.wavenumber)
    return ob.BornOracle(chi, s, sol.alpha, sol.beta, k, sol.vectorial, E0=E)


@pytest.mark.parametrize('case', ['aniso3d', 'iso3d', 'scalar2d', 'vec2d_aniso'])
@pytest.mark.parametrize('dtype', [np.complex64, np.complex128])
def test_trajectory_matches_oracle(case, dtype):
    """Per-iteration field E_k of the HIP Solution against the complex128 oracle, same alpha and inputs."""
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'trajectory {case} {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(7)
    if case.endswith('3d'):
        shape = (24, 20, 32)
    else:
        shape = (48, 40)
    grid = mm.Grid(shape=list(shape), step=0.25e-6)
    nd = len(shape)
    n = 1.0 + 0.5 * rng.random(shape)
    if case in ('aniso3d', 'vec2d_aniso'):
        a = rng.standard_normal((3, *shape))
        a /= np.linalg.norm(a, axis=0)
        eps = 1.5 ** 2 * np.eye(3).reshape(3, 3, *[1] * nd) + (1.65 ** 2 - 1.5 ** 2) * a[:, None] * a[None, :]
    else:
        eps = n ** 2
    C = 1 if case == 'scalar2d' else 3
    J = np.zeros((C, *shape), dtype=np.complex128)
    J[0, ..., 5] = 1.0
    if C == 3:
        J[1, ..., 5] = 0.5j
    bound = LinearBound(grid, thickness=2e-6)
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J.astype(dtype), bound=bound,
                      dtype=dtype)
    o = _oracle_from_solution(sol)
    tol = 1e-5 if dtype == np.complex64 else 1e-10
    for it in range(1, 11):
        sol.step()
        o.step()
        Eg = sol.field_array.cpu().numpy()
        err = np.linalg.norm((Eg - o.E).ravel()) / np.linalg.norm(o.E.ravel())
        assert err < tol, (it, err)
    npt.assert_allclose(sol.residue, o.relative_residue, rtol=1e-4 if dtype == np.complex64 else 1e-9)


def test_divergence_guard_matches_oracle():
    """solver.py:1083-1096 on the device path: a too-small alpha_i makes updates grow; they are rejected,
    alpha_i grows by 1.5 per rejection (applied as chi_eff = a chi + c I), and the accepted trajectory follows
    the oracle, which rescales chi and s in memory as the reference does."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    rng = np.random.default_rng(3)
    shape = (32, 24, 20)
    grid = mm.Grid(shape=list(shape), step=0.25e-6)
    eps = (1.0 + 0.6 * rng.random(shape)) ** 2
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, :, :, 4] = 1.0
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J,
                      bound=LinearBound(grid, thickness=2e-6), dtype=np.complex128)
    # force an under-damped preconditioner on both sides
    small = complex(sol.alpha.real, sol.alpha.imag * 0.02)
    sol._Solution__update_operators(small)
    o = _oracle_from_solution(sol)
    rejected = 0
    for it in range(40):
        acc = sol.step()
        acc_o = o.step()
        assert acc == acc_o, it
        rejected += not acc
        npt.assert_allclose(sol.alpha, o.alpha, rtol=1e-12)
        Eg = sol.field_array.cpu().numpy()
        assert np.linalg.norm((Eg - o.E).ravel()) <= 1e-9 * max(np.linalg.norm(o.E.ravel()), 1e-300)
    assert rejected >= 1


def test_derived_fields_of_plane_wave():
    """Derived fields (solver.py:834-991) of the solved vectorial plane wave, checked against what the physics fixes
    for a y-polarised wave travelling along +x: in vacuum S_x = |E|^2 / (2 mu_0 c), D = eps_0 E,
    u = eps_0 |E|^2 / 2, stress diag(-u, 0, 0) with zero off-diagonal (torque) entries; inside the absorbing layer
    the force density equals d sigma_xx / dx and div S equals (d S_x / dx) / k0 (the reference's divS is taken
    with respect to k0 x), both against central differences of the same arrays. Tolerance 2 % of the scale of each
    quantity (complex64 solve to residue 1e-6); 5 % for the two derivatives (spectral against finite differences)."""
    print('derived fields of the vectorial plane wave', flush=True)
    sol, ref, sel = _plane_wave('vectorial')
    E = sol.E
    eps0, mu0, c = const.epsilon_0, const.mu_0, const.c
    k0 = sol.wavenumber
    x = np.asarray(sol.grid[0]).ravel()
    Ey2 = np.abs(E[1, sel]) ** 2
    S = sol.S
    assert S.shape == E.shape
    npt.assert_allclose(S[0, sel], Ey2 / (2 * mu0 * c), rtol=0, atol=2e-2 * np.max(Ey2) / (2 * mu0 * c))
    assert np.max(np.abs(S[1:, sel])) <= 2e-2 * np.max(np.abs(S[0, sel]))
    D = sol.D
    npt.assert_allclose(D[1, sel], eps0 * E[1, sel], rtol=0, atol=2e-2 * eps0 * np.max(np.abs(E[1, sel])))
    u = sol.energy_density
    assert u.shape == E.shape[1:]
    npt.assert_allclose(u[sel], 0.5 * eps0 * Ey2, rtol=0, atol=2e-2 * 0.5 * eps0 * np.max(Ey2))
    sigma = sol.stress_tensor
    assert sigma.shape == (3, 3, *E.shape[1:])
    npt.assert_array_equal(sigma, sigma.transpose(1, 0, *range(2, sigma.ndim)))
    scale = np.max(u[sel])
    npt.assert_allclose(sigma[0, 0, sel], -u[sel], rtol=0, atol=2e-2 * scale)
    assert np.max(np.abs(sigma[1, 1, sel])) <= 2e-2 * scale and np.max(np.abs(sigma[2, 2, sel])) <= 2e-2 * scale
    torque = sol.torque
    assert torque.shape == E.shape and np.max(np.abs(torque[:, sel])) <= 2e-2 * scale
    # inside the absorbing layer (away from its start and from the periodic wrap); the spectral discretisation
    # leaves a small alternating (Nyquist-frequency) tail on the field, which the derivatives amplify, so both
    # sides are compared after averaging neighbouring samples
    layer = (x > x[-1] - 9e-6) & (x < x[-1] - 2e-6)

    def pairs(a):
        a = a[..., layer]
        return 0.5 * (a[..., 1:] + a[..., :-1])
    force = sol.f
    assert force.shape == E.shape
    dsig = pairs(np.gradient(sigma[0, 0], x))
    npt.assert_allclose(pairs(force[0]), dsig, rtol=0, atol=5e-2 * np.max(np.abs(dsig)))
    assert np.max(np.abs(pairs(force[1:]))) <= 5e-2 * np.max(np.abs(dsig))
    assert np.min(pairs(force[0])) > 0  # the absorbed wave pushes along +x
    divS = sol.divS
    assert divS.shape == E.shape[1:]
    dS = pairs(np.gradient(S[0], x) / k0)
    npt.assert_allclose(pairs(divS.real), dS, rtol=0, atol=5e-2 * np.max(np.abs(dS)))
    assert np.max(pairs(divS.real)) < 0  # absorption
    assert np.max(np.abs(divS[sel])) <= 5e-2 * np.max(np.abs(S[0, sel]))  # no sources or losses in vacuum


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128], ids=['c64', 'c128'])
def test_uniaxial_storage(dtype):
    """Uniaxial media (eps = a I + b n n^T per voxel, as grains of a birefringent crystal) are stored as
    [u, w, n] (chi kind 4); a biaxial medium keeps the packed symmetric storage (kind 3). The stored operator equals
    chi = (i/alpha_i)(eps/beta - alpha I) to 8 rounding units of the element type."""
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'uniaxial storage {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(7)
    grid = mm.Grid([16, 20, 12], 0.1e-6)
    a = rng.standard_normal((3, *grid.shape))
    a /= np.linalg.norm(a, axis=0)
    eps = 1.5 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.65 ** 2 - 1.5 ** 2) * a[:, None] * a[None, :]
    J = np.zeros((3, *grid.shape), dtype=dtype)
    J[0, :, :, 3] = 1
    bound = LinearBound(grid, thickness=0.4e-6)
    sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J, bound=bound, dtype=dtype)
    assert sol.chi_kind == 4
    al, beta = sol.alpha, sol.beta
    e = eps + np.eye(3).reshape(3, 3, 1, 1, 1) * bound.electric_susceptibility
    ref = (e / beta - al * np.eye(3).reshape(3, 3, 1, 1, 1)) * (1j / al.imag)
    chi = sol.chi.cpu().numpy()
    assert np.max(np.abs(chi - ref)) <= 8 * np.finfo(dtype).eps * np.max(np.abs(ref))
    eps_b = eps.copy()
    eps_b[0, 0] += 0.2  # biaxial
    sol_b = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps_b, current_density=J, bound=bound, dtype=dtype)
    assert sol_b.chi_kind == 3


def test_chi_storage_option():
    """chi_storage forces the full (kind 2) or symmetric (kind 3) storage of a uniaxial medium: the same operator, so
    the same iterates (complex64, 1e-5); a storage that cannot hold the medium raises ValueError."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print('chi storage option', flush=True)
    rng = np.random.default_rng(17)
    grid = mm.Grid([16, 20, 24], 0.1e-6)
    a = rng.standard_normal((3, *grid.shape))
    a /= np.linalg.norm(a, axis=0)
    eps = 1.5 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.65 ** 2 - 1.5 ** 2) * a[:, None] * a[None, :]
    J = np.zeros((3, *grid.shape), dtype=np.complex64)
    J[0, :, :, 3] = 1
    J[1, :, :, 3] = 0.5j
    bound = LinearBound(grid, thickness=0.4e-6)
    fields = {}
    for storage, kind in (('auto', 4), ('uniaxial', 4), ('symmetric', 3), ('full', 2)):
        sol = mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J, bound=bound,
                          dtype=np.complex64, chi_storage=storage)
        assert sol.chi_kind == kind, storage
        sol.solve(lambda s_: s_.iteration < 8)
        fields[storage] = sol.E
    for storage in ('uniaxial', 'symmetric', 'full'):
        err = np.linalg.norm(fields[storage] - fields['auto']) / np.linalg.norm(fields['auto'])
        assert err < 1e-5, (storage, err)
    eps_b = eps.copy()
    eps_b[0, 0] += 0.2  # biaxial
    with pytest.raises(ValueError):
        mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps_b, current_density=J, bound=bound,
                    dtype=np.complex64, chi_storage='uniaxial')
    eps_n = eps_b.astype(np.complex128)
    eps_n[0, 1] += 0.05j  # non-reciprocal
    with pytest.raises(ValueError):
        mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps_n, current_density=J, bound=bound,
                    dtype=np.complex64, chi_storage='symmetric')
    with pytest.raises(ValueError):
        mm.Solution(grid, vacuum_wavelength=0.5e-6, epsilon=eps, current_density=J, bound=bound,
                    dtype=np.complex64, chi_storage='packed')


@pytest.mark.parametrize('shape,dtype', [((19, 23, 16), np.complex128), ((19, 23, 16), np.complex64),
                                         ((101, 48), np.complex128)], ids=str)
def test_prime_axis_solve_matches_oracle(shape, dtype):
    """Grids with axes of prime length (Bluestein's algorithm, the unfused iteration): the whole solve to residue
    1e-6 (1e-5 in complex64) against the oracle run with the same alpha and stopping rule: the same iteration
    count (within 2 in complex64), fields within 1e-9 (c128) / 1e-3 (c64) relative L2. The reference transforms any axis length
    (backend/numpy.py:90-94, backend/torch.py:183-205)."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'prime-axis solve {shape} {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(31)
    grid = mm.Grid(shape=list(shape), step=0.2e-6)
    nd = len(shape)
    C = 3 if nd == 3 else 1
    if nd == 3:
        a = rng.standard_normal((3, *shape))
        a /= np.linalg.norm(a, axis=0)
        eps = 1.4 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (1.6 ** 2 - 1.4 ** 2) * a[:, None] * a[None, :]
    else:
        eps = (1.0 + 0.5 * rng.random(shape)) ** 2
    J = np.zeros((C, *shape), dtype=np.complex128)
    J[0, ..., 4] = 1.0
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J.astype(dtype),
                      bound=LinearBound(grid, thickness=0.8e-6), dtype=dtype)
    o = _oracle_from_solution(sol)
    tol = 1e-6 if dtype == np.complex128 else 1e-5
    sol.solve(lambda s_: s_.residue > tol and s_.iteration < 3000)
    while True:
        o.step()
        if not (o.relative_residue > tol and o.iteration < 3000):
            break
    if dtype == np.complex128:
        assert sol.iteration == o.iteration
    else:  # single-precision rounding may move the stopping iteration at the threshold
        assert abs(sol.iteration - o.iteration) <= 2, (sol.iteration, o.iteration)
    Eg = sol.field_array.cpu().numpy()
    err = np.linalg.norm((Eg - o.E).ravel()) / np.linalg.norm(o.E.ravel())
    assert err < (1e-9 if dtype == np.complex128 else 1e-3), err
