"""Magnetic media (macromax_amd/magnetic.py) on the GPU.

- The reference's magnetic solve (python/tests/test_solution.py:360-418): a plate of refractive index -1.5 (permittivity
  and permeability both negative) lit at 30 degrees, complex64, residue < 1e-3 within 1000 iterations, and the
  golden field norms in front of and behind the plate (0.0002582 and 0.0001864, decimal 5), with the derived
  fields' dtypes.
- The magnetic iteration (curl terms, its own alpha and beta) converges to the same field as the fused
  non-magnetic kernels when the medium is non-magnetic up to a bi-anisotropy xi = 1e-7 at one voxel (complex128,
  both to residue 1e-8; fields within 1e-5 relative).
"""
import numpy as np
import numpy.testing as npt
import pytest

pytestmark = pytest.mark.gpu


def test_negative_index_plate_matches_reference_norms():
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print('magnetic: negative-index plate', flush=True)
    wavelength = 500e-9
    boundary_thickness = 2e-6
    beam_diameter = 5e-6
    plate_thickness = 5e-6
    k0 = 2 * np.pi / wavelength
    grid = mm.Grid(np.ones(2) * 128, wavelength / 4)
    angle = 30 * np.pi / 180

    def rot_z(a):
        return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    k_inc = rot_z(angle) * k0 @ np.array([1, 0, 0])
    pol = (rot_z(angle) @ np.array([0, 1, 1j]) / np.sqrt(2))[:, np.newaxis, np.newaxis]
    J = np.exp(1j * (k_inc[0] * grid[0] + k_inc[1] * grid[1]))
    src = grid.shape[0] - int(boundary_thickness / grid.step[0])
    J[:src, :] = 0
    J[src + 1:, :] = 0
    J = J * np.exp(-0.5 * ((grid[1] - grid[1].ravel()[grid.shape[0] // 3]) / (beam_diameter / 2)) ** 2)
    J = J[np.newaxis, ...] * pol
    n = 1 + (-1.5 - 1) * np.ones(grid[1].shape) * (np.abs(grid[0]) < plate_thickness / 2)
    bound = LinearBound(grid, thickness=boundary_thickness, max_extinction_coefficient=0.50)
    sol = mm.solve(grid, vacuum_wavelength=wavelength, current_density=J, refractive_index=n, bound=bound,
                   callback=lambda s: s.iteration < 1e4 and s.residue > 1e-3, dtype=np.complex64)
    assert sol.magnetic
    assert sol.residue < 1e-3, sol.residue
    assert sol.iteration <= 1000, sol.iteration
    E = sol.E
    front = E[:, src + 1:src + 8, :]
    back = E[:, -src - 8:-src - 1, :]
    npt.assert_almost_equal(np.linalg.norm(back.ravel()), 0.0001864, decimal=5)
    npt.assert_almost_equal(np.linalg.norm(front.ravel()), 0.0002582, decimal=5)
    assert E.dtype == np.complex64
    assert sol.B.dtype == np.complex64 and sol.D.dtype == np.complex64 and sol.H.dtype == np.complex64
    assert sol.S.dtype == np.float32


@pytest.mark.parametrize('anisotropic', [False, True], ids=['iso', 'aniso'])
def test_magnetic_iteration_matches_fused_kernels(anisotropic):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'magnetic vs non-magnetic path {"aniso" if anisotropic else "iso"}', flush=True)
    wavelength = 500e-9
    grid = mm.Grid([64, 48], wavelength / 4)
    rng = np.random.default_rng(4)
    eps = (1.0 + 0.4 * rng.random(grid.shape)) ** 2
    if anisotropic:
        a = rng.standard_normal((3, *grid.shape))
        a /= np.linalg.norm(a, axis=0)
        eps = eps * np.eye(3).reshape(3, 3, 1, 1) + 0.3 * a[:, None] * a[None, :]
    J = np.zeros((3, *grid.shape), dtype=np.complex128)
    J[1, 12, :] = np.exp(-((np.arange(grid.shape[1]) - 24) / 8.0) ** 2)
    J[2, 12, :] = 0.5j * J[1, 12, :]
    bound = LinearBound(grid, thickness=2e-6, max_extinction_coefficient=0.25)
    xi = np.zeros(grid.shape, dtype=np.complex128)
    xi[30, 20] = 1e-7  # above the allclose threshold that makes a medium magnetic, physically negligible
    cb = lambda s: s.iteration < 20000 and s.residue > 1e-8
    plain = mm.solve(grid, vacuum_wavelength=wavelength, current_density=J, epsilon=eps, bound=bound,
                     dtype=np.complex128, callback=cb)
    mag = mm.solve(grid, vacuum_wavelength=wavelength, current_density=J, epsilon=eps, xi=xi, bound=bound,
                   dtype=np.complex128, callback=cb)
    assert mag.magnetic and not plain.magnetic
    assert mag.residue <= 1e-8 and plain.residue <= 1e-8
    err = np.linalg.norm(mag.E - plain.E) / np.linalg.norm(plain.E)
    assert err <= 1e-5, err
    # H of the magnetic path (mu^-1 curl E) equals the non-magnetic one for mu = 1
    errH = np.linalg.norm(mag.H - plain.H) / np.linalg.norm(plain.H)
    assert errH <= 1e-5, errH
    with pytest.raises(NotImplementedError):
        mag.refine()


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128], ids=['c64', 'c128'])
def test_magnetic_trajectory_matches_oracle(dtype):
    """A medium with varying permeability and bi-isotropy (mu, xi, zeta per voxel): alpha and beta against the
    oracle's search (solver.py:405-443 restated in NumPy), then the field after each of 10 iterations against
    oracle/born_oracle.py MagneticBornOracle from the same alpha and beta (relative L2 <= 1e-10 complex128,
    <= 2e-5 complex64)."""
    import scipy.constants as const
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    print(f'magnetic trajectory {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(11)
    shape = (32, 24)
    pitch, wavelength = 0.1e-6, 0.5e-6
    k0 = 2 * np.pi / wavelength
    grid = mm.Grid(list(shape), pitch)
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2
    mu = 1.0 + 0.3 * rng.random(shape)
    xi = 0.05 * (rng.random(shape) + 1j * rng.random(shape))
    zeta = 0.05 * rng.random(shape)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, 8, :] = np.exp(-((np.arange(shape[1]) - 12) / 5.0) ** 2)
    J[2, 8, :] = 0.4j * J[1, 8, :]
    bound = LinearBound(grid, thickness=0.6e-6)
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, current_density=J, epsilon=eps, mu=mu, xi=xi, zeta=zeta,
                      bound=bound, dtype=dtype)
    assert sol.magnetic
    eps_b = eps + bound.electric_susceptibility
    thickness = np.asarray(bound.thickness)
    a_o, b_o, _, _, _ = ob.magnetic_setup(shape, pitch, k0, eps_b, xi, zeta, mu, thickness.ravel())
    tol_ab = 1e-9 if dtype == np.complex128 else 1e-4
    assert abs(sol.alpha - a_o) <= tol_ab * abs(a_o) and abs(sol.beta - b_o) <= tol_ab * abs(b_o), \
        (sol.alpha, a_o, sol.beta, b_o)
    alpha, beta = sol.alpha, sol.beta
    _, _, ops, _, _ = ob.magnetic_setup(shape, pitch, k0, eps_b, xi, zeta, mu, thickness.ravel(),
                                        alpha_beta=(alpha, beta))
    S = J.reshape(3, 1, *shape) * (-1j * k0 * const.c * const.mu_0)
    k = ob.k_vectors(shape, np.array([pitch, pitch]) * k0)
    o = ob.MagneticBornOracle(ops, S * (1j / alpha.imag / beta), alpha, beta, k)
    tol = 1e-10 if dtype == np.complex128 else 2e-5
    for it in range(10):
        sol.step()
        o.step()
        assert sol.alpha == o.alpha
        Eg = sol.field_array.cpu().numpy()
        err = np.linalg.norm(Eg - o.E) / np.linalg.norm(o.E)
        assert err <= tol, (it, err)


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128], ids=['c64', 'c128'])
def test_magnetic_graph_replays_match_eager_launches(dtype):
    """The magnetic iteration replayed from HIP graphs (Solution.magnetic_graphs) gives the same fields, norms and
    alpha as the same iteration launched kernel by kernel, over 14 iterations that include eager work and host
    reads between replays (derived fields, norms) and a rejected update (the field scaled by 50 in place), after
    which alpha changes and the graphs are recorded again."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'magnetic graphs vs eager {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(12)
    shape = (32, 24)
    grid = mm.Grid(list(shape), 0.1e-6)
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2
    mu = 1.0 + 0.3 * rng.random(shape)
    xi = 0.05 * (rng.random(shape) + 1j * rng.random(shape))
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, 8, :] = np.exp(-((np.arange(shape[1]) - 12) / 5.0) ** 2)
    bound = LinearBound(grid, thickness=0.6e-6)

    def make(graphs):
        s = mm.Solution(grid, vacuum_wavelength=0.5e-6, current_density=J, epsilon=eps, mu=mu, xi=xi,
                        bound=bound, dtype=dtype)
        s.magnetic_graphs = graphs
        return s
    eager, graphed = make(False), make(True)
    assert graphed.magnetic and graphed.alpha == eager.alpha
    rejected = 0
    for it in range(14):
        if it == 8:
            for s in (eager, graphed):
                s.field_array.mul_(50.0)  # in place: the next update is larger than the last one
        acc_e, acc_g = eager.step(), graphed.step()
        assert acc_e == acc_g, it
        rejected += not acc_g
        npt.assert_array_equal(graphed.field_array.cpu().numpy(), eager.field_array.cpu().numpy(), err_msg=str(it))
        assert graphed.residue == eager.residue and graphed.alpha == eager.alpha, it
        if it % 3 == 1:
            npt.assert_array_equal(graphed.H, eager.H)
    assert rejected >= 1


@pytest.mark.parametrize('shape', [(64, 128), (64, 20, 64)], ids=str)
def test_magnetic_split_norms_match_the_launch_stream_reduction(shape):
    """The magnetic iteration with its norm reduction and host copy on a second stream
    (Solution.magnetic_split_norms, mmx_magnetic_iterate_split; fused passes) gives the same fields, residues and
    alpha as with both on the launch stream, over 12 iterations that include a rejected update."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print(f'magnetic split norms {shape}', flush=True)
    rng = np.random.default_rng(4)
    grid = mm.Grid(list(shape), 0.1e-6)
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2
    mu = 1.0 + 0.3 * rng.random(shape)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, 8] = 1.0
    bound = LinearBound(grid, thickness=0.6e-6)

    def make(split):
        s = mm.Solution(grid, vacuum_wavelength=0.5e-6, current_density=J, epsilon=eps, mu=mu, bound=bound,
                        dtype=np.complex64)
        s.magnetic_split_norms = split
        return s
    a, b = make(False), make(True)
    assert a.magnetic and b.magnetic
    rejected = 0
    for it in range(12):
        if it == 7:
            for s in (a, b):
                s.field_array.mul_(50.0)
        acc_a, acc_b = a.step(), b.step()
        assert acc_a == acc_b, it
        rejected += not acc_b
        npt.assert_array_equal(b.field_array.cpu().numpy(), a.field_array.cpu().numpy(), err_msg=str(it))
        assert b.residue == a.residue and b.alpha == a.alpha, it
    assert rejected >= 1


def test_fused_magnetic_graph_replays_match_direct_launches():
    """On a grid with fused magnetic passes (complex64, 64 x 128: the two-group passes with more than 64 KiB of LDS,
    enabled per kernel on first launch), iterations replayed from HIP graphs give the fields, residues and alpha of
    direct launches bit for bit, through a rejected update."""
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print('fused magnetic graphs vs direct', flush=True)
    shape = (64, 128)
    rng = np.random.default_rng(8)
    grid = mm.Grid(list(shape), 0.1e-6)
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2
    mu = 1.0 + 0.3 * rng.random(shape)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, 8] = 1.0
    bound = LinearBound(grid, thickness=0.6e-6)

    def make(graphs):
        s = mm.Solution(grid, vacuum_wavelength=0.5e-6, current_density=J, epsilon=eps, mu=mu, bound=bound,
                        dtype=np.complex64)
        s.magnetic_graphs = graphs
        return s
    direct, graphed = make(False), make(True)
    rejected = 0
    for it in range(12):
        if it == 7:
            for s in (direct, graphed):
                s.field_array.mul_(50.0)
        acc_d, acc_g = direct.step(), graphed.step()
        assert acc_d == acc_g, it
        rejected += not acc_g
        npt.assert_array_equal(graphed.field_array.cpu().numpy(), direct.field_array.cpu().numpy(), err_msg=str(it))
        assert graphed.residue == direct.residue and graphed.alpha == direct.alpha, it
    assert rejected >= 1


def test_magnetic_graphs_follow_a_reallocated_plan_scratch():
    """The plan's FP64 scratch buffer grows when a larger array is reduced on the same plan (MMX_PARAM_SCRATCH_GEN
    counts the allocations); graphs recorded before hold the old address, so the next magnetic step records them
    again, and the iterates stay equal to eager launches."""
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    print('magnetic graphs after scratch growth', flush=True)
    rng = np.random.default_rng(13)
    shape = (32, 24)
    grid = mm.Grid(list(shape), 0.1e-6)
    mu = 1.0 + 0.3 * rng.random(shape)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[1, 8, :] = 1.0
    bound = LinearBound(grid, thickness=0.6e-6)

    def make(graphs):
        s = mm.Solution(grid, vacuum_wavelength=0.5e-6, current_density=J, epsilon=2.0, mu=mu, bound=bound,
                        dtype=np.complex64)
        s.magnetic_graphs = graphs
        return s
    eager, graphed = make(False), make(True)
    for _ in range(4):
        eager.step()
        graphed.step()
    plan = graphed._Solution__BE.plan
    gen = plan.get_param('scratch_gen')
    dev = graphed.field_array.device
    V = int(np.prod(shape))
    chi = torch.full((1, 1, *shape), 0.1, dtype=torch.complex64, device=dev)
    s1 = torch.zeros((3, 1, *shape), dtype=torch.complex64, device=dev)
    for nbatch in (64, 256, 1024):  # batched Born steps keep per-tile partial sums of every problem in the scratch
        E = torch.ones((nbatch, 3, 1, *shape), dtype=torch.complex64, device=dev)
        Eo = torch.empty_like(E)
        coef = torch.tensor([1.0, 1.0, 1.0, 0.0, 0.0, 0.0] * nbatch, dtype=torch.float32, device=dev)
        norms = torch.zeros(2 * nbatch, dtype=torch.float64, device=dev)
        plan.born_step(E, Eo, s1, chi, 3, 1, coef, nbatch, 0, 0, norms)
        torch.cuda.synchronize()
        if plan.get_param('scratch_gen') > gen:
            break
    assert plan.get_param('scratch_gen') > gen, 'the scratch buffer was not reallocated'
    assert np.all(np.isfinite(norms.cpu().numpy())) and V > 0
    for it in range(4):
        eager.step()
        graphed.step()
        npt.assert_array_equal(graphed.field_array.cpu().numpy(), eager.field_array.cpu().numpy(), err_msg=str(it))
        assert graphed.residue == eager.residue


@pytest.mark.parametrize('dtype', [np.complex64, np.complex128], ids=['c64', 'c128'])
def test_magnetic_3d_tensor_media_match_oracle(dtype):
    """3D medium with a constant anisotropic permeability tensor and a bi-isotropy xi per voxel, so the fused
    iteration (mmx_magnetic_iterate) sees a constant 3x3 chi_HH, 3x3 fields chi_EH and chi_HE and a scalar chi_EE:
    10 iterates against MagneticBornOracle (relative L2 <= 1e-10 complex128, <= 2e-5 complex64), and the first update
    against the composed operator (MagneticOperator.update: curls, Green's operator and element-wise products)."""
    import scipy.constants as const
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    print(f'magnetic 3d tensors {np.dtype(dtype).name}', flush=True)
    rng = np.random.default_rng(12)
    shape = (12, 10, 16)
    pitch, wavelength = 0.1e-6, 0.5e-6
    k0 = 2 * np.pi / wavelength
    grid = mm.Grid(list(shape), pitch)
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2
    m = rng.standard_normal((3, 3)) * 0.1
    mu = (np.eye(3) * 1.2 + m @ m.T).reshape(3, 3, 1, 1, 1)
    xi = (0.05 * (rng.random(shape) + 1j * rng.random(shape)))
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[0, :, :, 4] = 1.0
    J[1, :, :, 4] = 0.3j
    bound = LinearBound(grid, thickness=0.3e-6)
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, current_density=J, epsilon=eps, mu=mu, xi=xi,
                      bound=bound, dtype=dtype)
    assert sol.magnetic
    mag = sol._Solution__mag
    E0 = torch.zeros((3, 1, *shape), dtype=mag.s.dtype, device=mag.s.device)
    d_composed = mag.update(E0)
    E1 = torch.empty_like(E0)
    norms = torch.zeros(2, dtype=torch.float64, device=E0.device)
    mag.iterate(E0, E1, norms)
    rel = float(torch.linalg.vector_norm(E1 - d_composed) / torch.linalg.vector_norm(d_composed))
    assert rel <= (1e-12 if dtype == np.complex128 else 2e-6), rel
    npt.assert_allclose(norms[0].item(), float(torch.linalg.vector_norm(d_composed)) ** 2, rtol=1e-5)
    eps_b = eps + bound.electric_susceptibility
    alpha, beta = sol.alpha, sol.beta
    _, _, ops, _, _ = ob.magnetic_setup(shape, pitch, k0, eps_b, xi, 0.0, mu, np.asarray(bound.thickness).ravel(),
                                        alpha_beta=(alpha, beta))
    S = J.reshape(3, 1, *shape) * (-1j * k0 * const.c * const.mu_0)
    k = ob.k_vectors(shape, np.array([pitch] * 3) * k0)
    o = ob.MagneticBornOracle(ops, S * (1j / alpha.imag / beta), alpha, beta, k)
    tol = 1e-10 if dtype == np.complex128 else 2e-5
    for it in range(10):
        sol.step()
        o.step()
        assert sol.alpha == o.alpha
        Eg = sol.field_array.cpu().numpy()
        err = np.linalg.norm(Eg - o.E) / np.linalg.norm(o.E)
        assert err <= tol, (it, err)


def test_magnetic_fine_grid_extreme_alpha_matches_oracle():
    """A negative-index plate on a fine grid (pitch lambda / 16): the reference's two-parameter search (restated by
    the oracle) drives beta towards 0 and alpha_i to about 4.5e19, whose square overflows single precision. The
    complex64 Green's operator is evaluated without alpha_i^2; 10 iterates against MagneticBornOracle within 2e-5."""
    import scipy.constants as const
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    print('magnetic plate at lambda / 16, extreme alpha', flush=True)
    wavelength = 500e-9
    k0 = 2 * np.pi / wavelength
    size, pitch = 64, wavelength / 16
    grid = mm.Grid([size, size], pitch)
    nidx = 1 + (-1.5 - 1) * np.ones(grid[1].shape) * (np.abs(grid[0]) < 0.5e-6)
    J = np.zeros((3, size, size), dtype=np.complex128)
    J[1, 8, :] = np.exp(-((np.arange(size) - 32) / 10.0) ** 2)
    J[2, 8, :] = 0.5j * J[1, 8, :]
    bound = LinearBound(grid, thickness=0.5e-6, max_extinction_coefficient=0.50)
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, current_density=J, refractive_index=nidx, bound=bound,
                      dtype=np.complex64)
    assert sol.magnetic and sol.alpha.imag > 1e18, sol.alpha
    sign = 1 - 2 * (nidx < 0)
    eps_b = (nidx ** 2 * sign)[None, None] + bound.electric_susceptibility
    _, _, ops, _, _ = ob.magnetic_setup((size, size), pitch, k0, eps_b, 0.0, 0.0, sign[None, None].astype(complex),
                                        np.asarray(bound.thickness).ravel(), alpha_beta=(sol.alpha, sol.beta))
    S = J.reshape(3, 1, size, size) * (-1j * k0 * const.c * const.mu_0)
    k = ob.k_vectors((size, size), np.array([pitch, pitch]) * k0)
    o = ob.MagneticBornOracle(ops, S * (1j / sol.alpha.imag / sol.beta), sol.alpha, sol.beta, k)
    for it in range(10):
        sol.step()
        o.step()
        Eg = sol.field_array.cpu().numpy()
        err = np.linalg.norm(Eg - o.E) / np.linalg.norm(o.E)
        assert err <= 2e-5, (it, err)


@pytest.mark.parametrize('shape', [(64, 128), (128, 64), (64, 20, 64), (128, 12, 256)], ids=lambda s: 'x'.join(map(str, s)))
def test_magnetic_fused_passes_match_oracle_and_unfused(shape):
    """Grids whose first and last axes have fused magnetic passes (complex64, 64..1024 points): per-voxel anisotropic
    permittivity, a constant 3x3 permeability and per-voxel bi-isotropy xi and zeta, so every susceptibility kind
    reaches the pass kernels. The first update equals the unfused kernel sequence (register kernels off) and the
    composed operator (MagneticOperator.update) to 2e-6, and 10 iterates follow MagneticBornOracle within 2e-5."""
    import scipy.constants as const
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    print(f'magnetic fused passes {shape}', flush=True)
    rng = np.random.default_rng(21)
    nd = len(shape)
    pitch, wavelength = 0.1e-6, 0.5e-6
    k0 = 2 * np.pi / wavelength
    grid = mm.Grid(list(shape), pitch)
    a = rng.standard_normal((3, *shape))
    a /= np.linalg.norm(a, axis=0)
    eye = np.eye(3).reshape(3, 3, *([1] * nd))
    eps = (1.0 + 0.3 * rng.random(shape)) ** 2 * eye + 0.2 * a[:, None] * a[None, :]
    m = rng.standard_normal((3, 3)) * 0.1
    mu = (np.eye(3) * 1.2 + m @ m.T).reshape(3, 3, *([1] * nd))
    xi = 0.05 * (rng.random(shape) + 1j * rng.random(shape))
    zeta = 0.03 * rng.random(shape)
    J = np.zeros((3, *shape), dtype=np.complex128)
    J[(slice(0, 1),) + (slice(None),) * (nd - 1) + (slice(4, 5),)] = 1.0
    J[(slice(1, 2),) + (slice(None),) * (nd - 1) + (slice(4, 5),)] = 0.3j
    bound = LinearBound(grid, thickness=0.4e-6)
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, current_density=J, epsilon=eps, mu=mu, xi=xi, zeta=zeta,
                      bound=bound, dtype=np.complex64)
    assert sol.magnetic
    mag = sol._Solution__mag
    g = torch.Generator(device=mag.s.device).manual_seed(3)
    E0 = (torch.randn((3, 1, *shape), generator=g, device=mag.s.device, dtype=torch.float32)
          + 1j * torch.randn((3, 1, *shape), generator=g, device=mag.s.device, dtype=torch.float32)) * 1e-3
    E0 = E0.to(mag.s.dtype)
    norms = torch.zeros(2, dtype=torch.float64, device=E0.device)
    E_fused = torch.empty_like(E0)
    mag.iterate(E0, E_fused, norms)
    n_fused = norms.clone()
    d_composed = mag.update(E0)
    mag.plan.set_param('reg_fft', 0)
    E_unfused = torch.empty_like(E0)
    mag.iterate(E0, E_unfused, norms)
    mag.plan.set_param('reg_fft', 1)
    d_f, d_u = E_fused - E0, E_unfused - E0
    rel = float(torch.linalg.vector_norm(d_f - d_u) / torch.linalg.vector_norm(d_u))
    assert rel <= 2e-6, rel
    rel = float(torch.linalg.vector_norm(d_f - d_composed) / torch.linalg.vector_norm(d_composed))
    assert rel <= 2e-6, rel
    npt.assert_allclose(n_fused.cpu().numpy(), norms.cpu().numpy(), rtol=1e-5)
    eps_b = eps + bound.electric_susceptibility * eye
    alpha, beta = sol.alpha, sol.beta
    _, _, ops, _, _ = ob.magnetic_setup(shape, pitch, k0, eps_b, xi, zeta, mu, np.asarray(bound.thickness).ravel(),
                                        alpha_beta=(alpha, beta))
    S = J.reshape(3, 1, *shape) * (-1j * k0 * const.c * const.mu_0)
    k = ob.k_vectors(shape, np.array([pitch] * nd) * k0)
    o = ob.MagneticBornOracle(ops, S * (1j / alpha.imag / beta), alpha, beta, k)
    for it in range(10):
        sol.step()
        o.step()
        assert sol.alpha == o.alpha
        Eg = sol.field_array.cpu().numpy()
        err = np.linalg.norm(Eg - o.E) / np.linalg.norm(o.E)
        assert err <= 2e-5, (it, err)
