"""Batched solves (macromax_amd.batch) against sequential Solution.solve on the same problems.

Every problem of a batch follows the rules of Solution.__iter__ / Solution.solve (solver.py:1064-1123) with the
decision taken on the device, so each problem must reach the same iteration count, alpha and field as its own
Solution: complex128 fields within 1e-10 relative L2, complex64 within 1e-5 (the FP64 norm partials are summed over
differently shaped tiles in a batch, which can only move the accept decision at an exact tie).
"""
import numpy as np
import pytest
import scipy.constants as const

pytestmark = pytest.mark.gpu


def _disc_medium(shape, seed, n_discs=6):
    rng = np.random.default_rng(seed)
    n = np.ones(shape)
    yy, xx = np.meshgrid(*[np.arange(s) for s in shape], indexing='ij')
    for _ in range(n_discs):
        c = rng.uniform(0.2, 0.8, 2) * np.asarray(shape)
        r = rng.uniform(0.05, 0.15) * min(shape)
        n[(yy - c[0]) ** 2 + (xx - c[1]) ** 2 < r ** 2] = rng.uniform(1.2, 1.5)
    return n


def _scalar2d(shape, seed, dtype, src_row=6, amp=1.0):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    grid = mm.Grid(shape=list(shape), step=125e-9)
    n = _disc_medium(shape, seed)
    J = np.zeros((1, *shape), dtype=np.complex128)
    J[0, src_row, :] = amp * np.exp(-((np.arange(shape[1]) - shape[1] / 2) / (shape[1] / 6)) ** 2)
    return mm.Solution(grid, vacuum_wavelength=500e-9, epsilon=n ** 2, current_density=J.astype(dtype),
                       bound=LinearBound(grid, thickness=8 * 125e-9, max_extinction_coefficient=0.2), dtype=dtype)


def _rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def _sequential(sols, tol, max_iter):
    for s in sols:
        s.solve(lambda s_: s_.residue > tol and s_.iteration < max_iter)


@pytest.mark.parametrize('dtype', [np.complex128, np.complex64])
@pytest.mark.parametrize('graph', [True, False])
def test_batch_of_media_matches_sequential(dtype, graph):
    from macromax_amd.batch import BatchSolution
    print(f'batch of media {np.dtype(dtype).name} graph={graph}', flush=True)
    tol, max_iter = 1e-5, 400
    shape = (64, 96)
    ref = [_scalar2d(shape, seed, dtype) for seed in range(5)]
    _sequential(ref, tol, max_iter)
    sols = [_scalar2d(shape, seed, dtype) for seed in range(5)]
    bt = BatchSolution(sols).solve(tol=tol, max_iter=max_iter, steps_per_sync=7, graph=graph)
    assert bt.done.all()
    ftol = 1e-10 if dtype == np.complex128 else 1e-5
    for b, (s, r) in enumerate(zip(sols, ref)):
        assert s.iteration == r.iteration, b
        assert bt.iteration[b] == r.iteration
        assert s.alpha == r.alpha
        assert _rel(s.E, r.E) < ftol, b
        np.testing.assert_allclose(s.residue, r.residue, rtol=1e-6 if dtype == np.complex64 else 1e-9)
    lat = bt.latency()
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)


def test_batch_of_sources_matches_sequential():
    from macromax_amd.batch import BatchSolution
    print('batch of right-hand sides', flush=True)
    tol, max_iter = 1e-5, 400
    shape = (96, 64)
    base = _scalar2d(shape, 11, np.complex64)
    rng = np.random.default_rng(4)
    srcs = []
    for b in range(6):
        S = np.zeros((1, *shape), dtype=np.complex64)
        theta = rng.uniform(-np.pi / 3, np.pi / 3)
        S[0, 10, :] = np.exp(1j * 2 * np.pi / 4 * np.sin(theta) * np.arange(shape[1]))
        srcs.append(S)
    bt = BatchSolution.from_sources(base, np.stack(srcs)).solve(tol=tol, max_iter=max_iter, steps_per_sync=5)
    E = bt.E
    for b in range(len(srcs)):
        r = _scalar2d(shape, 11, np.complex64)
        r.source_distribution = srcs[b]
        r.solve(lambda s_: s_.residue > tol and s_.iteration < max_iter)
        assert bt.iteration[b] == r.iteration, b
        assert _rel(E[b], r.E) < 1e-5, b
    # the same right-hand sides given as current densities J, S = -i omega mu_0 J
    J = np.stack(srcs) / (-1j * base.angular_frequency * const.mu_0)
    bj = BatchSolution.from_sources(base, current_densities=J).solve(tol=tol, max_iter=max_iter, steps_per_sync=5)
    assert np.array_equal(bj.iteration, bt.iteration)
    assert _rel(bj.E, E) < 1e-5


def test_batch_rejections_match_sequential():
    """Under-damped alpha: the device raises alpha_i per problem exactly as Solution.step does."""
    from macromax_amd.batch import BatchSolution
    print('batch with rejected updates', flush=True)
    tol, max_iter = 1e-6, 150
    shape = (48, 64)

    def make():
        out = []
        for seed in range(3):
            s = _scalar2d(shape, 20 + seed, np.complex128)
            s._Solution__update_operators(complex(s.alpha.real, s.alpha.imag * (0.02 + 0.01 * seed)))
            out.append(s)
        return out
    ref = make()
    _sequential(ref, tol, max_iter)
    sols = make()
    bt = BatchSolution(sols).solve(tol=tol, max_iter=max_iter, steps_per_sync=4)
    assert bt.rejected.sum() >= 1
    for b, (s, r) in enumerate(zip(sols, ref)):
        assert s.iteration == r.iteration
        assert s.alpha == r.alpha, b
        assert _rel(s.E, r.E) < 1e-10, b


def test_solve_many_uneven_sizes():
    from macromax_amd.batch import solve_many
    print('solve_many uneven sizes', flush=True)
    tol, max_iter = 1e-5, 500
    shapes = [(64, 64), (48, 48), (64, 64), (96, 96), (48, 48), (80, 64)]
    ref = [_scalar2d(sh, 30 + i, np.complex64) for i, sh in enumerate(shapes)]
    _sequential(ref, tol, max_iter)
    sols = [_scalar2d(sh, 30 + i, np.complex64) for i, sh in enumerate(shapes)]
    batches = solve_many(sols, tol=tol, max_iter=max_iter, steps_per_sync=8)
    assert sorted(b.nbatch for b in batches) == [1, 1, 2, 2]
    for b, (s, r) in enumerate(zip(sols, ref)):
        assert s.iteration == r.iteration, b
        assert _rel(s.E, r.E) < 1e-5, b


def test_batch_3d_anisotropic():
    """3D vectorial anisotropic problems: the single-component middle passes skip finished problems."""
    import macromax_amd as mm
    from macromax_amd.batch import BatchSolution
    from macromax_amd.bound import LinearBound
    print('batch 3d anisotropic', flush=True)
    tol = 1e-4

    def make(seed):
        rng = np.random.default_rng(seed)
        shape = (16, 32, 24)
        grid = mm.Grid(shape=list(shape), step=0.25e-6)
        a = rng.standard_normal((3, *shape))
        a /= np.linalg.norm(a, axis=0)
        ne = 1.6 + 0.1 * seed
        eps = 1.5 ** 2 * np.eye(3).reshape(3, 3, 1, 1, 1) + (ne ** 2 - 1.5 ** 2) * a[:, None] * a[None, :]
        J = np.zeros((3, *shape), dtype=np.complex64)
        J[0, :, :, 4] = 1.0
        J[1, :, :, 4] = 0.5j * (1 + seed)
        return mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J,
                           bound=LinearBound(grid, thickness=1e-6), dtype=np.complex64)
    seeds = (1, 2, 3)
    ref = [make(sd) for sd in seeds]
    _sequential(ref, tol, 300)
    sols = [make(sd) for sd in seeds]
    BatchSolution(sols).solve(tol=tol, max_iter=300, steps_per_sync=6)
    for b, (s, r) in enumerate(zip(sols, ref)):
        assert s.iteration == r.iteration, b
        assert _rel(s.E, r.E) < 1e-5, b


def test_batch_controller_matches_oracle():
    """The device-side accept/reject rule of mmx_batch_step (batch_control_kernel) against the oracle directly:
    per problem and per iteration the accept flag, alpha and the field of BornOracle.step (solver.py:1064-1098), on
    under-damped problems that reject updates."""
    from macromax_amd.batch import BatchSolution
    from oracle import born_oracle as ob
    print('batch controller against the oracle, 20 iterations', flush=True)
    shape = (48, 64)
    sols = []
    for seed in range(3):
        s = _scalar2d(shape, 40 + seed, np.complex128)
        s._Solution__update_operators(complex(s.alpha.real, s.alpha.imag * (0.02 + 0.01 * seed)))
        sols.append(s)
    oracles = []
    for s in sols:
        k = ob.k_vectors(shape, np.asarray(s.grid.step) * s.wavenumber)
        oracles.append(ob.BornOracle(s.chi_effective.cpu().numpy(), s.source_normalized.cpu().numpy(), s.alpha,
                                     s.beta, k, False))
    bt = BatchSolution(sols)
    prev_rej = np.zeros(3, dtype=np.int64)
    n_rej = 0
    for it in range(1, 21):
        bt.step()
        rej = bt.rejected
        alpha = bt.alpha
        E = bt.field_arrays().cpu().numpy()
        assert np.all(bt.iteration == it)
        for b, o in enumerate(oracles):
            acc = o.step()
            assert acc == (rej[b] == prev_rej[b]), (it, b)
            assert alpha[b] == pytest.approx(o.alpha, rel=1e-12), (it, b)
            assert _rel(E[b], o.E) < 1e-10, (it, b, _rel(E[b], o.E))
        n_rej += int(np.sum(rej - prev_rej))
        prev_rej = rej
    assert n_rej >= 2
    np.testing.assert_allclose(bt.residue, [o.relative_residue for o in oracles], rtol=1e-9)
