"""The batched and the adjoint workloads at their full sizes (the workloads of bench.py's batched and adjoint parts).

Batched (2D scalar complex64 disc media as in bench.disc_problem: 125 nm pixels, 500 nm wavelength, 32-pixel
LinearBound with kappa up to 0.2, Gaussian line source on row 40, residue 1e-6):
- one medium of every size {128, 192, 256, 384, 512}^2 in one solve_many call (one graph, one stream per size),
  10 iterations, each against the complex128 oracle: relative L2 <= 1e-5 per problem;
- ten media covering every size, solve_many to residue 1e-6 against sequential Solution.solve: the same iteration
  counts, fields within 1e-5 (the reference solves each medium with a Solution of its own);
- one 512^2 medium with 16 plane-wave right-hand sides (angles in [-60, 60] degrees), the batch's first 5 iterates
  of three of them against the oracle, and the whole batched solve against one Solution.solve per right-hand side
  (the reference's column loop, python/macromax/matrix.py:789-819).
Adjoint (96^3 isotropic vectorial, 48^3 design with eps uniform in [1, 2.25], y-polarised Gaussian source, loss
sum |E|^2 over a 6^3 box behind the design; bench.adjoint_problem):
- the complex64 gradient (forward and adjoint solves to residue 1e-6) against central differences of complex128
  forward solves (residue 1e-11, h = 1e-4) on 10 design voxels: within 1e-3 of the largest gradient entry; the
  complex128 gradient against the same differences within 1e-5;
- at 48^3 with a 24^3 design: the complex128 gradient against the oracle's reciprocity adjoint of its own converged
  field (oracle/born_oracle.py adjoint_gradient) within 1e-6 of the largest entry. The differentiable restatement
  the reference offers, python/examples/scattering_neural_network.py:206-318, is a scalar unrolled model; the
  oracle's adjoint is pinned against autograd through its own unrolled restatement on the CPU.
"""
import numpy as np
import pytest
import scipy.constants as const

from oracle import born_oracle as ob

pytestmark = pytest.mark.gpu

SIZES = (128, 192, 256, 384, 512)


def _rel(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def _oracle_of(sol):
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s = sol.source_normalized.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors(tuple(sol.grid.shape), np.asarray(sol.grid.step) * sol.wavenumber)
    return ob.BornOracle(chi, s, sol.alpha, sol.beta, k, sol.vectorial)


def test_disc_media_of_every_size_first_iterates_match_oracle():
    import bench
    from macromax_amd.batch import solve_many
    print('disc media 128..512^2 in one solve_many, 10 iterations against the oracle', flush=True)
    sols = [bench.disc_problem(n, 1000 + i) for i, n in enumerate(SIZES)]
    oracles = [_oracle_of(s) for s in sols]
    batches = solve_many(sols, tol=0.0, max_iter=10, steps_per_sync=5)
    assert len(batches) == len(SIZES)
    for n, s, o in zip(SIZES, sols, oracles):
        for _ in range(10):
            o.step()
        assert s.iteration == 10
        err = _rel(s.field_array.cpu().numpy(), o.E)
        assert err < 1e-5, (n, err)
        assert s.alpha == pytest.approx(o.alpha, rel=1e-6), n


def test_disc_media_solve_many_matches_sequential():
    import bench
    from macromax_amd.batch import solve_many
    print('ten disc media covering 128..512^2: solve_many against Solution.solve', flush=True)
    sizes = list(SIZES) * 2
    tol, max_iter = 1e-6, 10000
    seq = [bench.disc_problem(n, 2000 + i) for i, n in enumerate(sizes)]
    for s in seq:
        s.solve(lambda s_: s_.residue > tol and s_.iteration < max_iter)
    bat = [bench.disc_problem(n, 2000 + i) for i, n in enumerate(sizes)]
    batches = solve_many(bat, tol=tol, max_iter=max_iter, steps_per_sync=16)
    assert sorted(b.nbatch for b in batches) == [2] * len(SIZES)
    for n, b, s in zip(sizes, bat, seq):
        assert b.iteration == s.iteration, n
        assert s.residue <= tol and b.residue <= tol
        assert b.alpha == s.alpha
        assert _rel(b.E, s.E) < 1e-5, n
    lat = np.concatenate([bt.latency() for bt in batches])
    assert np.all(np.isfinite(lat)) and np.all(lat > 0)


def test_plane_wave_right_hand_sides_512():
    import bench
    from macromax_amd.batch import BatchSolution
    print('512^2 medium, 16 plane-wave right-hand sides: batch against the oracle and the column loop', flush=True)
    n, nrhs, tol = 512, 16, 1e-6
    src = bench.plane_wave_sources(n, nrhs)
    medium = bench.disc_problem(n, 0)
    # first iterates against the oracle
    bt = BatchSolution.from_sources(medium, current_densities=src)
    chi = medium.chi_effective.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors((n, n), np.asarray(medium.grid.step) * medium.wavenumber)
    picks = (0, 7, 15)
    oracles = []
    for j in picks:
        S = src[j].astype(np.complex128)[:, None] * (-1j * medium.angular_frequency * const.mu_0)
        oracles.append(ob.BornOracle(chi, ob.normalized_source(S, medium.alpha, medium.beta), medium.alpha,
                                     medium.beta, k, False))
    for it in range(1, 6):
        bt.step()
        E = bt.field_arrays().cpu().numpy()
        for j, o in zip(picks, oracles):
            o.step()
            err = _rel(E[j], o.E)
            assert err < 1e-5, (it, j, err)
    # the whole solve against one Solution per right-hand side
    bt = BatchSolution.from_sources(medium, current_densities=src).solve(tol=tol, max_iter=10000)
    assert bt.done.all()
    Eb = bt.E
    its = bt.iteration
    for j in range(nrhs):
        s = bench.disc_problem(n, 0, source=src[j])
        s.solve(lambda s_: s_.residue > tol and s_.iteration < 10000)
        assert its[j] == s.iteration, j
        assert _rel(Eb[j], s.E) < 1e-5, j


def _finite_differences(eps, grid, J, bound, loss_of, vox, h, alpha):
    from macromax_amd.adjoint import fields_at_fixed_alpha
    epsilons = []
    for v in vox:
        for sgn in (+1, -1):
            e = eps.copy()
            e[v] += sgn * h
            epsilons.append(e)
    E = fields_at_fixed_alpha(epsilons, grid, 1e-6, alpha, current_density=J.astype(np.complex128), bound=bound,
                              dtype=np.complex128, tol=1e-11, max_iter=30000)
    losses = [float(loss_of(E[i])) for i in range(len(epsilons))]
    return np.array([(losses[2 * i] - losses[2 * i + 1]) / (2 * h) for i in range(len(vox))])


def _gradient(eps, grid, J, bound, loss_of, dtype, tol):
    import torch
    from macromax_amd.adjoint import solve_field
    eps_t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    info = {}
    E = solve_field(eps_t, grid, 1e-6, current_density=J.astype(dtype), bound=bound, dtype=dtype, tol=tol,
                    max_iter=30000, info=info)
    loss = loss_of(E)
    loss.backward()
    assert info['forward_residue'] <= tol and info['adjoint_residue'] <= tol, info
    return eps_t.grad.numpy(), info


def test_adjoint_96_gradient_against_complex128_finite_differences():
    import bench
    print('96^3 adjoint (48^3 design): c64 and c128 gradients against c128 finite differences', flush=True)
    grid, eps, J, bound, box, dsl, loss_of = bench.adjoint_problem(96, 48)
    g64, info = _gradient(eps, grid, J, bound, loss_of, np.complex64, 1e-6)
    g128, _ = _gradient(eps, grid, J, bound, loss_of, np.complex128, 1e-11)
    d0 = dsl[0].start
    rng = np.random.default_rng(6)
    vox = [tuple(int(d0 + rng.integers(0, 48)) for _ in range(3)) for _ in range(10)]
    fd = _finite_differences(eps, grid, J, bound, loss_of, vox, 1e-4, info['alpha'])
    scale = np.abs(g128[dsl]).max()
    assert np.abs(g64[dsl]).max() == pytest.approx(scale, rel=1e-2)
    for i, v in enumerate(vox):
        assert abs(g64[v] - fd[i]) <= 1e-3 * scale, (v, g64[v], fd[i], scale)
        assert abs(g128[v] - fd[i]) <= 1e-5 * scale, (v, g128[v], fd[i], scale)


def test_adjoint_48_complex128_gradient_matches_oracle():
    import bench
    import macromax_amd as mm
    print('48^3 adjoint (24^3 design): c128 gradient against the oracle adjoint', flush=True)
    grid, eps, J, bound, box, dsl, loss_of = bench.adjoint_problem(48, 24, bound_px=8, source_plane=6)
    tol = 1e-10
    grad, info = _gradient(eps, grid, J, bound, loss_of, np.complex128, tol)
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=eps, current_density=J.astype(np.complex128),
                      bound=bound, dtype=np.complex128)
    chi = sol.chi_effective.cpu().numpy()
    s = sol.source_normalized.cpu().numpy()
    k0 = sol.wavenumber
    k = ob.k_vectors(tuple(grid.shape), np.asarray(grid.step) * k0)
    o = ob.solve_to(chi, s, sol.alpha, sol.beta, k, True, tol)
    m = np.zeros(o.E.shape)
    m[(slice(None), slice(None)) + box] = 1.0
    g_int = 2.0 * m * o.E / k0 ** 4
    og, _ = ob.adjoint_gradient(chi, sol.alpha, sol.beta, k, True, o.E, g_int, tol)
    og = og.real[0, 0]
    scale = np.abs(og).max()
    np.testing.assert_allclose(grad, og, rtol=0, atol=1e-6 * scale)
