"""Benchmark: Born-series iterations of the 256^3 anisotropic vectorial complex64 solve on one MI355X.

Workload (the metric's configuration): 256^3 grid, 250 nm pitch, vacuum wavelength 1 um; permittivity per voxel
R diag(no^2, no^2, ne^2) R^T (no = 1.50, ne = 1.65) with seeded random optic axes constant within 2000 Voronoi
grains; an isotropic n = 1.33 slab for z in [64, 96); circularly polarised current sheet on z = 20; 16-pixel
LinearBound on z (periodic x, y); complex64 fields, FP64 norms.

A "step" is one Born iteration (Solution.step: the fused HIP kernel sequence plus the accept/reject decision).
`value` is voxel-iterations/s with all inputs resident in HBM. The JSON line also carries the HBM roofline of the
iteration: `achieved` = the bytes the fused four-pass schedule must move per voxel-iteration (268 B with the
uniaxial chi storage of this medium) x voxels / the iteration's HIP-event time; `traffic` is the HBM bytes per
iteration measured with rocprofv3 counters (profiles/born_iterate_256cubed_aniso.md); the reference operator
sequence's 480 B per voxel-iteration gives `alg_equivalent_frac`, which can exceed 1 because the schedule moves
fewer bytes. The CPU baseline is the NumPy complex128 oracle iteration on this host (numpy.fft, one core; min of 3
trials of 3 iterations), with a labelled scipy.fft multi-threaded variant beside it. Also reported: time to
residual 1e-5, time to 1e-4 relative L2 against a complex128 solve (FP64 refinement), the same iteration with the
symmetric and full chi storage, batched solves, the adjoint gradient, magnetic media and the other configurations.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--grid N]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = 'voxel-iter/s at 256³ vectorial c64, %HBM peak; batched+adjoint vs plain'


def build_anisotropic_medium(n, device, seed=2, n_seeds=2000, no=1.50, ne=1.65, slab=(64, 96), n_slab=1.33):
    """Permittivity [3,3,n,n,n] complex64 on the device: Voronoi grains with random optic axes + isotropic slab."""
    import torch
    rng = np.random.default_rng(seed)
    seeds = rng.uniform(0, n, size=(n_seeds, 3))
    axes = rng.standard_normal((n_seeds, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    sd = torch.as_tensor(seeds, dtype=torch.float32, device=device)
    s2 = (sd ** 2).sum(1)
    idx = torch.empty(n ** 3, dtype=torch.int64, device=device)
    coords = torch.arange(n, device=device, dtype=torch.float32)
    chunk = n * n
    for i0 in range(n):  # one plane of constant axis-0 index at a time
        yy, zz = torch.meshgrid(coords, coords, indexing='ij')
        pts = torch.stack([torch.full_like(yy, float(i0)), yy, zz], dim=-1).reshape(-1, 3)
        d2 = (pts ** 2).sum(1, keepdim=True) - 2.0 * pts @ sd.T + s2[None, :]
        idx[i0 * chunk:(i0 + 1) * chunk] = torch.argmin(d2, dim=1)
    a = torch.as_tensor(axes, dtype=torch.float64, device=device)[idx]  # [V,3]
    eps = no ** 2 * torch.eye(3, dtype=torch.float64, device=device)[None] + (ne ** 2 - no ** 2) * a[:, :, None] * a[:, None, :]
    eps = eps.reshape(n, n, n, 3, 3)
    eps[:, :, slab[0]:slab[1]] = n_slab ** 2 * torch.eye(3, dtype=torch.float64, device=device)
    eps = eps.permute(3, 4, 0, 1, 2).to(torch.complex64).contiguous()
    return eps


def setup_solution(n, device, dtype=np.complex64, chi_storage='auto'):
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    wavelength = 1e-6
    pitch = 250e-9
    grid = mm.Grid(shape=[n, n, n], step=pitch)
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    eps = build_anisotropic_medium(n, device).to(tdt)
    J = torch.zeros((3, n, n, n), dtype=tdt, device=device)
    J[0, :, :, 20] = 1.0 / math.sqrt(2)
    J[1, :, :, 20] = 1.0j / math.sqrt(2)
    bound = LinearBound(grid, thickness=[[0, 0], [0, 0], [16 * pitch, 16 * pitch]])
    t0 = time.perf_counter()
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, epsilon=eps, current_density=J, bound=bound,
                      dtype=dtype, chi_storage=chi_storage)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    del eps, J
    return sol, setup_s


def cpu_baseline(sol, n, trials=3, k=3):
    """Complex128 NumPy oracle iterations of the same problem on this host (the reference NumPy backend's operator
    sequence with numpy.fft, which runs on one core): one warm-up iteration, then the minimum over ``trials`` trials
    of ``k`` iterations each (about 30 s of CPU work at 256^3), reported per iteration. Beside it, labelled
    "tuned", the same iteration with scipy.fft on all cores of this process's affinity (the FFTs only; the
    susceptibility products stay single-threaded NumPy), one warm-up and ``k`` iterations."""
    from oracle import born_oracle as ob
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s = sol.source_normalized.cpu().numpy().astype(np.complex128)
    E = sol.field_array.cpu().numpy().astype(np.complex128)
    k_vec = ob.k_vectors((n, n, n), np.asarray(sol.grid.step) * sol.wavenumber)
    o = ob.BornOracle(chi, s, sol.alpha, sol.beta, k_vec, True, E0=E)
    t0 = time.perf_counter()
    o.step()
    total = time.perf_counter() - t0
    best = math.inf
    for _ in range(trials):
        t0 = time.perf_counter()
        for _ in range(k):
            o.step()
        dt = (time.perf_counter() - t0) / k
        total += dt * k
        best = min(best, dt)
    cores = len(os.sched_getaffinity(0))
    ot = ob.BornOracle(chi, s, sol.alpha, sol.beta, k_vec, True, E0=E, workers=cores)
    ot.step()
    t0 = time.perf_counter()
    for _ in range(k):
        ot.step()
    dt_tuned = (time.perf_counter() - t0) / k
    return {'value': n ** 3 / best, 'unit': 'voxel-iter/s', 'cores': 1, 'kind': 'port',
            'sample': f'{n}^3 anisotropic problem of the benchmark, complex128 oracle (oracle/born_oracle.py, numpy.fft): '
                      f'1 warm-up iteration, then the minimum over {trials} trials of {k} iterations; '
                      f'{total:.1f} s of CPU work', 'seconds_per_iteration': best,
            'tuned': {'value': n ** 3 / dt_tuned, 'unit': 'voxel-iter/s', 'cores': cores,
                      'seconds_per_iteration': dt_tuned,
                      'what': f'the same iteration with scipy.fft(workers={cores}); not the reference\'s backend'},
            'os_cpu_count': os.cpu_count(), 'affinity_cores': cores}


def must_move_bytes(ndim, ncomp, elem, chi_kind, s_fraction=1.0):
    """HBM bytes per voxel-iteration that the fused pass schedule must move (mmx_born_iterate): 2 ndim - 3 passes
    that read and write the work field (C e each way), then the update pass, which reads the work field, E and chi
    (and s on its support), and writes E' and the next work field. chi: e (isotropic), 9 e (full 3x3), 6 e
    (symmetric), 3.5 e (uniaxial: three complex planes and one real plane)."""
    ce = ncomp * elem
    b_chi = {1: 1.0, 2: 9.0, 3: 6.0, 4: 3.5}[chi_kind] * elem
    return (2 * ndim - 3) * 2 * ce + 4 * ce + b_chi + s_fraction * ce


def algorithmic_bytes(ndim, ncomp, elem, chi_elems):
    """The reference operator sequence's bytes per voxel-iteration: 4 d C e + 2 B_chi + B_s + C e (d forward and d
    inverse axis passes over the field, chi read twice, s and E once), B_chi = chi_elems x e."""
    return 4 * ndim * ncomp * elem + 2 * chi_elems * elem + 2 * ncomp * elem


def time_iterations(sol, steps, warmup):
    """Per-iteration time of Solution.step: wall clock and HIP events on the launch stream."""
    import torch
    for _ in range(warmup):
        sol.step()
    dev = sol.field_array.device
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        sol.step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall / steps, ev0.elapsed_time(ev1) * 1e-3 / steps


def storage_benchmark(n, dev, steps=50, warmup=5, log=lambda m: None):
    """The benchmark medium with chi stored as the full 3x3 (72 B/voxel in complex64) and as the 6 symmetric entries
    (48 B) instead of the uniaxial form (28 B): per-iteration time and bandwidth of the same iteration."""
    import torch
    out = {}
    for storage, kind in (('symmetric', 3), ('full', 2)):
        sol, _ = setup_solution(n, dev, chi_storage=storage)
        assert sol.chi_kind == kind
        sol.solve(callback=lambda s_: s_.iteration < 20)
        wall, ev = time_iterations(sol, steps, warmup)
        mm_b = must_move_bytes(3, 3, 8, kind, 1.0 / n)
        out[storage] = {'chi_kind': kind, 'chi_bytes_per_voxel': {3: 48, 2: 72}[kind], 'ms_per_iteration': round(wall * 1e3, 4),
                        'voxel_iter_per_s': round(n ** 3 / wall, 1), 'must_move_bytes_per_voxel_iter': round(mm_b, 1),
                        'GBps_must_move': round(mm_b * n ** 3 / ev / 1e9, 1),
                        'frac_must_move': round(mm_b * n ** 3 / ev / 1e9 / HBM_PEAK_GBS, 4)}
        log(f'chi storage {storage}: {wall * 1e3:.3f} ms per iteration')
        del sol
        torch.cuda.empty_cache()
    return out


# ---------------------------------------------------------------------------------------------------------------
# Batched form: 2D scalar disc-scatterer problems (125 nm pixels, 500 nm wavelength, 32-pixel LinearBound with
# kappa up to 0.2, line source on row 40), complex64, stopped at residual 1e-6.
def disc_medium(n, seed, n_discs=40, pitch=125e-9, bound_px=32):
    """Refractive index n x n: background 1 with n_discs discs (radius 0.5-2 um, index 1.2-1.5, later discs on top)
    whose centres lie in the interior outside the absorbing layer."""
    rng = np.random.default_rng(seed)
    idx = np.ones((n, n))
    x = np.arange(n) * pitch
    xx, yy = np.meshgrid(x, x, indexing='ij')
    lo, hi = bound_px * pitch, (n - bound_px) * pitch
    for _ in range(n_discs):
        cx, cy = rng.uniform(lo, hi, 2)
        r = rng.uniform(0.5e-6, 2e-6)
        idx[(xx - cx) ** 2 + (yy - cy) ** 2 < r ** 2] = rng.uniform(1.2, 1.5)
    return idx


def disc_problem(n, seed, source=None, pitch=125e-9, wavelength=500e-9, bound_px=32):
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    grid = mm.Grid(shape=[n, n], step=pitch)
    idx = disc_medium(n, seed, pitch=pitch, bound_px=bound_px)
    if source is None:
        y = (np.arange(n) - n / 2) * pitch
        source = np.zeros((1, n, n), dtype=np.complex64)
        source[0, 40, :] = np.exp(-(y / 3e-6) ** 2)
    bound = LinearBound(grid, thickness=bound_px * pitch, max_extinction_coefficient=0.2)
    return mm.Solution(grid, vacuum_wavelength=wavelength, epsilon=idx ** 2, current_density=source, bound=bound,
                       dtype=np.complex64)


def plane_wave_sources(n, nrhs, pitch=125e-9, wavelength=500e-9):
    k0 = 2 * np.pi / wavelength
    y = (np.arange(n) - n / 2) * pitch
    src = np.zeros((nrhs, 1, n, n), dtype=np.complex64)
    for j, th in enumerate(np.linspace(-np.pi / 3, np.pi / 3, nrhs)):
        src[j, 0, 40, :] = np.exp(1j * k0 * np.sin(th) * y)
    return src


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def _rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


def batched_benchmark(tol=1e-6, max_iter=10000, n_problems=64, nrhs=256, seq_rhs=256, steps_per_sync=16,
                      log=lambda m: None):
    """Part A: n_problems independent media with sizes drawn from {128,192,256,384,512}^2; Part B: one 512^2
    medium with nrhs plane-wave line sources at angles in [-60, 60] degrees. Batched (solve_many /
    BatchSolution, graph replays) against sequential Solution.solve, solves/s and per-problem latency."""
    import torch
    from macromax_amd.batch import BatchSolution, solve_many
    cb = lambda s_: s_.residue > tol and s_.iteration < max_iter
    out = {}
    # Part A
    sizes = np.random.default_rng(3).choice([128, 192, 256, 384, 512], n_problems)
    seq = [disc_problem(int(n), 1000 + i) for i, n in enumerate(sizes)]
    seq_t = []
    for s in seq:
        _, dt = _timed(lambda: s.solve(cb))
        seq_t.append(dt)
    log(f'part A sequential: {sum(seq_t):.2f} s')
    bat = [disc_problem(int(n), 1000 + i) for i, n in enumerate(sizes)]
    # autotune every size group for its batch (timed separately; cached per size)
    from macromax_amd.batch import BatchSolution as _BS, _group_key
    groups = {}
    for sol_ in bat:
        groups.setdefault(_group_key(sol_), []).append(sol_)
    tune_s = 0.0
    for g_ in groups.values():
        _, dt = _timed(lambda: _BS(g_).tune())
        tune_s += dt
    log(f'part A autotune: {tune_s:.2f} s')
    batches, bt_total = _timed(lambda: solve_many(bat, tol=tol, max_iter=max_iter, steps_per_sync=steps_per_sync))
    log(f'part A batched: {bt_total:.2f} s')
    lat = np.concatenate([b.latency() for b in batches])
    errs = [_rel_l2(b.E, s.E) for b, s in zip(bat, seq)]
    its = np.array([s.iteration for s in bat])
    out['part_a'] = {
        'problems': int(n_problems), 'sizes': {str(k): int(v) for k, v in zip(*np.unique(sizes, return_counts=True))},
        'solves_per_s': round(n_problems / bt_total, 2), 'sequential_solves_per_s': round(n_problems / sum(seq_t), 2),
        'speedup_vs_sequential': round(sum(seq_t) / bt_total, 2),
        'latency_p50_s': round(float(np.percentile(lat, 50)), 4), 'latency_p99_s': round(float(np.percentile(lat, 99)), 4),
        'sequential_latency_p50_s': round(float(np.percentile(seq_t, 50)), 4),
        'sequential_latency_p99_s': round(float(np.percentile(seq_t, 99)), 4),
        # the same 64 problems arriving together and solved one after the other: completion time of each
        'sequential_completion_p50_s': round(float(np.percentile(np.cumsum(seq_t), 50)), 4),
        'sequential_completion_p99_s': round(float(np.percentile(np.cumsum(seq_t), 99)), 4),
        'latency': 'latency_p*: completion time of each problem from the start of solve_many (all problems arrive '
                   'together); sequential_latency_p*: duration of each Solution.solve on its own; '
                   'sequential_completion_p*: completion time of each problem when the same set is solved one '
                   'after the other',
        'iterations_median': int(np.median(its)), 'iterations_max': int(its.max()),
        'autotune_s': round(tune_s, 2),
        'iterations_equal_to_sequential': bool(all(a.iteration == b.iteration for a, b in zip(bat, seq))),
        'max_rel_l2_vs_sequential': float(max(errs)),
        'timed': 'solve_many end to end (batch set-up, graph capture, replays, write-back); sequential: '
                 'Solution.solve per problem; medium set-up excluded from both'}
    del seq, bat, batches
    torch.cuda.empty_cache()
    # Part B
    n = 512
    src = plane_wave_sources(n, nrhs)
    medium = disc_problem(n, 0)
    bt = BatchSolution.from_sources(medium, current_densities=src)
    _, tune_b = _timed(lambda: bt.tune())
    bt, bt_total = _timed(lambda: BatchSolution.from_sources(medium, current_densities=src).solve(
        tol=tol, max_iter=max_iter, steps_per_sync=steps_per_sync))
    log(f'part B batched: {bt_total:.2f} s')
    Eb = bt.E
    lat = bt.latency()
    its_b = bt.iteration
    seq_t, errs, its_eq = [], [], []
    for j in range(min(seq_rhs, nrhs)):
        s = disc_problem(n, 0, source=src[j])
        _, dt = _timed(lambda: s.solve(cb))
        seq_t.append(dt)
        errs.append(_rel_l2(Eb[j], s.E))
        its_eq.append(s.iteration == its_b[j])
    log(f'part B sequential ({len(seq_t)} of {nrhs}): {sum(seq_t):.2f} s')
    seq_rate = len(seq_t) / sum(seq_t)
    out['part_b'] = {
        'grid': [n, n], 'rhs': int(nrhs), 'solves_per_s': round(nrhs / bt_total, 2),
        'sequential_solves_per_s': round(seq_rate, 2), 'speedup_vs_sequential': round(nrhs / bt_total / seq_rate, 2),
        'sequential_measured_on': f'the first {len(seq_t)} of {nrhs} right-hand sides',
        'latency_p50_s': round(float(np.percentile(lat, 50)), 4), 'latency_p99_s': round(float(np.percentile(lat, 99)), 4),
        'sequential_latency_p50_s': round(float(np.percentile(seq_t, 50)), 4),
        'sequential_latency_p99_s': round(float(np.percentile(seq_t, 99)), 4),
        'sequential_completion_p50_s': round(float(np.percentile(np.cumsum(seq_t), 50)), 4),
        'sequential_completion_p99_s': round(float(np.percentile(np.cumsum(seq_t), 99)), 4),
        'iterations_median': int(np.median(its_b)), 'iterations_max': int(its_b.max()), 'autotune_s': round(tune_b, 2),
        'iterations_equal_to_sequential': bool(all(its_eq)), 'max_rel_l2_vs_sequential': float(max(errs)),
        'timed': 'BatchSolution.from_sources(...).solve end to end; sequential: Solution.solve per right-hand side'}
    del bt, Eb
    torch.cuda.empty_cache()
    out['part_c'] = smatrix_benchmark(log=log)
    return out


def smatrix_benchmark(n0=512, nt=128, tol=1e-6, log=lambda m: None):
    """Part C: the scattering matrix of a 2D scalar multilayer thin film (seven one-pixel layers of index 1.5, the
    reference's unitarity case, python/tests/test_matrix.py:249-262) on an n0 x nt grid at 125 nm, 500 nm
    wavelength, 20-wavelength LinearBound along the propagation axis, periodic across it; complex128 and residue 1e-6
    for every column (the reference's ScatteringMatrix defaults). A layer with structure across the propagation axis
    would trap light travelling along it (nothing absorbs across), so that its columns do not reach 1e-6 within 10^4
    iterations. ScatteringMatrix.__array__ with one batched solve of all columns against the reference's
    column-by-column Solution loop (the same class with a callback), and the complex128 NumPy oracle iteration on the
    host, extrapolated to the columns' total iteration count."""
    import macromax_amd as mm
    from macromax_amd import matrix
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    wavelength = 500e-9
    grid = mm.Grid([n0, nt], wavelength / 4)
    thickness = np.zeros([2, 1])
    thickness[0, 0] = 20 * wavelength
    bound = LinearBound(grid, thickness=thickness, max_extinction_coefficient=0.25)
    n = np.ones(grid.shape)
    for off in (-11, -10, -9, -8, -1, 0, 2):
        n[n0 // 2 + off] = 1.5
    kw = dict(vectorial=False, vacuum_wavelength=wavelength, refractive_index=n, bound=bound, dtype=np.complex128)
    m_b = matrix.ScatteringMatrix(grid, tol=tol, **kw)
    arr_b, t_b = _timed(lambda: np.asarray(m_b))
    log(f'part C batched scattering matrix {arr_b.shape}: {t_b:.2f} s')
    seq_its = [0]

    def seq_callback(s_):
        seq_its[0] += 1
        return s_.iteration < 1e4 and s_.residue > tol
    m_s = matrix.ScatteringMatrix(grid, callback=seq_callback, **kw)
    arr_s, t_s = _timed(lambda: np.asarray(m_s))
    log(f'part C column-by-column: {t_s:.2f} s')
    sv = np.linalg.svd(arr_b, compute_uv=False)
    # iterations of every column (one more batched solve of the basis, for the count only; not timed)
    from macromax_amd.batch import BatchSolution
    eye = np.eye(arr_b.shape[1], dtype=np.complex128)
    srcs = m_b._sources(m_b._vectors(eye))
    bt = BatchSolution.from_sources(m_b.solution, sources=srcs)
    bt.solve(tol=tol, max_iter=10000)
    total_its = int(np.sum(bt.iteration))
    sol = m_b.solution
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s0 = srcs[0].cpu().numpy().astype(np.complex128)[:, None] * (1.0j / sol.alpha.imag / sol.beta)
    k = ob.k_vectors((n0, nt), np.asarray(grid.step) * sol.wavenumber)
    o = ob.BornOracle(chi, s0, sol.alpha, sol.beta, k, False)
    o.step()
    t0 = time.perf_counter()
    for _ in range(5):
        o.step()
    t_cpu_it = (time.perf_counter() - t0) / 5
    return {'grid': [n0, nt], 'columns': int(arr_b.shape[1]), 'dtype': 'c128', 'residue': tol,
            'medium': 'seven one-pixel layers of index 1.5 around the centre',
            'seconds': round(t_b, 3), 'sequential_seconds': round(t_s, 3), 'speedup_vs_sequential': round(t_s / t_b, 2),
            'max_rel_l2_vs_sequential': _rel_l2(arr_b, arr_s),
            'singular_values_min_max': [round(float(sv.min()), 4), round(float(sv.max()), 4)],
            'column_iterations_total': total_its, 'column_iterations_max': int(np.max(bt.iteration)),
            'sequential_iterations_total': int(seq_its[0]),
            'cpu_oracle_seconds_per_iteration': round(t_cpu_it, 5),
            'cpu_oracle_extrapolated_seconds': round(t_cpu_it * total_its, 1),
            'timed': 'ScatteringMatrix.__array__ end to end (sources, batched solve, detection); sequential: the '
                     'same class with a callback, one Solution.solve per column; medium set-up excluded; CPU: '
                     'one complex128 oracle iteration (numpy.fft, 1 core) x the columns\' total iterations'}


# ---------------------------------------------------------------------------------------------------------------
# The other single-problem configurations: 256^2 scalar complex128 disc medium to residue 1e-6 and 128^3 isotropic
# vectorial complex64 (Gaussian-smoothed random index in [1.0, 1.6], seed 1) to residue 1e-5.
def configs_benchmark(log=lambda m: None):
    import scipy.ndimage
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    out = {}
    # 256^2 scalar complex128
    n, pitch = 256, 125e-9
    grid = mm.Grid(shape=[n, n], step=pitch)
    idx = disc_medium(n, 0, pitch=pitch)
    y = (np.arange(n) - n / 2) * pitch
    J = np.zeros((1, n, n), dtype=np.complex128)
    J[0, 40, :] = np.exp(-(y / 3e-6) ** 2)
    sol = mm.Solution(grid, vacuum_wavelength=500e-9, epsilon=idx ** 2, current_density=J,
                      bound=LinearBound(grid, thickness=32 * pitch, max_extinction_coefficient=0.2),
                      dtype=np.complex128)
    _, t = _timed(lambda: sol.solve(lambda s_: s_.residue > 1e-6 and s_.iteration < 10000))
    rec = {'grid': [n, n], 'dtype': 'c128', 'residue': float(sol.residue), 'iterations': int(sol.iteration),
           'solve_s': round(t, 3), 'ms_per_iteration': round(1e3 * t / sol.iteration, 4),
           'voxel_iter_per_s': round(n * n * sol.iteration / t, 1)}
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    sv = sol.source_normalized.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors((n, n), np.asarray(grid.step) * sol.wavenumber)
    o = ob.BornOracle(chi, sv, sol.alpha, sol.beta, k, False)
    o.step()
    t0 = time.perf_counter()
    for _ in range(5):
        o.step()
    rec['cpu_oracle_seconds_per_iteration'] = round((time.perf_counter() - t0) / 5, 5)
    out['disc_256x256_scalar_c128'] = rec
    log(f'configs: 256^2 c128 {sol.iteration} iterations, {t:.3f} s')
    del sol
    # 128^3 isotropic vectorial complex64
    n, pitch = 128, 250e-9
    grid = mm.Grid(shape=[n, n, n], step=pitch)
    noise = np.random.default_rng(1).standard_normal((n, n, n))
    sm = scipy.ndimage.gaussian_filter(noise, 2.0, mode='wrap')
    idx = 1.0 + 0.6 * (sm - sm.min()) / (sm.max() - sm.min())
    x = (np.arange(n) - n / 2) * pitch
    J = np.zeros((3, n, n, n), dtype=np.complex64)
    J[0, :, :, 20] = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2e-6) ** 2)
    sol = mm.Solution(grid, vacuum_wavelength=1e-6, epsilon=idx ** 2, current_density=J,
                      bound=LinearBound(grid, thickness=16 * pitch), dtype=np.complex64)
    _, t = _timed(lambda: sol.solve(lambda s_: s_.residue > 1e-5 and s_.iteration < 10000))
    rec = {'grid': [n, n, n], 'dtype': 'c64', 'residue': float(sol.residue), 'iterations': int(sol.iteration),
           'solve_s': round(t, 3), 'ms_per_iteration': round(1e3 * t / sol.iteration, 4),
           'voxel_iter_per_s': round(n ** 3 * sol.iteration / t, 1)}
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    sv = sol.source_normalized.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors((n, n, n), np.asarray(grid.step) * sol.wavenumber)
    o = ob.BornOracle(chi, sv, sol.alpha, sol.beta, k, True)
    t0 = time.perf_counter()
    o.step()
    rec['cpu_oracle_seconds_per_iteration'] = round(time.perf_counter() - t0, 3)
    out['smooth_128cubed_iso_c64'] = rec
    log(f'configs: 128^3 c64 {sol.iteration} iterations, {t:.3f} s')
    del sol
    out['timed'] = 'Solution.solve to the residue (set-up excluded); CPU: complex128 oracle iteration, numpy.fft, 1 core'
    return out


# ---------------------------------------------------------------------------------------------------------------
# Magnetic medium: the reference's negative-index plate (tests/test_solution.py:360-398), 128^2 vectorial complex64,
# solved to residue 1e-3 as there; scaled to 512^2 (same physical layout, 4x the pitch count per axis) for timing.
def magnetic_benchmark(n=512, log=lambda m: None):
    import scipy.constants as const
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    from oracle import born_oracle as ob
    wavelength = 500e-9
    k0 = 2 * np.pi / wavelength
    out = {}
    for size in (128, n):
        pitch = wavelength / 4 * 128 / size
        grid = mm.Grid([size, size], pitch)
        ang = np.pi / 6
        rot = np.array([[np.cos(ang), -np.sin(ang), 0], [np.sin(ang), np.cos(ang), 0], [0, 0, 1]])
        k_inc = rot * k0 @ np.array([1, 0, 0])
        pol = (rot @ np.array([0, 1, 1j]) / np.sqrt(2))[:, np.newaxis, np.newaxis]
        J = np.exp(1j * (k_inc[0] * grid[0] + k_inc[1] * grid[1]))
        src = grid.shape[0] - int(2e-6 / grid.step[0])
        J[:src, :] = 0
        J[src + 1:, :] = 0
        J = J * np.exp(-0.5 * ((grid[1] - grid[1].ravel()[grid.shape[0] // 3]) / 2.5e-6) ** 2)
        J = J[np.newaxis, ...] * pol
        nidx = 1 + (-1.5 - 1) * np.ones(grid[1].shape) * (np.abs(grid[0]) < 2.5e-6)
        bound = LinearBound(grid, thickness=2e-6, max_extinction_coefficient=0.50)
        sol = mm.Solution(grid, vacuum_wavelength=wavelength, current_density=J, refractive_index=nidx, bound=bound,
                          dtype=np.complex64)
        _, t = _timed(lambda: sol.solve(lambda s_: s_.iteration < 1e4 and s_.residue > 1e-3))
        E = sol.E
        rec = {'grid': [size, size], 'iterations': int(sol.iteration), 'residue': float(sol.residue),
               'solve_s': round(t, 3), 'ms_per_iteration': round(1e3 * t / sol.iteration, 3)}
        if size == 128:
            rec['front_norm'] = float(np.linalg.norm(E[:, src + 1:src + 8, :].ravel()))
            rec['back_norm'] = float(np.linalg.norm(E[:, -src - 8:-src - 1, :].ravel()))
            rec['reference_norms'] = {'front': 0.0002582, 'back': 0.0001864}
        log(f'magnetic {size}^2: {sol.iteration} iterations, {t:.2f} s')
        # one complex128 oracle iteration on the host, same alpha and beta
        sign = 1 - 2 * (nidx < 0)
        eps = (nidx ** 2 * sign)[None, None] + bound.electric_susceptibility
        _, _, ops, _, _ = ob.magnetic_setup((size, size), pitch, k0, eps, 0.0, 0.0, sign[None, None].astype(complex),
                                            alpha_beta=(sol.alpha, sol.beta))
        S = J.reshape(3, 1, size, size) * (-1j * k0 * const.c * const.mu_0)
        kv = ob.k_vectors((size, size), np.array([pitch, pitch]) * k0)
        o = ob.MagneticBornOracle(ops, S * (1j / sol.alpha.imag / sol.beta), sol.alpha, sol.beta, kv)
        o.step()
        t0 = time.perf_counter()
        o.step()
        rec['cpu_oracle_seconds_per_iteration'] = round(time.perf_counter() - t0, 4)
        rec['cpu_oracle_extrapolated_s'] = round(rec['cpu_oracle_seconds_per_iteration'] * sol.iteration, 1)
        out[f'{size}'] = rec
        del sol
    out['timed'] = ('Solution.solve to residue 1e-3 (set-up excluded); one magnetic iteration is mmx_magnetic_iterate '
                    '(fused passes: the k-space products inside the passes along the first axis, the voxel-wise '
                    'products inside those along the last; 8 kernels per 2D iteration), the next iteration '
                    'launched before the norms of the current one are read; CPU: one '
                    'complex128 oracle iteration (numpy.fft, 1 core) x the GPU iteration count')
    return out


# ---------------------------------------------------------------------------------------------------------------
# Adjoint gradient: 96^3 isotropic vectorial, complex64; design cube 48^3 (eps uniform in [1, 2.25], seed 5),
# y-polarised Gaussian current on plane z = 12, loss = sum |E|^2 over a 6^3 box centred 6 pixels behind the design.
def adjoint_problem(n=96, design=48, bound_px=16, source_plane=12, seed=5):
    """The adjoint workload: n^3 isotropic grid at 250 nm (1 um wavelength), a centred design cube of ``design``^3
    voxels with eps uniform in [1, 2.25] (seeded), eps = 1 elsewhere, ``bound_px``-pixel LinearBound on all faces,
    y-polarised Gaussian current (waist 2 um) on the plane z = ``source_plane``; the loss is sum |E|^2 over a 6^3
    box centred 6 pixels behind the design region. Returns (grid, eps, J, bound, box, design slices, loss)."""
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    pitch = 250e-9
    grid = mm.Grid(shape=[n, n, n], step=pitch)
    bound = LinearBound(grid, thickness=bound_px * pitch)
    eps = np.ones((n, n, n))
    d0 = (n - design) // 2
    dsl = (slice(d0, d0 + design),) * 3
    eps[dsl] = np.random.default_rng(seed).uniform(1.0, 2.25, (design,) * 3)
    x = (np.arange(n) - n / 2) * pitch
    J = np.zeros((3, n, n, n), dtype=np.complex64)
    J[1, :, :, source_plane] = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / (2e-6) ** 2)
    zc = d0 + design + 6
    box = (slice(n // 2 - 3, n // 2 + 3), slice(n // 2 - 3, n // 2 + 3), slice(zc - 3, zc + 3))

    def loss_of(E):
        return torch.sum(torch.abs(E[(slice(None),) + box]) ** 2)
    return grid, eps, J, bound, box, dsl, loss_of


def adjoint_benchmark(n=96, design=48, tol=1e-6, max_iter=10000, n_fd=20, h=1e-4, log=lambda m: None):
    import torch
    import macromax_amd as mm
    from macromax_amd.adjoint import solve_field, fields_at_fixed_alpha
    wavelength = 1e-6
    grid, eps, J, bound, box, dsl, loss_of = adjoint_problem(n, design)
    d0 = dsl[0].start

    def gradient():
        eps_t = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
        info = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        E = solve_field(eps_t, grid, wavelength, current_density=J, bound=bound, tol=tol, max_iter=max_iter,
                        info=info)
        loss = loss_of(E)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        t0 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        return eps_t.grad.numpy(), info, t_fwd, time.perf_counter() - t0

    _, _, t_fwd0, t_bwd0 = gradient()  # first call in the process: includes one-time start-up costs
    grad, info, t_fwd, t_bwd = gradient()
    log(f'adjoint: forward {t_fwd:.3f} s ({info["forward_iterations"]} it), backward {t_bwd:.3f} s '
        f'({info["adjoint_iterations"]} it)')
    # forward solve alone (set-up excluded), for the ratio of iteration work
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, epsilon=eps, current_density=J, bound=bound,
                      dtype=np.complex64)
    _, t_iter = _timed(lambda: sol.solve(lambda s_: s_.residue > tol and s_.iteration < max_iter))
    # central finite differences at frozen alpha on n_fd random voxels of the design region, from complex128 forward
    # solves to residue 1e-11 (complex64 differences at h = 1e-3 are dominated by rounding where the gradient is small)
    rng = np.random.default_rng(6)
    vox = [tuple(int(d0 + rng.integers(0, design)) for _ in range(3)) for _ in range(n_fd)]
    epsilons = []
    for v in vox:
        for sgn in (+1, -1):
            e = eps.copy()
            e[v] += sgn * h
            epsilons.append(e)
    Efd, t_fd = _timed(lambda: fields_at_fixed_alpha(epsilons, grid, wavelength, info['alpha'],
                                                     current_density=J.astype(np.complex128), bound=bound,
                                                     dtype=np.complex128, tol=1e-11, max_iter=30000))
    losses = [float(loss_of(Efd[i])) for i in range(len(epsilons))]
    fd = np.array([(losses[2 * i] - losses[2 * i + 1]) / (2 * h) for i in range(n_fd)])
    ad = np.array([grad[v] for v in vox])
    gmax = float(np.abs(grad[dsl]).max())
    log(f'adjoint: {2 * n_fd} complex128 finite-difference solves in one batch, {t_fd:.2f} s')
    return {
        'grid': [n, n, n], 'design': [design] * 3, 'dtype': 'c64', 'tol': tol,
        'forward_s': round(t_fwd, 4), 'backward_s': round(t_bwd, 4),
        'first_call_forward_s': round(t_fwd0, 4), 'first_call_backward_s': round(t_bwd0, 4),
        'backward_over_forward': round(t_bwd / t_fwd, 3),
        'forward_solve_only_s': round(t_iter, 4), 'backward_over_forward_solve_only': round(t_bwd / t_iter, 3),
        'forward_iterations': int(info['forward_iterations']), 'adjoint_iterations': int(info['adjoint_iterations']),
        'fd_voxels': n_fd, 'fd_h': h, 'fd_dtype': 'c128 forward solves to residue 1e-11, alpha frozen',
        'fd_max_abs_err_over_max_grad': float(np.abs(ad - fd).max() / gmax),
        'fd_median_rel_err': float(np.median(np.abs(ad - fd) / np.maximum(np.abs(fd), 1e-300))),
        'fd_max_rel_err_where_grad_above_10pct': float(np.max(np.where(np.abs(ad) > 0.1 * gmax,
                                                                        np.abs(ad - fd) / np.abs(ad), 0.0))),
        'fd_batch_s': round(t_fd, 3),
        'timed': 'second forward/backward in the process (the first, with start-up costs, is first_call_*): '
                 'forward = Solution set-up + solve + loss; backward = adjoint solve of the transposed operator + '
                 'gradient assembly (loss.backward()); forward_solve_only = Solution.solve without set-up'}


def accuracy_benchmark(n, dev, E_plain, plain_residue, plain_iterations, log):
    """The complex64 field against a complex128 solve of the same problem converged to residual 1e-10, at residual
    1e-5 (the plain solve) and after FP64 refinement: a complex64 solve to residual 1e-4 and two refinement sweeps
    (correction solves to 1e-4), which is timed as time-to-accuracy. The plain field is also compared with the
    complex128 iterate after the same number of iterations (the same algorithm in double precision: what complex64
    arithmetic costs, apart from the distance of residual 1e-5 from convergence)."""
    import torch
    ref, _ = setup_solution(n, dev, dtype=np.complex128)
    snap = {}

    def callback(s_):
        if s_.iteration == plain_iterations:
            snap['E'] = s_.field_array.clone()
            snap['residual'] = s_.residue
        return s_.residue > 1e-10 and s_.iteration < 60000
    t0 = time.perf_counter()
    ref.solve(callback=callback)
    torch.cuda.synchronize()
    t128 = time.perf_counter() - t0
    Eref = ref.field_array.clone()
    nref = torch.linalg.vector_norm(Eref)
    out = {'reference': {'dtype': 'c128', 'residual': ref.residue, 'iterations': ref.iteration,
                         'solve_s': round(t128, 2)}}
    del ref
    torch.cuda.empty_cache()
    err = lambda E: float(torch.linalg.vector_norm(E.to(torch.complex128) - Eref) / nref)
    out['plain_c64'] = {'residual': plain_residue, 'iterations': plain_iterations, 'rel_l2_vs_c128': err(E_plain)}
    if 'E' in snap:
        E_same = snap.pop('E')
        out['plain_c64']['rel_l2_vs_c128_same_iteration'] = float(
            torch.linalg.vector_norm(E_plain.to(torch.complex128) - E_same) / torch.linalg.vector_norm(E_same))
        out['plain_c64']['c128_residual_at_same_iteration'] = snap['residual']
        del E_same
    sol, _ = setup_solution(n, dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.solve(callback=lambda s_: s_.residue > 1e-4 and s_.iteration < 20000)
    its = sol.iteration
    sweeps = []
    for _ in range(2):
        info = sol.refine(tol=1e-4)
        its += info['correction_iterations']
        sweeps.append({k: (float(v) if isinstance(v, float) else v) for k, v in info.items()})
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    out['refined_c64'] = {'schedule': 'solve to residual 1e-4, then 2 FP64 refinement sweeps (correction solves to '
                                      '1e-4)', 'rel_l2_vs_c128': err(sol.field_array_fp64), 'time_s': round(t_ref, 2),
                          'iterations': int(its), 'sweeps': sweeps}
    log(f"accuracy: plain c64 at 1e-5 {out['plain_c64']['rel_l2_vs_c128']:.2e} (vs the complex128 iterate of the "
        f"same iteration {out['plain_c64'].get('rel_l2_vs_c128_same_iteration', float('nan')):.2e}); refined "
        f"{out['refined_c64']['rel_l2_vs_c128']:.2e} in {t_ref:.1f} s; c128 reference {t128:.1f} s")
    return out


def stream_ceiling(n, dev, must_move_update, reps=20):
    """The bandwidth of kernels that only stream the planes of the passes (macromax_amd/csrc/mmx_stream_probe.hip:
    rows of 256 complex64 values, non-temporal, no arithmetic): 3 read + 3 written planes (a plain or Green pass of three
    components), 10 + 6 (the update pass with uniaxial chi), 6 + 6; and the iteration time the passes would take at
    those rates (3 passes of 48 B at the 3 + 3 rate, the update pass's bytes at the 10 + 6 rate)."""
    import ctypes
    import torch
    from macromax_amd import build as _b
    if not os.path.exists(_b.PROBE_LIB_PATH) or n ** 3 % 2048:
        return None
    lib = ctypes.CDLL(_b.PROBE_LIB_PATH)
    lib.mmx_stream_probe.restype = ctypes.c_int
    lib.mmx_stream_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p]
    V = n ** 3
    src = torch.zeros(10 * V, dtype=torch.complex64, device=dev)
    dst = torch.zeros(6 * V, dtype=torch.complex64, device=dev)
    stream = torch.cuda.current_stream(dev)
    out = {}
    for nr, nw in ((3, 3), (6, 6), (10, 6)):
        call = lambda: lib.mmx_stream_probe(nr, nw, V, src.data_ptr(), dst.data_ptr(), stream.cuda_stream)
        for _ in range(3):
            if call() != 0:
                return None
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            call()
        e1.record(stream)
        e1.synchronize()
        t = e0.elapsed_time(e1) * 1e-3 / reps
        out[f'rows_{nr}r{nw}w_GBps'] = round((nr + nw) * 8 * V / t / 1e9, 1)
    del src, dst
    r33, r106 = out['rows_3r3w_GBps'] * 1e9, out['rows_10r6w_GBps'] * 1e9
    floor = V * (3 * 48 / r33 + must_move_update / r106)
    out['iteration_floor_ms'] = round(floor * 1e3, 4)
    out['what'] = ('streaming-only kernels with the passes\' plane counts; iteration_floor_ms = the axis-0, Green and '
                   'update passes\' must-move bytes at those rates (the norm reduction excluded)')
    return out


def measured_traffic(n):
    """HBM bytes per voxel-iteration from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (256^3 only)."""
    path = os.path.join(ROOT, 'profiles', 'born_iterate_256cubed_aniso.md')
    if n != 256 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith('traffic_bytes_per_voxel_iteration:'):
            return float(line.split(':')[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--grid', type=int, default=256)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--residual', type=float, default=1e-5)
    ap.add_argument("--max-iter", type=int, default=10000)
    ap.add_argument('--no-batched', action='store_true', help='skip the batched-solve comparison')
    ap.add_argument('--no-tune', action='store_true', help='keep the library default tile shapes (no autotuning)')
    ap.add_argument('--no-accuracy', action='store_true', help='skip the complex128 reference solve')
    ap.add_argument('--no-storage', action='store_true', help='skip the timing of the full and symmetric chi storage')
    ap.add_argument('--batched-only', action='store_true', help='only the batched-solve comparison (own JSON line)')
    ap.add_argument('--no-adjoint', action='store_true', help='skip the adjoint-gradient measurement')
    ap.add_argument('--no-magnetic', action='store_true', help='skip the magnetic-medium measurement')
    ap.add_argument('--magnetic-only', action='store_true', help='only the magnetic-medium measurement')
    ap.add_argument('--no-configs', action='store_true', help='skip the 256^2 and 128^3 configurations')
    ap.add_argument('--configs-only', action='store_true', help='only the 256^2 and 128^3 configurations')
    ap.add_argument('--adjoint-only', action='store_true', help='only the adjoint-gradient measurement')
    ap.add_argument('--copy-probe-mb', type=int, default=0,
                    help='also run one device-to-device copy of this many MiB (calibrates HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus N > 1 is not supported (the multi-GPU '
                                   'slab decomposition is a design, not a measured path)'}))
        sys.exit(2)
    import torch
    dev = torch.device('cuda:0')
    log = lambda m: print(f'bench: {m}', file=sys.stderr, flush=True)
    if args.batched_only:
        print(json.dumps({'batched': batched_benchmark(log=log)}), flush=True)
        return
    if args.magnetic_only:
        print(json.dumps({'magnetic': magnetic_benchmark(log=log)}), flush=True)
        return
    if args.configs_only:
        print(json.dumps({'configs': configs_benchmark(log=log)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({'adjoint': adjoint_benchmark(log=log)}), flush=True)
        return
    n = args.grid
    V = n ** 3
    print(f'bench: building the {n}^3 anisotropic workload', file=sys.stderr, flush=True)
    sol, setup_s = setup_solution(n, dev)
    print(f'bench: set-up {setup_s:.1f} s, alpha = {sol.alpha}', file=sys.stderr, flush=True)
    tuned = None
    if not args.no_tune:
        from macromax_amd.autotune import tune_solution
        tuned = tune_solution(sol)
        log(f'autotune: {tuned}')

    # time-to-residual from E = 0 (includes the per-iteration host read of the two norms)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.solve(callback=lambda s: s.residue > args.residual and s.iteration < args.max_iter)
    torch.cuda.synchronize()
    ttr = time.perf_counter() - t0
    iters_to_res = sol.iteration
    residue = sol.residue
    print(f'bench: residual {residue:.2e} after {iters_to_res} iterations, {ttr:.3f} s', file=sys.stderr, flush=True)
    E_plain = None if args.no_accuracy else sol.field_array.clone()  # the field at residual 1e-5

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb * 2 ** 20 // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f'bench: copy probe of {args.copy_probe_mb} MiB done', file=sys.stderr, flush=True)
        del src, dst
    wall_per_iter, gpu_per_iter = time_iterations(sol, args.steps, args.warmup)
    ms_per_step = wall_per_iter * 1e3
    value = V / wall_per_iter

    chi_kind = sol.chi_kind
    chi_bytes = {1: 8, 2: 72, 3: 48, 4: 28}[chi_kind]
    # bytes the fused schedule must move (the source is the plane z = 20: 1/n of the voxels)
    b_must = must_move_bytes(3, 3, 8, chi_kind, 1.0 / n)
    b_alg = algorithmic_bytes(3, 3, 8, 9)  # the reference's sequence with its 3x3 chi: 480 B
    achieved = b_must * V / gpu_per_iter / 1e9
    traffic = measured_traffic(n)
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4),
                'traffic': None if traffic is None else round(traffic * V),
                'kernel': 'mmx_born_iterate_split: one Born iteration = 4 pencil-pass kernels on the launch '
                          'stream (its FP64 norm reduction runs on a second stream, beside the next iteration), '
                          'timed with HIP events on the launch stream; achieved = bytes the fused schedule must '
                          'move per iteration / event time',
                'bytes_per_voxel_iter': round(b_must, 2), 'gpu_ms_per_iter_events': round(gpu_per_iter * 1e3, 4),
                'alg_bytes_per_voxel_iter': b_alg,
                'alg_equivalent_GBps': round(b_alg * V / gpu_per_iter / 1e9, 1),
                'alg_equivalent_frac': round(b_alg * V / gpu_per_iter / 1e9 / HBM_PEAK_GBS, 4)}
    ceiling = stream_ceiling(n, dev, b_must - 3 * 48)
    if ceiling is not None:
        ceiling['fraction_of_floor'] = round(ceiling['iteration_floor_ms'] / (gpu_per_iter * 1e3), 4)
        roofline['stream_ceiling'] = ceiling
        log(f"stream ceiling: {ceiling}")
    if traffic is not None:
        roofline['traffic_bytes_per_voxel_iter'] = traffic
        roofline['traffic_GBps'] = round(traffic * V / gpu_per_iter / 1e9, 1)
        roofline['traffic_frac'] = round(roofline['traffic_GBps'] / HBM_PEAK_GBS, 4)

    out = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'voxel-iter/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms_per_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'c64', 'data': 'synthetic',
        'config': {'workload': f'{n}^3 anisotropic vectorial Born iteration (Voronoi grains, 2000 seeds), complex64',
                   'grid': [n, n, n], 'components': 3, 'chi_kind': chi_kind,
                   'chi': {1: 'isotropic', 2: 'full 3x3', 3: 'symmetric 3x3 (6 entries)',
                           4: 'uniaxial u I + w n n^T'}[chi_kind] + f', {chi_bytes} B per voxel',
                   'must_move_bytes_per_voxel_iter': round(b_must, 2), 'pitch_nm': 250, 'wavelength_nm': 1000},
        'hbm_fraction': roofline['frac'],
        'roofline': roofline,
        'time_to_residual_s': round(ttr, 4), 'iterations_to_residual': iters_to_res, 'residual': residue,
        'setup_s': round(setup_s, 2), 'alpha': [sol.alpha.real, sol.alpha.imag],
    }
    if tuned is not None:
        out['autotune'] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in tuned.items()}
    if not args.no_cpu_baseline:
        log('CPU baseline (complex128 oracle iterations)')
        cb = cpu_baseline(sol, n)
        out['cpu_baseline'] = cb
        out['cpu_time_to_residual_s_extrapolated'] = round(cb['seconds_per_iteration'] * iters_to_res, 1)
    if not args.no_accuracy:
        del sol
        torch.cuda.empty_cache()
        out['accuracy'] = accuracy_benchmark(n, dev, E_plain, residue, iters_to_res, log)
        out['time_to_accuracy_1e-4_s'] = out['accuracy']['refined_c64']['time_s']
        out['accuracy_at_time_to_accuracy'] = out['accuracy']['refined_c64']['rel_l2_vs_c128']
        del E_plain
        torch.cuda.empty_cache()
    else:
        del sol
    if not args.no_storage:
        out['chi_storage'] = storage_benchmark(n, dev, log=log)
    if not args.no_batched:
        torch.cuda.empty_cache()
        out['batched'] = batched_benchmark(log=log)
    if not args.no_adjoint:
        torch.cuda.empty_cache()
        out['adjoint'] = adjoint_benchmark(log=log)
    if not args.no_magnetic:
        torch.cuda.empty_cache()
        out['magnetic'] = magnetic_benchmark(log=log)
    if not args.no_configs:
        torch.cuda.empty_cache()
        out['configs'] = configs_benchmark(log=log)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
