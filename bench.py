"""Benchmark: Born-series iterations of the 256^3 anisotropic vectorial complex64 solve on one MI355X.

Workload (the metric's configuration): 256^3 grid, 250 nm pitch, vacuum wavelength 1 um; permittivity per voxel
R diag(no^2, no^2, ne^2) R^T (no = 1.50, ne = 1.65) with seeded random optic axes constant within 2000 Voronoi
grains; an isotropic n = 1.33 slab for z in [64, 96); circularly polarised current sheet on z = 20; 16-pixel
LinearBound on z (periodic x, y); complex64 fields, FP64 norms.

A "step" is one Born iteration (Solution.step: the fused HIP kernel sequence plus the accept/reject decision).
`value` is voxel-iterations/s with all inputs resident in HBM. The JSON line also carries the HBM roofline of the
iteration (algorithmic bytes per voxel-iteration, 480 B for this workload; `traffic` is the HBM bytes per iteration
measured with rocprofv3 counters, from profiles/born_iterate_256cubed_aniso.md) and a CPU baseline: the NumPy
complex128 oracle iteration on this host (single-threaded numpy.fft), timed for one iteration of the same grid.

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


def setup_solution(n, device):
    import torch
    import macromax_amd as mm
    from macromax_amd.bound import LinearBound
    wavelength = 1e-6
    pitch = 250e-9
    grid = mm.Grid(shape=[n, n, n], step=pitch)
    eps = build_anisotropic_medium(n, device)
    J = torch.zeros((3, n, n, n), dtype=torch.complex64, device=device)
    J[0, :, :, 20] = 1.0 / math.sqrt(2)
    J[1, :, :, 20] = 1.0j / math.sqrt(2)
    bound = LinearBound(grid, thickness=[[0, 0], [0, 0], [16 * pitch, 16 * pitch]])
    t0 = time.perf_counter()
    sol = mm.Solution(grid, vacuum_wavelength=wavelength, epsilon=eps, current_density=J, bound=bound,
                      dtype=np.complex64)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    del eps, J
    return sol, setup_s


def cpu_baseline(sol, n):
    """One complex128 NumPy oracle iteration of the same problem on this host (numpy.fft is single-threaded)."""
    from oracle import born_oracle as ob
    chi = sol.chi_effective.cpu().numpy().astype(np.complex128)
    s = sol.source_normalized.cpu().numpy().astype(np.complex128)
    E = sol.field_array.cpu().numpy().astype(np.complex128)
    k = ob.k_vectors((n, n, n), np.asarray(sol.grid.step) * sol.wavenumber)
    o = ob.BornOracle(chi, s, sol.alpha, sol.beta, k, True, E0=E)
    t0 = time.perf_counter()
    o.step()
    dt = time.perf_counter() - t0
    return {'value': n ** 3 / dt, 'unit': 'voxel-iter/s', 'cores': 1, 'kind': 'port',
            'sample': f'1 complex128 iteration of the same {n}^3 anisotropic problem (oracle/born_oracle.py, '
                      f'numpy.fft), {dt:.1f} s', 'seconds_per_iteration': dt,
            'os_cpu_count': os.cpu_count(), 'affinity_cores': len(os.sched_getaffinity(0))}


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
    ap.add_argument('--copy-probe-mb', type=int, default=0,
                    help='also run one device-to-device copy of this many MiB (calibrates HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({'error': 'this build runs on one GPU; --gpus N > 1 is not supported (the multi-GPU '
                                   'slab decomposition is a design, not a measured path)'}))
        sys.exit(2)
    import torch
    dev = torch.device('cuda:0')
    n = args.grid
    V = n ** 3
    print(f'bench: building the {n}^3 anisotropic workload', file=sys.stderr, flush=True)
    sol, setup_s = setup_solution(n, dev)
    print(f'bench: set-up {setup_s:.1f} s, alpha = {sol.alpha}', file=sys.stderr, flush=True)

    # time-to-residual from E = 0 (includes the per-iteration host read of the two norms)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sol.solve(callback=lambda s: s.residue > args.residual and s.iteration < args.max_iter)
    torch.cuda.synchronize()
    ttr = time.perf_counter() - t0
    iters_to_res = sol.iteration
    residue = sol.residue
    print(f'bench: residual {residue:.2e} after {iters_to_res} iterations, {ttr:.3f} s', file=sys.stderr, flush=True)

    if args.copy_probe_mb:
        src = torch.ones(args.copy_probe_mb * 2 ** 20 // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f'bench: copy probe of {args.copy_probe_mb} MiB done', file=sys.stderr, flush=True)
        del src, dst
    for _ in range(args.warmup):
        sol.step()
    stream = torch.cuda.current_stream(dev)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        sol.step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = wall * 1e3 / args.steps
    value = V * args.steps / wall

    b_alg = 480.0  # bytes per voxel-iteration: 4*d*C*e + 2*B_chi + B_s + C*e with d=3, C=3, e=8, B_chi=72
    gpu_ms_per_iter = gpu_ms / args.steps
    achieved = b_alg * V / (gpu_ms_per_iter * 1e-3) / 1e9
    traffic = measured_traffic(n)
    roofline = {'bound': 'hbm', 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                'frac': round(achieved / HBM_PEAK_GBS, 4),
                'traffic': None if traffic is None else round(traffic * V),
                'kernel': 'mmx_born_iterate: one Born iteration = 4 pencil-pass kernels + FP64 norm reduction, '
                          'timed with HIP events on its launch stream',
                'bytes_per_voxel_iter_alg': b_alg, 'gpu_ms_per_iter_events': round(gpu_ms_per_iter, 4)}
    if traffic is not None:
        roofline['traffic_bytes_per_voxel_iter'] = traffic
        roofline['actual_GBps'] = round(traffic * V / (gpu_ms_per_iter * 1e-3) / 1e9, 1)

    out = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'voxel-iter/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms_per_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'c64', 'data': 'synthetic',
        'config': {'workload': f'{n}^3 anisotropic vectorial Born iteration (Voronoi grains, 2000 seeds), complex64',
                   'grid': [n, n, n], 'components': 3, 'chi': '3x3 per voxel', 'pitch_nm': 250,
                   'wavelength_nm': 1000},
        'hbm_fraction': roofline['frac'],
        'roofline': roofline,
        'time_to_residual_s': round(ttr, 4), 'iterations_to_residual': iters_to_res, 'residual': residue,
        'setup_s': round(setup_s, 2), 'alpha': [sol.alpha.real, sol.alpha.imag],
    }
    if not args.no_cpu_baseline:
        print('bench: CPU baseline (one complex128 oracle iteration)', file=sys.stderr, flush=True)
        cb = cpu_baseline(sol, n)
        out['cpu_baseline'] = cb
        out['cpu_time_to_residual_s_extrapolated'] = round(cb['seconds_per_iteration'] * iters_to_res, 1)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
