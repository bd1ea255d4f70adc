"""Autotuner for the pencil-pass kernels: picks the tile shapes, the axis of the fused Green pass, the kernel family
(register-resident radix-16/8 or LDS-staged radix-8/4/2), the radix order of the LDS-staged transforms and the XCD
block order for one grid shape, element type, field and chi kind and batch size, by timing the fused iteration
(``mmx_born_iterate``) on the device.

No reference counterpart: the reference calls FFT libraries (pyFFTW plans, ``backend/numpy.py:46-89``;
PyTorch's FFT module). The search is coordinate descent over ``mmx_plan_set_param`` parameters (each parameter in turn, the
others at their best values so far), timed with HIP events on the launch stream, median of ``repeats`` runs of
``iterations`` iterations. The choice depends only on sizes, so results are cached per key in memory and, when the
environment variable ``MACROMAX_AMD_TUNE_CACHE`` names a JSON file, on disk.
"""
from __future__ import annotations

import json
import os
import threading
import time

import numpy as np
import torch

__all__ = ['tune_plan', 'tune_solution', 'apply', 'cache_key', 'PARAMETERS']

# parameter -> candidate values; the first is the library's default (tile sizes 0 = automatic; for xcd_remap None
# stands for the library's default mask)
PARAMETERS = {
    'g_axis': (1, 0),
    'reg_fft': (1, 0),
    'radix_order': (0, 1, 2),
    'p_rows': (0, 4, 8, 16, 32),
    'p_strided': (0, 8, 16, 32, 64),
    'p_green': (0, 4, 8, 16),
    'xcd_remap': (None, 0, 255),
}

_cache: dict = {}
_lock = threading.Lock()


def cache_key(shape, dtype, ncomp: int, chi_kind: int, nbatch: int, chi_shared: bool) -> str:
    return json.dumps([list(int(_) for _ in shape), str(dtype), int(ncomp), int(chi_kind), int(nbatch),
                       bool(chi_shared)])


def _cache_file():
    return os.environ.get('MACROMAX_AMD_TUNE_CACHE')


def _load_disk():
    path = _cache_file()
    if path and os.path.exists(path):
        try:
            with open(path) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def _store_disk(key, value):
    path = _cache_file()
    if not path:
        return
    data = _load_disk()
    data[key] = value
    tmp = path + '.tmp'
    with open(tmp, 'w') as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def apply(plan, params: dict):
    """Set the parameters of a tuning result on a plan."""
    for name, value in params.items():
        if value is None:
            continue
        plan.set_param(name, int(value))


def _lds_axes(plan) -> bool:
    """Whether some axis of the plan runs on the LDS-staged kernels whatever reg_fft says (complex128, or an axis
    length without register kernels)."""
    from . import _native
    return plan.dtype != torch.complex64 or any(not _native.register_fft_length(n) for n in plan.shape if n > 1)


def _candidates(plan):
    out = {}
    from . import _native
    reg = [n for n in plan.shape if n > 1 and _native.register_fft_length(n)]
    for name, values in PARAMETERS.items():
        if name == 'g_axis' and sum(1 for n in plan.shape if n > 1) < 3:
            continue
        if name == 'reg_fft' and (not reg or plan.dtype != torch.complex64):
            continue
        if name == 'radix_order' and not _lds_axes(plan) and 'reg_fft' not in out:
            continue  # every axis runs on the register kernels, whose radices are fixed
        if name == 'p_strided' and sum(1 for n in plan.shape if n > 1) < 3:
            continue
        out[name] = values
    return out


def _time_iterate(plan, bufs, ncomp, chi_kind, nbatch, chi_bstride, s_bstride, s_box, iterations, repeats):
    E, Eo, W, s, chi, coef, norms = bufs
    plan.born_prologue(E, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, s_box=s_box)
    plan.born_iterate(E, Eo, W, s, chi, ncomp, chi_kind, coef, norms, nbatch, s_bstride, chi_bstride, s_box=s_box)
    times = []
    for _ in range(repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(iterations):
            plan.born_iterate(E, Eo, W, s, chi, ncomp, chi_kind, coef, norms, nbatch, s_bstride, chi_bstride,
                              s_box=s_box)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1) / iterations)
    return float(np.median(times))


def tune_plan(plan, ncomp: int, chi_kind: int, nbatch: int = 1, chi_shared: bool = True, iterations: int = 5,
              repeats: int = 5, use_cache: bool = True, verbose: bool = False, operands=None) -> dict:
    """Tune ``plan`` for ``nbatch`` problems of ``ncomp`` components and chi kind ``chi_kind``; the best parameters
    are set on the plan and returned (with the measured milliseconds per iteration under ``'ms'``).

    ``operands`` = (s, chi, coef, s_box) of a real problem (one problem, ``nbatch`` 1) times the kernels that problem
    uses (e.g. the source-box variant); otherwise random dense operands of the right sizes are used."""
    key = cache_key(plan.shape, plan.dtype, ncomp, chi_kind, nbatch, chi_shared)
    if use_cache:
        with _lock:
            hit = _cache.get(key) or _load_disk().get(key)
        if hit is not None:
            apply(plan, {k: v for k, v in hit.items() if k in PARAMETERS})
            return dict(hit)
    dev = plan.device
    V = plan.V
    g = torch.Generator(device=dev).manual_seed(0)

    def crand(*shape):
        return torch.complex(torch.randn(*shape, device=dev, generator=g),
                             torch.randn(*shape, device=dev, generator=g)).to(plan.dtype) * 0.1
    nchi = {1: 1, 2: 9, 3: 6, 4: 4}[chi_kind]
    E = crand(nbatch, ncomp, V)
    if operands is not None:
        s_op, chi_op, coef_op, s_box = operands
        bufs = (E, torch.empty_like(E), torch.empty_like(E), s_op, chi_op, coef_op,
                torch.zeros(2 * nbatch, dtype=torch.float64, device=dev))
        s_bstride = 0
    else:
        s_box = None
        bufs = (E, torch.empty_like(E), torch.empty_like(E), crand(nbatch, ncomp, V),
                crand(1 if chi_shared else nbatch, nchi, V),
                torch.tensor([1.2, 0.4, 1.0, 0.0, 0.0, 0.0] * nbatch, dtype=plan.real_dtype, device=dev),
                torch.zeros(2 * nbatch, dtype=torch.float64, device=dev))
        s_bstride = ncomp * V
    chi_bstride = 0 if chi_shared else nchi * V
    default_xcd = plan.get_param('xcd_remap')
    defaults = {name: (default_xcd if values[0] is None else values[0]) for name, values in PARAMETERS.items()}
    cands = {name: tuple(default_xcd if v is None else v for v in values)
             for name, values in _candidates(plan).items()}
    t_start = time.perf_counter()

    def measure(params):
        apply(plan, defaults)
        apply(plan, params)
        return _time_iterate(plan, bufs, ncomp, chi_kind, nbatch, chi_bstride, s_bstride, s_box, iterations,
                             repeats)

    best = {name: values[0] for name, values in cands.items()}
    initial = dict(best)
    measure(best)  # warm-up: the first timed launches run at lower clocks and with cold caches
    best_ms = measure(best)
    for name, values in cands.items():
        if name == 'radix_order' and best.get('reg_fft', 1) and not _lds_axes(plan):
            continue  # the register kernels won: no axis uses the radix order
        for v in values:
            if v == best[name]:
                continue
            trial = dict(best)
            trial[name] = v
            ms = measure(trial)
            if verbose:
                print(f'autotune {plan.shape} {name}={v}: {ms:.4f} ms', flush=True)
            if ms < best_ms * 0.97:  # require a clear gain over the current choice (timing noise is ~1-2 %)
                best, best_ms = trial, ms
    if best != initial:
        # confirm the winner against the defaults, timed back to back, before keeping it
        ms_initial, ms_best = measure(initial), measure(best)
        if not ms_best < ms_initial * 0.97:
            best, best_ms = initial, ms_initial
        else:
            best_ms = ms_best
    measure(best)  # leave the plan configured with the winner
    result = dict(defaults)
    result.update(best)
    result['ms'] = best_ms
    result['tuning_s'] = time.perf_counter() - t_start
    with _lock:
        _cache[key] = result
        _store_disk(key, result)
    return dict(result)


def tune_solution(solution, **kwargs) -> dict:
    """Tune the plan of a ``Solution``'s backend on its own chi, source (and source box) and coefficients."""
    o = solution._batch_operands()
    return tune_plan(solution.backend.plan, o['ncomp'], o['chi_kind'], nbatch=1,
                     operands=(o['s'], o['chi'], o['coef'], o['s_box']), **kwargs)
