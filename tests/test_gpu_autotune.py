"""The autotuner (macromax_amd.autotune): it returns a complete parameter set, leaves the plan configured with it,
caches per size (memory and an optional JSON file), and the tuned plan computes the same iteration as the
default one (complex64, relative L2 <= 1e-5: other tile shapes or kernel families round differently)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('shape,C,kind', [((64, 64, 128), 3, 3), ((96, 80), 1, 1), ((128, 128), 3, 1)],
                         ids=lambda v: str(v))
def test_autotune_result_and_parity(shape, C, kind, tmp_path, monkeypatch):
    import torch
    from macromax_amd import autotune
    from macromax_amd.backend.hip import Plan
    print(f'autotune {shape} C={C} kind={kind}', flush=True)
    monkeypatch.setenv('MACROMAX_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    dev = torch.device('cuda:0')
    step = [0.4] * len(shape)
    tuned = Plan(shape, step, torch.complex64, dev)
    res = autotune.tune_plan(tuned, C, kind, nbatch=2, chi_shared=True, iterations=2, repeats=2, use_cache=False)
    for name in autotune.PARAMETERS:
        assert name in res
        assert tuned.get_param(name) == res[name], name
    assert res['ms'] > 0
    on_disk = json.loads((tmp_path / 'tune.json').read_text())
    assert autotune.cache_key(shape, torch.complex64, C, kind, 2, True) in on_disk
    # a cached call configures a fresh plan the same way without timing anything
    other = Plan(shape, step, torch.complex64, dev)
    res2 = autotune.tune_plan(other, C, kind, nbatch=2, chi_shared=True)
    assert res2['tuning_s'] == res['tuning_s']
    for name in autotune.PARAMETERS:
        assert other.get_param(name) == res[name]
    # parity of one batched iteration: tuned plan vs default plan
    default = Plan(shape, step, torch.complex64, dev)
    V = int(np.prod(shape))
    g = torch.Generator(device=dev).manual_seed(3)
    crand = lambda *s: torch.complex(torch.randn(*s, device=dev, generator=g),
                                     torch.randn(*s, device=dev, generator=g)).to(torch.complex64) * 0.2
    nchi = {1: 1, 3: 6}[kind]
    E, s, chi = crand(2, C, V), crand(2, C, V), crand(1, nchi, V)
    coef = torch.tensor([1.1, 0.5, 1.0, 0.0, 0.0, 0.0] * 2, dtype=torch.float32, device=dev)
    outs = []
    for plan in (tuned, default):
        Eo = torch.empty_like(E)
        norms = torch.zeros(4, dtype=torch.float64, device=dev)
        plan.born_step(E, Eo, s, chi, C, kind, coef, 2, C * V, 0, norms)
        outs.append((Eo.cpu().numpy(), norms.cpu().numpy()))
    d = outs[0][0] - outs[1][0]
    assert np.linalg.norm(d) <= 1e-5 * np.linalg.norm(outs[1][0] - E.cpu().numpy())
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-5)


def test_autotune_times_radix_orders_on_lds_axes(tmp_path, monkeypatch, capsys):
    """A complex128 plan runs every axis on the LDS-staged kernels: the tuner times each radix order, and the plan it
    leaves computes the iteration of the default plan (to complex128 rounding)."""
    import torch
    from macromax_amd import autotune
    from macromax_amd.backend.hip import Plan
    monkeypatch.setenv('MACROMAX_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    dev = torch.device('cuda:0')
    shape, C, kind = (48, 40, 24), 3, 2
    step = [0.45, 0.5, 0.55]
    tuned = Plan(shape, step, torch.complex128, dev)
    res = autotune.tune_plan(tuned, C, kind, nbatch=1, iterations=2, repeats=2, use_cache=False, verbose=True)
    out = capsys.readouterr().out
    assert 'radix_order=1' in out and 'radix_order=2' in out
    assert tuned.get_param('radix_order') == res['radix_order'] in (0, 1, 2)
    default = Plan(shape, step, torch.complex128, dev)
    V = int(np.prod(shape))
    g = torch.Generator(device=dev).manual_seed(5)
    crand = lambda *s: torch.complex(torch.randn(*s, device=dev, generator=g, dtype=torch.float64),
                                     torch.randn(*s, device=dev, generator=g, dtype=torch.float64)) * 0.2
    E, s, chi = crand(1, C, V), crand(1, C, V), crand(1, 9, V)
    coef = torch.tensor([1.1, 0.5, 1.0, 0.0, 0.0, 0.0], dtype=torch.float64, device=dev)
    outs = []
    for plan in (tuned, default):
        Eo = torch.empty_like(E)
        norms = torch.zeros(2, dtype=torch.float64, device=dev)
        plan.born_step(E, Eo, s, chi, C, kind, coef, 1, C * V, 0, norms)
        outs.append(Eo.cpu().numpy())
    d = outs[0] - outs[1]
    assert np.linalg.norm(d) <= 1e-11 * np.linalg.norm(outs[1] - E.cpu().numpy())


def test_autotune_skips_radix_order_on_register_axes(tmp_path, monkeypatch, capsys):
    """Every axis of a complex64 256 x 256 grid has register kernels; once they win, the radix order (which only
    the LDS-staged kernels use) is not timed."""
    import torch
    from macromax_amd import autotune
    from macromax_amd.backend.hip import Plan
    monkeypatch.setenv('MACROMAX_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    plan = Plan((256, 256), [0.4, 0.4], torch.complex64, torch.device('cuda:0'))
    res = autotune.tune_plan(plan, 1, 1, nbatch=4, iterations=2, repeats=2, use_cache=False, verbose=True)
    out = capsys.readouterr().out
    if res['reg_fft'] == 1:
        assert 'radix_order=' not in out
    else:
        assert 'radix_order=1' in out
