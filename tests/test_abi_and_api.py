"""CPU checks of the C ABI library (loads, exports every declared symbol) and of the host-side API."""
import os
import re
import subprocess

import numpy as np
import numpy.testing as npt
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'macromax_hip.h')


def _declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(mmx_\w+)\s*\(', text, flags=re.M)))


def test_header_declares_the_abi():
    names = _declared_functions()
    for required in ('mmx_plan_create', 'mmx_plan_destroy', 'mmx_fft', 'mmx_born_step', 'mmx_norm2',
                     'mmx_chi_mul', 'mmx_longitudinal_projection_ft', 'mmx_version', 'mmx_last_error'):
        assert required in names


def test_library_loads_and_exports_every_declared_symbol():
    from macromax_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail('libmacromax_hip.so is not built (run python -m macromax_amd.build)')
    L = _native.lib()
    for name in _declared_functions():
        assert hasattr(L, name), name
    assert set(_native.EXPORTED) == set(_declared_functions())
    out = subprocess.run(['nm', '-D', '--defined-only', _native.LIB_PATH], capture_output=True, text=True).stdout
    for name in _declared_functions():
        assert re.search(rf'\bT {name}\b', out), name
    assert L.mmx_version() >= 100


def test_header_constants_match_the_binding():
    """Every MMX_PARAM_*, MMX_STATE_* and MMX_TRACK_* value of the header is the one the ctypes binding uses."""
    from macromax_amd import _native
    text = open(os.path.join(ROOT, 'include', 'macromax_hip.h')).read()
    defs = {m.group(1): int(m.group(2)) for m in re.finditer(r'#define\s+(MMX_\w+)\s+(-?\d+)u?\b', text)}

    def group(prefix):
        return {k[len(prefix):].lower(): v for k, v in defs.items() if k.startswith(prefix)}
    assert group('MMX_PARAM_') == _native.PARAMS
    track = group('MMX_TRACK_')
    assert track.pop('len') == _native.TRACK_LEN
    assert track == _native.TRACK
    assert (defs['MMX_STATE_PARITY'], defs['MMX_STATE_DONE'], defs['MMX_STATE_STALE']) == \
        (_native.STATE_PARITY, _native.STATE_DONE, _native.STATE_STALE)
    assert defs['MMX_C64'] == _native.MMX_C64 and defs['MMX_C128'] == _native.MMX_C128


def test_backend_load_without_gpu_fails_like_the_reference():
    import torch
    from macromax_amd import backend
    from macromax_amd.grid import Grid
    if torch.cuda.is_available():
        pytest.skip('a GPU is present')
    with pytest.raises(TypeError, match='No viable backend detected!'):
        backend.load(3, Grid([4, 4, 4], 1.0), np.complex64)
    # other backend types are not provided by this package: skipped, then the same error
    with pytest.raises(TypeError, match='No viable backend detected!'):
        backend.load(3, Grid([4, 4, 4], 1.0), np.complex64, config_list=[{'type': 'numpy'}, {'type': 'torch'}])


def test_backend_config_validation():
    from macromax_amd import backend
    backend.config(type='hip')
    backend.config([{'type': 'hip'}])
    with pytest.raises(TypeError):
        backend.config(['hip'])
    backend.config()


def test_grid_matches_reference_behaviour():
    from macromax_amd.grid import Grid
    g = Grid([5, 10, 20], 100e-9)
    npt.assert_array_equal(g.shape, [5, 10, 20])
    npt.assert_allclose((g * 1e7).step, [1, 1, 1])
    # a grid defined by scalars is the sequence of its coordinates
    npt.assert_allclose(np.asarray(Grid(5).f), [0, 1 / 5, 2 / 5, -2 / 5, -1 / 5])
    npt.assert_allclose(np.asarray(Grid(5, 2).k), np.array([0, 1 / 10, 2 / 10, -2 / 10, -1 / 10]) * 2 * np.pi)
    assert Grid(5).f[1] == 1 / 5
    npt.assert_allclose(Grid([5, 2]).f[0], np.array([[0], [1 / 5], [2 / 5], [-2 / 5], [-1 / 5]]))
    r = np.arange(10) * 0.5 - 1
    g1 = Grid.from_ranges(r)
    npt.assert_allclose(g1[0].ravel(), r)
    assert g1 == Grid.from_ranges(r)
    # shapes of the open-grid axis vectors
    assert Grid([3, 4, 5])[1].shape == (1, 4, 1)


def test_bound_matches_reference_tables():
    from macromax_amd.bound import Electric, LinearBound, Magnetic, PeriodicBound
    from macromax_amd.grid import Grid
    grid = Grid([10], 1e-6)
    b = LinearBound(grid, thickness=4e-6, max_extinction_coefficient=0.1, background_permittivity=1.33 ** 2)
    kappa = np.array([0.100, 0.075, 0.050, 0.025, 0.0, 0.0, 0.025, 0.050, 0.075, 0.100])
    npt.assert_array_almost_equal(b.extinction, kappa)
    assert isinstance(b, Electric) and not isinstance(b, Magnetic)
    npt.assert_array_almost_equal(b.permittivity, (1.33 + 1j * kappa) ** 2)
    b = LinearBound(grid, thickness=[4e-6, 2e-6], max_extinction_coefficient=(0.1, 0.2))
    npt.assert_array_almost_equal(b.extinction, [0.100, 0.075, 0.050, 0.025, 0.0, 0.0, 0.0, 0.0, 0.100, 0.200])
    grid2 = Grid([2, 10], 1e-6)
    b = LinearBound(grid2, thickness=[4e-6, 2e-6], max_extinction_coefficient=[[0, 0], [0.1, 0.1]])
    k1 = [0.100, 0.075, 0.050, 0.025, 0.0, 0.0, 0.0, 0.0, 0.050, 0.100]
    npt.assert_array_almost_equal(b.extinction, [k1, k1])
    p = PeriodicBound(grid2)
    npt.assert_array_equal(p.thickness, [[0, 0], [0, 0]])
    npt.assert_array_equal(p.electric_susceptibility, 0)


def test_hermitian3_eigvals_closed_form():
    import torch
    from macromax_amd.backend.hip import hermitian3_eigvals
    rng = np.random.default_rng(0)
    A = rng.standard_normal((3, 3, 50)) + 1j * rng.standard_normal((3, 3, 50))
    H = A + np.conj(np.swapaxes(A, 0, 1))
    H[:, :, 0] = np.diag([2.7, 2.7, 2.7])  # degenerate
    H[:, :, 1] = np.diag([4.0, 5.0, -6.0])
    got = hermitian3_eigvals(torch.as_tensor(H)).numpy()
    for v in range(50):
        npt.assert_allclose(got[:, v], np.linalg.eigvalsh(H[:, :, v]), atol=1e-9)


def test_no_oracle_or_cpu_fallback_in_the_product():
    """The product package never imports the oracle and has no CPU compute path."""
    pkg = os.path.join(ROOT, 'macromax_amd')
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith('.py'):
                text = open(os.path.join(dirpath, f)).read()
                assert 'oracle' not in text.replace('oracle_', ''), f
                assert 'numpy.fft' not in text and 'np.fft.fft' not in text and 'torch.fft' not in text, f


def test_stream_probe_library_exports_its_entry_point():
    """The streaming-ceiling probe of bench.py is built beside the library and exports mmx_stream_probe (loaded with
    ctypes; no GPU call)."""
    import ctypes
    import os
    from macromax_amd import build
    if not os.path.exists(build.PROBE_LIB_PATH):
        pytest.skip('libmmx_stream_probe.so not built (python -m macromax_amd.build)')
    lib = ctypes.CDLL(build.PROBE_LIB_PATH)
    assert hasattr(lib, 'mmx_stream_probe')


def test_grid_import_paths_of_the_reference():
    """Grid is reachable under the reference's import paths: macromax.utils, macromax.utils.ft and
    macromax.utils.ft.grid (with MutableGrid), all the same class."""
    import macromax_amd
    from macromax_amd import grid
    from macromax_amd.utils import Grid as G1
    from macromax_amd.utils.ft import Grid as G2, MutableGrid as M2
    from macromax_amd.utils.ft.grid import Grid as G3, MutableGrid as M3
    assert macromax_amd.Grid is G1 is G2 is G3 is grid.Grid
    assert M2 is M3 is grid.MutableGrid
