"""MacroMax-compatible Born-series solver for AMD Instinct MI355X (gfx950).

``solve``, ``Solution``, ``Grid``, ``backend`` and ``bound`` mirror the reference package's public names
(MacroMax ``python/macromax/__init__.py:73-82``). The iteration runs in hand-written HIP kernels from
``libmacromax_hip.so``; build it with ``python -m macromax_amd.build``.
"""
from . import backend, bound
from .grid import Grid
from .solver import Solution, solve

__version__ = '0.1.0'

__all__ = ['solve', 'Solution', 'Grid', 'backend', 'bound', '__version__']
