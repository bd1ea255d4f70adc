"""MacroMax-compatible Born-series solver for AMD Instinct MI355X (gfx950).

``solve``, ``Solution``, ``ScatteringMatrix``, ``Grid``, ``log`` and ``backend`` mirror the reference package's public
names (MacroMax ``python/macromax/__init__.py:73-82``), with ``bound`` and ``utils`` as in its subpackages. The
iteration runs in hand-written HIP kernels from ``libmacromax_hip.so``; build it with ``python -m macromax_amd.build``.
"""
import logging

from . import backend, bound, utils
from .grid import Grid
from .matrix import ScatteringMatrix
from .solver import Solution, solve

log = logging.getLogger(__name__)

__version__ = '0.1.0'

__all__ = ['__version__', 'solve', 'Solution', 'ScatteringMatrix', 'Grid', 'log', 'backend', 'bound', 'utils']
