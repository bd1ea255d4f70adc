"""The import path ``macromax.utils.ft`` (MacroMax python/macromax/utils/ft/__init__.py) for the sampling grids:
``Grid`` and ``MutableGrid``, the same classes as ``macromax_amd.grid``."""
import logging

from .grid import Grid, MutableGrid

log = logging.getLogger(__name__)

__all__ = ['Grid', 'MutableGrid']
