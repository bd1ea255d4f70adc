"""The import path ``macromax.utils.ft.grid`` (MacroMax python/macromax/utils/ft/grid.py): the sampling grids of
``macromax_amd.grid``."""
from ...grid import Grid, MutableGrid

__all__ = ['Grid', 'MutableGrid']
