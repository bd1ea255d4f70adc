"""Calculation-volume boundaries (MacroMax ``bound.py``): periodic, and absorbing with a linear kappa ramp.

``AbsorbingBound.extinction`` follows ``bound.py:159-183`` and ``electric_susceptibility`` ``:186-193``:
chi_bound = (n_bg + i kappa)^2 - n_bg^2, added to the permittivity before the iteration (``solver.py:242-247``).
"""
from __future__ import annotations

import numpy as np

from .grid import Grid

__all__ = ['Electric', 'Magnetic', 'Bound', 'PeriodicBound', 'AbsorbingBound', 'LinearBound']


class Electric:
    """Mixin: the bound has a non-zero electric susceptibility."""

    @property
    def background_permittivity(self):
        return 0.0

    @property
    def electric_susceptibility(self) -> np.ndarray:
        return NotImplemented

    @property
    def permittivity(self) -> np.ndarray:
        return self.background_permittivity + self.electric_susceptibility


class Magnetic:
    """Mixin: the bound has a non-zero magnetic susceptibility (not used by the solver in this package)."""

    @property
    def magnetic_susceptibility(self) -> np.ndarray:
        return NotImplemented

    @property
    def permeability(self) -> np.ndarray:
        return 1.0 + self.magnetic_susceptibility


class Bound:
    def __init__(self, grid=None, thickness=0.0, background_permittivity=1.0):
        if not isinstance(grid, Grid):
            grid = Grid.from_ranges(grid)
        self.__grid = grid
        self.__thickness = np.broadcast_to(thickness, (grid.ndim, 2)).astype(float)
        self.__background_permittivity = background_permittivity

    @property
    def grid(self) -> Grid:
        return self.__grid

    @property
    def thickness(self) -> np.ndarray:
        """Thickness per ``[axis, front/back]``."""
        return self.__thickness.copy()

    @property
    def background_permittivity(self):
        return self.__background_permittivity

    @property
    def electric_susceptibility(self) -> np.ndarray:
        return np.zeros(self.grid.shape)

    @property
    def magnetic_susceptibility(self) -> np.ndarray:
        return np.zeros(self.grid.shape)

    @property
    def between(self) -> np.ndarray:
        result = np.asarray(True)
        for axis in range(self.grid.ndim):
            rng = self.grid[axis]
            inside = (rng.ravel()[0] + self.thickness[axis, 0] <= rng) & (rng <= rng.ravel()[-1] - self.thickness[axis, 1])
            result = np.logical_and(result, inside)
        return result

    @property
    def beyond(self) -> np.ndarray:
        return np.logical_not(self.between)


class PeriodicBound(Bound):
    def __init__(self, grid):
        super().__init__(grid=grid, thickness=0.0)


class AbsorbingBound(Bound, Electric):
    def __init__(self, grid, thickness=0.0, extinction_coefficient_function=lambda rel_depth: rel_depth,
                 background_permittivity=1.0):
        super().__init__(grid=grid, thickness=thickness, background_permittivity=background_permittivity)
        self.__kappa_functions = np.broadcast_to(np.asarray(extinction_coefficient_function, dtype=object),
                                                 (self.grid.ndim, 2))

    @property
    def is_electric(self) -> bool:
        return True

    @property
    def extinction(self) -> np.ndarray:
        """kappa on the plaid grid; the maximum of overlapping ramps at corners (``bound.py:159-183``)."""
        kappa = 0.0
        for axis in range(self.grid.ndim):
            rng = self.grid[axis]
            sgn = np.sign(self.grid.step[axis])
            for back_side in range(2):
                t = self.thickness[axis, back_side] * sgn
                depth = (rng - (rng.ravel()[-1] - t)) if back_side else ((rng.ravel()[0] + t) - rng)
                depth = depth * sgn
                in_boundary = depth > 0
                if np.any(in_boundary):
                    rel_depth = in_boundary * depth / t
                    kappa = np.maximum(kappa, self.__kappa_functions[axis, back_side](rel_depth) * in_boundary)
        return kappa

    @property
    def electric_susceptibility(self) -> np.ndarray:
        n = np.lib.scimath.sqrt(self.background_permittivity)
        return (n + 1j * self.extinction) ** 2 - self.background_permittivity


class LinearBound(AbsorbingBound):
    """Absorbing bound whose kappa rises linearly to ``max_extinction_coefficient`` at the outer edge."""

    def __init__(self, grid, thickness=0.0, max_extinction_coefficient=0.25, background_permittivity=1.0):
        self._max_extinction_coefficient = max_extinction_coefficient
        kmax = np.asarray(max_extinction_coefficient, dtype=float)
        funcs = np.empty(kmax.shape, dtype=object)
        for idx in np.ndindex(kmax.shape):
            funcs[idx] = (lambda km: (lambda rel_depth: km * rel_depth))(float(kmax[idx]))
        super().__init__(grid=grid, thickness=thickness, extinction_coefficient_function=funcs,
                         background_permittivity=background_permittivity)
