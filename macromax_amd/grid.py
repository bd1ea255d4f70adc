"""A uniformly spaced plaid Cartesian sampling grid (the subset of MacroMax's Grid the solver uses).

Mirrors ``utils/ft/grid.py`` of the reference: ``Grid(shape, step, center=..., first=...)``,
``Grid.from_ranges(*ranges)`` (``:201-236``), open-grid axis vectors ``grid[axis]`` (``:532-563``), the
frequency grids ``f`` and ``k`` in ifftshifted (numpy ``fftfreq``) order (``:426-439``) and scaling by a
number (``:453-473``, used to express lengths in units of 1/k0).
"""
from __future__ import annotations

from collections.abc import Sequence

import numpy as np

__all__ = ['Grid']


class Grid(Sequence):
    """Immutable plaid grid. ``center`` is the coordinate at index ``shape // 2`` of each axis."""

    def __init__(self, shape=None, step=None, *, center=None, first=None, extent=None,
                 origin_at_center=True, ndim: int | None = None):
        vals = [v for v in (shape, step, center, first, extent) if v is not None]
        if ndim is None:
            ndim = max([np.asarray(v).size for v in vals] + [0])
        self.__multidimensional = any(not np.isscalar(v) for v in vals)
        if shape is None:
            if extent is None:
                raise ValueError('Grid needs a shape or an extent.')
            step = 1 if step is None else step
            shape = np.maximum(1, np.ceil(np.asarray(extent, dtype=float) / np.asarray(step, dtype=float))).astype(int)
        shape = np.broadcast_to(np.asarray(shape, dtype=int), (ndim,)).copy()
        if np.any(shape < 1):
            raise ValueError(f'shape {shape} must be at least 1 on every axis.')
        if step is None:
            step = np.asarray(extent) / shape if extent is not None else 1
        step = np.broadcast_to(np.asarray(step), (ndim,)).copy()
        if first is not None:
            first = np.broadcast_to(np.asarray(first), (ndim,))
            center = first + (shape // 2) * step
        if center is None:
            center = np.zeros(ndim, dtype=step.dtype)
        center = np.broadcast_to(np.asarray(center), (ndim,)).copy()
        dtype = (step[0] + center[0]).dtype if ndim > 0 else float
        self._shape = shape
        self._step = step.astype(dtype)
        self._center = center.astype(dtype)
        self._origin_at_center = np.broadcast_to(np.asarray(origin_at_center, dtype=bool), (ndim,)).copy()

    @staticmethod
    def from_ranges(*ranges) -> 'Grid':
        """Grid from one or more uniformly spaced (possibly ifftshifted) coordinate vectors (``grid.py:201-236``)."""
        ranges = [np.arange(r.start, r.stop, r.step) if isinstance(r, slice) else r for r in ranges]
        ranges = [np.array([r] if np.isscalar(r) else r) for r in ranges]
        if any(r.size < 1 for r in ranges):
            raise AttributeError('All input ranges should have at least one element.')
        ranges = [(r.swapaxes(0, axis).reshape(r.shape[axis], -1)[:, 0] if r.ndim > 1 else r)
                  for axis, r in zip(range(-len(ranges), 0), ranges)]
        shape = np.array([r.size for r in ranges])
        singleton = shape <= 1
        odd = np.mod(shape, 2) == 1
        first = np.array([r[0] for r in ranges])
        before_center = np.array([r[int((r.size - 1) / 2)] for r in ranges])
        after_center = np.array([r[-int(r.size / 2)] for r in ranges])
        last = np.array([r[-1] for r in ranges])
        origin_at_center = np.abs(last - first) >= np.abs(before_center - after_center)
        extent_m1 = origin_at_center * (last - first) + (1 - origin_at_center) * (before_center - after_center)
        step = extent_m1 / (shape - 1 + singleton)
        center = origin_at_center * (odd * before_center + (1 - odd) * after_center) + (1 - origin_at_center) * first
        g = Grid(shape=shape, step=step, center=center, origin_at_center=origin_at_center)
        return g

    # properties
    @property
    def ndim(self) -> int:
        return int(self._shape.size)

    @property
    def shape(self) -> np.ndarray:
        return self._shape.copy()

    @property
    def step(self) -> np.ndarray:
        return self._step.copy()

    @property
    def center(self) -> np.ndarray:
        return self._center.copy()

    @property
    def first(self) -> np.ndarray:
        return self._center - self._step * (self._shape // 2)

    @property
    def extent(self) -> np.ndarray:
        return self._shape * self._step

    @property
    def size(self) -> int:
        return int(np.prod(self._shape))

    @property
    def multidimensional(self) -> bool:
        return self.__multidimensional

    @property
    def immutable(self) -> 'Grid':
        return self

    @property
    def origin_at_center(self) -> np.ndarray:
        return self._origin_at_center.copy()

    # Sequence interface: open-grid axis vectors
    def __len__(self) -> int:
        return self.ndim

    def __getitem__(self, key):
        if isinstance(key, slice):
            return [self[a] for a in range(self.ndim)[key]]
        axis = int(key)
        if axis < -self.ndim or axis >= self.ndim:
            raise IndexError(f'axis {key} out of range for a {self.ndim}-dimensional grid')
        axis %= self.ndim
        n = int(self._shape[axis])
        rng = self._center[axis] + self._step[axis] * (np.arange(n) - n // 2)
        if not self._origin_at_center[axis]:
            rng = np.fft.ifftshift(rng)
        return rng.reshape([-1 if a == axis else 1 for a in range(self.ndim)])

    def __array__(self, dtype=None, copy=None):
        if self.ndim == 1:
            return np.asarray(self[0].ravel(), dtype=dtype)
        return np.asarray([np.broadcast_to(self[a], self.shape) for a in range(self.ndim)], dtype=dtype)

    # frequency grids (grid.py:426-439)
    @property
    def f(self) -> 'Grid':
        with np.errstate(divide='ignore'):
            step = 1.0 / self.extent
        return Grid(shape=self._shape, step=step, center=np.zeros(self.ndim), origin_at_center=False)

    @property
    def k(self) -> 'Grid':
        return self.f * (2 * np.pi)

    # arithmetic (grid.py:453-473)
    def __mul__(self, factor) -> 'Grid':
        if isinstance(factor, Grid):
            raise TypeError('A Grid cannot be multiplied with a Grid.')
        factor = np.asarray(factor)
        return Grid(shape=self._shape, step=self._step * factor, center=self._center * factor,
                    origin_at_center=self._origin_at_center)

    __rmul__ = __mul__

    def __truediv__(self, denominator) -> 'Grid':
        return self * (1 / np.asarray(denominator))

    def __eq__(self, other) -> bool:
        return isinstance(other, Grid) and self.ndim == other.ndim and np.all(self._shape == other._shape) \
            and np.allclose(self._step, other._step) and np.allclose(self._center, other._center) \
            and np.all(self._origin_at_center == other._origin_at_center)

    def __hash__(self):
        return hash((tuple(self._shape), tuple(np.round(self._step, 12))))

    def __repr__(self) -> str:
        return f'Grid(shape={self._shape.tolist()}, step={self._step.tolist()}, center={self._center.tolist()})'
