"""Uniformly spaced plaid Cartesian sampling grids (MacroMax ``utils/ft/grid.py``: ``Grid`` and ``MutableGrid``).

A grid is defined per axis by its number of samples, spacing and central coordinate, plus three per-axis flags:
``flat`` (``grid[axis]`` as a flat vector or as an open-grid array along that axis), ``origin_at_center`` (monotonous
coordinates, or ifftshifted so that the central sample comes first) and ``center_at_index`` (for an even number of
samples, whether the central coordinate is a sample, index shape // 2, or lies half a step before it). Whatever is
not given is derived from the rest as the reference does (``grid.py:17-199``): shape from extent, first, last or center
and the step; the step from the extent, or from two of first, center and last; the center from first or last; an
integer step or center stays integer. Grids defined with a scalar behave as one sequence of coordinates; grids
defined with any vector behave as a sequence of axis vectors (``multidimensional``). The frequency grids ``f`` and ``k``
are in the order of numpy's ``fftfreq`` (``:426-439``); ``Grid.from_ranges`` recovers a grid from coordinate vectors
(``:201-236``).
"""
from __future__ import annotations

import warnings
from collections.abc import Sequence

import numpy as np

__all__ = ['Grid', 'MutableGrid']


def _index_offset(shape, center_at_index):
    """Index distance from the first sample to the central coordinate: shape // 2 when the center is a sample,
    (shape - 1) / 2 otherwise."""
    return shape // 2 * center_at_index + (shape - 1) / 2 * (1 - center_at_index)


def _keep_integer(value, reference, dtype_from):
    """``value`` cast to ``dtype_from``'s type when an integer-valued result of arguments of another type became
    floating point only through the division (so integer grids stay integer)."""
    if reference.dtype != value.dtype and np.allclose(value, np.round(value)):
        return value.astype(dtype_from.dtype)
    return value


class Grid(Sequence):
    """An immutable plaid grid; see the module documentation for the parameters."""

    def __init__(self, shape=None, step=None, *, extent=None, first=None, center=None, last=None, include_last=False,
                 ndim: int | None = None, flat=False, origin_at_center=True, center_at_index=True):
        given = (shape, step, extent, first, center, last)
        if ndim is None:
            ndim = max([np.asarray(v).size for v in given if v is not None] + [0])
        self.__ndim = int(ndim)
        self.__multidimensional = any(v is not None and not np.isscalar(v) for v in given)
        vec = self._to_ndim
        shape, step, extent, first, center, last = (vec(v) for v in given)
        flat, origin_at_center, include_last, center_at_index = (vec(v) for v in (flat, origin_at_center,
                                                                                  include_last, center_at_index))

        if shape is None:
            shape, step, extent, center = self.__derive_shape(step, extent, first, center, last, include_last,
                                                              center_at_index)
        shape = np.maximum(1, np.ceil(shape).astype(int))
        if step is None:
            step = self.__derive_step(shape, extent, first, center, last, include_last, center_at_index)
        if center is None:
            center = self.__derive_center(shape, step, first, last, include_last, center_at_index)

        if extent is not None and not np.allclose(extent / step, shape) \
                and np.any(np.maximum(1, np.ceil(extent / step)) != shape):
            raise ValueError(f'Extent {extent} and step {step} are not compatible with shape {shape}: '
                             f'extent / step = {extent / step}.')
        if last is not None and first is not None and np.any(shape * step != last + step * include_last - first):
            raise ValueError(f'First={first} and last={last} (include_last={include_last}) do not correspond to a step '
                             f'{step} and shape {shape}.')
        if center is not None and first is not None \
                and np.any(center != first + _index_offset(shape, center_at_index) * step):
            raise ValueError(f'First={first} and center={center} do not correspond to a step {step} and shape {shape} '
                             f'(center_at_index={center_at_index}).')
        if np.any(shape < 1):
            warnings.warn(f'shape = {shape}. All input ranges should have at least one element.')
            shape = np.maximum(1, shape)

        dtype = (step[0] + center[0]).dtype if self.__ndim > 0 else float
        self._shape = shape
        self._step = np.asarray(step).astype(dtype)
        self._center = np.asarray(center).astype(dtype)
        self._flat = flat
        self._origin_at_center = origin_at_center
        self._center_at_index = center_at_index

    # ---- derivation of the unspecified parameters (grid.py:96-179)
    def __derive_shape(self, step, extent, first, center, last, include_last, center_at_index):
        if extent is not None:
            if step is None:
                step = extent
            extent = np.sign(step) * np.abs(extent)  # the extent takes the sign of the step
            return extent / step, step, extent, center
        if step is None:
            step = 1
        if last is None:
            if first is None:
                if center is None:
                    center = self._to_ndim(0)
                first = center  # one sample at the center
            if center is None:
                center = (first + 1 / step) % step - step / 2  # the sample position closest to 0
            shape = 2 * (center - first) / step - (1 - center_at_index)  # even when the center is a sample
        elif first is None:
            if center is None:
                center = (last + 1 / step) % step - step / 2
            shape = 2 * (last - center) / step + include_last - (1 - center_at_index)
        else:
            shape = (last - first) / step + include_last
        return shape, step, extent, center

    def __derive_step(self, shape, extent, first, center, last, include_last, center_at_index):
        if extent is not None:
            step = extent / shape
        elif last is None:
            if first is None or center is None:
                step = self._to_ndim(1)
            else:
                step = _keep_integer((center - first) / _index_offset(shape, center_at_index), center - first, center)
        elif first is None:
            if center is None:
                step = self._to_ndim(1)
            else:
                step = (last - center) / (shape - include_last - _index_offset(shape, center_at_index))
                step = _keep_integer(step, last - center, center)
        else:
            step = (last - first) / (shape - include_last) if np.all(shape > include_last) else np.ones(1)
            step = _keep_integer(step, last - first, first)
        if np.all(step == step.astype(int)):
            step = step.astype(int)
        return step

    def __derive_center(self, shape, step, first, last, include_last, center_at_index):
        if first is not None:
            center = first + _index_offset(shape, center_at_index) * step
            return _keep_integer(center, first + step, step)
        if last is not None:
            center = last - step * (shape - include_last - _index_offset(shape, center_at_index))
            return _keep_integer(center, last - step, step)
        return self._to_ndim(0)

    @staticmethod
    def from_ranges(*ranges) -> Grid:
        """The grid of one or more uniformly spaced coordinate vectors (monotonous, or ifftshifted); slices,
        scalars (singletons) and arrays with the axis in position -len(ranges) + i are accepted (``grid.py:201-236``)."""
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
        # a monotonous range spans more from its first to its last value than across its middle
        origin_at_center = np.abs(last - first) >= np.abs(before_center - after_center)
        extent_m1 = origin_at_center * (last - first) + (1 - origin_at_center) * (before_center - after_center)
        step = extent_m1 / (shape - 1 + singleton)
        center = origin_at_center * (odd * before_center + (1 - odd) * after_center) + (1 - origin_at_center) * first
        return Grid(shape=shape, step=step, center=center, flat=False, origin_at_center=origin_at_center)

    # ---- defining properties
    @property
    def ndim(self) -> int:
        return self.__ndim

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
    def center_at_index(self) -> np.ndarray:
        return self._center_at_index.copy()

    @property
    def flat(self) -> np.ndarray:
        return self._flat.copy()

    @property
    def origin_at_center(self) -> np.ndarray:
        return self._origin_at_center.copy()

    @property
    def multidimensional(self) -> bool:
        return self.__multidimensional

    # ---- derived properties
    @property
    def first(self) -> np.ndarray:
        result = self._center - self._step * (self._shape // 2)
        between = ~self._center_at_index.astype(bool) & (self._shape % 2 == 0)
        if np.any(between):
            half = self._step // 2 if np.all(self._step % 2 == 0) else self._step / 2
            result = result + between * half
        return result

    @property
    def extent(self) -> np.ndarray:
        return self._shape * self._step

    @property
    def size(self) -> int:
        return int(np.prod(self._shape))

    @property
    def dtype(self):
        return (self._step[0] + self._center[0]).dtype if self.__ndim > 0 else float

    # ---- conversions
    def _kwargs(self, **changes) -> dict:
        """Constructor arguments that reproduce this grid (scalars for a grid that is not multidimensional)."""
        d = dict(shape=self.shape, step=self.step, center=self.center, flat=self.flat,
                 center_at_index=self.center_at_index, origin_at_center=self.origin_at_center)
        if not self.__multidimensional:
            d = {k: v[0] for k, v in d.items()}
        d.update(changes)
        return d

    @property
    def as_flat(self) -> Grid:
        return Grid(**self._kwargs(flat=True))

    @property
    def as_non_flat(self) -> Grid:
        return Grid(**self._kwargs(flat=False))

    @property
    def as_origin_at_0(self) -> Grid:
        return Grid(**self._kwargs(origin_at_center=False))

    @property
    def as_origin_at_center(self) -> Grid:
        return Grid(**self._kwargs(origin_at_center=True))

    def swapaxes(self, axes) -> Grid:
        """The grid with the order of the given axes reversed."""
        axes = np.array(axes).flatten()
        order = np.arange(self.__ndim)
        order[axes] = axes[::-1]
        return self.transpose(order)

    def transpose(self, axes=None) -> Grid:
        """The grid with its axes in the given order (default: reversed)."""
        if axes is None:
            axes = np.arange(self.__ndim - 1, -1, -1)
        return self.project(axes)

    def project(self, axes_to_keep=None, axes_to_remove=None) -> Grid:
        """The grid of the kept axes (in the given order), without the removed ones."""
        def as_list(axes, default):
            if axes is None:
                return list(default)
            if isinstance(axes, slice):
                return list(np.arange(self.__ndim)[axes])
            return [axes] if np.isscalar(axes) else list(axes)
        keep = np.array(as_list(axes_to_keep, range(self.__ndim)))
        if np.any(keep >= self.__ndim) or np.any(keep < -self.__ndim):
            raise IndexError(f'Axis range {keep} requested from a Grid of dimension {self.__ndim}.')
        remove = [a % self.__ndim for a in as_list(axes_to_remove, [])]
        keep = np.array([a % self.__ndim for a in keep if a % self.__ndim not in remove])
        if keep.size == 0:
            return Grid([])
        return Grid(shape=self.shape[keep], step=self.step[keep], center=self.center[keep], flat=self.flat[keep],
                    origin_at_center=self.origin_at_center[keep], center_at_index=self.center_at_index[keep])

    @property
    def immutable(self) -> Grid:
        return Grid(**self._kwargs())

    @property
    def mutable(self) -> MutableGrid:
        return MutableGrid(**self._kwargs())

    # ---- frequency grids (grid.py:426-439)
    @property
    def f(self) -> Grid:
        """The frequency grid: spacing 1 / extent, zero frequency first (numpy fftfreq order)."""
        with np.errstate(divide='ignore'):
            step = 1 / self.extent
        shape, flat = self.shape, self.flat
        if not self.__multidimensional:
            shape, step, flat = shape[0], step[0], flat[0]
        return Grid(shape=shape, step=step, flat=flat, origin_at_center=False, center_at_index=True)

    @property
    def k(self) -> Grid:
        """The angular frequency grid, 2 pi f."""
        return self.f * (2 * np.pi)

    # ---- arithmetic
    def __add__(self, term) -> Grid:
        center = self.center + np.asarray(term)
        return Grid(**self._kwargs(center=center if self.__multidimensional else center[0]))

    def __sub__(self, term) -> Grid:
        return self + (-np.asarray(term))

    def __mul__(self, factor) -> Grid:
        if isinstance(factor, Grid):
            raise TypeError("A Grid object can't be multiplied with a Grid object. "
                            "Use matmul @ to determine the tensor space.")
        factor = np.asarray(factor)
        step, center = self.step * factor, self.center * factor
        if not self.__multidimensional:
            step, center = step[0], center[0]
        return Grid(**self._kwargs(step=step, center=center))

    def __rmul__(self, factor) -> Grid:
        return self * factor

    def __truediv__(self, denominator) -> Grid:
        return self * (1 / np.asarray(denominator))

    def __neg__(self) -> Grid:
        return self * -1

    def __matmul__(self, other: Grid) -> Grid:
        """The grid of the tensor-product space: this grid's axes, then the other's."""
        return Grid(shape=(*self.shape, *other.shape), step=(*self.step, *other.step),
                    center=(*self.center, *other.center), flat=(*self.flat, *other.flat),
                    origin_at_center=(*self.origin_at_center, *other.origin_at_center),
                    center_at_index=(*self.center_at_index, *other.center_at_index))

    # ---- sequence interface (grid.py:522-572)
    def __len__(self) -> int:
        return self.__ndim if self.__multidimensional else int(self._shape[0])

    def _axis_vector(self, axis: int) -> np.ndarray:
        n = int(self._shape[axis])
        rng = np.arange(n) - n // 2
        if n > 1:  # a singleton's coordinate is its center, whatever the step (which may be inf)
            rng = rng * self._step[axis]
        rng = rng + self._center[axis]
        if not self._center_at_index[axis] and n % 2 == 0:
            rng = rng + self._step[axis] / 2
        if not self._origin_at_center[axis]:
            rng = np.fft.ifftshift(rng)
        if not self._flat[axis]:
            rng = rng.reshape([-1 if a == axis else 1 for a in range(self.__ndim)])
        return rng

    def __getitem__(self, key):
        """The coordinate vector of an axis (or a list of them for a slice or sequence of axes); for a grid that is
        not multidimensional, the coordinates themselves."""
        scalar_key = np.isscalar(key)
        n = self.__ndim if self.__multidimensional else int(self._shape[0])
        indices = np.atleast_1d(np.arange(n)[key])
        out = []
        for idx in indices.ravel():
            if self.__multidimensional:
                if idx >= self.__ndim:
                    raise IndexError(f'Axis range {idx} requested from a Grid of dimension {self.__ndim}.')
                out.append(self._axis_vector(int(idx)))
            else:
                out.append(self._axis_vector(0)[idx])
        return out[0] if scalar_key else out

    def __iter__(self):
        for idx in range(len(self)):
            yield self[idx]

    def __array__(self, dtype=None, copy=None):
        """A grid that is not multidimensional converts to its coordinates; a multidimensional one to the stacked
        coordinate arrays of all axes, broadcast to the grid's shape."""
        if not self.__multidimensional:
            return np.asarray(self._axis_vector(0).ravel(), dtype=dtype)
        return np.asarray([np.broadcast_to(self[a], self.shape) for a in range(self.__ndim)], dtype=dtype)

    # ---- object protocol
    def __str__(self) -> str:
        return f"{type(self).__name__}({','.join(f'{k}={v}' for k, v in self._kwargs().items())})"

    def __repr__(self) -> str:
        d = self._kwargs()
        d['dtype'] = self.dtype
        return f"{type(self).__name__}({','.join(f'{k}={v!r}' for k, v in d.items())})"

    def __eq__(self, other) -> bool:
        return type(self) == type(other) and np.all(self.shape == other.shape) and np.all(self.step == other.step) \
            and np.all(self.center == other.center) and np.all(self.flat == other.flat) \
            and np.all(self.center_at_index == other.center_at_index) \
            and np.all(self.origin_at_center == other.origin_at_center) and self.dtype == other.dtype

    def __hash__(self) -> int:
        return hash(repr(self))

    def _to_ndim(self, arg):
        """``arg`` as a vector of ndim elements (a scalar or one element repeated); None stays None."""
        if arg is None:
            return None
        arg = np.array(arg).flatten()
        if arg.size == 1:
            return np.repeat(arg, repeats=self.__ndim)
        if arg.size != self.__ndim:
            raise ValueError(f'All input arguments should be scalar or of length {self.__ndim}, not {arg.size} as '
                             f'{arg}.')
        return arg


class MutableGrid(Grid):
    """A grid whose shape, step, center, first, flags and dtype can be set (``grid.py:652-765``)."""

    def __init__(self, shape=None, step=None, extent=None, first=None, center=None, last=None, include_last=False,
                 ndim: int | None = None, flat=False, origin_at_center=True, center_at_index=True):
        super().__init__(shape=shape, step=step, extent=extent, first=first, center=center, last=last,
                         include_last=include_last, ndim=ndim, flat=flat, origin_at_center=origin_at_center,
                         center_at_index=center_at_index)

    @property
    def shape(self) -> np.ndarray:
        return self._shape.copy()

    @shape.setter
    def shape(self, value):
        if value is not None:
            self._shape = self._to_ndim(value)

    @property
    def step(self) -> np.ndarray:
        return self._step.copy()

    @step.setter
    def step(self, value):
        self._step = self._to_ndim(value)
        self._center = self._center.astype(self.dtype)

    @property
    def center(self) -> np.ndarray:
        return self._center.copy()

    @center.setter
    def center(self, value):
        self._center = self._to_ndim(value).astype(self.dtype)

    @property
    def flat(self) -> np.ndarray:
        return self._flat.copy()

    @flat.setter
    def flat(self, value):
        self._flat = self._to_ndim(value)

    @property
    def origin_at_center(self) -> np.ndarray:
        return self._origin_at_center.copy()

    @origin_at_center.setter
    def origin_at_center(self, value):
        self._origin_at_center = self._to_ndim(value)

    @property
    def first(self) -> np.ndarray:
        return Grid.first.fget(self)

    @first.setter
    def first(self, value):
        self._center = self._center + self._to_ndim(value) - self.first

    @property
    def dtype(self):
        return (self._step[0] + self._center[0]).dtype

    @dtype.setter
    def dtype(self, new_type):
        self._step = self._step.astype(new_type)
        self._center = self._center.astype(new_type)

    def __iadd__(self, number):
        self.center = self.center + np.asarray(number)
        return self

    def __isub__(self, number):
        self.center = self.center - np.asarray(number)
        return self

    def __imul__(self, number):
        self.step = self.step * np.asarray(number)
        self.center = self.center * np.asarray(number)
        return self

    def __itruediv__(self, number):
        self.step = self.step / np.asarray(number)
        self.center = self.center / np.asarray(number)
        return self
