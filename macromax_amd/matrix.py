"""Scattering, transfer and deposition matrices of a medium (MacroMax ``python/macromax/matrix.py``).

Same classes, methods, vector layout and normalisation as the reference:

- ``Matrix`` / ``SquareMatrix`` / ``LiteralScatteringMatrix`` (``matrix.py:100-353``): matrices as
  ``scipy.sparse.linalg.LinearOperator`` objects that convert to ``numpy.ndarray``;
- ``ScatteringMatrix`` (``:356-844``): the matrix of a medium between two source planes; input and output vectors
  hold the propagating plane-wave modes as ``[side (front, back), mode (raster order over the transverse k grid,
  negative to positive), polarisation (1 scalar, 2 vectorial: rotated y', z')]``;
- the quarter matrices (``:850-917``), ``DepositionMatrix`` (``:920-989``), ``inv`` and ``convert`` (``:996-1055``).

Where the reference solves one right-hand side after the other (``ScatteringMatrix._matvec``, ``:770-820``, and
``__array__``, ``:829-844``), this module builds the sources of all requested columns on the device, solves them in
one ``BatchSolution`` (every problem follows the iterates a ``Solution`` would produce for it) and converts the
detected fields of all columns together. The plane-wave bookkeeping (``srcvec2freespace`` ``:528-610``,
``srcvec2source`` ``:612-648``, ``detfield2detvec`` ``:682-753``) runs on the device too; its transverse Fourier
transforms use this package's HIP FFT (``mmx_fft``) on a plan of the transverse grid.

The default stop rule (iteration < 1e4 and residue > 1e-6, per column as in the reference, whose ``Solution``
restarts its iteration counter for every solve) is applied on the device; a user ``callback`` selects the
reference's column-by-column ``Solution.solve`` loop instead of the batched solve.
"""
from __future__ import annotations

from numbers import Complex
from typing import Callable, Optional, Sequence, Union

import numpy as np
import scipy.constants as const
import torch
from scipy.sparse.linalg import LinearOperator

from .bound import Bound, PeriodicBound
from .grid import Grid
from .solver import Solution

__all__ = ['CachingMatrix', 'Matrix', 'SquareMatrix', 'LiteralScatteringMatrix', 'ScatteringMatrix',
           'QuarterMatrix', 'ForwardTransmissionMatrix', 'FrontReflectionMatrix', 'BackReflectionMatrix',
           'BackwardTransmissionMatrix', 'DepositionMatrix', 'inv', 'convert']

array_like = Union[Complex, Sequence, np.ndarray, LinearOperator]


def _as_columns(right) -> tuple[np.ndarray, tuple]:
    """A vector or matrix as a 2D array of columns, with the original shape."""
    right = np.asarray(right)
    shape = right.shape
    return right.reshape(-1, right.shape[-1] if right.ndim > 1 else 1), shape


class CachingMatrix:
    """Column cache for matrices whose products are expensive (``matrix.py:23-97``): the product with a multiple
    of basis vector i is the cached column i times that multiple."""

    def __init__(self, caching: bool = True):
        self.__caching = bool(caching)
        self.__columns: dict[int, np.ndarray] = {}

    @property
    def caching(self) -> bool:
        """Use cached columns for products with (multiples of) basis vectors."""
        return self.__caching

    @caching.setter
    def caching(self, new_value: bool):
        self.__caching = bool(new_value)

    def _store_column(self, index: int, column: np.ndarray):
        self.__columns[int(index)] = np.array(column, copy=True)

    def _cached_product(self, right: np.ndarray, columns_function: Callable[[np.ndarray], np.ndarray],
                        nrows: int, dtype) -> np.ndarray:
        """``self @ right`` for a 2D ``right``: zero columns give zeros, cached basis columns come from the cache,
        and all other columns go to ``columns_function`` (2D in, 2D out) in one call."""
        out = np.zeros((nrows, right.shape[1]), dtype=dtype)
        todo = []
        for j in range(right.shape[1]):
            nz = np.flatnonzero(right[:, j])
            if nz.size == 0:
                continue
            if self.__caching and nz.size == 1 and int(nz[0]) in self.__columns:
                out[:, j] = self.__columns[int(nz[0])] * right[nz[0], j]
            else:
                todo.append(j)
        if todo:
            res = columns_function(right[:, todo])
            out[:, todo] = res
            if self.__caching:
                for i, j in enumerate(todo):
                    nz = np.flatnonzero(right[:, j])
                    if nz.size == 1:
                        self._store_column(nz[0], res[:, i] / right[nz[0], j])
        return out


class Matrix(LinearOperator):
    """A rectangular matrix from an array (or computed by a subclass's ``_matvec``), ``matrix.py:100-226``."""

    def __init__(self, array: Optional[array_like] = None, shape: Optional[Sequence[int]] = None, dtype=np.complex128):
        if shape is None and array is not None:
            shape = np.asarray(array).shape
        super().__init__(shape=tuple(int(_) for _ in shape), dtype=np.dtype(dtype))
        self.__array = array

    def __len__(self) -> int:
        return self.shape[0]

    def __getitem__(self, item) -> np.ndarray:
        """Rows (one index) or elements (two indices; only the selected columns are computed)."""
        if not isinstance(item, tuple):
            item = (item, slice(None))
        new_dims, slices, dropped = [], [], 0
        for pos, s in enumerate(item):
            if s is np.newaxis:
                new_dims.append(pos - dropped)
            else:
                slices.append(s)
                if np.isscalar(s):
                    dropped += 1
        if len(slices) > 2:
            raise TypeError(f'Cannot index {len(slices)} in a 2 dimensional Matrix!')
        if len(slices) > 1:
            arr = (self @ np.eye(self.shape[1])[:, slices[1]])[slices[0]]
            return np.expand_dims(arr, new_dims)
        return self.__array__()[item]

    def __setitem__(self, key, value):
        self.__array.__setitem__(key, value)

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        if self.__array is not None:
            arr = np.asarray(self.__array)
        else:
            arr = np.asarray(self @ np.eye(self.shape[1], dtype=self.dtype))
        return arr if dtype is None else arr.astype(dtype)

    def __array_ufunc__(self, ufunc, method, *args, **kwargs):
        return getattr(ufunc, method)(self.__array__(), *args[1:], **kwargs)

    def __array_function__(self, func, types, args, kwargs):
        def to_numpy(arg):
            if isinstance(arg, Matrix):
                return arg.__array__()
            if isinstance(arg, (list, tuple)):
                return type(arg)(to_numpy(_) for _ in arg)
            return arg
        return func(*(to_numpy(_) for _ in args), **{k: to_numpy(v) for k, v in kwargs.items()})

    def _matvec(self, right: array_like) -> np.ndarray:
        return np.asarray(self.__array) @ right

    def _matmat(self, right: array_like) -> np.ndarray:
        return self._matvec(right)

    def _adjoint(self):
        return Matrix(self.__array__().conj().T, dtype=self.dtype)

    def __repr__(self) -> str:
        return f'{self.__class__.__name__}({self.__array__()}, side={self.shape[0]}, dtype={self.dtype})'

    def __str__(self) -> str:
        return f'{self.__class__.__name__}(side={self.shape[0]}, dtype={self.dtype})'

    def inv(self, noise_level: float = 0.0) -> np.ndarray:
        """The (Tikhonov-regularised) inverse."""
        return inv(self, noise_level=noise_level)


class SquareMatrix(Matrix):
    """A square matrix (``matrix.py:229-255``)."""

    def __init__(self, array: Optional[array_like] = None, side: Optional[int] = None, dtype=np.complex128):
        if array is not None:
            array = np.asarray(array)
            if side is None:
                side = array.shape[0]
            if array.ndim < 2 or array.shape[0] != side or array.shape[1] != side:
                raise ValueError(f'array should be {side}x{side} square but it has shape {array.shape}.')
        super().__init__(array=array, shape=(side, side), dtype=dtype)

    @property
    def side(self) -> int:
        return self.shape[0]

    @property
    def size(self) -> int:
        return self.shape[0] * self.shape[1]


class LiteralScatteringMatrix(SquareMatrix):
    """A scattering matrix given as numbers, with its transfer matrix and quarters (``matrix.py:258-353``)."""

    def transfer(self, noise_level: float = 0.0) -> np.ndarray:
        """The transfer matrix (``convert``), optionally Tikhonov-regularised for the given noise level."""
        return convert(self, noise_level=noise_level)

    @property
    def forward_transmission(self) -> 'ForwardTransmissionMatrix':
        return ForwardTransmissionMatrix(self)

    @property
    def front_reflection(self) -> 'FrontReflectionMatrix':
        return FrontReflectionMatrix(self)

    @property
    def back_reflection(self) -> 'BackReflectionMatrix':
        return BackReflectionMatrix(self)

    @property
    def backward_transmission(self) -> 'BackwardTransmissionMatrix':
        return BackwardTransmissionMatrix(self)


class _Modes:
    """Propagating plane-wave modes of a grid and their polarisation frames (``matrix.py:451-473, 558-572``).

    ``index[j]``: raveled (fftfreq-ordered) transverse k index of mode j; ``kdir``: [3, M] unit k-vectors of the
    forward modes; ``radial``: [3, M] unit vectors orthogonal to k in the (axial, k) plane, pointing against the
    propagation axis; ``radial_mode``: [2, M] unit transverse directions of k (zero for normal incidence)."""

    def __init__(self, grid: Grid, k0: float, dtype):
        nd = grid.ndim
        tshape = tuple(int(_) for _ in grid.shape[1:])
        k_t = [_centred(np.asarray(grid.k[a]).ravel()) for a in range(1, nd)]
        tol = np.sqrt(np.finfo(np.float64 if dtype is None else dtype).eps)
        propagating = np.ones(tshape, dtype=bool)
        k_rem = np.full(tshape, k0)
        for a, k in enumerate(k_t):
            kb = k.reshape([-1 if b == a else 1 for b in range(nd - 1)])
            propagating = propagating & (np.abs(kb) < k_rem - tol)
            k_rem = np.sqrt(np.maximum(k_rem ** 2 - kb ** 2, 0.0))
        centred = np.flatnonzero(propagating.ravel())
        if nd > 1:
            sub = np.array(np.unravel_index(centred, tshape))
            half = (np.array(tshape) // 2)[:, None]
            self.index = np.ravel_multi_index((sub - half) % np.array(tshape)[:, None], tshape)
        else:
            self.index = centred  # the single mode along the axis
        M = self.index.size
        kdir = np.zeros((3, M))
        for a in range(1, nd):
            ka = np.broadcast_to(np.asarray(grid.k[a]).reshape([-1 if b == a - 1 else 1 for b in range(nd - 1)]),
                                 tshape).ravel()
            kdir[a] = ka[self.index] / k0
        kdir[0] = np.sqrt(np.maximum(1.0 - np.sum(kdir[1:] ** 2, axis=0), 0.0))
        self.kdir = kdir
        axial = np.array([1.0, 0.0, 0.0])[:, None]
        self.radial = _unit(kdir[0] * kdir - axial)
        self.radial_mode = _unit(kdir[1:])
        self.tshape = tshape

    @property
    def count(self) -> int:
        return int(self.index.size)


def _centred(v: np.ndarray) -> np.ndarray:
    """A 1D array in fftfreq order reordered from the most negative frequency up (numpy's fftshift)."""
    return np.roll(v, v.shape[0] // 2)


def _unit(v: np.ndarray) -> np.ndarray:
    """Unit vectors along axis 0 (zero vectors stay zero), ``matrix.py:518-521``."""
    n = np.linalg.norm(v, axis=0)
    return v / (n + (n == 0))


class ScatteringMatrix(LiteralScatteringMatrix):
    """The scattering matrix of a medium (``matrix.py:356-844``).

    Constructor arguments as the reference's (``:357-449``); ``tol`` and ``max_iter`` state its default stop rule
    (``callback=lambda s: s.iteration < 1e4 and s.residue > 1e-6``) for the batched solve. ``batch_size`` limits
    the number of columns solved together (default: as many as fit in about 16 GiB of device memory)."""

    def __init__(self, grid: Union[Grid, Sequence, np.ndarray], vectorial: Optional[bool] = True,
                 wavenumber: Optional[float] = None, angular_frequency: Optional[float] = None,
                 vacuum_wavelength: Optional[float] = None, epsilon: Optional[array_like] = None,
                 xi: Optional[array_like] = 0.0, zeta: Optional[array_like] = 0.0, mu: Optional[array_like] = 1.0,
                 refractive_index: Optional[array_like] = None, bound: Bound = None, dtype=None,
                 callback: Optional[Callable] = None, caching: bool = True, array: Optional[array_like] = None,
                 tol: float = 1e-6, max_iter: int = 10000, batch_size: Optional[int] = None):
        if not isinstance(grid, Grid):
            if np.isscalar(grid[0]):
                grid = [grid]
            grid = Grid.from_ranges(*grid)
        self.__grid = grid
        self.__vectorial = bool(vectorial)
        self.__callback = callback
        self.__tol = float(tol)
        self.__max_iter = int(max_iter)
        self.__solution = Solution(grid=grid, vectorial=self.__vectorial,
                                   wavenumber=wavenumber if wavenumber is not None else 1.0,
                                   angular_frequency=angular_frequency, vacuum_wavelength=vacuum_wavelength,
                                   epsilon=epsilon, xi=xi, zeta=zeta, mu=mu, refractive_index=refractive_index,
                                   bound=bound, dtype=dtype)
        sol = self.__solution
        self.__k0 = sol.wavenumber
        self.__modes = _Modes(grid, self.__k0, dtype)
        if bound is None:
            bound = PeriodicBound(grid)
        thick = np.asarray(bound.thickness)
        step0 = float(grid.step[0])
        n0 = int(grid.shape[0])
        # source planes just inside the boundary layers, detection planes a wavelength further in (matrix.py:475-486)
        self.__source_planes = (int(thick[0, 0] / step0) + 1, n0 - int(thick[0, -1] / step0) - 1)
        front_det = int((thick[0, 0] + sol.wavelength) / step0) + 1
        back_det = n0 - int((thick[0, -1] + sol.wavelength) / step0) - 1
        self.__detector_planes = (back_det, front_det)  # outputs of the front side (forward) are detected at the back
        self.__x0 = np.asarray(grid[0]).ravel()
        self.__ncomp = 1 + 2 * self.__vectorial
        self.__npol = 1 + self.__vectorial
        side = 2 * self.__npol * self.__modes.count
        super().__init__(side=side, dtype=np.dtype(sol.dtype))
        self.__cache = CachingMatrix(caching)
        if array is not None:
            array = np.asarray(array)
            if array.shape != (side, side):
                raise ValueError(f'Number of modes {side} is different from the number of rows or columns of the '
                                 f'specified matrix {array.shape}!')
            for j in range(side):
                self.__cache._store_column(j, array[:, j])
        self.__device = sol.backend.device
        self.__cdtype = sol.backend.hardware_dtype
        self.__tplan = None
        if grid.ndim > 1 and int(np.prod(self.__modes.tshape)) > 1:
            from .backend.hip import Plan
            self.__tplan = Plan(self.__modes.tshape, np.asarray(grid.step[1:]) * self.__k0, self.__cdtype,
                                self.__device)
        V = int(np.prod(grid.shape))
        if batch_size is None:
            per_problem = 6 * self.__ncomp * V * torch.empty((), dtype=self.__cdtype).element_size()
            batch_size = max(1, (16 << 30) // per_problem)
        self.__batch_size = int(batch_size)

    # ------------------------------------------------------------------ properties
    @property
    def grid(self) -> Grid:
        return self.__grid

    @property
    def vectorial(self) -> bool:
        return self.__vectorial

    @property
    def solution(self) -> Solution:
        """The ``Solution`` of the medium (its chi, alpha and beta are shared by all columns)."""
        return self.__solution

    @property
    def caching(self) -> bool:
        return self.__cache.caching

    @caching.setter
    def caching(self, new_value: bool):
        self.__cache.caching = new_value

    @property
    def source_planes(self) -> tuple:
        """Indices along axis 0 of the front and back source planes."""
        return self.__source_planes

    @property
    def detector_planes(self) -> tuple:
        """Indices along axis 0 of the back and front detection planes (outputs of the front and back inputs)."""
        return self.__detector_planes

    # ------------------------------------------------------------------ device helpers
    def _const(self, a: np.ndarray) -> torch.Tensor:
        return torch.as_tensor(np.asarray(a), device=self.__device).to(self.__cdtype)

    def _vectors(self, vecs) -> torch.Tensor:
        """Source vectors [B, N] -> [B, 2 (side), M, P] on the device."""
        v = torch.as_tensor(np.asarray(vecs), device=self.__device).to(self.__cdtype)
        return v.reshape(v.shape[0], 2, self.__modes.count, self.__npol)

    def _transverse_ft(self, x: torch.Tensor, direction: int) -> torch.Tensor:
        """FFT over the transverse axes of [..., T] (forward unnormalised: -1; inverse with 1/T: +1)."""
        if self.__tplan is None:
            return x
        x = x.contiguous()
        out = torch.empty_like(x)
        return self.__tplan.fft(x, out, direction)

    def _freespace_planes(self, v: torch.Tensor, planes: Sequence[int]) -> torch.Tensor:
        """Free-space field of source vectors v [B, 2, M, P] at the given axis-0 planes, [B, C, nplanes, T]
        (``srcvec2freespace``, ``matrix.py:528-610``, evaluated on selected planes)."""
        md = self.__modes
        B = v.shape[0]
        M = md.count
        T = int(np.prod(md.tshape)) if md.tshape else 1
        kdir = md.kdir
        amp = v / self._const(np.sqrt(kdir[0]))[None, None, :, None]  # angle of incidence: brighter oblique waves
        amp = amp.permute(0, 1, 3, 2)  # [B, side, P, M]
        first_t = np.asarray(self.__grid.first[1:], dtype=float)
        origin = np.exp(1j * self.__k0 * np.einsum('i...,i...->...', kdir[1:self.__grid.ndim], first_t[:, None]))
        x0 = self.__x0[list(planes)]
        prop = np.exp(1j * self.__k0 * kdir[0][None, :] * x0[:, None])  # [nplanes, M]
        rev = torch.as_tensor(md.index[::-1].copy(), device=self.__device)
        out = torch.zeros((B, self.__ncomp, len(planes), T), dtype=self.__cdtype, device=self.__device)
        for side, direction in ((0, 1), (1, -1)):
            vs = amp[:, side]  # [B, P, M]
            if direction < 0:
                vs = vs.conj()
            if self.__vectorial:
                rm = self._const(md.radial_mode)  # [2, M]
                rn = torch.sum(vs * rm[None], dim=1)  # [B, M]
                f = rn[:, None, :] * self._const(md.radial)[None]  # [B, 3, M]
                f = f.clone()
                f[:, 1:] += vs - rn[:, None, :] * rm[None]
            else:
                f = vs  # [B, 1, M]
            f = f * self._const(origin)[None, None, :]
            f = f[:, :, None, :] * self._const(prop)[None, None]  # [B, C, nplanes, M]
            grid_ift = torch.zeros((B, self.__ncomp, len(planes), T), dtype=self.__cdtype, device=self.__device)
            grid_ift[..., rev] = f
            field = self._transverse_ft(grid_ift, -1)
            if direction < 0:
                field = field.conj()
            out += field
        return out

    def _sources(self, v: torch.Tensor) -> torch.Tensor:
        """Source distributions [B, C, *grid] of source vectors v [B, 2, M, P] (``srcvec2source``, :612-648)."""
        md = self.__modes
        sol = self.__solution
        current_amp = 1.0 / (const.c * const.mu_0 * self.__grid.step[0] / 2)  # unit E-field plane waves
        source_amp = (-1.0j * sol.angular_frequency * const.mu_0) * current_amp
        scale = self._const(md.kdir[0] * source_amp)  # oblique waves: smaller current for the same field
        v = v * scale[None, None, :, None]
        B = v.shape[0]
        shape = tuple(int(_) for _ in self.__grid.shape)
        out = torch.zeros((B, self.__ncomp, *shape), dtype=self.__cdtype, device=self.__device)
        for side in range(2):
            mask = torch.zeros(2, dtype=self.__cdtype, device=self.__device)
            mask[side] = 1
            plane = self.__source_planes[side]
            f = self._freespace_planes(v * mask[None, :, None, None], [plane])[:, :, 0]
            out[:, :, plane] = f.reshape(B, self.__ncomp, *shape[1:])
        return out

    def _detvecs(self, det: torch.Tensor) -> torch.Tensor:
        """Detection vectors [B, N] of the outward fields at the (back, front) detection planes, det [B, C, 2, T]
        (``detfield2detvec``, ``matrix.py:682-753``)."""
        md = self.__modes
        B = det.shape[0]
        det = det.clone()
        det[:, :, 1] = det[:, :, 1].conj()  # the backward field, conjugated so that the transform flips directions
        ft = self._transverse_ft(det, +1)  # [B, C, 2, T]
        rev = torch.as_tensor(md.index[::-1].copy(), device=self.__device)
        f = ft[..., rev]  # [B, C, 2, M]
        kdir = self._const(md.kdir)[:, None, :]  # [3, 1, M]
        if self.__vectorial:
            f = f - torch.sum(f * kdir[None], dim=1, keepdim=True) * kdir[None]  # keep the transverse part
            radial = self._const(md.radial)[:, None, :]
            radial_mode = self._const(md.radial_mode)[:, None, :]
            rn = torch.sum(f * radial[None], dim=1)  # [B, 2, M]
            azim = f[:, 1:] - rn[:, None] * radial[None, 1:]
            vec = azim + rn[:, None] * radial_mode[None]  # [B, 2 (pol), 2 (side), M]
        else:
            vec = f  # [B, 1, 2, M]
        first_t = np.asarray(self.__grid.first[1:], dtype=float)
        origin = np.exp(-1j * self.__k0 * np.einsum('i...,i...->...', md.kdir[1:self.__grid.ndim], first_t[:, None]))
        dist = self.__x0[list(self.__detector_planes)]
        back = np.exp(-1j * self.__k0 * md.kdir[0][None, :] * dist[:, None])  # [2, M]
        fac = origin[None, :] * back * np.sqrt(md.kdir[0])[None, :]
        vec = vec * self._const(fac)[None, None]
        vec = vec.permute(0, 2, 3, 1).contiguous()  # [B, side, M, P]
        vec[:, 1] = vec[:, 1].conj()
        return vec.reshape(B, -1)

    def _incident_at_detectors(self, v: torch.Tensor) -> torch.Tensor:
        """Free-space fields of the inputs at the detection planes they cross first, [B, C, 2, T]: the back input
        at the back plane, the front input at the front plane (``matrix.py:792-798``)."""
        out = []
        for d, plane in enumerate(self.__detector_planes):
            mask = torch.ones(2, dtype=self.__cdtype, device=self.__device)
            mask[d] = 0  # detector 0 (back) sees the back input (side 1) incident on it, detector 1 the front input
            out.append(self._freespace_planes(v * mask[None, :, None, None], [plane])[:, :, 0])
        return torch.stack(out, dim=2)

    def _solve_sources(self, sources: torch.Tensor) -> torch.Tensor:
        """Fields [B, C, *grid] (SI units, device) for source distributions [B, C, *grid] of this medium."""
        from .batch import BatchSolution
        sol = self.__solution
        if self.__callback is None:
            outs = []
            for b0 in range(0, sources.shape[0], self.__batch_size):
                chunk = sources[b0:b0 + self.__batch_size]
                batch = BatchSolution.from_sources(sol, sources=chunk)
                batch.solve(tol=self.__tol, max_iter=self.__max_iter)
                outs.append(batch.field_arrays()[:, :, 0] / (self.__k0 ** 2))
            return torch.cat(outs)
        outs = []
        for b in range(sources.shape[0]):  # the reference's loop (matrix.py:676-680) for a user stop rule
            sol.source_distribution = sources[b]
            sol.E = 0
            sol.solve(self.__callback)
            outs.append(sol.field_array[:, 0] / (self.__k0 ** 2))
        return torch.stack(outs)

    def _columns(self, right: np.ndarray) -> np.ndarray:
        """``self @ right`` for columns that need a solve: right [N, K] -> [N, K]."""
        outs = []
        for k0 in range(0, right.shape[1], self.__batch_size):
            v = self._vectors(right[:, k0:k0 + self.__batch_size].T)
            fields = self._solve_sources(self._sources(v))  # [B, C, *grid]
            B = fields.shape[0]
            det = fields[:, :, list(self.__detector_planes)].reshape(B, self.__ncomp, 2, -1)
            det = det - self._incident_at_detectors(v)
            outs.append(self._detvecs(det).cpu().numpy().T)
        return np.concatenate(outs, axis=1).astype(self.dtype, copy=False)

    # ------------------------------------------------------------------ reference API
    def srcvec2freespace(self, input_vector: array_like) -> np.ndarray:
        """Superposition of free-space plane waves over the whole volume, [C, *grid] (``matrix.py:528-610``)."""
        v = self._vectors(np.asarray(input_vector).reshape(1, -1))
        f = self._freespace_planes(v, range(int(self.__grid.shape[0])))[0]
        return f.reshape(self.__ncomp, *self.__grid.shape).cpu().numpy().astype(self.dtype, copy=False)

    def srcvec2source(self, input_vector: array_like, out: np.ndarray = None) -> np.ndarray:
        """Source distribution on the front and back source planes, [C, *grid] (``matrix.py:612-648``)."""
        src = self._sources(self._vectors(np.asarray(input_vector).reshape(1, -1)))[0].cpu().numpy()
        if out is None:
            return src.astype(self.dtype, copy=False)
        out[...] = src
        return out

    def source2detfield(self, source: array_like, out: np.ndarray = None) -> np.ndarray:
        """Field in the whole volume for a source distribution, [C, *grid] (``matrix.py:650-669``)."""
        sol = self.__solution
        sol.source_distribution = source
        sol.E = 0
        if self.__callback is None:
            sol.solve(lambda s: s.iteration < self.__max_iter and s.residue > self.__tol)
        else:
            sol.solve(self.__callback)
        if out is None:
            return sol.E
        out[...] = sol.E
        return out

    def srcvec2detfield(self, input_vector: array_like, out: np.ndarray = None) -> np.ndarray:
        """Field in the whole volume for a source vector (``matrix.py:671-680``)."""
        return self.source2detfield(self.srcvec2source(input_vector), out=out)

    def detfield2detvec(self, field: array_like) -> np.ndarray:
        """Detection vector of an outward-propagating field given in the whole volume (``matrix.py:682-753``)."""
        f = torch.as_tensor(np.asarray(field), device=self.__device).to(self.__cdtype)
        f = f.reshape(self.__ncomp, *self.__grid.shape)
        det = f[:, list(self.__detector_planes)].reshape(1, self.__ncomp, 2, -1)
        return self._detvecs(det)[0].cpu().numpy().astype(self.dtype, copy=False)

    def detvec2srcvec(self, vec: array_like) -> np.ndarray:
        """Time-reversed source vector of a detection vector: sides swapped, values conjugated (``:755-765``)."""
        vec = np.asarray(vec).conj()
        for axis in range(vec.ndim):
            vec = np.roll(vec, vec.shape[axis] // 2, axis=axis)
        return vec

    def _matvec(self, right: array_like) -> np.ndarray:
        right, shape = _as_columns(right)
        if self.shape[1] != right.shape[0]:
            raise TypeError(f'Incorrect number of modes specified for this {self.shape} scattering matrix.')
        out = self.__cache._cached_product(right, self._columns, self.shape[0], self.dtype)
        return out.reshape(shape)

    def _matmat(self, right: array_like) -> np.ndarray:
        return self._matvec(right)

    def __setitem__(self, key, value):
        raise KeyError('The values of this ScatteringMatrix are calculated from the scattering and cannot be changed '
                       'directly. Use a Matrix object instead.')

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        """All columns (batched solves of the basis vectors)."""
        arr = self._matvec(np.eye(self.shape[1], dtype=self.dtype))
        return arr if dtype is None else arr.astype(dtype)


class QuarterMatrix(SquareMatrix):
    """A quarter of a scattering matrix (``matrix.py:850-884``)."""

    def __init__(self, matrix: SquareMatrix, backwards_output: bool, backwards_input: bool):
        if matrix.shape[0] % 2 > 0:
            raise TypeError(f'A QuarterMatrix object can only be constructed using a Matrix with an even number of '
                            f'rows and columns. The specified matrix has dimensions {matrix.shape}.')
        side = matrix.shape[0] // 2
        super().__init__(side=side, dtype=matrix.dtype)
        self.__matrix = matrix
        self.__rows = slice(side, None) if backwards_output else slice(None, side)
        self.__cols = slice(side, None) if backwards_input else slice(None, side)

    @property
    def full_matrix(self) -> SquareMatrix:
        return self.__matrix

    def _matvec(self, vec: array_like) -> np.ndarray:
        vec = np.asarray(vec)
        cols, shape = _as_columns(vec)
        full = np.zeros((2 * self.shape[1], cols.shape[1]), dtype=np.result_type(cols.dtype, self.dtype))
        full[self.__cols] = cols
        res = np.asarray(self.__matrix @ full)[self.__rows]
        return res.reshape(shape)

    def _matmat(self, right):
        return self._matvec(right)

    def __array__(self, dtype=None, copy=None) -> np.ndarray:
        arr = np.asarray(self.__matrix)[self.__rows, self.__cols]
        return arr if dtype is None else arr.astype(dtype)


class ForwardTransmissionMatrix(QuarterMatrix):
    """Light from negative infinity transmitted to positive infinity."""

    def __init__(self, scattering_matrix: SquareMatrix):
        super().__init__(scattering_matrix, False, False)


class FrontReflectionMatrix(QuarterMatrix):
    """Light from negative infinity reflected back (``matrix.py:897-903``)."""

    def __init__(self, scattering_matrix: SquareMatrix):
        super().__init__(scattering_matrix, True, False)


class BackReflectionMatrix(QuarterMatrix):
    """Light from positive infinity reflected back (``matrix.py:906-912``)."""

    def __init__(self, scattering_matrix: SquareMatrix):
        super().__init__(scattering_matrix, False, True)


class BackwardTransmissionMatrix(QuarterMatrix):
    """Light from positive infinity transmitted to negative infinity."""

    def __init__(self, scattering_matrix: SquareMatrix):
        super().__init__(scattering_matrix, True, True)


class DepositionMatrix(Matrix, CachingMatrix):
    """Free-space input vectors to any linear function of the field inside the medium (``matrix.py:920-989``).

    ``input_operator`` maps input vectors to source distributions (default: ``srcvec2source``), ``output_operator``
    maps the raveled field to outputs (default: ``detfield2detvec``; a plain function is wrapped). The columns of
    one product are solved together."""

    def __init__(self, scattering_matrix: ScatteringMatrix, input_operator: Optional[LinearOperator] = None,
                 output_operator: Union[LinearOperator, Callable[[array_like], np.ndarray], None] = None,
                 caching: bool = True):
        s = scattering_matrix
        n_field = (1 + 2 * s.vectorial) * s.grid.size
        if input_operator is None:
            input_operator = LinearOperator(shape=(n_field, s.shape[1]), dtype=s.dtype,
                                            matvec=lambda v: s.srcvec2source(v).ravel())
        if output_operator is None:
            output_operator = LinearOperator(
                shape=(s.shape[0], n_field), dtype=s.dtype,
                matvec=lambda f: s.detfield2detvec(np.asarray(f).reshape(1 + 2 * s.vectorial, *s.grid.shape)))
        elif not isinstance(output_operator, LinearOperator) and callable(output_operator):
            func = output_operator
            out_shape = np.shape(func(np.zeros(n_field, dtype=s.dtype)))
            output_operator = LinearOperator(shape=(int(np.prod(out_shape, dtype=int)), n_field), dtype=s.dtype,
                                             matvec=lambda f: np.ravel(func(f)))
        self.__s = s
        self.__in = input_operator
        self.__out = output_operator
        Matrix.__init__(self, shape=(output_operator.shape[0], input_operator.shape[1]), dtype=s.dtype)
        CachingMatrix.__init__(self, caching=caching)

    def _products(self, right: np.ndarray) -> np.ndarray:
        s = self.__s
        shape = (1 + 2 * s.vectorial, *[int(_) for _ in s.grid.shape])
        sources = np.stack([np.asarray(self.__in @ right[:, j]).reshape(shape) for j in range(right.shape[1])])
        src_t = torch.as_tensor(sources, device=s.solution.backend.device).to(s.solution.backend.hardware_dtype)
        fields = s._solve_sources(src_t).cpu().numpy()
        return np.stack([np.asarray(self.__out @ fields[j].ravel()).ravel() for j in range(fields.shape[0])], axis=1)

    def _matvec(self, right: array_like) -> np.ndarray:
        right = np.asarray(right)
        cols, _ = _as_columns(right)
        out = self._cached_product(cols, self._products, self.shape[0], self.dtype)
        return out.reshape((self.shape[0],) + right.shape[1:])

    def _matmat(self, right):
        return self._matvec(right)


# ---------------------------------------------------------------------- functions
def inv(mat: array_like, noise_level: float = 0.0) -> np.ndarray:
    """Matrix inverse, Tikhonov-regularised for ``noise_level`` > 0 (``matrix.py:996-1006``)."""
    mat = np.asarray(mat)
    if noise_level > 0.0:
        u, sigma, vh = np.linalg.svd(mat)
        sigma_inv = sigma.conj() / (np.abs(sigma) ** 2 + noise_level ** 2)
        return ((u * sigma_inv) @ vh).T.conj()
    return np.linalg.inv(mat)


def convert(s: array_like, noise_level: float = 0.0) -> np.ndarray:
    """Scattering matrix <-> transfer matrix (the map is its own inverse), ``matrix.py:1009-1055``.

    With quarters [[A, B], [C, D]] (rows: outputs forward/backward; columns: inputs forward/backward), the result is
    [[A - B D^-1 C, B D^-1], [-D^-1 C, D^-1]], D^-1 regularised for ``noise_level`` > 0."""
    s = np.asarray(s)
    orig = s.shape
    if s.ndim == 2:
        s = s.reshape(2, orig[0] // 2, 2, orig[1] // 2)
    t = np.empty_like(s)
    d_inv = inv(s[1, :, 1], noise_level)
    t[1, :, 1] = d_inv
    t[1, :, 0] = -d_inv @ s[1, :, 0]
    t[0, :, 1] = s[0, :, 1] @ d_inv
    t[0, :, 0] = s[0, :, 0] + s[0, :, 1] @ t[1, :, 0]
    return t.reshape(orig) if len(orig) == 2 else t
