"""MI355X backend: device memory and streams from PyTorch-ROCm, compute in libmacromax_hip.so.

``BackEndHip`` implements the ``BackEnd`` interface on ``torch.Tensor``s resident on a ROCm device. The operations
the Born iteration needs run as gfx950 kernels through the C ABI (``include/macromax_hip.h``):

==============================  ========================================  =====================================
method                          kernel (C ABI)                            reference
==============================  ========================================  =====================================
``ft`` / ``ift``                ``mmx_fft`` (Stockham pencil passes)       ``torch.py:183-205``, ``numpy.py:112-133``
``mul`` (3x3 / iso chi field)   ``mmx_chi_mul``                           ``backend/__init__.py:415-444``
``longitudinal_projection_ft``  ``mmx_longitudinal_projection_ft``        ``backend/__init__.py:606-649``
``norm``                        ``mmx_norm2`` (FP64 accumulation)         ``backend/__init__.py:872-874``
``born_step``                   ``mmx_born_step`` (fused iteration)       ``solver.py:1070-1085``
==============================  ========================================  =====================================

Set-up and readout helpers (``astype``, ``assign``, ``subtract``, ``ldivide``, ``mat3_eigh`` ...) are PyTorch tensor
plumbing on the same device; they are not on the iteration's path. Without a ROCm device or without the built
library the constructor raises ``ImportError`` and ``backend.load`` moves on (it then fails: there is no other
backend in this package).
"""
from __future__ import annotations

import ctypes
import logging
import math
from numbers import Complex

import numpy as np

from . import BackEnd
from ..grid import Grid

log = logging.getLogger(__name__)

try:
    import torch
except ImportError as _e:  # pragma: no cover - torch is a hard requirement of this backend
    raise ImportError('BackEndHip needs PyTorch') from _e

from .. import _native

__all__ = ['BackEndHip', 'Plan']


def _torch_dtype(dtype):
    if dtype in (np.complex64, torch.complex64):
        return torch.complex64
    if dtype in (np.complex128, complex, torch.complex128):
        return torch.complex128
    if dtype in (np.float32, torch.float32):
        return torch.float32
    if dtype in (np.float64, float, torch.float64):
        return torch.float64
    if dtype in (bool, np.bool_, torch.bool):
        return torch.bool
    if dtype in (int, np.int64, torch.int64):
        return torch.int64
    if isinstance(dtype, torch.dtype):
        return dtype
    return torch.from_numpy(np.zeros(1, dtype=dtype)).dtype


def _stream_ptr(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


class Plan:
    """Owner of one ``mmx_plan_t``: twiddles and k-vectors for a grid on one device."""

    def __init__(self, shape, step, complex_dtype, device):
        self.shape = tuple(int(_) for _ in shape)
        self.ndim = len(self.shape)
        self.dtype = _torch_dtype(complex_dtype)
        self.real_dtype = torch.float32 if self.dtype == torch.complex64 else torch.float64
        self.device = torch.device(device)
        self.V = int(np.prod(self.shape))
        L = _native.lib()
        handle = ctypes.c_void_p()
        shp = (ctypes.c_int64 * self.ndim)(*self.shape)
        stp = (ctypes.c_double * self.ndim)(*[float(_) for _ in step])
        code = _native.MMX_C64 if self.dtype == torch.complex64 else _native.MMX_C128
        with torch.cuda.device(self.device):
            _native.check(L.mmx_plan_create(ctypes.byref(handle), self.ndim, shp, stp, code), 'mmx_plan_create')
        self.handle = handle
        self._lib = L

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            try:
                self._lib.mmx_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self.handle = None

    def set_param(self, name: str, value: int):
        """Set a tuning parameter (see include/macromax_hip.h, MMX_PARAM_*); 0 = automatic."""
        _native.check(self._lib.mmx_plan_set_param(self.handle, _native.PARAMS[name], int(value)), 'mmx_plan_set_param')

    def set_source_support(self, lo=None, hi=None):
        """Box [lo, hi) outside which the sources of later Born calls are zero (None: the whole grid)."""
        key = None if lo is None else (tuple(int(_) for _ in lo), tuple(int(_) for _ in hi))
        if key == getattr(self, '_source_box', None):
            return
        if key is None:
            rc = self._lib.mmx_plan_set_source_support(self.handle, None, None)
        else:
            lo_a = (ctypes.c_int64 * self.ndim)(*key[0])
            hi_a = (ctypes.c_int64 * self.ndim)(*key[1])
            rc = self._lib.mmx_plan_set_source_support(self.handle, lo_a, hi_a)
        _native.check(rc, 'mmx_plan_set_source_support')
        self._source_box = key

    def get_param(self, name: str) -> int:
        v = ctypes.c_int64()
        _native.check(self._lib.mmx_plan_get_param(self.handle, _native.PARAMS[name], ctypes.byref(v)), 'mmx_plan_get_param')
        return int(v.value)

    def _check(self, t, what):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f'{what}: expected a torch.Tensor')
        if t.device != self.device:
            raise ValueError(f'{what}: tensor on {t.device}, plan on {self.device}')
        if t.dtype != self.dtype:
            raise TypeError(f'{what}: dtype {t.dtype}, plan dtype {self.dtype}')
        if not t.is_contiguous():
            raise ValueError(f'{what}: tensor must be contiguous')

    def fft(self, x, out, direction: int):
        self._check(x, 'fft input')
        self._check(out, 'fft output')
        nbatch = x.numel() // self.V
        if nbatch * self.V != x.numel() or out.numel() != x.numel():
            raise ValueError('fft: array size is not a multiple of the grid size')
        _native.check(self._lib.mmx_fft(self.handle, x.data_ptr(), out.data_ptr(), nbatch, int(direction),
                                        _stream_ptr(self.device)), 'mmx_fft')
        return out

    def born_step(self, E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms,
                  s_box=None, noalloc=False):
        """One Born iteration (``mmx_born_step``; ``noalloc``: ``mmx_born_step_noalloc``, which keeps its norm
        partials in the library's static device array and allocates nothing)."""
        for t, w in ((E, 'E'), (Eout, 'Eout'), (s, 's'), (chi, 'chi')):
            self._check(t, w)
        if coef.dtype != self.real_dtype or coef.device != self.device or coef.numel() < 6 * nbatch:
            raise ValueError('born_step: coef must hold 6 reals of the plan precision per problem')
        if norms.dtype != torch.float64 or norms.numel() < 2 * nbatch or norms.device != self.device:
            raise ValueError('born_step: norms must be a float64 device tensor of 2*nbatch elements')
        if E.numel() != nbatch * ncomp * self.V or Eout.numel() != E.numel():
            raise ValueError('born_step: E / Eout size mismatch')
        chi_size = {1: 1, 2: 9, 3: 6, 4: 4}[int(chi_kind)] * self.V
        if chi.numel() != (chi_size if chi_bstride == 0 else chi_size * nbatch):
            raise ValueError('born_step: chi size mismatch')
        if s.numel() != (ncomp * self.V if s_bstride == 0 else ncomp * self.V * nbatch):
            raise ValueError('born_step: s size mismatch')
        self.set_source_support(*(s_box if s_box is not None else (None, None)))
        fn, name = ((self._lib.mmx_born_step_noalloc, 'mmx_born_step_noalloc') if noalloc else
                    (self._lib.mmx_born_step, 'mmx_born_step'))
        with torch.cuda.device(self.device):
            _native.check(fn(self.handle, E.data_ptr(), Eout.data_ptr(), s.data_ptr(), chi.data_ptr(), int(ncomp),
                             int(chi_kind), coef.data_ptr(), int(nbatch), int(s_bstride), int(chi_bstride),
                             norms.data_ptr(), _stream_ptr(self.device)), name)
        return norms

    def _born_checks(self, E, others, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride):
        for t, w in ((E, 'E'), (s, 's'), (chi, 'chi')) + tuple(others):
            self._check(t, w)
        if coef.dtype != self.real_dtype or coef.device != self.device or coef.numel() < 6 * nbatch:
            raise ValueError('coef must hold 6 reals of the plan precision per problem')
        if E.numel() != nbatch * ncomp * self.V:
            raise ValueError('E size mismatch')
        for t, w in others:
            if t.numel() != E.numel():
                raise ValueError(f'{w} size mismatch')
        chi_size = {1: 1, 2: 9, 3: 6, 4: 4}[int(chi_kind)] * self.V
        if chi.numel() != (chi_size if chi_bstride == 0 else chi_size * nbatch):
            raise ValueError('chi size mismatch')
        if s.numel() != (ncomp * self.V if s_bstride == 0 else ncomp * self.V * nbatch):
            raise ValueError('s size mismatch')

    def magnetic_iterate(self, E, Eout, s, chis, alpha: complex, work, norms, norm_stream=None, slot=0,
                         norms_host=None):
        """One magnetic Born iteration (``mmx_magnetic_iterate``): Eout = E + d, norms = (|d|^2, |Eout|^2).

        ``chis`` = (chi_EE, chi_HE, chi_EH, chi_HH), each a [1, 1, ...] scalar or [3, 3, ...] tensor field that is
        either constant (one element per entry) or spans the whole grid (contiguous). ``norm_stream``, ``slot``,
        ``norms_host``: the norm reduction on a second stream, as for ``born_iterate``
        (``mmx_magnetic_iterate_split``)."""
        for t, w in ((E, 'E'), (Eout, 'Eout'), (s, 's'), (work, 'work')):
            self._check(t, w)
        if E.numel() != 3 * self.V or Eout.numel() != E.numel() or s.numel() != E.numel():
            raise ValueError('magnetic_iterate: E, Eout and s must be [3, 1, *grid]')
        if work.numel() < 12 * self.V:
            raise ValueError('magnetic_iterate: the work buffer needs 12 grid-sized complex components')
        if norms.dtype != torch.float64 or norms.numel() < 2 or norms.device != self.device:
            raise ValueError('magnetic_iterate: norms must be a float64 device tensor of 2 elements')
        ptrs, kinds, strides = [], [], []
        for t in chis:
            self._check(t, 'chi')
            n = int(t.shape[0])
            if n not in (1, 3) or int(t.shape[1]) != n:
                raise ValueError('magnetic_iterate: susceptibilities are [1, 1, ...] or [3, 3, ...] fields')
            per = t.numel() // (n * n)
            if per not in (1, self.V):
                raise ValueError('magnetic_iterate: a susceptibility must be constant or span the grid')
            ptrs.append(t.data_ptr())
            kinds.append(1 if n == 1 else 2)
            strides.append(0 if per == 1 and self.V > 1 else 1)
        c_p = (ctypes.c_void_p * 4)(*ptrs)
        c_k = (ctypes.c_int * 4)(*kinds)
        c_s = (ctypes.c_int64 * 4)(*strides)
        c_a = (ctypes.c_double * 2)(float(alpha.real), float(alpha.imag))
        args = (self.handle, E.data_ptr(), Eout.data_ptr(), s.data_ptr(), c_p, c_k, c_s, c_a, work.data_ptr(),
                norms.data_ptr())
        with torch.cuda.device(self.device):
            if norm_stream is None:
                _native.check(self._lib.mmx_magnetic_iterate(*args, _stream_ptr(self.device)), 'mmx_magnetic_iterate')
            else:
                host = None if norms_host is None else self._pinned_norms(norms_host, 1)
                _native.check(self._lib.mmx_magnetic_iterate_split(*args, host, int(slot), _stream_ptr(self.device),
                                                                   norm_stream.cuda_stream),
                              'mmx_magnetic_iterate_split')
        return norms

    @staticmethod
    def _pinned_norms(norms_host, nbatch: int) -> int:
        if (norms_host.dtype != torch.float64 or norms_host.numel() < 2 * nbatch or norms_host.device.type != 'cpu'
                or not norms_host.is_pinned()):
            raise ValueError('norms_host must be a pinned float64 CPU tensor of 2*nbatch elements')
        return norms_host.data_ptr()

    def born_prologue(self, E, W, s, chi, ncomp, chi_kind, coef, nbatch=1, s_bstride=0, chi_bstride=0,
                      s_box=None):
        """W = FFT along the innermost axis of chi_eff E + s_eff (``mmx_born_prologue``)."""
        self._born_checks(E, ((W, 'W'),), s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride)
        self.set_source_support(*(s_box if s_box is not None else (None, None)))
        _native.check(self._lib.mmx_born_prologue(self.handle, E.data_ptr(), W.data_ptr(), s.data_ptr(),
                                                  chi.data_ptr(), int(ncomp), int(chi_kind), coef.data_ptr(),
                                                  int(nbatch), int(s_bstride), int(chi_bstride),
                                                  _stream_ptr(self.device)), 'mmx_born_prologue')
        return W

    def born_iterate(self, E, Eout, W, s, chi, ncomp, chi_kind, coef, norms, nbatch=1, s_bstride=0, chi_bstride=0,
                     s_box=None, norm_stream=None, slot=0, norms_host=None):
        """Finish the iteration begun in W (``mmx_born_iterate``): Eout = E + d, norms, W = next prologue.

        With ``norm_stream`` (a torch stream of the same device other than the current one), the norms are reduced
        on that stream once the current one has run the passes, from partial buffer ``slot`` (0 or 1) of the plan,
        and copied there to ``norms_host`` (a pinned float64 CPU tensor, optional) (``mmx_born_iterate_split``): a
        slot may be used again once the reduction of its previous call has run."""
        self._born_checks(E, ((Eout, 'Eout'), (W, 'W')), s, chi, ncomp, chi_kind, coef, nbatch, s_bstride,
                          chi_bstride)
        if norms.dtype != torch.float64 or norms.numel() < 2 * nbatch or norms.device != self.device:
            raise ValueError('norms must be a float64 device tensor of 2*nbatch elements')
        self.set_source_support(*(s_box if s_box is not None else (None, None)))
        args = (self.handle, E.data_ptr(), Eout.data_ptr(), W.data_ptr(), s.data_ptr(), chi.data_ptr(), int(ncomp),
                int(chi_kind), coef.data_ptr(), int(nbatch), int(s_bstride), int(chi_bstride), norms.data_ptr())
        if norm_stream is None:
            _native.check(self._lib.mmx_born_iterate(*args, _stream_ptr(self.device)), 'mmx_born_iterate')
        else:
            host = None if norms_host is None else self._pinned_norms(norms_host, nbatch)
            _native.check(self._lib.mmx_born_iterate_split(*args, host, int(slot), _stream_ptr(self.device),
                                                           norm_stream.cuda_stream), 'mmx_born_iterate_split')
        return norms

    def batch_step(self, EA, EB, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, state, track,
                   norms, tol, max_iter, s_box=None):
        """One device-decided iteration of every unfinished problem of a batch (``mmx_batch_step``)."""
        self._born_checks(EA, ((EB, 'EB'), (W, 'W')), s, chi, ncomp, chi_kind, coef, nbatch, s_bstride,
                          chi_bstride)
        if state.dtype != torch.int32 or state.numel() < nbatch or state.device != self.device:
            raise ValueError('state must be an int32 device tensor of nbatch elements')
        if track.dtype != torch.float64 or track.numel() < nbatch * _native.TRACK_LEN or track.device != self.device:
            raise ValueError('track must be a float64 device tensor of nbatch x TRACK_LEN elements')
        if norms.dtype != torch.float64 or norms.numel() < 2 * nbatch or norms.device != self.device:
            raise ValueError('norms must be a float64 device tensor of 2*nbatch elements')
        self.set_source_support(*(s_box if s_box is not None else (None, None)))
        _native.check(self._lib.mmx_batch_step(self.handle, EA.data_ptr(), EB.data_ptr(), W.data_ptr(),
                                               s.data_ptr(), chi.data_ptr(), int(ncomp), int(chi_kind),
                                               coef.data_ptr(), int(nbatch), int(s_bstride), int(chi_bstride),
                                               state.data_ptr(), track.data_ptr(), norms.data_ptr(), float(tol),
                                               int(max_iter), _stream_ptr(self.device)), 'mmx_batch_step')

    def _setup_eps(self, eps):
        if not isinstance(eps, torch.Tensor) or eps.dtype != torch.complex128 or eps.device != self.device:
            raise TypeError('set-up reductions take a complex128 device tensor')
        eps = eps.contiguous()
        if eps.numel() == self.V:
            return eps, 1
        if eps.numel() == 9 * self.V:
            return eps, 2
        raise ValueError('eps must be [1, 1, *grid] or [3, 3, *grid]')

    def setup_sigma(self, eps, shift: float) -> float:
        """max over voxels of the largest eigenvalue of (eps - shift I)^H (eps - shift I) (``mmx_setup_sigma``)."""
        eps, kind = self._setup_eps(eps)
        out = torch.empty(2, dtype=torch.float64, device=self.device)
        _native.check(self._lib.mmx_setup_sigma(self.handle, eps.data_ptr(), kind, float(shift), out.data_ptr(),
                                                _stream_ptr(self.device)), 'mmx_setup_sigma')
        return float(out[1])

    def setup_gain(self, eps):
        """(min, max) over voxels of the eigenvalues of -i/2 (eps - eps^H) (``mmx_setup_gain``)."""
        eps, kind = self._setup_eps(eps)
        out = torch.empty(2, dtype=torch.float64, device=self.device)
        _native.check(self._lib.mmx_setup_gain(self.handle, eps.data_ptr(), kind, out.data_ptr(),
                                               _stream_ptr(self.device)), 'mmx_setup_gain')
        lo, hi = out.tolist()
        return lo, hi

    def norm2(self, x, out):
        self._check(x, 'norm2 input')
        _native.check(self._lib.mmx_norm2(self.handle, x.data_ptr(), x.numel(), out.data_ptr(),
                                          _stream_ptr(self.device)), 'mmx_norm2')
        return out

    def chi_mul(self, chi, chi_kind, x, y, ncomp, nbatch):
        for t, w in ((chi, 'chi'), (x, 'x'), (y, 'y')):
            self._check(t, w)
        _native.check(self._lib.mmx_chi_mul(self.handle, chi.data_ptr(), int(chi_kind), x.data_ptr(), y.data_ptr(),
                                            int(ncomp), int(nbatch), _stream_ptr(self.device)), 'mmx_chi_mul')
        return y

    def longitudinal_projection(self, F, out):
        self._check(F, 'F')
        self._check(out, 'out')
        _native.check(self._lib.mmx_longitudinal_projection_ft(self.handle, F.data_ptr(), out.data_ptr(),
                                                               _stream_ptr(self.device)), 'mmx_longitudinal_projection_ft')
        return out


class BackEndHip(BackEnd):
    """BackEnd on MI355X (gfx950). ``hardware_dtype`` complex64 (default for the kernels) or complex128."""

    def __init__(self, nb_dims: int, grid: Grid, hardware_dtype=np.complex64, device: str = None):
        hardware_dtype = _torch_dtype(hardware_dtype)
        if hardware_dtype not in (torch.complex64, torch.complex128):
            raise ValueError(f'hardware_dtype must be complex64 or complex128, not {hardware_dtype}')
        super().__init__(nb_dims, grid, hardware_dtype)
        if not torch.cuda.is_available() or torch.cuda.device_count() < 1:
            raise ImportError('BackEndHip needs a ROCm GPU (torch.cuda.is_available() is False).')
        if device is None or str(device).lower() in ('cuda', 'gpu', 'hip'):
            device = f'cuda:{torch.cuda.current_device()}'
        self.__device = torch.device(str(device).lower().replace('gpu', 'cuda').replace('hip', 'cuda'))
        try:
            _native.lib()
        except OSError as e:
            raise ImportError(f'libmacromax_hip.so could not be loaded: {e}') from e
        self.__plan = Plan(grid.shape, grid.step, hardware_dtype, self.__device)
        self.__transposed_plan = None
        self.__k = None
        self.__k2 = None
        self.__eye = None
        self.__ik3 = None

    # ------------------------------------------------------------------ properties
    @property
    def device(self) -> torch.device:
        return self.__device

    @property
    def plan(self) -> Plan:
        return self.__plan

    @property
    def transposed_plan(self) -> Plan:
        """A plan of the transposed operator (adjoint solves by reciprocity), created on first use."""
        if self.__transposed_plan is None:
            grid = self.grid
            self.__transposed_plan = Plan(grid.shape, grid.step, self.hardware_dtype, self.__device)
            self.__transposed_plan.set_param('transposed', 1)
        return self.__transposed_plan

    @property
    def numpy_dtype(self):
        return np.complex64 if self.hardware_dtype == torch.complex64 else np.complex128

    @property
    def real_dtype(self):
        return torch.float32 if self.hardware_dtype == torch.complex64 else torch.float64

    @property
    def eps(self) -> float:
        return torch.finfo(self.hardware_dtype).eps

    @property
    def eye(self) -> torch.Tensor:
        if self.__eye is None:
            n = self.vector_length
            self.__eye = self.astype(np.eye(n).reshape((n, n, *([1] * self.grid.ndim))))
        return self.__eye

    @property
    def k(self):
        """k-vector components along each axis as open-grid tensors (``backend/__init__.py:651-656``)."""
        if self.__k is None:
            self.__k = [self.astype(np.asarray(self.grid.k[a]), dtype=self.real_dtype) for a in range(self.grid.ndim)]
        return self.__k

    @property
    def k2(self) -> torch.Tensor:
        if self.__k2 is None:
            self.__k2 = sum(_ ** 2 for _ in self.k)
        return self.__k2

    # ------------------------------------------------------------------ conversions and plumbing
    def astype(self, arr, dtype=None) -> torch.Tensor:
        dtype = self.hardware_dtype if dtype is None else _torch_dtype(dtype)
        if not isinstance(arr, torch.Tensor):
            a = np.asarray(arr)
            if dtype.is_complex and a.dtype.kind in 'biuf':
                a = a.astype(np.complex128 if dtype == torch.complex128 else np.complex64)
            arr = torch.as_tensor(a, device=self.__device)
        if arr.dtype != dtype or arr.device != self.__device:
            arr = arr.to(dtype=dtype, device=self.__device)
        return arr

    def asnumpy(self, arr) -> np.ndarray:
        if isinstance(arr, torch.Tensor):
            return arr.detach().cpu().resolve_conj().numpy()
        return np.asarray(arr)

    def allocate_array(self, shape=None, dtype=None, fill_value: Complex = None) -> torch.Tensor:
        if shape is None:
            shape = (self.vector_length, 1, *self.grid.shape)
        dtype = self.hardware_dtype if dtype is None else _torch_dtype(dtype)
        arr = torch.empty(tuple(int(_) for _ in shape), dtype=dtype, device=self.__device)
        if fill_value is not None:
            arr.fill_(fill_value)
        return arr

    def assign(self, arr, out) -> torch.Tensor:
        arr = self.to_matrix_field(arr)
        if tuple(arr.shape) != tuple(out.shape):
            arr = arr.expand(*out.shape) if arr.ndim == out.ndim else arr.expand_as(out)
        return self.assign_exact(arr, out)

    def assign_exact(self, arr, out) -> torch.Tensor:
        out[...] = self.astype(arr)
        return out

    def copy(self, arr) -> torch.Tensor:
        if isinstance(arr, torch.Tensor):
            return arr.detach().clone()
        return self.astype(arr)

    def ravel(self, arr) -> torch.Tensor:
        return self.astype(arr).reshape(-1)

    def sign(self, arr):
        return torch.sign(arr)

    def first(self, arr):
        return self.ravel(arr)[0]

    @staticmethod
    def expand_dims(arr, axis: int):
        return arr.unsqueeze(axis)

    def swapaxes(self, arr, ax_from: int, ax_to: int):
        return arr.transpose(ax_from, ax_to)

    def abs(self, arr):
        return self.astype(arr).abs()

    def conj(self, arr):
        return self.astype(arr).conj().resolve_conj()

    def real(self, arr):
        return arr.real

    def any(self, arr) -> bool:
        return bool(torch.any(self.astype(arr, dtype=torch.bool)))

    def allclose(self, arr, other=0.0) -> bool:
        arr = self.astype(arr)
        other = self.astype(other)
        return bool(torch.allclose(arr, other.expand_as(arr) if other.numel() == 1 else other))

    def amax(self, arr) -> float:
        arr = arr if isinstance(arr, torch.Tensor) else self.astype(arr)
        if arr.is_complex():
            arr = arr.real
        return float(torch.amax(arr.to(torch.float64)))

    def sort(self, arr):
        arr = self.astype(arr)
        return torch.sort(arr.real, dim=0).values.to(arr.dtype)

    def is_scalar(self, arr) -> bool:
        if np.isscalar(arr):
            return True
        shape = tuple(arr.shape)
        return len(shape) == 0 or (shape[0] == 1 and (len(shape) == 1 or shape[1] == 1))

    def is_vector(self, arr) -> bool:
        return arr.ndim == 2 + self.grid.ndim - 1

    def is_matrix(self, arr) -> bool:
        return arr.ndim == 2 + self.grid.ndim

    def to_matrix_field(self, arr) -> torch.Tensor:
        """Add leading singleton dims up to ``2 + ndim``; 1xN becomes Nx1 (``backend/__init__.py:355-385``)."""
        if arr is None:
            arr = 0
        arr = self.astype(arr)
        while arr.ndim < self.grid.ndim + 2:
            arr = arr.unsqueeze(0)
        if arr.shape[0] == 1 and arr.shape[1] > 1:
            arr = arr.transpose(0, 1)
        return arr

    def adjoint(self, mat):
        return self.astype(mat).transpose(0, 1).conj().resolve_conj()

    def subtract(self, left_term, right_term):
        if not self.is_scalar(left_term) or not self.is_scalar(right_term):
            if self.is_scalar(left_term):
                left_term = self.eye * self.to_matrix_field(left_term)
            if self.is_scalar(right_term):
                right_term = self.eye * self.to_matrix_field(right_term)
        return self.astype(left_term) - self.astype(right_term)

    def _full(self, t) -> bool:
        return isinstance(t, torch.Tensor) and tuple(t.shape[2:]) == tuple(int(_) for _ in self.grid.shape) \
            and t.is_contiguous() and t.dtype == self.hardware_dtype and t.device == self.__device

    def mul(self, left_factor, right_factor, out=None):
        """Point-wise matrix product (``backend/__init__.py:415-444``). Overwrites ``out`` when given."""
        if self.is_scalar(left_factor) or self.is_scalar(right_factor):
            lf = left_factor if np.isscalar(left_factor) else self.astype(left_factor)
            rf = right_factor if np.isscalar(right_factor) else self.astype(right_factor)
            if out is not None:
                if self._full(out) and self._full(lf) and lf.shape[0] == 1 and out.shape[1] == 1 \
                        and out.shape[0] in (1, 3):
                    return self.__plan.chi_mul(lf, 1, out, out, out.shape[0], 1)
                out *= lf
                return out
            if isinstance(lf, torch.Tensor) or isinstance(rf, torch.Tensor):
                return (lf if isinstance(lf, torch.Tensor) else self.astype(lf)) * \
                       (rf if isinstance(rf, torch.Tensor) else self.astype(rf))
            return self.astype(lf * rf)
        lf = self.astype(left_factor)
        rf = self.astype(right_factor)
        if self._full(lf) and self._full(rf) and tuple(lf.shape[:2]) == (3, 3) and tuple(rf.shape[:2]) == (3, 1):
            y = out if (out is not None and self._full(out) and tuple(out.shape[:2]) == (3, 1)) else \
                torch.empty_like(rf)
            return self.__plan.chi_mul(lf, 2, rf, y, 3, 1)
        lm = lf.movedim(0, -1).movedim(0, -1)
        rm = rf.movedim(0, -1).movedim(0, -1)
        res = torch.matmul(lm, rm).movedim(-1, 0).movedim(-1, 0)
        if out is not None:
            out[...] = res
            return out
        return res

    def ldivide(self, denominator, numerator=1.0):
        """A^-1 B per voxel (``torch.py:251-289``), on the device."""
        den = self.to_matrix_field(denominator)
        num = self.to_matrix_field(numerator)
        if self.is_scalar(den):
            return num / den
        dm = den.movedim(0, -1).movedim(0, -1)
        if self.is_scalar(num):
            inv = torch.linalg.inv(dm) if den.shape[0] == den.shape[1] else torch.linalg.pinv(dm)
            return (inv * num.reshape(num.shape[2:] + (1, 1)) if num.ndim > 2 else inv * num).movedim(-1, 0).movedim(-1, 0)
        nm = num.movedim(0, -1).movedim(0, -1)
        dm, nm = torch.broadcast_tensors(dm, nm) if dm.shape[-2:] == nm.shape[-2:] else (dm, nm)
        res = torch.linalg.solve(dm, nm)
        return res.movedim(-1, 0).movedim(-1, 0)

    # ------------------------------------------------------------------ Fourier transforms (HIP)
    def _field(self, arr) -> torch.Tensor:
        arr = self.astype(arr)
        full = tuple(int(_) for _ in self.grid.shape)
        if tuple(arr.shape[-self.grid.ndim:]) != full or arr.ndim < self.grid.ndim:
            arr = arr.expand(*arr.shape[:-self.grid.ndim], *full)
        return arr.contiguous()

    def ft(self, arr) -> torch.Tensor:
        """Unnormalised forward FFT over the spatial axes (``mmx_fft``, direction -1)."""
        x = self._field(arr)
        out = torch.empty_like(x)
        return self.__plan.fft(x, out, -1)

    def ift(self, arr) -> torch.Tensor:
        """Inverse FFT with 1/N over the spatial axes (``mmx_fft``, direction +1)."""
        x = self._field(arr)
        out = torch.empty_like(x)
        return self.__plan.fft(x, out, +1)

    def longitudinal_projection_ft(self, field_array_ft) -> torch.Tensor:
        F = self._field(field_array_ft)
        if F.shape[0] != 3:
            raise ValueError('longitudinal_projection_ft needs a 3-vector field [3, 1, *grid]')
        out = torch.empty_like(F)
        return self.__plan.longitudinal_projection(F, out)

    def transversal_projection_ft(self, field_array_ft) -> torch.Tensor:
        F = self._field(field_array_ft)
        return F - self.longitudinal_projection_ft(F)

    def longitudinal_projection(self, field_array) -> torch.Tensor:
        """The longitudinal (curl-free) part of a vector field [3, 1, *grid] (``backend/__init__.py:582-591``): the
        k-space projection between a forward and an inverse FFT (HIP kernels)."""
        return self.convolve(self.longitudinal_projection_ft, field_array)

    def transversal_projection(self, field_array) -> torch.Tensor:
        """The transverse (divergence-free) part of a vector field (``backend/__init__.py:571-580``)."""
        return self.convolve(self.transversal_projection_ft, field_array)

    def curl_ft(self, field_array_ft) -> torch.Tensor:
        """i k x F (``backend/__init__.py:483-509``).

        A vector field [3, 1, *grid] takes one cross-product kernel with the cached i k [3, 1, *grid] (components
        along absent axes are 0); other shapes take the component-wise form."""
        F = self.astype(field_array_ft)
        full = tuple(int(_) for _ in self.grid.shape)
        if F.ndim == 2 + len(full) and tuple(F.shape[:2]) == (3, 1) and tuple(F.shape[2:]) == full:
            if self.__ik3 is None:
                ik = torch.zeros((3, 1, *full), dtype=self.hardware_dtype, device=self.__device)
                for a in range(self.grid.ndim):
                    ik[a, 0] = 1j * self.k[a].to(self.hardware_dtype)
                self.__ik3 = ik
            return torch.linalg.cross(self.__ik3, F, dim=0)
        k = [1j * _.to(self.hardware_dtype) for _ in self.k] + [0] * (3 - self.grid.ndim)
        comps = [F[c, 0] if c < F.shape[0] else 0 for c in range(3)]
        out = self.allocate_array(shape=(3, *F.shape[1:]))
        out[0, 0] = k[1] * comps[2] - k[2] * comps[1]
        out[1, 0] = k[2] * comps[0] - k[0] * comps[2]
        out[2, 0] = k[0] * comps[1] - k[1] * comps[0]
        return out

    def curl(self, field_array) -> torch.Tensor:
        return self.ift(self.curl_ft(self.ft(field_array)))

    def div_ft(self, field_array_ft) -> torch.Tensor:
        """sum_d i k_d F[d] over the spatial axes, [m, n, *grid] -> [n, 1, *grid] (``backend/__init__.py:555-569``)."""
        F = self.astype(field_array_ft)
        out = torch.zeros(F.shape[1:], dtype=self.hardware_dtype, device=self.device)
        for d in range(min(self.grid.ndim, F.shape[0])):
            out = out + 1j * self.k[d].to(self.hardware_dtype) * F[d]
        return out.unsqueeze(1)

    def div(self, field_array) -> torch.Tensor:
        """Divergence of a vector or tensor field [m, n, *grid] -> [n, 1, *grid] (HIP FFTs)."""
        return self.ift(self.div_ft(self.ft(field_array)))

    def outer(self, A, B) -> torch.Tensor:
        """Dyadic product A B^H of [3, 1, *grid] fields, [3, 3, *grid] (``backend/__init__.py:531-542``)."""
        return self.astype(A) * self.conj(self.astype(B))[None, :, 0]

    def cross(self, A, B):
        A = self.astype(A)
        B = self.astype(B)
        return torch.stack([A[1] * B[2] - A[2] * B[1], A[2] * B[0] - A[0] * B[2], A[0] * B[1] - A[1] * B[0]])

    def norm(self, arr) -> float:
        """l2 norm with FP64 accumulation on the device (one host sync)."""
        x = self.astype(arr)
        if not x.is_contiguous():
            x = x.contiguous()
        out = torch.empty(1, dtype=torch.float64, device=self.__device)
        self.__plan.norm2(x, out)
        return math.sqrt(float(out.item()))

    # ------------------------------------------------------------------ eigenvalues (set-up)
    def mat3_eigh(self, arr) -> torch.Tensor:
        """Eigenvalues of Hermitian 3x3 matrix fields (``backend/__init__.py:670-723``), as a [3, *grid] array.

        1x1 input returns the single value; 3x1 (diagonal-as-column) returns the diagonal. 3x3 uses the
        trigonometric closed form for Hermitian matrices, evaluated in float64.
        """
        arr = self.astype(arr)
        if arr.shape[0] == 1:
            return arr[0].clone()
        if arr.shape[1] == 1:
            return arr[:, 0].clone()
        return hermitian3_eigvals(arr).to(self.hardware_dtype)

    def calc_roots_of_low_order_polynomial(self, C) -> torch.Tensor:
        """Roots of polynomials of degree <= 3 in parallel (``backend/__init__.py:725-850``).

        C holds the coefficients from low to high order along axis 0, one polynomial per position along the other
        axes: sum_k C[k] x^k == 0. The result holds the roots along axis 0 (one entry fewer than C), sorted by real
        then imaginary part. When the leading coefficients are below 2 eps of the hardware type the polynomial has
        fewer roots; the missing ones are nan and sort last. Computed on the device in complex128 (see
        ``polynomial_roots``) and returned in the hardware type.
        """
        Ct = torch.as_tensor(np.asarray(C) if not isinstance(C, torch.Tensor) else C, device=self.__device)
        return polynomial_roots(Ct.to(torch.complex128), 2.0 * self.eps).to(self.hardware_dtype)

    @staticmethod
    def evaluate_polynomial(C, X):
        """sum_k C[k] X^k per polynomial (Horner's scheme; ``backend/__init__.py:853-866``), for NumPy arrays or
        tensors."""
        result = 0
        for k in range(C.shape[0] - 1, -1, -1):
            result = result * X + C[k]
        return result

    @staticmethod
    def clear_cache():
        pass

    # ------------------------------------------------------------------ fused iteration
    def born_step(self, E, Eout, s, chi, chi_kind: int, coef, norms, nbatch: int = 1, s_bstride: int = 0,
                  chi_bstride: int = 0, s_box=None):
        """d = chi [G(chi E + s) - E]; Eout = E + d; norms = [sum |d|^2, sum |Eout|^2] (FP64, on the device).
        ``s_box`` = (lo, hi): s is zero outside that box (None: anywhere)."""
        ncomp = E.shape[-self.grid.ndim - 2] if E.ndim >= self.grid.ndim + 2 else 1
        return self.__plan.born_step(E, Eout, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride, norms,
                                     s_box=s_box)

    def born_prologue(self, E, W, s, chi, chi_kind: int, coef, nbatch: int = 1, s_bstride: int = 0,
                      chi_bstride: int = 0, s_box=None):
        ncomp = E.shape[-self.grid.ndim - 2]
        return self.__plan.born_prologue(E, W, s, chi, ncomp, chi_kind, coef, nbatch, s_bstride, chi_bstride,
                                         s_box=s_box)

    def born_iterate(self, E, Eout, W, s, chi, chi_kind: int, coef, norms, nbatch: int = 1, s_bstride: int = 0,
                     chi_bstride: int = 0, s_box=None, norm_stream=None, slot: int = 0, norms_host=None):
        ncomp = E.shape[-self.grid.ndim - 2]
        return self.__plan.born_iterate(E, Eout, W, s, chi, ncomp, chi_kind, coef, norms, nbatch, s_bstride,
                                        chi_bstride, s_box=s_box, norm_stream=norm_stream, slot=slot,
                                        norms_host=norms_host)

    def __str__(self) -> str:
        return f'BackEndHip(vectorial={self.vectorial}, {self.grid}, dtype={self.numpy_dtype}, device={self.__device})'

    __repr__ = __str__


def polynomial_roots(C: torch.Tensor, tol: float) -> torch.Tensor:
    """Roots of the polynomials sum_k C[k] x^k (degree <= 3, complex128 [n, ...]) -> [n - 1, ...] complex128.

    The degree of each polynomial is the highest k with |C[k]| >= tol; its missing roots are nan. Monic linear:
    x = -c0. Monic quadratic: q = -(b + sigma sqrt(b^2 - 4 c0)) / 2 with the sign sigma that avoids cancellation,
    roots q and c0 / q (Numerical Recipes' form). Monic cubic: x = t - a / 3 turns it into t^3 + p t + q = 0;
    Cardano's u = cbrt(-q/2 +- sqrt(q^2/4 + p^3/27)) with the sign of larger magnitude and v = -p / (3 u) give
    t = w u + v / w for the three cube roots of unity w (t = 0 three times when u = 0). Every root then takes two
    Newton steps on the original polynomial, each kept only if it lowers |P(x)|. The roots of each polynomial are
    sorted lexicographically (real, then imaginary part; nan last).
    """
    n = C.shape[0]
    rest = tuple(C.shape[1:])
    if n < 2:
        return torch.empty((0, *rest), dtype=torch.complex128, device=C.device)
    if n > 4:
        raise ValueError('Orders above 3 not implemented!')
    d = n - 1
    P = C.reshape(n, -1)
    M = P.shape[1]
    deg = torch.zeros(M, dtype=torch.int64, device=C.device)
    for k in range(1, n):
        deg = torch.where(P[k].abs() >= tol, torch.full_like(deg, k), deg)
    nan = complex(float('nan'), float('nan'))
    roots = torch.full((d, M), nan, dtype=torch.complex128, device=C.device)

    def horner(c, x):  # c: [e + 1, K] low to high, x: [r, K]
        val = torch.zeros_like(x)
        der = torch.zeros_like(x)
        for k in range(c.shape[0] - 1, -1, -1):
            der = der * x + val
            val = val * x + c[k]
        return val, der

    for e in range(1, d + 1):
        sel = deg == e
        if not bool(sel.any()):
            continue
        c = P[:e + 1, sel] / P[e, sel]  # monic
        if e == 1:
            r = -c[0:1]
        elif e == 2:
            b, c0 = c[1], c[0]
            sq = torch.sqrt(b * b - 4.0 * c0)
            sigma = torch.where((b.conj() * sq).real >= 0, 1.0, -1.0).to(torch.complex128)
            q = -0.5 * (b + sigma * sq)
            q_zero = q == 0
            r2 = torch.where(q_zero, torch.zeros_like(q), c0 / torch.where(q_zero, torch.ones_like(q), q))
            r = torch.stack([q, r2])
        else:
            a, b, c0 = c[2], c[1], c[0]
            p = b - a * a / 3.0
            q = 2.0 * a ** 3 / 27.0 - a * b / 3.0 + c0
            sd = torch.sqrt(q * q / 4.0 + p ** 3 / 27.0)
            w1, w2 = -q / 2.0 + sd, -q / 2.0 - sd
            w = torch.where(w1.abs() >= w2.abs(), w1, w2)
            u = torch.pow(w, 1.0 / 3.0)
            u_zero = u == 0
            v = torch.where(u_zero, torch.zeros_like(u), -p / (3.0 * torch.where(u_zero, torch.ones_like(u), u)))
            unity = torch.tensor([1.0, complex(-0.5, math.sqrt(3) / 2), complex(-0.5, -math.sqrt(3) / 2)],
                                 dtype=torch.complex128, device=C.device)[:, None]
            r = unity * u[None] + unity.conj() * v[None] - a[None] / 3.0
        for _ in range(2):  # Newton polish, kept where it lowers |P|
            val, der = horner(c, r)
            ok = der != 0
            step = torch.where(ok, val / torch.where(ok, der, torch.ones_like(der)), torch.zeros_like(val))
            r_new = r - step
            better = horner(c, r_new)[0].abs() < val.abs()
            r = torch.where(better, r_new, r)
        block = torch.full((d, r.shape[1]), nan, dtype=torch.complex128, device=C.device)
        block[:e] = r
        roots[:, sel] = block
    # lexicographic sort along axis 0: stable sort by imaginary part, then by real part (nan sorts last)
    order = torch.sort(roots.imag, dim=0, stable=True).indices
    roots = torch.gather(roots, 0, order)
    order = torch.sort(roots.real, dim=0, stable=True).indices
    roots = torch.gather(roots, 0, order)
    return roots.reshape(d, *rest)


def hermitian3_eigvals(m: torch.Tensor) -> torch.Tensor:
    """Real eigenvalues (ascending) of Hermitian 3x3 fields ``m`` [3,3,...] by the trigonometric method (float64)."""
    m = m.to(torch.complex128)
    a00, a11, a22 = m[0, 0].real, m[1, 1].real, m[2, 2].real
    a01, a02, a12 = m[0, 1], m[0, 2], m[1, 2]
    q = (a00 + a11 + a22) / 3
    p1 = a01.abs() ** 2 + a02.abs() ** 2 + a12.abs() ** 2
    p2 = (a00 - q) ** 2 + (a11 - q) ** 2 + (a22 - q) ** 2 + 2 * p1
    p = torch.sqrt(p2 / 6)
    safe = p > 0
    ps = torch.where(safe, p, torch.ones_like(p))
    b00, b11, b22 = (a00 - q) / ps, (a11 - q) / ps, (a22 - q) / ps
    b01, b02, b12 = a01 / ps, a02 / ps, a12 / ps
    # det of the Hermitian matrix B (real)
    det = b00 * b11 * b22 + 2 * (b01 * b12 * b02.conj()).real - b00 * b12.abs() ** 2 - b11 * b02.abs() ** 2 \
        - b22 * b01.abs() ** 2
    r = torch.clamp(det / 2, -1.0, 1.0)
    phi = torch.acos(r) / 3
    l1 = q + 2 * p * torch.cos(phi)
    l3 = q + 2 * p * torch.cos(phi + 2 * math.pi / 3)
    l2 = 3 * q - l1 - l3
    l1 = torch.where(safe, l1, q)
    l2 = torch.where(safe, l2, q)
    l3 = torch.where(safe, l3, q)
    return torch.stack([l3, l2, l1])
