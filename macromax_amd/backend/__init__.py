"""Backend interface and registry, mirroring MacroMax ``backend/__init__.py``.

``BackEnd`` is the abstract interface the solver calls (reference ``backend/__init__.py:37-880``); ``config`` and
``load`` select an implementation (reference ``:883-1012``). This package registers exactly one implementation,
``{"type": "hip"}`` (``BackEndHip``): an MI355X backend whose FFTs, Green's operator, susceptibility multiply,
norms and fused Born iteration run as hand-written gfx950 kernels. Configurations of other types are skipped
(``load`` treats an unknown type as not loadable, as the reference does at ``:1007-1008``), and when nothing
loads ``load`` raises ``TypeError('No viable backend detected!')`` (``:1012``).

Deliberate differences from the reference's ``config``/``load`` quirks:
- ``config()`` with no arguments resets to the default list instead of appending an empty dict (``:916-919``);
- a single-dict ``backend_config.json`` is wrapped in a list (the reference tests the function ``config`` at
  ``:951``);
- the default candidates are not appended to the global list on every call (``:961-965``).
"""
from __future__ import annotations

import json
import logging
from abc import ABC, abstractmethod
from typing import Dict, List, Optional, Sequence

import numpy as np

from ..grid import Grid

log = logging.getLogger(__name__)

__all__ = ['config', 'load', 'BackEnd']

_config_list: Optional[List[Dict]] = None


class BackEnd(ABC):
    """Parallel operations on fields of matrices/vectors ``[rows, cols, *grid.shape]``.

    The first two dimensions are the matrix dimensions; Fourier transforms act on the remaining (spatial) axes.
    Interface of ``backend/__init__.py:37-880`` (reference).
    """

    def __init__(self, nb_dims: int, grid: Grid, hardware_dtype=np.complex128):
        self.__nb_rows = int(nb_dims)
        self.__grid = grid
        self.__hardware_dtype = hardware_dtype
        self.__array_ft_input = None
        self.__array_ft_output = None

    @property
    def vector_length(self) -> int:
        return self.__nb_rows

    @property
    def ft_axes(self) -> tuple:
        return tuple(range(2, 2 + self.grid.ndim))

    @property
    def grid(self) -> Grid:
        return self.__grid

    @property
    def vectorial(self) -> bool:
        return self.vector_length > 1

    @property
    def hardware_dtype(self):
        return self.__hardware_dtype

    @property
    def numpy_dtype(self):
        return self.hardware_dtype

    @property
    def array_ft_input(self):
        if self.__array_ft_input is None:
            self.__array_ft_input = self.allocate_array()
        return self.__array_ft_input

    @property
    def array_ft_output(self):
        if self.__array_ft_output is None:
            self.__array_ft_output = self.allocate_array()
        return self.__array_ft_output

    @property
    def array_ift_input(self):
        return self.array_ft_output

    @property
    def array_ift_output(self):
        return self.array_ft_input

    @abstractmethod
    def allocate_array(self, shape=None, dtype=None, fill_value=None):
        ...

    @abstractmethod
    def ft(self, arr):
        ...

    @abstractmethod
    def ift(self, arr):
        ...

    @abstractmethod
    def ldivide(self, denominator, numerator=1.0):
        ...

    def convolve(self, operation_ft, arr):
        """ft -> operation -> ift (``backend/__init__.py:306-319``)."""
        return self.ift(operation_ft(self.ft(arr)))

    def inv(self, mat):
        return self.ldivide(mat, 1.0)

    @staticmethod
    def clear_cache():
        pass

    def __str__(self) -> str:
        return f'{self.__class__.__name__}(vectorial={self.vectorial}, {self.grid}, dtype={self.numpy_dtype})'

    __repr__ = __str__


def config(*args: Sequence[Dict], **kwargs: str):
    """Set the list of backend configurations to try, e.g. ``config(type='hip')`` or ``config([{'type': 'hip'}])``."""
    global _config_list
    items = []
    for a in args:
        if isinstance(a, dict):
            items.append(a)
        elif isinstance(a, (list, tuple)):
            items.extend(a)
        else:
            raise TypeError("Configuration must be specified as a dictionary or a series of dictionaries with the key 'type'.")
    if kwargs:
        items.append(dict(kwargs))
    if any(not isinstance(_, dict) or 'type' not in _ for _ in items):
        raise TypeError("Configuration must be specified as a dictionary or a series of dictionaries with the key 'type'.")
    _config_list = items if items else None


def load(nb_pol_dims: int, grid: Grid, dtype, config_list: List[Dict] = None) -> BackEnd:
    """Return the first backend of the configuration list that can be constructed (``backend/__init__.py:922-1012``).

    The list comes from ``config_list``, else ``config()``, else ``./backend_config.json``; ``{"type": "hip"}`` is
    always tried last. Constructors that raise ``ImportError`` are skipped.
    """
    if config_list is None:
        if _config_list is None:
            try:
                with open('backend_config.json', 'r') as file:
                    config_list = json.load(file)
                if isinstance(config_list, dict):
                    config_list = [config_list]
            except FileNotFoundError:
                config_list = []
            except json.decoder.JSONDecodeError:
                log.warning('Could not read backend_config.json.')
                config_list = []
        else:
            config_list = _config_list
    candidates = list(config_list) + [{'type': 'hip'}]
    for c in candidates:
        config_type = str(c.get('type', 'unknown')).lower()
        try:
            if config_type in ('hip', 'mi355x'):
                from .hip import BackEndHip
                device = c.get('device', None)
                return BackEndHip(nb_pol_dims, grid, dtype, device=device)
            raise ImportError(f'Backend config type {config_type} not provided by this package.')
        except ImportError as ie:
            log.info(f'Backend type "{config_type}" did not load: {ie}.')
    raise TypeError('No viable backend detected!')
