"""Helpers beside the solver (MacroMax ``python/macromax/utils``): the sampling grid and the adaptive stop
criterion."""
import logging

from .ft import Grid
from .stop_criterion import Adaptive

log = logging.getLogger(__name__)

__all__ = ['Grid', 'Adaptive']
