"""Helpers beside the solver (MacroMax ``python/macromax/utils``): the adaptive stop criterion."""
from .stop_criterion import Adaptive

__all__ = ['Adaptive']
