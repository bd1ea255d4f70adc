"""Adaptive stop criterion (MacroMax ``utils/stop_criterion.py:7-52``).

A callback for ``Solution.solve`` / ``solve`` that reads ``Solution.residue`` only at a few iterations: two readings
25 and 50 iterations after it starts watching, then, from the observed geometric rate between the last two readings,
one reading at the iteration where the residue is predicted to reach the target (at most as many iterations ahead as
have already run). It stops at a reading below the target, or at the iteration cap. The decisions follow the
reference's rule step for step, so a solve stops at the same iteration as with the reference's callback; on this
backend the residue is a host value after every iteration, so the saving is only in the callback itself.
"""
from __future__ import annotations

import math

import numpy as np


class Adaptive:
    """Stop when a scheduled residue reading is below ``residue``, or at ``iteration`` iterations."""

    def __init__(self, residue: float = 1e-4, iteration: int = 10_000):
        self.residue = residue
        self.iteration = iteration
        self._at = np.ones(2, dtype=int)      # iterations of the two pending readings
        self._read = np.ones(2)               # residues read there
        self._first_offsets = np.asarray((25, 50))
        self._last_seen = None                # the iteration of the previous call (a smaller one: a new problem)

    def __call__(self, solution) -> bool:
        it = solution.iteration
        if self._last_seen is not None and it < self._last_seen:
            self._at = np.ones(2, dtype=int)  # applied to another problem: watch it from the start
        self._last_seen = it
        if it > self._at.max():  # past both readings: schedule two new ones
            self._at = it + self._first_offsets
        if it == self._at[0]:
            self._read[0] = solution.residue
        elif it == self._at[1]:
            self._read[1] = solution.residue
            if self._read[1] < self.residue:
                return False
            steps = self._at[1] - self._at[0]
            rate = (self._read[1] / self._read[0]) ** (1.0 / steps)
            # a residue that did not fall: read again at the next iteration (the reference's formula gives 1 for a
            # rising residue and fails for a constant one)
            to_go = math.log(self.residue / self._read[1]) / math.log(rate) if 0.0 < rate < 1.0 else 1.0
            # the later reading becomes the first of the next pair; the next one where the target is predicted
            self._at[0], self._read[0] = self._at[1], self._read[1]
            self._at[1] = it + min(max(1, int(to_go + 0.5)), it)
        elif it >= self.iteration:
            return False
        return True
