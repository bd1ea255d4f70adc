"""The adaptive stop criterion (macromax_amd.utils.stop_criterion.Adaptive, MacroMax utils/stop_criterion.py:7-52)
on synthetic residue sequences (CPU): where it stops, how often it reads the residue, the iteration cap, and its
reset when applied to another problem."""
import math

import pytest

from macromax_amd.utils import Adaptive


class _Sequence:
    """A stand-in for Solution: residue(n) given as a function, counting the readings."""

    def __init__(self, residue_of):
        self.residue_of = residue_of
        self.iteration = 0
        self.reads = 0

    @property
    def residue(self):
        self.reads += 1
        return self.residue_of(self.iteration)


def _run(stop, seq, limit=100000):
    while seq.iteration < limit:
        seq.iteration += 1
        if not stop(seq):
            return seq.iteration
    raise AssertionError('no stop')


@pytest.mark.parametrize('rate,tol', [(0.9, 1e-4), (0.97, 1e-4), (0.995, 1e-6), (0.9995, 1e-5)])
def test_geometric_decay_stops_at_the_crossing(rate, tol):
    seq = _Sequence(lambda n: rate ** n)
    n_stop = _run(Adaptive(tol, iteration=10 ** 6), seq)
    n_cross = math.ceil(math.log(tol) / math.log(rate))
    assert rate ** n_stop < tol
    assert n_cross <= n_stop <= n_cross + 1, (n_stop, n_cross)
    # two readings, then one at the predicted iteration; a prediction further ahead than the iterations run so far is
    # cut to that many, so the readings at most double the iteration until the prediction is in reach
    assert seq.reads <= 4 + max(0, math.ceil(math.log2(n_cross / 50))), seq.reads


def test_slower_than_predicted_reads_again():
    # the rate halves its speed after the first estimate: further readings follow until the target
    seq = _Sequence(lambda n: 0.98 ** n if n <= 60 else 0.98 ** 60 * 0.99 ** (n - 60))
    n_stop = _run(Adaptive(1e-4, iteration=10 ** 6), seq)
    assert seq.residue_of(n_stop) < 1e-4
    assert 3 <= seq.reads <= 12
    # it never stops before the residue is below the target at a reading
    n_true = next(n for n in range(1, 10 ** 6) if seq.residue_of(n) < 1e-4)
    assert n_stop >= n_true


def test_iteration_cap():
    seq = _Sequence(lambda n: 1.0 / (1.0 + 1e-3 * n))  # far from the target
    assert _run(Adaptive(1e-8, iteration=300), seq) == 300


def test_rising_or_constant_residue_is_read_again():
    seq = _Sequence(lambda n: 0.5 if n < 100 else 1e-9)
    n_stop = _run(Adaptive(1e-6, iteration=10 ** 6), seq)
    assert n_stop == 100  # read at every iteration once the residue stopped falling


def test_reset_for_another_problem():
    stop = Adaptive(1e-4, iteration=10 ** 6)
    first = _Sequence(lambda n: 0.9 ** n)
    _run(stop, first)
    second = _Sequence(lambda n: 0.95 ** n)
    n_stop = _run(stop, second)
    n_cross = math.ceil(math.log(1e-4) / math.log(0.95))
    assert n_cross <= n_stop <= n_cross + 1
