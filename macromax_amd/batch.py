"""Batched Born solves on one GPU: many right-hand sides of one medium, or many media of one grid shape, iterated
together with the accept/reject decision and the stop test taken on the device (``mmx_batch_step``), K iterations
per captured graph replay and one host synchronisation per replay.

The reference solves every right-hand side with a ``Solution`` of its own; ``ScatteringMatrix`` loops over them
(``python/macromax/matrix.py:789-819``). Each problem of a batch follows exactly the iterates ``Solution.solve``
would produce for it (``python/macromax/solver.py:1064-1123``: accept if ||d|| < ||d_prev||, else alpha_i grows by
1.5; stop when the callback ``residue > tol and iteration < max_iter`` turns false), so the results can be handed
back to the ``Solution`` objects the batch was built from.

``solve_many`` takes problems of uneven grid sizes: one batch per grid shape, the batches' launch sequences on
separate streams of one captured graph.
"""
from __future__ import annotations

import math
from collections import OrderedDict
from typing import Sequence

import numpy as np
import scipy.constants as const
import torch

from . import _native
from .solver import Solution, source_support

__all__ = ['BatchSolution', 'solve_many']


class BatchSolution:
    """nbatch problems of one grid shape, element type and chi storage, solved together on the device.

    :param solutions: ``Solution`` objects of identical grid shape and k0-scaled grid step; their current fields,
        alphas and iteration counts are the starting point. They share one chi on the device when they share the
        medium tensor (``from_sources``), otherwise chi is stacked per problem.
    """

    def __init__(self, solutions: Sequence[Solution], _sources=None, _transpose=False):
        sols = list(solutions)
        if len(sols) == 0:
            raise ValueError('a batch needs at least one problem')
        first = sols[0]
        ops0 = first._batch_operands()
        shape = tuple(int(_) for _ in first.grid.shape)
        kstep = np.asarray(first.grid.step) * first.wavenumber
        self.__sources_mode = _sources is not None
        if self.__sources_mode:
            ops = [ops0] * len(_sources)
        else:
            ops = [s._batch_operands() for s in sols]
            for s, o in zip(sols, ops):
                if tuple(int(_) for _ in s.grid.shape) != shape:
                    raise ValueError('all problems of a batch must have the same grid shape')
                if not np.allclose(np.asarray(s.grid.step) * s.wavenumber, kstep, rtol=1e-12, atol=0):
                    raise ValueError('all problems of a batch must have the same grid step in units of 1/k0')
                if s.dtype != first.dtype or o['ncomp'] != ops0['ncomp'] or o['chi_kind'] != ops0['chi_kind']:
                    raise ValueError('all problems of a batch must share dtype, polarisation and chi storage')
        self.__solutions = sols
        self.__backend = first.backend
        self.__plan = first.backend.transposed_plan if _transpose else first.backend.plan
        self.__shape = shape
        self.__wavenumber = first.wavenumber
        B = len(ops)
        self.__nbatch = B
        C = ops0['ncomp']
        V = int(np.prod(shape))
        self.__ncomp = C
        self.__chi_kind = ops0['chi_kind']
        dev = first.backend.device
        shared = self.__sources_mode or all(o['chi'] is ops0['chi'] for o in ops)
        if shared:
            self.__chi = ops0['chi']
            if _transpose and ops0['chi_kind'] == 2:  # kinds 1 and 3 are symmetric
                self.__chi = self.__chi.transpose(0, 1).contiguous()
            self.__chi_bstride = 0
        else:
            self.__chi = torch.stack([o['chi'] for o in ops]).contiguous()
            self.__chi_bstride = ops0['chi'].numel()
        if self.__sources_mode:
            self.__s = torch.stack([first._batch_source(src) for src in _sources]).contiguous()
            fields = torch.zeros((B, C, 1, *shape), dtype=first.backend.hardware_dtype, device=dev)
        else:
            self.__s = torch.stack([o['s'] for o in ops]).contiguous()
            fields = torch.stack([o['field'] for o in ops]).contiguous()
        self.__s_bstride = C * V
        # union of the problems' source supports (the passes read s only inside it)
        if self.__sources_mode:
            box = source_support(self.__s, len(shape))
        else:
            boxes = [o['s_box'] for o in ops if list(o['s_box'][0]) != list(o['s_box'][1])]
            if boxes:
                box = ([min(b[0][a] for b in boxes) for a in range(len(shape))],
                       [max(b[1][a] for b in boxes) for a in range(len(shape))])
            else:
                box = ([0] * len(shape), [0] * len(shape))
        self.__s_box = box
        self.__coef = torch.stack([o['coef'] for o in ops]).contiguous()
        self.__EA = fields
        self.__EB = torch.zeros_like(fields)
        self.__W = torch.zeros_like(fields)
        self.__state = torch.full((B,), _native.STATE_STALE, dtype=torch.int32, device=dev)
        T = _native.TRACK
        track = np.zeros((B, _native.TRACK_LEN))
        for b, o in enumerate(ops):
            track[b, T['prev']] = math.inf if self.__sources_mode else o['prev']
            track[b, T['iter']] = 0 if self.__sources_mode else o['iteration']
            track[b, T['alpha_r']] = o['alpha'].real
            track[b, T['alpha_i']] = o['alpha'].imag
            track[b, T['a']] = o['coef_a']
            track[b, T['c_r']] = o['coef_c'].real
            track[b, T['c_i']] = o['coef_c'].imag
        self.__track = torch.tensor(track, dtype=torch.float64, device=dev)
        self.__norms = torch.zeros(2 * B, dtype=torch.float64, device=dev)
        self.__clock0 = torch.zeros(1, dtype=torch.float64, device=dev)
        self.__graph = None
        self.__graph_key = None

    @classmethod
    def from_sources(cls, solution: Solution, sources=None, current_densities=None,
                     transpose: bool = False) -> 'BatchSolution':
        """Many right-hand sides of the medium of ``solution`` (its chi and alpha), all starting from E = 0.

        :param sources: source distributions S (as for ``Solution.source_distribution``), indexed by the first
            dimension: ``sources[b]`` is problem b's S.
        :param current_densities: alternatively the current densities J (S = -i omega mu_0 J, solver.py:183).
        :param transpose: solve with the transposed operator (chi^T and the transposed Green's function), as an
            adjoint solve by reciprocity needs.
        """
        if (sources is None) == (current_densities is None):
            raise ValueError('give either sources or current_densities')
        if sources is None:
            factor = -1j * solution.angular_frequency * const.mu_0
            srcs = [np.asarray(current_densities[b]) * factor for b in range(len(current_densities))]
        else:
            srcs = [sources[b] for b in range(len(sources))]
        return cls([solution], _sources=srcs, _transpose=transpose)

    # ------------------------------------------------------------------ properties
    @property
    def nbatch(self) -> int:
        return self.__nbatch

    @property
    def shape(self) -> tuple:
        return self.__shape

    @property
    def plan(self):
        return self.__plan

    @property
    def voxels(self) -> int:
        return int(np.prod(self.__shape))

    def _track(self) -> np.ndarray:
        return self.__track.cpu().numpy()

    @property
    def iteration(self) -> np.ndarray:
        return self._track()[:, _native.TRACK['iter']].astype(np.int64)

    @property
    def residue(self) -> np.ndarray:
        return self._track()[:, _native.TRACK['residue']]

    @property
    def alpha(self) -> np.ndarray:
        t = self._track()
        return t[:, _native.TRACK['alpha_r']] + 1j * t[:, _native.TRACK['alpha_i']]

    @property
    def rejected(self) -> np.ndarray:
        return self._track()[:, _native.TRACK['rejected']].astype(np.int64)

    @property
    def done(self) -> np.ndarray:
        return (self.__state.cpu().numpy() & _native.STATE_DONE) != 0

    def field_arrays(self) -> torch.Tensor:
        """The current k0^2-scaled fields on the device, [nbatch, C, 1, *grid] (a new tensor)."""
        parity = (self.__state & _native.STATE_PARITY) != 0
        sel = parity.reshape(-1, *([1] * (self.__EA.ndim - 1)))
        return torch.where(sel, self.__EB, self.__EA)

    @property
    def E(self) -> np.ndarray:
        """Electric fields in SI units, [nbatch, C, *grid]."""
        return (self.field_arrays()[:, :, 0] / (self.__wavenumber ** 2)).cpu().numpy()

    def latency(self) -> np.ndarray:
        """Seconds from the start of the last ``solve`` to each problem's stop (device clock); NaN if unfinished."""
        khz = _native.wall_clock_rate_khz()
        t = self._track()[:, _native.TRACK['done_clock']]
        t0 = float(self.__clock0.cpu()[0])
        out = (t - t0) / (khz * 1e3)
        out[~self.done] = np.nan
        return out

    # ------------------------------------------------------------------ iteration
    def _launch(self, tol: float, max_iter: int):
        self.__plan.batch_step(self.__EA, self.__EB, self.__W, self.__s, self.__chi, self.__ncomp, self.__chi_kind,
                               self.__coef, self.__nbatch, self.__s_bstride, self.__chi_bstride, self.__state,
                               self.__track, self.__norms, tol, max_iter, s_box=self.__s_box)

    def _prepare(self, tol: float, max_iter: int):
        """One launch with every problem marked finished: allocates the plan's scratch and sets kernel attributes
        (so that later launches can be captured) without changing any problem."""
        saved = self.__state.clone()
        self.__state.fill_(_native.STATE_DONE)
        self._launch(tol, max_iter)
        self.__state.copy_(saved)

    def _stamp_clock(self):
        _native.check(_native.lib().mmx_wall_clock(self.__clock0.data_ptr(),
                                                   torch.cuda.current_stream(self.__EA.device).cuda_stream),
                      'mmx_wall_clock')

    def _set_clock0(self, other: 'BatchSolution'):
        self.__clock0.copy_(other.__clock0)

    def step(self, tol: float = 0.0, max_iter: int = 2 ** 62):
        """One iteration of every unfinished problem (no host synchronisation)."""
        self._launch(tol, max_iter)

    def all_done(self) -> bool:
        return bool(((self.__state & _native.STATE_DONE) != 0).all())

    def tune(self, **kwargs) -> dict:
        """Autotune the batch's plan for this batch (``macromax_amd.autotune.tune_plan``)."""
        from .autotune import tune_plan
        return tune_plan(self.__plan, self.__ncomp, self.__chi_kind, nbatch=self.__nbatch,
                         chi_shared=self.__chi_bstride == 0, **kwargs)

    def solve(self, tol: float = 1e-4, max_iter: int = 10000, steps_per_sync: int = 16, graph: bool = True,
              write_back: bool = True, tune: bool = False) -> 'BatchSolution':
        """Iterate every problem until ``not (residue > tol and iteration < max_iter)``.

        :param steps_per_sync: iterations launched (or captured in one graph) between host checks.
        :param graph: capture ``steps_per_sync`` iterations in a graph once and replay it.
        :param write_back: hand each problem's final state back to its ``Solution`` (batches built from Solutions).
        :param tune: autotune the plan's tile shapes for this batch first (cached per size).
        """
        if tune:
            self.tune()
            self.__graph = None
        if graph:
            key = (float(tol), int(max_iter), int(steps_per_sync))
            if self.__graph is None or self.__graph_key != key:
                self._prepare(tol, max_iter)
                torch.cuda.synchronize(self.__EA.device)
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    for _ in range(steps_per_sync):
                        self._launch(tol, max_iter)
                self.__graph, self.__graph_key = g, key
            self._stamp_clock()
            while not self.all_done():
                self.__graph.replay()
        else:
            self._stamp_clock()
            while not self.all_done():
                for _ in range(steps_per_sync):
                    self._launch(tol, max_iter)
        if write_back and not self.__sources_mode:
            self.write_back()
        return self

    def write_back(self):
        """Give every ``Solution`` of the batch its problem's field, alpha, coefficients and counters."""
        if self.__sources_mode:
            raise ValueError('a batch of right-hand sides has no Solution per problem')
        fields = self.field_arrays()
        t = self._track()
        T = _native.TRACK
        for b, sol in enumerate(self.__solutions):
            sol._batch_result(fields[b], t[b, T['iter']], t[b, T['prev']], t[b, T['field_norm']],
                              complex(t[b, T['alpha_r']], t[b, T['alpha_i']]), t[b, T['a']],
                              complex(t[b, T['c_r']], t[b, T['c_i']]))


def _group_key(sol: Solution):
    o = sol._batch_operands()
    kstep = tuple(np.round(np.asarray(sol.grid.step) * sol.wavenumber, 12))
    return (tuple(int(_) for _ in sol.grid.shape), kstep, np.dtype(sol.dtype).name, o['ncomp'], o['chi_kind'])


def _stamp_all(batches):
    """One start time for every batch (latencies are measured from it)."""
    batches[0]._stamp_clock()
    for bt in batches[1:]:
        bt._set_clock0(batches[0])


def solve_many(solutions: Sequence[Solution], tol: float = 1e-4, max_iter: int = 10000, steps_per_sync: int = 16,
               concurrent: bool = True, graph: bool = True, tune: bool = False):
    """Solve problems of uneven grid sizes: one ``BatchSolution`` per grid shape (and element type / chi storage),
    the groups' iterations on separate streams of one captured graph when ``concurrent``; ``tune`` autotunes every
    group's plan for its batch first.

    Each ``Solution`` receives its final state. Returns the list of batches (their ``latency()`` gives the
    per-problem stop times from the common start).
    """
    groups = OrderedDict()
    for sol in solutions:
        groups.setdefault(_group_key(sol), []).append(sol)
    batches = [BatchSolution(g) for g in groups.values()]
    if not batches:
        return []
    if tune:
        for bt in batches:
            bt.tune()
    dev = batches[0].field_arrays().device
    if graph:
        for bt in batches:
            bt._prepare(tol, max_iter)
        torch.cuda.synchronize(dev)
        streams = [torch.cuda.Stream(device=dev) for _ in batches] if concurrent else None
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            if concurrent:
                cur = torch.cuda.current_stream(dev)
                for bt, st in zip(batches, streams):
                    st.wait_stream(cur)
                    with torch.cuda.stream(st):
                        for _ in range(steps_per_sync):
                            bt._launch(tol, max_iter)
                for st in streams:
                    cur.wait_stream(st)
            else:
                for bt in batches:
                    for _ in range(steps_per_sync):
                        bt._launch(tol, max_iter)
        _stamp_all(batches)
        while not all(bt.all_done() for bt in batches):
            g.replay()
    else:
        _stamp_all(batches)
        while not all(bt.all_done() for bt in batches):
            for bt in batches:
                for _ in range(steps_per_sync):
                    bt._launch(tol, max_iter)
    for bt in batches:
        bt.write_back()
    return batches
