"""Gradients of a loss of the converged field with respect to the permittivity, by a hand-written adjoint Born
solve behind ``torch.autograd.Function``.

The reference has no gradient (its one differentiable Born series, ``examples/scattering_neural_network.py:206-318``,
unrolls a scalar iteration through autograd). Here the forward pass is an ordinary ``Solution`` solve on the GPU;
only its converged field is kept (the checkpoint; no iterate is stored). The converged field solves

    L E~ = S / beta,   L = curl curl (or -laplacian) - eps / beta          (solver.py:566-584, 480-484, 702-720)

in the solver's units (E~ = k0^2 E). For a real loss with gradient g = dL/dRe(E) + i dL/dIm(E) (PyTorch's
convention for complex tensors), dE = L^-1 (d eps E~ / beta) / k0^2, so

    mu = L^-T conj(g)            one Born solve of the transposed operator with source S' = beta conj(g)
    dL/d eps = conj(sum_c mu_c E~_c) / (beta k0^2)       (isotropic eps; conj(mu_i E~_j) / (beta k0^2) for 3x3)

L^T is L with chi^T and, on even-length axes, the Nyquist entry of k negated (only the dyadic Green's cross terms
k_i k_j / k^2 on those planes differ); the transposed solve runs on the same kernels through a second plan. The
adjoint solve starts from zero and reuses the forward solve's alpha and chi (no new preconditioner set-up).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .batch import BatchSolution
from .solver import Solution

__all__ = ['solve_field', 'BornField', 'fields_at_fixed_alpha']


class BornField(torch.autograd.Function):
    """E = solve(epsilon) with d loss / d epsilon by one adjoint solve (see the module docstring)."""

    @staticmethod
    def forward(ctx, epsilon: torch.Tensor, problem: dict):
        sol = Solution(problem['grid'], vacuum_wavelength=problem['vacuum_wavelength'], epsilon=epsilon.detach(),
                       current_density=problem.get('current_density'),
                       source_distribution=problem.get('source_distribution'), bound=problem.get('bound'),
                       dtype=problem.get('dtype', np.complex64), vectorial=problem.get('vectorial'))
        tol, max_iter = problem.get('tol', 1e-6), problem.get('max_iter', 10000)
        sol.solve(lambda s: s.residue > tol and s.iteration < max_iter)
        k0 = sol.wavenumber
        ctx.sol = sol
        ctx.problem = problem
        ctx.field = sol.field_array.clone()  # the only state kept for the backward pass
        ctx.eps_shape = epsilon.shape
        ctx.eps_dtype = epsilon.dtype
        ctx.eps_device = epsilon.device
        problem['forward_iterations'] = sol.iteration
        problem['forward_residue'] = sol.residue
        problem['alpha'] = sol.alpha
        return ctx.field[:, 0] / (k0 ** 2)

    @staticmethod
    def backward(ctx, grad_E: torch.Tensor):
        sol, problem = ctx.sol, ctx.problem
        if grad_E is None:
            return None, None
        g = grad_E.to(ctx.field.device, ctx.field.dtype)
        C = ctx.field.shape[0]
        beta, k0 = sol.beta, sol.wavenumber
        scale = float(torch.linalg.vector_norm(g))
        if not math.isfinite(scale):
            raise FloatingPointError('the gradient with respect to the field is not finite')
        if scale == 0.0:
            return torch.zeros(ctx.eps_shape, dtype=ctx.eps_dtype, device=ctx.eps_device), None
        # S' = beta conj(g), normalised to unit norm (the solve is linear; the residue test is relative)
        src = (torch.conj(g) * (beta / scale)).reshape(1, C, *g.shape[1:])
        tol, max_iter = problem.get('tol', 1e-6), problem.get('max_iter', 10000)
        # one problem: eager launches (16 iterations between host checks) cost the same as graph replays here and
        # avoid a capture per backward pass
        adj = BatchSolution.from_sources(sol, sources=src, transpose=True).solve(tol=tol, max_iter=max_iter,
                                                                                graph=False)
        problem['adjoint_iterations'] = int(adj.iteration[0])
        problem['adjoint_residue'] = float(adj.residue[0])
        mu = adj.field_arrays()[0] * scale  # [C, 1, *grid], k0^2-scaled
        E = ctx.field
        factor = 1.0 / (beta * k0 ** 2)
        nd = len(ctx.eps_shape)
        grid_nd = E.ndim - 2
        if nd >= 2 + grid_nd and ctx.eps_shape[0] == 3 and ctx.eps_shape[1] == 3:
            grad = torch.conj(mu[:, None, 0] * E[None, :, 0]) * factor  # [3, 3, *grid]
        else:
            grad = torch.conj(torch.sum(mu[:, 0] * E[:, 0], dim=0)) * factor  # [*grid]
        if not ctx.eps_dtype.is_complex:
            grad = grad.real
        grad = grad.to(ctx.eps_device, ctx.eps_dtype)
        # sum out any broadcast dimensions of epsilon
        while grad.ndim > nd:
            grad = grad.sum(0)
        for ax, (ge, gg) in enumerate(zip(ctx.eps_shape, grad.shape)):
            if ge == 1 and gg != 1:
                grad = grad.sum(ax, keepdim=True)
        return grad.reshape(ctx.eps_shape), None


def solve_field(epsilon: torch.Tensor, grid, vacuum_wavelength: float, current_density=None,
                source_distribution=None, bound=None, dtype=np.complex64, tol: float = 1e-6, max_iter: int = 10000,
                vectorial=None, info: dict = None) -> torch.Tensor:
    """The converged electric field E (SI units, [C, *grid], on the GPU) for permittivity ``epsilon``,
    differentiable with respect to ``epsilon`` (a real or complex tensor of shape grid.shape or [3, 3, *grid]).

    ``info``, if given, receives the forward and adjoint iteration counts.
    """
    problem = {'grid': grid, 'vacuum_wavelength': vacuum_wavelength, 'current_density': current_density,
               'source_distribution': source_distribution, 'bound': bound, 'dtype': dtype, 'tol': tol,
               'max_iter': max_iter, 'vectorial': vectorial}
    if info is not None:
        problem = _Recorder(problem, info)
    return BornField.apply(epsilon, problem)


class _Recorder(dict):
    """A problem dict that also writes the iteration counts into the caller's ``info``."""

    def __init__(self, problem, info):
        super().__init__(problem)
        self.__info = info

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        self.__info[key] = value


def fields_at_fixed_alpha(epsilons, grid, vacuum_wavelength: float, alpha: complex, current_density=None,
                          bound=None, dtype=np.complex64, tol: float = 1e-6, max_iter: int = 10000) -> torch.Tensor:
    """Converged fields [n, C, *grid] (SI units) for several permittivities, all with the preconditioner constant
    ``alpha`` (instead of each one's own), solved as one batch. Finite-difference checks of gradients use it so
    that only the permittivity changes between the solves."""
    sols = []
    for eps in epsilons:
        s = Solution(grid, vacuum_wavelength=vacuum_wavelength, epsilon=eps, current_density=current_density,
                     bound=bound, dtype=dtype)
        s._set_alpha(alpha)
        sols.append(s)
    bt = BatchSolution(sols).solve(tol=tol, max_iter=max_iter, write_back=False)
    k0 = sols[0].wavenumber
    return bt.field_arrays()[:, :, 0] / (k0 ** 2)
