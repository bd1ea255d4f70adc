"""
Differentiable restatement of MacroMax's Born iteration in PyTorch (CPU, complex128).

TEST INFRASTRUCTURE ONLY. Only ``tests/`` may import this module; the product package ``macromax_amd`` never
does. It provides the autograd reference for gradients of a loss of the converged field with respect to the
permittivity: the iteration is unrolled out of place so that autograd differentiates through it, as the
reference's only differentiable Born series does (``examples/scattering_neural_network.py:206-318``). The
operators follow the reference torch backend and solver:

=================================  ==============================================================================
function                           reference
=================================  ==============================================================================
``chi_from_eps``                   ``solver.py:474-484`` (chi = (i/alpha_i)(eps/beta - alpha I))
``source_from_S``                  ``solver.py:702-720``
``ft`` / ``ift``                   ``backend/torch.py:183-205`` (torch.fft.fftn / ifftn over the spatial axes)
``longitudinal_projection_ft``     ``backend/__init__.py:606-649`` (DC kept as longitudinal, comps >= ndim zero)
``green``                          ``solver.py:564-584``
``chi_mul``                        ``backend/torch.py:220-249``
``born_field``                     ``solver.py:1070-1085`` (E <- E + chi [G (chi E + s) - E]), fixed alpha
=================================  ==============================================================================

The reference torch backend itself cannot be imported in this environment (its package import needs modules
that are not installed); this restatement is pinned against the NumPy oracle (same iterates) and against
central finite differences of the converged field in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import numpy as np
import torch

__all__ = ['chi_from_eps', 'source_from_S', 'ft', 'ift', 'longitudinal_projection_ft', 'green', 'chi_mul',
           'born_field']


def chi_from_eps(eps, alpha: complex, beta: float):
    """``solver.py:480-484``; eps [1,1,*shape] or [3,3,*shape] complex128 tensor."""
    chi = eps * (1.0j / alpha.imag / beta)
    if eps.shape[0] == 1:
        return chi - alpha * 1.0j / alpha.imag
    eye = torch.eye(3, dtype=eps.dtype).reshape(3, 3, *([1] * (eps.ndim - 2)))
    return chi - eye * (alpha * 1.0j / alpha.imag)


def source_from_S(S, alpha: complex, beta: float):
    """``solver.py:702-720``: s = i S / (alpha_i beta)."""
    return S * (1.0j / alpha.imag / beta)


def ft(x):
    return torch.fft.fftn(x, dim=tuple(range(2, x.ndim)))


def ift(x):
    return torch.fft.ifftn(x, dim=tuple(range(2, x.ndim)))


def _kb(k, axis, ndim):
    return torch.as_tensor(np.asarray(k[axis], dtype=float)).reshape([-1 if a == axis else 1 for a in range(ndim)])


def longitudinal_projection_ft(F, k):
    """Pi_L F = k (k . F) / k^2, with F at k = 0 kept as longitudinal (``backend/__init__.py:606-649``)."""
    shape = F.shape[2:]
    nd = min(len(shape), len(k))
    kb = [_kb(k, a, len(shape)).to(F.dtype) for a in range(nd)]
    proj = sum(kb[a] * F[a, 0] for a in range(nd))
    k2 = sum(_ ** 2 for _ in kb).expand(*shape)
    dc = torch.zeros(shape, dtype=torch.bool)
    dc[(0,) * len(shape)] = True
    k2safe = torch.where(dc, torch.ones_like(k2), k2)
    comps = []
    for a in range(F.shape[0]):
        if a < nd:
            comps.append(torch.where(dc, F[a, 0], kb[a] * proj / k2safe))
        else:
            comps.append(torch.zeros_like(F[a, 0]))
    return torch.stack(comps)[:, None]


def green(X, k, alpha: complex, vectorial: bool):
    """``solver.py:564-584``: G X = ift(g ft(X)), dyadic for vectorial fields."""
    F = ft(X)
    shape = X.shape[2:]
    k2 = sum(_kb(k, a, len(shape)) ** 2 for a in range(len(shape))).to(X.dtype)
    g = -1.0j * alpha.imag / (k2 - alpha)
    if vectorial:
        PL = longitudinal_projection_ft(F, k)
        F = g * (F - PL) + PL * (1.0j * alpha.imag / alpha)
    else:
        F = g * F
    return ift(F)


def chi_mul(chi, X):
    """``backend/torch.py:220-249``: scalar broadcast or per-voxel 3x3 mat-vec."""
    if chi.shape[0] == 1 and chi.shape[1] == 1:
        return chi * X
    return torch.einsum('ij...,jk...->ik...', chi, X)


def born_field(eps, S, alpha: complex, beta: float, k, vectorial: bool, n_iter: int, E0=None):
    """The k0^2-scaled field after n_iter Born iterations at fixed alpha, differentiable in eps and S."""
    chi = chi_from_eps(eps, alpha, beta)
    s = source_from_S(S, alpha, beta)
    E = torch.zeros_like(s) if E0 is None else E0
    for _ in range(n_iter):
        d = chi_mul(chi, green(chi_mul(chi, E) + s, k, alpha, vectorial) - E)
        E = E + d
    return E
