"""Born iteration for magnetic and bi-(an)isotropic media (MacroMax ``solver.py:270-493, 496-552``).

With permeability mu (not a constant), or coupling tensors xi, zeta, the susceptibility operator has curl terms:

    chi E = chi_EE E + chi_HE curl(E) + curl(chi_EH E + chi_HH curl(E)),

(curls with respect to k0 x, as the reference's backend, whose grid is in units of 1/k0), with

    chi_EE = (i/alpha_i) ((eps - xi mu^-1 zeta)/beta - alpha I),    chi_EH = -i xi mu^-1 (i/alpha_i)/beta,
    chi_HE = i mu^-1 zeta (i/alpha_i)/beta,                        chi_HH = (I - mu^-1/beta) (i/alpha_i),

and alpha_r, beta chosen by the reference's Nelder-Mead search over (alpha_r, sqrt(beta)) of
beta * (sigma_D^2 sigma_HH(beta) + sigma_D (sigma_EH + sigma_HE)/beta + sigma_EE(alpha_r, beta)) (``solver.py:405-443``).
The largest singular and eigen values over the volume come from this package's FP64 set-up reduction kernel
(``mmx_setup_sigma``: max over voxels of the largest eigenvalue of (M - shift I)^H (M - shift I)). Each iteration is
one ``mmx_magnetic_iterate`` call: the FFTs of 24 field components and five voxel-wise kernels, with chi E + s and
G (chi E + s) kept in k-space, and FP64 norms.
"""
from __future__ import annotations

import logging
import math

import numpy as np
import scipy.optimize
import torch

log = logging.getLogger(__name__)


def _full(t: torch.Tensor, shape) -> torch.Tensor:
    return t.expand(t.shape[0], t.shape[1], *shape).contiguous()


class MagneticOperator:
    """chi and G of a magnetic medium for one backend; owns the (current) alpha and the normalised source."""

    def __init__(self, BE, plan, grid_shape, sample_pitch, wavenumber, epsilon, xi, zeta, mu, bound_offset):
        """epsilon, xi, zeta, mu: complex128 [n, n, *grid-broadcastable] device tensors (n = 1 or 3);
        ``bound_offset(alpha) -> alpha`` applies the kernel-width bias of ``solver.py:591-627``."""
        self.BE = BE
        self.plan = plan
        shape = tuple(int(_) for _ in grid_shape)
        self.shape = shape
        eye3 = torch.eye(3, dtype=torch.complex128, device=BE.device).reshape(3, 3, *([1] * len(shape)))

        def mat(t):
            return t if t.shape[0] == 3 else eye3 * t

        def mmul(a, b):
            if a.shape[0] == 1 and b.shape[0] == 1:
                return a * b
            a, b = mat(a), mat(b)
            return torch.einsum('ij...,jk...->ik...', a, b)

        if mu.shape[0] == 1:
            mu_inv = 1.0 / mu
        else:
            m = _full(mu, shape)
            mu_inv = torch.linalg.inv(m.movedim(0, -1).movedim(0, -1)).movedim(-1, 0).movedim(-1, 0)
        chiEH_beta = -1.0j * mmul(xi, mu_inv)
        chiHE_beta = 1.0j * mmul(mu_inv, zeta)
        eps_eff = epsilon - mmul(xi, -1.0j * chiHE_beta)  # eps - xi mu^-1 zeta
        if eps_eff.shape[0] == 1 and epsilon.shape[0] == 3:
            eps_eff = mat(eps_eff)

        eps_full = _full(eps_eff, shape)
        mu_inv_full = _full(mu_inv, shape)
        sigma_d = float(np.linalg.norm(2.0 * np.pi / (2.0 * np.asarray(sample_pitch)))) / wavenumber
        sigma_he = math.sqrt(plan.setup_sigma(_full(chiHE_beta, shape), 0.0))
        sigma_eh = math.sqrt(plan.setup_sigma(_full(chiEH_beta, shape), 0.0))

        def sigma_hh(beta):  # calc_sigmaHH, solver.py:360-366: || mu^-1 - beta || / beta
            return math.sqrt(plan.setup_sigma(mu_inv_full, float(beta))) / abs(beta)

        def sigma_ee(alpha_r, beta):  # calc_sigmaEE, solver.py:392-401
            return math.sqrt(plan.setup_sigma(eps_full, float(alpha_r) * float(beta))) / abs(beta)

        def total(eta, beta):
            return sigma_d ** 2 * sigma_hh(beta) + sigma_d * (sigma_eh + sigma_he) / abs(beta) + sigma_ee(eta, beta)

        def target(vec):
            beta = vec[1] ** 2  # positive
            return total(vec[0], beta) * beta + max(0.0, beta - 1e-3) ** 2

        vec = scipy.optimize.fmin(target, [0, 1], initial_simplex=[[0, 1], [1, 1], [0, 0.9]], disp=False,
                                  full_output=False, ftol=0.01, xtol=0.01, maxiter=100, maxfun=200)
        beta = float(vec[1] ** 2)
        alpha = complex(vec[0], total(vec[0], beta))
        alpha = bound_offset(alpha)
        log.info(f'Magnetic preconditioner: alpha = {alpha.real:0.4g} + {alpha.imag:0.4g}i, beta = {beta:0.4g}')
        self.alpha = alpha
        self.beta = beta
        hw = BE.hardware_dtype

        def prep(t):  # 3x3 fields are kept expanded and contiguous for the HIP 3x3 multiply
            t = t.to(hw)
            return _full(t, shape) if t.shape[0] == 3 else t
        f = 1.0j / alpha.imag
        self.chiEH = prep(chiEH_beta * (f / beta))
        self.chiHE = prep(chiHE_beta * (f / beta))
        chiHH = (eye3 - mat(mu_inv) / beta) if (mu_inv.shape[0] == 3) else (1.0 - mu_inv / beta)
        self.chiHH = prep(chiHH * f)
        chiEE = eps_eff * (f / beta)
        chiEE = chiEE - alpha * f if chiEE.shape[0] == 1 else chiEE - eye3 * (alpha * f)
        self.chiEE = prep(chiEE)
        self.mu_inv = prep(mu_inv)
        self.chiHE_beta = prep(chiHE_beta)
        self.s = None
        self._g = None  # (alpha, g) of the Green's operator
        self._ops = None  # (version, susceptibilities as the fused kernels take them)
        self._work = None  # 12 grid-sized components of scratch for mmx_magnetic_iterate
        self.version = 0  # incremented whenever s, alpha or a susceptibility tensor is replaced

    # ------------------------------------------------------------------ operators
    def set_source(self, S: torch.Tensor):
        """s = S (i/alpha_i)/beta for the current alpha, [3, 1, *grid]."""
        self.s = (S.to(torch.complex128) * (1.0j / self.alpha.imag / self.beta)).to(self.BE.hardware_dtype).contiguous()
        self.version += 1

    def source_distribution(self) -> torch.Tensor:
        return self.s * (self.beta * self.alpha.imag / 1.0j)

    def _mul(self, m, x):
        if m.shape[0] == 1:
            return m * x
        return self.BE.mul(m, x.contiguous())

    def _mul_add(self, m1, x1, m2, x2):
        """m1 x1 + m2 x2; two element-wise kernels when both factors are scalar fields."""
        if m1.shape[0] == 1 and m2.shape[0] == 1:
            return torch.addcmul(m1 * x1, m2, x2)
        return self._mul(m1, x1) + self._mul(m2, x2)

    def chi_op(self, E: torch.Tensor) -> torch.Tensor:
        """chi_EE E + chi_HE curl(E) + curl(chi_EH E + chi_HH curl(E)) (``solver.py:532-552``)."""
        BE = self.BE
        ED = BE.curl(E)
        chiE = self._mul_add(self.chiEE, E, self.chiHE, ED)
        chiH = self._mul_add(self.chiEH, E, self.chiHH, ED)
        return chiE + BE.curl(chiH)

    def green(self, X: torch.Tensor) -> torch.Tensor:
        """G X = ift(g (F - P_L F) + (i alpha_i / alpha) P_L F), F = ft(X), g = -i alpha_i / (k^2 - alpha)
        (``solver.py:566-584``)."""
        BE = self.BE
        a = self.alpha
        F = BE.ft(X)
        if self._g is None or self._g[0] != a:  # g depends only on alpha: kept until alpha changes
            self._g = (a, ((-1.0j * a.imag) / (BE.k2.to(torch.complex128) - a)).to(BE.hardware_dtype))
        g = self._g[1]
        PL = BE.longitudinal_projection_ft(F)
        F = torch.addcmul(PL * complex(1.0j * a.imag / a), g, F - PL)  # g (F - PL) + (i alpha_i / alpha) PL
        return BE.ift(F)

    def update(self, E: torch.Tensor) -> torch.Tensor:
        """d = chi (G (chi E + s) - E), composed of the backend's curl, Green's and element-wise operations."""
        return self.chi_op(self.green(self.chi_op(E) + self.s) - E)

    def _kernel_operand(self, t: torch.Tensor) -> torch.Tensor:
        """A susceptibility as mmx_magnetic_iterate takes it: constant (one element per entry) or spanning the grid."""
        n = int(t.shape[0])
        if t.numel() == n * n:
            return t.reshape(n, n).contiguous()
        return _full(t, self.shape)

    def iterate(self, E: torch.Tensor, Eout: torch.Tensor, norms: torch.Tensor, norm_stream=None, slot: int = 0,
                norms_host=None):
        """Eout = E + d, norms = (|d|^2, |Eout|^2) in FP64 with the fused kernel sequence (mmx_magnetic_iterate:
        24 transformed components and five voxel-wise kernels per iteration). With ``norm_stream`` the norms are
        reduced on that stream from partial slot ``slot`` and copied to ``norms_host`` there
        (mmx_magnetic_iterate_split)."""
        if self._ops is None or self._ops[0] != self.version:
            self._ops = (self.version, [self._kernel_operand(t) for t in (self.chiEE, self.chiHE, self.chiEH,
                                                                          self.chiHH)])
        if self._work is None:
            self._work = torch.empty((12, *self.shape), dtype=self.BE.hardware_dtype, device=self.BE.device)
        self.plan.magnetic_iterate(E, Eout, self.s, self._ops[1], self.alpha, self._work, norms,
                                   norm_stream=norm_stream, slot=slot, norms_host=norms_host)

    def set_alpha(self, alpha_new: complex):
        """``solver.py:496-529`` for a new alpha (after a rejected update)."""
        a_old = self.alpha
        r = a_old.imag / alpha_new.imag
        self.s = self.s * r
        self.chiEH = self.chiEH * r
        self.chiHE = self.chiHE * r
        self.chiHH = self.chiHH * r
        chiEE = self.chiEE.to(torch.complex128) / (1.0j / a_old.imag)
        if chiEE.shape[0] == 1:
            chiEE = chiEE + (a_old - alpha_new)
        else:
            eye3 = torch.eye(3, dtype=torch.complex128, device=self.BE.device).reshape(3, 3, *([1] * len(self.shape)))
            chiEE = chiEE + eye3 * (a_old - alpha_new)
        self.chiEE = (chiEE * (1.0j / alpha_new.imag)).to(self.BE.hardware_dtype).contiguous()
        self.alpha = alpha_new
        self.version += 1

    def H(self, E: torch.Tensor, c_mu0: float) -> torch.Tensor:
        """i/(mu_0 c) (-mu^-1 curl E + chi_HE,beta E) for the SI field E [3, 1, *grid] (``solver.py:862-884``)."""
        curlE = self.BE.curl(E)
        return (1.0j / c_mu0) * (-self._mul(self.mu_inv, curlE) + self._mul(self.chiHE_beta, E))
